"""ctypes binding of libmoribs_pimc.so (include/moribs_pimc.h).

The library is the only compute path: if it is missing or no HIP device is present the
constructor raises; nothing here falls back to a CPU implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmoribs_pimc.so")

FLAG_FP64 = 0x1
FLAG_REF_ORDER = 0x2
FLAG_PROFILE = 0x4
MAX_TYPES = 4

_d = ctypes.c_double
_i = ctypes.c_int
_u = ctypes.c_uint
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)

# surrogate network (moribs-pimc_amd/csrc/pimc_mlp.hpp): 4 -> 128 -> 128 -> 128 -> 1, SiLU
MLP_IN, MLP_H = 4, 128
MLP_NW = MLP_H * MLP_IN + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) + MLP_H + 1
MLP_NNORM = 2 * MLP_IN + 2


class ReplicaDesc(ctypes.Structure):
    _fields_ = [
        ("ntypes", _i),
        ("numb", _i * MAX_TYPES),
        ("molecule", _i * MAX_TYPES),
        ("levels", _i * MAX_TYPES),
        ("mcstep", _d * MAX_TYPES),
        ("mass", _d * MAX_TYPES),
        ("pot1d", _i * MAX_TYPES),
        ("pot2d", _i * MAX_TYPES),
        ("rotdens", _i),
        ("P", _i),
        ("Prot", _i),
        ("T", _d),
        ("rtstep", _d),
        ("pot3d", _i),
        ("rot3d", _i),
    ]


# every entry point declared in include/moribs_pimc.h, with its ctypes signature
SIGNATURES = {
    "pimc_create": (_i, [_i, _u, _u, ctypes.POINTER(ctypes.c_void_p)]),
    "pimc_destroy": (None, [ctypes.c_void_p]),
    "pimc_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "pimc_device_name": (_i, [ctypes.c_char_p, _i]),
    "pimc_add_pot1d": (_i, [ctypes.c_void_p, _i, _dp, _dp]),
    "pimc_add_pot2d": (_i, [ctypes.c_void_p, _i, _i, _d, _d, _dp, _dp, _dp]),
    "pimc_add_rotdens": (_i, [ctypes.c_void_p, _i, _dp, _dp, _dp, _dp]),
    "pimc_add_pot3d": (_i, [ctypes.c_void_p, _i, _i, _i, _d, _d, _dp]),
    "pimc_add_rot3d": (_i, [ctypes.c_void_p, _dp, _dp, _dp]),
    "pimc_linden": (_i, [ctypes.c_void_p, ctypes.c_double, _i, ctypes.c_double, _i, _i, _i, _dp, _dp, _dp, _dp]),
    "pimc_asymrho": (_i, [ctypes.c_void_p, ctypes.c_double, _i, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i, _i,
                          _dp, _dp, _dp]),
    "pimc_set_replica": (_i, [ctypes.c_void_p, _i, ctypes.POINTER(ReplicaDesc)]),
    "pimc_commit": (_i, [ctypes.c_void_p]),
    "pimc_check_replica": (_i, [ctypes.POINTER(ReplicaDesc), ctypes.c_char_p, _i]),
    "pimc_set_minimum_image": (_i, [ctypes.c_void_p, _i, _dp]),
    "pimc_plan_rotor3d_split": (_i, [_i, _i, _i, _i, _i, _i, _i]),
    "pimc_plan_rotor_parts": (_i, [_i, _ip, _ip, _ip, _i, _i, _i, _i, _i, _i, _ip, _ip, ctypes.POINTER(ctypes.c_long),
                                   ctypes.c_char_p, _i]),
    "pimc_set_option": (_i, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_long]),
    "pimc_set_state": (_i, [ctypes.c_void_p, _i, _dp, _dp]),
    "pimc_get_state": (_i, [ctypes.c_void_p, _i, _dp, _dp, _dp]),
    "pimc_restore_state": (_i, [ctypes.c_void_p, _i, _dp, _dp, _dp, _u]),
    "pimc_get_pass": (_i, [ctypes.c_void_p, _i, ctypes.POINTER(ctypes.c_uint)]),
    "pimc_run": (_i, [ctypes.c_void_p, _i, _i]),
    "pimc_autotune": (_i, [ctypes.c_void_p, _i, ctypes.POINTER(ctypes.c_long)]),
    "pimc_get_samples": (_i, [ctypes.c_void_p, _i, _dp, _i]),
    "pimc_clear_samples": (_i, [ctypes.c_void_p]),
    "pimc_get_counters": (_i, [ctypes.c_void_p, _i, _dp, _dp]),
    "pimc_get_histograms": (_i, [ctypes.c_void_p, _i, _dp, _dp]),
    "pimc_get_histograms3d": (_i, [ctypes.c_void_p, _i, _dp, _dp]),
    "pimc_get_work": (_i, [ctypes.c_void_p, _dp, _dp]),
    "pimc_get_rcf": (_i, [ctypes.c_void_p, _i, _dp, _i]),
    "pimc_pot_energy_slice": (_i, [ctypes.c_void_p, _i, _ip, _ip, _ip, _dp, _dp]),
    "pimc_pot_energy_slice_noalloc": (_i, [ctypes.c_void_p, _i, _ip, _ip, _ip, _dp, _dp]),
    "pimc_estimate": (_i, [ctypes.c_void_p, _dp]),
    "pimc_kernel_time": (_i, [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_long)]),
    "pimc_synchronize": (_i, [ctypes.c_void_p]),
    "pimc_family_evals": (_i, [ctypes.c_void_p, ctypes.c_char_p, _dp]),
    "pimc_surrogate_collect": (_i, [ctypes.c_void_p, _i, _fp, _fp, _i]),
    "pimc_surrogate_train": (_i, [ctypes.c_void_p, _fp, _fp, _i, _i, _i, _d, _u, _dp]),
    "pimc_surrogate_grad": (_i, [ctypes.c_void_p, _fp, _fp, _fp, _i, _fp, _dp]),
    "pimc_surrogate_set": (_i, [ctypes.c_void_p, _fp, _fp]),
    "pimc_surrogate_get": (_i, [ctypes.c_void_p, _fp, _fp]),
    "pimc_surrogate_eval": (_i, [ctypes.c_void_p, _fp, _i, _fp]),
    "pimc_get_da_counters": (_i, [ctypes.c_void_p, _i, _dp]),
}


def replica_desc(ntypes, numb, molecule, levels, mcstep, mass, P, Prot, T, rtstep, pot1d=None, pot2d=None, rotdens=-1,
                 pot3d=-1, rot3d=-1):
    """A pimc_replica_desc from the qmc.input quantities of one replica."""
    d = ReplicaDesc()
    d.ntypes = ntypes
    for t in range(MAX_TYPES):
        d.numb[t] = numb[t] if t < ntypes else 0
        d.molecule[t] = molecule[t] if t < ntypes else 0
        d.levels[t] = levels[t] if t < ntypes else 0
        d.mcstep[t] = mcstep[t] if t < ntypes else 0.0
        d.mass[t] = mass[t] if t < ntypes else 1.0
        d.pot1d[t] = (pot1d[t] if pot1d is not None and t < ntypes else -1)
        d.pot2d[t] = (pot2d[t] if pot2d is not None and t < ntypes else -1)
    d.rotdens = rotdens
    d.pot3d, d.rot3d = pot3d, rot3d
    d.P, d.Prot, d.T, d.rtstep = P, Prot, T, rtstep
    return d


def check_replica(desc, lib=None):
    """pimc_check_replica: '' if the description is accepted, else the reason (no device needed)."""
    lib = lib or load_library()
    msg = ctypes.create_string_buffer(256)
    rc = lib.pimc_check_replica(ctypes.byref(desc), msg, len(msg))
    return "" if rc == 0 else msg.value.decode()


def plan_rotor_parts(prot, rot_ratio, natoms, fp64=False, lds_kb=52, part_slices_max=0, blocks_per_cu=8,
                     lds_per_cu=160 * 1024, cus_per_xcd=32, lib=None):
    """pimc_plan_rotor_parts: (parts per chain, slices per part, launch LDS bytes) of the linear-rotor
    sweep's split, for a device that holds blocks_per_cu parts per CU by registers (no device needed).
    Raises PimcError with the library's message when no plan fits."""
    lib = lib or load_library()
    n = len(prot)
    a = [np.ascontiguousarray(x, dtype=np.int32) for x in (prot, rot_ratio, natoms)]
    W = np.zeros(n, np.int32)
    ns = np.zeros(n, np.int32)
    lds = ctypes.c_long(0)
    msg = ctypes.create_string_buffer(256)
    ip = lambda x: x.ctypes.data_as(_ip)  # noqa: E731
    rc = lib.pimc_plan_rotor_parts(n, ip(a[0]), ip(a[1]), ip(a[2]), int(bool(fp64)), int(lds_kb), int(part_slices_max),
                                   int(blocks_per_cu), int(lds_per_cu), int(cus_per_xcd), ip(W), ip(ns),
                                   ctypes.byref(lds), msg, len(msg))
    if rc != 0:
        raise PimcError(msg.value.decode())
    return W, ns, lds.value


def plan_rotor3d_split(protmax, nchains, nrot, blocks_per_cu, lds_per_cu=160 * 1024, ncu=256, cus_per_xcd=32, lib=None):
    """pimc_plan_rotor3d_split: parts per non-linear rotor chain (no device needed)."""
    lib = lib or load_library()
    return lib.pimc_plan_rotor3d_split(int(protmax), int(nchains), int(nrot), int(blocks_per_cu), int(lds_per_cu),
                                       int(ncu), int(cus_per_xcd))


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _fptr(a):
    return a.ctypes.data_as(_fp)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class PimcError(RuntimeError):
    pass


class Context:
    """A set of replicas sampled on the GPU through the C ABI."""

    def __init__(self, n_replicas, flags=0, seed=985456376, lib=None):
        self.lib = lib or load_library()
        h = ctypes.c_void_p()
        rc = self.lib.pimc_create(int(n_replicas), int(flags), int(seed) & 0xFFFFFFFF, ctypes.byref(h))
        if rc != 0:
            raise PimcError(f"pimc_create failed ({rc}): no usable HIP device")
        self.h = h
        self.n = int(n_replicas)
        self.descs = [None] * self.n

    def close(self):
        if self.h:
            self.lib.pimc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise PimcError(self.lib.pimc_last_error(self.h).decode())
        return rc

    # tables
    def add_pot1d(self, r, v):
        r, v = _f64(r), _f64(v)
        return self._check(self.lib.pimc_add_pot1d(self.h, len(r), _dptr(r), _dptr(v)))

    def add_pot2d(self, rgrid, cgrid, dr, dc, v):
        rg, cg, vv = _f64(rgrid), _f64(cgrid), _f64(v).reshape(-1)
        return self._check(self.lib.pimc_add_pot2d(self.h, len(rg), len(cg), float(dr), float(dc), _dptr(rg), _dptr(cg), _dptr(vv)))

    def add_rotdens(self, cosg, rho, erot, esq):
        a = [_f64(x) for x in (cosg, rho, erot, esq)]
        return self._check(self.lib.pimc_add_rotdens(self.h, len(a[0]), *[_dptr(x) for x in a]))

    def add_pot3d(self, nr, nth, nch, rmin, rmax, v):
        v = _f64(v).reshape(-1)
        if v.size != nr * nth * nch:
            raise PimcError("3D table size does not match its header")
        return self._check(self.lib.pimc_add_pot3d(self.h, nr, nth, nch, float(rmin), float(rmax), _dptr(v)))

    def add_rot3d(self, rho, erot, esq):
        a = [_f64(x).reshape(-1) for x in (rho, erot, esq)]
        if any(x.size != 181 * 361 * 361 for x in a):
            raise PimcError("non-linear density tables hold 181*361*361 values each")
        return self._check(self.lib.pimc_add_rot3d(self.h, *[_dptr(x) for x in a]))

    # density-matrix generators on the device
    def linden(self, T, nslice, bconst, npt=2001, iodevn=-1, lmax=0):
        """The four .rot columns (cos, rho, E, E^2 numerators) of linden.f:88-122 and the l cut-off used."""
        out = [np.zeros(npt) for _ in range(4)]
        used = self._check(self.lib.pimc_linden(self.h, float(T), int(nslice), float(bconst), int(npt), int(iodevn),
                                                int(lmax), *[_dptr(x) for x in out]))
        return (*out, used)

    def asymrho(self, T, nslice, A, B, C, maxj, iodevn=-1):
        """rho, E, E^2 of asymrho.f:34-731 on the (181, 361, 361) whole-degree grid."""
        out = [np.zeros(181 * 361 * 361) for _ in range(3)]
        self._check(self.lib.pimc_asymrho(self.h, float(T), int(nslice), float(A), float(B), float(C), int(maxj),
                                          int(iodevn), *[_dptr(x) for x in out]))
        return tuple(x.reshape(181, 361, 361) for x in out)

    # replicas
    def set_replica(self, r, ntypes, numb, molecule, levels, mcstep, mass, P, Prot, T, rtstep, pot1d=None, pot2d=None,
                    rotdens=-1, pot3d=-1, rot3d=-1):
        d = replica_desc(ntypes, numb, molecule, levels, mcstep, mass, P, Prot, T, rtstep, pot1d, pot2d, rotdens, pot3d,
                         rot3d)
        self.descs[r] = d
        return self._check(self.lib.pimc_set_replica(self.h, r, ctypes.byref(d)))

    def set_minimum_image(self, r, box):
        """pimc_set_minimum_image: periodic box edges (3 values, or one for a cube), None for off."""
        if box is None:
            return self._check(self.lib.pimc_set_minimum_image(self.h, r, None))
        b = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)).copy()
        return self._check(self.lib.pimc_set_minimum_image(self.h, r, _dptr(b)))

    def set_option(self, name, value):
        self._check(self.lib.pimc_set_option(self.h, name.encode(), int(value)))

    def commit(self):
        return self._check(self.lib.pimc_commit(self.h))

    def natoms(self, r):
        d = self.descs[r]
        return sum(d.numb[t] for t in range(d.ntypes))

    def set_state(self, r, coords, angles=None):
        c = _f64(coords)
        a = _f64(angles) if angles is not None else None
        return self._check(self.lib.pimc_set_state(self.h, r, _dptr(c), _dptr(a) if a is not None else None))

    def get_state(self, r):
        n = self.natoms(r) * self.descs[r].P
        c = np.empty(3 * n)
        a = np.empty(3 * n)
        cs = np.empty(3 * n)
        self._check(self.lib.pimc_get_state(self.h, r, _dptr(c), _dptr(a), _dptr(cs)))
        return c.reshape(3, n), a.reshape(3, n), cs.reshape(3, n)

    def restore_state(self, r, coords, angles, cosines, pass_counter):
        coords, angles, cosines = _f64(coords), _f64(angles), _f64(cosines)
        self._check(self.lib.pimc_restore_state(self.h, int(r), _dptr(coords), _dptr(angles), _dptr(cosines),
                                                int(pass_counter)))

    def pass_counter(self, r):
        p = ctypes.c_uint()
        self._check(self.lib.pimc_get_pass(self.h, int(r), ctypes.byref(p)))
        return p.value

    def run(self, n_passes, sample_every=0):
        return self._check(self.lib.pimc_run(self.h, int(n_passes), int(sample_every)))

    def autotune(self, passes=2):
        """Launch shapes picked by timing candidates on this workload (state restored after)."""
        out = (ctypes.c_long * 5)()
        self._check(self.lib.pimc_autotune(self.h, int(passes), out))
        return dict(bisect_lanes_small=out[0], bisect_lanes_medium=out[1], bisect_lanes_large=out[2],
                    rotor_lds_kb=out[3], centroid_parts=out[4])

    def samples(self, r, max_samples=1 << 16):
        out = np.empty(8 * max_samples)
        n = self._check(self.lib.pimc_get_samples(self.h, r, _dptr(out), max_samples))
        return out[: 8 * n].reshape(n, 8)

    def clear_samples(self):
        return self._check(self.lib.pimc_clear_samples(self.h))

    def counters(self, r):
        nt = self.descs[r].ntypes
        tot = np.zeros(3 * nt)
        acc = np.zeros(3 * nt)
        self._check(self.lib.pimc_get_counters(self.h, r, _dptr(tot), _dptr(acc)))
        return tot.reshape(nt, 3), acc.reshape(nt, 3)

    def histograms(self, r):
        g1 = np.zeros(300)
        g2 = np.zeros(300 * 50)
        self._check(self.lib.pimc_get_histograms(self.h, r, _dptr(g1), _dptr(g2)))
        return g1, g2.reshape(300, 50)

    def histograms3d(self, r):
        """Bins around a non-linear rotor since the last clear_samples: the (r, theta, chi)
        counts as [300][50][100] (None without them) and the relative Euler angle counts as
        (theta[50], phi[100], chi[100]) (None without them)."""
        g3 = np.zeros(300 * 50 * 100)
        rel = np.zeros(250)
        has = self._check(self.lib.pimc_get_histograms3d(self.h, int(r), _dptr(g3), _dptr(rel)))
        return (g3.reshape(300, 50, 100) if has & 1 else None,
                (rel[:50], rel[50:150], rel[150:]) if has & 2 else None)

    def rcf(self, r):
        """[10][Prot] rotational correlation sums since the last clear_samples (None: no rotor)."""
        d = self.descs[r]
        prot = d.Prot if d.Prot > 0 else d.P
        out = np.zeros((10, prot))
        n = self._check(self.lib.pimc_get_rcf(self.h, int(r), _dptr(out), out.size))
        return out if n else None

    def work(self):
        pe, bm = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.pimc_get_work(self.h, ctypes.byref(pe), ctypes.byref(bm)))
        return pe.value, bm.value

    def pot_energy_slice(self, reps, atoms, its, pos, noalloc=False):
        """PotEnergy(atom, pos, it) per query; noalloc=True stages the queries in the library's
        per-device array (pimc_pot_energy_slice_noalloc) instead of allocating."""
        reps = np.ascontiguousarray(reps, dtype=np.int32)
        atoms = np.ascontiguousarray(atoms, dtype=np.int32)
        its = np.ascontiguousarray(its, dtype=np.int32)
        pos = _f64(pos).reshape(-1)
        out = np.empty(len(reps))
        ip = lambda a: a.ctypes.data_as(_ip)
        f = self.lib.pimc_pot_energy_slice_noalloc if noalloc else self.lib.pimc_pot_energy_slice
        self._check(f(self.h, len(reps), ip(reps), ip(atoms), ip(its), _dptr(pos), _dptr(out)))
        return out

    def estimate(self):
        out = np.empty(8 * self.n)
        self._check(self.lib.pimc_estimate(self.h, _dptr(out)))
        return out.reshape(self.n, 8)

    def kernel_time(self, family):
        ms, n = ctypes.c_double(), ctypes.c_long()
        self._check(self.lib.pimc_kernel_time(self.h, family.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def family_evals(self, family):
        e = ctypes.c_double()
        self._check(self.lib.pimc_family_evals(self.h, family.encode(), ctypes.byref(e)))
        return e.value

    def synchronize(self):
        return self._check(self.lib.pimc_synchronize(self.h))

    # surrogate network and delayed acceptance
    def surrogate_collect(self, r):
        """(features [n,4], exact pair energies [n]) of replica r's current configuration."""
        d = self.descs[r]
        nrot = sum(d.numb[t] for t in range(d.ntypes) if d.molecule[t] != 0)
        nat = sum(d.numb[t] for t in range(d.ntypes) if d.molecule[t] == 0)
        # rotor-atom rows with atoms, rotor-rotor rows without (pimc_surrogate_collect)
        n = nrot * d.Prot * (d.P // d.Prot) * (nat if nat > 0 else nrot - 1)
        x = np.zeros((n, MLP_IN), np.float32)
        v = np.zeros(n, np.float32)
        got = self._check(self.lib.pimc_surrogate_collect(self.h, int(r), _fptr(x), _fptr(v), n))
        return x[:got], v[:got]

    def surrogate_train(self, x, v, epochs=20, batch=4096, lr=2e-3, seed=1):
        x, v = _f32(x).reshape(-1, MLP_IN), _f32(v)
        loss = np.zeros(epochs)
        self._check(self.lib.pimc_surrogate_train(self.h, _fptr(x), _fptr(v), len(v), int(epochs), int(batch),
                                                  float(lr), int(seed), _dptr(loss)))
        return loss

    def surrogate_grad(self, theta, x, y):
        """Loss and gradient of the raw network (no normalisation) on rows x, targets y."""
        theta, x, y = _f32(theta), _f32(x).reshape(-1, MLP_IN), _f32(y)
        g = np.zeros(MLP_NW, np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.pimc_surrogate_grad(self.h, _fptr(theta), _fptr(x), _fptr(y), len(y), _fptr(g),
                                                 ctypes.byref(loss)))
        return loss.value, g

    def surrogate_set(self, theta, norm):
        theta, norm = _f32(theta), _f32(norm)
        assert theta.size == MLP_NW and norm.size == MLP_NNORM
        self._check(self.lib.pimc_surrogate_set(self.h, _fptr(theta), _fptr(norm)))

    def surrogate_get(self):
        theta, norm = np.zeros(MLP_NW, np.float32), np.zeros(MLP_NNORM, np.float32)
        self._check(self.lib.pimc_surrogate_get(self.h, _fptr(theta), _fptr(norm)))
        return theta, norm

    def surrogate_eval(self, x):
        x = _f32(x).reshape(-1, MLP_IN)
        out = np.zeros(len(x), np.float32)
        self._check(self.lib.pimc_surrogate_eval(self.h, _fptr(x), len(x), _fptr(out)))
        return out

    def da_counters(self, r):
        """Delayed acceptance: proposals, stage-1 passes, accepted moves, network evaluations."""
        out = np.zeros(4)
        self._check(self.lib.pimc_get_da_counters(self.h, int(r), _dptr(out)))
        return out
