"""ctypes access to the oracle (liboracle.so) and to the compiled reference harness
(_ref/libref.so).  TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, never by the product."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
REF_PIMC = os.path.join(HERE, "_ref", "pimc")

ORDER_REFERENCE = 0
ORDER_SEGMENTS = 1

_d, _i, _u, _l = ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_long
_dp, _ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype, f.argtypes = res, args


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        for name, res, args in [
            ("or_philox4x32_10", None, [u32p, u32p, u32p]),
            ("or_spline", None, [_dp, _dp, _i, _d, _d, _dp]),
            ("or_splint", _d, [_dp, _dp, _dp, _i, _d]),
            ("or_init_spline", None, [_dp, _dp, _dp, _i]),
            ("or_matpre", None, [_dp, _dp]),
            ("or_deleul", None, [_dp, _dp, _dp, _ip]),
            ("or_rotden", None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
            ("or_vcord", None, [_dp, _dp, _dp, _dp, _i, _i, _i, _d, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i]),
            ("or_vcalc", None, [_d, _d, _d, _d, _d, _d, _i, _i, _i, _dp, _dp]),
            ("or_caleng", None, [_dp, _dp, _dp, _dp, _dp]),
            ("or_linden_rho", _d, [_d, _d, _d, _i, _i, _i, _dp, _dp]),
            ("or_linden_lmax", _i, [_d, _d]),
            ("or_create", _vp, [_i, _ip, _ip, _ip, _dp, _dp, _i, _i, _d, _d, _i]),
            ("or_destroy", None, [_vp]),
            ("or_last_error", ctypes.c_char_p, [_vp]),
            ("or_set_pot1d", _i, [_vp, _i, _i, _dp, _dp]),
            ("or_set_pot2d", _i, [_vp, _i, _i, _i, _d, _d, _dp, _dp, _dp]),
            ("or_set_rotdens", _i, [_vp, _i, _i, _dp, _dp, _dp, _dp]),
            ("or_set_pot3d", _i, [_vp, _i, _i, _i, _d, _d, _dp]),
            ("or_set_rot3d", _i, [_vp, _dp, _dp, _dp]),
            ("or_set_minimage", None, [_vp, _i, _dp]),
            ("or_set_state", None, [_vp, _dp, _dp]),
            ("or_get_state", None, [_vp, _dp, _dp, _dp]),
            ("or_get_counters", None, [_vp, _dp, _dp]),
            ("or_reset_counters", None, [_vp]),
            ("or_natoms", _i, [_vp]),
            ("or_get_pot1d_extrap", None, [_vp, _i, _dp, _dp, _dp, _dp]),
            ("or_spot1d", _d, [_vp, _d, _i]),
            ("or_lpot2d", _d, [_vp, _d, _d, _i]),
            ("or_srotdens", _d, [_vp, _d, _i]),
            ("or_srotdens_deriv", _d, [_vp, _d, _i]),
            ("or_srotdens_esqrt", _d, [_vp, _d, _i]),
            ("or_pot_energy_slice", _d, [_vp, _i, _dp, _i]),
            ("or_pot_energy_path", _d, [_vp, _i, _dp]),
            ("or_pot_rot_energy", _d, [_vp, _i, _dp, _i]),
            ("or_pot_rot_e3d", _d, [_vp, _i, _dp, _i]),
            ("or_kin_energy", _d, [_vp]),
            ("or_pot_energy_densities", _d, [_vp, _dp, _dp]),
            ("or_rot_energy", _d, [_vp, _dp, _dp]),
            ("or_rot_e3d", _d, [_vp, _dp, _dp]),
            ("or_nl_densities", _i, [_vp, _dp, _dp]),
            ("or_sample", None, [_vp, _dp]),
            ("or_set_rng", None, [_vp, _u, _u]),
            ("or_molecular_move", _i, [_vp, _i, _u]),
            ("or_bisection_move", _i, [_vp, _i, _i, _u]),
            ("or_bisection_atom", _i, [_vp, _i, _i, _i, _u]),
            ("or_rotations_move", _i, [_vp, _i, _u, _u]),
            ("or_rotations_3d", _i, [_vp, _i, _u, _u]),
            ("or_pass", _i, [_vp, _u, _i]),
            ("or_run", _i, [_vp, _u, _i, _i, _i, _dp]),
            ("or_pass_sampled", _i, [_vp, _u, _i, _i, _dp]),
            ("or_set_tape", _i, [_vp, _i, _dp, _l]),
            ("or_tape_pos", _l, [_vp, _i]),
            ("or_tape_short", _i, [_vp]),
            ("or_tape_off", None, [_vp]),
            ("or_uniform", _d, [_u, _u, u32p, _i]),
            ("or_gauss", _d, [_u, _u, u32p, _d]),
        ]:
            _sig(L, name, res, args)
        _lib = L
    return _lib


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def I(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(_ip)


def philox(ctr, key):
    L = lib()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.or_philox4x32_10(c, k, o)
    return list(o)


class System:
    """One replica in the oracle."""

    def __init__(self, rp, tables, seed=985456376, replica=0):
        self.L = lib()
        self._keep = []
        nt = rp["ntypes"]
        self.h = self.L.or_create(nt, I(rp["numb"]), I(rp["molecule"]), I(rp["levels"]),
                                  P(rp["mcstep"]), P(rp["mass"]), rp["P"], rp["Prot"], rp["T"], rp["rtstep"], 0)
        if not self.h:
            raise RuntimeError("or_create failed")
        self.rp = rp
        self.natoms = sum(rp["numb"])
        for t in range(nt):
            if rp["molecule"][t] == 0 and rp["pot1d"][t] >= 0:
                r, v = tables["pot1d"][rp["pot1d"][t]]
                self.L.or_set_pot1d(self.h, t, len(r), P(r), P(v))
            if rp["molecule"][t] == 1:
                rg, cg, dr, dc, V = tables["pot2d"][rp["pot2d"][t]]
                self.L.or_set_pot2d(self.h, t, len(rg), len(cg), dr, dc, P(rg), P(cg), P(np.asarray(V).reshape(-1)))
                c, rho, er, es = tables["rot"][rp["rotdens"]]
                self.L.or_set_rotdens(self.h, t, len(c), P(c), P(rho), P(er), P(es))
            if rp["molecule"][t] == 2:
                # the oracle borrows 3D tables: keep them alive with the system
                if tables.get("pot3d") is not None:
                    nr, nth, nch, rmin, rmax, v = tables["pot3d"]
                    v = np.ascontiguousarray(v, dtype=np.float64)
                    self._keep.append(v)
                    self.L.or_set_pot3d(self.h, nr, nth, nch, rmin, rmax, v.ctypes.data_as(_dp))
                rho, er, es = [np.ascontiguousarray(x, dtype=np.float64) for x in tables["rot3d"]]
                self._keep += [rho, er, es]
                self.L.or_set_rot3d(self.h, rho.ctypes.data_as(_dp), er.ctypes.data_as(_dp), es.ctypes.data_as(_dp))
        self.L.or_set_rng(self.h, seed & 0xFFFFFFFF, replica)
        if rp.get("box") is not None:  # MINIMAGE with this box
            self.L.or_set_minimage(self.h, 1, P(np.broadcast_to(np.asarray(rp["box"], dtype=np.float64), (3,)).copy()))
        if "coords" in rp:
            self.set_state(rp["coords"], rp["angles"])

    def __del__(self):
        try:
            self.L.or_destroy(self.h)
        except Exception:
            pass

    def set_state(self, coords, angles):
        self.L.or_set_state(self.h, P(coords), P(angles))

    def state(self):
        n = self.natoms * self.rp["P"]
        c, a, cs = np.empty(3 * n), np.empty(3 * n), np.empty(3 * n)
        self.L.or_get_state(self.h, P(c) if False else c.ctypes.data_as(_dp), a.ctypes.data_as(_dp), cs.ctypes.data_as(_dp))
        return c.reshape(3, n), a.reshape(3, n), cs.reshape(3, n)

    def counters(self):
        nt = self.rp["ntypes"]
        t, a = np.zeros(3 * nt), np.zeros(3 * nt)
        self.L.or_get_counters(self.h, t.ctypes.data_as(_dp), a.ctypes.data_as(_dp))
        return t.reshape(nt, 3), a.reshape(nt, 3)

    def run(self, pass0, npasses, order=ORDER_SEGMENTS, every=0):
        ns = npasses // every if every else 0
        out = np.zeros(8 * max(ns, 1))
        rc = self.L.or_run(self.h, pass0, npasses, order, every, out.ctypes.data_as(_dp))
        if rc < 0:
            raise RuntimeError(self.L.or_last_error(self.h).decode())
        return out[: 8 * ns].reshape(ns, 8)

    def pass_sampled(self, pass_, order, per_pass):
        """One pass with per_pass samples inside it (or_pass_sampled): [per_pass, 8]."""
        out = np.zeros(8 * per_pass)
        rc = self.L.or_pass_sampled(self.h, pass_, order, per_pass, out.ctypes.data_as(_dp))
        if rc < 0:
            raise RuntimeError(self.L.or_last_error(self.h).decode())
        return out[: 8 * rc].reshape(rc, 8)

    def sample(self):
        out = np.zeros(8)
        self.L.or_sample(self.h, out.ctypes.data_as(_dp))
        return out

    def set_tapes(self, tapes):
        for k, v in tapes.items():
            v = np.ascontiguousarray(v, dtype=np.float64)
            self.L.or_set_tape(self.h, k, v.ctypes.data_as(_dp), len(v))


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The compiled reference harness (oracle/_ref/libref.so); only in the build container."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_SO)
        for name, res, args in [
            ("h_setup", _i, [_i, ctypes.c_char_p, _ip, _ip, _ip, _dp, _dp, _i, _i, _d, _d, ctypes.c_char_p, ctypes.c_char_p]),
            ("h_set_minimage", None, [_i, _dp]),
            ("h_set_state", None, [_dp, _dp]),
            ("h_set_newcoords", None, [_dp]),
            ("h_get_state", None, [_dp, _dp, _dp]),
            ("h_set_rot3d", None, [_dp, _dp, _dp]),
            ("h_spot1d", _d, [_d, _i]),
            ("h_lpot2d", _d, [_d, _d, _i]),
            ("h_srotdens", _d, [_d, _i]),
            ("h_srotdens_deriv", _d, [_d, _i]),
            ("h_srotdens_esqrt", _d, [_d, _i]),
            ("h_pot_energy_slice", _d, [_i, _i, _i]),
            ("h_pot_energy_path", _d, [_i, _i]),
            ("h_pot_rot_energy", _d, [_i, _i, _i]),
            ("h_pot_rot_e3d", _d, [_i, _dp, _i]),
            ("h_rotden", None, [_dp, _dp, _dp, _dp, _dp, _dp, _ip]),
            ("h_vcord", None, [_dp, _dp, _dp, _dp]),
            ("h_caleng", _d, [_dp, _dp, _dp, _dp]),
            ("h_kin_energy", _d, []),
            ("h_pot_energy_densities", _d, [_dp, _dp]),
            ("h_rot_energy", _d, [_dp, _dp]),
            ("h_rot_e3d", _d, [_dp, _dp]),
            ("h_nl_densities", _i, [_dp, _dp]),
            ("h_set_tape", None, [_i, _dp, _l]),
            ("h_tape_pos", _l, [_i]),
            ("h_tape_short", _i, []),
            ("h_molecular_move", None, [_i]),
            ("h_bisection_move", None, [_i, _i]),
            ("h_rotations_move", None, [_i]),
            ("h_rotations_3d", None, [_i]),
            ("h_pimc_pass", None, [_i, _i]),
            ("h_reset_counters", None, []),
            ("h_get_counters", None, [_dp, _dp]),
        ]:
            _sig(L, name, res, args)
        _ref = L
    return _ref
