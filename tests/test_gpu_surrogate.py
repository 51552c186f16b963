"""The surrogate network and the delayed-acceptance rotor move on the GPU.

Kernels against the FP64 oracle (oracle/mlp_oracle.py): the loss and gradient of the training
GEMMs (BF16 operands, FP32 sums), and the sampler's matrix-core evaluation.  Then the learned
part in use: rows collected from the run's own configurations, training, the delayed-acceptance
move (density, network and exact factors) keeping the plain move's ensemble (within two standard
errors) whatever the network, and what a trained network saves against an untrained one."""
import numpy as np
import pytest

import mlp_oracle as M
import pimc_configs as pc
import pimc_hip as ph
from golden import tables3d
from test_gpu_nonlinear import nl_workload

pytestmark = pytest.mark.gpu
SEED = 985456376


@pytest.fixture(scope="module")
def rot3d():
    return tables3d.rot3d()


@pytest.fixture(scope="module")
def ctx1():
    return ph.Context(1, seed=SEED)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def test_gradient_matches_oracle(ctx1):
    """Loss and all 33793 gradient entries of a batch of 512 rows against FP64; BF16 operands
    bound the error at a few 1e-3 of each block's norm."""
    rng = np.random.default_rng(5)
    theta = M.glorot(2) + rng.normal(0, 0.02, M.NW)
    X = rng.normal(0, 1, (512, 4))
    y = rng.normal(0, 1, 512)
    loss, g = ctx1.surrogate_grad(theta, X, y)
    lo, go = M.loss_grad(theta.astype(np.float32).astype(np.float64), X.astype(np.float32), y.astype(np.float32))
    assert abs(loss - lo) <= 1e-2 * lo
    assert rel(g, go) < 1e-2
    for k, (o, s) in M.OFF.items():
        n = int(np.prod(s))
        assert rel(g[o:o + n], go[o:o + n]) < 2e-2, k


def test_gradient_of_a_fitted_batch(ctx1):
    """Near a minimum the gradient is small and the BF16 products must still follow it."""
    rng = np.random.default_rng(6)
    theta = M.glorot(3)
    X = rng.normal(0, 1, (256, 4))
    y = M.forward(theta, X) + rng.normal(0, 1e-2, 256)
    loss, g = ctx1.surrogate_grad(theta, X, y)
    lo, go = M.loss_grad(theta, X, y)
    assert loss < 1e-3 and abs(loss - lo) < 2e-4
    assert rel(g, go) < 0.2


def test_sampler_evaluation_matches_oracle(ctx1):
    rng = np.random.default_rng(7)
    theta = M.glorot(4) + rng.normal(0, 0.02, M.NW)
    norm = np.array([6.0, 0.1, -0.2, 0.3, 1.5, 0.6, 0.6, 0.6, -12.0, 25.0])
    ctx1.surrogate_set(theta, norm)
    X = np.column_stack([rng.uniform(3, 9, 1000), rng.uniform(-1, 1, (1000, 3))])
    got = ctx1.surrogate_eval(X)
    xn = (X - norm[:4]) / norm[4:8]
    th32 = theta.astype(np.float32).astype(np.float64)
    want = norm[9] * M.forward(th32, xn.astype(np.float32), rounding="bf16") + norm[8]
    exact = norm[9] * M.forward(th32, xn) + norm[8]
    scale = np.std(exact)
    assert np.max(np.abs(got - want)) < 2e-3 * scale      # same roundings, FP32 sums
    assert np.max(np.abs(got - exact)) < 3e-2 * scale     # BF16 against FP64
    t2, n2 = ctx1.surrogate_get()
    assert np.array_equal(t2, theta.astype(np.float32)) and np.allclose(n2, norm)


def replicated(rot3d, n_rep, n_h2=4, P=8):
    base = nl_workload("h2o_h2", rot3d, n_h2=n_h2, P=P, Prot=P)
    return dict(tables=base["tables"], replicas=[dict(base["replicas"][0]) for _ in range(n_rep)])


def make_ctx(wl, da=False):
    c = ph.Context(len(wl["replicas"]), seed=SEED)
    if da:
        c.set_option("delayed_acceptance", 1)
    return pc.load_into_context(c, wl)


@pytest.fixture(scope="module")
def trained(rot3d):
    """Rows from a plain run of 64 chains (every 5th pass of 100), and a network trained on them."""
    wl = replicated(rot3d, 64)
    c = make_ctx(wl)
    xs, vs = [], []
    c.run(20)
    for _ in range(20):
        c.run(5)
        for r in range(len(wl["replicas"])):
            x, v = c.surrogate_collect(r)
            xs.append(x)
            vs.append(v)
    X, V = np.concatenate(xs), np.concatenate(vs)
    hold = np.arange(len(V)) % 10 == 0
    loss = c.surrogate_train(X[~hold], V[~hold], epochs=40, batch=512, lr=3e-3, seed=11)
    theta, norm = c.surrogate_get()
    pred = c.surrogate_eval(X[hold])
    return dict(X=X, V=V, hold=hold, loss=loss, theta=theta, norm=norm, pred=pred)


def test_collected_rows_are_exact_pair_energies(rot3d):
    """A collected row's energy is the vcord term of PotEnergy at that slice."""
    wl = nl_workload("h2o_h2", rot3d, n_h2=4, P=8, Prot=8)
    c = make_ctx(wl)
    X, V = c.surrogate_collect(0)
    assert X.shape == (8 * 4, 4) and V.shape == (32,)
    rp = wl["replicas"][0]
    co = rp["coords"]
    P = rp["P"]
    # the rotor (atom 4) at slice 3 against H2 partner 2: distance feature
    d = np.linalg.norm(co[:, 4 * P + 3] - co[:, 2 * P + 3])
    assert abs(X[3 * 4 + 2, 0] - d) < 1e-5 * d
    assert np.all(np.abs(np.linalg.norm(X[:, 1:3], axis=1)) <= 1.0 + 1e-6)
    # sum of a slice's rows = PotEnergy of the rotor at that slice minus nothing else: the rotor
    # has only H2 partners, so pot_energy_slice of the rotor at its own position is the sum
    e = c.pot_energy_slice(np.zeros(1, np.int32), [4], [3], co[:, 4 * P + 3][None, :])
    assert abs(e[0] - V[12:16].sum()) < 1e-4 * (abs(e[0]) + 1.0)


def test_training_fits_the_run_rows(trained):
    t = trained
    assert t["loss"][-1] < 0.2 * t["loss"][0]
    err = t["pred"] - t["V"][t["hold"]]
    assert np.sqrt(np.mean(err ** 2)) < 0.1 * np.std(t["V"]), (np.sqrt(np.mean(err ** 2)), np.std(t["V"]))


def test_da_requires_a_surrogate(rot3d):
    wl = nl_workload("h2o_h2", rot3d, n_h2=4, P=8, Prot=8)
    c = make_ctx(wl, da=True)
    with pytest.raises(ph.PimcError, match="surrogate"):
        c.run(1)


def test_collected_rotor_rows_are_tip4p_energies(rot3d):
    """Two rotors and no atoms: the rows are rotor-rotor pairs, (R, zq.u, zp.u, zq.zp) and the
    TIP4P energy of caleng, in the order (rotor q, slice k, other rotor)."""
    import oracle as O
    wl = nl_workload("water2", rot3d, P=8, Prot=8)
    c = make_ctx(wl)
    X, V = c.surrogate_collect(0)
    assert X.shape == (2 * 8, 4)
    rp = wl["replicas"][0]
    co, an = rp["coords"], rp["angles"]
    P = rp["P"]
    L = O.lib()
    for q in range(2):
        p = 1 - q
        for k in range(P):
            fr = []
            for a in (q, p):
                eul = np.array([an[0, a * P + k], np.arccos(an[1, a * P + k]), an[2, a * P + k]])
                m = np.zeros(9)
                L.or_matpre(O.P(eul), m.ctypes.data_as(O._dp))
                fr.append((eul, m))
            want = M.rotor_pair_features(fr[0][1], fr[1][1], co[:, q * P + k], co[:, p * P + k])
            row = q * P + k
            assert np.allclose(X[row], want, rtol=1e-5, atol=1e-6), (row, X[row], want)
            e = np.zeros(1)
            L.or_caleng(O.P(co[:, q * P + k]), O.P(co[:, p * P + k]), e.ctypes.data_as(O._dp), O.P(fr[0][0]),
                        O.P(fr[1][0]))
            assert abs(V[row] - e[0]) <= 1e-5 * max(1.0, abs(e[0])), (row, V[row], e[0])


def chain_means(c, n_rep, burn, passes):
    c.run(burn)
    c.clear_samples()
    c.run(passes, sample_every=1)
    m = []
    for r in range(n_rep):
        s = c.samples(r)
        m.append([s[:, 0].mean(), s[:, 1].mean(), s[:, 2].mean(), (s[:, 0] + s[:, 1] + s[:, 2]).mean()])
    return np.array(m)


def replicated_water(rot3d, n_rep, P=8):
    base = nl_workload("water2", rot3d, P=P, Prot=P)
    return dict(tables=base["tables"], replicas=[dict(base["replicas"][0]) for _ in range(n_rep)])


@pytest.fixture(scope="module")
def trained_water(rot3d):
    """A rotor-rotor network trained on rows of a plain run of 64 two-rotor chains."""
    wl = replicated_water(rot3d, 64)
    c = make_ctx(wl)
    xs, vs = [], []
    c.run(20)
    for _ in range(40):
        c.run(5)
        for r in range(len(wl["replicas"])):
            x, v = c.surrogate_collect(r)
            xs.append(x)
            vs.append(v)
    X, V = np.concatenate(xs), np.concatenate(vs)
    hold = np.arange(len(V)) % 10 == 0
    loss = c.surrogate_train(X[~hold], V[~hold], epochs=60, batch=512, lr=3e-3, seed=12)
    theta, norm = c.surrogate_get()
    pred = c.surrogate_eval(X[hold])
    return dict(X=X, V=V, hold=hold, loss=loss, theta=theta, norm=norm, pred=pred)


def test_rotor_network_fits_the_run_rows(trained_water):
    t = trained_water
    assert t["loss"][-1] < 0.3 * t["loss"][0]
    err = t["pred"] - t["V"][t["hold"]]
    rms = np.sqrt(np.mean(err ** 2))
    print("rotor-rotor network: hold-out rms", rms, "spread", np.std(t["V"]))
    assert rms < 0.5 * np.std(t["V"])


def assert_same_ensemble(plain, da, n_rep):
    for j, name in enumerate(("K", "V", "E_rot", "E")):
        d = da[:, j].mean() - plain[:, j].mean()
        se = np.hypot(da[:, j].std(ddof=1), plain[:, j].std(ddof=1)) / np.sqrt(n_rep)
        assert abs(d) <= 2.0 * se, (name, da[:, j].mean(), plain[:, j].mean(), se)


@pytest.mark.parametrize("P", [8, 64])
def test_da_two_rotors_keeps_the_ensemble(rot3d, trained_water, P):
    """Rotor-rotor pairs through the network as stage 1, caleng as stage 2: per-chain means of
    K, V, E_rot and E over 64 chains agree with plain moves within two standard errors, with
    whole chains (P = 8) and with chains split into parts (P = 64).  Both methods continue the
    same equilibrated chains (a plain run from the cold start): the four features of a rotor
    pair leave the twist of the hydrogens about the dipole axes out, so the network misjudges
    some downhill moves of the cold start and the delayed-acceptance move relaxes from it more slowly
    than the plain move, although both keep the same equilibrium."""
    n_rep, (equil, burn, passes) = 64, ((300, 20, 240) if P == 8 else (600, 20, 300))
    wl = replicated_water(rot3d, n_rep, P)
    c0 = make_ctx(wl)
    c0.run(equil)
    states = [c0.get_state(r)[:2] for r in range(n_rep)]
    c0.close()

    def continued(da, seed):
        c = ph.Context(n_rep, seed=seed)
        if da:
            c.set_option("delayed_acceptance", 1)
        pc.load_into_context(c, wl)
        for r, (co, an) in enumerate(states):
            c.set_state(r, co, an)
        if da:
            c.surrogate_set(trained_water["theta"], trained_water["norm"])
        return c

    plain = chain_means(continued(False, SEED + 21), n_rep, burn, passes)
    c = continued(True, SEED + 23)
    da = chain_means(c, n_rep, burn, passes)
    assert_same_ensemble(plain, da, n_rep)
    cnt = c.da_counters(0)
    assert cnt[0] == (burn + passes) * P * P * 2 and cnt[2] <= cnt[1] <= cnt[0]


def test_da_fp64_keeps_the_ensemble(rot3d, trained):
    """The FP64 build of the delayed-acceptance sweep (pair terms and table look-ups in FP64, the
    network unchanged) against FP64 plain moves, rotor-atom network, 64 chains."""
    n_rep, burn, passes = 64, 100, 300

    def ctx(da):
        c = ph.Context(n_rep, flags=ph.FLAG_FP64, seed=SEED)
        if da:
            c.set_option("delayed_acceptance", 1)
        pc.load_into_context(c, replicated(rot3d, n_rep))
        if da:
            c.surrogate_set(trained["theta"], trained["norm"])
        return c

    plain = chain_means(ctx(False), n_rep, burn, passes)
    c = ctx(True)
    da = chain_means(c, n_rep, burn, passes)
    assert_same_ensemble(plain, da, n_rep)
    cnt = c.da_counters(0)
    assert cnt[0] == (burn + passes) * 8 * 8 and cnt[2] <= cnt[1] <= cnt[0] and cnt[3] > 0


def test_da_split_chain_keeps_the_ensemble(rot3d, trained):
    """Rotor-atom network, chains of P = 64 split into parts that trade boundary records."""
    n_rep, burn, passes = 64, 40, 80
    plain = chain_means(make_ctx(replicated(rot3d, n_rep, P=64)), n_rep, burn, passes)
    c = make_ctx(replicated(rot3d, n_rep, P=64), da=True)
    c.surrogate_set(trained["theta"], trained["norm"])
    da = chain_means(c, n_rep, burn, passes)
    assert_same_ensemble(plain, da, n_rep)


@pytest.mark.parametrize("net", ["trained", "random"])
def test_da_keeps_the_ensemble(rot3d, trained, net):
    """Per-chain means of K, V, E_rot and E over 64 chains: delayed acceptance against the
    plain move within two standard errors, with the trained network and with a random one
    (the second stage makes the chain exact for any surrogate)."""
    n_rep, burn, passes = 64, 100, 400
    plain = chain_means(make_ctx(replicated(rot3d, n_rep)), n_rep, burn, passes)
    c = make_ctx(replicated(rot3d, n_rep), da=True)
    if net == "trained":
        c.surrogate_set(trained["theta"], trained["norm"])
    else:
        c.surrogate_set(M.glorot(9), trained["norm"])
    da = chain_means(c, n_rep, burn, passes)
    for j, name in enumerate(("K", "V", "E_rot", "E")):
        d = da[:, j].mean() - plain[:, j].mean()
        se = np.hypot(da[:, j].std(ddof=1), plain[:, j].std(ddof=1)) / np.sqrt(n_rep)
        assert abs(d) <= 2.0 * se, (name, da[:, j].mean(), plain[:, j].mean(), se)
    cnt = c.da_counters(0)
    assert cnt[0] == (burn + passes) * 8 * 8 and cnt[2] <= cnt[1] <= cnt[0]


def test_trained_network_saves_exact_evaluations(rot3d, trained):
    """Exact PotRotE3D sums per accepted move: a trained network lets through few proposals
    that the exact stage then rejects; with an untrained (constant) network the move is the
    density factor followed by the exact one, and every proposal that passes the density is
    charged the exact sums.  The network's factor can only lower the acceptance (a product of
    three Metropolis factors accepts at most as often as the product of two), and it lowers it
    little when the network is accurate."""
    out = {}
    for net in ("trained", "constant"):
        c = make_ctx(replicated(rot3d, 32), da=True)
        if net == "trained":
            c.surrogate_set(trained["theta"], trained["norm"])
        else:
            c.surrogate_set(np.zeros(M.NW), trained["norm"])
        c.run(200)
        cnt = np.sum([c.da_counters(r) for r in range(32)], axis=0)
        out[net] = (cnt[1] / cnt[2], cnt[2] / cnt[1], cnt[2] / cnt[0])
    print("exact sums per accepted move, stage-2 acceptance, acceptance:", out)
    assert out["trained"][1] > 0.9
    assert out["trained"][0] < 0.85 * out["constant"][0]
    assert out["constant"][2] - 0.03 <= out["trained"][2] <= out["constant"][2] + 0.005
