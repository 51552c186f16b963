"""Oracle (test infrastructure only) for the surrogate network of the delayed-acceptance rotor
move: an FP64 numpy statement of the network, its loss and gradient, the Adam step and the
pair features.  The reference has no surrogate; this states the design the GPU kernels
(moribs-pimc_amd/csrc/pimc_mlp.hpp) implement, so that the tests can check them.

Network: 4 -> 128 -> 128 -> 128 -> 1, SiLU after each hidden layer; flat weight vector theta
in the order W1[128][4] b1 W2[128][128] b2 W3[128][128] b3 w4[128] b4.  Loss (1/2n) sum (f - y)^2.
Pinned by tests/test_oracle_golden.py: central finite differences of the loss, the
constant-network closed form, and SiLU's derivative.
"""
import numpy as np

NIN, NH = 4, 128
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
         ("w4", (NH,)), ("b4", (1,))]
NW = sum(int(np.prod(s)) for _, s in SIZES)


def offsets():
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = (o, shape)
        o += n
    return out


OFF = offsets()


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    return {k: theta[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_prime(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, X, rounding=None):
    """Outputs [n] of the raw network.  rounding="bf16" rounds the operands of the three
    matrix products (weights and incoming activations) to bfloat16, as the matrix-core
    evaluation does; biases and the output row stay in full precision."""
    p = unpack(theta)
    r = bf16 if rounding == "bf16" else (lambda a: np.asarray(a, np.float64))
    h = np.asarray(X, np.float64)
    for W, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        h = silu(r(h) @ r(p[W]).T + p[b])
    return h @ p["w4"] + p["b4"][0]


def loss_grad(theta, X, y):
    """Loss (1/2n) sum (f(x) - y)^2 and its gradient in theta's layout (FP64)."""
    p = unpack(theta)
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    n = len(y)
    zs, hs = [], [X]
    for W, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        z = hs[-1] @ p[W].T + p[b]
        zs.append(z)
        hs.append(silu(z))
    out = hs[-1] @ p["w4"] + p["b4"][0]
    d = out - y
    loss = 0.5 * np.dot(d, d) / n
    dout = d / n
    g = {"w4": hs[3].T @ dout, "b4": np.array([dout.sum()])}
    dh = np.outer(dout, p["w4"])
    for li, (W, b) in reversed(list(enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3"))))):
        dz = dh * silu_prime(zs[li])
        g[W] = dz.T @ hs[li]
        g[b] = dz.sum(axis=0)
        dh = dz @ p[W]
    flat = np.zeros(NW)
    for k, (o, s) in OFF.items():
        flat[o:o + int(np.prod(s))] = np.asarray(g[k]).reshape(-1)
    return loss, flat


def adam_step(theta, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with bias correction; returns theta, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    return theta - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps), m, v


def pair_features(frame_m, zn, centre, partner):
    """(R, ux, uy, uz): distance and direction cosines of the partner on the rotor's body
    axes, the rows of the rotation matrix as vcord reads them (vcord.f:50-85; z axis
    normalised by its length zn)."""
    m = np.asarray(frame_m, np.float64).reshape(3, 3)
    r = np.asarray(partner, np.float64) - np.asarray(centre, np.float64)
    rn = np.linalg.norm(r)
    u = m @ r / rn
    return np.array([rn, u[0], u[1], u[2] / zn])


def rotor_pair_features(frame_q, frame_p, centre_q, centre_p):
    """(R, zq.u, zp.u, zq.zp) of a rotor-rotor pair: the distance of the centres, the cosines of
    each rotor's body z axis (the third row of the frame as stored, the third column of the
    rotation matrix) with the unit separation u from rotor q to rotor p, and the cosine between
    the two z axes."""
    zq = np.asarray(frame_q, np.float64).reshape(3, 3)[2]
    zp = np.asarray(frame_p, np.float64).reshape(3, 3)[2]
    r = np.asarray(centre_p, np.float64) - np.asarray(centre_q, np.float64)
    rn = np.linalg.norm(r)
    u = r / rn
    return np.array([rn, zq @ u, zp @ u, zq @ zp])


def glorot(seed=0):
    """A random weight vector (Glorot-uniform matrices, zero biases) for tests."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(NW)
    for k, (o, s) in OFF.items():
        if k[0] in "Ww":
            fan_out, fan_in = (s[0], s[1]) if len(s) == 2 else (1, s[0])
            a = np.sqrt(6.0 / (fan_in + fan_out))
            theta[o:o + int(np.prod(s))] = rng.uniform(-a, a, int(np.prod(s)))
    return theta
