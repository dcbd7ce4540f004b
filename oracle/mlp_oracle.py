"""FP64 CPU oracle of the surrogate MLP (5 -> 64 SiLU -> 64 SiLU -> 3) -- TEST INFRASTRUCTURE ONLY.

The checker of the HIP forward, gradient and Adam kernels of libmclgpu (mcl_mlp_*). The network has no reference
counterpart; its backpropagation is pinned against central finite differences of the loss in
tests/test_oracle_mlp.py. Parameter layout as in libmclgpu: W1[64][5], b1[64], W2[64][64], b2[64], W3[3][64], b3[3].
"""
import numpy as np

NI, NH, NO = 5, 64, 3
SHAPES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NO, NH)), ("b3", (NO,))]
NP = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, off = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[off:off + n].reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, x, keep=False):
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64).reshape(-1, NI)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = silu(z1)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = silu(z2)
    y = h2 @ w["W3"].T + w["b3"]
    if keep:
        return y, (x, z1, h1, z2, h2)
    return y


def loss_grad(p, x, t, w_out=None):
    """L = (1/n) sum_s sum_o w_o (y_so - t_so)^2 and dL/dp."""
    wo = np.ones(NO) if w_out is None else np.asarray(w_out, dtype=np.float64)
    y, (x, z1, h1, z2, h2) = forward(p, x, keep=True)
    t = np.asarray(t, dtype=np.float64).reshape(-1, NO)
    n = len(x)
    r = y - t
    loss = float((wo * r * r).sum() / n)
    w = unpack(p)
    dy = 2.0 * wo * r / n
    sg2 = 1.0 / (1.0 + np.exp(-z2))
    dz2 = (dy @ w["W3"]) * sg2 * (1.0 + z2 * (1.0 - sg2))
    sg1 = 1.0 / (1.0 + np.exp(-z1))
    dz1 = (dz2 @ w["W2"]) * sg1 * (1.0 + z1 * (1.0 - sg1))
    g = {"W3": dy.T @ h2, "b3": dy.sum(0), "W2": dz2.T @ h1, "b2": dz2.sum(0), "W1": dz1.T @ x, "b1": dz1.sum(0)}
    return loss, pack(g)


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, 2015) with bias correction; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
