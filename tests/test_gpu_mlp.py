"""The surrogate MLP's HIP kernels against the FP64 oracle (oracle/mlp_oracle.py), through the C ABI (needs an
MI355X). Tolerances are FP32 accumulation levels: forward 1e-5 relative, gradients 1e-4 of the largest component."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mo():
    import mlp_oracle
    return mlp_oracle


def rand(mo, seed, n=300):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.25, size=mo.NP).astype(np.float32)
    x = rng.normal(size=(n, 5)).astype(np.float32)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    return p, x, t


def test_forward_matches_oracle(mcl, mo):
    p, x, _ = rand(mo, 1, 1000)
    net = mcl.Mlp(seed=3)
    net.set_params(p)
    y = net.forward(x)
    ref = mo.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.allclose(y, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    assert np.array_equal(net.params(), p)


@pytest.mark.parametrize("n", [1, 31, 32, 77, 1000])
def test_gradient_matches_oracle(mcl, mo, n):
    p, x, t = rand(mo, 2, n)
    w = np.array([1.0, 0.3, 2.5], dtype=np.float32)
    net = mcl.Mlp()
    net.set_params(p)
    g, loss = net.gradient(x, t, w)
    rl, rg = mo.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), w.astype(np.float64))
    assert loss == pytest.approx(rl, rel=1e-5)
    assert np.abs(g - rg).max() <= 1e-4 * np.abs(rg).max()


def test_initial_parameters(mcl, mo):
    """He-uniform weights within sqrt(6 / fan_in), zero biases; a different seed gives different weights."""
    a, b = mcl.Mlp(seed=5).params(), mcl.Mlp(seed=6).params()
    d = mo.unpack(a)
    assert np.abs(d["W1"]).max() <= np.sqrt(6 / 5) and np.abs(d["W2"]).max() <= np.sqrt(6 / 64)
    assert np.all(d["b1"] == 0) and np.all(d["b2"] == 0) and np.all(d["b3"] == 0)
    assert not np.array_equal(a, b)


def test_adam_steps_match_oracle(mcl, mo):
    """Three full-batch Adam steps (one epoch each, batch = n) equal the oracle's Adam on its FP64 gradients."""
    p, x, t = rand(mo, 7, 256)
    net = mcl.Mlp()
    net.set_params(p)
    lr = 1e-3
    ref = p.astype(np.float64)
    m = np.zeros(mo.NP)
    v = np.zeros(mo.NP)
    for step in (1, 2, 3):
        net.train(x, t, epochs=1, batch=256, lr=lr)
        _, g = mo.loss_grad(ref, x.astype(np.float64), t.astype(np.float64))
        ref, m, v = mo.adam(ref, m, v, g, step, lr)
    got = net.params()
    # each Adam step moves a parameter by at most ~lr; FP32 gradients flip the direction only where g ~ 0
    assert np.abs(got - ref).max() < 3 * lr
    assert np.median(np.abs(got - ref)) < 1e-6


def test_training_fits_a_smooth_function(mcl, mo):
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, size=(4096, 5)).astype(np.float32)
    t = np.stack([np.tanh(2 * x[:, 0] - x[:, 1]), x[:, 2] * x[:, 3], np.exp(-x[:, 4] ** 2)], axis=1).astype(np.float32)
    net = mcl.Mlp(seed=2)
    l0 = net.gradient(x, t)[1]
    loss = net.train(x, t, epochs=150, batch=256, lr=3e-3, seed=4)
    assert loss < 0.01 * l0, (l0, loss)
    y = net.forward(x)
    assert np.abs(y - t).mean() < 0.05
