"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py): its backpropagation against central finite differences of the
loss, and the closed forms of a network with zero hidden weights."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def mo():
    import mlp_oracle
    return mlp_oracle


def rand_params(mo, seed, scale=0.3):
    return np.random.default_rng(seed).normal(scale=scale, size=mo.NP)


def test_gradient_matches_finite_differences(mo):
    rng = np.random.default_rng(1)
    p = rand_params(mo, 2)
    x = rng.normal(size=(20, 5))
    t = rng.normal(size=(20, 3))
    w = np.array([1.0, 0.5, 2.0])
    loss, g = mo.loss_grad(p, x, t, w)
    h = 1e-6
    for k in rng.choice(mo.NP, size=120, replace=False):
        e = np.zeros(mo.NP)
        e[k] = h
        fd = (mo.loss_grad(p + e, x, t, w)[0] - mo.loss_grad(p - e, x, t, w)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_zero_hidden_weights_closed_form(mo):
    """With W3 = 0 the output is b3 for every input, and the loss gradient of b3 is 2 mean(b3 - t) w."""
    p = rand_params(mo, 3)
    d = mo.unpack(p)
    d["W3"] = np.zeros_like(d["W3"])
    d["b3"] = np.array([0.1, -0.2, 0.3])
    p = mo.pack(d)
    x = np.random.default_rng(4).normal(size=(7, 5))
    t = np.random.default_rng(5).normal(size=(7, 3))
    assert np.allclose(mo.forward(p, x), d["b3"])
    _, g = mo.loss_grad(p, x, t)
    gb3 = mo.unpack(g)["b3"]
    assert np.allclose(gb3, 2 * (d["b3"] - t).mean(0))


def test_silu(mo):
    z = np.array([-3.0, 0.0, 2.0])
    assert np.allclose(mo.silu(z), z / (1 + np.exp(-z)))
    assert mo.silu(np.array([0.0]))[0] == 0.0


def test_adam_first_step_is_sign_times_lr(mo):
    """The bias-corrected first Adam step moves every parameter by lr * g / (|g| + eps) ~ lr * sign(g)."""
    p = np.zeros(4)
    g = np.array([0.5, -2.0, 1e-3, 3.0])
    p1, m, v = mo.adam(p, np.zeros(4), np.zeros(4), g, 1, 0.01)
    assert np.allclose(p1, -0.01 * np.sign(g), rtol=1e-4)
