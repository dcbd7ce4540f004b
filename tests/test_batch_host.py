"""Host side of the batched driver (no GPU): geometry, specs and the replica combination."""
import numpy as np
import pytest

from mclics_amd import batch


@pytest.mark.parametrize("n", [8, 11, 16, 21])
def test_sphere_geometry_matches_oracle(orc, n):
    pt, ns = batch.sphere_geometry(n)
    opt, ons = orc.geometry("sphere", n, n, n)
    assert np.array_equal(pt, opt)
    assert np.allclose(ns, ons, atol=1e-6)


@pytest.mark.parametrize("dims", [(4, 4, 3), (6, 4, 8), (16, 16, 8)])
def test_slab_geometry_matches_oracle(orc, dims):
    pt, ns = batch.slab_geometry(*dims)
    opt, ons = orc.geometry("slab", *dims)
    assert np.array_equal(pt, opt)
    assert np.array_equal(ns, ons)


def test_c4_specs():
    specs = batch.c4_specs(512, 1)
    L = np.array([s["L"] for s in specs])
    T = np.array([s["T"] for s in specs])
    assert len(specs) == 512 and set(L % 2) == {0} and L.min() >= 16 and L.max() <= 64
    assert T.min() >= 0.9 and T.max() <= 1.3
    nsph = sum(s["geometry"] == "sphere" for s in specs)
    assert 200 < nsph < 312
    assert len({s["seed"] for s in specs}) == 512
    assert specs == batch.c4_specs(512, 1)


def test_combine_replicas_law_of_total_variance():
    """tools/average_replicas.py:86-97: Var = mean(var_r) + var(mean_r) with ddof = 0, checked against pooling the
    raw samples of equal-length runs (the identity the law states)."""
    rng = np.random.default_rng(0)
    samples = [rng.normal(loc=mu, scale=sd, size=400) for mu, sd in ((0.5, 0.1), (0.55, 0.2), (0.45, 0.05))]
    S = np.array([[x.mean()] for x in samples])
    varS = np.array([[x.var()] for x in samples])
    m, v, _, _ = batch.combine_replicas(S, varS, S, varS)
    pooled = np.concatenate(samples)
    assert m[0] == pytest.approx(pooled.mean(), rel=1e-12)
    assert v[0] == pytest.approx(pooled.var(), rel=1e-12)


def test_random_directors_are_unit_and_isotropic():
    pt = np.ones(20000, dtype=np.int32)
    pt[:100] = 0
    ni = batch.random_directors(pt, 3).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(ni, axis=1), 1, atol=1e-6)
    assert np.all(ni[:100] == [0, 0, 1])
    assert abs(ni[100:, 2].mean()) < 0.03 and abs((ni[100:, 2] ** 2).mean() - 1 / 3) < 0.02


@pytest.mark.parametrize("k", [(1.0, 0.6, 1.4, 0.0), (1.48, 1.0, 2.28, 0.0), (1.0, 1.0, 1.0, 10.0), (0.9, 1.3, 0.7, 8.0)])
def test_ghrl_coefficients_match_oracle(orc, k):
    """The product's setGHRL restatement (io.cpp:340-360) equals the oracle's bit for bit in FP32."""
    from mclics_amd import capi
    got = capi.ghrl_coefficients(*k)
    ref = orc.ghrl_coeffs(*k)
    for name in ("lam", "mu", "nu", "rho", "sigma", "scale"):
        assert np.float32(got[name]) == np.float32(ref[name]), name
