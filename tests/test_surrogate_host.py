"""Host side of the surrogate and of the T_NI estimators (no GPU): C_V-peak vertex, coexistence sign change,
temperature grids, Maier-Saupe warm starts, placement and the training distribution."""
import math

import numpy as np
import pytest

from mclics_amd import surrogate, tni


def test_cv_peak_recovers_a_parabola_vertex():
    """Three samples of a parabola fix its vertex exactly (make_figures.py:444-451); C_V = L^3 varE / (4 T^2)."""
    L = 10
    T = tni.uniform_grid(0.004)
    assert len(T) == 56 and T[0] == pytest.approx(1.24) and T[-1] == pytest.approx(1.02)
    t0 = 1.12345
    cv = 50.0 - 3e4 * (T - t0) ** 2
    varE = cv * 4 * T * T / L ** 3
    got_cv, est = tni.cv_peak(T, varE, L)
    assert est == pytest.approx(t0, abs=1e-9)
    assert np.allclose(got_cv, cv[np.argsort(T)])
    # maximum at the end of the grid: the grid point itself
    _, end = tni.cv_peak(T, np.exp(-T), L)
    assert end == pytest.approx(T.min())


def test_cv_peak_is_within_half_a_step_of_the_largest_point():
    rng = np.random.default_rng(3)
    T = tni.uniform_grid(0.004)
    for _ in range(200):
        v = rng.random(len(T))
        cv, est = tni.cv_peak(T, v, 16)
        Ts = np.sort(T)
        i = int(np.argmax(cv))
        assert abs(est - Ts[i]) <= 0.002 + 1e-12


def test_coexistence_point_step_with_noise():
    rng = np.random.default_rng(5)
    T = np.round(np.arange(1.107, 1.1385, 0.001), 4)
    t0 = 1.1226
    drift = 0.02 * np.tanh((t0 - T) / 0.002) + rng.normal(scale=0.002, size=len(T))
    est, ok, side = tni.coexistence_point(T, drift)
    assert ok and side == 0 and abs(est - t0) < 6e-4
    # order of the points does not matter
    p = rng.permutation(len(T))
    assert tni.coexistence_point(T[p], drift[p])[0] == pytest.approx(est)
    # an outlier on the wrong side does not move the split
    d2 = drift.copy()
    d2[2] = -0.01
    assert abs(tni.coexistence_point(T, d2)[0] - t0) < 1e-3


def test_coexistence_point_outside_the_window():
    T = np.linspace(1.0, 1.05, 20)
    assert tni.coexistence_point(T, np.full(20, 0.01)) == (pytest.approx(1.05), False, 1)
    assert tni.coexistence_point(T, np.full(20, -0.01)) == (pytest.approx(1.0), False, -1)
    # one noisy point of the other sign at an edge is no bracket: the window moves towards the majority
    d = np.full(20, 0.01)
    d[-1] = -0.002
    assert tni.coexistence_point(T, d) == (pytest.approx(1.05), False, 1)
    d = np.full(20, -0.01)
    d[0] = 0.003
    assert tni.coexistence_point(T, d) == (pytest.approx(1.0), False, -1)
    # two points beyond the split are a bracket
    d = np.full(20, 0.01)
    d[-2:] = -0.01
    est, ok, side = tni.coexistence_point(T, d)
    assert ok and side == 0 and T[-3] <= est <= T[-2]


@pytest.mark.parametrize("S", [0.0, 0.15, 0.33, 0.6, 0.9])
def test_warm_start_has_the_requested_order(S):
    pt = np.ones(200000, dtype=np.int32)
    v = surrogate.warm_start(pt, S, 7).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 1, atol=1e-6)
    p2 = (1.5 * v[:, 2] ** 2 - 0.5).mean()
    assert p2 == pytest.approx(S, abs=0.005)
    # uniform azimuth: no order in the xy plane
    assert abs((v[:, 0] ** 2 - v[:, 1] ** 2).mean()) < 0.01


def test_warm_start_axis_and_empty_sites():
    pt = np.ones(100000, dtype=np.int32)
    pt[:10] = 0
    v = surrogate.warm_start(pt, 0.5, 2, axis=(1.0, 1.0, 0.0)).reshape(-1, 3)
    a = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    assert (1.5 * (v[10:] @ a) ** 2 - 0.5).mean() == pytest.approx(0.5, abs=0.01)
    assert np.all(v[:10] == [0, 0, 1])


def test_maier_saupe_table_is_monotone():
    assert np.all(np.diff(surrogate._P2) > 0)
    assert surrogate._P2[0] < 0 < surrogate._P2[-1] and surrogate._P2[-1] > 0.97


class _Bump:
    """Stand-in network: log chi peaked at t0 for every size."""

    def __init__(self, t0):
        self.t0 = t0

    def predict(self, X):
        X = np.asarray(X)
        T = X[:, 0]
        return np.stack([np.clip(1.2 - T, 0, 1), -T, -1e3 * (T - self.t0) ** 2], axis=1)


def test_placement_centres_on_the_predicted_peak():
    temps, t_peak = surrogate.place_points(_Bump(1.127), 96, n=32, spacing=1e-3)
    assert t_peak == pytest.approx(1.127, abs=2e-4)
    assert len(temps) == 32 and np.allclose(np.diff(temps), 1e-3)
    assert temps.mean() == pytest.approx(t_peak)
    temps, s_nem, s_iso, tp = tni.guided_window(_Bump(1.127), 96)
    assert np.all(s_nem >= s_iso) and np.all(s_nem > 0.08) and np.all(s_iso < 0.07)


def test_features_layout():
    X = surrogate.features(np.array([1.0, 1.1]), 0.9, 1.5, 0.2, 1000.0)
    assert X.shape == (2, 5)
    assert np.allclose(X[1], [1.1, 0.9, 1.5, 0.2, math.log(1000.0)])


def test_training_distribution():
    rng = np.random.default_rng(1)
    specs = surrogate.random_lattices(2000, rng)
    L = np.array([s["L"] for s in specs])
    A = np.array([s["A"] for s in specs])
    W = np.array([s["W"] for s in specs])
    f = np.array([s["field"] for s in specs])
    sph = np.array([s["geometry"] == "sphere" for s in specs])
    heat = np.array([s["heating"] for s in specs])
    # small lattices by default (the network places points and warm-starts; sizes beyond are asked at the largest)
    assert set(L % 2) == {0} and L.min() == 12 and L.max() == 20
    L2 = np.array([s["L"] for s in surrogate.random_lattices(500, rng, lmin=16, lmax=64)])
    assert set(L2 % 2) == {0} and L2.min() == 16 and L2.max() == 64
    assert A.min() >= 0.8 and A.max() <= 1.2
    assert np.all(W[~sph] == 0) and W[sph].min() >= 0.5 and W[sph].max() <= 1.5
    assert 0.4 < (f == 0).mean() < 0.6 and f.max() <= 0.3
    assert 0.4 < sph.mean() < 0.6 and 0.4 < heat.mean() < 0.6


def test_slab_directors_layout():
    L = 6
    lo = np.tile(np.array([0, 0, 1], np.float32), L ** 3 // 2)
    hi = np.tile(np.array([1, 0, 0], np.float32), L ** 3 // 2)
    v = tni.slab_directors(L, lo, hi).reshape(L, L, L, 3)  # (k, j, i, 3): i fastest
    assert np.all(v[: L // 2, ..., 2] == 1) and np.all(v[L // 2:, ..., 0] == 1)
