"""What the trained surrogate gains over the method without it, on small lattices (needs an MI355X).

A network is trained with the HIP kernels on pairs from batched plain ladders of 8^3..16^3 lattices, then
  - placement: its predicted susceptibility peak for a 12^3 bulk lattice lies near the peak of a dense plain scan,
    where an untrained network (the same architecture and normalization, initial weights) puts it far away;
  - warm start: at T = 1.0 a 32^3 lattice started from the network's predicted order has an equilibrium energy
    (within 4 standard deviations of the equilibrium fluctuation) after 100 sweeps, where the reference's random
    start is still more than 10 standard deviations away.
Every number compared here is a plain Metropolis average; the network only chooses temperatures and starts.
"""
import numpy as np
import pytest

from mclics_amd import batch, capi, surrogate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(mcl):
    X, Y, _ = surrogate.generate(runs=2, lattices=128, rungs=16, mct=100, mcs=300, every=2, seed=5, lmin=8,
                                 lmax=16, log=None)
    sur, info = surrogate.Surrogate.fit(X, Y, epochs=200, batch_size=256, lr=3e-3, seed=2)
    return sur, info, X, Y


def dense_chi_peak(L, temps, seeds=4, sweeps=3000):
    """Plain runs at every temperature (independent lattices, random starts), chi = N varS / T averaged over
    seeds; the parabola vertex through the three largest values."""
    specs = [dict(L=L, geometry="bulk", T=float(np.float32(T)), seed=1000 * s + k)
             for s in range(seeds) for k, T in enumerate(temps)]
    rows, _ = batch.run_batch(specs, sweeps // 3, sweeps - sweeps // 3, every=2, autotune=False, graph=50)
    chi = rows[:, 5].reshape(seeds, len(temps)).mean(0)
    i = int(np.argmax(chi))
    i = min(max(i, 1), len(temps) - 2)
    y0, y1, y2 = chi[i - 1:i + 2]
    h = temps[1] - temps[0]
    return float(temps[i] + 0.5 * h * (y0 - y2) / (y0 - 2 * y1 + y2)), rows


def test_training_reaches_a_useful_fit(trained):
    sur, info, X, Y = trained
    assert len(X) == 2 * 128 * 16
    assert info["val_rmse"]["S"] < 0.06, info


def test_placement_gain(trained):
    sur, _, _, _ = trained
    temps = np.round(np.arange(1.00, 1.301, 0.01), 4)
    t_true, rows = dense_chi_peak(12, temps)
    t_trained = surrogate.predicted_peak(sur, 12)[0]
    untrained = surrogate.Surrogate(sur.xm, sur.xs, sur.ym, sur.ys, seed=12345, xlo=sur.xlo, xhi=sur.xhi)
    t_untrained = surrogate.predicted_peak(untrained, 12)[0]
    print("chi peak: plain scan %.4f, trained %.4f, untrained %.4f" % (t_true, t_trained, t_untrained))
    assert abs(t_trained - t_true) < 0.025
    assert abs(t_untrained - t_true) > 2 * abs(t_trained - t_true)
    # the trained network's S curve follows the plain scan's; the untrained one does not
    S_pred = sur.predict(surrogate.features(temps, 1.0, 0.0, 0.0, 12.0 ** 3))[:, 0]
    S_meas = rows[:, 1].reshape(4, len(temps)).mean(0)
    S_raw = untrained.predict(surrogate.features(temps, 1.0, 0.0, 0.0, 12.0 ** 3))[:, 0]
    assert np.abs(S_pred - S_meas).mean() < 0.05
    assert np.abs(S_raw - S_meas).mean() > 2 * np.abs(S_pred - S_meas).mean()


def test_warm_start_gain(trained):
    sur, _, _, _ = trained
    L, T = 32, 1.0
    S0 = float(sur.predict(surrogate.features(np.array([T]), 1.0, 0.0, 0.0, float(L) ** 3))[0, 0])
    pt = np.ones(L ** 3, dtype=np.int32)
    hom = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), L ** 3)
    ics = [None, None, surrogate.warm_start(pt, S0, 11), surrogate.warm_start(pt, S0, 12), hom]
    lats = [batch.build_lattice(dict(L=L, geometry="bulk", T=T, seed=70 + k), ic=ic) for k, ic in enumerate(ics)]
    with capi.Engine(lats) as eng:
        eng.sweeps(100)
        E100 = [eng.measure(r)["E"] for r in range(len(lats))]
        eng.sweeps(1900)
        eng.reset_moments()
        eng.sweeps(1000, 5)
        m = eng.moments(4)
        E_eq = m["E"] / m["n"]
        sd = (m["E2"] / m["n"] - E_eq * E_eq) ** 0.5
    print("S0 %.3f  E_eq %.4f +- %.4f  E(100): random %s warm %s" % (S0, E_eq, sd, E100[:2], E100[2:4]))
    cold = min(abs(E100[0] - E_eq), abs(E100[1] - E_eq))
    warm = max(abs(E100[2] - E_eq), abs(E100[3] - E_eq))
    assert warm < 4 * sd
    assert cold > 10 * sd
