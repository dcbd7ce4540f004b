"""Statistical parity of the GPU sweep with the CPU oracle: the 16^3 Lebwohl-Lasher thermal run and the GHRL
Freedericksz slab under an electric-field sweep.

Case: A = 1, 16x16x16 bulk, periodic x/y/z, random initial directors from the replica's seed, T = 1.2 -> 0.1 with
dT = -0.02 (56 points), MCT = 500 thermalization and MCS = 200 measurement sweeps per point, 64 replicas
(seeds 1..64). The oracle's run is the committed fixture tests/golden/c1_oracle.npz (tests/golden/
make_c1_fixture.py). The GPU runs the 64 replicas as one batched context, starting from the same initial
directors. (The oracle's seeds 1..16 alone put S at T = 0.44 four standard errors below seeds 17..32 and 33..64,
which agree with each other and with the GPU; 64 replicas per side keep such a fluctuation of one group of seeds
from deciding the test.)

Criterion (S, varS, E, varE at each of the 56 temperatures, 224 comparisons): z = (mean_gpu - mean_cpu) /
sqrt(s_gpu^2 / n_gpu + s_cpu^2 / n_cpu) with s the replica-to-replica standard deviation. With 224 comparisons about
one |z| > 3 is expected by chance, so the test allows at most 4 (p ~ 0.5 % of a false failure) and none beyond 5.
"""
import os

import numpy as np
import pytest

from helpers import z_scores

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "c1_oracle.npz")

pytestmark = pytest.mark.gpu


def run_gpu_c1(mcl, orc, seeds):
    N = 16
    pt, _ = orc.geometry("bulk", N, N, N)
    lats = [mcl.Lattice(N, N, N, orc.ic_random(pt, s), periodic=(1, 1, 1), model=mcl.make_model("ll", A=1.0),
                        T=1.2, ang_var=0.5, seed=s) for s in seeds]
    f32 = lambda x: float(np.float32(x))
    Ti, Tf, dT = f32(1.2), f32(0.1), f32(-0.02)
    nT = orc.sweep_points(Ti, Tf, dT)
    rows = np.zeros((len(seeds), nT, 5))
    with mcl.Engine(lats) as eng:
        for it in range(nT):
            T = float(np.float32(orc.sweep_value(Ti, Tf, dT, it, nT)))
            eng.set_temperature(T)
            eng.sweeps(500)
            eng.reset_moments()
            eng.sweeps(200, 1)
            for r in range(len(seeds)):
                m = eng.moments(r)
                S1, S2, E1, E2 = m["S"] / m["n"], m["S2"] / m["n"], m["E"] / m["n"], m["E2"] / m["n"]
                rows[r, it] = (T, S1, S2 - S1 * S1, E1, E2 - E1 * E1)
    return rows


def test_c1_thermal_parity_64_replicas(mcl, orc):
    ref = np.load(FIXTURE)["rows"]  # (64, 56, 6): T S varS E varE Sr
    seeds = list(range(1, ref.shape[0] + 1))
    gpu = run_gpu_c1(mcl, orc, seeds)
    assert np.allclose(gpu[0, :, 0], ref[0, :, 0], atol=1e-6)
    names = ["S", "varS", "E", "varE"]
    zs = np.stack([z_scores(gpu[:, :, c], ref[:, :, c]) for c in (1, 2, 3, 4)])  # (4, 56)
    worst = np.unravel_index(np.argmax(np.abs(zs)), zs.shape)
    report = "; ".join("%s max|z| %.2f at T=%.2f" % (names[c], np.abs(zs[c]).max(),
                                                     ref[0, int(np.argmax(np.abs(zs[c]))), 0]) for c in range(4))
    print(report)
    n3 = int((np.abs(zs) > 3).sum())
    assert n3 <= 4, "%d of 224 comparisons beyond 3 sigma: %s" % (n3, report)
    assert np.abs(zs).max() < 5, "worst %s at T=%.2f: z=%.2f (%s)" % (names[worst[0]], ref[0, worst[1], 0],
                                                                      zs[worst], report)
    # the sampled physics: ordered at low T, disordered at high T
    assert gpu[:, -1, 1].mean() > 0.95 and gpu[:, 0, 1].mean() < 0.15


def test_c3_electric_slab_parity_16_replicas(mcl, orc):
    """GHRL (K11, K22, K33) = (1.0, 0.6, 1.4) 16x16x8 slab, homeotropic W = 1.5 on both surfaces, T = 0.5, field
    along x with elecA = 0.25 swept 0 -> 2.0 in steps of 0.2, MCT = 300, MCS = 200, 16 replicas (the oracle's run is
    tests/golden/c3_oracle.npz, tests/golden/make_c3_fixture.py). Same z criterion as the thermal case over
    4 x 11 comparisons: at most 2 beyond 3 sigma (0.1 expected by chance), none beyond 5."""
    import importlib.util
    from helpers import lattice_from_orc
    spec = importlib.util.spec_from_file_location("make_c3_fixture", os.path.join(HERE, "golden", "make_c3_fixture.py"))
    fx = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fx)
    c = fx.CASE
    ref = np.load(os.path.join(HERE, "golden", "c3_oracle.npz"))["rows"]  # (16, 11, 6): field S varS E varE -
    seeds = list(range(1, ref.shape[0] + 1))
    pt, ns = orc.geometry("slab", c["Nx"], c["Ny"], c["Nz"])
    p = fx.case_params(orc)
    lats = [lattice_from_orc(mcl, p, orc.ic_random(pt, s), pt, ns, T=c["T"], seed=s) for s in seeds]
    f32 = lambda x: float(np.float32(x))
    Ei, Ef, dE = f32(c["Ei"]), f32(c["Ef"]), f32(c["dE"])
    nE = orc.sweep_points(Ei, Ef, dE)
    gpu = np.zeros((len(seeds), nE, 5))
    with mcl.Engine(lats) as eng:
        for ie in range(nE):
            E = float(np.float32(orc.sweep_value(Ei, Ef, dE, ie, nE)))
            eng.set_field(E)
            eng.sweeps(c["MCT"])
            eng.reset_moments()
            eng.sweeps(c["MCS"], 1)
            for r in range(len(seeds)):
                m = eng.moments(r)
                S1, S2, E1, E2 = m["S"] / m["n"], m["S2"] / m["n"], m["E"] / m["n"], m["E2"] / m["n"]
                gpu[r, ie] = (E, S1, S2 - S1 * S1, E1, E2 - E1 * E1)
    assert np.allclose(gpu[0, :, 0], ref[0, :, 0], atol=1e-6)
    names = ["S", "varS", "E", "varE"]
    zs = np.stack([z_scores(gpu[:, :, k], ref[:, :, k]) for k in (1, 2, 3, 4)])
    report = "; ".join("%s max|z| %.2f at field %.1f" % (names[k], np.abs(zs[k]).max(),
                                                          ref[0, int(np.argmax(np.abs(zs[k]))), 0]) for k in range(4))
    print(report)
    assert int((np.abs(zs) > 3).sum()) <= 2, report
    assert np.abs(zs).max() < 5, report
