"""Batching, launch shapes, graph replay and the autotuner (needs an MI355X).

The RNG counter is (site, sweep) and the acceptance count is an exact integer sum, so a lattice's trajectory does
not depend on the launch shape, on graph replay or on the other lattices of its context: all of these must give
bit-identical directors, ang_var and counts. The FP64 moments of a context that measures in the sweep (these do)
come from partial sums that follow the sweep's workgroups, finalized either in the next pass or by a trailing
kernel, so their last bits may depend on the launch shape and on graph replay: they must agree to 1e-12.
"""
import numpy as np
import pytest

from mclics_amd import batch

pytestmark = pytest.mark.gpu

SPECS = [dict(L=24, geometry="bulk", T=1.05, seed=11), dict(L=10, geometry="bulk", T=0.7, seed=12),
         dict(L=14, geometry="sphere", T=0.9, seed=13, W=1.0), dict(L=16, geometry="bulk", T=1.25, seed=14)]


def state(eng, n):
    return [eng.directors(r) for r in range(n)], [eng.moments(r) for r in range(n)], \
           [eng.ang_var(r) for r in range(n)], [eng.sweep_stats(r) for r in range(n)]


def close_moments(ma, mb):
    return all(a["n"] == b["n"] and all(b[k] == pytest.approx(a[k], rel=1e-12, abs=1e-15) for k in a if k != "n")
               for a, b in zip(ma, mb))


def same(a, b):
    dirs_a, mom_a, av_a, st_a = a
    dirs_b, mom_b, av_b, st_b = b
    return all(np.array_equal(x, y) for x, y in zip(dirs_a, dirs_b)) and close_moments(mom_a, mom_b) and \
        av_a == av_b and st_a == st_b


def run(mcl, shape=None, graph=0, calls=((30, 5),)):
    with mcl.Engine([batch.build_lattice(s) for s in SPECS]) as eng:
        if shape:
            eng.set_launch(*shape)
        if graph:
            eng.set_graph(graph)
        out = []
        for n, every in calls:
            eng.sweeps(n, every)
            out.append(state(eng, len(SPECS)))
        return out


@pytest.mark.parametrize("shape", [(64, 0), (128, 1), (512, 0), (1024, 1)])
def test_launch_shape_does_not_change_the_trajectory(mcl, shape):
    ref = run(mcl, (256, 1))
    got = run(mcl, shape)
    assert same(ref[-1], got[-1])


def test_graph_replay_matches_eager(mcl):
    """Several graph replays with eager calls and host synchronizations between them; every call's state must
    equal the eager run's."""
    calls = ((40, 4), (7, 0), (24, 4), (33, 3), (64, 0))
    eager = run(mcl, calls=calls)
    graph = run(mcl, graph=8, calls=calls)
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert same(a, b), "call %d differs" % k


BULK_SPECS = [dict(L=24, geometry="bulk", T=1.05, seed=21), dict(L=16, geometry="bulk", T=0.8, seed=22)]


@pytest.mark.parametrize("specs", [SPECS, BULK_SPECS], ids=["mixed", "bulk-fused"])
def test_short_calls_replay_graphs_like_eager(mcl, specs):
    """With a 50-sweep graph block, calls shorter than the block replay a graph of their own length (the largest
    multiple of measure_every; the rest eagerly), calls longer than it add such a tail after the whole blocks, and
    mcl_prepare_sweeps captures them ahead without running anything. More distinct lengths than the context's 8
    cached graphs evict the least recently used ones. Every call's state equals the eager run's. The bulk-only
    context measures inside the sweep kernel (its finalize then ends each graph)."""
    calls = ((20, 10), (5, 0), (13, 4), (9, 3), (11, 0), (22, 2), (3, 1), (17, 0), (6, 6), (20, 10), (14, 7),
             (1, 0), (120, 10), (20, 10))
    prepared = {0, 6, 12}

    def go(graph):
        with mcl.Engine([batch.build_lattice(s) for s in specs]) as eng:
            if graph:
                eng.set_graph(graph)
            out = []
            for k, (n, every) in enumerate(calls):
                if k in prepared:
                    before = state(eng, len(specs))
                    eng.prepare(n, every)
                    assert same(before, state(eng, len(specs))), "prepare changed the state"
                eng.sweeps(n, every)
                out.append(state(eng, len(specs)))
            return out

    eager, graph = go(0), go(50)
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert same(a, b), "call %d %s differs" % (k, calls[k])
    assert graph[-1][3][0][1] == sum(n for n, _ in calls)


def test_batched_lattice_equals_lattice_alone(mcl):
    with mcl.Engine([batch.build_lattice(s) for s in SPECS]) as eng:
        eng.sweeps(25, 5)
        together = state(eng, len(SPECS))
    for r, s in enumerate(SPECS):
        with mcl.Engine(batch.build_lattice(s)) as eng:
            eng.sweeps(25, 5)
            alone = state(eng, 1)
        assert np.array_equal(alone[0][0], together[0][r])
        # the measurement kernels of the two contexts are different template instances (the batch has a sphere:
        # masked, radial), which may contract the FP32 site energy differently by an ulp (1e-7 relative per site)
        for k, v in alone[1][0].items():
            assert together[1][r][k] == pytest.approx(v, rel=1e-6, abs=1e-9)


def test_autotune_keeps_the_state(mcl):
    with mcl.Engine([batch.build_lattice(s) for s in SPECS]) as eng, \
            mcl.Engine([batch.build_lattice(s) for s in SPECS]) as twin:
        eng.sweeps(12, 3)
        twin.sweeps(12, 3)
        before = state(eng, len(SPECS))
        res = eng.autotune(5)
        assert res["threads"] in (64, 128, 256, 512, 1024) and res["us_per_sweep"] > 0
        assert same(before, state(eng, len(SPECS)))
        again = eng.autotune(5)
        assert again["from_cache"] and again["threads"] == res["threads"]
        eng.sweeps(20, 4)
        twin.sweeps(20, 4)
        assert same(state(eng, len(SPECS)), state(twin, len(SPECS)))
        # trials measured every 2 sweeps (mcl_autotune_ex): the moments they add are restored as well
        before = state(eng, len(SPECS))
        res = eng.autotune(6, 2)
        assert res["threads"] in (64, 128, 256, 512, 1024) and not res["from_cache"]
        assert same(before, state(eng, len(SPECS)))
        eng.sweeps(9, 3)
        twin.sweeps(9, 3)
        assert same(state(eng, len(SPECS)), state(twin, len(SPECS)))


def test_small_batch_physics(mcl):
    """A small version of the batched configuration: ordered lattices at low T, disordered at high T, FP64 moments
    and chi from every lattice. The lattices are quenched from random directors for 600 sweeps, so a small periodic
    lattice may still hold a defect: the test separates ordered from disordered, the equilibrium values are the
    statistical parity tests' business."""
    specs = batch.c4_specs(24, seed=3, lmin=10, lmax=16, tmin=0.5, tmax=1.8)
    rows, info = batch.run_batch(specs, 300, 300, every=5, autotune=True, graph=16)
    T, S = rows[:, 0], rows[:, 1]
    bulk = np.array([s["geometry"] == "bulk" for s in specs])
    assert np.all(S[bulk & (T < 0.8)] > 0.5)
    assert np.all(S[bulk & (T > 1.5)] < 0.4)
    assert np.all(rows[:, 2] >= 0) and np.all(rows[:, 4] >= 0) and np.all(np.isfinite(rows))
    assert info["site_updates_per_ns"] > 0


def test_c4_batch_of_512_lattices(mcl, orc):
    """The batched configuration at full size: 512 LL lattices of 16^3-64^3, bulk or homeotropic sphere (W = 1.0),
    T uniform in [0.9, 1.3], one context with the autotuner's launch shape (trials at the run's cadence) and graph
    replay; MCT = 300, then MCS = 300 measured every 10. Every lattice gets 30 measurements with 0 < S < 1 and a mean
    site energy between the ground state and the disordered value; the bulk lattices below T = 1.0 are ordering (mean
    S above 0.1 and four times that of the bulk lattices above 1.2). For the smallest and the largest lattice and one sphere: the FP64 energy measured in the
    batch equals the oracle's energy of the same final directors (1e-6), and the lattice run alone (same seed and
    sweeps) ends with the same directors, bit for bit."""
    specs = batch.c4_specs(512, 1)
    sizes = np.array([s["L"] for s in specs])
    spheres = [r for r, s in enumerate(specs) if s["geometry"] == "sphere"]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), spheres[len(spheres) // 2]})

    with mcl.Engine([batch.build_lattice(s) for s in specs]) as eng:
        eng.autotune(10, 10)
        eng.set_graph(32)
        eng.sweeps(300)
        eng.reset_moments()
        eng.sweeps(300, 10)
        moms = [eng.moments(r) for r in range(len(specs))]
        finals = {r: eng.directors(r) for r in picks}
        measured = {r: eng.measure(r)["E"] for r in picks}
    S = np.array([m["S"] / m["n"] for m in moms])
    E = np.array([m["E"] / m["n"] for m in moms])
    assert all(m["n"] == 30 for m in moms)
    assert np.all((S > 0) & (S < 1)) and np.all((E > -10.0) & (E < -1.0))
    T = np.array([s["T"] for s in specs])
    bulk = np.array([s["geometry"] == "bulk" for s in specs])
    # from random directors, 300 sweeps order the cold lattices only partly (domains coarsen slowly at 64^3)
    cold, hot = S[bulk & (T < 1.0)].mean(), S[bulk & (T > 1.2)].mean()
    assert cold > 0.1 and cold > 4 * hot, (cold, hot)
    for r in picks:
        s = specs[r]
        L = s["L"]
        if s["geometry"] == "sphere":
            pt, ns = batch.sphere_geometry(L)
            p = orc.params(L, L, L, periodic=(0, 0, 0), potential="ll", A=1.0,
                           surfaces=[("homeotropic", 1.0, (0, 0, 0))])
            e = orc.energy_mean(p, finals[r], np.array(pt, dtype=np.int32), np.array(ns, dtype=np.float32))
        else:
            p = orc.params(L, L, L, periodic=(1, 1, 1), potential="ll", A=1.0)
            e = orc.energy_mean(p, finals[r], np.ones(L ** 3, dtype=np.int32))
        assert measured[r] == pytest.approx(e, rel=1e-6), (r, s)
        with mcl.Engine(batch.build_lattice(s)) as eng:
            eng.sweeps(300)
            eng.sweeps(300, 10)
            assert np.array_equal(eng.directors(), finals[r]), (r, s)
