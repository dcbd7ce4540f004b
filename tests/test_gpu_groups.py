"""Lattice groups (mcl_set_lattice_groups): a context larger than the Infinity Cache is swept group by group, each
group through all the sweeps of a call before the next. A lattice's trajectory depends only on its own sites and
sweep numbers, so the grouped order must give bit-identical directors, ang_var and acceptance counts to the
ungrouped one: eager and graph-replayed, a batch with spheres and a bulk batch (both measure in the sweep), under two
launch shapes. The moments agree to 1e-12: a group's last measurement of a call is finalized by a trailing kernel,
whose compiled arithmetic may round the last bit of an FP64 sum differently from the in-sweep finalize."""
import numpy as np
import pytest

from mclics_amd import batch

pytestmark = pytest.mark.gpu

MIXED = [dict(L=24, geometry="bulk", T=1.05, seed=31), dict(L=30, geometry="sphere", T=0.9, seed=32, W=1.0),
         dict(L=16, geometry="bulk", T=0.7, seed=33), dict(L=36, geometry="bulk", T=1.2, seed=34),
         dict(L=22, geometry="sphere", T=1.1, seed=35, W=1.0), dict(L=28, geometry="bulk", T=1.0, seed=36),
         dict(L=40, geometry="sphere", T=0.95, seed=37, W=1.0)]
BULK = [dict(L=s["L"], geometry="bulk", T=s["T"], seed=s["seed"]) for s in MIXED]
# more than 2^18 colour-block positions: the pair-per-thread instance of the real batch (the sets above take the
# one-position-per-thread instance of small contexts)
LARGE = [dict(L=48, geometry="bulk", T=1.05, seed=41), dict(L=56, geometry="sphere", T=0.9, seed=42, W=1.0),
         dict(L=64, geometry="bulk", T=1.15, seed=43), dict(L=40, geometry="sphere", T=1.0, seed=44, W=1.0),
         dict(L=44, geometry="bulk", T=0.8, seed=45)]
CALLS = ((20, 0), (30, 5), (7, 0), (40, 10))


def state(eng, n):
    return ([eng.directors(r) for r in range(n)], [eng.ang_var(r) for r in range(n)],
            [eng.sweep_stats(r) for r in range(n)], [eng.moments(r) for r in range(n)])


def run(mcl, specs, group_bytes, graph, shape):
    with mcl.Engine([batch.build_lattice(s) for s in specs]) as eng:
        eng.set_launch(*shape)
        eng.set_lattice_groups(group_bytes)
        ngroups = eng.lattice_groups()
        if graph:
            eng.set_graph(graph)
        out = []
        for n, every in CALLS:
            eng.sweeps(n, every)
            out.append(state(eng, len(specs)))
        return ngroups, out


@pytest.mark.parametrize("specs", [MIXED, BULK, LARGE], ids=["mixed", "bulk", "large"])
@pytest.mark.parametrize("graph", [0, 16], ids=["eager", "graph"])
@pytest.mark.parametrize("shape", [(256, 1), (512, 0)], ids=["256x", "512"])
def test_groups_give_the_ungrouped_trajectories(mcl, specs, graph, shape):
    ng0, ref = run(mcl, specs, 0, graph, shape)
    # about a fifth of the context per group: several groups of one to three lattices
    total = sum(2 * -(-s["L"] // 2) * s["L"] ** 2 * 13 for s in specs)
    ng, got = run(mcl, specs, total // 5, graph, shape)
    assert ng0 == 1 and 3 <= ng <= len(specs), (ng0, ng)
    for k, (a, b) in enumerate(zip(ref, got)):
        assert all(np.array_equal(x, y) for x, y in zip(a[0], b[0])), (k, "directors")
        assert a[1] == b[1] and a[2] == b[2], (k, "ang_var / counts")
        for ma, mb in zip(a[3], b[3]):
            assert ma["n"] == mb["n"], (k, "moments")
            for q in ma:
                assert mb[q] == pytest.approx(ma[q], rel=1e-12, abs=1e-15), (k, "moments", q)


def test_the_c4_batch_is_grouped_and_timed_in_whole_passes(mcl):
    """The 512-lattice 16^3-64^3 batch holds about 600 MB of directors and point types, so it is swept in groups
    by default; kernel timing still counts whole-context colour passes (2 per sweep)."""
    specs = batch.c4_specs(512, 1)
    with mcl.Engine([batch.build_lattice(s) for s in specs]) as eng:
        assert eng.lattice_groups() >= 3
        eng.set_kernel_timing(5, 40)
        eng.sweeps(20, 0)
        launches, ms = eng.kernel_timing()
        assert launches == 40 and ms > 0
