"""One position per thread (the SPLIT sweep instance) for small contexts: a context of at most 2^18 colour-block
positions without the flip trial and not row-aligned runs it. The draws and trials are those of the pair path, so a
lattice swept alone (split) must follow the trajectory it follows inside a larger context (pairs) bit for bit."""
import numpy as np
import pytest

from mclics_amd import batch


def _ghrl(capi, k=(1.0, 0.6, 1.4, 0.0), **kw):
    c = capi.ghrl_coefficients(*k)
    return capi.make_model("ghrl", ghrl=(c["lam"], c["mu"], c["nu"], c["rho"], c["sigma"]), **kw)


def _lattice(capi, name):
    ones = lambda n: np.ones(n, dtype=np.int32)
    if name in ("ll_32", "ll_64"):
        n = int(name[3:])
        return capi.Lattice(n, n, n, batch.random_directors(ones(n ** 3), n), model=capi.make_model("ll"), T=1.0,
                            seed=n + 1)
    if name == "ll_sphere_32":
        pt, ns = batch.sphere_geometry(32)
        return capi.Lattice(32, 32, 32, batch.random_directors(pt, 5), periodic=(0, 0, 0), pt=pt, ns=ns,
                            surfaces=[("homeotropic", 1.0, (0, 0, 0))], model=capi.make_model("ll"), T=0.9, seed=6)
    if name == "ll_odd_free_33x30x20":
        # odd nx: the colour blocks carry padding positions; free x and z, a field
        return capi.Lattice(33, 30, 20, batch.random_directors(ones(33 * 30 * 20), 11), periodic=(0, 1, 0),
                            model=capi.make_model("ll", elec_dir=(0.0, 0.6, 0.8), elecA=0.4), T=0.7, seed=12)
    assert name == "ghrl_slab_24x24x20"
    pt, ns = batch.slab_geometry(24, 24, 20)
    return capi.Lattice(24, 24, 20, batch.random_directors(pt, 7), periodic=(1, 1, 0), pt=pt, ns=ns,
                        surfaces=[("homeotropic", 1.5, (0, 0, 0)), ("homeotropic", 1.5, (0, 0, 0))],
                        model=_ghrl(capi, elec_dir=(1.0, 0.0, 0.0), elecA=0.25, elecE=0.8), T=0.5, seed=8)


def _run(capi, lats):
    with capi.Engine(lats) as eng:
        eng.sweeps(25)
        eng.reset_moments()
        eng.sweeps(6, 2)
        eng.sweeps(3)
        return eng.directors(0), eng.ang_var(0), eng.sweep_stats(0), eng.moments(0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ll_32", "ll_64", "ll_sphere_32", "ll_odd_free_33x30x20", "ghrl_slab_24x24x20"])
def test_split_sweeps_follow_the_pair_path(mcl, name):
    print("split case", name, flush=True)
    lat = _lattice(mcl, name)
    # a 96 x 96 x 64 partner makes the context too large for the split instance (294,912 positions per colour)
    big = mcl.Lattice(96, 96, 64, batch.random_directors(np.ones(96 * 96 * 64, dtype=np.int32), 3),
                      model=lat.model, T=1.0, seed=9)
    da, ava, sa, ma = _run(mcl, lat)
    db, avb, sb, mb = _run(mcl, [lat, big])
    assert np.array_equal(da, db), (name, int(np.sum(da != db)))
    assert ava == avb and sa == sb, (name, ava, avb, sa, sb)
    # the FP64 measurement sums group differently in the two contexts
    for k in ma:
        assert ma[k] == pytest.approx(mb[k], rel=1e-6, abs=1e-9), (name, k)
    assert not np.array_equal(da, lat.ni)
