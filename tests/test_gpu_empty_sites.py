"""Sweeps of lattices with empty sites (point type 0; needs an MI355X).

Empty sites are skipped by the sweep and are never neighbours. At a very high temperature nearly every trial is
accepted: after a few sweeps every non-empty site must have moved, every empty site must still hold its input
director bit for bit, and the acceptance count of a sweep must lie just below the number of non-empty sites. The
trajectory does not depend on the launch shape nor on the other lattices of the context (the FP64 moments, whose
partial sums follow the sweep's workgroups when the measurement is folded into the sweep, agree to 1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def holes_lattice(mcl, dims, frac, seed, T, periodic=(1, 1, 1)):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    pt = (rng.random(nx * ny * nz) >= frac).astype(np.int32)
    v = rng.normal(size=(nx * ny * nz, 3))
    ni = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)
    return mcl.Lattice(nx, ny, nz, ni, periodic=periodic, pt=pt, model=mcl.make_model("ll"), T=T, seed=seed), pt, ni


@pytest.mark.parametrize("dims,periodic", [((20, 18, 16), (1, 1, 1)), ((15, 9, 7), (0, 0, 0)), ((64, 4, 4), (1, 1, 1))])
def test_every_non_empty_site_and_no_empty_site_is_updated(mcl, dims, periodic):
    lat, pt, ni0 = holes_lattice(mcl, dims, 0.3, 5, 50.0, periodic)
    sweeps = 12
    with mcl.Engine(lat) as eng:
        eng.set_ang_var(1.0)
        eng.sweeps(sweeps)
        d = eng.directors().reshape(-1, 3)
        acc, done = eng.sweep_stats()
    d0 = ni0.reshape(-1, 3)
    empty = pt == 0
    assert np.array_equal(d[empty], d0[empty])
    moved = np.any(d != d0, axis=1)
    assert moved[~empty].all(), int((~moved[~empty]).sum())
    assert np.abs(np.linalg.norm(d[~empty], axis=1) - 1).max() < 1e-5
    nvalid = int((~empty).sum())
    assert done == sweeps
    # the last sweep's accepted rotations: one trial per non-empty site, and at T = 50 almost all are accepted
    assert 0.8 * nvalid <= acc <= nvalid


def test_list_sweep_is_launch_shape_and_batch_independent(mcl):
    def run(shape, extra):
        lats = [holes_lattice(mcl, (24, 20, 18), 0.25, 9, 1.0)[0]]
        if extra:
            lats.append(holes_lattice(mcl, (12, 10, 8), 0.5, 10, 0.8)[0])
        with mcl.Engine(lats) as eng:
            eng.set_launch(*shape)
            eng.sweeps(30, 5)
            return eng.directors(0), eng.moments(0)
    ref_d, ref_m = run((256, 1), False)
    for shape, extra in (((64, 0), False), ((1024, 1), True), ((128, 0), True)):
        d, m = run(shape, extra)
        assert np.array_equal(d, ref_d), (shape, extra)
        assert m["n"] == ref_m["n"]
        for k in ref_m:
            assert m[k] == pytest.approx(ref_m[k], rel=1e-12, abs=1e-15), (shape, extra, k)


def test_set_directors_mid_run_keeps_the_empty_sites(mcl):
    """An LL context keeps its empty sites as zero vectors on the device (their bonds add nothing to the LL sum) and
    the caller's directors for them on the host: mcl_set_directors in the middle of a run must return them unchanged
    and must not change the sweep of the non-empty sites."""
    lat, pt, ni0 = holes_lattice(mcl, (18, 16, 14), 0.3, 3, 1.0)
    rng = np.random.default_rng(4)
    with mcl.Engine(lat) as eng, mcl.Engine(holes_lattice(mcl, (18, 16, 14), 0.3, 3, 1.0)[0]) as twin:
        eng.sweeps(10, 5)
        twin.sweeps(10, 5)
        d = eng.directors().reshape(-1, 3)
        empty = pt == 0
        new_empty = rng.normal(size=(int(empty.sum()), 3)).astype(np.float32)
        d[empty] = new_empty
        eng.set_directors(d.reshape(-1))
        got = eng.directors().reshape(-1, 3)
        assert np.array_equal(got[empty], new_empty)
        eng.sweeps(10, 5)
        twin.sweeps(10, 5)
        a, b = eng.directors().reshape(-1, 3), twin.directors().reshape(-1, 3)
        assert np.array_equal(a[~empty], b[~empty])
        assert np.array_equal(a[empty], new_empty)
        assert eng.moments() == twin.moments()
