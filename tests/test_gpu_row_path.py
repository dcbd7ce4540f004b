"""Row-aligned gathering against the generic one (needs an MI355X).

A context whose lattices all have nx a multiple of 128 sweeps with gather_row: the row, j, k, the y / z wraps and
the y / z free faces of each 64-lane group are scalar work. Batched with a lattice that is not row-aligned, the same
lattice is swept by the generic gather. Batching never changes a lattice's trajectory, and the trial's FP32
arithmetic is written with explicit fmaf, so the two must be bit-identical: any wrong neighbour index, wrap or
existence bit in either path shows up as a different director field. The cases cover the wraps in every direction
(ny and nz of 2 to 8 rows), nxh of 64 and 128 (one and two lane groups per row), free faces in x, y and z, surface
point types and empty sites, measured and unmeasured sweeps, the field term, the flip trial and the three
potentials.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_dirs(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


GHRL = dict(ghrl=(0.7272726, -1.6363635, -1.0, -0.1818182, 0.0))
CASES = {
    "ll-128x6x4": ((128, 6, 4), "ll", {}, 1.0),
    "ll-256x2x6": ((256, 2, 6), "ll", {}, 0.7),
    "ll-128x2x2": ((128, 2, 2), "ll", {}, 1.2),
    "ll-128x4x8": ((128, 4, 8), "ll", {}, 0.9),
    "ll-field-flip": ((128, 4, 4), "ll", dict(elec_dir=(0.0, 0.6, 0.8), elecA=0.4, elecP=0.3), 1.0),
    "ghrl-field": ((128, 4, 6), "ghrl", dict(GHRL, elec_dir=(1.0, 0.0, 0.0), elecA=0.25), 0.6),
    "pear": ((128, 6, 4), "pear", dict(A=1.0, B1=0.2, C=0.5), 0.8),
}


def lattice(mcl, dims, pot, kw, T, seed):
    nx, ny, nz = dims
    return mcl.Lattice(nx, ny, nz, rand_dirs(nx * ny * nz, seed), periodic=(1, 1, 1), model=mcl.make_model(pot, **kw),
                       T=T, seed=seed)


def run(mcl, lats, calls):
    out = []
    with mcl.Engine(lats) as eng:
        for n, every in calls:
            eng.sweeps(n, every)
            out.append((eng.directors(0), eng.moments(0)))
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_row_path_equals_generic_path(mcl, name):
    dims, pot, kw, T = CASES[name]
    calls = ((12, 4), (5, 0), (9, 3))
    alone = run(mcl, [lattice(mcl, dims, pot, kw, T, 11)], calls)
    mixed = run(mcl, [lattice(mcl, dims, pot, kw, T, 11), lattice(mcl, (16, 12, 10), pot, kw, T, 12)], calls)
    for k, ((da, ma), (db, mb)) in enumerate(zip(alone, mixed)):
        assert np.array_equal(da, db), (name, k, int(np.sum(da != db)))
        assert ma["n"] == mb["n"]
        for q in ("S", "S2", "E", "E2"):
            assert mb[q] == pytest.approx(ma[q], rel=1e-12, abs=1e-15), (name, k, q)
    # the run did something: directors moved and stayed unit vectors
    d = alone[-1][0].reshape(-1, 3)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
    assert not np.array_equal(alone[-1][0], rand_dirs(dims[0] * dims[1] * dims[2], 11))


def masked_lattice(mcl, dims, per, pot, kw, surfaces, geom, T, seed):
    """A slab (surface planes at k = 0 and nz - 1), or a bulk box with 12 % of its sites empty ('holes')."""
    from mclics_amd import batch
    nx, ny, nz = dims
    if geom == "slab":
        pt, ns = batch.slab_geometry(nx, ny, nz)
    elif geom == "holes":
        rng = np.random.default_rng(seed)
        pt = (rng.random(nx * ny * nz) >= 0.12).astype(np.int32)
        ns = None
    else:
        pt, ns = None, None
    return mcl.Lattice(nx, ny, nz, rand_dirs(nx * ny * nz, seed), periodic=per, pt=pt, ns=ns, surfaces=surfaces,
                       model=mcl.make_model(pot, **kw), T=T, seed=seed)


# (dims, periodic, potential, model, surfaces, geometry, T): free faces in every direction, surface point types,
# empty sites, the field, the flip trial; measured sweeps of slabs go through the in-sweep measurement
MASKED = {
    "ghrl-slab-256x4x6": ((256, 4, 6), (1, 1, 0), "ghrl", dict(GHRL, elec_dir=(1.0, 0.0, 0.0), elecA=0.25),
                          [("homeotropic", 1.5, (0, 0, 0)), ("homeotropic", 1.5, (0, 0, 0))], "slab", 0.5),
    "ll-slab-128x6x4-planar-flip": ((128, 6, 4), (1, 1, 0), "ll", dict(elec_dir=(0.0, 0.6, 0.8), elecA=0.3,
                                                                       elecP=0.2),
                                    [("rp", 1.0, (1.0, 0.0, 0.0)), ("rp", 0.7, (0.0, 1.0, 0.0))], "slab", 0.9),
    "ll-free-x-128x4x4": ((128, 4, 4), (0, 1, 1), "ll", {}, [], "bulk", 1.0),
    "ghrl-free-all-128x4x6": ((128, 4, 6), (0, 0, 0), "ghrl", GHRL, [], "bulk", 0.7),
    "ll-holes-128x4x6": ((128, 4, 6), (1, 1, 1), "ll", {}, [], "holes", 1.1),
}


@pytest.mark.parametrize("name", list(MASKED))
def test_row_path_equals_generic_path_masked_and_free(mcl, name):
    dims, per, pot, kw, surfaces, geom, T = MASKED[name]
    calls = ((12, 4), (5, 0), (9, 1))
    alone = run(mcl, [masked_lattice(mcl, dims, per, pot, kw, surfaces, geom, T, 21)], calls)
    mixed = run(mcl, [masked_lattice(mcl, dims, per, pot, kw, surfaces, geom, T, 21),
                      masked_lattice(mcl, (16, 12, 10), per, pot, kw, surfaces, geom, T, 22)], calls)
    for k, ((da, ma), (db, mb)) in enumerate(zip(alone, mixed)):
        assert np.array_equal(da, db), (name, k, int(np.sum(da != db)))
        assert ma["n"] == mb["n"]
        for q in ("S", "S2", "E", "E2"):
            assert mb[q] == pytest.approx(ma[q], rel=1e-12, abs=1e-15), (name, k, q)
    d = alone[-1][0].reshape(-1, 3)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
