"""The measurement folded into the colour-1 sweep pass against the separate measurement kernels (needs an MI355X).

After a measured sweep the moments hold one sample of the lattice that sweep left behind; mcl_measure then measures the
same lattice with k_measure. The energies come from different FP32 site sums (colour-1 final energies in the sweep,
colour-0 bond sums in k_measure), so they agree to FP32 rounding (1e-6 relative); S comes from the same directors
summed in FP64 in a different order (1e-10). Cases cover the fully periodic path, the field term, a free x axis of odd
length (the last column of colour 0 is counted from its -x side), the flip kernel and several lattices in one
context.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_dirs(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)


def lattice(mcl, dims, periodic, model, seed, T=1.0):
    nx, ny, nz = dims
    return mcl.Lattice(nx, ny, nz, rand_dirs(nx * ny * nz, seed), periodic=periodic, model=model, T=T, seed=seed)


CASES = {
    "ll-periodic": [((16, 16, 16), (1, 1, 1), ("ll", {}))],
    "ll-field": [((12, 10, 8), (1, 1, 1), ("ll", dict(elec_dir=(0.7071, 0.7071, 0.0), elecA=0.5)))],
    "ghrl-free-odd-x": [((9, 8, 6), (0, 1, 0), ("ghrl", dict(ghrl=(0.7272726, -1.6363635, -1.0, -0.1818182, 0.0))))],
    "ll-flip-free-z": [((10, 10, 10), (1, 1, 0), ("ll", dict(elec_dir=(0.0, 0.0, 1.0), elecA=0.3, elecP=0.2)))],
    "ll-three-lattices": [((8, 8, 8), (1, 1, 1), ("ll", {})), ((16, 12, 10), (1, 1, 1), ("ll", {})),
                          ((6, 14, 4), (1, 1, 1), ("ll", {}))],
}


@pytest.mark.parametrize("name", list(CASES))
def test_fused_measurement_equals_measure_kernel(mcl, name):
    lats = [lattice(mcl, dims, per, mcl.make_model(pot, **kw), 100 + k)
            for k, (dims, per, (pot, kw)) in enumerate(CASES[name])]
    with mcl.Engine(lats) as eng:
        assert eng.measure_in_sweep()
        eng.sweeps(3)
        for rep in range(2):
            eng.reset_moments()
            eng.sweeps(1, 1)
            for r in range(len(lats)):
                m = eng.moments(r)
                o = eng.measure(r)
                assert m["n"] == 1
                assert m["E"] == pytest.approx(o["E"], rel=1e-6, abs=1e-7), (name, r)
                assert m["S"] == pytest.approx(o["S"], rel=1e-10, abs=1e-12), (name, r)
                assert m["E2"] == pytest.approx(o["E"] ** 2, rel=2e-6)


def run_bulk(mcl, shape=None, graph=0, calls=((40, 4), (7, 0), (24, 1))):
    lats = [lattice(mcl, (12, 12, 12), (1, 1, 1), mcl.make_model("ll"), 200 + k, T=0.9 + 0.2 * k) for k in range(3)]
    with mcl.Engine(lats) as eng:
        if shape:
            eng.set_launch(*shape)
        if graph:
            eng.set_graph(graph)
        out = []
        for n, every in calls:
            eng.sweeps(n, every)
            out.append(([eng.directors(r) for r in range(3)], [eng.moments(r) for r in range(3)]))
        return out


def test_fused_measurement_under_graphs_and_launch_shapes(mcl):
    """Directors stay bit-identical across eager launches, graph replay and launch shapes; the moments, whose FP64
    partial sums follow the sweep's workgroups, agree to 1e-12 (measurements finalized in the next sweep's extra
    workgroups, in a trailing finalize, or inside a graph)."""
    ref = run_bulk(mcl)
    for kw in (dict(graph=8), dict(shape=(64, 0)), dict(shape=(1024, 1)), dict(shape=(128, 1), graph=12)):
        got = run_bulk(mcl, **kw)
        for (da, ma), (db, mb) in zip(ref, got):
            assert all(np.array_equal(x, y) for x, y in zip(da, db)), kw
            for a, b in zip(ma, mb):
                assert a["n"] == b["n"]
                for k in ("S", "S2", "E", "E2"):
                    assert b[k] == pytest.approx(a[k], rel=1e-12, abs=1e-15), (kw, k)


# Masked lattices without empty sites (slabs): the one-site terms of the colour-0 sites a colour-1 site counts come
# from their own point type and anchoring vector. Cases: the C3 physics (GHRL, homeotropic anchoring on both z
# surfaces, field along x), LL with planar (rp) anchoring, a field with a polar term and the flip trial, GHRL with the
# GHRL form of the homeotropic and planar anchorings, and a free x axis of odd length (last column of colour 0 counted
# from its -x side).
GHRL_K = (0.7272726, -1.6363635, -1.0, -0.1818182, 0.0)
MASKED = {
    "ghrl-slab-homeotropic-field": ((12, 10, 6), (1, 1, 0), ("ghrl", dict(ghrl=GHRL_K, elec_dir=(1.0, 0.0, 0.0),
                                                                           elecA=0.25)),
                                    [("homeotropic", 1.5, (0, 0, 0)), ("homeotropic", 1.5, (0, 0, 0))]),
    "ll-slab-planar-flip": ((10, 8, 6), (1, 1, 0), ("ll", dict(elec_dir=(0.0, 0.6, 0.8), elecA=0.3, elecP=0.2)),
                            [("rp", 1.0, (1.0, 0.0, 0.0)), ("rp", 0.7, (0.0, 1.0, 0.0))]),
    "ghrl-slab-ghrl-anchorings": ((8, 8, 8), (1, 1, 0), ("ghrl", dict(ghrl=GHRL_K)),
                                  [("homeotropic_ghrl", 1.2, (0, 0, 0)), ("rp_ghrl", 0.8, (1.0, 0.0, 0.0))]),
    "ll-slab-free-odd-x": ((9, 8, 6), (0, 1, 0), ("ll", {}),
                           [("homeotropic", 1.0, (0, 0, 0)), ("fg", 0.5, (0, 0, 0))]),
}


@pytest.mark.parametrize("name", list(MASKED))
def test_fused_measurement_on_masked_slabs(mcl, name):
    from mclics_amd import batch
    dims, per, (pot, kw), surfaces = MASKED[name]
    pt, ns = batch.slab_geometry(*dims)
    lat = mcl.Lattice(*dims, rand_dirs(dims[0] * dims[1] * dims[2], 7), periodic=per, pt=pt, ns=ns,
                      surfaces=surfaces, model=mcl.make_model(pot, **kw), T=0.8, seed=7)
    with mcl.Engine(lat) as eng:
        assert eng.measure_in_sweep()
        eng.sweeps(3)
        for rep in range(2):
            eng.reset_moments()
            eng.sweeps(1, 1)
            m = eng.moments()
            o = eng.measure()
            assert m["n"] == 1
            assert m["E"] == pytest.approx(o["E"], rel=1e-6, abs=1e-7), name
            assert m["S"] == pytest.approx(o["S"], rel=1e-10, abs=1e-12), name


# LL contexts with empty sites (stored as zero vectors on the device) and the radial order. An empty colour-1 position
# still counts the colour-0 sites it owns (its -x neighbour, and its +x neighbour at the end of a free x axis), an
# empty colour-0 site is not counted, and a lattice measuring a radial order carries its three sums in the compact
# partials. Cases: homeotropic spheres (even and odd size, with a field, with the flip trial, one above 2^18
# positions so that it takes the pair-per-thread instance), random holes in periodic, free odd-x and
# row-aligned (128-wide) lattices, cylinder order on a lattice with holes, and a batch mixing bulk lattices, spheres
# and a cylinder-order lattice.
def holes_pt(dims, seed, frac=0.2):
    rng = np.random.default_rng(seed)
    return (rng.random(dims[0] * dims[1] * dims[2]) >= frac).astype(np.int32)


def empty_case_lattice(mcl, spec, seed):
    from mclics_amd import batch
    kind = spec[0]
    if kind == "bulk":
        n = spec[1]
        return lattice(mcl, (n, n, n), (1, 1, 1), mcl.make_model("ll"), seed, T=1.1)
    if kind == "sphere":
        n, kw = spec[1], spec[2]
        pt, ns = batch.sphere_geometry(n)
        return mcl.Lattice(n, n, n, rand_dirs(n ** 3, seed), periodic=(0, 0, 0), pt=np.array(pt), ns=np.array(ns),
                           surfaces=[("homeotropic", 1.0, (0, 0, 0))], model=mcl.make_model("ll", **kw), T=1.0,
                           seed=seed, radial="sphere")
    dims, per, radial = spec[1], spec[2], spec[3]
    return mcl.Lattice(*dims, rand_dirs(dims[0] * dims[1] * dims[2], seed), periodic=per, pt=holes_pt(dims, seed),
                       model=mcl.make_model("ll"), T=0.9, seed=seed, radial=radial)


EMPTY = {
    "ll-sphere-radial": [("sphere", 14, {})],
    "ll-sphere-odd-field": [("sphere", 13, dict(elec_dir=(0.0, 0.0, 1.0), elecA=0.4))],
    "ll-sphere-flip": [("sphere", 12, dict(elec_dir=(0.0, 0.0, 1.0), elecA=0.2, elecP=0.3))],
    "ll-sphere-66": [("sphere", 66, {})],
    "ll-holes-periodic-cylinder": [("holes", (12, 10, 8), (1, 1, 1), "cylinder")],
    "ll-holes-free-odd-x": [("holes", (9, 8, 6), (0, 1, 1), "none")],
    "ll-holes-row-aligned": [("holes", (128, 4, 4), (1, 1, 1), "none")],
    "ll-batch-bulk-spheres-cylinder": [("bulk", 10), ("sphere", 12, {}), ("sphere", 9, {}),
                                       ("holes", (8, 6, 6), (1, 1, 0), "cylinder")],
}


@pytest.mark.parametrize("name", list(EMPTY))
def test_fused_measurement_with_empty_sites_and_radial_order(mcl, name):
    lats = [empty_case_lattice(mcl, spec, 300 + k) for k, spec in enumerate(EMPTY[name])]
    with mcl.Engine(lats) as eng:
        assert eng.measure_in_sweep()
        eng.sweeps(3)
        for rep in range(2):
            eng.reset_moments()
            eng.sweeps(1, 1)
            for r, lat in enumerate(lats):
                m = eng.moments(r)
                o = eng.measure(r)
                assert m["n"] == 1
                assert m["E"] == pytest.approx(o["E"], rel=1e-6, abs=1e-7), (name, r)
                assert m["S"] == pytest.approx(o["S"], rel=1e-10, abs=1e-12), (name, r)
                for k in ("Sr_sph", "Pr_sph", "Sr_cyl", "Pr_cyl"):
                    assert m[k] == pytest.approx(o[k], rel=1e-10, abs=1e-12), (name, r, k)
                if lat.radial == 1:
                    assert m["Sr_sph"] != 0.0 and m["Sr_cyl"] == 0.0
                if lat.radial == 2:
                    assert m["Sr_cyl"] != 0.0 and m["Sr_sph"] == 0.0


def test_both_radial_orders_measure_after_the_sweep(mcl):
    """A lattice measuring the sphere and the cylinder order keeps the separate measurement kernels."""
    lat = empty_case_lattice(mcl, ("holes", (10, 10, 10), (1, 1, 1), "both"), 5)
    with mcl.Engine(lat) as eng:
        assert not eng.measure_in_sweep()
        eng.sweeps(2, 1)
        m = eng.moments()
        assert m["n"] == 2 and m["Sr_sph"] != 0.0 and m["Sr_cyl"] != 0.0


def run_mixed(mcl, shape=None, graph=0, calls=((40, 4), (7, 0), (24, 1))):
    specs = [("bulk", 10), ("sphere", 12, {}), ("holes", (8, 6, 6), (1, 1, 0), "cylinder")]
    lats = [empty_case_lattice(mcl, spec, 400 + k) for k, spec in enumerate(specs)]
    with mcl.Engine(lats) as eng:
        if shape:
            eng.set_launch(*shape)
        if graph:
            eng.set_graph(graph)
        out = []
        for n, every in calls:
            eng.sweeps(n, every)
            out.append(([eng.directors(r) for r in range(3)], [eng.moments(r) for r in range(3)]))
        return out


def test_fused_radial_measurement_under_graphs_and_launch_shapes(mcl):
    """With empty sites and radial sums, directors stay bit-identical across eager launches, graph replay and
    launch shapes. The moments agree to 1e-12: a measurement finalized in the next sweep's extra workgroups and one
    finalized by the trailing k_finalize (every graph ends with one) sum the compact partials, radial rows included,
    in the same order, but the two compiled copies of the final arithmetic may contract different products into
    FMAs (the last bit of E), and across launch shapes the partials themselves follow the workgroups."""
    ref = run_mixed(mcl)
    for kw in (dict(graph=8), dict(shape=(64, 0)), dict(shape=(512, 1), graph=12)):
        got = run_mixed(mcl, **kw)
        for (da, ma), (db, mb) in zip(ref, got):
            assert all(np.array_equal(x, y) for x, y in zip(da, db)), kw
            for a, b in zip(ma, mb):
                assert a["n"] == b["n"]
                for k in ("S", "S2", "E", "E2", "Sr_sph", "Pr_sph", "Sr_cyl", "Pr_cyl"):
                    assert b[k] == pytest.approx(a[k], rel=1e-12, abs=1e-15), (kw, k)
