"""Resident sweeps (k_sweep_resident): small lattices held in the LDS of one workgroup run whole sweeps in one launch.
They must follow the two-pass path's trajectory bit for bit, and leave the same sweep state (ang_var, accepted
counts, sweep number) for the passes and measurements that follow."""
import numpy as np
import pytest

from mclics_amd import batch, freedericksz


def _ghrl(capi, k=(1.0, 0.6, 1.4, 0.0), **kw):
    c = capi.ghrl_coefficients(*k)
    return capi.make_model("ghrl", ghrl=(c["lam"], c["mu"], c["nu"], c["rho"], c["sigma"]), **kw)


def _cases(capi):
    """name -> list of Lattice: every kernel path the resident sweep has (potential, mask, surfaces, free
    boundaries, field, flip trial, odd nx with padding positions, a batch of unequal lattices)."""
    rnd = lambda pt, seed: batch.random_directors(pt, seed)
    ones = lambda n: np.ones(n, dtype=np.int32)
    out = {}
    out["ll_bulk_16"] = [capi.Lattice(16, 16, 16, rnd(ones(16 ** 3), 1), model=capi.make_model("ll"), T=1.1, seed=3)]
    out["fz_splay_slab"] = [freedericksz.cell_lattice("splay", 2, 0.0008, 10.0, 0.25)]
    out["fz_bend_slab_T"] = [freedericksz.cell_lattice("bend", 1, 0.3, 10.0, 0.25)]
    out["pear_flip_12"] = [capi.Lattice(12, 12, 12, rnd(ones(12 ** 3), 2), model=capi.make_model("pear", A=1.5),
                                        T=0.8, seed=4)]
    pt, ns = batch.sphere_geometry(14)
    out["ll_sphere_14"] = [capi.Lattice(14, 14, 14, rnd(pt, 5), periodic=(0, 0, 0), pt=pt, ns=ns,
                                        surfaces=[("homeotropic", 1.0, (0, 0, 0))], model=capi.make_model("ll"),
                                        T=0.9, seed=6)]
    pt, ns = batch.sphere_geometry(16)
    out["ghrl_chiral_sphere_16"] = [capi.Lattice(16, 16, 16, rnd(pt, 7), periodic=(0, 0, 0), pt=pt, ns=ns,
                                                 surfaces=[("homeotropic", 1.5, (0, 0, 0))],
                                                 model=_ghrl(capi, (1.0, 0.6, 1.4, 12.0)), T=0.5, seed=8)]
    out["ll_odd_free_9x10x12"] = [capi.Lattice(9, 10, 12, rnd(ones(9 * 10 * 12), 9), periodic=(0, 1, 0),
                                               model=capi.make_model("ll", elec_dir=(0.0, 0.6, 0.8), elecA=0.4),
                                               T=0.7, seed=10)]
    out["ll_batch"] = [capi.Lattice(n, n, n, rnd(ones(n ** 3), 20 + n), model=capi.make_model("ll"),
                                    T=0.9 + 0.05 * n / 4, seed=30 + n) for n in (4, 8, 10, 12, 14, 16)]
    # at most 1024 positions per colour: one position per thread (its own instance)
    out["ll_batch_small"] = [capi.Lattice(n, n, n, rnd(ones(n ** 3), 40 + n), model=capi.make_model("ll"),
                                          T=1.0, seed=50 + n) for n in (4, 6, 8, 10, 12)]
    pt, ns = batch.sphere_geometry(12)
    out["ghrl_chiral_sphere_12"] = [capi.Lattice(12, 12, 12, rnd(pt, 11), periodic=(0, 0, 0), pt=pt, ns=ns,
                                                 surfaces=[("homeotropic", 1.5, (0, 0, 0))],
                                                 model=_ghrl(capi, (1.0, 0.6, 1.4, 12.0)), T=0.5, seed=12)]
    return out


def _run(capi, lats, resident):
    """A run that mixes plain blocks (resident when on), a host-set ang_var and measured blocks (always two-pass)."""
    with capi.Engine(lats) as eng:
        eng.set_resident(resident)
        assert eng.resident() == resident
        eng.sweeps(37)
        eng.reset_moments()
        eng.sweeps(6, 1)
        eng.sweeps(1)
        eng.set_ang_var(0.31)
        eng.sweeps(20)
        eng.sweeps(4, 2)
        eng.sweeps(2)
        out = []
        for r in range(len(lats)):
            out.append((eng.directors(r), eng.ang_var(r), eng.sweep_stats(r), eng.moments(r)))
        return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ll_bulk_16", "fz_splay_slab", "fz_bend_slab_T", "pear_flip_12", "ll_sphere_14",
                                  "ghrl_chiral_sphere_16", "ll_odd_free_9x10x12", "ll_batch", "ll_batch_small",
                                  "ghrl_chiral_sphere_12"])
def test_resident_sweeps_are_bit_identical(mcl, name):
    print("resident case", name, flush=True)
    lats = _cases(mcl)[name]
    a = _run(mcl, lats, False)
    b = _run(mcl, lats, True)
    for r, ((da, ava, sa, ma), (db, avb, sb, mb)) in enumerate(zip(a, b)):
        assert np.array_equal(da, db), (name, r, int(np.sum(da != db)))
        assert ava == avb and sa == sb, (name, r, ava, avb, sa, sb)
        # the measured sweeps run the same two-pass kernels from the same states; only the FP64 sums' grouping
        # could differ, and it does not, since the launch shape is the same
        assert ma == mb, (name, r)
    assert any(not np.array_equal(d, lat.ni) for (d, _, _, _), lat in zip(b, lats))


@pytest.mark.gpu
def test_resident_eligibility(mcl):
    """Contexts whose every lattice has at most 2048 colour-block positions per colour (1024 with the flip trial)
    run resident; larger ones do not, since one workgroup would update more sites per lane than the two passes."""
    ones = lambda n: np.ones(n, dtype=np.int32)
    for dims, pot, ok in (((16, 16, 16), "ll", True), ((17, 16, 16), "ll", False), ((18, 18, 18), "ghrl", False),
                          ((10, 10, 16), "ghrl", True), ((128, 2, 2), "ll", True), ((12, 12, 12), "pear", True),
                          ((14, 14, 14), "pear", False)):
        lat = mcl.Lattice(*dims, batch.random_directors(ones(dims[0] * dims[1] * dims[2]), 1), periodic=(0, 0, 0),
                          model=mcl.make_model(pot))
        with mcl.Engine(lat) as eng:
            assert eng.resident() == ok, (dims, pot)
    small = mcl.Lattice(8, 8, 8, batch.random_directors(ones(512), 1), model=mcl.make_model("ll"))
    big = mcl.Lattice(24, 24, 24, batch.random_directors(ones(24 ** 3), 2), model=mcl.make_model("ll"))
    with mcl.Engine([small, big]) as eng:
        assert not eng.resident()
