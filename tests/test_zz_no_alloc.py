"""mcl_site_energies_noalloc, the site-energy entry point that allocates no device memory (its staging lives in a
static device array per GPU), against mcl_site_energies on the same state: bit-identical energies on two calls in
a row, the second with fewer sites, on lattices that exercise every kernel instance of the energy function (masked
surfaces with a field, GHRL and pear). Sizes beyond the array are rejected before any device work."""
import ctypes

import numpy as np
import pytest

from helpers import lattice_from_orc, rand_unit

pytestmark = pytest.mark.gpu

MAX_SITES = 65536  # MCL_SCRATCH_SITES


def cases(orc):
    out = []
    pt, ns = orc.geometry("slab", 24, 20, 12)
    p = orc.params(24, 20, 12, periodic=(1, 1, 0), potential="ghrl", k=(1.0, 0.6, 1.4, 0.0), elecA=0.3,
                   elec=(1, 0, 0), T=0.5, surfaces=[("homeotropic", 1.5, (0, 0, 0)), ("rp", 1.0, (1, 0, 0))])
    out.append((p, pt, ns))
    pt, ns = orc.geometry("sphere", 18, 18, 18)
    p = orc.params(18, 18, 18, periodic=(0, 0, 0), potential="pear", A=1.5, B1=0.09, C=0.3, T=0.8,
                   surfaces=[("homeotropic", 1.0, (0, 0, 0))])
    out.append((p, pt, ns))
    pt, ns = orc.geometry("bulk", 48, 48, 48)
    p = orc.params(48, 48, 48, potential="ll", T=1.1)
    out.append((p, pt, None))
    return out


@pytest.mark.parametrize("which", [0, 1, 2], ids=["ghrl-slab-field", "pear-sphere", "ll-bulk"])
def test_noalloc_site_energies_equal_the_allocating_ones(mcl, orc, which):
    p, pt, ns = cases(orc)[which]
    ni = orc.ic_random(pt, 7)
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, ns)) as eng:
        eng.sweeps(3)
        valid = int((pt > 0).sum())
        first = min(valid, MAX_SITES)
        for n, seed in ((first, 1), (max(1, first // 7), 2)):
            rng = np.random.default_rng(seed)
            sites = rng.choice(np.flatnonzero(pt), size=n, replace=n > valid).astype(np.int32)
            trial = rand_unit(rng, n)
            ec, et = eng.site_energies(sites, trial)
            ec2, et2 = eng.site_energies_noalloc(sites, trial)
            assert np.array_equal(ec, ec2) and np.array_equal(et, et2), (n, seed)


def test_noalloc_rejects_more_sites_than_its_array(mcl, orc):
    p, pt, ns = cases(orc)[2]
    with mcl.Engine(lattice_from_orc(mcl, p, orc.ic_random(pt, 3), pt, None)) as eng:
        n = MAX_SITES + 1
        sites = np.zeros(n, dtype=np.int32)
        trial = np.tile(np.array([0, 0, 1], dtype=np.float32), n)
        ec = np.empty(n, dtype=np.float32)
        et = np.empty(n, dtype=np.float32)
        F = ctypes.POINTER(ctypes.c_float)
        st = eng.lib.mcl_site_energies_noalloc(eng.h, 0, sites.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                               trial.ctypes.data_as(F), n, ec.ctypes.data_as(F),
                                               et.ctypes.data_as(F))
        assert st == 1 and b"at most 65536 per call" in eng.lib.mcl_last_error()
        # the context still works
        e1, e2 = eng.site_energies_noalloc(sites[:10], trial[:30])
        assert np.all(np.isfinite(e1)) and np.all(np.isfinite(e2))
