"""Parity of the libmclgpu kernels with the CPU oracle, through the C ABI (needs an MI355X).

Tolerances: energies and dE within 1e-5 relative to the energy scale |E_old| + |E_new| (FP32 site energies);
FP64 reductions within 1e-9 relative; integer and RNG outputs bit-exact.
"""
import math

import numpy as np
import pytest

from helpers import lattice_from_orc, rand_unit, sites_and_trials
from test_oracle_known_answers import (CASES, FLIP_CASES, PHILOX_KAT, base_run, boltzmann_cos_hist, build_case,
                                      chi2_of, flip_case, hedgehog)

pytestmark = pytest.mark.gpu


def test_philox_device_kat(mcl, orc):
    ctr = np.array([c for c, _, _ in PHILOX_KAT], dtype=np.uint32)
    key = np.array([k for _, k, _ in PHILOX_KAT], dtype=np.uint32)
    out = mcl.philox(ctr, key)
    assert [tuple(int(x) for x in r) for r in out] == [w for _, _, w in PHILOX_KAT]
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2 ** 32, size=(1000, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, size=(1000, 2), dtype=np.uint64).astype(np.uint32)
    out = mcl.philox(ctr, key)
    for q in range(0, 1000, 97):
        assert list(out[q]) == orc.philox([int(x) for x in ctr[q]], [int(x) for x in key[q]])


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (c["potential"], c["geometry"],
                                                                    "".join(map(str, c["periodic"]))))
def test_site_energies_and_delta_e(mcl, orc, case):
    """E_old, E_new and dE of the sweep kernel's energy function against the oracle's site energy."""
    p, ni, pt, ns = build_case(orc, case)
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, ns)) as eng:
        sites, trial = sites_and_trials(pt, 200, 4)
        ec, et = eng.site_energies(sites, trial)
    for s, t, a, b in zip(sites, trial, ec, et):
        ref_c = orc.site_energy(p, ni, pt, ns, s, ni[3 * s:3 * s + 3])
        ref_t = orc.site_energy(p, ni, pt, ns, s, t)
        scale = abs(ref_c) + abs(ref_t) + 1e-3
        assert abs(a - ref_c) <= 1e-5 * scale
        assert abs(b - ref_t) <= 1e-5 * scale
        assert abs((b - a) - (ref_t - ref_c)) <= 1e-5 * scale


GROUND = [
    (dict(), -9.0),
    (dict(potential="ghrl", k=(1, 1, 1, 0)), -6.0),
    (dict(potential="pear", A=1.5, B1=0.09, C=0.3), 6 * (-1.5 - 0.09)),
    (dict(Nz=6, geometry="slab", periodic=(1, 1, 0),
          surfaces=[("homeotropic", 1.0, (0, 0, 0)), ("homeotropic", 1.0, (0, 0, 0))]),
     (-1.5 * (256 * 6 + 128 * 5) - 128) / 384),
    (dict(elecA=0.5, elec=(0, 0, 2)), -9.5),
]


@pytest.mark.parametrize("kw,E", GROUND, ids=["ll", "ghrl", "pear", "slab", "field"])
def test_ground_states_through_sweep(mcl, orc, kw, E):
    """run_checks.py:113-161 through the GPU sweep: T = 0, uniform along z, one sweep with one measurement."""
    r, p, ni, pt, ns = base_run(orc, **kw)
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    pt2, ns2 = orc.geometry(kw.get("geometry", "bulk"), Nx, Ny, Nz)
    ni2 = orc.ic_homogeneous(pt2, 0, 0)
    with mcl.Engine(lattice_from_orc(mcl, p, ni2, pt2, ns2, T=0.0)) as eng:
        eng.sweeps(1, 1)
        m = eng.moments()
    assert m["n"] == 1
    assert m["E"] == pytest.approx(E, abs=1e-4)
    assert m["S"] == pytest.approx(1.0, abs=1e-5)
    assert m["E2"] - m["E"] ** 2 == pytest.approx(0.0, abs=1e-8)
    if kw.get("potential") == "pear":
        assert m["P"] == pytest.approx(1.0, abs=1e-5)


@pytest.mark.parametrize("case", CASES[:6] + CASES[7:], ids=lambda c: "%s-%s" % (c["potential"], c["geometry"]))
def test_measure_vs_oracle(mcl, orc, case):
    """FP64 measurement (energy, Q, S, director, radial, pear observables) against the oracle and numpy."""
    p, ni, pt, ns = build_case(orc, case, seed=2)
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, ns, radial=3)) as eng:
        o = eng.measure()
    assert o["nvalid"] == int((pt > 0).sum())
    assert o["E"] == pytest.approx(orc.energy_mean(p, ni, pt, ns), rel=1e-6, abs=1e-7)
    n = ni.reshape(-1, 3)[pt > 0].astype(np.float64)
    Q = (3 * n.T @ n / len(n) - np.eye(3)) / 2
    assert np.allclose(o["Q"][:5], [Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2]], atol=1e-12)
    Qz = Q.copy()
    Qz[2, 2] = -Q[0, 0] - Q[1, 1]
    w, v = np.linalg.eigh(Qz)
    assert o["S"] == pytest.approx(w[-1], abs=1e-10)
    assert abs(abs(np.dot(o["director"], v[:, -1])) - 1) < 1e-8
    for cyl, ks, kp in ((False, "Sr_sph", "Pr_sph"), (True, "Sr_cyl", "Pr_cyl")):
        Sr, Pr = orc.radial_order(p, ni, pt, cyl)
        assert o[ks] == pytest.approx(Sr, abs=1e-6) and o[kp] == pytest.approx(Pr, abs=1e-6)
    if case["potential"] == "pear":
        assert o["P"] == pytest.approx(orc.polarization(ni, pt, o["director"]), abs=1e-9)
        assert o["dtheta"] == pytest.approx(orc.splay_magnitude(p, ni, pt), rel=1e-5, abs=1e-7)
        assert o["divn"] == pytest.approx(orc.splay_divergence(p, ni, pt), rel=1e-5, abs=1e-7)


def test_hedgehog_observables(mcl, orc):
    n = 21
    pt, ns, ni = hedgehog(orc, n)
    p = orc.params(n, n, n, periodic=(0, 0, 0), potential="pear", A=1.5, B1=0.09, C=0.3,
                   surfaces=[("homeotropic", 1.5, (0, 0, 0))])
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, ns, radial=1)) as eng:
        o = eng.measure()
    assert o["Sr_sph"] == pytest.approx(1.0, abs=0.03) and o["Pr_sph"] == pytest.approx(1.0, abs=0.03)
    assert o["divn"] == pytest.approx(orc.splay_divergence(p, ni, pt), abs=1e-6)
    assert o["divn"] == pytest.approx(0.275, abs=0.02)


def test_local_order_vs_oracle(mcl, orc):
    for case in (CASES[0], CASES[4], CASES[7]):
        p, ni, pt, ns = build_case(orc, case, seed=6)
        with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, ns)) as eng:
            lo = eng.local_order()
        for idx in range(0, len(pt), max(1, len(pt) // 60)):
            i, j, k = idx % p.Nx, (idx // p.Nx) % p.Ny, idx // (p.Nx * p.Ny)
            if pt[idx] == 0:
                assert math.isnan(lo[idx])
            else:
                assert lo[idx] == pytest.approx(orc.local_order(p, ni, pt, i, j, k), abs=2e-6)


@pytest.mark.parametrize("dims,per", [((8, 6, 4), (1, 1, 1)), ((7, 5, 3), (0, 0, 0)), ((1, 1, 1), (0, 0, 0)),
                                      ((2, 9, 5), (1, 0, 0))])
def test_director_roundtrip(mcl, orc, dims, per):
    rng = np.random.default_rng(1)
    nt = dims[0] * dims[1] * dims[2]
    ni = rand_unit(rng, nt).reshape(-1)
    p = orc.params(*dims, periodic=per)
    pt = np.ones(nt, dtype=np.int32)
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None)) as eng:
        assert np.array_equal(eng.directors(), ni)
        ni2 = rand_unit(rng, nt).reshape(-1)
        eng.set_directors(ni2)
        assert np.array_equal(eng.directors(), ni2)


def test_sweep_keeps_unit_norm_and_is_deterministic(mcl, orc):
    N = 16
    pt = np.ones(N ** 3, dtype=np.int32)
    p = orc.params(N, N, N, T=1.1)
    ni = orc.ic_random(pt, 5)
    outs = []
    for seed in (1, 1, 2):
        with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None, seed=seed)) as eng:
            eng.sweeps(100, 10)
            outs.append(eng.directors())
            m = eng.moments()
            assert m["n"] == 10
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], outs[2])
    norms = np.linalg.norm(outs[0].reshape(-1, 3), axis=1)
    assert np.abs(norms - 1).max() < 1e-5


def test_ang_var_rule_on_device(mcl, orc):
    """The device applies the ang_var rule of evolve.cu:430-436 to the accepted count of the sweep."""
    N = 8
    pt = np.ones(N ** 3, dtype=np.int32)
    for T, av0 in ((0.3, 0.5), (5.0, 0.5), (5.0, 0.995), (0.05, 0.0101)):
        p = orc.params(N, N, N, T=T)
        ni = orc.ic_random(pt, 2)
        with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None, ang_var=av0)) as eng:
            av = av0
            for s in range(5):
                eng.sweeps(1)
                acc, done = eng.sweep_stats()
                assert done == s + 1
                want = orc.lib().orc_ang_var_update(av, acc, N ** 3)
                assert eng.ang_var() == np.float32(want)
                av = eng.ang_var()


@pytest.mark.parametrize("shape,holes", [((512, 0), False), ((64, 1), False), ((256, 1), True)],
                         ids=["unmasked-512", "unmasked-64", "holes-256"])
def test_ang_var_rule_two_pass(mcl, orc, shape, holes):
    """The rule in the two-pass sweep, sweep by sweep against the oracle's rule, including a host-set value in the
    middle of the run (evolve.cu:430-436): in an unmasked lattice only the first wave of each workgroup reads the
    acceptance counters and hands ang_var to the others through LDS; a lattice with empty sites (masked, not
    row-aligned) uses two counters at 32^3 and every colour-0 wave reads them."""
    N = 32
    pt = np.ones(N ** 3, dtype=np.int32)
    if holes:
        pt[np.random.default_rng(7).random(N ** 3) < 0.25] = 0
    p = orc.params(N, N, N, T=1.1)
    ni = orc.ic_random(pt, 5)
    nvalid = int(pt.sum())
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None, ang_var=0.5)) as eng:
        eng.set_resident(False)
        eng.set_launch(*shape)
        av = 0.5
        for s in range(12):
            if s == 6:
                eng.set_ang_var(0.9)
                av = 0.9
            eng.sweeps(1)
            acc, done = eng.sweep_stats()
            assert done == s + 1 and 0 < acc <= nvalid
            want = orc.lib().orc_ang_var_update(av, acc, nvalid)
            assert eng.ang_var() == np.float32(want), (s, acc)
            av = eng.ang_var()


def test_detailed_balance_independent_spins_gpu(mcl, orc):
    """Independent spins (LL with A = 0) in a dielectric field: the GPU chain samples exp(elecA cos^2 / T)."""
    N = 16
    pt = np.ones(N ** 3, dtype=np.int32)
    p = orc.params(N, N, N, potential="ll", A=0.0, elecA=1.0, elec=(0, 0, 1), T=0.5)
    ni = orc.ic_random(pt, 11)
    edges = np.linspace(-1, 1, 21)
    counts = np.zeros(20)
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None, seed=3)) as eng:
        eng.sweeps(50)
        for _ in range(200):
            eng.sweeps(1)
            counts += np.histogram(eng.directors()[2::3], bins=edges)[0]
    prob = boltzmann_cos_hist(1.0, 0.5, edges)
    tot = counts.sum()
    chi2 = (((counts - tot * prob) ** 2) / (tot * prob)).sum()
    assert chi2 < 200, chi2


@pytest.mark.parametrize("case", FLIP_CASES, ids=lambda c: c[0])
def test_detailed_balance_flip_trial_gpu(mcl, orc, case):
    """The device's flip trial at T > 0 (metropolis_trial's FLIP branch, its own Philox draw per site and the fourth
    uniform): independent spins in a polar field sample exp((elecP c + elecA c^2) / T). The pear case runs the masked
    kernel on isolated sites."""
    p, pt, ni, edges, prob = flip_case(orc, *case)
    counts = np.zeros(20)
    valid = pt == 1
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None, seed=5)) as eng:
        eng.sweeps(50)
        for _ in range(200):
            eng.sweeps(1)
            counts += np.histogram(eng.directors()[2::3][valid], bins=edges)[0]
    assert chi2_of(counts, prob) < 200, chi2_of(counts, prob)


def test_detailed_balance_dimers_gpu(mcl, orc):
    """Isolated LL dimers (pairs of sites separated by empty sites): the relative angle b = n1.n2 must follow
    exp(1.5 A b^2 / T) for each dimer (b uniform a priori). Exercises the bond path and the point-type mask."""
    Nx, Ny, Nz = 3, 24, 24  # sites (0, j, k) and (1, j, k) for even j, k form a dimer; the rest is empty
    nt = Nx * Ny * Nz
    pt = np.zeros(nt, dtype=np.int32)
    for k in range(0, Nz, 2):
        for j in range(0, Ny, 2):
            pt[0 + Nx * (j + Ny * k)] = 1
            pt[1 + Nx * (j + Ny * k)] = 1
    T = 0.6
    p = orc.params(Nx, Ny, Nz, periodic=(0, 0, 0), potential="ll", A=1.0, T=T)
    ni = orc.ic_random(pt, 4)
    edges = np.linspace(-1, 1, 21)
    counts = np.zeros(20)
    a_idx = np.array([0 + Nx * (j + Ny * k) for k in range(0, Nz, 2) for j in range(0, Ny, 2)])
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None, seed=9)) as eng:
        eng.sweeps(100)
        for _ in range(400):
            eng.sweeps(1)
            d = eng.directors().reshape(-1, 3)
            b = np.sum(d[a_idx] * d[a_idx + 1], axis=1)
            counts += np.histogram(b, bins=edges)[0]
    prob = boltzmann_cos_hist(1.5, T, edges)
    tot = counts.sum()
    chi2 = (((counts - tot * prob) ** 2) / (tot * prob)).sum()
    assert chi2 < 200, chi2


def test_large_lattice_properties(mcl, orc):
    """128^3 (the throughput case): ground state E = -9 and S = 1 at T = 0 after a sweep; at T = 1.1 the directors
    stay unit vectors, the acceptance stays near 50 % and E lies between the ground state and the disordered
    value."""
    N = 128
    nt = N ** 3
    ni = np.tile(np.array([0, 0, 1], dtype=np.float32), nt)
    p = orc.params(N, N, N, T=0.0)
    pt = np.ones(nt, dtype=np.int32)
    with mcl.Engine(lattice_from_orc(mcl, p, ni, pt, None)) as eng:
        eng.sweeps(1, 1)
        m = eng.moments()
        assert m["E"] == pytest.approx(-9.0, abs=1e-6) and m["S"] == pytest.approx(1.0, abs=1e-9)
        eng.set_temperature(1.1)
        eng.reset_moments()
        eng.sweeps(200, 10)
        m = eng.moments()
        acc, done = eng.sweep_stats()
        d = eng.directors().reshape(-1, 3)
    assert done == 201 and m["n"] == 20
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
    assert 0.4 < acc / nt < 0.6
    assert -9.0 < m["E"] / m["n"] < -2.0
