"""Pins the CPU oracle (oracle/mcl_oracle.c) before it is trusted as the checker of the GPU path.

Each check is a known answer of the reference's own test suite (tests/run_checks.py), a published test vector, a
closed form or an invariant fixed by the mathematics, chosen so that a plausible restatement mistake (a dropped
term, a wrong sign or index, a transposed operand) fails at least one of them.
"""
import math

import numpy as np
import pytest


def base_run(orc, potential="ll", Nx=8, Ny=8, Nz=8, geometry="bulk", periodic=(1, 1, 1), surfaces=(), **kw):
    """run_checks.py:97-103: periodic bulk at T = 0, uniform along z, MCT = 0, MCS = 1."""
    pt, ns = orc.geometry(geometry, Nx, Ny, Nz)
    p = orc.params(Nx, Ny, Nz, periodic=periodic, potential=potential, surfaces=surfaces, T=0.0, **kw)
    ni = orc.ic_homogeneous(pt, 0.0, 0.0)
    rows = orc.thermal_run(p, ni, pt, ns, 0.0, 0.0, -0.1, 0, 1, 1, 1)
    return rows[0], p, ni, pt, ns


# ---- Random123 known-answer vectors for Philox4x32-10 (kat_vectors of the Random123 distribution) ----
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(orc, ctr, key, want):
    assert tuple(orc.philox(ctr, key)) == want


# ---- ground states of run_checks.py:113-161 (exact energies at T = 0) ----

def test_ground_state_ll(orc):
    r, *_ = base_run(orc)
    T, S, varS, E, varE, _ = r
    assert abs(E + 9.0) <= 1e-4 and abs(S - 1.0) <= 1e-5 and abs(varE) <= 1e-8


def test_ground_state_ghrl_equal_k(orc):
    r, *_ = base_run(orc, potential="ghrl", k=(1, 1, 1, 0))
    assert abs(r[3] + 6.0) <= 1e-4


def test_ground_state_pear(orc):
    r, p, ni, pt, ns = base_run(orc, potential="pear", A=1.5, B1=0.09, C=0.3)
    assert abs(r[3] - 6 * (-1.5 - 0.09)) <= 1e-4
    Q = orc.q_tensor(ni, pt)
    lam, v, _ = orc.eig_sym3(Q)
    assert abs(orc.polarization(ni, pt, v) - 1.0) <= 1e-5


def test_ground_state_slab_homeotropic(orc):
    surf = [("homeotropic", 1.0, (0, 0, 0)), ("homeotropic", 1.0, (0, 0, 0))]
    r, *_ = base_run(orc, Nz=6, geometry="slab", periodic=(1, 1, 0), surfaces=surf)
    assert abs(r[3] - (-1.5 * (256 * 6 + 128 * 5) - 128) / 384) <= 1e-4


def test_ground_state_field(orc):
    r, *_ = base_run(orc, elecA=0.5, elec=(0, 0, 2))
    assert abs(r[3] + 9.5) <= 1e-4


# ---- GHRL coefficients (setGHRL, io.cpp:346-371) and their defining elastic property ----

def test_ghrl_coefficients_config3(orc):
    """K = (1.0, 0.6, 1.4), p0 = 0: Scale = 9/2.2 and the closed forms of io.cpp:351-357."""
    c = orc.ghrl_coeffs(1.0, 0.6, 1.4, 0.0)
    S = 9 / 2.2
    assert c["scale"] == pytest.approx(S, rel=1e-6)
    assert c["lam"] == pytest.approx(S / 9 * (2 - 1.8 + 1.4), rel=1e-6)
    assert c["mu"] == pytest.approx(S * (0.6 - 1.0), rel=1e-6)
    assert c["rho"] == pytest.approx(S / 9 * (1.0 - 1.4), rel=1e-6)
    assert c["nu"] == pytest.approx(-1.0, rel=1e-6)
    assert c["sigma"] == 0.0
    assert (c["lam"], c["mu"], c["rho"]) == pytest.approx((0.727273, -1.636364, -0.181818), abs=1e-6)


def _bond_curvature(orc, p, ni, nj_of_eps, r, h=2e-2):
    """Second derivative of the bond energy at eps = 0 by a central difference (FP32 potential, so h is large)."""
    f = lambda e: orc.pair(p, ni, nj_of_eps(e), r)
    return (f(h) - 2 * f(0.0) + f(-h)) / (h * h)


@pytest.mark.parametrize("K", [(1.0, 0.6, 1.4), (1.48, 1.0, 2.28), (2.0, 1.0, 0.5)])
def test_ghrl_elastic_constants(orc, K):
    """The GHRL potential is built so that one bond along the gradient of a splay, twist or bend deformation costs
    (Scale K_ii / 2) eps^2: splay (r = z, n_j tilted in xz), twist (r = z, n_j rotated in xy), bend (r = x, n_j
    tilted in xz). A dropped or sign-flipped term of lambda, mu, nu or rho breaks the ratios K11 : K22 : K33."""
    p = orc.params(potential="ghrl", k=K + (0.0,))
    scale = orc.ghrl_coeffs(*K, 0.0)["scale"]
    x = (1, 0, 0)
    splay = _bond_curvature(orc, p, x, lambda e: (math.cos(e), 0, math.sin(e)), (0, 0, 1))
    twist = _bond_curvature(orc, p, x, lambda e: (math.cos(e), math.sin(e), 0), (0, 0, 1))
    bend = _bond_curvature(orc, p, x, lambda e: (math.cos(e), 0, math.sin(e)), (1, 0, 0))
    for got, k in ((splay, K[0]), (twist, K[1]), (bend, K[2])):
        assert got == pytest.approx(scale * k, rel=2e-3)


@pytest.mark.parametrize("p0", [10.0, 16.0, -12.0])
def test_ghrl_chiral_twist_angle(orc, p0):
    """With sigma = -Scale k22 2 pi / p0 the twist bond is minimal where |sin(phi)| = 2 pi / |p0| (the lattice pitch
    2 pi / arcsin(2 pi / p0) of README.md:289-290)."""
    p = orc.params(potential="ghrl", k=(1.0, 1.0, 1.0, p0))
    phis = np.linspace(-1.2, 1.2, 24001)
    e = [orc.pair(p, (1, 0, 0), (math.cos(f), math.sin(f), 0), (0, 0, 1)) for f in phis]
    fmin = phis[int(np.argmin(e))]
    assert abs(math.sin(fmin)) == pytest.approx(2 * math.pi / abs(p0), abs=2e-3)


def test_pear_closed_forms(orc):
    """Pear: parallel U = -A - B1, antiparallel U = -A + B1 (the C term vanishes), and the polar-splay slope
    dU/d(eps) = -C at eps = 0 for n_i = z, n_j tilted towards r = x."""
    p = orc.params(potential="pear", A=1.3, B1=0.2, C=0.7)
    z = (0, 0, 1)
    assert orc.pair(p, z, z, (1, 0, 0)) == pytest.approx(-1.3 - 0.2, abs=1e-6)
    assert orc.pair(p, z, (0, 0, -1), (1, 0, 0)) == pytest.approx(-1.3 + 0.2, abs=1e-6)
    h = 1e-3
    f = lambda e: orc.pair(p, z, (math.sin(e), 0, math.cos(e)), (1, 0, 0))
    assert (f(h) - f(-h)) / (2 * h) == pytest.approx(-0.7, rel=1e-2)


def test_ll_closed_form(orc):
    p = orc.params(potential="ll", A=1.7)
    for th in (0.0, 0.3, 1.1, math.pi / 2):
        assert orc.pair(p, (0, 0, 1), (math.sin(th), 0, math.cos(th)), (1, 0, 0)) == pytest.approx(
            -1.5 * 1.7 * math.cos(th) ** 2, abs=1e-5)


def _rand_unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("pot,kw", [("ll", {}), ("ghrl", {"k": (1.48, 1.0, 2.28, 10.0)}),
                                    ("pear", {"A": 1.5, "B1": 0.09, "C": 0.3})])
def test_pair_symmetry_and_rotation_invariance(orc, pot, kw):
    """U(i, j, r) = U(j, i, -r) (so each bond is one energy), and U is invariant under a common rotation of n_i, n_j
    and r (a transposed cross product or a wrong axis breaks it)."""
    p = orc.params(potential=pot, **kw)
    rng = np.random.default_rng(5)
    a, b, r = _rand_unit(rng, 50), _rand_unit(rng, 50), _rand_unit(rng, 50)
    q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    for x, y, rr in zip(a, b, r):
        u = orc.pair(p, x, y, rr)
        assert orc.pair(p, y, x, -rr) == pytest.approx(u, abs=2e-5 * (1 + abs(u)))
        assert orc.pair(p, q @ x, q @ y, q @ rr) == pytest.approx(u, abs=5e-5 * (1 + abs(u)))


# ---- dE of a single-site change = change of the total energy (brute force on small lattices) ----

CASES = [
    dict(potential="ll", geometry="bulk", dims=(4, 4, 4), periodic=(1, 1, 1)),
    dict(potential="ll", geometry="bulk", dims=(5, 3, 4), periodic=(0, 0, 0), elecA=0.7, elec=(1, 2, 2)),
    dict(potential="ghrl", geometry="bulk", dims=(4, 4, 6), periodic=(1, 1, 1), k=(1.48, 1.0, 2.28, 8.0)),
    dict(potential="pear", geometry="bulk", dims=(4, 6, 4), periodic=(1, 1, 1), A=1.5, B1=0.09, C=0.3,
         elecP=0.4, elec=(0, 0, 1)),
    dict(potential="ll", geometry="slab", dims=(4, 4, 5), periodic=(1, 1, 0),
         surfaces=[("rp", 1.5, "axis"), ("fg", 0.7, None)]),
    dict(potential="ghrl", geometry="slab", dims=(4, 4, 5), periodic=(1, 1, 0), k=(1.0, 0.6, 1.4, 0.0),
         surfaces=[("homeotropic_ghrl", 1.5, None), ("rp_ghrl", 1.0, "axis")]),
    dict(potential="ghrl", geometry="slab", dims=(4, 4, 4), periodic=(0, 1, 0), k=(1.0, 0.6, 1.4, 0.0),
         surfaces=[("fg_ghrl", 1.2, None), ("homeotropic", 0.5, None)], elecA=0.3, elec=(1, 0, 0)),
    dict(potential="ll", geometry="sphere", dims=(9, 9, 9), periodic=(0, 0, 0),
         surfaces=[("rp_tangent", 1.0, "axis")]),
    dict(potential="pear", geometry="sphere", dims=(8, 8, 8), periodic=(0, 0, 0), A=1.5, B1=0.09, C=0.3,
         surfaces=[("homeotropic", 1.5, None)]),
]


def build_case(orc, case, seed=0):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = case["dims"]
    pt, ns = orc.geometry(case["geometry"], Nx, Ny, Nz)
    surf = []
    for kind, W, ax in case.get("surfaces", []):
        axis = orc.lab_axis(*rng.uniform((0, 0), (180, 360))) if ax == "axis" else (0, 0, 0)
        surf.append((kind, W, axis))
    kw = {k: v for k, v in case.items() if k not in ("geometry", "dims", "surfaces")}
    p = orc.params(Nx, Ny, Nz, surfaces=surf, **kw)
    ni = _rand_unit(rng, Nx * Ny * Nz).reshape(-1)
    ni[np.repeat(pt == 0, 3)] = np.tile([0, 0, 1], int((pt == 0).sum())).astype(np.float32)
    return p, ni, pt, ns


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (c["potential"], c["geometry"],
                                                                    "".join(map(str, c["periodic"]))))
def test_delta_e_equals_total_energy_change(orc, case):
    p, ni, pt, ns = build_case(orc, case)
    rng = np.random.default_rng(1)
    E0 = orc.energy_total(p, ni, pt, ns)
    valid = np.flatnonzero(pt)
    for idx in rng.choice(valid, size=min(40, len(valid)), replace=False):
        new = _rand_unit(rng, 1)[0]
        old = ni[3 * idx:3 * idx + 3].copy()
        dE_site = orc.site_energy(p, ni, pt, ns, idx, new) - orc.site_energy(p, ni, pt, ns, idx, old)
        ni[3 * idx:3 * idx + 3] = new
        E1 = orc.energy_total(p, ni, pt, ns)
        ni[3 * idx:3 * idx + 3] = old
        scale = abs(orc.site_energy(p, ni, pt, ns, idx, old)) + abs(orc.site_energy(p, ni, pt, ns, idx, new)) + 1
        assert dE_site == pytest.approx(E1 - E0, abs=1e-5 * scale)


def test_energy_mean_counts_bonds_twice(orc):
    """E in po.dat counts every bond twice and anchoring / field once (README.md:423-425)."""
    case = dict(potential="ll", geometry="slab", dims=(4, 4, 5), periodic=(1, 1, 0),
                surfaces=[("homeotropic", 1.1, None), ("fg", 0.4, None)], elecA=0.3, elec=(1, 0, 0))
    p, ni, pt, ns = build_case(orc, case, seed=3)
    nv = int((pt > 0).sum())
    p0 = orc.params(4, 4, 5, periodic=(1, 1, 0), potential="ll", surfaces=[("homeotropic", 0.0, (0, 0, 0)),
                                                                           ("fg", 0.0, (0, 0, 0))])
    bonds = orc.energy_total(p0, ni, pt, ns)  # bonds only, once
    onsite = orc.energy_total(p, ni, pt, ns) - bonds
    assert orc.energy_mean(p, ni, pt, ns) * nv == pytest.approx(2 * bonds + onsite, rel=1e-5)


# ---- order parameter, eigen-solve and radial observables ----

def test_eig_matches_numpy(orc):
    rng = np.random.default_rng(2)
    for _ in range(200):
        a = rng.normal(size=(3, 3))
        a = a + a.T
        if rng.random() < 0.3:  # near-degenerate cases
            w, v = np.linalg.eigh(a)
            w[1] = w[2] - 1e-9
            a = (v * w) @ v.T
        lam, vec, ev = orc.eig_sym3(a)
        w, v = np.linalg.eigh(a)
        assert lam == pytest.approx(w[-1], abs=1e-10 * (1 + abs(w).max()))
        assert np.allclose(ev, w[::-1], atol=1e-10 * (1 + abs(w).max()))
        if w[-1] - w[-2] > 1e-6:
            assert abs(abs(vec @ v[:, -1]) - 1) < 1e-8


def test_q_tensor_uniform_and_isotropic(orc):
    pt = np.ones(16 ** 3, dtype=np.int32)
    for th, ph in ((0, 0), (90, 0), (37, 112)):
        ni = orc.ic_homogeneous(pt, th, ph)
        Q = orc.q_tensor(ni, pt)
        assert abs(np.trace(Q)) < 1e-6
        assert orc.eig_sym3(Q)[0] == pytest.approx(1.0, abs=1e-6)
    ni = orc.ic_random(pt, 7)
    S = orc.eig_sym3(orc.q_tensor(ni, pt))[0]
    assert 0 < S < 5 / math.sqrt(16 ** 3)


def test_q_tensor_matches_numpy(orc):
    rng = np.random.default_rng(9)
    pt = (rng.random(6 ** 3) < 0.8).astype(np.int32)
    ni = _rand_unit(rng, 6 ** 3).reshape(-1)
    n = ni.reshape(-1, 3)[pt > 0].astype(np.float64)
    want = (3 * n.T @ n / len(n) - np.eye(3)) / 2
    Q = orc.q_tensor(ni, pt)
    assert np.allclose(Q[:2, :], want[:2, :], atol=1e-12)
    assert Q[2, 2] == pytest.approx(-Q[0, 0] - Q[1, 1], abs=1e-15)


def hedgehog(orc, n, sign=1):
    c = n // 2
    pt, ns = orc.geometry("sphere", n, n, n)
    ni = np.zeros((n, n, n, 3), dtype=np.float32)  # [k, j, i]
    for k in range(n):
        for j in range(n):
            for i in range(n):
                v = np.array([i - c, j - c, k - c], dtype=np.float64)
                m = np.linalg.norm(v)
                ni[k, j, i] = (0, 0, 1) if m == 0 else sign * v / m
    return pt, ns, ni.reshape(-1)


def test_radial_and_pear_observables_hedgehog(orc):
    """README.md:459-460 reference values for a 21^3 droplet: outward hedgehog Sr_sph = 1, Pr_sph = 1,
    divn = 0.275; inward divn = -0.275 (tolerances of run_checks.py:191-193)."""
    n = 21
    p = orc.params(n, n, n, periodic=(0, 0, 0), potential="pear")
    pt, ns, ni = hedgehog(orc, n)
    Sr, Pr = orc.radial_order(p, ni, pt)
    assert Sr == pytest.approx(1.0, abs=0.03) and Pr == pytest.approx(1.0, abs=0.03)
    assert orc.splay_divergence(p, ni, pt) == pytest.approx(0.275, abs=0.02)
    pt, ns, ni = hedgehog(orc, n, -1)
    assert orc.splay_divergence(p, ni, pt) == pytest.approx(-0.275, abs=0.02)
    assert orc.radial_order(p, ni, pt)[1] == pytest.approx(-1.0, abs=0.03)


def test_radial_uniform_z(orc):
    """Uniform along z: Sr_cyl = -1/2 exactly (n is perpendicular to every in-plane radius)."""
    n = 21
    pt, ns = orc.geometry("sphere", n, n, n)
    p = orc.params(n, n, n, periodic=(0, 0, 0))
    ni = orc.ic_homogeneous(pt, 0, 0)
    Sr_c, Pr_c = orc.radial_order(p, ni, pt, cylinder=True)
    assert Sr_c == pytest.approx(-0.5, abs=1e-6) and Pr_c == pytest.approx(0.0, abs=1e-6)
    Sr_s, Pr_s = orc.radial_order(p, ni, pt)
    assert abs(Pr_s) < 1e-6 and abs(Sr_s) < 0.05


def test_splay_divergence_periodic_zero(orc):
    """divn telescopes to exactly 0 in a fully periodic box (README.md:442)."""
    pt = np.ones(6 ** 3, dtype=np.int32)
    p = orc.params(6, 6, 6)
    ni = orc.ic_random(pt, 3)
    assert abs(orc.splay_divergence(p, ni, pt)) < 1e-6


def test_local_order(orc):
    pt = np.ones(4 ** 3, dtype=np.int32)
    p = orc.params(4, 4, 4)
    ni = orc.ic_homogeneous(pt, 20, 30)
    assert orc.local_order(p, ni, pt, 1, 2, 3) == pytest.approx(1.0, abs=1e-6)
    ni = orc.ic_random(pt, 4)
    i, j, k = 1, 2, 3
    pts = [(i, j, k), (2, 2, 3), (0, 2, 3), (1, 3, 3), (1, 1, 3), (1, 2, 0), (1, 2, 2)]
    n = np.array([ni[3 * (a + 4 * (b + 4 * c)):3 * (a + 4 * (b + 4 * c)) + 3] for a, b, c in pts], dtype=np.float64)
    Q = (3 * n.T @ n / len(n) - np.eye(3)) / 2
    assert orc.local_order(p, ni, pt, i, j, k) == pytest.approx(np.linalg.eigvalsh(Q)[-1], abs=1e-6)


# ---- proposal, ang_var rule and the sweep schedule ----

def test_rotation_closed_form(orc):
    out = (np.ctypeslib.ctypes.c_float * 3)()
    for rot, axis in ((0.1, 0), (0.5, 1), (0.9, 2)):
        for va in (0.3, -0.7):
            n = np.array([0.48, -0.6, 0.64], dtype=np.float32)
            orc.lib().orc_rotate(orc.vec(n), va, rot, out)
            th = math.pi * va
            c, s = math.cos(th), math.sin(th)
            R = {0: [[1, 0, 0], [0, c, s], [0, -s, c]], 1: [[c, 0, -s], [0, 1, 0], [s, 0, c]],
                 2: [[c, s, 0], [-s, c, 0], [0, 0, 1]]}[axis]
            assert np.allclose(np.array(out), np.array(R) @ n, atol=1e-6)


def test_ang_var_rule(orc):
    f = orc.lib().orc_ang_var_update
    assert f(0.5, 40, 100) == pytest.approx(0.5 * 0.99, rel=1e-6)
    assert f(0.5, 60, 100) == pytest.approx(0.5 / 0.99, rel=1e-6)
    assert f(0.5, 50, 100) == pytest.approx(0.5)
    assert f(0.0100001, 0, 100) == pytest.approx(0.01, rel=1e-6)
    assert f(0.999, 100, 100) == pytest.approx(0.999 / 0.99 - 0.5, rel=1e-5)


def test_sweep_schedule_names(orc):
    """run_checks.py:213-261: temperature and field points and snapshot names."""
    n = orc.sweep_points(1.0, 0.5, np.float32(-0.1))
    T = [orc.sweep_value(1.0, 0.5, float(np.float32(-0.1)), i, n) for i in range(n)]
    assert n == 6 and np.allclose(T, [1.0, 0.9, 0.8, 0.7, 0.6, 0.5], atol=1e-6)
    S = orc.snapshot_scale(float(np.float32(-0.1)))
    assert [orc.snapshot_index(t, S) for t in T] == [100, 90, 80, 70, 60, 50]
    d = float(np.float32(-0.004))
    a, b = float(np.float32(1.24)), float(np.float32(1.224))
    n = orc.sweep_points(a, b, d)
    S = orc.snapshot_scale(d)
    assert n == 5 and [orc.snapshot_index(orc.sweep_value(a, b, d, i, n), S) for i in range(n)] == [
        1240, 1236, 1232, 1228, 1224]
    d = float(np.float32(-0.01))
    a, b = float(np.float32(1.2)), float(np.float32(1.1))
    n = orc.sweep_points(a, b, d)
    S = orc.snapshot_scale(d)
    assert S == 100 and sorted(orc.snapshot_index(orc.sweep_value(a, b, d, i, n), S) for i in range(n)) == list(
        range(110, 121))
    d = float(np.float32(0.1))
    n = orc.sweep_points(0.0, float(np.float32(1.2)), d)
    assert n == 13 and abs(orc.sweep_value(0.0, float(np.float32(1.2)), d, n - 1, n) - 1.2) < 1e-6


# ---- the Metropolis chain samples the Boltzmann distribution (textbook result, no reference needed) ----

def boltzmann_cos_hist(elecA, T, edges):
    """Probabilities of cos(theta) bins for E = -elecA cos^2(theta) on the sphere (cos uniform a priori)."""
    x = np.linspace(-1, 1, 200001)
    w = np.exp(elecA * x * x / T)
    cdf = np.concatenate([[0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(x))])
    cdf /= cdf[-1]
    return np.diff(np.interp(edges, x, cdf))


def test_detailed_balance_independent_spins(orc):
    """LL with A = 0 and a dielectric field: independent spins; the sampled cos(theta) histogram must follow
    exp(elecA cos^2 / T) (chi-square, 20 bins)."""
    N = 16
    pt = np.ones(N ** 3, dtype=np.int32)
    p = orc.params(N, N, N, potential="ll", A=0.0, elecA=1.0, elec=(0, 0, 1), T=0.5)
    ni = orc.ic_random(pt, 11)
    rngs = orc.make_rngs(2, 3)
    av = 0.5
    for _ in range(50):
        _, av = orc.sweep(p, ni, pt, None, av, rngs, 2)
    edges = np.linspace(-1, 1, 21)
    counts = np.zeros(20)
    for _ in range(200):
        _, av = orc.sweep(p, ni, pt, None, av, rngs, 2)
        counts += np.histogram(ni[2::3], bins=edges)[0]
    # samples of successive sweeps are correlated: use the effective number of independent sweeps
    prob = boltzmann_cos_hist(1.0, 0.5, edges)
    tot = counts.sum()
    chi2 = (((counts - tot * prob) ** 2) / (tot * prob)).sum()
    # with correlated samples chi2 is inflated by ~ the autocorrelation time (~2-4 sweeps here); a biased
    # acceptance rule (e.g. exp(+dE/T) or a missing factor) gives chi2 in the 1e4-1e5 range
    assert chi2 < 20 * 10, chi2


def boltzmann_hist(energy, T, edges):
    """Probabilities of cos(theta) bins for a site energy energy(cos theta) on the sphere (cos uniform a priori)."""
    x = np.linspace(-1, 1, 200001)
    w = np.exp(-(energy(x) - energy(x).min()) / T)
    cdf = np.concatenate([[0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(x))])
    cdf /= cdf[-1]
    return np.diff(np.interp(edges, x, cdf))


# Independent spins with the flip trial on (flipMove: pear, or elecP != 0; main.cpp:88): (potential, isolated sites,
# elecP, elecA, T). The polar field makes n and -n differ, so the n -> -n trial (evolve.cpp:197-213, evolve.cu:225-239)
# moves weight between the hemispheres; the sampled cos(theta) must follow exp((elecP c + elecA c^2) / T). The pear
# case empties every odd site, so no site has a neighbour (a masked lattice of isolated spins).
FLIP_CASES = [("ll", False, 1.0, 0.0, 0.5), ("pear", True, 0.6, 0.8, 0.4)]


def flip_case(orc, potential, isolated, elecP, elecA, T, N=16):
    idx = np.arange(N ** 3)
    i, j, k = idx % N, (idx // N) % N, idx // (N * N)
    pt = ((i + j + k) % 2 == 0).astype(np.int32) if isolated else np.ones(N ** 3, dtype=np.int32)
    p = orc.params(N, N, N, potential=potential, A=0.0 if potential == "ll" else 1.5, B1=0.09, C=0.3, elecP=elecP,
                   elecA=elecA, elec=(0, 0, 1), T=T)
    assert p.flipMove == 1
    ni = orc.ic_random(pt, 11)
    edges = np.linspace(-1, 1, 21)
    prob = boltzmann_hist(lambda c: -elecP * c - elecA * c * c, T, edges)
    return p, pt, ni, edges, prob


def chi2_of(counts, prob):
    tot = counts.sum()
    return (((counts - tot * prob) ** 2) / (tot * prob)).sum()


@pytest.mark.parametrize("case", FLIP_CASES, ids=lambda c: c[0])
def test_detailed_balance_flip_trial(orc, case):
    """The oracle's flip trial at T > 0 (chi-square over 20 bins, as above)."""
    p, pt, ni, edges, prob = flip_case(orc, *case)
    rngs = orc.make_rngs(2, 7)
    av = 0.5
    for _ in range(50):
        _, av = orc.sweep(p, ni, pt, None, av, rngs, 2)
    counts = np.zeros(20)
    valid = pt == 1
    for _ in range(200):
        _, av = orc.sweep(p, ni, pt, None, av, rngs, 2)
        counts += np.histogram(ni[2::3][valid], bins=edges)[0]
    assert chi2_of(counts, prob) < 200, chi2_of(counts, prob)


# ---- electric mode loop (evolve_electric.cpp:31-103) ----

def test_electric_run_zero_field_point_equals_thermal_point(orc):
    """At field 0 the first point of a field sweep is the thermal point at the same T: same RNG seeding, ang_var
    start and MCT/MCS schedule, and the field term vanishes, so the rows agree bit for bit."""
    surf = [("homeotropic", 1.5, (0, 0, 0)), ("homeotropic", 1.5, (0, 0, 0))]
    pt, ns = orc.geometry("slab", 8, 8, 6)
    kw = dict(periodic=(1, 1, 0), potential="ghrl", k=(1.0, 0.6, 1.4, 0.0), elec=(1, 0, 0), elecA=0.25, elecE=0.0,
              T=0.5, surfaces=surf)
    p1, p2 = orc.params(8, 8, 6, **kw), orc.params(8, 8, 6, **kw)
    n1, n2 = orc.ic_random(pt, 3), orc.ic_random(pt, 3)
    th = orc.thermal_run(p1, n1, pt, ns, 0.5, 0.5, -0.1, 20, 10, 3, 1)
    el = orc.electric_run(p2, n2, pt, ns, 0.0, 0.4, 0.2, 20, 10, 3, 1)
    assert el.shape == (3, 6) and list(el[:, 0]) == [0.0, pytest.approx(0.2), pytest.approx(0.4)]
    assert np.array_equal(th[0, 1:5], el[0, 1:5])
    assert np.isnan(el[0, 5])


def test_electric_run_aligns_with_a_strong_field(orc):
    """LL bulk, dielectric coupling 1 along x: at field 2 the field energy -(2 n_x)^2 dominates T = 0.5, the directors
    align along x and the mean site energy drops by about 4 below the zero-field value."""
    pt, ns = orc.geometry("bulk", 6, 6, 6)
    p = orc.params(6, 6, 6, potential="ll", elec=(1, 0, 0), elecA=1.0, elecE=0.0, T=0.5)
    ni = orc.ic_random(pt, 4)
    rows = orc.electric_run(p, ni, pt, ns, 0.0, 2.0, 2.0, 300, 100, 4, 1)
    assert rows.shape[0] == 2
    assert rows[1, 1] > 0.9
    assert rows[0, 3] - rows[1, 3] > 3.0
    lam, v, _ = orc.eig_sym3(orc.q_tensor(ni, pt))
    assert abs(v[0]) > 0.99


def test_thermal_loop_reproduces_the_c1_fixture_prefix(orc):
    """The oracle's thermal loop is deterministic for a fixed seed and thread count: the first temperatures of the
    committed 16^3 fixture (replica 1, one thread) come out bit for bit, so a change to the loop code that alters the
    trajectory is caught here."""
    import os
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_oracle.npz"))["rows"][0]
    N = 16
    pt, _ = orc.geometry("bulk", N, N, N)
    p = orc.params(N, N, N, periodic=(1, 1, 1), potential="ll", A=1.0)
    ni = orc.ic_random(pt, 1)
    f32 = lambda x: float(np.float32(x))
    rows = orc.thermal_run(p, ni, pt, None, f32(1.2), f32(1.1), f32(-0.02), 500, 200, 1, 1)
    assert np.array_equal(rows[:5], ref[:5])


# ---- anchoring energies in closed form (anchoring_homeotropic.cpp, anchoring_rp.cpp, anchoring_rp_tangent.cpp,
# anchoring_fg.cpp): each kind is evaluated at orientations where the value follows from the definition by hand ----

def _surf_params(orc, kind, W, axis=(0.0, 0.0, 0.0), k=None):
    kw = dict(potential="ghrl", k=k) if k is not None else {}
    return orc.params(4, 4, 4, surfaces=[(kind, W, axis)], **kw)


X, Y, Z = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
R2 = 1 / math.sqrt(2)


@pytest.mark.parametrize("W", [1.0, 2.5])
def test_anchoring_quadratic_kinds(orc, W):
    """homeotropic -W (n.s)^2; fg +W (n.s)^2; rp -W (n.e)^2 with e the lab axis, whatever the normal."""
    p = _surf_params(orc, "homeotropic", W)
    assert orc.surface(p, 0, Z, Z) == pytest.approx(-W, abs=1e-6)
    assert orc.surface(p, 0, (0, 0, -1.0), Z) == pytest.approx(-W, abs=1e-6)
    assert orc.surface(p, 0, X, Z) == pytest.approx(0.0, abs=1e-7)
    assert orc.surface(p, 0, (R2, 0, R2), Z) == pytest.approx(-W / 2, abs=1e-6)
    p = _surf_params(orc, "fg", W)
    assert orc.surface(p, 0, Z, Z) == pytest.approx(+W, abs=1e-6)
    assert orc.surface(p, 0, Y, Z) == pytest.approx(0.0, abs=1e-7)
    assert orc.surface(p, 0, (0, R2, R2), Z) == pytest.approx(+W / 2, abs=1e-6)
    axis = orc.lab_axis(90.0, 0.0)  # x
    p = _surf_params(orc, "rp", W, axis)
    for s in (Z, X, (0.6, 0.0, 0.8)):  # the normal plays no part in rp
        assert orc.surface(p, 0, X, s) == pytest.approx(-W, abs=1e-6)
        assert orc.surface(p, 0, Y, s) == pytest.approx(0.0, abs=1e-6)
        assert orc.surface(p, 0, Z, s) == pytest.approx(0.0, abs=1e-6)
        assert orc.surface(p, 0, (R2, R2, 0), s) == pytest.approx(-W / 2, abs=1e-6)


def test_rp_tangent_easy_axis(orc):
    """The easy axis is the lab axis d projected on the plane normal to s and normalized; none when d is parallel
    (or antiparallel) to s. On the flat surface with d in-plane it is d itself."""
    e, ok = orc.rp_tangent_axis(orc.lab_axis(45.0, 0.0), Z)
    assert ok == 1 and e == pytest.approx(X, abs=1e-6)
    e, ok = orc.rp_tangent_axis(orc.lab_axis(90.0, 30.0), Z)
    assert ok == 1 and e == pytest.approx((math.cos(math.pi / 6), 0.5, 0.0), abs=1e-6)
    e, ok = orc.rp_tangent_axis(Z, (0.6, 0.0, 0.8))  # z projected on the plane normal to (0.6, 0, 0.8)
    assert ok == 1 and e == pytest.approx((-0.8, 0.0, 0.6), abs=1e-6)
    for s in (Z, (0, 0, -1.0)):
        e, ok = orc.rp_tangent_axis(Z, s)
        assert ok == 0 and e == (0.0, 0.0, 0.0)
    rng = np.random.default_rng(5)
    for _ in range(50):
        d, s = rng.normal(size=3), rng.normal(size=3)
        d, s = d / np.linalg.norm(d), s / np.linalg.norm(s)
        e, ok = orc.rp_tangent_axis(d, s)
        assert ok == 1
        e = np.array(e)
        assert abs(np.linalg.norm(e) - 1) < 1e-6 and abs(e @ s) < 1e-6 and e @ d > 0
        assert np.linalg.norm(np.cross(e, d - (d @ s) * s)) < 1e-5


@pytest.mark.parametrize("W", [1.0, 3.0])
def test_rp_tangent_energy(orc, W):
    p = _surf_params(orc, "rp_tangent", W, orc.lab_axis(45.0, 0.0))
    assert orc.surface(p, 0, X, Z) == pytest.approx(-W, abs=1e-6)
    assert orc.surface(p, 0, Y, Z) == pytest.approx(0.0, abs=1e-6)
    assert orc.surface(p, 0, Z, Z) == pytest.approx(0.0, abs=1e-6)
    assert orc.surface(p, 0, (R2, R2, 0), Z) == pytest.approx(-W / 2, abs=1e-6)
    # on a tilted site the axis follows the surface: d = (1, 0, 1)/sqrt 2 on s = (0.6, 0, 0.8) gives e = (0.8, 0, -0.6)
    assert orc.surface(p, 0, (0.8, 0.0, -0.6), (0.6, 0.0, 0.8)) == pytest.approx(-W, abs=1e-5)
    assert orc.surface(p, 0, (0.6, 0.0, 0.8), (0.6, 0.0, 0.8)) == pytest.approx(0.0, abs=1e-5)
    assert orc.surface(p, 0, Y, (0.6, 0.0, 0.8)) == pytest.approx(0.0, abs=1e-6)
    p = _surf_params(orc, "rp_tangent", W, Z)
    assert orc.surface(p, 0, (-0.8, 0.0, 0.6), (0.6, 0.0, 0.8)) == pytest.approx(-W, abs=1e-6)
    for n in (X, Y, Z, (R2, 0, R2)):  # axis parallel to the normal: no anchoring
        assert orc.surface(p, 0, n, Z) == 0.0


def _ghrl_k_forms(k11, k22, k33, p0=0.0):
    """The GHRL bond U(n, m, r) at the orientations used below, written with the Frank constants only. With
    L = 9 / |k11 - 3 k22 - k33| and l = L / 9 (io.cpp:346-357; nu = -1 when k11 < 3 k22 + k33):
      parallel    m = n, n perpendicular to r (or m = n = r) -> -2 l (k11 - k22) - 1
      end-on      m = r, n perpendicular to r           ->  l (1.75 k11 - 2.5 k22 + 0.75 k33) + 1/2
      twisted     m, n perpendicular to r and each other -> l (-0.5 k11 + 2 k22 - 1.5 k33) + 1/2 - sigma
      chiral part of U(n, m, r) - U(n', m, r) for n, n' = m rotated by +-phi about r: 2 sigma sin(phi)
    with sigma = -L k22 2 pi / p0. The end-on form is also U(n, r, r) with n parallel to r (the bond is symmetric
    under exchanging its ends)."""
    L = 9.0 / abs(k11 - 3 * k22 - k33)
    l = L / 9
    sigma = -L * k22 * 2 * math.pi / p0 if p0 else 0.0
    return dict(parallel=-2 * l * (k11 - k22) - 1, end_on=l * (1.75 * k11 - 2.5 * k22 + 0.75 * k33) + 0.5,
                twisted=l * (-0.5 * k11 + 2 * k22 - 1.5 * k33) + 0.5 - sigma, sigma=sigma)


GHRL_KS = [(1.0, 1.0, 1.0, 0.0), (1.0, 0.6, 1.4, 0.0), (1.48, 1.0, 2.28, 0.0), (1.0, 0.6, 1.4, 12.0)]


@pytest.mark.parametrize("k", GHRL_KS)
def test_homeotropic_ghrl_closed_form(orc, k):
    """homeotropic_ghrl: W U(n, s, s), a phantom neighbour along the normal pointing along it."""
    f, W = _ghrl_k_forms(*k), 1.7
    p = _surf_params(orc, "homeotropic_ghrl", W, k=k)
    for s in (Z, (0, 0, -1.0), (0.6, 0.0, 0.8)):
        s = np.array(s)
        t = np.cross(s, Y if abs(s[1]) < 0.9 else X)
        t /= np.linalg.norm(t)
        # n along s: the parallel form (b = 1, a_i = a_j = 1); n in the plane: the end-on form
        assert orc.surface(p, 0, s, s) == pytest.approx(W * f["parallel"], abs=2e-5)
        assert orc.surface(p, 0, t, s) == pytest.approx(W * f["end_on"], abs=2e-5)


@pytest.mark.parametrize("k", GHRL_KS)
def test_rp_ghrl_closed_form(orc, k):
    """rp_ghrl: W U(n, e, s), a phantom neighbour along the normal s with the lab easy axis e (here x, in the
    plane of a z surface)."""
    f, W = _ghrl_k_forms(*k), 0.8
    p = _surf_params(orc, "rp_ghrl", W, orc.lab_axis(90.0, 0.0), k=k)
    assert orc.surface(p, 0, X, Z) == pytest.approx(W * f["parallel"], abs=2e-5)
    assert orc.surface(p, 0, Z, Z) == pytest.approx(W * f["end_on"], abs=2e-5)
    assert orc.surface(p, 0, Y, Z) == pytest.approx(W * f["twisted"], abs=2e-5)
    phi = 0.3
    up = orc.surface(p, 0, (math.cos(phi), math.sin(phi), 0), Z)
    down = orc.surface(p, 0, (math.cos(phi), -math.sin(phi), 0), Z)
    assert up - down == pytest.approx(W * 2 * f["sigma"] * math.sin(phi), abs=5e-5)


@pytest.mark.parametrize("k", GHRL_KS)
def test_fg_ghrl_closed_form(orc, k):
    """fg_ghrl: W U(n, t, s) with t the normalized projection of n on the surface plane; zero when n is along the
    normal. It is degenerate in the plane: only the tilt of n out of the plane matters."""
    f, W = _ghrl_k_forms(*k), 1.3
    p = _surf_params(orc, "fg_ghrl", W, k=k)
    for n in (X, Y, (R2, -R2, 0)):
        assert orc.surface(p, 0, n, Z) == pytest.approx(W * f["parallel"], abs=2e-5)
    assert orc.surface(p, 0, Z, Z) == 0.0
    assert orc.surface(p, 0, (0, 0, -1.0), Z) == 0.0
    L = 9.0 / abs(k[0] - 3 * k[1] - k[2])
    lam, mu, rho, nu = (L / 9 * (2 * k[0] - 3 * k[1] + k[2]), L * (k[1] - k[0]), L / 9 * (k[0] - k[2]), -1.0)
    for theta in (0.2, 0.7, 1.2):  # n tilted by theta out of the plane: a_i = sin, a_j = 0, b = cos, c = 0
        c, s_ = math.cos(theta), math.sin(theta)
        P2 = 1.5 * c * c - 0.5
        want = (1.5 * s_ * s_ - 1) * (rho * P2 + lam) - mu / 9 + nu * P2
        for az in (0.0, 1.0, 2.5):
            n = (c * math.cos(az), c * math.sin(az), s_)
            assert orc.surface(p, 0, n, Z) == pytest.approx(W * want, abs=3e-5)


# ---- Splay_magnitude (parameter_order.cpp:147-176) on constructed fields ----

def _xz_rotation(Nx, Ny, Nz, q):
    """n(x) = (sin qx, 0, cos qx): the director turns in the xz plane along x."""
    x = np.arange(Nx)
    n = np.zeros((Nz, Ny, Nx, 3), dtype=np.float32)
    n[..., 0] = np.sin(q * x)
    n[..., 2] = np.cos(q * x)
    return n.reshape(-1)


def test_splay_magnitude_closed_forms(orc):
    """Only x bonds carry splay here. Summed over a ring of Nx sites the vector sum is Nt sin q cos q along z over
    6 Nt directed bonds, so <dtheta> = 1.5 |sin q cos q| / 6 = |sin 2q| / 8. On a free x axis the end terms remain:
    s_x = NyNz (sin^2 q(Nx-1) - 0), s_z = NyNz (sin(2q(Nx-1)) / 2 + (Nx-1) sin q cos q), over 6 Nt - 2 NyNz bonds."""
    for Nx, m in ((12, 1), (16, 1), (16, 3), (10, 2)):
        q = 2 * math.pi * m / Nx
        p = orc.params(Nx, 4, 6, periodic=(1, 1, 1), potential="pear")
        pt = np.ones(Nx * 24, dtype=np.int32)
        got = orc.splay_magnitude(p, _xz_rotation(Nx, 4, 6, q), pt)
        assert got == pytest.approx(abs(math.sin(2 * q)) / 8, rel=1e-5, abs=1e-7)
    for Nx, q in ((9, 0.3), (6, 1.1)):
        Ny, Nz = 4, 4
        p = orc.params(Nx, Ny, Nz, periodic=(0, 1, 1), potential="pear")
        pt = np.ones(Nx * Ny * Nz, dtype=np.int32)
        got = orc.splay_magnitude(p, _xz_rotation(Nx, Ny, Nz, q), pt)
        a = q * (Nx - 1)
        sx = Ny * Nz * math.sin(a) ** 2
        sz = Ny * Nz * (math.sin(2 * a) / 2 + (Nx - 1) * math.sin(q) * math.cos(q))
        want = 1.5 * math.hypot(sx, sz) / (6 * Nx * Ny * Nz - 2 * Ny * Nz)
        assert got == pytest.approx(want, rel=1e-5)


def test_splay_magnitude_uniform_and_empty(orc):
    for periodic in ((1, 1, 1), (0, 0, 0)):
        p = orc.params(6, 6, 6, periodic=periodic, potential="pear")
        pt = np.ones(216, dtype=np.int32)
        n = np.tile(np.array([0.36, 0.48, 0.8], dtype=np.float32), 216)
        assert orc.splay_magnitude(p, n, pt) == pytest.approx(0.0, abs=1e-7)
    p = orc.params(6, 6, 6, potential="pear")
    assert orc.splay_magnitude(p, np.zeros(3 * 216, dtype=np.float32), np.zeros(216, dtype=np.int32)) == 0.0
