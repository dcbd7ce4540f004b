"""Writes tests/golden/scenarios_oracle.npz: the oracle's runs of the scenarios of tests/scenarios.py (the reference's
seven regression inputs and an LL sphere of the batched configuration), 16 replicas each (seeds 1..16). With --only,
the named scenario is (re)computed and the others already in the file are kept.

The mode loops follow the reference's CPU modes: thermal (evolve_thermal.cpp:61-100: at each T of sweep_points, MCT
sweeps, then MCS sweeps each measured), step (evolve_step.cpp: fn blocks at Ti), quench (evolve_quench.cpp: per
block MCT sweeps at Ti, MCQ at Tf, then MCS measured at Tf) and electric (evolve_electric.cpp:31-103: the field
magnitude swept at T = Ti). ang_var starts at 0.5 and carries over from point to point. Each measurement is the
oracle's energy mean, Q tensor and largest eigenvalue, and the extras of ExtraObservables
(parameter_order.cpp:249-295): P, dtheta and divn for the pear model, the radial order about the grid centre. Each
po.dat-like row holds the point's label, S, varS, E, varE and the extras' means, in the column order of po.dat.
Replica r uses seed r for the initial directors (oracle generator) and the sweep RNG (one thread).

    python tests/golden/make_scenario_fixtures.py [--reps 16] [--only NAME] [--out tests/golden/scenarios_oracle.npz]
"""
import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, TESTS)

import scenarios as sc  # noqa: E402

f32 = lambda x: float(np.float32(x))  # input values are floats in the reference (parameters.h:44-46)


def build(orc, name, seed):
    """(params, pt, ns, ni) of the scenario for the oracle."""
    v = sc.values(name)
    Nx, Ny, Nz = sc.SCENARIOS[name]["dims"]
    geom = v.get("geometry", "bulk")
    per = [1 if v.get(a + "bound", "free") == "periodic" else 0 for a in "xyz"]
    if geom == "slab":
        per[2] = 0
    if geom == "sphere":
        per = [0, 0, 0]
    if geom == "custom":
        text = sc.files(name)["cyl.csv"]
        t = sc.parse_table(text, ["x", "y", "z", "nx", "ny", "nz", "pt"])
        nt = Nx * Ny * Nz
        pt = np.ones(nt, dtype=np.int32)
        ns = np.zeros(3 * nt, dtype=np.float32)
        idx = (t[:, 0] + Nx * (t[:, 1] + Ny * t[:, 2])).astype(int)
        pt[idx] = t[:, 6].astype(np.int32)
        ns.reshape(-1, 3)[idx] = t[:, 3:6]
    else:
        pt, ns = orc.geometry(geom, Nx, Ny, Nz)
    surfaces = []
    for s in sorted(v["surfaces"]):
        e = v["surfaces"][s]
        kind = e["anchoring_type"]
        axis = (0.0, 0.0, 0.0)
        if kind in ("rp", "rp_tangent", "rp_ghrl"):
            axis = orc.lab_axis(f32(e["theta_s"]), f32(e["phi_s"]))
        surfaces.append((kind, f32(e["W"]), axis))
    pot = v.get("potential", "ll")
    kw = dict(potential=pot, A=f32(v.get("a", 1.0)), B1=f32(v.get("b1", 0.04)), C=f32(v.get("c", 0.3)),
              surfaces=surfaces, elecA=f32(v.get("eleca", 0)), elecP=f32(v.get("elecp", 0)),
              elec=(f32(v.get("elecx", 0)), f32(v.get("elecy", 0)), f32(v.get("elecz", 0))), T=f32(v.get("ti", 1.2)))
    if pot == "ghrl":
        kw["k"] = (f32(v.get("k11", 1)), f32(v.get("k22", 1)), f32(v.get("k33", 1)), f32(v.get("p0", 0)))
    p = orc.params(Nx, Ny, Nz, periodic=tuple(per), **kw)
    if v.get("ic", "random") == "ic_file":
        t = sc.parse_table(sc.files(name)[v["ic_file"]], ["x", "y", "z", "nx", "ny", "nz"])
        ni = np.zeros(3 * Nx * Ny * Nz, dtype=np.float32)
        idx = (t[:, 0] + Nx * (t[:, 1] + Ny * t[:, 2])).astype(int)
        n = t[:, 3:6].copy()
        norm = np.linalg.norm(n, axis=1)
        off = np.abs(norm - 1) > 1e-3  # ic.cpp's rule: only clearly non-unit directors are rescaled
        n[off] /= norm[off, None]
        ni.reshape(-1, 3)[idx] = n
        ni.reshape(-1, 3)[pt == 0] = (0, 0, 1)
    else:
        ni = orc.ic_random(pt, seed)
    return p, pt, ns, ni


def measure(orc, p, ni, pt, ns, pear, radial):
    E = orc.energy_mean(p, ni, pt, ns)
    S, director, _ = orc.eig_sym3(orc.q_tensor(ni, pt))
    extras = []
    if pear:
        extras += [orc.polarization(ni, pt, director), orc.splay_magnitude(p, ni, pt), orc.splay_divergence(p, ni, pt)]
    if radial in ("sphere", "both"):
        sr, pr = orc.radial_order(p, ni, pt, False)
        extras += [sr, pr] if pear else [sr]
    if radial in ("cylinder", "both"):
        sr, pr = orc.radial_order(p, ni, pt, True)
        extras += [sr, pr] if pear else [sr]
    return S, E, extras


def one(job):
    name, seed = job
    os.environ["OMP_NUM_THREADS"] = "1"
    import pyoracle as orc
    v = sc.values(name)
    mode = sc.SCENARIOS[name]["mode"]
    p, pt, ns, ni = build(orc, name, seed)
    rngs = orc.make_rngs(1, seed)
    av = 0.5
    pear = v.get("potential") == "pear"
    radial = v.get("radial", "sphere" if v.get("geometry") == "sphere" else "none")
    MCT, MCS = int(v.get("mct", 100)), int(v.get("mcs", 100))
    MCQ = int(v.get("mcq", MCT))

    def run(n):
        nonlocal av
        for _ in range(n):
            _, av = orc.sweep(p, ni, pt, ns, av, rngs, 1)

    def block(label):
        nonlocal av
        samples = []
        for _ in range(MCS):
            _, av = orc.sweep(p, ni, pt, ns, av, rngs, 1)
            S, E, extras = measure(orc, p, ni, pt, ns, pear, radial)
            samples.append([S, E] + extras)
        a = np.array(samples)
        S, E = a[:, 0], a[:, 1]
        return [label, S.mean(), (S * S).mean() - S.mean() ** 2, E.mean(), (E * E).mean() - E.mean() ** 2] + \
            list(a[:, 2:].mean(0))

    rows = []
    t0 = time.time()
    if mode in ("thermal", "electric"):
        a, b, d = (("ti", "tf", "dt") if mode == "thermal" else ("ei", "ef", "de"))
        lo, hi, step = f32(v[a]), f32(v[b]), f32(v[d])
        n = orc.sweep_points(lo, hi, step)
        for it in range(n):
            x = f32(orc.sweep_value(lo, hi, step, it, n))
            if mode == "thermal":
                p.T = x
            else:
                p.T = f32(v["ti"])
                p.elecE = x
            run(MCT)
            rows.append(block(x))
    else:
        first, fn = int(v.get("first_file_number", 0)), int(v.get("fn", 1))
        for ii in range(first, first + fn):
            p.T = f32(v["ti"])
            run(MCT)
            if mode == "quench":
                p.T = f32(v["tf"])
                run(MCQ)
            rows.append(block(ii))
    print("%s seed %d done in %.1f s" % (name, seed, time.time() - t0), flush=True)
    return name, seed, np.array(rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=16)
    ap.add_argument("--only", default=None)
    ap.add_argument("--out", default=os.path.join(HERE, "scenarios_oracle.npz"))
    ap.add_argument("--workers", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    names = [a.only] if a.only else list(sc.SCENARIOS)
    jobs = [(n, s) for n in names for s in range(1, a.reps + 1)]
    with mp.Pool(a.workers) as pool:
        done = pool.map(one, jobs)
    out = dict(np.load(a.out)) if os.path.exists(a.out) else {}  # other scenarios of an existing file are kept
    for n in names:
        out["rows_" + n] = np.stack([r for (m, s, r) in sorted(done, key=lambda t: t[1]) if m == n])
        out["columns_" + n] = np.array(sc.columns(n))
    np.savez_compressed(a.out, **out)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
