#!/usr/bin/env python3
"""Throughput benchmark of the checkerboard Metropolis sweep on one MI355X.

Workload: Lebwohl-Lasher (A = 1) 128x128x128 bulk lattice, periodic in x, y and z, FP32 directors initialised
homogeneous along z, fixed T = 1.10 (just below T_NI ~ 1.123), Metropolis rotation proposals with the adaptive
maximum angle (ang_var rule, about 50 % acceptance), FP64 lattice reductions for S and E every 10 sweeps.
A step is one sweep: one Metropolis trial per site of the lattice. The default run times 20000 sweeps, the
length of the benchmark configuration, after 2000 warmup sweeps; the directors stay resident in HBM throughout.
The sweeps are launched eagerly (two kernel launches per sweep), with the launch shape from the library's
autotuner; one pair of HIP events brackets the timed call on the library's stream. (--graph 50 replays captured
50-sweep HIP graphs instead, the mclics program's default, with one event pair per replay: the same rate over 20000
sweeps, 1-2 % slower for a 20-sweep call.) The measurement every 10 sweeps is folded into the sweep kernel itself.

    python bench.py [--gpus 1] [--steps 20000] [--warmup 2000] [--no-cpu-baseline]

Prints one JSON line: site-updates/ns over the timed sweeps, the roofline of the sweep kernel (algorithmic bytes
per launch over its mean launch duration, measured with HIP events on the library's stream) and the CPU baseline
(the oracle's OpenMP restatement of the reference sweep, timed on this host's cores on a bounded sample).

Two further workloads print a line of the same form (the default run prints only the 128^3 one):
  --workload slab    GHRL (K11, K22, K33) = (1.0, 0.6, 1.4) 256x256x64 slab, periodic x/y, free z, homeotropic
                     W = 1.5 on both z surfaces, T = 0.5, random directors, field along x (dielectric coupling
                     0.25) swept 0 -> 2.0 in 40 steps, MCT = 200 and MCS = 100 sweeps per field value, each MCS
                     sweep measured (electric mode); --warmup sweeps at field 0 first.
  --workload batch   512 Lebwohl-Lasher lattices in one context, sizes 16^3..64^3, bulk or homeotropic sphere
                     (W = 1.0), T uniform in [0.9, 1.3], per-lattice seeds, autotuned launch shape, MCT = 1000 and
                     MCS = 2000 with a measurement every 10 sweeps (graph replay is off while kernels are timed).
                     The batch's directors exceed the 256 MB Infinity Cache, so the library sweeps it in lattice
                     groups of at most 192 MiB, each through all the sweeps of a call (identical results).
  --workload refcases  the reference's own benchmark cases (paper/benchmark/run.sh:38-87): periodic bulk from a
                     random IC (seed 1), LL (A = 1) at L = 32, 64, 128, 256 and T = 1.0, GHRL with the E7 constants
                     at L = 64 and T = 0.1, run through the mclics program in thermalGPU mode with MCT = N, MCS = 1.
                     As in the reference's make_table.py, the cost is the best wall time of the N-sweep runs minus
                     the best of the 0-sweep runs (same setup, ic.csv and snapshot), so setup and output are left
                     out; N is larger than the reference's so that the sweeps dominate on this GPU. One JSON line
                     with a row per case, beside the reference's RTX 4070 figures from its results.csv.
  --workload smalllat  the small lattices of the reference's paper campaigns, whose sweeps are bound by launch
                     latency: the six 10x10x16 GHRL slabs of the Freedericksz campaign (E7 constants, W = 10,
                     T = 0.0008, field 0.6) in one context, and one LL 16^3 bulk at T = 1.1, --steps plain sweeps
                     each (graph replay as --graph sets), timed with the resident sweeps (each lattice held in one
                     workgroup's LDS) and with the two-pass kernels; the directors are compared bit for bit.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Metropolis site-updates/ns at 128^3 LL; GPU sweeps to locate T_NI to ±0.002"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_UPDATE = 24          # read + write of one FP32 director triplet per site update
BYTES_PER_MASKED_UPDATE = 25   # plus the one-byte point type of a masked lattice (slab, sphere)
N = 128
T_BENCH = 1.10
MEASURE_EVERY = 10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(threads, xcd_order):
    """HBM bytes per sweep-kernel launch at the given launch shape, measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy kernel calibration), read from the committed summary if there is one."""
    path = os.path.join(ROOT, "profiles", "sweep_ll128_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch_%d_%d:\s*([0-9.eE+]+)" % (int(threads), int(xcd_order)),
                  open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(budget_s=15.0):
    """The oracle's restatement of the reference OpenMP sweep (evolve.cpp:85-221) on this host's cores, same
    lattice and temperature, a bounded number of sweeps (about budget_s seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = max(1, min(threads, os.cpu_count() or 1))
    pt, _ = orc.geometry("bulk", N, N, N)
    p = orc.params(N, N, N, periodic=(1, 1, 1), potential="ll", A=1.0, T=T_BENCH)
    ni = orc.ic_homogeneous(pt, 0.0, 0.0)
    t1, av = orc.time_sweeps(p, ni, pt, None, 2, 1, threads)  # warm-up and calibration
    nsw = max(3, int(budget_s / max(t1 / 2, 1e-3)))
    t, _ = orc.time_sweeps(p, ni, pt, None, nsw, 2, threads, av)
    rate = nsw * N ** 3 / (t * 1e9)
    return {"value": rate, "unit": "site-updates/ns", "cores": threads, "kind": "port",
            "sample": "%d sweeps of the same 128^3 LL lattice at T = 1.10 (%.1f s), oracle OpenMP sweep with %d "
                      "threads" % (nsw, t, threads)}


def pick_launch(eng, a, trial_sweeps, measure_every):
    """The autotuner's launch shape for trials measured at the workload's cadence (the state is restored after the
    trials), or the one --launch fixes."""
    if a.launch:
        th, xo = (int(v) for v in a.launch.split(","))
        eng.set_launch(th, xo)
        return {"threads": th, "xcd_order": xo, "fixed": True}
    return eng.autotune(trial_sweeps, measure_every)


def kernel_roofline(launches, kms, bytes_per_launch, copy_gbs):
    avg_launch_s = kms / launches / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_sweep (one colour pass)",
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "timed_launches": launches, "copy_measured_gbs": None if copy_gbs is None else round(copy_gbs, 1)}


def copy_check(nbytes):
    from mclics_amd import capi
    if nbytes <= 0:
        return None
    gbs, _ = capi.copy_bandwidth(nbytes, 20)
    log("bench: device copy of %d MiB: %.0f GB/s (read + write)" % (nbytes >> 20, gbs))
    return gbs


def line(value, steps, warmup, elapsed, config, roofline, cpu, dtype="f32"):
    return {"metric": METRIC, "value": round(value, 4), "unit": "site-updates/ns", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(elapsed * 1e3 / steps, 6), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config,
            "roofline": roofline, "cpu_baseline": cpu}


SLAB = dict(dims=(256, 256, 64), k=(1.0, 0.6, 1.4), W=1.5, T=0.5, elecA=0.25, Ei=0.0, Ef=2.0, steps=40, MCT=200,
            MCS=100, seed=1)


def slab_model(capi):
    """The slab workload's physics through libmclgpu (GHRL coefficients of setGHRL, io.cpp:340-360)."""
    c = capi.ghrl_coefficients(*SLAB["k"], 0.0)
    return capi.make_model("ghrl", ghrl=(c["lam"], c["mu"], c["nu"], c["rho"], c["sigma"]), elec_dir=(1.0, 0.0, 0.0),
                           elecA=SLAB["elecA"], elecE=0.0)


def run_slab(a):
    from mclics_amd import batch, capi
    nx, ny, nz = SLAB["dims"]
    pt, ns = batch.slab_geometry(nx, ny, nz)
    nvalid = int((pt > 0).sum())
    lat = capi.Lattice(nx, ny, nz, batch.random_directors(pt, SLAB["seed"]), periodic=(1, 1, 0), pt=pt, ns=ns,
                       surfaces=[("homeotropic", SLAB["W"], (0, 0, 0)), ("homeotropic", SLAB["W"], (0, 0, 0))],
                       model=slab_model(capi), T=SLAB["T"], seed=SLAB["seed"])
    fields = [float(np.float32(SLAB["Ei"] + (SLAB["Ef"] - SLAB["Ei"]) * i / SLAB["steps"]))
              for i in range(SLAB["steps"] + 1)]
    per = SLAB["MCT"] + SLAB["MCS"]
    steps = len(fields) * per
    log("bench: slab %dx%dx%d GHRL, %d field values, %d sweeps" % (nx, ny, nz, len(fields), steps))
    rows = []
    with capi.Engine(lat) as eng:
        tune = pick_launch(eng, a, 12, 3)  # the electric mode measures one sweep in three (MCS of MCT + MCS)
        eng.sweeps(a.warmup)
        eng.synchronize()
        eng.set_kernel_timing(a.timing_stride, (steps + a.timing_stride - 1) // a.timing_stride + len(fields) * 2)
        t0 = time.perf_counter()
        for E in fields:
            eng.set_field(E)
            eng.sweeps(SLAB["MCT"])
            eng.reset_moments()
            eng.sweeps(SLAB["MCS"], 1)
            m = eng.moments()
            rows.append((E, m["S"] / m["n"], m["E"] / m["n"]))
        eng.synchronize()
        t1 = time.perf_counter()
        launches, kms = eng.kernel_timing()
        eng.set_kernel_timing(0)
    elapsed = t1 - t0
    value = steps * nvalid / (elapsed * 1e9)
    roof = kernel_roofline(launches, kms, (nvalid // 2) * BYTES_PER_MASKED_UPDATE, copy_check(a.copy_bytes))
    log("bench: %.3f s, %.2f site-updates/ns; E(field 0) = %.4f, E(field 2) = %.4f" % (elapsed, value, rows[0][2],
                                                                                      rows[-1][2]))
    cpu = None if a.no_cpu_baseline else cpu_baseline_slab()
    cfg = {"workload": "GHRL (1.0, 0.6, 1.4) 256x256x64 slab, periodic x/y, homeotropic W=1.5 on both z "
                       "surfaces, T=0.5, field along x (elecA 0.25) swept 0->2.0 in 40 steps, MCT 200 + MCS 100 "
                       "per field, every MCS sweep measured",
           "lattice": list(SLAB["dims"]), "fields": len(fields), "sweeps": steps, "launch": tune,
           "S_E": [[round(x, 5) for x in r] for r in rows[::10]]}
    return line(value, steps, a.warmup, elapsed, cfg, roof, cpu)


def cpu_baseline_slab(budget_s=15.0):
    """The oracle's OpenMP sweep on the same slab (same physics, field 1.0), bounded to about budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1, os.cpu_count() or 1))
    nx, ny, nz = SLAB["dims"]
    pt, ns = orc.geometry("slab", nx, ny, nz)
    p = orc.params(nx, ny, nz, periodic=(1, 1, 0), potential="ghrl", k=SLAB["k"] + (0.0,), elec=(1, 0, 0),
                   elecA=SLAB["elecA"], elecE=1.0, T=SLAB["T"],
                   surfaces=[("homeotropic", SLAB["W"], (0, 0, 0)), ("homeotropic", SLAB["W"], (0, 0, 0))])
    ni = orc.ic_random(pt, 1)
    t1, av = orc.time_sweeps(p, ni, pt, ns, 2, 1, threads)
    nsw = max(3, int(budget_s / max(t1 / 2, 1e-3)))
    t, _ = orc.time_sweeps(p, ni, pt, ns, nsw, 2, threads, av)
    rate = nsw * nx * ny * nz / (t * 1e9)
    log("bench: CPU %.4f site-updates/ns on %d threads" % (rate, threads))
    return {"value": rate, "unit": "site-updates/ns", "cores": threads, "kind": "port",
            "sample": "%d sweeps of the same slab at field 1.0 (%.1f s), oracle OpenMP sweep with %d threads" % (
                nsw, t, threads)}


def run_batch_workload(a):
    from mclics_amd import batch, capi
    specs = batch.c4_specs(512, 1)
    lats = [batch.build_lattice(s) for s in specs]
    nvalid = np.array([int((l.pt > 0).sum()) if l.pt is not None else int(np.prod(l.dims)) for l in lats])
    mct, mcs, every = 1000, 2000, 10
    steps = mct + mcs
    log("bench: batch of %d lattices, %d valid sites, %d sweeps" % (len(lats), nvalid.sum(), steps))
    with capi.Engine(lats) as eng:
        tune = pick_launch(eng, a, 10, MEASURE_EVERY)
        eng.sweeps(a.warmup)
        eng.synchronize()
        # the batch is larger than the Infinity Cache: the library sweeps it in lattice groups, each with event
        # pairs of its own around its runs of sweep launches
        ngroups = eng.lattice_groups()
        eng.set_kernel_timing(a.timing_stride, ngroups * (steps // a.timing_stride + 2 * (mcs // every)) + 4)
        t0 = time.perf_counter()
        eng.sweeps(mct)
        eng.reset_moments()
        eng.sweeps(mcs, every)
        eng.synchronize()
        t1 = time.perf_counter()
        launches, kms = eng.kernel_timing()
        eng.set_kernel_timing(0)
        S = np.array([eng.moments(r)["S"] / eng.moments(r)["n"] for r in range(len(lats))])
    elapsed = t1 - t0
    value = steps * float(nvalid.sum()) / (elapsed * 1e9)
    masked_sites = float(sum(n for n, l in zip(nvalid, lats) if l.pt is not None))
    bulk_sites = float(nvalid.sum()) - masked_sites
    bytes_per_launch = int((bulk_sites * BYTES_PER_UPDATE + masked_sites * BYTES_PER_MASKED_UPDATE) / 2)
    roof = kernel_roofline(launches, kms, bytes_per_launch, copy_check(a.copy_bytes))
    log("bench: %.3f s, %.2f site-updates/ns, launch shape %s" % (elapsed, value, tune))
    cfg = {"workload": "512 LL lattices, sizes 16^3..64^3 bulk or homeotropic sphere (W=1.0), T uniform in "
                       "[0.9, 1.3], per-lattice seeds, MCT 1000 + MCS 2000, measured every 10 sweeps",
           "lattices": len(lats), "valid_sites": int(nvalid.sum()), "sweeps": steps, "launch": tune,
           "lattice_groups": ngroups,
           "S_range": [round(float(S.min()), 4), round(float(S.max()), 4)]}
    cpu = None if a.no_cpu_baseline else cpu_baseline_batch(specs[:16])
    return line(value, steps, a.warmup, elapsed, cfg, roof, cpu)


def cpu_baseline_batch(specs, sweeps=20):
    """The oracle's OpenMP sweep on the first lattices of the batch (same geometry, physics and temperature),
    `sweeps` sweeps each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1, os.cpu_count() or 1))
    tot_t, tot_u = 0.0, 0.0
    for s in specs:
        L = s["L"]
        pt, ns = orc.geometry(s["geometry"], L, L, L)
        surf = [("homeotropic", 1.0, (0, 0, 0))] if s["geometry"] == "sphere" else []
        per = (0, 0, 0) if s["geometry"] == "sphere" else (1, 1, 1)
        p = orc.params(L, L, L, periodic=per, potential="ll", A=1.0, T=s["T"], surfaces=surf)
        ni = orc.ic_random(pt, s["seed"])
        t, _ = orc.time_sweeps(p, ni, pt, ns if surf else None, sweeps, s["seed"], threads)
        tot_t += t
        tot_u += sweeps * float((pt > 0).sum())
    rate = tot_u / (tot_t * 1e9)
    log("bench: CPU %.4f site-updates/ns on %d threads" % (rate, threads))
    return {"value": rate, "unit": "site-updates/ns", "cores": threads, "kind": "port",
            "sample": "%d sweeps each of the first %d lattices of the batch (%.1f s), oracle OpenMP sweep with %d "
                      "threads" % (sweeps, len(specs), tot_t, threads)}


# (case, potential lines, L, T, N sweeps, reference RTX 4070 site-updates/ns from its results.csv)
REF_CASES = (("ll32", ("potential ll", "A 1"), 32, 1.0, 1200000, 0.420),
             ("ll64", ("potential ll", "A 1"), 64, 1.0, 300000, 0.962),
             ("ll128", ("potential ll", "A 1"), 128, 1.0, 200000, 1.010),
             ("ll256", ("potential ll", "A 1"), 256, 1.0, 40000, 1.154),
             ("ghrl64", ("potential ghrl", "k11 1.48", "k22 1", "k33 2.28"), 64, 0.1, 200000, 1.043))


def run_refcases(a, repeat=2):
    """The reference's benchmark method on the mclics program (see the module docstring)."""
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "mclics_amd", "bin", "mclics")
    if not os.path.exists(exe):
        raise RuntimeError("mclics is not built (python -m mclics_amd.build)")
    rows = []
    work = tempfile.mkdtemp(prefix="mcl_refcases_")
    try:
        for name, pot, L, T, n, ref in REF_CASES:
            best = {}
            for rep in range(repeat):
                for steps in (n, 0):
                    d = os.path.join(work, "%s_%d" % (name, steps))
                    os.makedirs(d, exist_ok=True)
                    lines = ["Nx %d" % L, "Ny %d" % L, "Nz %d" % L, "geometry bulk", "xbound periodic",
                             "ybound periodic", "zbound periodic"] + list(pot) + [
                        "evol thermalGPU", "Ti %g" % T, "Tf %g" % T, "dT -0.1", "MCT %d" % steps, "MCS 1",
                        "ic random", "seed 1"]
                    with open(os.path.join(d, "input_parameters.txt"), "w") as f:
                        f.write("\n".join(lines) + "\n")
                    t0 = time.perf_counter()
                    r = subprocess.run([exe, "input_parameters.txt"], cwd=d, stdout=subprocess.DEVNULL,
                                       stderr=subprocess.PIPE, text=True)
                    t = time.perf_counter() - t0
                    if r.returncode != 0:
                        raise RuntimeError("mclics failed on %s (%d sweeps): %s" % (name, steps, r.stderr[-500:]))
                    best[steps] = min(best.get(steps, 1e30), t)
                    shutil.rmtree(d)
                    log("refcases: %s L=%d steps=%d %.2f s" % (name, L, steps, t))
            cost = best[n] - best[0]
            value = n * float(L) ** 3 / (cost * 1e9)
            rows.append({"case": name, "L": L, "T": T, "steps": n, "t_steps_s": round(best[n], 3),
                         "t_0_s": round(best[0], 3), "value": round(value, 3),
                         "reference_rtx4070": ref, "ratio": round(value / ref, 1)})
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return {"metric": "site-updates/ns on the reference's benchmark cases (mclics thermalGPU, N-sweep minus "
                      "0-sweep wall time)", "unit": "site-updates/ns", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "paper/benchmark/run.sh cases: periodic bulk, random IC seed 1, LL L=32..256 "
                                   "T=1.0, GHRL E7 L=64 T=0.1; best of %d" % repeat,
                       "program_defaults": "graph replay of 50-sweep blocks, deterministic launch shape"},
            "cases": rows}


def run_smalllat(a):
    """Plain sweeps of small lattices, resident against two-pass (the same trajectories)."""
    from mclics_amd import batch, capi, freedericksz
    rng_dirs = lambda n, seed: batch.random_directors(np.ones(n, dtype=np.int32), seed)
    cases = [("six 10x10x16 GHRL slabs (Freedericksz cells, seeds 1 and 2)",
              lambda: [freedericksz.cell_lattice(c, s, 0.0008, 10.0, 0.25) for s in (1, 2)
                       for c in ("splay", "twist", "bend")], 0.6),
             ("LL 16^3 bulk, T = 1.1", lambda: [capi.Lattice(16, 16, 16, rng_dirs(4096, 1), model=capi.make_model("ll"),
                                                             T=1.1, seed=1)], None)]
    rows = []
    for name, make, field in cases:
        row = {"case": name}
        dirs = {}
        for resident in (True, False):
            lats = make()
            sites = sum(l.dims[0] * l.dims[1] * l.dims[2] for l in lats)
            with capi.Engine(lats) as eng:
                eng.set_graph(a.graph)
                eng.set_resident(resident)
                if field is not None:
                    eng.set_field(field)
                eng.sweeps(a.warmup)
                eng.synchronize()
                t0 = time.perf_counter()
                eng.sweeps(a.steps)
                eng.synchronize()
                dt = time.perf_counter() - t0
                dirs[resident] = [eng.directors(r) for r in range(len(lats))]
            key = "resident" if resident else "two_pass"
            row[key + "_us_per_sweep"] = round(dt / a.steps * 1e6, 3)
            row[key + "_site_updates_per_ns"] = round(a.steps * sites / (dt * 1e9), 4)
        row["speedup"] = round(row["two_pass_us_per_sweep"] / row["resident_us_per_sweep"], 3)
        row["bit_identical"] = all(np.array_equal(x, y) for x, y in zip(dirs[True], dirs[False]))
        log("bench: %s: %.2f us per sweep resident, %.2f two-pass (x%.2f), identical %s" % (
            name, row["resident_us_per_sweep"], row["two_pass_us_per_sweep"], row["speedup"], row["bit_identical"]))
        rows.append(row)
    first = rows[0]
    return {"metric": "Metropolis site-updates/ns of small lattices (resident sweeps)",
            "value": first["resident_site_updates_per_ns"], "unit": "site-updates/ns", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(first["resident_us_per_sweep"] / 1e3, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "small lattices of the paper campaigns, plain sweeps", "cases": rows},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--timing-stride", type=int, default=None,
                    help="sweeps per event pair of the kernel timing (default: the whole timed call for the 128^3 "
                         "run, 10 for the slab and batch runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=256 << 20,
                    help="size of the STREAM-style copy that checks the HBM peak (0: skip)")
    ap.add_argument("--workload", choices=("ll128", "slab", "batch", "refcases", "smalllat"), default="ll128")
    ap.add_argument("--launch", default=None, metavar="THREADS,XCD_ORDER",
                    help="fixed launch shape instead of the autotuner's choice (for profiling runs, whose tracer "
                         "changes the timings the autotuner compares)")
    ap.add_argument("--graph", type=int, default=None,
                    help="sweeps per captured HIP graph (default: 0, eager launches, for the 128^3 run; 50 for the "
                         "smalllat run)")
    a = ap.parse_args()
    if a.graph is None:
        a.graph = 50 if a.workload == "smalllat" else 0
    if a.timing_stride is None:
        a.timing_stride = max(1, a.steps) if a.workload == "ll128" else 10
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the benchmark lattice is one 128^3 system on one MI355X "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    from mclics_amd import capi
    capi.load()
    if a.workload != "ll128":
        out = {"slab": run_slab, "batch": run_batch_workload, "refcases": run_refcases,
               "smalllat": run_smalllat}[a.workload](a)
        print(json.dumps(out), flush=True)
        return
    nt = N ** 3
    ni = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), nt)
    lat = capi.Lattice(N, N, N, ni, periodic=(1, 1, 1), model=capi.make_model("ll", A=1.0), T=T_BENCH,
                       ang_var=0.5, seed=1)
    log("bench: 128^3 LL periodic, T = %.2f, warmup %d sweeps" % (T_BENCH, a.warmup))
    with capi.Engine(lat) as eng:
        tune = pick_launch(eng, a, 20, MEASURE_EVERY)
        log("bench: launch shape %s" % tune)
        if a.graph:
            eng.set_graph(a.graph)  # replay captured graphs; the events then bracket each replay
            # capture the timed call's graphs now (nothing runs), before the warmup, so that the warmup sweeps run
            # right up to the timed region
            eng.prepare(a.steps, MEASURE_EVERY)
        eng.sweeps(a.warmup, MEASURE_EVERY)
        eng.synchronize()
        eng.reset_moments()
        n_timed = (a.steps + a.timing_stride - 1) // a.timing_stride + 2
        eng.set_kernel_timing(a.timing_stride, n_timed)
        log("bench: timing %d sweeps" % a.steps)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.sweeps(a.steps, MEASURE_EVERY)
        eng.synchronize()
        t1 = time.perf_counter()
        launches, kms = eng.kernel_timing()
        eng.set_kernel_timing(0)
        mom = eng.moments()
        acc, done = eng.sweep_stats()
        av = eng.ang_var()
    copy_gbs = None
    if a.copy_bytes > 0:
        copy_gbs, copy_ms = capi.copy_bandwidth(a.copy_bytes, 20)
        log("bench: device copy of %d MiB: %.0f GB/s (read + write)" % (a.copy_bytes >> 20, copy_gbs))
    elapsed = t1 - t0
    value = a.steps * nt / (elapsed * 1e9)
    avg_launch_s = kms / launches / 1e3
    bytes_per_launch = (nt // 2) * BYTES_PER_UPDATE  # one colour: nt / 2 site updates
    achieved = bytes_per_launch / avg_launch_s / 1e9
    S = mom["S"] / mom["n"]
    E = mom["E"] / mom["n"]
    log("bench: %.3f s, %.2f site-updates/ns, sweep kernel %.2f us/launch, S = %.4f, E = %.4f, acceptance %.3f, "
        "ang_var %.3f" % (elapsed, value, avg_launch_s * 1e6, S, E, acc / nt, av))
    cpu = None
    if not a.no_cpu_baseline:
        log("bench: CPU baseline (oracle OpenMP sweep), ~15 s")
        cpu = cpu_baseline()
        log("bench: CPU %.4f site-updates/ns on %d threads" % (cpu["value"], cpu["cores"]))
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "site-updates/ns", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(elapsed * 1e3 / a.steps, 6), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "LL 128^3 periodic bulk, homogeneous IC along z, T=1.10, adaptive ang_var, "
                               "FP64 S/E reductions every 10 sweeps",
                   "lattice": [N, N, N], "T": T_BENCH, "measure_every": MEASURE_EVERY, "sweeps": a.steps,
                   "launch": tune, "graph_sweeps": a.graph,
                   "S_mean": round(S, 6), "E_mean": round(E, 6), "measurements": int(mom["n"])},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(tune["threads"], tune["xcd_order"]),
                     "kernel": "k_sweep (one colour pass)", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": round(avg_launch_s * 1e6, 3), "timed_launches": launches,
                     "copy_measured_gbs": None if copy_gbs is None else round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
