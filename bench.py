#!/usr/bin/env python3
"""Throughput benchmark of the checkerboard Metropolis sweep on one MI355X.

Workload: Lebwohl-Lasher (A = 1) 128x128x128 bulk lattice, periodic in x, y and z, FP32 directors initialised
homogeneous along z, fixed T = 1.10 (just below T_NI ~ 1.123), Metropolis rotation proposals with the adaptive
maximum angle (ang_var rule, about 50 % acceptance), FP64 lattice reductions for S and E every 10 sweeps.
A step is one sweep: one Metropolis trial per site of the lattice. The default run times 20000 sweeps, the
length of the benchmark configuration, after 2000 warmup sweeps; the directors stay resident in HBM throughout.

    python bench.py [--gpus 1] [--steps 20000] [--warmup 2000] [--no-cpu-baseline]

Prints one JSON line: site-updates/ns over the timed sweeps, the roofline of the sweep kernel (algorithmic bytes
per launch over its mean launch duration, measured with HIP events on the library's stream) and the CPU baseline
(the oracle's OpenMP restatement of the reference sweep, timed on this host's cores on a bounded sample).
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
N = 128
T_BENCH = 1.10
MEASURE_EVERY = 10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per sweep-kernel launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    kernel calibration), read from the committed summary if there is one."""
    path = os.path.join(ROOT, "profiles", "sweep_ll128_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--timing-stride", type=int, default=10, help="time the kernels of every n-th sweep")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=256 << 20,
                    help="size of the STREAM-style copy that checks the HBM peak (0: skip)")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the benchmark lattice is one 128^3 system on one MI355X "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    from mclics_amd import capi
    capi.load()
    nt = N ** 3
    ni = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), nt)
    lat = capi.Lattice(N, N, N, ni, periodic=(1, 1, 1), model=capi.make_model("ll", A=1.0), T=T_BENCH,
                       ang_var=0.5, seed=1)
    log("bench: 128^3 LL periodic, T = %.2f, warmup %d sweeps" % (T_BENCH, a.warmup))
    with capi.Engine(lat) as eng:
        eng.sweeps(a.warmup, MEASURE_EVERY)
        eng.synchronize()
        eng.reset_moments()
        n_timed = (a.steps + a.timing_stride - 1) // a.timing_stride
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
                   "S_mean": round(S, 6), "E_mean": round(E, 6), "measurements": int(mom["n"])},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": "k_sweep (one colour pass)", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": round(avg_launch_s * 1e6, 3), "timed_launches": launches,
                     "copy_measured_gbs": None if copy_gbs is None else round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
