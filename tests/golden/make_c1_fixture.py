"""Writes tests/golden/c1_oracle.npz: the oracle's thermal run of the 16^3 LL case, 64 replicas.

Case (the CPU configuration of the benchmark): Lebwohl-Lasher A = 1, 16x16x16 bulk, periodic x/y/z, random
initial directors from the replica's seed, T = 1.2 -> 0.1 with dT = -0.02 (56 points), MCT = 500, MCS = 200.
Replica r uses seed r + 1 for both the initial condition and the sweep RNG. One process per replica, one thread
each, so the fixture does not depend on the machine's core count.

    python tests/golden/make_c1_fixture.py [--reps 64] [--out tests/golden/c1_oracle.npz]
"""
import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CASE = dict(N=16, Ti=1.2, Tf=0.1, dT=-0.02, MCT=500, MCS=200)


def one(seed):
    os.environ["OMP_NUM_THREADS"] = "1"  # one thread per replica process, also in the energy reduction
    import pyoracle as orc
    N = CASE["N"]
    pt, ns = orc.geometry("bulk", N, N, N)
    p = orc.params(N, N, N, periodic=(1, 1, 1), potential="ll", A=1.0)
    ni = orc.ic_random(pt, seed)
    f32 = lambda x: float(np.float32(x))  # the input values are floats in the reference (parameters.h:44-46)
    t0 = time.time()
    rows = orc.thermal_run(p, ni, pt, None, f32(CASE["Ti"]), f32(CASE["Tf"]), f32(CASE["dT"]), CASE["MCT"],
                           CASE["MCS"], seed, 1)
    print("seed %d done in %.1f s" % (seed, time.time() - t0), flush=True)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=64)
    ap.add_argument("--out", default=os.path.join(HERE, "c1_oracle.npz"))
    ap.add_argument("--workers", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    seeds = list(range(1, a.reps + 1))
    with mp.Pool(a.workers) as pool:
        rows = pool.map(one, seeds)
    np.savez_compressed(a.out, rows=np.array(rows), seeds=np.array(seeds),
                        columns=np.array(["T", "S", "varS", "E", "varE", "Sr_sph"]),
                        case=np.array([str(CASE)]))
    print("wrote", a.out)


if __name__ == "__main__":
    main()
