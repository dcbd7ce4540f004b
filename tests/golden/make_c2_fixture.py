"""Writes tests/golden/c2_oracle.npz: the oracle's run of the 128^3 LL throughput case, 8 replicas.

Case (the GPU throughput configuration at a length the oracle finishes in about two hours on 8 cores):
Lebwohl-Lasher A = 1, 128x128x128 bulk, periodic x/y/z, homogeneous initial directors along z, fixed T = 1.10,
ang_var starting at 0.5 with the adaptive rule, 6000 sweeps. S and E are measured after every 10th sweep (the
FP64 reductions of the throughput case), over the whole run, so a test can score the measurement window
[3000, 6000) and follow the relaxation before it. Replica r uses seed r + 1 for the sweep RNG of each of its
8 threads (evolve_thermal.cpp:49-50: one generator per thread); the thread count is fixed at 8 so the fixture
does not depend on the machine's core count.

The oracle's sweep is evolve.cpp:85-221 (eight sublattice ticks); its energy and Q-tensor measurement are
evolve.cpp:33-75 and parameter_order.cpp:19-71.

    python tests/golden/make_c2_fixture.py [--reps 8] [--out tests/golden/c2_oracle.npz]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CASE = dict(N=128, T=1.10, sweeps=6000, measure_every=10, threads=8)


def one(orc, seed):
    N, T = CASE["N"], np.float32(CASE["T"])
    pt = np.ones(N ** 3, dtype=np.int32)
    p = orc.params(N, N, N, periodic=(1, 1, 1), potential="ll", A=1.0, T=float(T))
    ni = orc.ic_homogeneous(pt, 0.0, 0.0)
    rngs = orc.make_rngs(CASE["threads"], seed)
    av = 0.5
    every = CASE["measure_every"]
    S, E, acc = [], [], []
    t0 = time.time()
    for s in range(1, CASE["sweeps"] + 1):
        a, av = orc.sweep(p, ni, pt, None, av, rngs, CASE["threads"])
        if s % every == 0:
            E.append(orc.energy_mean(p, ni, pt))
            lam, _, _ = orc.eig_sym3(orc.q_tensor(ni, pt))
            S.append(lam)
            acc.append(a / N ** 3)
        if s % 500 == 0:
            print("seed %d sweep %d: S %.4f E %.4f acc %.3f ang_var %.3f (%.0f s)" % (
                seed, s, S[-1], E[-1], acc[-1], av, time.time() - t0), flush=True)
    return np.array(S), np.array(E), np.array(acc)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "c2_oracle.npz"))
    a = ap.parse_args()
    import pyoracle as orc
    seeds = list(range(1, a.reps + 1))
    S, E, acc = [], [], []
    for seed in seeds:
        s, e, c = one(orc, seed)
        S.append(s)
        E.append(e)
        acc.append(c)
        # written after every replica, so an interrupted run keeps what it finished
        np.savez_compressed(a.out, S=np.array(S), E=np.array(E), acc=np.array(acc), seeds=np.array(seeds[:len(S)]),
                            case=np.array([str(CASE)]))
    print("wrote", a.out)


if __name__ == "__main__":
    main()
