"""Writes tests/golden/c3_oracle.npz: the oracle's electric-mode run of a small Freedericksz slab, 16 replicas.

Case (the slab configuration at a size the oracle runs in seconds): GHRL potential with K11 = 1.0, K22 = 0.6,
K33 = 1.4 (io.cpp:setGHRL), 16x16x8 slab periodic in x/y and free in z, homeotropic anchoring W = 1.5 on both z
surfaces, T = 0.5, field along x with dielectric coupling elecA = 0.25 (the reference's Freedericksz default,
paper/freedericksz/run.sh:44), field magnitude swept 0 -> 2.0 in steps of 0.2 (11 points), MCT = 300, MCS = 200,
random initial directors from the replica's seed. Replica r uses seed r + 1 for the initial condition and the sweep
RNG. One process per replica, one thread each.

    python tests/golden/make_c3_fixture.py [--reps 16] [--out tests/golden/c3_oracle.npz]
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

CASE = dict(Nx=16, Ny=16, Nz=8, k=(1.0, 0.6, 1.4, 0.0), W=1.5, T=0.5, elecA=0.25, elec=(1.0, 0.0, 0.0), Ei=0.0,
            Ef=2.0, dE=0.2, MCT=300, MCS=200)


def case_params(orc):
    c = CASE
    return orc.params(c["Nx"], c["Ny"], c["Nz"], periodic=(1, 1, 0), potential="ghrl", k=c["k"], elec=c["elec"],
                      elecA=c["elecA"], elecE=0.0, T=c["T"],
                      surfaces=[("homeotropic", c["W"], (0, 0, 0)), ("homeotropic", c["W"], (0, 0, 0))])


def one(seed):
    os.environ["OMP_NUM_THREADS"] = "1"
    import pyoracle as orc
    c = CASE
    pt, ns = orc.geometry("slab", c["Nx"], c["Ny"], c["Nz"])
    p = case_params(orc)
    ni = orc.ic_random(pt, seed)
    f32 = lambda x: float(np.float32(x))
    t0 = time.time()
    rows = orc.electric_run(p, ni, pt, ns, f32(c["Ei"]), f32(c["Ef"]), f32(c["dE"]), c["MCT"], c["MCS"], seed, 1)
    print("seed %d done in %.1f s" % (seed, time.time() - t0), flush=True)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=16)
    ap.add_argument("--out", default=os.path.join(HERE, "c3_oracle.npz"))
    a = ap.parse_args()
    with mp.Pool(min(a.reps, 8)) as pool:
        rows = pool.map(one, range(1, a.reps + 1))
    np.savez_compressed(a.out, rows=np.stack(rows), case=str(CASE))
    print("wrote", a.out)


if __name__ == "__main__":
    main()
