"""Statistical parity of the mclics CPU modes (the program's own OpenMP sweep, evol thermal / electric without the
GPU suffix) with the CPU oracle, on the two configurations the GPU is tested on (the oracle's runs are the committed
fixtures tests/golden/c1_oracle.npz and c3_oracle.npz).

Each replica is one run of the mclics binary with its own seed (1..16) and one OpenMP thread; 8 run at a time.
The program draws its own random initial directors (the oracle's replicas start from the oracle's), which the
500 (300) thermalization sweeps at the first point forget.

Criterion, as for the GPU (tests/test_gpu_stat_parity.py): z = (mean_cpu - mean_oracle) / sqrt(s_cpu^2 / n_cpu +
s_oracle^2 / n_oracle) over S, varS, E and varE at every point; C1: at most 4 of 224 beyond 3 sigma and none beyond
5; C3: at most 2 of 44 beyond 3 sigma and none beyond 5.
"""
import concurrent.futures as cf
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import z_scores

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PROGRAM = os.path.join(ROOT, "mclics_amd", "bin", "mclics")
REPLICAS = 16


def run_replica(lines, seed):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "input.txt"), "w") as f:
            f.write("\n".join(lines + ["seed %d" % seed]) + "\n")
        p = subprocess.run([PROGRAM, "input.txt"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           universal_newlines=True, timeout=600, env=dict(os.environ, OMP_NUM_THREADS="1"))
        assert p.returncode == 0, p.stderr[-2000:]
        return np.loadtxt(os.path.join(d, "po.dat"), skiprows=1)


def run_replicas(lines):
    workers = max(1, min(8, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(workers) as pool:
        return np.stack(list(pool.map(lambda s: run_replica(lines, s), range(1, REPLICAS + 1))))


def check(cpu, ref, col_name, max_beyond_3):
    names = ["S", "varS", "E", "varE"]
    zs = np.stack([z_scores(cpu[:, :, c], ref[:, :, c]) for c in (1, 2, 3, 4)])
    report = "; ".join("%s max|z| %.2f at %s=%.2f" % (names[c], np.abs(zs[c]).max(), col_name,
                                                      ref[0, int(np.argmax(np.abs(zs[c]))), 0]) for c in range(4))
    print(report)
    assert int((np.abs(zs) > 3).sum()) <= max_beyond_3, report
    assert np.abs(zs).max() < 5, report


def test_c1_thermal_cpu_mode_parity():
    """LL A = 1, 16^3 periodic bulk, random start, T = 1.2 -> 0.1 in steps of -0.02, MCT = 500, MCS = 200."""
    ref = np.load(os.path.join(HERE, "golden", "c1_oracle.npz"))["rows"]
    lines = ["Nx 16", "Ny 16", "Nz 16", "geometry bulk", "xbound periodic", "ybound periodic", "zbound periodic",
             "potential ll", "A 1", "evol thermal", "Ti 1.2", "Tf 0.1", "dT -0.02", "MCT 500", "MCS 200", "ic random"]
    cpu = run_replicas(lines)
    assert cpu.shape == (REPLICAS, 56, 5)
    assert np.allclose(cpu[0, :, 0], ref[0, :, 0], atol=1e-5)
    check(cpu, ref, "T", 4)
    assert cpu[:, -1, 1].mean() > 0.95 and cpu[:, 0, 1].mean() < 0.15


def test_c3_electric_slab_cpu_mode_parity():
    """GHRL (1.0, 0.6, 1.4) 16x16x8 slab, homeotropic W = 1.5 on both surfaces, T = 0.5, field along x with
    elecA = 0.25 swept 0 -> 2.0 in steps of 0.2, MCT = 300, MCS = 200."""
    ref = np.load(os.path.join(HERE, "golden", "c3_oracle.npz"))["rows"]
    lines = ["Nx 16", "Ny 16", "Nz 8", "geometry slab", "xbound periodic", "ybound periodic", "potential ghrl",
             "k11 1.0", "k22 0.6", "k33 1.4", "anchoring_type 0 homeotropic", "W 0 1.5",
             "anchoring_type 1 homeotropic", "W 1 1.5", "evol electric", "Ti 0.5", "elecA 0.25", "elecX 1",
             "elecY 0", "elecZ 0", "Ei 0", "Ef 2.0", "dE 0.2", "MCT 300", "MCS 200", "ic random"]
    cpu = run_replicas(lines)
    assert cpu.shape == (REPLICAS, 11, 5)
    assert np.allclose(cpu[0, :, 0], ref[0, :, 0], atol=1e-5)
    check(cpu, ref, "field", 2)
