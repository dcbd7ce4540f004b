"""The reference's seven regression scenarios (tests/scenarios.py) through the mclics program, in the GPU modes (evol
<mode>GPU, as the reference's inputs are written) and in the CPU modes (the suffix removed, as the reference's
run_regression.sh --cpu does), each against the oracle's runs of the same scenario (tests/golden/
scenarios_oracle.npz, written by tests/golden/make_scenario_fixtures.py).

Every run must write po.dat with the reference's header for the scenario (mode column, moments, pear and radial
extras) and one snapshot per po.dat line. Then the 16 replicas (seeds 1..16, one process each) are compared with the
oracle's 16: z = (mean_mclics - mean_oracle) / sqrt(s_mclics^2 / 16 + s_oracle^2 / 16), s the replica-to-replica
standard deviation, for every column and every point. With k comparisons in a scenario, at most 1 + k // 50 may lie
beyond 3 sigma (about k / 250 are expected by chance with 30 degrees of freedom) and none beyond 5. The scenarios
follow the same schedule from matched starts (random, or the cylinder's helix), so points still relaxing (the quench
and step blocks) compare like with like.
"""
import concurrent.futures as cf
import os
import subprocess
import tempfile

import numpy as np
import pytest

import scenarios as sc
from helpers import z_scores

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PROGRAM = os.path.join(ROOT, "mclics_amd", "bin", "mclics")
FIXTURE = os.path.join(HERE, "golden", "scenarios_oracle.npz")
REPLICAS = 16


def run_one(name, suffix, seed):
    mode = sc.SCENARIOS[name]["mode"]
    with tempfile.TemporaryDirectory() as d:
        lines = sc.SCENARIOS[name]["lines"] + ["evol %s%s" % (mode, suffix), "seed %d" % seed]
        with open(os.path.join(d, "input.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")
        for fname, text in sc.files(name).items():
            with open(os.path.join(d, fname), "w") as f:
                f.write(text)
        print("%s %s seed %d" % (name, suffix or "CPU", seed), flush=True)
        p = subprocess.run([PROGRAM, "input.txt"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           universal_newlines=True, timeout=300, env=dict(os.environ, OMP_NUM_THREADS="1"))
        assert p.returncode == 0, "%s seed %d: exit %d\n%s" % (name, seed, p.returncode, p.stderr[-2000:])
        with open(os.path.join(d, "po.dat")) as f:
            header = f.readline().split()
        rows = np.loadtxt(os.path.join(d, "po.dat"), skiprows=1, ndmin=2)
        snaps = [f for f in os.listdir(d) if f.startswith("director_field_")]
        return header, rows, len(snaps)


@pytest.mark.parametrize("name", list(sc.SCENARIOS))
@pytest.mark.parametrize("suffix", [pytest.param("GPU", marks=pytest.mark.gpu, id="gpu"), pytest.param("", id="cpu")])
def test_scenario_against_oracle(name, suffix):
    fx = np.load(FIXTURE)
    ref, cols = fx["rows_" + name], list(fx["columns_" + name])
    assert cols == sc.columns(name)
    workers = 4 if suffix else max(1, min(8, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(workers) as pool:
        out = list(pool.map(lambda s: run_one(name, suffix, s), range(1, REPLICAS + 1)))
    for header, rows, nsnap in out:
        assert header == cols
        assert rows.shape == ref.shape[1:] and nsnap == rows.shape[0]
    got = np.stack([rows for _, rows, _ in out])
    assert np.allclose(got[:, :, 0], ref[:, :, 0], atol=1e-5)
    zs = np.stack([z_scores(got[:, :, c], ref[:, :, c]) for c in range(1, len(cols))])  # (columns, points)
    k = zs.size
    worst = np.unravel_index(np.argmax(np.abs(zs)), zs.shape)
    report = "%s %s: %d comparisons, %d beyond 3 sigma, worst %s at %s = %g: z = %.2f" % (
        name, suffix or "CPU", k, int((np.abs(zs) > 3).sum()), cols[1 + worst[0]], cols[0], ref[0, worst[1], 0],
        zs[worst])
    print(report)
    assert int((np.abs(zs) > 3).sum()) <= 1 + k // 50, report
    assert np.abs(zs).max() < 5, report
