"""The mclics program as a drop-in for the reference's modes: the known-answer checks of the reference's
tests/run_checks.py, re-expressed as pytest, in the GPU modes (evol <mode>GPU, run_checks.py --gpu) and in the CPU
modes (evol <mode>, the program's OpenMP sweep).

The input-error checks stop before any device call and run without a GPU; the known-answer checks run once per
mode family (the GPU variants are marked gpu). Each check writes an input file into a fresh folder, runs the program
there and reads po.dat and the snapshots.
"""
import math
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "mclics_amd", "bin", "mclics")


class Run:
    def __init__(self, d, proc):
        self.dir, self.proc = d, proc

    def po(self):
        path = os.path.join(self.dir, "po.dat")
        assert os.path.exists(path), "no po.dat; stdout:\n%s\nstderr:\n%s" % (self.proc.stdout[-2000:],
                                                                            self.proc.stderr[-2000:])
        with open(path) as f:
            lines = [l.split() for l in f if l.strip()]
        return lines[0], [dict(zip(lines[0], map(float, l))) for l in lines[1:]]

    def ok(self):
        assert self.proc.returncode == 0, "exit %d\nstdout:\n%s\nstderr:\n%s" % (
            self.proc.returncode, self.proc.stdout[-3000:], self.proc.stderr[-3000:])
        return self

    def snapshots(self):
        return sorted(f for f in os.listdir(self.dir) if f.startswith("director_field_"))


@pytest.fixture
def run(tmp_path):
    assert os.access(PROGRAM, os.X_OK), "mclics is not built (python -m mclics_amd.build)"
    count = [0]

    def go(params, files=None, setup=None, raw=False, env=None, suffix="GPU"):
        count[0] += 1
        d = tmp_path / ("run%03d" % count[0])
        d.mkdir()
        lines = []
        for k, v in (params.items() if isinstance(params, dict) else params):
            if k.lower() in ("evol", "mode") and not raw and not str(v).lower().endswith("gpu"):
                v = str(v) + suffix
            lines.append("%s %s" % (k, v))
        (d / "input.txt").write_text("\n".join(lines) + "\n")
        for name, text in (files or {}).items():
            (d / name).write_text(text)
        if setup:
            setup(str(d))
        print("mclics run in %s: %s" % (d, dict(params) if isinstance(params, dict) else params), flush=True)
        p = subprocess.run([PROGRAM, "input.txt"], cwd=str(d), stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           universal_newlines=True, errors="replace", timeout=300,
                           env=dict(os.environ, **(env or {})))
        return Run(str(d), p)

    return go


@pytest.fixture(params=[pytest.param("GPU", marks=pytest.mark.gpu, id="gpu"), pytest.param("", id="cpu")])
def krun(request, run):
    """run() in the GPU modes and, as a second case, in the CPU modes."""
    def go(params, **kw):
        return run(params, suffix=request.param, **kw)
    return go


def base(**kw):
    """Small periodic bulk at T = 0 starting from the uniform state along z: one measurement sweep
    (run_checks.py:97-103)."""
    p = dict(Nx=8, Ny=8, Nz=8, geometry="bulk", xbound="periodic", ybound="periodic", zbound="periodic",
             potential="ll", A=1, evol="thermal", Ti=0, Tf=0, dT=-0.1, MCT=0, MCS=1, ic="homogeneous",
             theta_0=0, phi_0=0, seed=1)
    p.update(kw)
    return p


def plus(params, *lines):
    return list(params.items()) + [tuple(l.split(None, 1)) for l in lines]


def close(name, value, expected, tol):
    assert abs(value - expected) <= tol, "%s = %.7g, expected %.7g (tolerance %g)" % (name, value, expected, tol)


# ---- input errors: the run stops, reports the problem and leaves the old output alone (run_checks.py:328-380) ----

def expect_error(run, params, text, status=None, raw=False, files=None):
    r = run(params, files=dict(files or {}, **{"po.dat": "old results\n"}), raw=raw)
    assert r.proc.returncode != 0, "the run did not stop"
    if status is not None:
        assert r.proc.returncode == status
    assert text in r.proc.stderr, r.proc.stderr[-500:]
    with open(os.path.join(r.dir, "po.dat")) as f:
        assert f.read() == "old results\n", "po.dat of the previous run was changed"


def test_error_odd_periodic(run):
    expect_error(run, base(Nx=5), "requires an even number of sites", 2)


def test_error_wrong_dT(run):
    expect_error(run, base(Ti=0.5, Tf=1.0, dT=-0.1), "does not lead from Ti", 1)


def test_error_unknown_parameter(run):
    expect_error(run, plus(base(), "Tff 0.1"), "Unknown parameter", 3)


def test_error_unknown_anchoring(run):
    expect_error(run, plus(base(Nz=6, geometry="slab"), "anchoring_type 0 strong", "W 0 1", "anchoring_type 1 rp",
                           "W 1 1", "theta_s 1 90", "phi_s 1 0"), "not implemented", 2)


def test_error_missing_anchoring(run):
    expect_error(run, plus(base(Nz=6, geometry="slab"), "anchoring_type 0 homeotropic", "W 0 1"),
                 "anchoring_type 1 is not defined", 2)


def test_error_ghrl_constants(run):
    expect_error(run, base(potential="ghrl", k11=5, k22=1, k33=1), "needs k11 < 3 k22 + k33", 1)


def test_error_negative_temperature(run):
    expect_error(run, base(Ti=-0.1, Tf=-0.1), "Ti must be >= 0", 1)


def test_error_invalid_value(run):
    expect_error(run, plus(base(), "MCS many"), "Invalid value for parameter: MCS", 1)


def test_error_radial(run):
    expect_error(run, base(radial="spiral"), "radial must be sphere, cylinder, both or none", 1)


def test_error_unknown_mode(run):
    expect_error(run, base(evol="anneal"), "Evol mode 'annealGPU' not implemented", 2)


def test_error_missing_input_file(run, tmp_path):
    p = subprocess.run([PROGRAM, str(tmp_path / "nope.txt")], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       universal_newlines=True)
    assert p.returncode == 5


def test_error_ic_file_unlisted_site(run):
    r = run(base(Nx=2, Ny=2, Nz=2, ic="ic_file", ic_file="start.csv"),
            files={"start.csv": "x,y,z,nx,ny,nz\n0,0,0,0,0,1\n", "po.dat": "old results\n"})
    assert r.proc.returncode == 3 and "are not listed" in r.proc.stderr
    assert open(os.path.join(r.dir, "po.dat")).read() == "old results\n"


# Format errors of the two column-table files (geometry_custom.cpp:62-163 with status 10, ic.cpp:103-200 with
# status 3): each case is (file text, the message's line number, the message). The table reader reports them as
# "  [ERROR] <file>, line <n>: <message>" and stops before any output file is touched.
GEOM_ERRORS = [
    ("", 1, "empty file (a header line is expected)."),
    ("x,y,nx,ny,nz\n", 1, "missing column(s) in the header: z, pt (required: x,y,z,nx,ny,nz,pt)."),
    ("x,y,z,nx,ny,nz,pt\n\n0,0,0,0,0,1\n", 3, "6 fields, but the header has 7."),
    ("pt, X ,y,z,nx,ny,nz\n1,0,0,0,0,zz,1\n", 2, "column ny is not a number: 'zz'."),
    ("x,y,z,nx,ny,nz,pt\n0,0,0,0,0,1,1.5\n", 2, "x, y, z and pt must be integers."),
    ("x,y,z,nx,ny,nz,pt\n0,4,0,0,0,1,1\n", 2, "point (0, 4, 0) outside the 4 x 4 x 6 grid."),
    ("x,y,z,nx,ny,nz,pt\n0,0,0,0,0,1,-1\n", 2, "negative point type (use 0 empty, 1 bulk, 2+ surface)."),
    ("x,y,z,nx,ny,nz,pt\n1,2,3,0,0,1,1\n1,2,3,0,0,1,1\n", 3, "point (1, 2, 3) is listed twice."),
    ("x,y,z,nx,ny,nz,pt\n0,0,0,0,0,0,2\n", 2, "surface point with a zero normal vector."),
]
IC_ERRORS = [
    ("", 1, "empty file (a header line is expected)."),
    ("x,y,z,nx,ny\n", 1, "missing column(s) in the header: nz (required: x,y,z,nx,ny,nz)."),
    ("x,y,z,nx,ny,nz\n0,0,0,0,0\n", 2, "5 fields, but the header has 6."),
    ("x,y,z,nx,ny,nz\n0,0,0,0,,1\n", 2, "column ny is not a number: ''."),
    ("x,y,z,nx,ny,nz\n0,0.5,0,0,0,1\n", 2, "x, y and z must be integers."),
    ("x,y,z,nx,ny,nz\n0,0,-1,0,0,1\n", 2, "point (0, 0, -1) outside the 4 x 4 x 6 grid."),
    ("x,y,z,nx,ny,nz\n3,1,5,0,0,1\n\n3,1,5,0,0,1\n", 4, "point (3, 1, 5) is listed twice."),
    ("x,y,z,nx,ny,nz\n0,0,0,0,0,0\n", 2, "zero director on a non-empty site."),
]


@pytest.mark.parametrize("text,line,msg", GEOM_ERRORS)
def test_error_custom_geometry_file(run, text, line, msg):
    params = plus(base(Nx=4, Ny=4, Nz=6, geometry="custom"), "boundary_file geom.csv")
    expect_error(run, params, "  [ERROR] geom.csv, line %d: %s\n" % (line, msg), 10, files={"geom.csv": text})


@pytest.mark.parametrize("text,line,msg", IC_ERRORS)
def test_error_ic_file_format(run, text, line, msg):
    expect_error(run, base(Nx=4, Ny=4, Nz=6, ic="ic_file", ic_file="start.csv"),
                 "  [ERROR] start.csv, line %d: %s\n" % (line, msg), 3, files={"start.csv": text})


# ---- known answers in the GPU and CPU modes (run_checks.py:113-323) ----

pytestmark_gpu = pytest.mark.gpu


def test_ground_state_ll(krun):
    r = krun(base()).ok()
    _, rows = r.po()
    close("E", rows[0]["E"], -9.0, 1e-4)
    close("S", rows[0]["S"], 1.0, 1e-5)
    close("varE", rows[0]["varE"], 0.0, 1e-8)


def test_ground_state_ghrl_equal_k(krun):
    _, rows = krun(base(potential="ghrl", k11=1, k22=1, k33=1)).ok().po()
    close("E", rows[0]["E"], -6.0, 1e-4)


def test_ground_state_pear(krun):
    _, rows = krun(base(potential="pear", A=1.5, B1=0.09, C=0.3)).ok().po()
    close("E", rows[0]["E"], 6 * (-1.5 - 0.09), 1e-4)
    close("P", rows[0]["P"], 1.0, 1e-5)


def test_ground_state_slab_homeotropic(krun):
    _, rows = krun(plus(base(Nz=6, geometry="slab"), "anchoring_type 0 homeotropic", "W 0 1",
                       "anchoring_type 1 homeotropic", "W 1 1")).ok().po()
    close("E", rows[0]["E"], (-1.5 * (256 * 6 + 128 * 5) - 128) / 384, 1e-4)


def test_ground_state_field(krun):
    _, rows = krun(base(elecA=0.5, elecX=0, elecY=0, elecZ=2)).ok().po()
    close("E", rows[0]["E"], -9.5, 1e-4)


def hedgehog_csv(n, sign=1):
    c = n // 2
    out = ["x,y,z,nx,ny,nz"]
    for k in range(n):
        for j in range(n):
            for i in range(n):
                v = (i - c, j - c, k - c)
                m = math.sqrt(sum(x * x for x in v))
                u = (0, 0, 1) if m == 0 else tuple(sign * x / m for x in v)
                out.append("%d,%d,%d,%.6f,%.6f,%.6f" % ((i, j, k) + u))
    return "\n".join(out) + "\n"


def test_hedgehog_radial_order(krun):
    p = plus(base(Nx=21, Ny=21, Nz=21, geometry="sphere", potential="pear", A=1.5, B1=0.09, C=0.3, ic="ic_file",
                  ic_file="hedgehog.csv"), "anchoring_type 0 homeotropic", "W 0 1.5")
    header, rows = krun(p, files={"hedgehog.csv": hedgehog_csv(21)}).ok().po()
    for col in ("P", "dtheta", "divn", "Sr_sph", "Pr_sph"):
        assert col in header
    close("Sr_sph", rows[0]["Sr_sph"], 1.0, 0.03)
    close("Pr_sph", rows[0]["Pr_sph"], 1.0, 0.03)
    close("divn", rows[0]["divn"], 0.275, 0.02)


def test_columns_ll_sphere(krun):
    header, _ = krun(plus(base(Nx=11, Ny=11, Nz=11, geometry="sphere", ic="random"), "anchoring_type 0 homeotropic",
                         "W 0 1")).ok().po()
    assert header == ["T", "S", "varS", "E", "varE", "Sr_sph"]


def test_thermal_sweep_points(krun):
    r = krun(base(Nx=4, Ny=4, Nz=4, Ti=1.0, Tf=0.5, dT=-0.1, ic="random")).ok()
    _, rows = r.po()
    T = [row["T"] for row in rows]
    assert len(T) == 6 and all(abs(a - b) <= 1e-6 for a, b in zip(T, [1.0, 0.9, 0.8, 0.7, 0.6, 0.5]))
    assert r.snapshots() == sorted("director_field_%d.csv" % v for v in (100, 90, 80, 70, 60, 50))


def test_thermal_fine_step_names(krun):
    r = krun(base(Nx=4, Ny=4, Nz=4, Ti=1.24, Tf=1.224, dT=-0.004, ic="random")).ok()
    _, rows = r.po()
    assert len(rows) == 5
    assert r.snapshots() == sorted("director_field_%d.csv" % v for v in (1240, 1236, 1232, 1228, 1224))


def test_thermal_step_001_names(krun):
    r = krun(base(Nx=4, Ny=4, Nz=4, Ti=1.2, Tf=1.1, dT=-0.01, ic="random")).ok()
    assert r.snapshots() == sorted("director_field_%d.csv" % v for v in range(110, 121))


def test_electric_sweep(krun):
    header, rows = krun(base(Nx=4, Ny=4, Nz=4, evol="electric", Ti=0.5, elecA=0.1, elecZ=1, Ei=0, Ef=1.2, dE=0.1,
                            ic="random")).ok().po()
    assert header[0] == "field"
    E = [row["field"] for row in rows]
    assert len(E) == 13 and abs(E[-1] - 1.2) <= 1e-6


def test_step_blocks(krun):
    r = krun(base(Nx=4, Ny=4, Nz=4, evol="step", Ti=0.5, fn=3, first_file_number=5, MCT=2, ic="random")).ok()
    header, rows = r.po()
    assert header[0] == "ii" and [row["ii"] for row in rows] == [5, 6, 7]
    assert r.snapshots() == ["director_field_5.csv", "director_field_6.csv", "director_field_7.csv"]


def test_quench_relaxes(krun):
    _, rows = krun(base(evol="quench", Ti=2.0, Tf=0.5, MCT=50, MCQ=300, MCS=20, fn=1, ic="random")).ok().po()
    assert rows[0]["S"] > 0.8


def test_snapshot_as_ic_file(krun):
    r1 = krun(base(Nx=6, Ny=6, Nz=6, Ti=0.3, Tf=0.3, MCT=5, ic="random")).ok()
    snap = os.path.join(r1.dir, "director_field_30.csv")

    def copy(d):
        shutil.copy(snap, os.path.join(d, "start.csv"))
    r2 = krun(base(Nx=6, Ny=6, Nz=6, ic="ic_file", ic_file="start.csv"), setup=copy).ok()

    def directors(path):
        with open(path) as f:
            next(f)
            return [tuple(round(float(x), 5) for x in l.split(",")[3:6]) for l in f]
    assert directors(snap) == directors(os.path.join(r2.dir, "ic.csv"))


def test_same_seed_same_run(krun):
    p = base(Nx=6, Ny=6, Nz=6, Ti=0.8, Tf=0.8, MCT=20, MCS=20, ic="random")
    a, b = krun(p).ok(), krun(p).ok()
    p["seed"] = 2
    c = krun(p).ok()
    po = [open(os.path.join(x.dir, "po.dat")).read() for x in (a, b, c)]
    assert po[0] == po[1]
    assert po[0] != po[2]


def test_snapshot_format(krun):
    """x,y,z,nx,ny,nz,S,pt with x fastest, NaN for S on empty sites (io.cpp:56-89)."""
    r = krun(plus(base(Nx=7, Ny=7, Nz=7, geometry="sphere", ic="random", MCS=2), "anchoring_type 0 homeotropic",
                 "W 0 1")).ok()
    lines = open(os.path.join(r.dir, "ic.csv")).read().splitlines()
    assert lines[0] == "x,y,z,nx,ny,nz,S,pt" and len(lines) == 1 + 7 ** 3
    f1 = lines[1].split(",")
    assert f1[:3] == ["0", "0", "0"] and f1[6] == "nan" and f1[7] == "0"
    assert lines[2].split(",")[:3] == ["1", "0", "0"]
    mid = lines[1 + 3 + 7 * (3 + 7 * 3)].split(",")
    assert mid[:3] == ["3", "3", "3"] and mid[7] == "1" and 0 < float(mid[6]) <= 1.0


def read_ic(r, name="ic.csv"):
    with open(os.path.join(r.dir, name)) as f:
        assert next(f).strip() == "x,y,z,nx,ny,nz,S,pt"
        return [l.strip().split(",") for l in f if l.strip()]


def test_cholesteric_ic(krun):
    """ic cholesteric (ic.cpp:202-225): on every non-empty site phi = phi_0 - 2 pi k / p0 and n = (sin theta_0 cos phi,
    sin theta_0 sin phi, cos theta_0), with the reference's float / double steps; p0_i replaces p0 when set."""
    import numpy as np
    f32 = np.float32
    for extra, pitch in (((), 8.0), (("p0_i 5",), 5.0)):
        r = krun(plus(base(Nx=4, Ny=2, Nz=12, ic="cholesteric", theta_0=70, phi_0=30), "p0 8", *extra)).ok()
        rows = read_ic(r)
        assert len(rows) == 4 * 2 * 12
        phi_0 = f32(30 * math.pi / 180)
        st, ct = math.sin(f32(70) * math.pi / 180), math.cos(f32(70) * math.pi / 180)
        for row in rows:
            k = int(row[2])
            phi = float(f32(phi_0 - 2 * math.pi * k / f32(pitch)))
            want = (f32(st * math.cos(phi)), f32(st * math.sin(phi)), f32(ct))
            for a in range(3):
                close("n%d at k=%d" % (a, k), float(row[3 + a]), float(want[a]), 2e-6)


def test_custom_geometry_ground_state(krun):
    """geometry custom with a boundary file (geometry_custom.cpp:70-163): plane k = 0 empty, plane k = 3 a homeotropic
    surface with normal z (pt 2), the rest unlisted and therefore bulk. Periodic 4 x 4 x 6, uniform along z at T = 0:
    the sites next to the empty plane have 5 bonds (-7.5), the surface plane adds -W, so
    E = (2 (-7.5) + 2 (-9) + (-9 - W)) / 5 per site with W = 2, and S = 1."""
    lines = ["x,y,z,nx,ny,nz,pt"]
    for j in range(4):
        for i in range(4):
            lines.append("%d,%d,0,0,0,1,0" % (i, j))
            lines.append("%d,%d,3,0,0,1,2" % (i, j))
    r = krun(plus(base(Nx=4, Ny=4, Nz=6, geometry="custom"), "boundary_file geom.csv", "anchoring_type 0 homeotropic",
                 "W 0 2"), files={"geom.csv": "\n".join(lines) + "\n"}).ok()
    _, rows = r.po()
    close("E", rows[0]["E"], (2 * -7.5 + 2 * -9 + (-9 - 2)) / 5, 1e-4)
    close("S", rows[0]["S"], 1.0, 1e-5)
    pts = {int(row[2]): int(row[7]) for row in read_ic(r)}
    assert pts == {0: 0, 1: 1, 2: 1, 3: 2, 4: 1, 5: 1}


@pytestmark_gpu
def test_launch_switches_leave_outputs_identical(run):
    """The program's launch switches change how sweeps are launched, never their results: graph replay off
    (MCLICS_GRAPH=0) and resident sweeps off (MCLICS_RESIDENT=0) give the same po.dat and snapshots, byte for
    byte, on a lattice small enough for the resident sweeps (12^3: 864 positions per colour) and a GHRL slab."""
    cases = [base(Nx=12, Ny=12, Nz=12, ic="random", Ti=1.2, Tf=1.0, dT=-0.1, MCT=300, MCS=40),
             plus(base(Nx=10, Ny=10, Nz=16, geometry="slab", zbound="free", ic="random", Ti=0.6, Tf=0.5, dT=-0.1,
                       MCT=200, MCS=30, potential="ghrl"), "k11 1.48", "k22 1", "k33 2.28",
                  "anchoring_type 0 homeotropic", "W 0 2", "anchoring_type 1 homeotropic", "W 1 2")]
    for params in cases:
        outs = []
        for env in ({}, {"MCLICS_RESIDENT": "0"}, {"MCLICS_GRAPH": "0", "MCLICS_RESIDENT": "0"}):
            r = run(params, env=env).ok()
            files = ["po.dat"] + r.snapshots()
            outs.append({f: open(os.path.join(r.dir, f)).read() for f in files})
        assert len(outs[0]) >= 3
        assert outs[0] == outs[1] == outs[2]


def test_cpu_modes_do_not_depend_on_the_thread_count(run):
    """The CPU engine keys every trial's random numbers by (seed, site, sweep) and sums its measurements per plane in
    plane order, so po.dat and every snapshot are byte-identical for any number of OpenMP threads (the reference's
    per-thread generators make its CPU trajectories change with OMP_NUM_THREADS). Cases: a GHRL slab with planar
    and homeotropic anchoring and a field in the electric mode; a pear sphere (flip trial, radial and pear columns)
    in the thermal mode."""
    cases = [plus(base(Nx=8, Ny=6, Nz=10, geometry="slab", zbound="free", ic="random", potential="ghrl", evol="electric",
                       Ti=0.5, elecA=0.3, elecX=1, Ei=0, Ef=0.4, dE=0.2, MCT=20, MCS=10), "k11 1.0", "k22 0.6",
                  "k33 1.4", "anchoring_type 0 rp", "W 0 1.5", "theta_s 0 90", "phi_s 0 30",
                  "anchoring_type 1 homeotropic", "W 1 1"),
             plus(base(Nx=9, Ny=9, Nz=9, geometry="sphere", ic="random", potential="pear", A=1.5, B1=0.09, C=0.3,
                       Ti=0.6, Tf=0.5, dT=-0.1, MCT=20, MCS=10, radial="both"), "anchoring_type 0 homeotropic",
                  "W 0 1")]
    for params in cases:
        outs = []
        for threads in ("1", "3", "8"):
            r = run(params, suffix="", env={"OMP_NUM_THREADS": threads}).ok()
            assert ("Threads:     %s" % threads) in r.proc.stdout
            files = ["po.dat", "ic.csv"] + r.snapshots()
            outs.append({f: open(os.path.join(r.dir, f)).read() for f in files})
        assert len(outs[0]) >= 4
        assert outs[0] == outs[1] == outs[2]
