"""The Freedericksz campaign driver (mclics_amd/freedericksz.py): the critical-field estimator, the continuum theory and
the snapshot reader on constructed inputs (CPU), and a short campaign through the drop-in program (GPU)."""
import math
import os

import numpy as np
import pytest

from mclics_amd import freedericksz as fz


def test_continuum_limits():
    """Strong anchoring: E_c^2 ratios are exactly the elastic-constant ratios and E_c = (pi / d) sqrt(K* / 2 deps)
    with K* = Lambda* K_i / K2; finite anchoring lowers E_c by d / (d + 2b) with b = K* / 2W, and W -> inf recovers
    the strong value."""
    k11, k22, k33, nz, deps = 1.48, 1.0, 2.28, 16, 0.25
    th = fz.continuum_fields(k11, k22, k33, nz, 10.0, deps)
    assert (th["splay"][1] / th["twist"][1]) ** 2 == pytest.approx(k11 / k22, rel=1e-12)
    assert (th["bend"][1] / th["twist"][1]) ** 2 == pytest.approx(k33 / k22, rel=1e-12)
    lam = 9.0 / abs(k11 - 3 - k33)
    assert lam == pytest.approx(2.368421052631579, rel=1e-12)  # 9 / 3.8
    # one-constant case: Lambda* = 3, so K* = 3 and the strong-anchoring field is (pi / 15) sqrt(6)
    one = fz.continuum_fields(1.0, 1.0, 1.0, nz, 10.0, deps)
    for cell in one:
        assert one[cell][1] == pytest.approx(math.pi / 15 * math.sqrt(6.0), rel=1e-12)
        assert one[cell][0] == pytest.approx(one[cell][1] * 15 / (15 + 2 * 3.0 / 20), rel=1e-12)
    far = fz.continuum_fields(k11, k22, k33, nz, 1e12, deps)
    for cell in far:
        assert far[cell][0] == pytest.approx(far[cell][1], rel=1e-9)
        assert th[cell][0] < th[cell][1]
    # the sign of the dielectric anisotropy does not enter (the bend cell uses -deps)
    assert fz.continuum_fields(k11, k22, k33, nz, 10.0, -deps) == th


def test_critical_field_recovers_a_known_threshold():
    """A baseline of 0.5 degrees below E0 and theta^2 = 0.25 + 900 (E - E0) above it: the fitted line meets the
    baseline theta^2 exactly at E0; the 20-degree crossing is interpolated between the bracketing fields."""
    E0 = 0.5437
    E = np.round(np.arange(0.1, 1.2001, 0.005), 6)
    th = np.where(E < E0, 0.5, np.sqrt(0.25 + 900.0 * np.clip(E - E0, 0, None)))
    ec, e20 = fz.critical_field(E, th)
    assert ec == pytest.approx(E0, abs=1e-9)
    i = int(np.nonzero(th >= 20)[0][0])
    want = E[i - 1] + (20 - th[i - 1]) * (E[i] - E[i - 1]) / (th[i] - th[i - 1])
    assert e20 == pytest.approx(want, abs=1e-12) and E[i - 1] < e20 <= E[i]
    # no onset: nothing determined; onset too close to the start: no baseline, E_c undetermined
    assert fz.critical_field(E, np.full(len(E), 0.5)) == (None, None)
    early = np.where(np.arange(len(E)) < 3, 0.5, 25.0)  # onset at the fourth field: no point 0.05 below it
    assert fz.critical_field(E, early)[0] is None


def test_critical_field_rejects_a_noisy_baseline_point():
    """The onset threshold is max(3 x median of the first five angles, 2 degrees): one high point among the first
    five does not move it, and a thermal baseline above 2/3 degree raises it."""
    E = np.round(np.arange(0.1, 1.2001, 0.01), 6)
    th = np.where(E < 0.6, 0.5, np.sqrt(0.25 + 900.0 * np.clip(E - 0.6, 0, None)))
    th2 = th.copy()
    th2[1] = 1.9
    assert fz.critical_field(E, th2)[0] == pytest.approx(fz.critical_field(E, th)[0], abs=0.002)


def _write_run(d, fields, angles, axis, de=0.05, nx=2, ny=2, nz=6):
    """A run directory as mclics writes it: po.dat (field first), one snapshot per field with the middle layers
    tilted by `angle` from the axis and the other layers along it, plus an empty site in a middle layer."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "input_parameters.txt"), "w") as f:
        f.write("dE %g\nNz %d\n" % (de, nz))
    with open(os.path.join(d, "po.dat"), "w") as f:
        f.write("field S varS E varE\n")
        for E in fields:
            f.write("%g 1 0 -1 0\n" % E)
    a = "xyz".index(axis)
    b = (a + 1) % 3
    scale = fz.snapshot_scale(de)
    for E, ang in zip(fields, angles):
        with open(os.path.join(d, "director_field_%d.csv" % int(scale * (E + 1e-3 / scale))), "w") as f:
            f.write("x,y,z,nx,ny,nz,S,pt\n")
            for k in range(nz):
                for j in range(ny):
                    for i in range(nx):
                        n = [0.0, 0.0, 0.0]
                        t = math.radians(ang) if k in (nz // 2 - 1, nz // 2) else 0.0
                        n[a], n[b] = math.cos(t), math.sin(t)
                        if k % 2:
                            n = [-v for v in n]  # the sign of a director does not matter
                        pt = 0 if (i, j, k) == (0, 0, nz // 2) else 1
                        if pt == 0:
                            n = [0.0, 0.0, 1.0]
                        f.write("%d,%d,%d,%g,%g,%g,1,%d\n" % (i, j, k, n[0], n[1], n[2], pt))


def test_angle_curve_reads_snapshots(tmp_path):
    fields, angles = [0.1, 0.15, 0.2], [0.0, 12.5, 61.0]
    for axis in ("x", "z"):
        d = str(tmp_path / axis)
        _write_run(d, fields, angles, axis)
        E, th = fz.angle_curve(d, axis)
        assert list(E) == fields
        np.testing.assert_allclose(th, angles, atol=1e-4)


def test_input_lines_follow_the_campaign():
    """The three cells of paper/freedericksz/run.sh:71-86: planar rp anchoring along x with the field along z (splay)
    or y (twist) and +deps; homeotropic anchoring, field along z and -deps (bend)."""
    kv = lambda lines: {tuple(l.split()[:-1]): l.split()[-1] for l in lines}
    s, t, b = (kv(fz.input_lines(c, 3, **fz.DEFAULTS)) for c in ("splay", "twist", "bend"))
    for p in (s, t, b):
        assert p[("potential",)] == "ghrl" and p[("evol",)] == "electricGPU" and p[("seed",)] == "3"
        assert (p[("k11",)], p[("k22",)], p[("k33",)]) == ("1.48", "1", "2.28")
        assert (p[("Nx",)], p[("Ny",)], p[("Nz",)], p[("geometry",)]) == ("10", "10", "16", "slab")
        assert (p[("Ti",)], p[("Ei",)], p[("Ef",)], p[("dE",)]) == ("0.0008", "0.1", "1.2", "0.005")
        assert (p[("MCT",)], p[("MCS",)]) == ("40000", "1000")
        assert p[("W", "0")] == p[("W", "1")] == "10"
    for p in (s, t):
        assert p[("anchoring_type", "0")] == p[("anchoring_type", "1")] == "rp"
        assert (p[("theta_s", "0")], p[("phi_s", "0")], p[("theta_0",)]) == ("90", "0", "90")
        assert p[("elecA",)] == "0.25"
    assert (s[("elecX",)], s[("elecY",)], s[("elecZ",)]) == ("0", "0", "1")
    assert (t[("elecX",)], t[("elecY",)], t[("elecZ",)]) == ("0", "1", "0")
    assert b[("anchoring_type", "0")] == "homeotropic" and b[("theta_0",)] == "0"
    assert b[("elecA",)] == "-0.25" and (b[("elecX",)], b[("elecY",)], b[("elecZ",)]) == ("0", "0", "1")


@pytest.mark.gpu
def test_short_campaign_through_mclics(tmp_path):
    """Three field values per cell (0.1, 0.65, 1.2) with 2000 + 50 sweeps each through the drop-in program: well
    below every threshold the cells stay undeformed (thermal angle under 2 degrees), at twice the thresholds every
    cell has deformed (above 30 degrees). The snapshot names follow the reference's scale (100 for dE = 0.55)."""
    root = str(tmp_path / "fz")
    print("short Freedericksz campaign in", root, flush=True)
    failed = fz.run_campaign(root, seeds=(1,), jobs=3, MCT=2000, MCS=50, dE=0.55)
    assert failed == []
    for cell, axis in fz.CELL_AXIS.items():
        d = os.path.join(root, cell, "seed1")
        assert os.path.exists(os.path.join(d, "director_field_65.csv"))
        E, th = fz.angle_curve(d, axis)
        np.testing.assert_allclose(E, [0.1, 0.65, 1.2])
        assert th[0] < 2.0, (cell, th)
        assert th[2] > 30.0, (cell, th)


def test_field_values_follow_the_schedule():
    """The electric mode's field values (include/sweep.h:13-34): 221 values 0.1 .. 1.2 in steps of 0.005, the last one
    exactly Ef, each the FP32 rounding of the double schedule."""
    E = fz.field_values(0.1, 1.2, 0.005)
    assert len(E) == 221 and E[-1] == float(np.float32(1.2)) and E[0] == float(np.float32(0.1))
    assert all(np.float32(e) == e for e in E)
    assert E[100] == float(np.float32(float(np.float32(0.1)) + 100 * float(np.float32(0.005))))
    f = lambda v: float(np.float32(v))
    assert fz.field_values(0.1, 1.2, 0.55) == [f(0.1), f(f(0.1) + f(0.55)), f(1.2)]  # a + n d, then the end exactly


def test_setup_matches_the_host_program():
    """The batched runner's lattice inputs are the host program's: the rp axis of theta_s = 90, phi_s = 0 from FP32
    sinf / cosf (x = 1, the FP32 cosine of the rounded pi/2 in z), and the homogeneous directors from double sin / cos
    (z = cos(pi/2) rounded to FP32, not 0)."""
    ax = fz.lab_axis(90, 0)
    assert ax[0] == 1.0 and ax[1] == 0.0 and ax[2] == float(np.float32(math.cos(float(np.float32(math.pi / 2)))))
    n = fz.homogeneous(90, 0)
    assert n[0] == 1.0 and n[1] == 0.0 and n[2] == np.float32(math.cos(math.pi / 2)) and n[2] != 0
    assert list(fz.homogeneous(0, 0)) == [0.0, 0.0, 1.0]


@pytest.mark.gpu
def test_batched_campaign_reproduces_the_mclics_runs(tmp_path):
    """The three cells in one batched context against three `mclics` runs of the same input: every snapshot is the
    same text (the directors are bit-identical), and the po.dat rows agree to the printed digits."""
    root = str(tmp_path / "fz")
    prot = dict(MCT=2000, MCS=50, dE=0.55)
    print("mclics runs in", root, flush=True)
    assert fz.run_campaign(root, seeds=(1,), jobs=3, **prot) == []
    print("batched runs", flush=True)
    fields, angles, po, kept = fz.run_batched(seeds=(1,), keep_directors=True, **prot)
    assert fields == fz.field_values(0.1, 1.2, 0.55)
    for cell in fz.CELL_AXIS:
        d = os.path.join(root, cell, "seed1")
        with open(os.path.join(d, "po.dat")) as f:
            next(f)
            rows = [[float(x) for x in l.split()] for l in f if l.strip()]
        assert len(rows) == len(fields)
        for ie, E in enumerate(fields):
            idx = int(100 * (E + 1e-3 / 100))
            with open(os.path.join(d, "director_field_%d.csv" % idx)) as f:
                next(f)
                cli = [l.split(",")[3:6] for l in f]
            mine = kept[ie][(cell, 1)].reshape(-1, 3)
            assert [["%g" % v for v in row] for row in mine] == cli, (cell, E)
            np.testing.assert_allclose(po[(cell, 1)][ie], rows[ie], rtol=2e-5, atol=1e-9)
        # the same angles, up to the 6 printed digits of the snapshots (acos is steep near |n . n0| = 1)
        np.testing.assert_allclose(angles[(cell, 1)], fz.angle_curve(d, fz.CELL_AXIS[cell])[1], atol=0.01)
