"""The Freedericksz campaign of McLics (paper/freedericksz/run.sh) run through the drop-in `mclics` program, as a
physics check of field-sweep (electricGPU) runs of the GHRL potential with finite anchoring.

Three 10 x 10 x 16 slabs (x, y periodic) of the GHRL potential with the E7 constants K1/K2 = 1.48, K3/K2 = 2.28, at
T = 0.0008, each select one deformation at the threshold (paper/freedericksz/README.md, run.sh:54-86):

    splay  planar anchoring along x (rp, W on both plates), field along z, dielectric anisotropy +0.25
    twist  planar anchoring along x,                       field along y, +0.25
    bend   homeotropic anchoring,                           field along z, -0.25

The field is swept from Ei to Ef in steps of dE, MCT sweeps per value before MCS measurement sweeps, one snapshot
per value. The deformation angle is the mean of acos(|n . n0|) over the non-empty sites of the two middle layers
(n0 the initial easy axis), averaged over the seeds (make_figures.py:59-82). The critical field is where the line
fitted to theta^2 versus E above the onset reaches the mean theta^2 of the points well below it
(make_figures.py:108-128); E_c^2 ~ K_i gives K1/K2 = (E_c splay / E_c twist)^2 and K3/K2 = (E_c bend / E_c twist)^2.
Both are compared with the input constants and with the continuum theory for the same cell,
E_c = pi / (d + 2b) sqrt(K* / (2 deps)), K* = Lambda* K_i / K2, Lambda* = 9 / |K1/K2 - 3 - K3/K2|, b = K* / (2W),
d = Nz - 1 (make_figures.py:85-105), which has no free parameter.

    python -m mclics_amd.freedericksz run --out runs            # the reference's protocol, 2 seeds, all cells at once
    python -m mclics_amd.freedericksz analyze runs --json out.json
    python -m mclics_amd.freedericksz run --batched --json out.json   # the same runs in one batched context
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MCLICS = os.path.join(HERE, "bin", "mclics")

# elastic constants (run.sh:67) and the axis of each cell's initial easy direction n0 (make_figures.py:35)
K11, K22, K33 = 1.48, 1.0, 2.28
CELL_AXIS = {"splay": "x", "twist": "x", "bend": "z"}

# the reference's defaults (run.sh:37-45)
DEFAULTS = dict(T=0.0008, W=10.0, MCT=40000, MCS=1000, Ei=0.1, Ef=1.2, dE=0.005, deps=0.25)


def log(m):
    print(m, file=sys.stderr, flush=True)


def input_lines(cell, seed, T, W, MCT, MCS, Ei, Ef, dE, deps, Nx=10, Ny=10, Nz=16):
    """The input file of one run of the campaign, with the parameters run.sh:71-86 writes."""
    lines = ["seed %d" % seed, "potential ghrl", "k11 %g" % K11, "k22 %g" % K22, "k33 %g" % K33,
             "Nx %d" % Nx, "Ny %d" % Ny, "Nz %d" % Nz, "geometry slab", "xbound periodic", "ybound periodic",
             "evol electricGPU", "Ti %g" % T, "Ei %g" % Ei, "Ef %g" % Ef, "dE %g" % dE, "MCT %d" % MCT,
             "MCS %d" % MCS]
    if cell in ("splay", "twist"):
        lines += ["ic homogeneous", "theta_0 90", "phi_0 0"]
        for s in (0, 1):
            lines += ["anchoring_type %d rp" % s, "W %d %g" % (s, W), "theta_s %d 90" % s, "phi_s %d 0" % s]
        field = (0, 0, 1) if cell == "splay" else (0, 1, 0)
        sign = 1
    elif cell == "bend":
        lines += ["ic homogeneous", "theta_0 0", "phi_0 0"]
        for s in (0, 1):
            lines += ["anchoring_type %d homeotropic" % s, "W %d %g" % (s, W)]
        field, sign = (0, 0, 1), -1
    else:
        raise ValueError("unknown cell %r (splay, twist or bend)" % cell)
    lines += ["elecA %g" % (sign * deps), "elecX %d" % field[0], "elecY %d" % field[1], "elecZ %d" % field[2]]
    return lines


def read_input(run_dir):
    """First value of each key of a run's input file, lower-cased keys (make_figures.py:40-46)."""
    p, W = {}, None
    with open(os.path.join(run_dir, "input_parameters.txt")) as f:
        for line in f:
            w = line.split("#")[0].split()
            if len(w) >= 2:
                p.setdefault(w[0].lower(), w[1])
            if len(w) == 3 and w[0].lower() == "w":
                W = float(w[2])
    p["_W"] = W
    return p


def snapshot_scale(step):
    """The snapshot naming scale of include/sweep.h:13-34: the smallest power of ten >= 100 with scale x step >= 1."""
    s = 100
    while s * abs(step) < 1 - 1e-9 and s < 100000000:
        s *= 10
    return s


def sweep_points(a, b, d):
    """Number of points of a sweep from a to b in steps of d (include/sweep.h:13-34)."""
    return int(math.floor((b - a) / d + 1e-3)) + 1


def sweep_value(a, b, d, n, npoints):
    """Point n of that sweep; the last point is b exactly when the steps reach it."""
    if n == npoints - 1 and abs((b - a) / d - (npoints - 1)) < 1e-3:
        return b
    return a + n * d


def field_values(Ei, Ef, dE):
    """The FP32 field values of an electric run: the FP32 inputs, the schedule in double, each value rounded to FP32."""
    f = np.float32
    a, b, d = float(f(Ei)), float(f(Ef)), float(f(dE))
    n = sweep_points(a, b, d)
    return [float(f(sweep_value(a, b, d, i, n))) for i in range(n)]


# ---------------------------------------------------------------------------------------------------------------
# Analysis

def angle_curve(run_dir, axis):
    """(fields, mid-cell deformation angle in degrees) of one run, from po.dat and the snapshots."""
    p = read_input(run_dir)
    scale = snapshot_scale(float(p["de"]))
    with open(os.path.join(run_dir, "po.dat")) as f:
        next(f)
        fields = [float(l.split()[0]) for l in f if l.strip()]
    a = "xyz".index(axis)
    out = []
    for E in fields:
        idx = int(scale * (E + 1e-3 / scale))
        d = np.loadtxt(os.path.join(run_dir, "director_field_%d.csv" % idx), delimiter=",", skiprows=1, ndmin=2)
        z, n, pt = d[:, 2].astype(int), d[:, 3 + a], d[:, 7].astype(int)
        nz = int(z.max()) + 1
        mid = (z == nz // 2 - 1) | (z == nz // 2) if nz % 2 == 0 else (z == nz // 2)
        sel = mid & (pt > 0)
        out.append(float(np.degrees(np.arccos(np.minimum(1.0, np.abs(n[sel])))).mean()))
    return np.array(fields), np.array(out)


def critical_field(E, th):
    """(E_c, E at 20 degrees) of one angle curve, either None where not determined.

    Baseline: the median angle of the first five fields. Onset: the first field whose angle exceeds
    lo = max(3 x baseline, 2 degrees). E_c: where the least-squares line through theta^2 versus E, over the points at
    or above the onset with lo < theta < 40, reaches the mean theta^2 of the points more than 0.05 below the onset
    (at least three points on each side, else None). E20: linear interpolation at the first crossing of 20 degrees."""
    E, th = np.asarray(E, dtype=np.float64), np.asarray(th, dtype=np.float64)
    lo = max(3.0 * float(np.median(th[:5])), 2.0)
    above = np.nonzero(th > lo)[0]
    if len(above) == 0:
        return None, None
    e20 = None
    for i in range(1, len(E)):
        if th[i - 1] < 20.0 <= th[i]:
            e20 = float(E[i - 1] + (20.0 - th[i - 1]) * (E[i] - E[i - 1]) / (th[i] - th[i - 1]))
            break
    e_on = E[above[0]]
    base = th[E < e_on - 0.05]
    sel = (E >= e_on) & (th > lo) & (th < 40.0)
    if len(base) < 3 or int(sel.sum()) < 3:
        return None, e20
    slope, icpt = np.polyfit(E[sel], th[sel] ** 2, 1)
    return float((np.mean(base ** 2) - icpt) / slope), e20


def continuum_fields(k11, k22, k33, Nz, W, deps):
    """{cell: (E_c with anchoring strength W, E_c with strong anchoring)} of the continuum theory for the slab."""
    lam = 9.0 / abs(k11 / k22 - 3.0 - k33 / k22)
    d = Nz - 1
    out = {}
    for cell, ki in (("splay", k11), ("twist", k22), ("bend", k33)):
        K = lam * ki / k22
        b = K / (2.0 * W)
        root = math.sqrt(K / (2.0 * abs(deps)))
        out[cell] = (math.pi / (d + 2.0 * b) * root, math.pi / d * root)
    return out


def analyze_curves(curves, k11, k22, k33, Nz, W, deps):
    """Critical fields and elastic-constant ratios from {cell: [(fields, angles) per seed]}, and the continuum
    theory for the same slab."""
    res = {"cells": {}}
    for cell in CELL_AXIS:
        if not curves.get(cell):
            continue
        n = min(len(c[0]) for c in curves[cell])
        E = np.asarray(curves[cell][0][0][:n], dtype=np.float64)
        th = np.array([c[1][:n] for c in curves[cell]], dtype=np.float64)
        ec, e20 = critical_field(E, th.mean(axis=0))
        per_seed = [x for x in (critical_field(E, t)[0] for t in th) if x is not None]
        err = float(np.std(per_seed, ddof=1) / math.sqrt(len(per_seed))) if len(per_seed) > 1 else None
        res["cells"][cell] = dict(seeds=len(th), E_c=ec, E_c_err=err, E20=e20, per_seed=per_seed,
                                  fields=E.tolist(), angle=th.mean(axis=0).tolist())
    c = res["cells"]
    if all(k in c and c[k]["E_c"] for k in CELL_AXIS):
        es, et, eb = (c[k]["E_c"] for k in ("splay", "twist", "bend"))
        res["K1/K2"], res["K3/K2"] = (es / et) ** 2, (eb / et) ** 2
        if all(c[k]["E20"] for k in CELL_AXIS):
            s, t, b = (c[k]["E20"] for k in ("splay", "twist", "bend"))
            res["K1/K2 (20 deg)"], res["K3/K2 (20 deg)"] = (s / t) ** 2, (b / t) ** 2
    th = continuum_fields(k11, k22, k33, Nz, W, deps)
    res["continuum"] = {k: dict(finite_anchoring=v[0], strong_anchoring=v[1]) for k, v in th.items()}
    res["continuum K1/K2"] = (th["splay"][0] / th["twist"][0]) ** 2
    res["continuum K3/K2"] = (th["bend"][0] / th["twist"][0]) ** 2
    return res


def analyze(root):
    """analyze_curves over a campaign directory root/<cell>/seed<s>/ of finished `mclics` runs."""
    curves, any_run = {}, None
    for cell, axis in CELL_AXIS.items():
        d = os.path.join(root, cell)
        runs = sorted(os.path.join(d, r) for r in os.listdir(d) if os.path.exists(os.path.join(d, r, "DONE"))) \
            if os.path.isdir(d) else []
        if runs:
            curves[cell] = [angle_curve(r, axis) for r in runs]
            any_run = any_run or runs[0]
    if any_run is None:
        return {"cells": {}}
    p = read_input(any_run)
    return analyze_curves(curves, float(p["k11"]), float(p["k22"]), float(p["k33"]), int(p["nz"]), p["_W"],
                          float(p["eleca"]))


# ---------------------------------------------------------------------------------------------------------------
# The campaign as one batched context

def lab_axis(theta_deg, phi_deg):
    """The easy axis of an rp surface from theta_s, phi_s in degrees, as the host program computes it: FP32 radians
    and the C library's sinf / cosf (called here through ctypes, so that the bits are the same)."""
    import ctypes
    import ctypes.util
    m = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
    m.sinf.restype = m.cosf.restype = ctypes.c_float
    m.sinf.argtypes = m.cosf.argtypes = [ctypes.c_float]
    f = np.float32
    to_rad = f(math.pi / 180)
    th, ph = f(to_rad * f(theta_deg)), f(to_rad * f(phi_deg))
    st = f(m.sinf(th))
    return (float(st * f(m.cosf(ph))), float(st * f(m.sinf(ph))), float(f(m.cosf(th))))


def homogeneous(theta_deg, phi_deg):
    """The homogeneous initial director (ic.cpp:72-83): double sin / cos of the FP32 angles, rounded to FP32."""
    t, p = float(np.float32(theta_deg)) * math.pi / 180, float(np.float32(phi_deg)) * math.pi / 180
    return np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)], dtype=np.float32)


def cell_lattice(cell, seed, T, W, deps, Nx=10, Ny=10, Nz=16):
    """The lattice of one campaign run as `mclics` sets it up from input_lines(cell, seed, ...): the same slab
    geometry, anchoring, GHRL coefficients, initial directors, temperature, ang_var (0.5) and seed."""
    from . import batch, capi
    pt, ns = batch.slab_geometry(Nx, Ny, Nz)
    if cell in ("splay", "twist"):
        n0 = homogeneous(90, 0)
        surf = [("rp", W, lab_axis(90, 0)), ("rp", W, lab_axis(90, 0))]
        field, sign = ((0.0, 0.0, 1.0) if cell == "splay" else (0.0, 1.0, 0.0)), 1.0
    else:
        n0 = homogeneous(0, 0)
        surf = [("homeotropic", W, (0.0, 0.0, 0.0)), ("homeotropic", W, (0.0, 0.0, 0.0))]
        field, sign = (0.0, 0.0, 1.0), -1.0
    c = capi.ghrl_coefficients(K11, K22, K33, 0.0)
    model = capi.make_model("ghrl", ghrl=(c["lam"], c["mu"], c["nu"], c["rho"], c["sigma"]), elec_dir=field,
                            elecA=sign * deps, elecE=0.0)
    return capi.Lattice(Nx, Ny, Nz, np.tile(n0, Nx * Ny * Nz), periodic=(1, 1, 0), pt=pt, ns=ns, surfaces=surf,
                        model=model, T=T, ang_var=0.5, seed=seed)


def mid_angle(ni, axis, Nx, Ny, Nz):
    """Mean acos(|n . n0|) in degrees over the two middle layers (one for odd Nz) of a slab without empty sites."""
    n = np.asarray(ni, dtype=np.float64).reshape(Nz, Ny, Nx, 3)[..., "xyz".index(axis)]
    mid = [Nz // 2 - 1, Nz // 2] if Nz % 2 == 0 else [Nz // 2]
    return float(np.degrees(np.arccos(np.minimum(1.0, np.abs(n[mid])))).mean())


def run_batched(cells=("splay", "twist", "bend"), seeds=(1, 2), Nx=10, Ny=10, Nz=16, keep_directors=False,
                **protocol):
    """All runs of the campaign side by side in one context: one launch per colour pass sweeps every lattice.

    Each lattice is set up as `mclics` sets up its run and keeps the same seed, so its directors follow the run's
    trajectory bit for bit; the electric mode's schedule is the same (field set, MCT sweeps, MCS measured sweeps).
    Returns (fields, {(cell, seed): angles}, {(cell, seed): po.dat rows}, directors per field if kept)."""
    from . import capi
    prm = dict(DEFAULTS)
    prm.update({k: v for k, v in protocol.items() if v is not None})
    runs = [(cell, s) for s in seeds for cell in cells]
    lats = [cell_lattice(c, s, prm["T"], prm["W"], prm["deps"], Nx, Ny, Nz) for c, s in runs]
    fields = field_values(prm["Ei"], prm["Ef"], prm["dE"])
    angles = {r: [] for r in runs}
    po = {r: [] for r in runs}
    kept = []
    t0 = last = time.time()
    with capi.Engine(lats) as eng:
        eng.set_graph(50)
        for ie, E in enumerate(fields):
            eng.set_field(E)
            eng.sweeps(prm["MCT"])
            eng.reset_moments()
            eng.sweeps(prm["MCS"], 1)
            snap = {}
            for r, run in enumerate(runs):
                m = eng.moments(r)
                S1, S2, E1, E2 = m["S"] / m["n"], m["S2"] / m["n"], m["E"] / m["n"], m["E2"] / m["n"]
                po[run].append((E, S1, S2 - S1 * S1, E1, E2 - E1 * E1))
                d = eng.directors(r)
                angles[run].append(mid_angle(d, CELL_AXIS[run[0]], Nx, Ny, Nz))
                if keep_directors:
                    snap[run] = d
            if keep_directors:
                kept.append(snap)
            if time.time() - last > 30:
                last = time.time()
                log("  %.0f s: %d of %d field values" % (last - t0, ie + 1, len(fields)))
    log("batched campaign: %d runs, %d field values, %.0f s" % (len(runs), len(fields), time.time() - t0))
    return fields, angles, po, kept




def report(res):
    lines = []
    for cell, r in res["cells"].items():
        lines.append("%-6s seeds %d  E_c = %s +- %s   E(20 deg) = %s   per seed: %s" % (
            cell, r["seeds"], "%.4f" % r["E_c"] if r["E_c"] else "n/a",
            "%.4f" % r["E_c_err"] if r["E_c_err"] else "n/a", "%.4f" % r["E20"] if r["E20"] else "n/a",
            " ".join("%.4f" % x for x in r["per_seed"])))
    if "K1/K2" in res:
        lines.append("K1/K2 = (E_c splay / E_c twist)^2 = %.3f   (input %.2f)" % (res["K1/K2"], K11 / K22))
        lines.append("K3/K2 = (E_c bend  / E_c twist)^2 = %.3f   (input %.2f)" % (res["K3/K2"], K33 / K22))
    if "K1/K2 (20 deg)" in res:
        lines.append("from the 20-degree fields: K1/K2 = %.3f, K3/K2 = %.3f"
                     % (res["K1/K2 (20 deg)"], res["K3/K2 (20 deg)"]))
    for cell, v in res.get("continuum", {}).items():
        m = res["cells"].get(cell, {}).get("E_c")
        lines.append("continuum %-6s E_c = %.4f (finite anchoring), %.4f (strong anchoring)%s" % (
            cell, v["finite_anchoring"], v["strong_anchoring"],
            "   measured %.4f (%+.1f%%)" % (m, 100 * (m / v["finite_anchoring"] - 1)) if m else ""))
    if "continuum K1/K2" in res:
        lines.append("continuum ratios with this anchoring: K1/K2 -> %.3f, K3/K2 -> %.3f"
                     % (res["continuum K1/K2"], res["continuum K3/K2"]))
    return "\n".join(lines)


# ---------------------------------------------------------------------------------------------------------------
# Runs

def run_campaign(root, cells=("splay", "twist", "bend"), seeds=(1, 2), jobs=6, **protocol):
    """Write root/<cell>/seed<s>/input_parameters.txt and run `mclics` there, `jobs` runs at a time (each run is a
    small lattice that leaves the GPU mostly idle, so the runs share it); runs with DONE are skipped."""
    if not os.path.exists(MCLICS):
        raise RuntimeError("%s is not built (python -m mclics_amd.build)" % MCLICS)
    prm = dict(DEFAULTS)
    prm.update({k: v for k, v in protocol.items() if v is not None})
    todo = []
    for s in seeds:
        for cell in cells:
            d = os.path.join(root, cell, "seed%d" % s)
            if os.path.exists(os.path.join(d, "DONE")):
                continue
            os.makedirs(d, exist_ok=True)
            with open(os.path.join(d, "input_parameters.txt"), "w") as f:
                f.write("\n".join(input_lines(cell, s, **prm)) + "\n")
            todo.append(d)
    log("%d runs to execute, %d at a time" % (len(todo), jobs))
    t0, live, failed, last = time.time(), [], [], time.time()
    while todo or live:
        while todo and len(live) < jobs:
            d = todo.pop(0)
            fout = open(os.path.join(d, "run.log"), "w")
            live.append((d, fout, subprocess.Popen([MCLICS, "input_parameters.txt"], cwd=d, stdout=fout,
                                                   stderr=subprocess.STDOUT)))
            log("[start] %s" % d)
        time.sleep(0.5)
        if time.time() - last > 30:
            last = time.time()
            log("  %.0f s: field values done %s" % (last - t0, " ".join(
                "%s=%d" % (os.path.relpath(d, root), _lines(os.path.join(d, "po.dat")) - 1) for d, _, _ in live)))
        for item in list(live):
            d, fout, proc = item
            if proc.poll() is None:
                continue
            fout.close()
            live.remove(item)
            if proc.returncode == 0:
                open(os.path.join(d, "DONE"), "w").close()
                log("[ok]    %s  %.0f s" % (d, time.time() - t0))
            else:
                failed.append(d)
                log("[FAIL]  %s exit %d (see run.log)" % (d, proc.returncode))
    return failed


def _lines(path):
    try:
        with open(path) as f:
            return sum(1 for _ in f)
    except OSError:
        return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--out", help="campaign directory (runs of mclics)")
    r.add_argument("--cells", nargs="+", default=["splay", "twist", "bend"])
    r.add_argument("--seeds", nargs="+", type=int, default=[1, 2])
    r.add_argument("--jobs", type=int, default=6)
    for k, v in DEFAULTS.items():
        r.add_argument("--" + k, type=type(v), default=None)
    r.add_argument("--json")
    r.add_argument("--batched", action="store_true",
                   help="all runs in one batched context (no mclics processes, no snapshot files)")
    a = sub.add_parser("analyze")
    a.add_argument("root")
    a.add_argument("--json")
    args = ap.parse_args(argv)
    if args.cmd == "run" and args.batched:
        prot = {k: getattr(args, k) for k in DEFAULTS}
        prm = dict(DEFAULTS)
        prm.update({k: v for k, v in prot.items() if v is not None})
        fields, angles, po, _ = run_batched(args.cells, args.seeds, **prot)
        curves = {c: [(fields, angles[(c, s)]) for s in args.seeds] for c in args.cells}
        res = analyze_curves(curves, K11, K22, K33, 16, prm["W"], prm["deps"])
        res["po"] = {"%s/seed%d" % k: v for k, v in po.items()}
    elif args.cmd == "run":
        if not args.out:
            ap.error("run needs --out (or --batched)")
        prot = {k: getattr(args, k) for k in DEFAULTS}
        t0 = time.time()
        failed = run_campaign(args.out, args.cells, args.seeds, args.jobs, **prot)
        log("campaign: %.0f s" % (time.time() - t0))
        if failed:
            return 1
        res = analyze(args.out)
    else:
        res = analyze(args.root)
    print(report(res), flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
