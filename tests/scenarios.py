"""The reference's regression scenarios (tests/regression/inputs/1..7_*.txt of the reference), and one LL sphere of the
batched configuration, as data, shared by the
fixture writer tests/golden/make_scenario_fixtures.py and the tests that run them through the mclics program.

Each scenario keeps the reference's input lines except `evol` (given without the GPU suffix; the tests add it for
the GPU modes) and `seed` (one per replica). Together they cover the paths the throughput and slab cases leave out:
the flip trial at T > 0 (pear sphere), lattices with empty sites (sphere, custom cylinder), the rp / fg and GHRL
anchorings at T > 0, a chiral GHRL with an initial condition from a file, and the step, quench and electric loops.

The cylinder of scenario 2 (23 x 23 x 24, free x/y, periodic z) is built here by the rules of the reference's
tests/regression/make_cylinder.py: bulk within (N - 1)/2 - 1 of the axis, a homeotropic surface shell out to
(N - 1)/2 + 1 with radial normals, empty beyond; the initial directors are two +1/2 disclination lines at +-d from
the axis (d = ((N - 1)/2 - 1)/2) that turn once about the axis over the length L. Both files are written with six
significant digits, as that script writes them.
"""
import math

import numpy as np

SCENARIOS = {
    "1_ll_bulk_thermal": dict(
        mode="thermal", dims=(16, 16, 16),
        lines=["Nx 16", "Ny 16", "Nz 16", "geometry bulk", "xbound periodic", "ybound periodic", "zbound periodic",
               "potential ll", "A 1", "Ti 1.5", "Tf 0.5", "dT -0.25", "MCT 500", "MCS 300"]),
    "2_ghrl_cylinder_custom": dict(
        mode="thermal", dims=(23, 23, 24), files="cylinder",
        lines=["Nx 23", "Ny 23", "Nz 24", "geometry custom", "boundary_file cyl.csv", "xbound free", "ybound free",
               "zbound periodic", "potential ghrl", "k11 1.48", "k22 1", "k33 2.28", "p0 25",
               "anchoring_type 0 homeotropic_ghrl", "W 0 5", "radial cylinder", "ic ic_file", "ic_file helix.csv",
               "Ti 0.375", "Tf 0.125", "dT -0.125", "MCT 500", "MCS 300"]),
    "3_pear_sphere": dict(
        mode="thermal", dims=(21, 21, 21),
        lines=["Nx 21", "Ny 21", "Nz 21", "geometry sphere", "potential pear", "A 1.5", "B1 0.09", "C 0.3",
               "anchoring_type 0 homeotropic", "W 0 1.5", "Ti 1.0", "Tf 0.5", "dT -0.25", "MCT 500", "MCS 300"]),
    "4_slab_rp_fg_step": dict(
        mode="step", dims=(16, 16, 12),
        lines=["Nx 16", "Ny 16", "Nz 12", "geometry slab", "xbound periodic", "ybound periodic", "potential ll",
               "anchoring_type 0 rp", "anchoring_type 1 fg", "W 0 1", "W 1 1", "theta_s 0 90", "phi_s 0 0",
               "Ti 0.5", "MCT 300", "MCS 100", "fn 3"]),
    "5_slab_ghrl_anchoring": dict(
        mode="thermal", dims=(16, 16, 12),
        lines=["Nx 16", "Ny 16", "Nz 12", "geometry slab", "xbound periodic", "ybound periodic", "potential ghrl",
               "k11 1.48", "k22 1", "k33 2.28", "anchoring_type 0 rp_ghrl", "anchoring_type 1 fg_ghrl", "W 0 1",
               "W 1 1", "theta_s 0 90", "phi_s 0 30", "Ti 0.5", "Tf 0.3", "dT -0.1", "MCT 500", "MCS 300"]),
    "6_quench": dict(
        mode="quench", dims=(16, 16, 16),
        lines=["Nx 16", "Ny 16", "Nz 16", "geometry bulk", "xbound periodic", "ybound periodic", "zbound periodic",
               "potential ll", "Ti 1.5", "Tf 0.5", "dT 1.0", "MCT 300", "MCS 100", "fn 2"]),
    # not one of the reference's inputs: a sphere of the batched configuration (LL, homeotropic W = 1.0 on the shell
    # of radius Nz/2, temperatures across [0.9, 1.3])
    "c4_ll_sphere": dict(
        mode="thermal", dims=(24, 24, 24),
        lines=["Nx 24", "Ny 24", "Nz 24", "geometry sphere", "potential ll", "A 1", "anchoring_type 0 homeotropic",
               "W 0 1.0", "Ti 1.3", "Tf 0.9", "dT -0.2", "MCT 500", "MCS 300"]),
    "7_electric": dict(
        mode="electric", dims=(16, 16, 12),
        lines=["Nx 16", "Ny 16", "Nz 12", "geometry slab", "xbound periodic", "ybound periodic", "potential ghrl",
               "k11 1.48", "k22 1", "k33 2.28", "anchoring_type 0 rp_ghrl", "anchoring_type 1 rp_ghrl", "W 0 1",
               "W 1 1", "theta_s 0 90", "phi_s 0 0", "theta_s 1 90", "phi_s 1 0", "elecA 0.25", "elecX 0",
               "elecY 0", "elecZ 1", "Ti 0.1", "Ei 0.2", "Ef 1.0", "dE 0.2", "MCT 500", "MCS 200"]),
}


def values(name):
    """The scenario's input lines as a dict of lower-case names (single values) and per-surface entries."""
    out, surf = {}, {}
    for line in SCENARIOS[name]["lines"]:
        key, rest = line.split(None, 1)
        parts = rest.split()
        if key in ("anchoring_type", "W", "theta_s", "phi_s"):
            surf.setdefault(int(parts[0]), {})[key] = parts[1]
        else:
            out[key.lower()] = rest
    out["surfaces"] = surf
    return out


def cylinder_files(N=23, L=24):
    """(geometry CSV text, initial-condition CSV text) of the cylinder scenario."""
    c = (N - 1) / 2.0
    d = (c - 1) / 2.0
    geom, helix = ["x,y,z,nx,ny,nz,pt"], ["x,y,z,nx,ny,nz"]
    for k in range(L):
        turn = -2 * math.pi * k / L
        ct, st = math.cos(turn), math.sin(turn)
        for j in range(N):
            for i in range(N):
                x, y = i - c, j - c
                r = math.hypot(x, y)
                pt = 1 if r <= c - 1 else (2 if r <= c + 1 else 0)
                if pt:
                    rr = r if r > 0 else 1.0
                    nrm = (x / rr, y / rr, 0.0)
                else:
                    nrm = (0.0, 0.0, 1.0)
                geom.append("%d,%d,%d,%.6g,%.6g,%.6g,%d" % (i, j, k, nrm[0], nrm[1], nrm[2], pt))
                # the frame turned with the helix: each +1/2 line adds half the polar angle about it
                u, w = ct * x + st * y, -st * x + ct * y
                angle = 0.5 * (math.atan2(w, u - d) + math.atan2(w, u + d)) + turn
                helix.append("%d,%d,%d,%.6g,%.6g,0" % (i, j, k, math.cos(angle), math.sin(angle)))
    return "\n".join(geom) + "\n", "\n".join(helix) + "\n"


def parse_table(text, cols):
    """Columns `cols` of a CSV text with a header, as a float array (rows in file order)."""
    lines = text.strip().split("\n")
    head = [h.strip().lower() for h in lines[0].split(",")]
    idx = [head.index(c) for c in cols]
    return np.array([[float(l.split(",")[i]) for i in idx] for l in lines[1:]])


def files(name):
    """Extra input files the scenario reads, {file name: text}."""
    if SCENARIOS[name].get("files") == "cylinder":
        geom, helix = cylinder_files()
        return {"cyl.csv": geom, "helix.csv": helix}
    return {}


def columns(name):
    """po.dat header of the scenario (mode column, moments, then the extras of parameter_order.cpp:249-295)."""
    v = values(name)
    first = {"thermal": "T", "step": "ii", "quench": "ii", "electric": "field"}[SCENARIOS[name]["mode"]]
    pear = v.get("potential") == "pear"
    radial = v.get("radial", "sphere" if v.get("geometry") == "sphere" else "none")
    cols = [first, "S", "varS", "E", "varE"]
    if pear:
        cols += ["P", "dtheta", "divn"]
    if radial in ("sphere", "both"):
        cols += ["Sr_sph", "Pr_sph"] if pear else ["Sr_sph"]
    if radial in ("cylinder", "both"):
        cols += ["Sr_cyl", "Pr_cyl"] if pear else ["Sr_cyl"]
    return cols
