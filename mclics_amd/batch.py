"""Batched Metropolis runs: many independent lattices of uneven size, geometry and temperature in one context.

This replaces the process-per-seed and per-temperature loops around McLics (tools/average_replicas.py:19-21,
paper/examples/run.sh:157-163): every lattice of a batch is swept by the same kernel launches (one work list over
all lattices), the launch shape comes from the autotuner, and the sweeps are replayed from captured HIP graphs.
Each lattice keeps its own temperature, seed, ang_var and FP64 moments.

    python -m mclics_amd.batch --replicas 512 --mct 1000 --mcs 2000 --every 10 --out batch.npz

The default batch is 512 Lebwohl-Lasher lattices with sizes drawn uniformly from the even integers 16..64, half
periodic bulk and half spheres (radius N/2, homeotropic W = 1.0, free boundaries), temperatures uniform in
[0.9, 1.3] and per-lattice seeds.
"""
import argparse
import functools
import json
import math
import sys
import time

import numpy as np

from . import capi


@functools.lru_cache(maxsize=64)
def sphere_geometry(n):
    """Point types and outward normals of the sphere geometry (geometry_sphere.cpp:32-58): bulk for
    r < Nz/2 - 1, surface 0 for r < Nz/2 + 1, empty outside, centre (N/2, N/2, N/2) with integer division.
    Cached per size; the arrays are read-only."""
    h = np.float32(n // 2)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rx, ry, rz = (i - h).astype(np.float32), (j - h).astype(np.float32), (k - h).astype(np.float32)
    r = np.sqrt(rx * rx + ry * ry + rz * rz).astype(np.float32)
    pt = np.where(r < h - 1, 1, np.where(r < h + 1, 2, 0)).astype(np.int32)
    ns = np.zeros((n, n, n, 3), dtype=np.float32)
    surf = pt == 2
    with np.errstate(invalid="ignore", divide="ignore"):
        ns[..., 0] = np.where(surf, rx / r, 0)
        ns[..., 1] = np.where(surf, ry / r, 0)
        ns[..., 2] = np.where(surf, rz / r, 0)
    pt, ns = pt.reshape(-1), ns.reshape(-1)
    pt.flags.writeable = False
    ns.flags.writeable = False
    return pt, ns


@functools.lru_cache(maxsize=16)
def slab_geometry(nx, ny, nz):
    """Point types and outward normals of the slab geometry (geometry_slab.cpp:51-71): surface 0 (type 2, normal
    -z) on the plane k = 0, surface 1 (type 3, normal +z) on k = Nz - 1, bulk between. Cached; read-only."""
    pt = np.ones((nz, ny, nx), dtype=np.int32)
    ns = np.zeros((nz, ny, nx, 3), dtype=np.float32)
    pt[0], pt[-1] = 2, 3
    ns[0, ..., 2], ns[-1, ..., 2] = -1.0, 1.0
    pt, ns = pt.reshape(-1), ns.reshape(-1)
    pt.flags.writeable = False
    ns.flags.writeable = False
    return pt, ns


def random_directors(pt, seed):
    """Isotropic unit vectors on non-empty sites, (0, 0, 1) on empty ones."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, len(pt))
    phi = rng.uniform(0.0, 2.0 * math.pi, len(pt))
    s = np.sqrt(1.0 - z * z)
    ni = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1).astype(np.float32)
    ni[pt == 0] = (0.0, 0.0, 1.0)
    return ni.reshape(-1)


def c4_specs(n=512, seed=1, lmin=16, lmax=64, tmin=0.9, tmax=1.3):
    """The batched configuration: sizes uniform over even integers, bulk or sphere with probability 1/2,
    temperatures uniform, one seed per lattice."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice(np.arange(lmin, lmax + 1, 2), size=n)
    geoms = rng.choice(["bulk", "sphere"], size=n)
    temps = rng.uniform(tmin, tmax, size=n)
    return [dict(L=int(L), geometry=str(g), T=float(np.float32(T)), seed=seed * 100003 + r, A=1.0, W=1.0,
                 field=0.0)
            for r, (L, g, T) in enumerate(zip(sizes, geoms, temps))]


def build_lattice(spec, ic=None):
    """libmclgpu lattice for a spec dict: L, geometry (bulk | sphere), T, seed, A, W (sphere homeotropic strength),
    field (dielectric coupling along z, elecA)."""
    L = spec["L"]
    field = float(spec.get("field", 0.0))
    model = capi.make_model("ll", A=float(spec.get("A", 1.0)), elec_dir=(0.0, 0.0, 1.0), elecA=field)
    if spec["geometry"] == "sphere":
        pt, ns = sphere_geometry(L)
        ni = random_directors(pt, spec["seed"]) if ic is None else ic
        return capi.Lattice(L, L, L, ni, periodic=(0, 0, 0), pt=pt, ns=ns,
                            surfaces=[("homeotropic", float(spec.get("W", 1.0)), (0, 0, 0))], model=model,
                            T=spec["T"], seed=spec["seed"], radial="sphere")
    pt = np.ones(L ** 3, dtype=np.int32)
    ni = random_directors(pt, spec["seed"]) if ic is None else ic
    return capi.Lattice(L, L, L, ni, periodic=(1, 1, 1), model=model, T=spec["T"], seed=spec["seed"])


def run_batch(specs, mct, mcs, every=1, autotune=True, graph=32, log=None):
    """Thermalizes every lattice for mct sweeps, then measures every `every` sweeps for mcs sweeps.
    Returns (rows, info): rows[r] = (T, S, varS, E, varE, chi, nvalid) with variances over the samples and
    chi = nvalid varS / T; info holds timings and the launch shape."""
    lats = [build_lattice(s) for s in specs]
    nvalid = np.array([int((l.pt > 0).sum()) if l.pt is not None else int(np.prod(l.dims)) for l in lats])
    info = {}
    with capi.Engine(lats) as eng:
        if autotune:
            # trials at the run's cadence: mct plain sweeps, then one measurement every `every` sweeps
            info["tune"] = eng.autotune(10, max(1, int(round(every * (mct + mcs) / max(1, mcs)))) if mcs else 0)
        if graph:
            eng.set_graph(graph)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.sweeps(mct)
        eng.reset_moments()
        eng.sweeps(mcs, every)
        eng.synchronize()
        t1 = time.perf_counter()
        rows = np.zeros((len(specs), 7))
        for r, s in enumerate(specs):
            m = eng.moments(r)
            S1, S2, E1, E2 = m["S"] / m["n"], m["S2"] / m["n"], m["E"] / m["n"], m["E2"] / m["n"]
            varS = S2 - S1 * S1
            rows[r] = (s["T"], S1, varS, E1, E2 - E1 * E1, nvalid[r] * varS / s["T"], nvalid[r])
    info["seconds"] = t1 - t0
    info["sweeps"] = mct + mcs
    info["site_updates"] = float(nvalid.sum()) * (mct + mcs)
    info["site_updates_per_ns"] = info["site_updates"] / (info["seconds"] * 1e9)
    if log:
        log("batch: %d lattices, %.3g site updates in %.3f s = %.2f site-updates/ns" % (
            len(specs), info["site_updates"], info["seconds"], info["site_updates_per_ns"]))
    return rows, info


def combine_replicas(S, varS, E, varE):
    """Replicas of one point combined by the law of total variance (tools/average_replicas.py:86-97):
    mean of the means, Var = mean of the variances + variance of the means (ddof = 0). Arrays: (replicas, points)."""
    S, varS, E, varE = map(np.asarray, (S, varS, E, varE))
    return S.mean(0), varS.mean(0) + S.var(0), E.mean(0), varE.mean(0) + E.var(0)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--mct", type=int, default=1000)
    ap.add_argument("--mcs", type=int, default=2000)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--graph", type=int, default=32, help="sweeps per captured graph (0: eager launches)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    specs = c4_specs(a.replicas, a.seed)
    log("batch: %d lattices, %d sites" % (len(specs), sum(s["L"] ** 3 for s in specs)))
    rows, info = run_batch(specs, a.mct, a.mcs, a.every, autotune=not a.no_autotune, graph=a.graph, log=log)
    if a.out:
        np.savez_compressed(a.out, rows=rows, specs=json.dumps(specs), info=json.dumps(info))
    print(json.dumps({"lattices": len(specs), **info}))


if __name__ == "__main__":
    main()
