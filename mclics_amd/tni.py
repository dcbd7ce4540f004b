"""Locating the nematic-isotropic transition T_NI of bulk Lebwohl-Lasher lattices with plain Metropolis runs.

The uniform scan (the reference's protocol, paper/examples/run.sh:97-129): one lattice per seed cooled through a
uniform temperature grid over [1.02, 1.24] from random directors (or heated from directors along z), directors and
ang_var carried from point to point as thermal mode does, MCT thermalization and MCS measurement sweeps per point.
T_NI is the vertex of the parabola through the three largest values of C_V = L^3 varE / (4 T^2)
(paper/examples/make_figures.py:444-451).

The guided scan: 32 temperature points placed by the surrogate (surrogate.py) around its predicted peak of the
susceptibility, one independent lattice per point, all points of a scan side by side in one batched context. Each
lattice starts as two slabs, the lower half (k < L/2) warm-started from the network's predicted nematic order and
the upper half from its predicted isotropic order, and the plain dynamics decide which phase grows: the drift of
the order parameter S over the run is positive below T_NI and negative above it, and T_NI is where it changes sign
(direct coexistence). A C_V peak over independently warm-started points would instead report wherever the warm
starts switch phase, since at this size neither phase converts within a run. When the sign change is not inside the
window the window is moved and run again (the plain method as fallback); every sweep is counted.

    python -m mclics_amd.tni experiment --net net.npz --L 96 --seeds 1 2 3 4 --out tni.json
"""
import argparse
import json
import sys
import time

import numpy as np

from . import batch, capi

T_NI_LITERATURE = 1.1232  # README.md:284 (Fabbri and Zannoni 1986), the value make_figures.py compares with


def log(m):
    print(m, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------------------------
# Estimators

def cv_peak(T, varE, L):
    """C_V = L^3 varE / (4 T^2) (E counts every bond twice) and the vertex of the parabola through its three
    largest values; a maximum at an end of the grid is returned as is (make_figures.py:444-451). T is a uniform
    grid in either order."""
    T = np.asarray(T, dtype=np.float64)
    order = np.argsort(T)
    T, v = T[order], np.asarray(varE, dtype=np.float64)[order]
    cv = float(L) ** 3 * v / (4.0 * T * T)
    i = int(np.argmax(cv))
    if 0 < i < len(cv) - 1:
        y0, y1, y2 = cv[i - 1:i + 2]
        h = T[1] - T[0]
        return cv, float(T[i] + 0.5 * h * (y0 - y2) / (y0 - 2 * y1 + y2))
    return cv, float(T[i])


def coexistence_point(T, drift, half=3):
    """Temperature where the nematic drift changes sign (positive below, negative above).

    The split between neighbouring points with the fewest points on the wrong side is located first (ties: the
    middle of the tied splits); a straight line fitted to the `half` points on each side of it then gives the zero.
    A split with fewer than two points on one side is no bracket: one noisy point at the window's edge does not
    make a sign change. Returns (T0, bracketed, side): bracketed is False when every drift has one sign or the
    split lies at an edge, and side is then +1 when T_NI lies above the window (drifts positive) and -1 when it
    lies below."""
    T = np.asarray(T, dtype=np.float64)
    d = np.asarray(drift, dtype=np.float64)
    order = np.argsort(T)
    T, d = T[order], d[order]
    n = len(T)
    if np.all(d > 0):
        return float(T[-1]), False, +1
    if np.all(d < 0):
        return float(T[0]), False, -1
    cost = np.array([np.sum(d[:j + 1] < 0) + np.sum(d[j + 1:] > 0) for j in range(n - 1)])
    ties = np.flatnonzero(cost == cost.min())
    j = int(ties[len(ties) // 2])
    if j + 1 < 2:  # at most one point below the split: T_NI lies below the window
        return float(T[0]), False, -1
    if n - (j + 1) < 2:  # at most one point above it: T_NI lies above the window
        return float(T[-1]), False, +1
    split = 0.5 * (T[j] + T[j + 1])
    lo, hi = max(0, j + 1 - half), min(n, j + 1 + half)
    b, a = np.polyfit(T[lo:hi], d[lo:hi], 1)
    if b < 0:
        root = -a / b
        if T[lo] <= root <= T[hi - 1]:
            return float(root), True, 0
    return float(split), True, 0


# ---------------------------------------------------------------------------------------------------------------
# Runs

def uniform_grid(dT, hi=1.24, lo=1.02):
    """Cooling grid hi, hi - dT, ... down to lo (paper/examples/run.sh:102: Ti 1.24, Tf 1.02)."""
    n = int(np.floor((hi - lo) / dT + 1e-9)) + 1
    return hi - dT * np.arange(n)


def bulk_ll(L, seed, T, ni=None):
    spec = dict(L=L, geometry="bulk", T=float(T), seed=int(seed), A=1.0, field=0.0)
    return batch.build_lattice(spec, ic=ni)


def _engine(lats, graph):
    eng = capi.Engine(lats)
    eng.autotune(5)
    if graph:
        eng.set_graph(graph)
    return eng


def _read(eng, n):
    out = np.zeros((n, 4))
    for r in range(n):
        m = eng.moments(r)
        S1, E1 = m["S"] / m["n"], m["E"] / m["n"]
        out[r] = (S1, m["S2"] / m["n"] - S1 * S1, E1, m["E2"] / m["n"] - E1 * E1)
    return out


def uniform_scan(L, temps, seeds, mct, mcs, every=2, heating=False, graph=50):
    """Thermal-mode scans over temps, one lattice per seed, all seeds in one context; cooling scans start from
    random directors, heating scans from directors along z. Returns (S, varS, E, varE) of shape (seeds, points, 4)
    and the lattice sweeps per scan."""
    temps = [float(np.float32(t)) for t in temps]
    hom = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), L ** 3) if heating else None
    lats = [bulk_ll(L, s, temps[0], hom) for s in seeds]
    res = np.zeros((len(seeds), len(temps), 4))
    with _engine(lats, graph) as eng:
        for p, T in enumerate(temps):
            eng.set_temperature(T)
            eng.sweeps(mct)
            eng.reset_moments()
            eng.sweeps(mcs, every)
            res[:, p] = _read(eng, len(seeds))
    return res, len(temps) * (mct + mcs)


def slab_directors(L, lower, upper):
    """Directors of a two-slab start: `lower` (array of L^3/2 directors, sites with k < L/2) then `upper`."""
    return np.concatenate([np.asarray(lower, dtype=np.float32).reshape(-1),
                           np.asarray(upper, dtype=np.float32).reshape(-1)])


def coexistence_runs(L, temps, ics, seeds, relax=1000, windows=5, window=800, every=2, graph=50):
    """Independent lattices, lattice k at temps[k] from directors ics[k], all in one context: `relax` sweeps, then
    `windows` measurement windows of `window` sweeps. The drift is the least-squares slope of the window means of S
    per window. Returns (drift, window means (lattices, windows, 4)) and the lattice sweeps per lattice."""
    temps = [float(np.float32(t)) for t in temps]
    lats = [bulk_ll(L, s, T, ni) for T, ni, s in zip(temps, ics, seeds)]
    win = np.zeros((len(lats), windows, 4))
    with _engine(lats, graph) as eng:
        eng.sweeps(relax)
        for w in range(windows):
            eng.reset_moments()
            eng.sweeps(window, every)
            win[:, w] = _read(eng, len(lats))
    t = np.arange(windows, dtype=np.float64)
    drift = np.polyfit(t, win[:, :, 0].T, 1)[0]
    return drift, win, relax + windows * window


def guided_window(sur, L, n=32, spacing=1e-3, center=None, margin=0.03):
    """Points and warm-start orders of one guided window: n points of the given spacing centred on the network's
    predicted chi peak (or on `center`), and for each point the predicted S at min(T, T_peak - margin) (nematic
    slab) and at max(T, T_peak + margin) (isotropic slab)."""
    from . import surrogate
    t_peak = surrogate.predicted_peak(sur, L)[0]
    c = t_peak if center is None else center
    temps = c + spacing * (np.arange(n) - (n - 1) / 2.0)
    temps = np.array([float(np.float32(t)) for t in temps])
    N = float(L) ** 3
    s_nem = sur.predict(surrogate.features(np.minimum(temps, t_peak - margin), 1.0, 0.0, 0.0, N))[:, 0]
    s_iso = sur.predict(surrogate.features(np.maximum(temps, t_peak + margin), 1.0, 0.0, 0.0, N))[:, 0]
    return temps, s_nem, s_iso, t_peak


def guided_scan(sur, L, seeds, n=32, spacing=1e-3, max_moves=3, warm=True, **run):
    """Guided coexistence scans, one per seed, the scans of all seeds side by side in one context. warm=False
    starts the slabs from directors along z and random directors instead of the network's orders. A scan whose
    window does not contain the sign change is moved by a window width towards it and run again. Returns a dict
    with per-seed estimates, sweeps per seed and the windows run."""
    from . import surrogate
    temps0, s_nem0, s_iso0, t_peak = guided_window(sur, L, n, spacing)
    pt = np.ones(L ** 3 // 2, dtype=np.int32)
    state = {s: dict(center=float(np.mean(temps0)), sweeps=0, moves=0, done=False, est=None, windows=[])
             for s in seeds}
    for attempt in range(max_moves + 1):
        todo = [s for s in seeds if not state[s]["done"]]
        if not todo:
            break
        ics, all_T, all_seeds, owner = [], [], [], []
        for s in todo:
            temps, s_nem, s_iso, _ = guided_window(sur, L, n, spacing, center=state[s]["center"])
            for k, T in enumerate(temps):
                sk = int(s) * 1000003 + 7919 * attempt + k
                if warm:
                    lo = surrogate.warm_start(pt, s_nem[k], sk)
                    hi = surrogate.warm_start(pt, s_iso[k], sk + 500009)
                else:
                    lo = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), len(pt))
                    hi = batch.random_directors(pt, sk + 500009)
                ics.append(slab_directors(L, lo, hi))
                all_T.append(T)
                all_seeds.append(sk)
                owner.append(s)
        drift, _, per = coexistence_runs(L, all_T, ics, all_seeds, **run)
        del ics
        owner = np.array(owner)
        all_T = np.array(all_T)
        for s in todo:
            sel = owner == s
            st = state[s]
            st["sweeps"] += int(sel.sum()) * per
            st["windows"].append((float(all_T[sel].min()), float(all_T[sel].max())))
            t0, ok, side = coexistence_point(all_T[sel], drift[sel])
            if ok:
                st.update(done=True, est=t0)
            elif attempt < max_moves:
                st["moves"] += 1
                st["center"] += side * n * spacing
            else:
                st.update(done=True, est=None)
    return {"t_peak_pred": t_peak, "seeds": list(seeds), "estimates": [state[s]["est"] for s in seeds],
            "sweeps": [state[s]["sweeps"] for s in seeds], "moves": [state[s]["moves"] for s in seeds],
            "windows": [state[s]["windows"] for s in seeds]}


def reference_protocol(L, seeds, dT=0.004, mct=4000, mcs=16000, every=2, graph=50):
    """The reference's LL bulk study at one size (paper/examples/run.sh:97-129, make_figures.py:494-545): per seed a
    cooling scan from random directors and a heating scan from directors along z over [1.02, 1.24] with step dT,
    MCT 4000 + MCS 16000 sweeps per point, both in one context; T_NI is the midpoint of the two C_V peaks. Returns
    per-seed (cooling, heating, midpoint) estimates and the sweeps per estimate (both scans)."""
    grid = uniform_grid(dT)
    temps = [float(np.float32(t)) for t in grid]
    hom = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), L ** 3)
    n = len(seeds)
    lats = [bulk_ll(L, s, temps[0]) for s in seeds] + [bulk_ll(L, 1000 + s, temps[-1], hom) for s in seeds]
    res = np.zeros((2 * n, len(temps), 4))
    with _engine(lats, graph) as eng:
        for p in range(len(temps)):
            for r in range(2 * n):
                eng.set_temperature(temps[p] if r < n else temps[len(temps) - 1 - p], r)
            eng.sweeps(mct)
            eng.reset_moments()
            eng.sweeps(mcs, every)
            res[:, p] = _read(eng, 2 * n)
    cool = [cv_peak(temps, r[:, 3], L)[1] for r in res[:n]]
    heat = [cv_peak(temps[::-1], r[:, 3], L)[1] for r in res[n:]]
    return cool, heat, [0.5 * (c + h) for c, h in zip(cool, heat)], 2 * len(temps) * (mct + mcs)


def uniform_coexistence(L, temps, seeds, **run):
    """The coexistence estimate without the network: points of a uniform grid, slabs from directors along z and
    random directors. Returns per-seed estimates and the sweeps per seed."""
    pt = np.ones(L ** 3 // 2, dtype=np.int32)
    hom = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), len(pt))
    ests = []
    per_seed = 0
    for s in seeds:
        ics, sk = [], []
        for k, T in enumerate(temps):
            sk.append(int(s) * 1000003 + k)
            ics.append(slab_directors(L, hom, batch.random_directors(pt, sk[-1] + 500009)))
        drift, _, per = coexistence_runs(L, temps, ics, sk, **run)
        t0, ok, _ = coexistence_point(temps, drift)
        ests.append(t0 if ok else None)
        per_seed = len(temps) * per
    return ests, per_seed


# ---------------------------------------------------------------------------------------------------------------
# The experiment

def within(ests, tol, ref=T_NI_LITERATURE):
    return all(e is not None and abs(e - ref) <= tol for e in ests)


def experiment(net, L=96, seeds=(1, 2, 3, 4), dTs=(0.004, 0.002, 0.001, 0.0005), npp=5000, tol=0.002,
               data_updates=None, ablations=True):
    """The comparison: guided scans with the trained network; uniform cooling scans at each dT with npp sweeps per
    point (1/5 of them thermalization), the guided scan's budget per point; the reference's own protocol (cooling
    and heating, 20000 sweeps per point); and, with `ablations`, the guided scan without warm starts, with an
    untrained network, and the coexistence estimate on uniform grids without the network. An estimate counts when
    every seed's value lies within tol of the literature T_NI."""
    from . import surrogate
    sur = surrogate.Surrogate.load(net)
    out = {"L": L, "seeds": list(seeds), "tolerance": tol, "reference_T_NI": T_NI_LITERATURE,
           "sweeps_per_point": npp}
    run = dict(relax=npp // 5, windows=5, window=(npp - npp // 5) // 5)
    t0 = time.perf_counter()
    g = guided_scan(sur, L, seeds, **run)
    g["seconds"] = time.perf_counter() - t0
    g["within"] = within(g["estimates"], tol)
    out["guided"] = g
    log("guided: %s (%.1f s)" % (json.dumps(g), g["seconds"]))
    # the same window at half the points (spacing 0.002): the placement alone must then carry the estimate
    t0 = time.perf_counter()
    g16 = guided_scan(sur, L, seeds, n=16, spacing=2e-3, **run)
    g16["seconds"] = time.perf_counter() - t0
    g16["within"] = within(g16["estimates"], tol)
    out["guided_16"] = g16
    log("guided, 16 points: %s (%.1f s)" % (json.dumps(g16), g16["seconds"]))
    mct, mcs = npp // 5, npp - npp // 5
    out["uniform"] = []
    for dT in dTs:
        t0 = time.perf_counter()
        grid = uniform_grid(dT)
        res, sweeps = uniform_scan(L, grid, seeds, mct, mcs)
        est = [cv_peak(grid, r[:, 3], L)[1] for r in res]
        u = {"dT": dT, "points": len(grid), "sweeps": sweeps, "estimates": est, "within": within(est, tol),
             "seconds": time.perf_counter() - t0}
        out["uniform"].append(u)
        log("uniform cooling: %s" % json.dumps(u))
    t0 = time.perf_counter()
    cool, heat, mid, sweeps = reference_protocol(L, seeds)
    ref = {"dT": 0.004, "sweeps_per_point": 20000, "sweeps": sweeps, "cooling": cool, "heating": heat,
           "estimates": mid, "within": within(mid, tol), "seconds": time.perf_counter() - t0}
    out["reference_protocol"] = ref
    log("reference protocol (cooling + heating, 20000 sweeps per point): %s" % json.dumps(ref))
    gs = max(g["sweeps"])
    out["guided_sweeps"] = gs
    passing = [u for u in out["uniform"] if u["within"]]
    base = max(passing, key=lambda u: u["dT"]) if passing else None
    out["baseline_same_sweeps_per_point"] = base
    if base is None:  # no uniform scan at npp sweeps per point reached the tolerance: a lower bound
        out["ratio_same_sweeps_per_point_at_least"] = max(u["sweeps"] for u in out["uniform"]) / gs
    else:
        out["ratio_same_sweeps_per_point"] = base["sweeps"] / gs
    out["ratio_reference_protocol"] = sweeps / gs if ref["within"] and g["within"] else None
    gs16 = max(g16["sweeps"])
    out["guided_16_sweeps"] = gs16
    out["ratio_reference_protocol_16"] = sweeps / gs16 if ref["within"] and g16["within"] else None
    if data_updates is not None:
        eq = data_updates / float(L) ** 3
        out["training_sweeps_equivalent"] = eq
        out["ratio_reference_protocol_with_training"] = sweeps / (gs + eq) if ref["within"] and g["within"] else None
        out["ratio_reference_protocol_16_with_training"] = (sweeps / (gs16 + eq) if ref["within"] and g16["within"]
                                                            else None)
    if ablations:
        t0 = time.perf_counter()
        cold = guided_scan(sur, L, seeds, warm=False, **run)
        cold["within"] = within(cold["estimates"], tol)
        out["guided_no_warm_start"] = cold
        log("guided without warm starts: %s (%.1f s)" % (json.dumps(cold), time.perf_counter() - t0))
        t0 = time.perf_counter()
        raw = surrogate.Surrogate(sur.xm, sur.xs, sur.ym, sur.ys, seed=12345, xlo=sur.xlo, xhi=sur.xhi)
        un = guided_scan(raw, L, seeds, **run)
        un["within"] = within(un["estimates"], tol)
        out["guided_untrained"] = un
        log("guided with an untrained network: %s (%.1f s)" % (json.dumps(un), time.perf_counter() - t0))
        out["uniform_coexistence"] = []
        for n in (16, 32, 40, 48, 56, 111):
            t0 = time.perf_counter()
            grid = np.linspace(1.02, 1.24, n)
            est, sweeps = uniform_coexistence(L, grid, seeds, **run)
            u = {"points": n, "spacing": float(grid[1] - grid[0]), "sweeps": sweeps, "estimates": est,
                 "within": within(est, tol), "seconds": time.perf_counter() - t0}
            out["uniform_coexistence"].append(u)
            log("coexistence on a uniform grid: %s" % json.dumps(u))
        # the network's own gain: the fewest uniform points at which the same estimator meets the tolerance for
        # every seed, against the guided scan's points
        ok = [u for u in out["uniform_coexistence"] if u["within"]]
        if ok:
            least = min(ok, key=lambda u: u["points"])
            out["network_gain_points_32"] = least["points"] / 32.0 if g["within"] else None
            out["network_gain_points_16"] = least["points"] / 16.0 if g16["within"] else None
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    u = sub.add_parser("uniform", help="uniform thermal-mode scans and their C_V-peak estimates")
    u.add_argument("--L", type=int, default=96)
    u.add_argument("--dT", type=float, default=0.004)
    u.add_argument("--seeds", type=int, nargs="+", default=[1, 2, 3, 4])
    u.add_argument("--npp", type=int, default=5000, help="sweeps per point (1/5 thermalization)")
    u.add_argument("--heating", action="store_true")
    u.add_argument("--out", default=None)
    e = sub.add_parser("experiment", help="guided against uniform scans")
    e.add_argument("--net", required=True, help="trained surrogate (surrogate.py train)")
    e.add_argument("--data", default=None, help="training data file, for the sweeps it cost")
    e.add_argument("--L", type=int, default=96)
    e.add_argument("--seeds", type=int, nargs="+", default=[1, 2, 3, 4])
    e.add_argument("--dT", type=float, nargs="+", default=[0.004, 0.002, 0.001, 0.0005])
    e.add_argument("--npp", type=int, default=5000)
    e.add_argument("--no-ablations", action="store_true")
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "uniform":
        grid = uniform_grid(a.dT)
        if a.heating:
            grid = grid[::-1]
        t0 = time.perf_counter()
        res, sweeps = uniform_scan(a.L, grid, a.seeds, a.npp // 5, a.npp - a.npp // 5, heating=a.heating)
        est = [cv_peak(grid, r[:, 3], a.L)[1] for r in res]
        summary = {"L": a.L, "dT": a.dT, "heating": a.heating, "points": len(grid), "sweeps_per_scan": sweeps,
                   "estimates": est, "seconds": time.perf_counter() - t0}
        if a.out:
            np.savez_compressed(a.out, temps=grid, res=res, summary=json.dumps(summary))
        print(json.dumps(summary))
        return
    upd = float(np.load(a.data)["site_updates"]) if a.data else None
    out = experiment(a.net, a.L, a.seeds, a.dT, a.npp, data_updates=upd, ablations=not a.no_ablations)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k.startswith("ratio") or k == "guided_sweeps"}))


if __name__ == "__main__":
    main()
