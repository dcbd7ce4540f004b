"""Learned surrogate of the plain Metropolis method: where to sample and how to start, never what is reported.

An MLP (inputs T, A, W, field, log N; hidden 64x64 SiLU; outputs S, E, chi) is trained on the GPU by the project's
own kernels (libmclgpu mcl_mlp_*) on (input, observable) pairs produced by batched plain Metropolis runs. It is
then used to
  - place temperature points around its predicted peak of the susceptibility chi = N varS / T, and
  - warm-start the directors of each point from the predicted order parameter S (a Maier-Saupe distribution
    around z with <P2> = S).
Every point is afterwards run with the plain method, which produces every reported number and measures the
surrogate's error (tni.py). The Metropolis rule itself is never changed.

    python -m mclics_amd.surrogate generate --runs 25 --out data.npz     # batched plain runs -> pairs (12^3-20^3)
    python -m mclics_amd.surrogate train data.npz --out net.npz           # HIP training
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import batch, capi

FEATURES = ("T", "A", "W", "field", "logN")
TARGETS = ("S", "E", "log_chi")


def log(m):
    print(m, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------------------------
# Training data from batched plain Metropolis runs

def random_lattices(n, rng, lmin=12, lmax=20):
    """Lattices of the training distribution, the batched configuration's physics widened over the network's inputs:
    size L uniform over even integers (small lattices: the network only places points and warm-starts, and a query
    beyond the largest trained size is asked at that size), bulk (W = 0) or homeotropic sphere (W uniform in
    [0.5, 1.5]) with probability 1/2,
    nematic coupling A uniform in [0.8, 1.2], dielectric coupling along z uniform in [0, 0.3] on half the lattices,
    and a ladder direction: cooling from random directors or heating from directors along z."""
    out = []
    for _ in range(n):
        L = int(rng.choice(np.arange(lmin, lmax + 1, 2)))
        sphere = rng.random() < 0.5
        field = float(rng.uniform(0.0, 0.3)) if rng.random() < 0.5 else 0.0
        out.append(dict(L=L, geometry="sphere" if sphere else "bulk", A=float(rng.uniform(0.8, 1.2)),
                        W=float(rng.uniform(0.5, 1.5)) if sphere else 0.0, field=field,
                        seed=int(rng.integers(1, 2 ** 31)), T=1.0, heating=bool(rng.random() < 0.5)))
    return out


def features(T, A, W, field, nvalid):
    T = np.asarray(T, dtype=np.float64)
    return np.stack([T, np.full(T.shape, A, dtype=np.float64), np.full(T.shape, W, dtype=np.float64),
                     np.full(T.shape, field, dtype=np.float64), np.full(T.shape, np.log(nvalid))], axis=-1)


def generate(runs=25, lattices=512, rungs=16, mct=150, mcs=450, every=3, seed=1, lmin=12, lmax=20, tmin=0.9,
             tmax=1.3, log=log):
    """Pairs (features, observables) from `runs` batched runs. Each lattice walks a ladder of `rungs` temperatures
    drawn uniformly in [tmin A, tmax A] (downwards from random directors or upwards from directors along z),
    thermal-mode style: directors and ang_var carried from rung to rung, MCT sweeps then MCS sweeps measured every
    `every` sweeps per rung. Returns X (n, 5) = (T, A, W, field, log N), Y (n, 3) = (S, E, log chi) with
    chi = N varS / T, and the total number of site updates spent."""
    rng = np.random.default_rng(seed)
    X, Y = [], []
    updates = 0.0
    for run in range(runs):
        t0 = time.perf_counter()
        specs = random_lattices(lattices, rng, lmin, lmax)
        lats = []
        for s in specs:
            ic = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), s["L"] ** 3) if s["heating"] else None
            lats.append(batch.build_lattice(s, ic=ic))
        nvalid = np.array([int((l.pt > 0).sum()) if l.pt is not None else int(np.prod(l.dims)) for l in lats])
        A = np.array([s["A"] for s in specs])
        ladders = np.sort(rng.uniform(tmin, tmax, size=(lattices, rungs)), axis=1) * A[:, None]
        for r, s in enumerate(specs):
            if not s["heating"]:
                ladders[r] = ladders[r, ::-1]
        ladders = ladders.astype(np.float32).astype(np.float64)
        with capi.Engine(lats) as eng:
            eng.autotune(5)
            eng.set_graph(50)
            for k in range(rungs):
                for r in range(lattices):
                    eng.set_temperature(float(ladders[r, k]), r)
                eng.sweeps(mct)
                eng.reset_moments()
                eng.sweeps(mcs, every)
                for r, s in enumerate(specs):
                    m = eng.moments(r)
                    S1, S2, E1 = m["S"] / m["n"], m["S2"] / m["n"], m["E"] / m["n"]
                    chi = nvalid[r] * max(S2 - S1 * S1, 1e-12) / ladders[r, k]
                    X.append((ladders[r, k], s["A"], s["W"], s["field"], math.log(nvalid[r])))
                    Y.append((S1, E1, math.log(chi)))
        updates += float(nvalid.sum()) * rungs * (mct + mcs)
        if log:
            log("surrogate data: run %d/%d, %d pairs, %.1f s" % (run + 1, runs, len(X), time.perf_counter() - t0))
    return np.array(X), np.array(Y), updates


# ---------------------------------------------------------------------------------------------------------------
# The network

class Surrogate:
    """Normalization + the HIP MLP. Queries are clamped to the range of each input seen in training (a lattice
    larger than the largest trained size is asked about at that size): the network is not trusted outside it."""

    def __init__(self, xm, xs, ym, ys, params=None, seed=1, xlo=None, xhi=None):
        self.xm, self.xs, self.ym, self.ys = (np.asarray(a, dtype=np.float64) for a in (xm, xs, ym, ys))
        self.xlo = None if xlo is None else np.asarray(xlo, dtype=np.float64)
        self.xhi = None if xhi is None else np.asarray(xhi, dtype=np.float64)
        self.net = capi.Mlp(seed=seed)
        if params is not None:
            self.net.set_params(params)

    @classmethod
    def fit(cls, X, Y, epochs=300, batch_size=1024, lr=2e-3, seed=1, val_frac=0.1):
        rng = np.random.default_rng(seed)
        perm = rng.permutation(len(X))
        nval = int(len(X) * val_frac)
        val, tr = perm[:nval], perm[nval:]
        xm, xs = X[tr].mean(0), X[tr].std(0) + 1e-12
        ym, ys = Y[tr].mean(0), Y[tr].std(0) + 1e-12
        sur = cls(xm, xs, ym, ys, seed=seed, xlo=X[tr].min(0), xhi=X[tr].max(0))
        xn = ((X - xm) / xs).astype(np.float32)
        yn = ((Y - ym) / ys).astype(np.float32)
        t0 = time.perf_counter()
        # three stages with a decreasing learning rate
        loss = None
        for frac, f_lr in ((0.5, lr), (0.3, lr / 4), (0.2, lr / 16)):
            loss = sur.net.train(xn[tr], yn[tr], epochs=max(1, int(epochs * frac)), batch=batch_size, lr=f_lr,
                                 seed=seed)
        train_s = time.perf_counter() - t0
        pred = sur.predict(X[val]) if nval else None
        rmse = None if pred is None else np.sqrt(((pred - Y[val]) ** 2).mean(0))
        info = {"train_loss": loss, "train_seconds": train_s, "n_train": int(len(tr)), "n_val": int(nval),
                "val_rmse": None if rmse is None else dict(zip(TARGETS, map(float, rmse)))}
        return sur, info

    def predict(self, X):
        """(S, E, log chi) for raw features X (n, 5)."""
        X = np.asarray(X, dtype=np.float64)
        if self.xlo is not None:
            X = np.clip(X, self.xlo, self.xhi)
        xn = ((X - self.xm) / self.xs).astype(np.float32)
        return self.net.forward(xn).astype(np.float64) * self.ys + self.ym

    def save(self, path, info=None):
        extra = {} if self.xlo is None else {"xlo": self.xlo, "xhi": self.xhi}
        np.savez(path, xm=self.xm, xs=self.xs, ym=self.ym, ys=self.ys, params=self.net.params(),
                 info=json.dumps(info or {}), **extra)

    @classmethod
    def load(cls, path):
        d = np.load(path)
        return cls(d["xm"], d["xs"], d["ym"], d["ys"], params=d["params"],
                   xlo=d["xlo"] if "xlo" in d else None, xhi=d["xhi"] if "xhi" in d else None)


# ---------------------------------------------------------------------------------------------------------------
# Placement and warm start

def predicted_peak(sur, L, A=1.0, W=0.0, field=0.0, lo=0.95, hi=1.30, step=2e-4):
    """Temperature of the predicted maximum of chi for an L^3 lattice, and the predicted curves."""
    T = np.arange(lo, hi + step / 2, step)
    pred = sur.predict(features(T, A, W, field, float(L) ** 3))
    k = int(np.argmax(pred[:, 2]))
    return float(T[k]), T, pred


def place_points(sur, L, n=32, spacing=1e-3, **kw):
    """n temperature points on a uniform grid of the given spacing centred on the predicted chi peak."""
    t_peak, _, _ = predicted_peak(sur, L, **kw)
    return t_peak + spacing * (np.arange(n) - (n - 1) / 2.0), t_peak


def _p2_table():
    """<P2(cos)> of f(c) ~ exp(m P2(c)) on [-1, 1] as a function of m (Maier-Saupe)."""
    c = np.linspace(-1.0, 1.0, 4001)
    p2 = 1.5 * c * c - 0.5
    ms = np.linspace(-2.0, 60.0, 1241)
    mean = []
    for m in ms:
        w = np.exp(m * p2 - m)
        mean.append(float((w * p2).sum() / w.sum()))
    return ms, np.array(mean)


_MS, _P2 = _p2_table()


def warm_start(pt, S, seed, axis=(0.0, 0.0, 1.0)):
    """Directors with order parameter <P2> = S about the axis: cos(theta) drawn from exp(m P2(cos)) with m set by
    S (inverse transform on a fine grid), azimuth uniform. S <= 0 gives an isotropic field."""
    rng = np.random.default_rng(seed)
    n = len(pt)
    S = float(np.clip(S, 0.0, 0.995))
    m = float(np.interp(S, _P2, _MS)) if S > 0 else 0.0
    c = np.linspace(-1.0, 1.0, 20001)
    w = np.exp(m * (1.5 * c * c - 0.5) - max(m, 0.0))
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cos = np.interp(rng.random(n), cdf, c)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sin = np.sqrt(np.maximum(0.0, 1.0 - cos * cos))
    v = np.stack([sin * np.cos(phi), sin * np.sin(phi), cos], axis=1)
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    if abs(a[2] - 1.0) > 1e-12:  # rotate z onto the axis (Rodrigues)
        k = np.cross([0.0, 0.0, 1.0], a)
        s, cth = np.linalg.norm(k), a[2]
        k /= s
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - cth) * K @ K
        v = v @ R.T
    v = v.astype(np.float32)
    v[np.asarray(pt) == 0] = (0.0, 0.0, 1.0)
    return v.reshape(-1)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--runs", type=int, default=25)
    g.add_argument("--lattices", type=int, default=512)
    g.add_argument("--rungs", type=int, default=16)
    g.add_argument("--mct", type=int, default=150)
    g.add_argument("--mcs", type=int, default=450)
    g.add_argument("--lmin", type=int, default=12)
    g.add_argument("--lmax", type=int, default=20)
    g.add_argument("--every", type=int, default=3)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("data")
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        X, Y, upd = generate(a.runs, a.lattices, a.rungs, a.mct, a.mcs, a.every, a.seed, a.lmin, a.lmax)
        np.savez_compressed(a.out, X=X, Y=Y, site_updates=upd)
        print(json.dumps({"pairs": len(X), "site_updates": upd}))
    else:
        d = np.load(a.data)
        sur, info = Surrogate.fit(d["X"], d["Y"], epochs=a.epochs)
        sur.save(a.out, info)
        print(json.dumps(info))


if __name__ == "__main__":
    main()
