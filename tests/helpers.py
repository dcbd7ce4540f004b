"""Shared helpers of the tests: build the same lattice for the oracle and for libmclgpu."""
import numpy as np


def rand_unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def model_from_orc(capi, p):
    """libmclgpu model with the physics of an oracle orc_params."""
    pot = {0: "ll", 1: "ghrl", 2: "pear"}[p.potential]
    return capi.make_model(pot, A=p.A, B1=p.B1, C=p.C,
                           ghrl=(p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_rho, p.ghrl_sigma),
                           elec_dir=(p.elecX, p.elecY, p.elecZ), elecA=p.elecA, elecP=p.elecP, elecE=p.elecE,
                           flip_move=p.flipMove)


def lattice_from_orc(capi, p, ni, pt, ns, T=None, seed=1, radial=0, ang_var=0.5):
    surfaces = [(int(p.surf_kind[s]), float(p.surf_W[s]), tuple(p.surf_axis[s][a] for a in range(3)))
                for s in range(p.nSurf)]
    has_surf = p.nSurf > 0
    return capi.Lattice(p.Nx, p.Ny, p.Nz, ni, periodic=tuple(p.periodic), pt=pt,
                        ns=ns if has_surf else None, surfaces=surfaces, model=model_from_orc(capi, p),
                        T=p.T if T is None else T, seed=seed, radial=radial, ang_var=ang_var)


def sites_and_trials(pt, n, seed):
    rng = np.random.default_rng(seed)
    valid = np.flatnonzero(pt)
    sites = rng.choice(valid, size=min(n, len(valid)), replace=False).astype(np.int32)
    return sites, rand_unit(rng, len(sites))


def z_scores(a, b):
    """a, b: (replicas, points) samples of one quantity from two samplers -> z per point:
    (mean_a - mean_b) / sqrt(s_a^2 / n_a + s_b^2 / n_b), s the replica-to-replica standard deviation."""
    n_a, n_b = a.shape[0], b.shape[0]
    d = a.mean(0) - b.mean(0)
    s = np.sqrt(a.var(0, ddof=1) / n_a + b.var(0, ddof=1) / n_b)
    s = np.where(s > 0, s, 1e-300)
    return d / s
