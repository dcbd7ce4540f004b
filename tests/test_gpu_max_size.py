"""A context at the pool limit: the kernels address the pool with 32-bit byte offsets, so a context holds fewer than
2^30 sites (mcl_create rejects more). 1024 x 1024 x 1022 = 1,071,644,672 sites is the largest box of that shape under
the limit; its director components (3.2e9) and colour blocks (5.4e8 positions) exceed what a 32-bit signed index
reaches, so a host or device index computed in int would show here."""
import numpy as np
import pytest

DIMS = (1024, 1024, 1022)


def _threads():
    import os
    return max(1, min(16, os.cpu_count() or 1))


def ll_energy_mean(ni, dims, A=1.0, chunk=16):
    """Mean LL site energy -1.5 A sum_j (n_i . n_j)^2 of a periodic lattice (every bond counted from both ends), from
    periodic shifts of the director field; FP32 products summed in FP64, over chunks of z-planes in a thread pool
    (numpy releases the interpreter lock in its element-wise loops)."""
    from concurrent.futures import ThreadPoolExecutor
    nx, ny, nz = dims
    n = np.asarray(ni, dtype=np.float32).reshape(nz, ny, nx, 3)

    def part(k0):
        k1 = min(nz, k0 + chunk)
        a = n[k0:k1]
        up = n[np.arange(k0 + 1, k1 + 1) % nz]
        tot = 0.0
        for b in (np.roll(a, -1, axis=2), np.roll(a, -1, axis=1), up):
            c = a[..., 0] * b[..., 0]
            c += a[..., 1] * b[..., 1]
            c += a[..., 2] * b[..., 2]
            tot += float(np.square(c).sum(dtype=np.float64))
        return tot

    with ThreadPoolExecutor(_threads()) as ex:
        tot = sum(ex.map(part, range(0, nz, chunk)))
    return -1.5 * A * 2.0 * tot / (nx * ny * nz)


def test_energy_helper_matches_the_oracle(orc):
    """The helper against the oracle's mean site energy on random periodic lattices (CPU)."""
    rng = np.random.default_rng(3)
    for dims in ((6, 8, 10), (12, 4, 6), (2, 2, 2)):
        nt = dims[0] * dims[1] * dims[2]
        v = rng.normal(size=(nt, 3))
        ni = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32).reshape(-1)
        p = orc.params(*dims, periodic=(1, 1, 1), potential="ll", A=1.3)
        pt = np.ones(nt, dtype=np.int32)
        assert ll_energy_mean(ni, dims, A=1.3) == pytest.approx(orc.energy_mean(p, ni, pt), rel=1e-6)


def structured_directors(dims):
    """Unit vectors (sin t cos f, sin t sin f, cos t) with t varying along y (period 8) and f along x (period 128), the
    same in every z-plane, so that all planes are alike (a random field of 1e9 sites would take minutes to draw)."""
    nx, ny, nz = dims
    assert ny % 8 == 0 and nx % 128 == 0
    f = 2 * np.pi * np.arange(nx) / 128.0
    t = 0.5 + 0.3 * np.cos(2 * np.pi * np.arange(8) / 8.0)
    blk = np.empty((8, nx, 3), dtype=np.float32)
    blk[..., 0] = np.sin(t)[:, None] * np.cos(f)[None, :]
    blk[..., 1] = np.sin(t)[:, None] * np.sin(f)[None, :]
    blk[..., 2] = np.cos(t)[:, None]
    plane = np.tile(blk, (ny // 8, 1, 1))
    ni = np.empty((nz, ny, nx, 3), dtype=np.float32)
    for k in range(nz):
        ni[k] = plane
    return ni.reshape(-1)


@pytest.mark.gpu
def test_lattice_at_the_pool_limit(mcl):
    """1024 x 1024 x 1022 periodic LL at T = 1.0: the energy measured on the GPU equals the host's energy of the same
    directors before and after 4 sweeps, every director stays a unit vector, and the sites at both ends of the pool
    (the first and last planes, the last row) change as often as those of a middle plane."""
    nx, ny, nz = DIMS
    assert 2 * ((nx + 1) // 2) * ny * nz < (1 << 30)
    print("building %dx%dx%d directors" % DIMS, flush=True)
    ni0 = structured_directors(DIMS)
    lat = mcl.Lattice(nx, ny, nz, ni0, periodic=(1, 1, 1), model=mcl.make_model("ll", A=1.0), T=1.0, seed=5)
    print("creating the context", flush=True)
    with mcl.Engine(lat) as eng:
        e0 = eng.measure()
        assert e0["nvalid"] == nx * ny * nz
        print("measured", e0["E"], flush=True)
        assert e0["E"] == pytest.approx(ll_energy_mean(ni0, DIMS), rel=1e-6)
        eng.sweeps(4)
        e1 = eng.measure()
        print("after 4 sweeps", e1["E"], flush=True)
        ni = eng.directors()
    assert ni.shape == ni0.shape
    assert e1["E"] == pytest.approx(ll_energy_mean(ni, DIMS), rel=1e-6)
    n, n0 = ni.reshape(nz, ny, nx, 3), ni0.reshape(nz, ny, nx, 3)
    from concurrent.futures import ThreadPoolExecutor

    def norm_err(k0):
        a = n[k0:k0 + 16]
        return float(np.abs(np.sqrt(a[..., 0] ** 2 + a[..., 1] ** 2 + a[..., 2] ** 2) - 1.0).max())

    with ThreadPoolExecutor(_threads()) as ex:
        assert max(ex.map(norm_err, range(0, nz, 16))) < 1e-5
    changed = lambda k: float(np.mean(np.any(n[k] != n0[k], axis=-1)))
    mid = changed(nz // 2)
    assert mid > 0.2, mid
    for k in (0, 1, nz - 2, nz - 1):
        assert abs(changed(k) - mid) < 0.01, (k, changed(k), mid)
    last_row = float(np.mean(np.any(n[nz - 1, ny - 1] != n0[nz - 1, ny - 1], axis=-1)))
    assert abs(last_row - mid) < 0.1, (last_row, mid)
