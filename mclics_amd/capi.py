"""ctypes binding of libmclgpu (include/mclgpu.h).

The structures mirror the C declarations field by field. `Engine` wraps one context: a set of lattices swept
together on the GPU. It is the Python-side mirror of the reference's EvolveNGPU interface
(include/evolve.cuh:25-28): `sweeps` is Monte_Carlo_Step_GPU repeated (src/evolve.cu:317-440), `measure` is
energy_calculator_GPU + Q_tensor + Eigen_value_evaluation (src/evolve.cu:442-478, src/parameter_order.cpp:19-71).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmclgpu.so")

POT = {"ll": 0, "lebwohl-lasher": 0, "ghrl": 1, "gruhn-hess": 1, "pear": 2}
ANCH = {"homeotropic": 0, "rp": 1, "rp_tangent": 2, "fg": 3, "homeotropic_ghrl": 4, "rp_ghrl": 5, "fg_ghrl": 6}
RADIAL = {"none": 0, "sphere": 1, "cylinder": 2, "both": 3}

MCL_OK = 0


class MclError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libmclgpu error %d: %s" % (code, msg))
        self.code = code


class Surface(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("W", ctypes.c_float), ("axis", ctypes.c_float * 3)]


class Model(ctypes.Structure):
    _fields_ = [("potential", ctypes.c_int32), ("A", ctypes.c_float), ("B1", ctypes.c_float), ("C", ctypes.c_float),
                ("ghrl_lambda", ctypes.c_float), ("ghrl_mu", ctypes.c_float), ("ghrl_nu", ctypes.c_float),
                ("ghrl_rho", ctypes.c_float), ("ghrl_sigma", ctypes.c_float), ("elec_dir", ctypes.c_float * 3),
                ("elecA", ctypes.c_float), ("elecP", ctypes.c_float), ("elecE", ctypes.c_float),
                ("flip_move", ctypes.c_int32)]


class LatticeDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3), ("ni", ctypes.POINTER(ctypes.c_float)),
                ("pt", ctypes.POINTER(ctypes.c_int32)), ("ns", ctypes.POINTER(ctypes.c_float)),
                ("nsurf", ctypes.c_int32), ("surf", ctypes.POINTER(Surface)), ("model", Model),
                ("T", ctypes.c_float), ("ang_var", ctypes.c_float), ("seed", ctypes.c_uint64),
                ("radial", ctypes.c_int32)]


class Moments(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(k, ctypes.c_double) for k in
                                          ("S", "S2", "E", "E2", "Sr_sph", "Pr_sph", "Sr_cyl", "Pr_cyl", "P",
                                           "dtheta", "divn")]


class TuneResult(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int32), ("xcd_order", ctypes.c_int32), ("us_per_sweep", ctypes.c_double),
                ("from_cache", ctypes.c_int32)]


class Observables(ctypes.Structure):
    _fields_ = [("nvalid", ctypes.c_int64), ("E", ctypes.c_double), ("Q", ctypes.c_double * 6),
                ("S", ctypes.c_double), ("director", ctypes.c_double * 3)] + \
               [(k, ctypes.c_double) for k in ("Sr_sph", "Pr_sph", "Sr_cyl", "Pr_cyl", "P", "dtheta", "divn")]


ABI = {
    "mcl_create": (ctypes.c_int, [ctypes.POINTER(LatticeDesc), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "mcl_destroy": (None, [ctypes.c_void_p]),
    "mcl_sweeps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "mcl_set_temperature": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float]),
    "mcl_set_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float]),
    "mcl_set_ang_var": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float]),
    "mcl_get_ang_var": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "mcl_get_sweep_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64)]),
    "mcl_reset_moments": (ctypes.c_int, [ctypes.c_void_p]),
    "mcl_get_moments": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Moments)]),
    "mcl_measure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Observables)]),
    "mcl_get_directors": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "mcl_set_directors": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "mcl_local_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "mcl_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "mcl_last_error": (ctypes.c_char_p, []),
    "mcl_abi_version": (ctypes.c_int, []),
    "mcl_site_energies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "mcl_site_energies_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                                 ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                                 ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "mcl_philox": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_uint32)]),
    "mcl_set_kernel_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "mcl_kernel_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_double)]),
    "mcl_get_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "mcl_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(TuneResult)]),
    "mcl_autotune_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(TuneResult)]),
    "mcl_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "mcl_set_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "mcl_prepare_sweeps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "mcl_set_resident": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "mcl_resident": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "mcl_measure_in_sweep": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "mcl_set_lattice_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "mcl_get_lattice_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "mcl_mlp_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_void_p)]),
    "mcl_mlp_destroy": (None, [ctypes.c_void_p]),
    "mcl_mlp_num_params": (ctypes.c_int, []),
    "mcl_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32]),
    "mcl_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32]),
    "mcl_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_float)]),
    "mcl_mlp_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)]),
    "mcl_mlp_set_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.c_int32]),
    "mcl_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                     ctypes.POINTER(ctypes.c_float), ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]),
    "mcl_mlp_last_error": (ctypes.c_char_p, []),
    "mcl_copy_bandwidth": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libmclgpu and declares every entry point of include/mclgpu.h. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libmclgpu is not built (%s missing): run `python -c 'import __graft_entry__ as g; "
                           "g.build()'` or `python -m mclics_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in ABI.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(code):
    if code != MCL_OK:
        raise MclError(code, load().mcl_last_error().decode(errors="replace"))


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def ghrl_coefficients(k11, k22, k33, p0=0.0):
    """GHRL coefficients from the Frank constants (setGHRL, io.cpp:340-360), in the reference's FP32 arithmetic:
    Scale = |9 / (k11 - 3 k22 - k33)| so that |nu| = 1, lambda = (Scale/9)(2 k11 - 3 k22 + k33),
    mu = Scale (k22 - k11), rho = (Scale/9)(k11 - k33), nu = (Scale/9)(k11 - 3 k22 - k33),
    sigma = -Scale k22 2 pi / p0 (0 without chirality)."""
    f = np.float32
    k11, k22, k33, p0 = f(k11), f(k22), f(k33), f(p0)
    scale = f(abs(9.0 / float(k11 - f(3) * k22 - k33)))
    s9 = scale / f(9)
    sigma = f(-float(scale) * float(k22) * (2 * math.pi / float(p0))) if p0 else f(0)
    return {"lam": float(s9 * (f(2) * k11 - f(3) * k22 + k33)), "mu": float(scale * (k22 - k11)),
            "nu": float(s9 * (k11 - f(3) * k22 - k33)), "rho": float(s9 * (k11 - k33)), "sigma": float(sigma),
            "scale": float(scale)}


def make_model(potential="ll", A=1.0, B1=0.04, C=0.3, ghrl=(0.0, 0.0, 1.0, 0.0, 0.0), elec_dir=(0.0, 0.0, 0.0),
               elecA=0.0, elecP=0.0, elecE=1.0, flip_move=None):
    """Model parameters; ghrl = (lambda, mu, nu, rho, sigma). flip_move defaults to the reference's rule
    (pear or elecP != 0, main.cpp:88)."""
    m = Model()
    m.potential = POT[potential] if isinstance(potential, str) else int(potential)
    m.A, m.B1, m.C = A, B1, C
    m.ghrl_lambda, m.ghrl_mu, m.ghrl_nu, m.ghrl_rho, m.ghrl_sigma = ghrl
    m.elec_dir[:] = list(elec_dir)
    m.elecA, m.elecP, m.elecE = elecA, elecP, elecE
    if flip_move is None:
        flip_move = 1 if (m.potential == 2 or elecP != 0) else 0
    m.flip_move = int(flip_move)
    return m


class Lattice:
    """Host description of one lattice (arrays kept alive while the engine is created)."""

    def __init__(self, nx, ny, nz, ni, periodic=(1, 1, 1), pt=None, ns=None, surfaces=(), model=None, T=1.0,
                 ang_var=0.5, seed=1, radial=0):
        self.dims = (int(nx), int(ny), int(nz))
        nt = nx * ny * nz
        self.ni = np.ascontiguousarray(ni, dtype=np.float32).reshape(3 * nt)
        self.pt = None if pt is None else np.ascontiguousarray(pt, dtype=np.int32).reshape(nt)
        self.ns = None if ns is None else np.ascontiguousarray(ns, dtype=np.float32).reshape(3 * nt)
        self.surfaces = list(surfaces)  # (type, W, axis)
        self.periodic = tuple(int(p) for p in periodic)
        self.model = model if model is not None else make_model()
        self.T, self.ang_var, self.seed = float(T), float(ang_var), int(seed)
        self.radial = RADIAL[radial] if isinstance(radial, str) else int(radial)

    def desc(self):
        d = LatticeDesc()
        d.nx, d.ny, d.nz = self.dims
        d.periodic[:] = list(self.periodic)
        d.ni = _fptr(self.ni)
        d.pt = self.pt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if self.pt is not None else None
        d.ns = _fptr(self.ns) if self.ns is not None else None
        self._surf = (Surface * max(1, len(self.surfaces)))()
        for k, (typ, W, axis) in enumerate(self.surfaces):
            self._surf[k].type = ANCH[typ] if isinstance(typ, str) else int(typ)
            self._surf[k].W = W
            self._surf[k].axis[:] = list(axis)
        d.nsurf = len(self.surfaces)
        d.surf = ctypes.cast(self._surf, ctypes.POINTER(Surface))
        d.model = self.model
        d.T, d.ang_var, d.seed, d.radial = self.T, self.ang_var, self.seed, self.radial
        return d


class Engine:
    """One libmclgpu context holding one or more lattices."""

    def __init__(self, lattices):
        self.lib = load()
        if isinstance(lattices, Lattice):
            lattices = [lattices]
        self.lattices = list(lattices)
        arr = (LatticeDesc * len(self.lattices))(*[l.desc() for l in self.lattices])
        h = ctypes.c_void_p()
        _check(self.lib.mcl_create(arr, len(self.lattices), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mcl_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sweeps(self, n, measure_every=0):
        _check(self.lib.mcl_sweeps(self.h, int(n), int(measure_every)))

    def set_temperature(self, T, lat=-1):
        _check(self.lib.mcl_set_temperature(self.h, lat, T))

    def set_field(self, E, lat=-1):
        _check(self.lib.mcl_set_field(self.h, lat, E))

    def set_ang_var(self, a, lat=-1):
        _check(self.lib.mcl_set_ang_var(self.h, lat, a))

    def ang_var(self, lat=0):
        v = ctypes.c_float()
        _check(self.lib.mcl_get_ang_var(self.h, lat, ctypes.byref(v)))
        return v.value

    def sweep_stats(self, lat=0):
        a, s = ctypes.c_int64(), ctypes.c_int64()
        _check(self.lib.mcl_get_sweep_stats(self.h, lat, ctypes.byref(a), ctypes.byref(s)))
        return a.value, s.value

    def reset_moments(self):
        _check(self.lib.mcl_reset_moments(self.h))

    def moments(self, lat=0):
        m = Moments()
        _check(self.lib.mcl_get_moments(self.h, lat, ctypes.byref(m)))
        return {k: getattr(m, k) for k, _ in Moments._fields_}

    def measure(self, lat=0):
        o = Observables()
        _check(self.lib.mcl_measure(self.h, lat, ctypes.byref(o)))
        d = {k: getattr(o, k) for k, _ in Observables._fields_}
        d["Q"] = list(o.Q)
        d["director"] = list(o.director)
        return d

    def directors(self, lat=0):
        nx, ny, nz = self.lattices[lat].dims
        out = np.empty(3 * nx * ny * nz, dtype=np.float32)
        _check(self.lib.mcl_get_directors(self.h, lat, _fptr(out)))
        return out

    def set_directors(self, ni, lat=0):
        a = np.ascontiguousarray(ni, dtype=np.float32).reshape(-1)
        _check(self.lib.mcl_set_directors(self.h, lat, _fptr(a)))

    def local_order(self, lat=0):
        nx, ny, nz = self.lattices[lat].dims
        out = np.empty(nx * ny * nz, dtype=np.float32)
        _check(self.lib.mcl_local_order(self.h, lat, _fptr(out)))
        return out

    def site_energies(self, sites, trial, lat=0):
        s = np.ascontiguousarray(sites, dtype=np.int32)
        t = np.ascontiguousarray(trial, dtype=np.float32).reshape(-1)
        ec = np.empty(len(s), dtype=np.float32)
        et = np.empty(len(s), dtype=np.float32)
        _check(self.lib.mcl_site_energies(self.h, lat, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fptr(t),
                                          len(s), _fptr(ec), _fptr(et)))
        return ec, et

    def site_energies_noalloc(self, sites, trial, lat=0):
        """mcl_site_energies_noalloc: the same energies without any device allocation (at most MCL_SCRATCH_SITES
        sites per call)."""
        s = np.ascontiguousarray(sites, dtype=np.int32)
        t = np.ascontiguousarray(trial, dtype=np.float32).reshape(-1)
        ec = np.empty(len(s), dtype=np.float32)
        et = np.empty(len(s), dtype=np.float32)
        _check(self.lib.mcl_site_energies_noalloc(self.h, lat, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                  _fptr(t), len(s), _fptr(ec), _fptr(et)))
        return ec, et

    def autotune(self, trial_sweeps=20, measure_every=0):
        """mcl_autotune_ex: the fastest launch shape for trial sweeps measured every measure_every sweeps."""
        r = TuneResult()
        _check(self.lib.mcl_autotune_ex(self.h, int(trial_sweeps), int(measure_every), ctypes.byref(r)))
        return {"threads": r.threads, "xcd_order": r.xcd_order, "us_per_sweep": r.us_per_sweep,
                "from_cache": bool(r.from_cache)}

    def set_launch(self, threads, xcd_order=1):
        _check(self.lib.mcl_set_launch(self.h, int(threads), int(xcd_order)))

    def set_graph(self, block_sweeps):
        _check(self.lib.mcl_set_graph(self.h, int(block_sweeps)))

    def prepare(self, n, measure_every=0):
        """mcl_prepare_sweeps: captures the graphs of sweeps(n, measure_every) without running them."""
        _check(self.lib.mcl_prepare_sweeps(self.h, int(n), int(measure_every)))

    def set_resident(self, enable):
        """mcl_set_resident: plain sweeps of small lattices in one launch per block (bit-identical trajectories)."""
        _check(self.lib.mcl_set_resident(self.h, 1 if enable else 0))

    def resident(self):
        v = ctypes.c_int32()
        _check(self.lib.mcl_resident(self.h, ctypes.byref(v)))
        return bool(v.value)

    def measure_in_sweep(self):
        v = ctypes.c_int32()
        _check(self.lib.mcl_measure_in_sweep(self.h, ctypes.byref(v)))
        return bool(v.value)

    def set_lattice_groups(self, max_bytes):
        """mcl_set_lattice_groups: sweep groups of lattices of at most max_bytes one after the other (0: one group;
        identical results)."""
        _check(self.lib.mcl_set_lattice_groups(self.h, int(max_bytes)))

    def lattice_groups(self):
        v = ctypes.c_int32()
        _check(self.lib.mcl_get_lattice_groups(self.h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        _check(self.lib.mcl_synchronize(self.h))

    def set_kernel_timing(self, stride, max_timed_sweeps=0):
        _check(self.lib.mcl_set_kernel_timing(self.h, int(stride), int(max_timed_sweeps)))

    def kernel_timing(self):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        _check(self.lib.mcl_kernel_timing(self.h, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value


def philox(ctr, key):
    """Device Philox4x32-10: ctr (n, 4) and key (n, 2) uint32 arrays -> (n, 4)."""
    lib = load()
    c = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    k = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.empty_like(c)
    P = ctypes.POINTER(ctypes.c_uint32)
    _check(lib.mcl_philox(c.ctypes.data_as(P), k.ctypes.data_as(P), len(c), out.ctypes.data_as(P)))
    return out


def copy_bandwidth(nbytes, reps=20):
    """STREAM-style device copy: (GB/s counting read + write bytes, ms per copy)."""
    g, ms = ctypes.c_double(), ctypes.c_double()
    _check(load().mcl_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


def _mlp_check(code):
    if code != MCL_OK:
        raise MclError(code, load().mcl_mlp_last_error().decode(errors="replace"))


class Mlp:
    """The surrogate network (5 -> 64 SiLU -> 64 SiLU -> 3) with HIP forward, gradient and Adam kernels."""

    def __init__(self, seed=1):
        self.lib = load()
        h = ctypes.c_void_p()
        _mlp_check(self.lib.mcl_mlp_create(5, 64, 3, int(seed), ctypes.byref(h)))
        self.h = h
        self.nparams = self.lib.mcl_mlp_num_params()

    def close(self):
        if getattr(self, "h", None):
            self.lib.mcl_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        _mlp_check(self.lib.mcl_mlp_get_params(self.h, _fptr(p), self.nparams))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _mlp_check(self.lib.mcl_mlp_set_params(self.h, _fptr(p), self.nparams))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
        y = np.empty((len(x), 3), dtype=np.float32)
        _mlp_check(self.lib.mcl_mlp_forward(self.h, _fptr(x), len(x), _fptr(y)))
        return y

    def gradient(self, x, t, w_out=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
        w = None if w_out is None else np.ascontiguousarray(w_out, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        _mlp_check(self.lib.mcl_mlp_gradient(self.h, _fptr(x), _fptr(t), None if w is None else _fptr(w), len(x),
                                             _fptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs, batch=512, lr=1e-3, w_out=None, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
        w = None if w_out is None else np.ascontiguousarray(w_out, dtype=np.float32)
        _mlp_check(self.lib.mcl_mlp_set_data(self.h, _fptr(x), _fptr(t), len(x)))
        loss = ctypes.c_float()
        _mlp_check(self.lib.mcl_mlp_train(self.h, int(epochs), int(batch), float(lr), None if w is None else _fptr(w),
                                          int(seed), ctypes.byref(loss)))
        return loss.value
