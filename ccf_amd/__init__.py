"""Python binding of ccf_amd, the MI355X critical-curve / caustic finder.

Loads the in-tree ``libccf_amd.so`` (built by ``make -C ccf_amd``) through
ctypes and mirrors the C-ABI of ``include/ccf_amd.h``. The class
:class:`CCF` mirrors the reference's ``CCF<T>`` (weisluke/ccf
include/ccf.cuh:34-1289): set attributes, then ``run(verbose)`` and
``save(verbose)``, both returning bool.

There is no CPU fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccf_amd.so")
CLI_PATH = os.path.join(HERE, "ccf_gpu")
PATH_MAX = 1024

ARRAY_STARS = 0
ARRAY_CCS = 1
ARRAY_CAUSTICS = 2
ARRAY_MU_LENGTH_SCALES = 3


class Params(ctypes.Structure):
    """ccf_params (include/ccf_amd.h)."""

    _fields_ = [
        ("kappa_tot", ctypes.c_double),
        ("shear", ctypes.c_double),
        ("kappa_star", ctypes.c_double),
        ("theta_star", ctypes.c_double),
        ("mass_function", ctypes.c_char * 32),
        ("m_solar", ctypes.c_double),
        ("m_lower", ctypes.c_double),
        ("m_upper", ctypes.c_double),
        ("rectangular", ctypes.c_int),
        ("approx", ctypes.c_int),
        ("safety_scale", ctypes.c_double),
        ("num_stars", ctypes.c_int),
        ("starfile", ctypes.c_char * PATH_MAX),
        ("num_phi", ctypes.c_int),
        ("num_branches", ctypes.c_int),
        ("random_seed", ctypes.c_int),
        ("write_stars", ctypes.c_int),
        ("write_critical_curves", ctypes.c_int),
        ("write_caustics", ctypes.c_int),
        ("write_mu_length_scales", ctypes.c_int),
        ("outfile_prefix", ctypes.c_char * PATH_MAX),
        ("fmm_order", ctypes.c_int),
        ("direct", ctypes.c_int),
        ("binomial_exact", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("init_iters", ctypes.c_int),
        ("step_iters", ctypes.c_int),
        ("roots_per_thread", ctypes.c_int),
        ("timing", ctypes.c_int),
        ("predictor", ctypes.c_int),
        ("root_fmm", ctypes.c_int),
        ("root_fmm_order", ctypes.c_int),
        ("root_fmm_leaf", ctypes.c_int),
        ("autotune", ctypes.c_int),
    ]


class Info(ctypes.Structure):
    """ccf_info (include/ccf_amd.h)."""

    _fields_ = [
        ("num_stars", ctypes.c_int),
        ("num_roots", ctypes.c_int),
        ("tree_levels", ctypes.c_int),
        ("expansion_order", ctypes.c_int),
        ("taylor_smooth", ctypes.c_int),
        ("num_nodes", ctypes.c_int),
        ("random_seed", ctypes.c_int),
        ("kappa_star", ctypes.c_double),
        ("kappa_star_actual", ctypes.c_double),
        ("mu_ave", ctypes.c_double),
        ("corner_re", ctypes.c_double),
        ("corner_im", ctypes.c_double),
        ("alpha_error", ctypes.c_double),
        ("max_error", ctypes.c_double),
        ("t_tree", ctypes.c_double),
        ("t_init_roots", ctypes.c_double),
        ("t_ccs", ctypes.c_double),
        ("t_caustics", ctypes.c_double),
        ("points_total", ctypes.c_longlong),
        ("points_converged", ctypes.c_longlong),
        ("updates_full", ctypes.c_longlong),
        ("updates_early", ctypes.c_longlong),
        ("flops_update", ctypes.c_double),
        ("aberth_kernel_ms", ctypes.c_double),
        ("aberth_launches", ctypes.c_longlong),
        ("sweeps", ctypes.c_longlong),
        ("cluster_chunks", ctypes.c_longlong),
        ("flops_sums", ctypes.c_double),
        ("root_fmm_updates", ctypes.c_longlong),
        ("root_fmm_rows", ctypes.c_longlong),
        ("root_fmm_order", ctypes.c_int),
        ("root_fmm_leaf_level", ctypes.c_int),
        ("max_witt_residual", ctypes.c_double),
        ("witt_kernel_ms", ctypes.c_double),
        ("witt_launches", ctypes.c_longlong),
        ("flops_witt", ctypes.c_double),
        ("updates_newton", ctypes.c_longlong),
        ("tail_launches", ctypes.c_longlong),
        ("tail_sweeps", ctypes.c_longlong),
        ("tail_kernel_ms", ctypes.c_double),
        ("flops_tail", ctypes.c_double),
        ("points_outside_tree", ctypes.c_longlong),
        ("roots_per_thread", ctypes.c_int),
        ("root_fmm_leaf", ctypes.c_int),
        ("t_autotune", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libccf_amd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C ccf_amd` (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(Params)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "ccf_create": (vp, []),
        "ccf_destroy": (None, [vp]),
        "ccf_default_params": (None, [P]),
        "ccf_set_params": (ctypes.c_int, [vp, P]),
        "ccf_get_params": (ctypes.c_int, [vp, P]),
        "ccf_run": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_save": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_check_input_params": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_calculate_derived_params": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_prepare": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_load": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_build_tree_fmm": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_find_initial_roots": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_find_ccs": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_find_caustics": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_find_mu_length_scales": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_get_info": (ctypes.c_int, [vp, ctypes.POINTER(Info)]),
        "ccf_get_array": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_longlong,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "ccf_last_error": (ctypes.c_char_p, [vp]),
        "ccf_eval_points": (ctypes.c_int, [vp, dp, ctypes.c_int, ctypes.c_double, dp]),
        "ccf_aberth_sweep": (ctypes.c_int, [vp, dp, ctypes.c_int, ctypes.c_double, dp]),
        "ccf_aberth_sweep_noalloc": (ctypes.c_int, [vp, dp, ctypes.c_int, ctypes.c_double, dp]),
        "ccf_root_sums": (ctypes.c_int, [vp, dp, ctypes.c_int, dp]),
        "ccf_batch_create": (vp, []),
        "ccf_batch_destroy": (None, [vp]),
        "ccf_batch_add": (ctypes.c_int, [vp, P]),
        "ccf_batch_size": (ctypes.c_int, [vp]),
        "ccf_batch_load": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_batch_build": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_batch_solve": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_batch_run": (ctypes.c_int, [vp, ctypes.c_int]),
        "ccf_batch_save": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "ccf_batch_get_info": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(Info)]),
        "ccf_batch_get_array": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_longlong,
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "ccf_batch_workgroup": (ctypes.c_int, [vp]),
        "ccf_batch_last_error": (ctypes.c_char_p, [vp]),
        "ccf_write_synthetic_stars": (ctypes.c_int, [P, ctypes.c_ulonglong, ctypes.c_char_p]),
        "ccf_mass_function": (ctypes.c_double, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_params() -> Params:
    p = Params()
    load().ccf_default_params(ctypes.byref(p))
    return p


def make_params(**kw) -> Params:
    """Reference defaults overridden by keyword arguments (strings are encoded)."""
    p = default_params()
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        if isinstance(v, str):
            v = v.encode()
        setattr(p, k, v)
    return p


def mass_function(name: str, which: int, p: float, m_lower: float, m_upper: float, m_solar: float = 1.0) -> float:
    return load().ccf_mass_function(name.encode(), which, p, m_lower, m_upper, m_solar)


def write_synthetic_stars(path: str, seed: int, **kw) -> None:
    """Seeded synthetic star file for the field the reference derives from kw."""
    p = make_params(**kw)
    if load().ccf_write_synthetic_stars(ctypes.byref(p), seed, path.encode()) != 0:
        raise RuntimeError(f"could not write {path}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One ccf_ctx: a star field and its device buffers."""

    def __init__(self, params: Params | None = None, **kw):
        self.lib = load()
        self.ptr = self.lib.ccf_create()
        if not self.ptr:
            raise MemoryError("ccf_create failed")
        self.params = params if params is not None else make_params(**kw)
        self.lib.ccf_set_params(self.ptr, ctypes.byref(self.params))

    def close(self):
        if self.ptr:
            self.lib.ccf_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v.encode() if isinstance(v, str) else v)
        self.lib.ccf_set_params(self.ptr, ctypes.byref(self.params))

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {self.last_error()}")

    def last_error(self) -> str:
        return self.lib.ccf_last_error(self.ptr).decode()

    def run(self, verbose: int = 0) -> bool:
        return self.lib.ccf_run(self.ptr, verbose) == 0

    def save(self, verbose: int = 0) -> bool:
        return self.lib.ccf_save(self.ptr, verbose) == 0

    def check(self, verbose: int = 0) -> bool:
        return self.lib.ccf_check_input_params(self.ptr, verbose) == 0

    def prepare(self, verbose: int = 0):
        self._check(self.lib.ccf_prepare(self.ptr, verbose), "ccf_prepare")

    def stage(self, name: str, verbose: int = 0):
        self._check(getattr(self.lib, "ccf_" + name)(self.ptr, verbose), "ccf_" + name)

    def info(self) -> dict:
        i = Info()
        self.lib.ccf_get_info(self.ptr, ctypes.byref(i))
        return i.as_dict()

    def array(self, which: int) -> np.ndarray:
        inf = self.info()
        R = inf["num_roots"]
        npb = self.params.num_phi + self.params.num_branches
        if which == ARRAY_STARS:
            out = np.empty((inf["num_stars"], 3), dtype=np.float64)
        elif which == ARRAY_MU_LENGTH_SCALES:
            out = np.empty(R * npb, dtype=np.float64)
        else:
            out = np.empty(R * npb, dtype=np.complex128)
        nr = ctypes.c_int()
        nc = ctypes.c_int()
        self._check(self.lib.ccf_get_array(self.ptr, which, out.ctypes.data_as(ctypes.c_void_p), out.nbytes,
                                           ctypes.byref(nr), ctypes.byref(nc)), "ccf_get_array")
        return out.reshape(nr.value, nc.value)

    def eval_points(self, z, phi: float) -> np.ndarray:
        """(n, 8): f0.re, f0.im, f1.re, f1.im, w.re, w.im, mu_length_scale, 1/mu bound."""
        z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128).ravel())
        out = np.empty((z.size, 8), dtype=np.float64)
        self._check(self.lib.ccf_eval_points(self.ptr, _dptr(z.view(np.float64)), z.size, phi, _dptr(out)),
                    "ccf_eval_points")
        return out

    def aberth_sweep(self, roots, phi: float) -> np.ndarray:
        r = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel())
        out = np.empty_like(r)
        self._check(self.lib.ccf_aberth_sweep(self.ptr, _dptr(r.view(np.float64)), r.size, phi,
                                              _dptr(out.view(np.float64))), "ccf_aberth_sweep")
        return out

    def aberth_sweep_noalloc(self, roots, phi: float) -> np.ndarray:
        """aberth_sweep without device allocations (the library's fixed scratch array)"""
        r = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel())
        out = np.empty_like(r)
        self._check(self.lib.ccf_aberth_sweep_noalloc(self.ptr, _dptr(r.view(np.float64)), r.size, phi,
                                                      _dptr(out.view(np.float64))), "ccf_aberth_sweep_noalloc")
        return out

    def root_sums(self, roots) -> np.ndarray:
        """sum_{j != k} 1/(z_k - z_j) for every root, by the root-sum FMM."""
        r = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel())
        out = np.empty_like(r)
        self._check(self.lib.ccf_root_sums(self.ptr, _dptr(r.view(np.float64)), r.size, _dptr(out.view(np.float64))),
                    "ccf_root_sums")
        return out


class Batch:
    """Several star fields solved together (ccf_batch_* of include/ccf_amd.h)."""

    def __init__(self):
        self.lib = load()
        self.ptr = self.lib.ccf_batch_create()
        if not self.ptr:
            raise MemoryError("ccf_batch_create failed")
        self.params = []

    def close(self):
        if self.ptr:
            self.lib.ccf_batch_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def add(self, params: Params | None = None, **kw) -> int:
        p = params if params is not None else make_params(**kw)
        i = self.lib.ccf_batch_add(self.ptr, ctypes.byref(p))
        if i < 0:
            raise RuntimeError("ccf_batch_add failed")
        self.params.append(p)
        return i

    def last_error(self) -> str:
        return self.lib.ccf_batch_last_error(self.ptr).decode()

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.last_error().strip()}")

    def load(self, verbose: int = 0):
        self._check(self.lib.ccf_batch_load(self.ptr, verbose), "ccf_batch_load")

    def build(self, verbose: int = 0):
        self._check(self.lib.ccf_batch_build(self.ptr, verbose), "ccf_batch_build")

    def solve(self, verbose: int = 0):
        self._check(self.lib.ccf_batch_solve(self.ptr, verbose), "ccf_batch_solve")

    def run(self, verbose: int = 0) -> bool:
        return self.lib.ccf_batch_run(self.ptr, verbose) == 0

    def save(self, field: int, verbose: int = 0) -> bool:
        return self.lib.ccf_batch_save(self.ptr, field, verbose) == 0

    def info(self, field: int = -1) -> dict:
        i = Info()
        self._check(self.lib.ccf_batch_get_info(self.ptr, field, ctypes.byref(i)), "ccf_batch_get_info")
        return i.as_dict()

    def workgroup(self) -> int:
        return self.lib.ccf_batch_workgroup(self.ptr)

    def array(self, field: int, which: int) -> np.ndarray:
        inf = self.info(field)
        p = self.params[field]
        R = inf["num_roots"]
        npb = p.num_phi + p.num_branches
        if which == ARRAY_STARS:
            out = np.empty((inf["num_stars"], 3), dtype=np.float64)
        elif which == ARRAY_MU_LENGTH_SCALES:
            out = np.empty(R * npb, dtype=np.float64)
        else:
            out = np.empty(R * npb, dtype=np.complex128)
        nr = ctypes.c_int()
        nc = ctypes.c_int()
        self._check(self.lib.ccf_batch_get_array(self.ptr, field, which, out.ctypes.data_as(ctypes.c_void_p), out.nbytes,
                                                 ctypes.byref(nr), ctypes.byref(nc)), "ccf_batch_get_array")
        return out.reshape(nr.value, nc.value)


class CCF:
    """Mirror of the reference's CCF<T>: public fields, run(verbose), save(verbose)."""

    _fields = [f for f, _ in Params._fields_]

    def __init__(self, **kw):
        p = default_params()
        for f in self._fields:
            v = getattr(p, f)
            object.__setattr__(self, f, v.decode() if isinstance(v, bytes) else v)
        self.mass_function_str = self.mass_function
        for k, v in kw.items():
            setattr(self, k, v)
        self._ctx = Context(default_params())

    def _push(self):
        kw = {f: getattr(self, f) for f in self._fields}
        kw["mass_function"] = self.mass_function_str
        self._ctx.set(**kw)

    def run(self, verbose: int = 1) -> bool:
        self._push()
        return self._ctx.run(verbose)

    def save(self, verbose: int = 1) -> bool:
        return self._ctx.save(verbose)

    @property
    def context(self) -> Context:
        return self._ctx
