"""ctypes binding of the FP64 CPU oracle (oracle/libccf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product (ccf_amd) never
imports it. See oracle/oracle.h for the restated algorithm and its
documented differences from the reference.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccf_oracle.so")


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("kappa_tot", ctypes.c_double),
        ("shear", ctypes.c_double),
        ("kappa_star", ctypes.c_double),
        ("theta_star", ctypes.c_double),
        ("m_solar", ctypes.c_double),
        ("m_lower", ctypes.c_double),
        ("m_upper", ctypes.c_double),
        ("safety_scale", ctypes.c_double),
        ("rectangular", ctypes.c_int),
        ("approx", ctypes.c_int),
        ("num_stars", ctypes.c_int),
        ("num_phi", ctypes.c_int),
        ("num_branches", ctypes.c_int),
        ("random_seed", ctypes.c_int),
        ("write_stars", ctypes.c_int),
        ("write_critical_curves", ctypes.c_int),
        ("write_caustics", ctypes.c_int),
        ("write_mu_length_scales", ctypes.c_int),
        ("verbose", ctypes.c_int),
        ("fmm_order", ctypes.c_int),
        ("direct", ctypes.c_int),
        ("binomial_exact", ctypes.c_int),
        ("init_iters", ctypes.c_int),
        ("step_iters", ctypes.c_int),
        ("step_limit", ctypes.c_int),
        ("branch_limit", ctypes.c_int),
        ("num_threads", ctypes.c_int),
        ("compact_output", ctypes.c_int),
        ("plant", ctypes.c_int),
        ("mass_function", ctypes.c_char * 32),
        ("starfile", ctypes.c_char * 512),
        ("outfile_prefix", ctypes.c_char * 512),
    ]


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("num_stars", ctypes.c_int),
        ("num_roots", ctypes.c_int),
        ("tree_levels", ctypes.c_int),
        ("expansion_order", ctypes.c_int),
        ("taylor_smooth", ctypes.c_int),
        ("num_nodes", ctypes.c_int),
        ("kappa_star", ctypes.c_double),
        ("corner_re", ctypes.c_double),
        ("corner_im", ctypes.c_double),
        ("max_error", ctypes.c_double),
        ("t_tree", ctypes.c_double),
        ("t_init_roots", ctypes.c_double),
        ("t_ccs", ctypes.c_double),
        ("t_caustics", ctypes.c_double),
        ("points_solved", ctypes.c_longlong),
        ("points_converged", ctypes.c_longlong),
        ("updates_full", ctypes.c_longlong),
        ("updates_early", ctypes.c_longlong),
        ("flops", ctypes.c_double),
        ("max_witt_residual", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# planted mistakes (oracle.h ORACLE_PLANT_*): tests of the tests only
PLANT_W_STAR_UNCONJUGATED = 1
PLANT_BOUND_SMOOTH_SIGN = 2
PLANT_MU_SCALE_FACTOR = 3
PLANT_DISC_OUTSIDE_Z = 4
PLANT_DISC_DZBAR_INSIDE = 5

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} missing: build with `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(OracleParams)
        S = ctypes.POINTER(OracleStats)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_default_params.argtypes = [P]
        lib.oracle_run.argtypes = [P, S]
        lib.oracle_run.restype = ctypes.c_int
        lib.oracle_binomial.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_binomial.restype = ctypes.c_double
        lib.oracle_mass_function.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double]
        lib.oracle_mass_function.restype = ctypes.c_double
        lib.oracle_eval.argtypes = [P, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp, S]
        lib.oracle_eval.restype = ctypes.c_int
        lib.oracle_eval_phis.argtypes = [P, ctypes.c_int, dp, dp, ctypes.c_int, dp]
        lib.oracle_eval_phis.restype = ctypes.c_int
        lib.oracle_aberth_sweep.argtypes = [P, dp, ctypes.c_int, ctypes.c_double, dp]
        lib.oracle_aberth_sweep.restype = ctypes.c_int
        _lib = lib
    return _lib


def params(**kw) -> OracleParams:
    p = OracleParams()
    load().oracle_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v.encode() if isinstance(v, str) else v)
    return p


def run(p: OracleParams) -> dict:
    s = OracleStats()
    rc = load().oracle_run(ctypes.byref(p), ctypes.byref(s))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed ({rc})")
    return s.as_dict()


def binomial(n: int, k: int, wrapped: bool) -> float:
    return load().oracle_binomial(n, k, int(wrapped))


def mass_function(name: str, which: int, p: float, m_lower: float, m_upper: float, m_solar: float = 1.0) -> float:
    return load().oracle_mass_function(name.encode(), which, p, m_lower, m_upper, m_solar)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def evaluate(p: OracleParams, z, phi: float, use_tree: bool = True):
    """(values (n, 8), stats): f0, f1, w, mu length scale, 1/mu bound per point."""
    z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128).ravel())
    out = np.empty((z.size, 8))
    s = OracleStats()
    rc = load().oracle_eval(ctypes.byref(p), int(use_tree), _dp(z.view(np.float64)), z.size, phi, _dp(out),
                            ctypes.byref(s))
    if rc != 0:
        raise RuntimeError("oracle_eval failed")
    return out, s.as_dict()


def evaluate_phis(p: OracleParams, z, phis, use_tree: bool = True) -> np.ndarray:
    """as evaluate, with one angle per point: values (n, 8)"""
    z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128).ravel())
    ph = np.ascontiguousarray(np.broadcast_to(np.asarray(phis, dtype=np.float64), z.shape))
    out = np.empty((z.size, 8))
    if load().oracle_eval_phis(ctypes.byref(p), int(use_tree), _dp(z.view(np.float64)), _dp(ph), z.size, _dp(out)) != 0:
        raise RuntimeError("oracle_eval_phis failed")
    return out


def newton_polish(p: OracleParams, z, phis, steps: int = 3) -> np.ndarray:
    """the roots of Witt's function nearest z (at angles phis), by FP64
    Newton steps z - f0/f1 of the oracle's field: the exact root to rounding"""
    z = np.asarray(z, dtype=np.complex128).ravel().copy()
    for _ in range(steps):
        v = evaluate_phis(p, z, phis)
        z -= (v[:, 0] + 1j * v[:, 1]) / (v[:, 2] + 1j * v[:, 3])
    return z


def aberth_sweep(p: OracleParams, roots, phi: float) -> np.ndarray:
    r = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel())
    out = np.empty_like(r)
    if load().oracle_aberth_sweep(ctypes.byref(p), _dp(r.view(np.float64)), r.size, phi, _dp(out.view(np.float64))) != 0:
        raise RuntimeError("oracle_aberth_sweep failed")
    return out
