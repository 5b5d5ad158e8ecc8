"""Shared helpers of the test suite: synthetic fields, oracle runs and
set matching of root columns."""
from __future__ import annotations

import os

import numpy as np
from scipy.optimize import linear_sum_assignment

import ccf_amd
import oracle
from ccf_amd import io

# a few small fields covering the four smooth-matter variants
FIELDS = {
    # 100 equal masses, kappa = kappa_star = 0.3, shear 0.1
    "cfg1": dict(num_stars=100, kappa_tot=0.3, kappa_star=0.3, shear=0.1, mass_function="equal", seed=0),
    "salpeter300": dict(num_stars=300, kappa_tot=0.45, kappa_star=0.45, shear=0.35, mass_function="salpeter",
                        m_lower=0.08, m_upper=1.0, seed=1),
    "rect200": dict(num_stars=200, kappa_tot=0.7, kappa_star=0.35, shear=0.25, mass_function="kroupa",
                    m_lower=0.08, m_upper=1.0, rectangular=1, seed=3),
    "smooth150": dict(num_stars=150, kappa_tot=0.6, kappa_star=0.4, shear=0.2, mass_function="uniform",
                      m_lower=0.1, m_upper=1.0, seed=5),
}


def make_field(dirpath, name: str) -> str:
    spec = dict(FIELDS[name])
    seed = spec.pop("seed")
    path = os.path.join(str(dirpath), f"{name}.bin")
    if not os.path.exists(path):
        ccf_amd.write_synthetic_stars(path, seed, **spec)
    return path


def oracle_run(starfile: str, prefix: str, **kw) -> dict:
    p = oracle.params(starfile=starfile, outfile_prefix=prefix, **kw)
    return oracle.run(p)


def match_column(a: np.ndarray, b: np.ndarray):
    """Match root set a to root set b; returns |a_i - b_pi(i)| (index match
    first, Hungarian assignment on the entries that do not agree by index)."""
    d = np.abs(a - b)
    scale = np.maximum(np.abs(b), 1.0)
    bad = np.nonzero(d > 1e-9 * scale)[0]
    if bad.size:
        aa, bb = a[bad], b[bad]
        cost = np.abs(aa[:, None] - bb[None, :])
        r, c = linear_sum_assignment(cost)
        d = d.copy()
        d[bad[r]] = cost[r, c]
    return d


def compare_arrays(x: np.ndarray, y: np.ndarray, num_roots: int, num_branches: int, details: bool = False):
    """Max |dz| / max(|z|, 1) over all phi columns after set matching, and the
    number of points whose relative difference exceeds 1e-9; with details,
    also (branch, column, reference root, matched difference) of those points."""
    cx = io.columns_by_phi(x, num_roots, num_branches)
    cy = io.columns_by_phi(y, num_roots, num_branches)
    worst = 0.0
    nbad = 0
    bad = []
    for c in range(cx.shape[0]):
        for k in range(cx.shape[1]):
            d = match_column(cx[c, k], cy[c, k])
            rel = d / np.maximum(np.abs(cy[c, k]), 1.0)
            worst = max(worst, float(rel.max()))
            idx = np.nonzero(rel > 1e-9)[0]
            nbad += idx.size
            if details:
                bad += [(c, k, cy[c, k][i], d[i]) for i in idx]
    if details:
        return worst, nbad, bad
    return worst, nbad


def phi_of(c: int, k: int, num_phi: int, num_branches: int) -> float:
    """angle of column k of branch c (the file layout of ccf.cuh:1072-1080)"""
    return 2 * np.pi * (c * num_phi // num_branches + k) / num_phi


def newton_step(starfile: str, points, num_phi: int, num_branches: int, **kw) -> np.ndarray:
    """|f0 / f1| of the oracle at (branch, column, z) points: how far a root
    accepted by the reference's convergence tests can sit from the exact root."""
    if not len(points):
        return np.empty(0)
    z = np.array([p[2] for p in points], dtype=np.complex128)
    ph = np.array([phi_of(c, k, num_phi, num_branches) for c, k, _, _ in points])
    v = oracle.evaluate_phis(oracle.params(starfile=starfile, num_phi=num_phi, num_branches=num_branches, **kw), z, ph)
    return np.abs((v[:, 0] + 1j * v[:, 1]) / (v[:, 2] + 1j * v[:, 3]))


def compare_param_files(gpath: str, opath: str):
    """The product's parameter file against the oracle's (ccf.cuh:1135-1192):
    the same keys in the same order and the same printed values, except the
    timings (t_*) and max_error_1/mu, returned as (product, oracle) for a
    numeric comparison."""
    def lines(p):
        with open(p) as f:
            return [ln.split() for ln in f if ln.strip()]

    a, b = lines(gpath), lines(opath)
    assert [k for k, _ in a] == [k for k, _ in b], ([k for k, _ in a], [k for k, _ in b])
    err = None
    for (k, va), (_, vb) in zip(a, b):
        if k.startswith("t_"):
            continue
        if k == "max_error_1/mu":
            err = (float(va), float(vb))
            continue
        assert va == vb, (k, va, vb)
    return err


def differing_pairs(x: np.ndarray, y: np.ndarray, num_roots: int, num_branches: int, rel: float = 1e-9):
    """Matched pairs (branch, column, x root, y root) of two root arrays whose
    distance exceeds rel * max(|y|, 1) after set matching per phi column
    (index match, then Hungarian assignment on the entries that disagree)."""
    cx = io.columns_by_phi(x, num_roots, num_branches)
    cy = io.columns_by_phi(y, num_roots, num_branches)
    out = []
    d = np.abs(cx - cy)
    for c, k in np.argwhere((d / np.maximum(np.abs(cy), 1.0)).max(axis=2) > rel):
        a, b = cx[c, k], cy[c, k]
        bad = np.nonzero(np.abs(a - b) > rel * np.maximum(np.abs(b), 1.0))[0]
        cost = np.abs(a[bad][:, None] - b[bad][None, :])
        r, s = linear_sum_assignment(cost)
        for i, j in zip(r, s):
            if cost[i, j] > rel * max(abs(b[bad[j]]), 1.0):
                out.append((int(c), int(k), a[bad[i]], b[bad[j]]))
    return out


def set_mismatch(x: np.ndarray, y: np.ndarray, num_roots: int, num_branches: int, branches: int):
    """For the first `branches` branches, every root of each phi column of x
    that differs from y's root of the same index by more than 1e-9 relative,
    matched to the nearest root of y's column (k-d tree: for columns whose
    roots are permuted at scale, where match_column's assignment would not
    finish). Returns (max relative distance of those matches, count over
    1e-9)."""
    from scipy.spatial import cKDTree

    cx = io.columns_by_phi(x, num_roots, num_branches)
    cy = io.columns_by_phi(y, num_roots, num_branches)
    worst, nbad = 0.0, 0
    for c in range(branches):
        for k in range(cx.shape[1]):
            a, b = cx[c, k], cy[c, k]
            idx = np.nonzero(np.abs(a - b) > 1e-9 * np.maximum(np.abs(b), 1.0))[0]
            if idx.size == 0:
                continue
            d, j = cKDTree(np.c_[b.real, b.imag]).query(np.c_[a[idx].real, a[idx].imag])
            rel = d / np.maximum(np.abs(b[j]), 1.0)
            worst = max(worst, float(rel.max()))
            nbad += int((rel > 1e-9).sum())
    return worst, nbad


def branch_slice(arr: np.ndarray, num_roots: int, num_branches: int, branches: int) -> np.ndarray:
    """the rows of the first `branches` branches of a file-layout array, in
    the layout of a run with that many branches"""
    cols = arr.shape[1]
    return arr.reshape(num_roots, num_branches, cols)[:, :branches, :].reshape(num_roots * branches, cols)
