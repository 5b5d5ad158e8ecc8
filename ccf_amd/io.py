"""Readers and writers of cc_finder's file layout.

Layouts as written by weisluke/ccf: parameter text (include/ccf.cuh:1135-1192),
star files (include/star.cuh:498-534) and 2-D arrays with an int32
(nrows, ncols) header (include/array_functions.cuh:280-306). Row
``root * num_branches + branch`` of a critical-curve or caustic array holds
column k at phi = 2 pi (branch * num_phi / num_branches + k) / num_phi.
"""
from __future__ import annotations

import numpy as np


def read_params(fname: str) -> dict:
    out = {}
    with open(fname) as f:
        for line in f:
            key, val = line.split()
            try:
                v = float(val)
                out[key] = int(v) if v == int(v) and "." not in val and "e" not in val.lower() else v
            except ValueError:
                out[key] = val
    return out


def read_stars(fname: str):
    """(nstars, rectangular, corner complex, theta, stars (n, 3) float64)."""
    with open(fname, "rb") as f:
        raw = f.read()
    n, rect = np.frombuffer(raw[:8], dtype=np.int32)
    if len(raw) == 8 + 16 + 8 + 24 * n:
        dt, sz = np.float64, 8
    elif len(raw) == 8 + 8 + 4 + 12 * n:
        dt, sz = np.float32, 4
    else:
        raise ValueError(f"{fname}: size {len(raw)} matches neither precision")
    corner = np.frombuffer(raw[8:8 + 2 * sz], dtype=dt).astype(np.float64)
    theta = float(np.frombuffer(raw[8 + 2 * sz:8 + 3 * sz], dtype=dt)[0])
    stars = np.frombuffer(raw[8 + 3 * sz:], dtype=dt).astype(np.float64).reshape(n, 3)
    return int(n), int(rect), complex(corner[0], corner[1]), theta, stars


def write_stars(fname: str, stars, rectangular: int, corner: complex, theta: float = 1.0) -> None:
    stars = np.ascontiguousarray(stars, dtype=np.float64).reshape(-1, 3)
    with open(fname, "wb") as f:
        f.write(np.array([len(stars), rectangular], dtype=np.int32).tobytes())
        f.write(np.array([corner.real, corner.imag, theta], dtype=np.float64).tobytes())
        f.write(stars.tobytes())


def read_array(fname: str, complex_values: bool = True) -> np.ndarray:
    with open(fname, "rb") as f:
        raw = f.read()
    nrows, ncols = np.frombuffer(raw[:8], dtype=np.int32)
    dt = np.complex128 if complex_values else np.float64
    return np.frombuffer(raw[8:], dtype=dt).reshape(int(nrows), int(ncols)).copy()


def columns_by_phi(arr: np.ndarray, num_roots: int, num_branches: int) -> np.ndarray:
    """(num_roots, num_branches * ncols) -> for branch c column k: all roots at that phi.

    Returns shape (num_branches, ncols, num_roots)."""
    nrows, ncols = arr.shape
    assert nrows == num_roots * num_branches
    a = arr.reshape(num_roots, num_branches, ncols)
    return np.transpose(a, (1, 2, 0))
