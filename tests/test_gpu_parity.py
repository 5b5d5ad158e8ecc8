"""Parity of the HIP path with the FP64 oracle on identical star files.

Every test calls the product through its C-ABI (libccf_amd.so) and compares
with the oracle (oracle/libccf_oracle.so) run on the same inputs:
  - field evaluation (f0, f1, lens map, length scale, 1/mu bound) at random
    points, tree and direct (relative 1e-11);
  - one Jacobi Aberth sweep (relative 1e-11);
  - whole runs: critical curves and caustics per phi column as sets,
    |dz| <= 1e-9 max(|z|, theta_E) (the parity bar), plus the parameter
    file, the star file and the 1/mu error.
"""
import ctypes
import os

import numpy as np
import pytest

import ccf_amd
import oracle
from ccf_amd import io
from helpers import FIELDS, compare_arrays, compare_param_files, make_field, newton_step, oracle_run

pytestmark = pytest.mark.gpu


def _ctx(starfile, prefix, **kw):
    return ccf_amd.Context(starfile=starfile, outfile_prefix=prefix, **kw)


@pytest.mark.parametrize("name,kw", [
    ("cfg1", dict(kappa_tot=0.3, shear=0.1)),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, approx=0)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, direct=1)),
])
def test_eval_matches_oracle(workdir, name, kw):
    sf = make_field(workdir, name)
    rng = np.random.default_rng(7)
    n, rect, corner, theta, st = io.read_stars(sf)
    scale = max(abs(corner.real), abs(corner.imag), abs(corner))
    # points near stars, inside and outside the field
    z = st[rng.integers(0, n, 300), 0] + 1j * st[rng.integers(0, n, 300), 1]
    z = np.concatenate([z + 0.3 * (rng.standard_normal(300) + 1j * rng.standard_normal(300)),
                        1.3 * scale * (rng.random(50) - 0.5 + 1j * (rng.random(50) - 0.5))])
    phi = 1.234
    with _ctx(sf, str(workdir) + "/", num_phi=10, **kw) as c:
        c.prepare()
        g = c.eval_points(z, phi)
    o, _ = oracle.evaluate(oracle.params(starfile=sf, num_phi=10, **kw), z, phi)
    for col, what in [(0, "f0"), (2, "f1"), (4, "w")]:
        a = g[:, col] + 1j * g[:, col + 1]
        b = o[:, col] + 1j * o[:, col + 1]
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
        assert rel.max() < 1e-11, (what, rel.max())
    rel = np.abs(g[:, 6] - o[:, 6]) / np.abs(o[:, 6])
    assert rel.max() < 1e-9
    assert np.allclose(g[:, 7], o[:, 7], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("name,kw", [
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, approx=0)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, direct=1)),
])
def test_aberth_sweep_matches_oracle(workdir, name, kw):
    sf = make_field(workdir, name)
    n, rect, corner, theta, st = io.read_stars(sf)
    # extra roots off the disc boundary |z| = |corner|, where the exact
    # circular smooth term is discontinuous
    roots = np.concatenate([st[:, 0] + 1 + 1j * st[:, 1], st[:, 0] - 1 + 1j * st[:, 1],
                            1.05 * abs(corner) * np.exp(2j * np.pi * (np.arange(7) + 0.5) / 7)])
    phi = np.pi / 3
    kw = dict(kw, num_phi=10)
    with _ctx(sf, str(workdir) + "/", **kw) as c:
        c.prepare()
        g = c.aberth_sweep(roots, phi)
    o = oracle.aberth_sweep(oracle.params(starfile=sf, **kw), roots, phi)
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1.0)
    assert rel.max() < 1e-11


@pytest.mark.parametrize("count", [20, None])
def test_mixed_sweep_step_matches_oracle(workdir, count):
    """precision 1, direct star sums: the Aberth step of one sweep with the
    Cauchy sums in FP32 (the sparse kernel for a row of at most 32 listed
    roots, the tiled one above) against the oracle's FP64 step. The sums only
    scale the step, so the two steps agree to FP32 rounding of S relative to
    the step's length; a wrong term, sign or index changes it by order one."""
    sf = make_field(workdir, "cfg1")
    n, rect, corner, theta, st = io.read_stars(sf)
    roots = np.concatenate([st[:, 0] + 1 + 1j * st[:, 1], st[:, 0] - 1 + 1j * st[:, 1],
                            1.05 * abs(corner) * np.exp(2j * np.pi * (np.arange(7) + 0.5) / 7)])
    if count:
        roots = roots[np.random.default_rng(7).choice(roots.size, count, replace=False)]
    phi = np.pi / 3
    kw = dict(kappa_tot=0.3, shear=0.1, direct=1, num_phi=10)
    with _ctx(sf, str(workdir) + "/", precision=1, **kw) as c:
        c.prepare()
        g = c.aberth_sweep(roots, phi)
    o = oracle.aberth_sweep(oracle.params(starfile=sf, **kw), roots, phi)
    step = np.abs(o - roots)
    assert step.min() > 1e-6  # every root took a full update (none at the 1/mu exit)
    assert np.all(np.abs(g - o) <= 1e-3 * step + 1e-12 * np.maximum(np.abs(o), 1.0)), \
        np.max(np.abs(g - o) / step)


def _direct_root_sums(z):
    """sum_{j != k} 1/(z_k - z_j) in FP64, blockwise, and the scale of each
    sum, sum_{j != k} 1/|z_k - z_j| (the far-field error bound's unit)."""
    S = np.empty_like(z)
    scale = np.empty(z.size)
    for a in range(0, z.size, 2048):
        d = z[a:a + 2048, None] - z[None, :]
        d[np.arange(d.shape[0]), a + np.arange(d.shape[0])] = np.inf
        S[a:a + 2048] = (1.0 / d).sum(axis=1)
        scale[a:a + 2048] = (1.0 / np.abs(d)).sum(axis=1)
    return S, scale


@pytest.mark.parametrize("name,kw,n,tol", [
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35), 6000, 5e-5),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, root_fmm_order=14), 6000, 2e-7),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, root_fmm_order=6, root_fmm_leaf=3), 6000, 5e-4),
    ("rect200", dict(kappa_tot=0.7, shear=0.25), 5000, 5e-5),
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, root_fmm_order=24, root_fmm_leaf=12), 3000, 1e-9),
])
def test_root_sums_fmm_matches_direct(workdir, name, kw, n, tol):
    """The root-sum FMM against direct FP64 sums: roots clustered at the stars
    and spread over the field, a few outside the grid, a close pair."""
    sf = make_field(workdir, name)
    rng = np.random.default_rng(11)
    _, rect, corner, _, st = io.read_stars(sf)
    ext = np.array([abs(corner.real), abs(corner.imag)]) if rect else np.array([abs(corner), abs(corner)])
    s = st[rng.integers(0, st.shape[0], n // 2), :2]
    z = np.concatenate([s[:, 0] + 1j * s[:, 1] + 0.05 * (rng.standard_normal(n // 2) + 1j * rng.standard_normal(n // 2)),
                        ext[0] * rng.uniform(-1, 1, n - n // 2) + 1j * ext[1] * rng.uniform(-1, 1, n - n // 2)])
    z[:4] = [2.0 * ext[0], -1.5j * ext[1], 0.3 + 0.2j, 0.3 + 0.2j + 1e-7]
    with _ctx(sf, str(workdir) + "/", num_phi=10, **kw) as c:
        c.prepare()
        S = c.root_sums(z)
    D, scale = _direct_root_sums(z)
    # truncation error of the far field, in units of sum 1/|z_k - z_j|
    # (Greengard-Rokhlin bound ~ 0.71^(p+1) for the closest well-separated
    # cells; measured on uniform sets: 3e-8 at p = 14, 8e-6 at p = 8, 5e-5 at p = 6, 1e-11 at p = 24;
    # the default order is 8)
    rel = np.abs(S - D) / scale
    assert np.all(np.isfinite(S))
    assert rel.max() <= tol, rel.max()


def test_root_sums_fmm_binning_without_lds(workdir):
    """34 000 roots at one root per leaf: the grid's leaf level 8 has 65 536
    cells per row, whose counts (256 KB) exceed a workgroup's LDS, so the
    roots are binned by the multi-launch kernels instead of the one-launch
    row kernel. Checked against direct FP64 sums on 3000 targets."""
    sf = make_field(workdir, "salpeter300")
    rng = np.random.default_rng(5)
    _, rect, corner, _, st = io.read_stars(sf)
    h = abs(corner)
    n = 34000
    z = h * (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n))
    z[:3] = [2.5 * h, -2.0j * h, 0.1 + 0.1j]
    with _ctx(sf, str(workdir) + "/", num_phi=10, kappa_tot=0.45, shear=0.35, root_fmm_leaf=1) as c:
        c.prepare()
        S = c.root_sums(z)
    pick = np.concatenate([np.arange(3), rng.choice(np.arange(3, n), 2997, replace=False)])
    d = z[pick, None] - z[None, :]
    d[np.arange(pick.size), pick] = np.inf
    D = (1.0 / d).sum(axis=1)
    scale = (1.0 / np.abs(d)).sum(axis=1)
    rel = np.abs(S[pick] - D) / scale
    assert np.all(np.isfinite(S))
    assert rel.max() <= 5e-5, rel.max()


def check_error_maxima(ginfo, ostats, file_err=None, oracle_file_err=None, exact=True):
    """max_error_1/mu (ccf.cuh:972,1068) and the max Witt residual |f0| of a
    run against the oracle's. Where the product follows the oracle's iterates
    (the FP64 path) they agree to the rounding of f0 at the worst point: that
    |f0| is 1e-9 to 1e-5, and the two last iterates differ by rounding
    (1e-15 |z|), which moves f0 by |f1| times that (|f1| reaches 1e3 beside
    a star), so 1e-3 relative or 1e-12 absolute. The other forms take
    other paths to the roots and stop elsewhere inside the convergence band:
    an early exit anywhere below the 1/mu bound's 1e-9 (ccf_functions.cuh:
    207-211), where |f0| < 1e-9 / (2 (1 - kappa - d alpha_smooth/dz)), below
    1e-8 on these fields; their maxima are held to the larger of the
    oracle's and that band."""
    g, o = ginfo["max_error"], ostats["max_error"]
    gr, orr = ginfo["max_witt_residual"], ostats["max_witt_residual"]
    assert np.isfinite(g) and np.isfinite(gr) and 0 < o < 1e-7 and 0 < orr
    if file_err is not None:
        assert file_err == pytest.approx(g, rel=1e-8) and oracle_file_err == pytest.approx(o, rel=1e-8)
    if exact:
        assert g == pytest.approx(o, rel=1e-3, abs=1e-12), (g, o)
        assert gr == pytest.approx(orr, rel=1e-3, abs=1e-12), (gr, orr)
    else:
        assert g <= max(o, 1e-9) * (1 + 1e-2) + 1e-12, (g, o)
        assert gr <= max(orr, 1e-8) * (1 + 1e-2) + 1e-12, (gr, orr)


RUNS = [
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, num_phi=100, num_branches=1, direct=1)),
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, num_phi=100, num_branches=1)),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, num_phi=80, num_branches=2)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, num_phi=60, num_branches=1, approx=0)),
]
# the mixed-precision form and the continuation predictors change only the
# path to each root: the same oracle run is the reference for them
MIXED = [
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, num_phi=100, num_branches=1, direct=1), dict(precision=1)),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20), dict(precision=1)),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20), dict(precision=1, predictor=1)),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20), dict(predictor=2)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, num_phi=80, num_branches=2), dict(precision=1, predictor=2)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, num_phi=60, num_branches=1, approx=0), dict(precision=1, predictor=1)),
    # root sums of the dense sweeps by the FMM
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, num_phi=100, num_branches=1, direct=1), dict(root_fmm=1)),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20),
     dict(precision=1, predictor=2, root_fmm=1)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, num_phi=80, num_branches=2), dict(precision=1, predictor=2, root_fmm=1)),
    # higher-degree predictors (guarded by the quadratic) with Newton-ended roots
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20),
     dict(precision=1, predictor=4, root_fmm=1)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, num_phi=80, num_branches=2), dict(precision=1, predictor=3)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, num_phi=60, num_branches=1, approx=0), dict(precision=1, root_fmm=1)),
    # the autotuned choices (tile of the sums; root-sum FMM order and cells)
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20),
     dict(precision=1, predictor=4, root_fmm=1, autotune=1)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, num_phi=80, num_branches=2), dict(autotune=1)),
]


@pytest.mark.parametrize("name,kw,gkw", [(n, k, {}) for n, k in RUNS] + MIXED)
def test_run_matches_oracle(workdir, name, kw, gkw):
    sf = make_field(workdir, name)
    tag = f"{name}_{abs(hash(tuple(sorted(kw.items())) + tuple(sorted(gkw.items())))) % 100000}"
    gp = os.path.join(str(workdir), f"g_{tag}_")
    op = os.path.join(str(workdir), f"o_{tag}_")
    with _ctx(sf, gp, write_mu_length_scales=1, **kw, **gkw) as c:
        assert c.run(0), c.last_error()
        assert c.save(0), c.last_error()
        ginfo = c.info()
    ostats = oracle_run(sf, op, write_mu_length_scales=1, **kw)
    R, nb = ostats["num_roots"], kw["num_branches"]
    assert ginfo["num_roots"] == R
    assert ginfo["tree_levels"] == ostats["tree_levels"]
    assert ginfo["expansion_order"] == ostats["expansion_order"]
    gccs = io.read_array(gp + "ccf_ccs.bin")
    occs = io.read_array(op + "ccf_ccs.bin")
    assert gccs.shape == occs.shape == (R * nb, kw["num_phi"] // nb + 1)
    worst, nbad, bad = compare_arrays(gccs, occs, R, nb, details=True)
    if not gkw:
        # the FP64 path follows the oracle's iterates: 1e-9 relative everywhere
        assert nbad == 0, (worst, nbad)
    else:
        # other paths may stop elsewhere inside the reference's convergence
        # band: within twice the oracle's own Newton step |f0/f1| of its root
        allowed = 2 * newton_step(sf, bad, kw["num_phi"], nb, **{k: v for k, v in kw.items()
                                                                  if k not in ("num_phi", "num_branches")})
        over = [(b, a) for b, a in zip(bad, allowed) if b[3] > 1e-9 * max(abs(b[2]), 1) + a]
        assert not over, over[:5]
        assert nbad <= max(5, 1e-3 * R * (kw["num_phi"] + nb)), nbad
    gc = io.read_array(gp + "ccf_caustics.bin")
    oc = io.read_array(op + "ccf_caustics.bin")
    worst_c, nbad_c = compare_arrays(gc, oc, R, nb)
    if not gkw:
        assert nbad_c == 0, (worst_c, nbad_c)
    else:
        # the lens map amplifies a within-band root difference (|dw| reaches
        # about 2|1 - kappa| |dz| on a critical curve), so the caustics are
        # pinned pointwise instead: each one is the oracle's lens map of the
        # product's own critical-curve point
        okw = {k: v for k, v in kw.items() if k not in ("num_phi", "num_branches")}
        o, _ = oracle.evaluate(oracle.params(starfile=sf, **okw), gccs.ravel(), 0.0)
        w = o[:, 4] + 1j * o[:, 5]
        rel = np.abs(gc.ravel() - w) / np.maximum(np.abs(w), 1.0)
        assert rel.max() < 1e-11, rel.max()
        assert nbad_c <= max(5, 1e-3 * R * (kw["num_phi"] + nb)) and worst_c < 1e-6, (worst_c, nbad_c)
    gm = io.read_array(gp + "ccf_mu_length_scales.bin", complex_values=False)
    assert gm.shape == gccs.shape and np.all(np.isfinite(gm)) and np.all(gm > 0)
    # parameter files agree line by line except for the timings; the
    # error maximum is compared numerically
    ge, oe = compare_param_files(gp + "ccf_parameter_info.txt", op + "ccf_parameter_info.txt")
    check_error_maxima(ginfo, ostats, ge, oe, exact=not gkw)
    # both write the same star set (tree order may differ)
    _, _, _, _, gs = io.read_stars(gp + "ccf_stars.bin")
    _, _, _, _, os_ = io.read_stars(op + "ccf_stars.bin")
    assert np.array_equal(np.sort(gs.view("f8,f8,f8"), axis=0), np.sort(os_.view("f8,f8,f8"), axis=0))
    assert ginfo["points_converged"] == ostats["points_converged"] == R * (kw["num_phi"] + nb)
    if gkw.get("root_fmm"):
        assert ginfo["root_fmm_updates"] > 0 and ginfo["root_fmm_rows"] > 0


def test_cli_run_matches_oracle(workdir):
    """ccf_gpu end to end (main_ccf.cu's flags, -sf star file): the files it
    writes hold the oracle's critical curves and caustics"""
    import subprocess

    sf = make_field(workdir, "salpeter300")
    gp = os.path.join(str(workdir), "cli_")
    op = os.path.join(str(workdir), "cli_o_")
    r = subprocess.run([ccf_amd.CLI_PATH, "-k", "0.45", "-y", "0.35", "-sf", sf, "-np", "60", "-nb", "3", "-ws", "1",
                        "-wcc", "1", "-wc", "1", "-wls", "1", "-o", gp, "-v", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Done writing critical curve positions" in r.stdout
    kw = dict(kappa_tot=0.45, shear=0.35, num_phi=60, num_branches=3)
    ost = oracle_run(sf, op, write_mu_length_scales=1, **kw)
    R = ost["num_roots"]
    for f in ("ccf_ccs.bin", "ccf_caustics.bin"):
        worst, nbad = compare_arrays(io.read_array(gp + f), io.read_array(op + f), R, 3)
        assert nbad == 0, (f, worst, nbad)
    a = io.read_params(gp + "ccf_parameter_info.txt")
    assert a["num_roots"] == R and a["num_phi"] == 60 and a["max_error_1/mu"] < 1e-8


def test_config2_full_size_matches_oracle(workdir):
    """BASELINE config 2 at its own size: 2,000 Salpeter stars, kappa 0.45,
    shear 0.35, num_phi 1000, num_branches 10, FMM order 20; every one of
    the 4,040,000 critical-curve points against the oracle's full run"""
    sf = os.path.join(str(workdir), "cfg2.bin")
    ccf_amd.write_synthetic_stars(sf, 1, num_stars=2000, kappa_tot=0.45, kappa_star=0.45, shear=0.35,
                                  mass_function="salpeter", m_lower=0.08, m_upper=1.0)
    kw = dict(kappa_tot=0.45, shear=0.35, num_phi=1000, num_branches=10, fmm_order=20)
    gp = os.path.join(str(workdir), "c2g_")
    op = os.path.join(str(workdir), "c2o_")
    with _ctx(sf, gp, **kw) as c:
        assert c.run(0), c.last_error()
        assert c.save(0), c.last_error()
        gi = c.info()
    ost = oracle_run(sf, op, num_threads=0, **kw)
    R = ost["num_roots"]
    assert gi["num_roots"] == R == 4000 and gi["points_converged"] == ost["points_converged"]
    for f in ("ccf_ccs.bin", "ccf_caustics.bin"):
        worst, nbad = compare_arrays(io.read_array(gp + f), io.read_array(op + f), R, 10)
        assert nbad == 0, (f, worst, nbad)
    assert gi["max_error"] < 1e-8 and np.isfinite(gi["max_witt_residual"])


@pytest.mark.parametrize("gkw", [{}, dict(precision=1, predictor=4, root_fmm=1)])
def test_context_reuse_is_exact(workdir, gkw):
    """One context solving a field, a larger one, then the first again (its
    device buffers come back from the context's pool in other roles) gives
    bit-identical results to fresh contexts: the solve is deterministic and
    nothing it reads survives from an earlier run."""
    runs = [("salpeter300", dict(kappa_tot=0.45, shear=0.35, num_phi=60, num_branches=3, fmm_order=20)),
            ("rect200", dict(kappa_tot=0.7, shear=0.25, num_phi=40, num_branches=2, fmm_order=0))]

    def solve(c):
        assert c.run(0), c.last_error()
        return c.array(ccf_amd.ARRAY_CCS), c.array(ccf_amd.ARRAY_CAUSTICS), c.info()

    fresh = []
    for name, kw in runs:
        with _ctx(make_field(workdir, name), str(workdir) + "/reuse_", write_stars=0, write_critical_curves=0,
                  write_caustics=1, **kw, **gkw) as c:
            fresh.append(solve(c))
    (n0, kw0), (n1, kw1) = runs
    with _ctx(make_field(workdir, n0), str(workdir) + "/reuse_", write_stars=0, write_critical_curves=0,
              write_caustics=1, **kw0, **gkw) as c:
        seq = [solve(c)]
        c.set(starfile=make_field(workdir, n1), **kw1)
        seq.append(solve(c))
        c.set(starfile=make_field(workdir, n0), **kw0)
        seq.append(solve(c))
    for (a, ac, ai), (b, bc, bi) in zip(seq, fresh + fresh[:1]):
        assert a.shape == b.shape and np.array_equal(a, b)
        assert np.array_equal(ac, bc)
        assert ai["max_error"] == bi["max_error"] and ai["points_converged"] == bi["points_converged"]


def test_autotune_choices_persist(workdir):
    """params.autotune: the first run tunes (time spent, choices among the
    candidates), the next run of the context reuses them without tuning and
    reproduces the first run's roots bit for bit; ccf_set_params keeps them
    when only the timing switch changes and forgets them otherwise"""
    sf = make_field(workdir, "salpeter300")
    kw = dict(kappa_tot=0.45, shear=0.35, num_phi=120, num_branches=3, fmm_order=20, precision=1, predictor=4,
              root_fmm=1, autotune=1)
    with _ctx(sf, str(workdir) + "/", **kw) as c:
        assert c.run(0), c.last_error()
        a, ia = c.array(ccf_amd.ARRAY_CCS), c.info()
        assert c.run(0), c.last_error()
        b, ib = c.array(ccf_amd.ARRAY_CCS), c.info()
        c.set(timing=1)
        assert c.run(0), c.last_error()
        ic = c.info()
        c.set(timing=0, init_iters=59)
        assert c.run(0), c.last_error()
        idd = c.info()
    assert ia["t_autotune"] > 0 and ib["t_autotune"] == 0 and ic["t_autotune"] == 0 and idd["t_autotune"] > 0
    assert ia["roots_per_thread"] in (1, 2, 4) and ia["root_fmm_order"] in (6, 8, 10)
    assert ia["root_fmm_leaf"] in (4, 6, 10)
    assert (ib["roots_per_thread"], ib["root_fmm_order"], ib["root_fmm_leaf"]) == \
        (ia["roots_per_thread"], ia["root_fmm_order"], ia["root_fmm_leaf"])
    assert ia["points_converged"] == ia["points_total"] == ib["points_converged"]
    assert ia["sweeps"] == ib["sweeps"] and ia["updates_full"] == ib["updates_full"]
    assert np.array_equal(a, b)


def test_error_pass_takes_points_outside_the_tree(workdir):
    """The rectangular field's critical curves leave the star tree's root box:
    the error pass evaluates those points with a whole workgroup each, and
    their caustics and the error maximum match the oracle (the full file
    comparison is test_run_matches_oracle)."""
    sf = make_field(workdir, "rect200")
    kw = dict(kappa_tot=0.7, shear=0.25, num_phi=80, num_branches=2)
    with _ctx(sf, str(workdir) + "/out_", write_stars=0, write_critical_curves=0, write_caustics=1, **kw) as c:
        assert c.run(0), c.last_error()
        inf = c.info()
        ccs = c.array(ccf_amd.ARRAY_CCS).ravel()
        cau = c.array(ccf_amd.ARRAY_CAUSTICS).ravel()
    nstar, rect, corner, _, _ = io.read_stars(sf)
    hl = 1.1 * max(abs(corner.real), abs(corner.imag))
    outside = (np.abs(ccs.real) > hl) | (np.abs(ccs.imag) > hl)
    assert inf["points_outside_tree"] == int(outside.sum()) > 0
    okw = {k: v for k, v in kw.items() if k not in ("num_phi", "num_branches")}
    o, _ = oracle.evaluate(oracle.params(starfile=sf, **okw), ccs[outside], 0.0)
    w = o[:, 4] + 1j * o[:, 5]
    assert np.max(np.abs(cau[outside] - w) / np.maximum(np.abs(w), 1.0)) < 1e-11


def test_non_finite_errors_abort_the_run(workdir):
    """ccf.cuh:960-967: a non-finite 1/mu error aborts the run with the
    reference's message and no files are written. Two stars exactly one
    theta_E apart put one star's seed root (star + 1, ccf.cuh:716) on the
    other star, where Witt's function is infinite; the NaN that the Aberth
    update makes of it spreads through the root sums of its row. The oracle
    fails the same way on the same file."""
    import subprocess

    rng = np.random.default_rng(41)
    st = np.column_stack([rng.uniform(-6, 6, 30), rng.uniform(-6, 6, 30), np.ones(30)])
    st[0, :2] = [0.25, 0.5]
    st[1, :2] = [1.25, 0.5]
    sf = os.path.join(str(workdir), "nanfield.bin")
    io.write_stars(sf, st, 0, complex(np.sqrt(30 / 0.4), 0.0))
    kw = dict(kappa_tot=0.4, shear=0.2, num_phi=8, num_branches=1)
    msg = "Error. Errors in 1/mu contain values which are not positive real numbers."
    pre = os.path.join(str(workdir), "nan_g_")
    with _ctx(sf, pre, **kw) as c:
        assert not c.run(0)
        assert c.last_error().strip() == msg
        assert not c.save(0)
    r = subprocess.run([ccf_amd.CLI_PATH, "-k", "0.4", "-y", "0.2", "-sf", sf, "-np", "8", "-nb", "1", "-o", pre + "cli_",
                        "-v", "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and msg in r.stderr
    assert not [f for f in os.listdir(str(workdir)) if f.startswith("nan_g_")]
    with pytest.raises(RuntimeError):
        oracle_run(sf, os.path.join(str(workdir), "nan_o_"), **kw)


RS_FIELDS = [
    dict(num_stars=4000, kappa_tot=0.45, kappa_star=0.45, shear=0.35, mass_function="salpeter", m_lower=0.08, m_upper=1.0),
    dict(num_stars=4000, kappa_tot=0.7, kappa_star=0.35, shear=0.25, mass_function="kroupa", m_lower=0.08, m_upper=1.0,
         rectangular=1),
    dict(num_stars=3000, kappa_tot=0.3, kappa_star=0.3, shear=0.1, mass_function="uniform", m_lower=0.1, m_upper=2.0),
]


def _generated(seed, **kw):
    with ccf_amd.Context(random_seed=seed, num_phi=10, outfile_prefix="./", **kw) as c:
        c.prepare()
        inf = c.info()
        st = c.array(ccf_amd.ARRAY_STARS)   # tree order: a deterministic permutation of the drawn stars
    return st, inf


@pytest.mark.parametrize("spec", RS_FIELDS)
def test_random_star_field_generation(spec):
    """-rs star fields (star.cuh:40-107, ccf.cuh:602-704) drawn on the
    device: every position inside the disc or rectangle the reference draws
    in (the field of the mass function's mean mass, before the corner
    follows the drawn mean mass, ccf.cuh:686-700), positions uniform in it
    and masses distributed as the mass function (Kolmogorov-Smirnov at
    p > 1e-4, moments within 5 sigma of the closed forms), the same seed
    giving the same bits and another seed another field."""
    from scipy import stats

    st, inf = _generated(1234, **spec)
    n = spec["num_stars"]
    assert st.shape == (n, 3) and np.all(np.isfinite(st)) and inf["random_seed"] == 1234
    ml, mu = spec["m_lower"], spec["m_upper"]
    mean = ccf_amd.mass_function(spec["mass_function"], 1, 0.0, ml, mu, 1.0)
    mean2 = ccf_amd.mass_function(spec["mass_function"], 2, 0.0, ml, mu, 1.0)
    m = st[:, 2]
    assert m.min() >= ml * (1 - 1e-12) and m.max() <= mu * (1 + 1e-12)
    sd = np.sqrt(mean2 - mean ** 2)
    assert abs(m.mean() - mean) < 5 * sd / np.sqrt(n), (m.mean(), mean)
    # quantiles of the inverse CDF: the fraction of masses below m(p) is p
    cdf_pts = [ccf_amd.mass_function(spec["mass_function"], 0, p, ml, mu, 1.0) for p in (0.1, 0.3, 0.5, 0.7, 0.9)]
    for p, mp in zip((0.1, 0.3, 0.5, 0.7, 0.9), cdf_pts):
        assert abs((m < mp).mean() - p) < 5 * np.sqrt(p * (1 - p) / n), (p, (m < mp).mean())
    corner = complex(inf["corner_re"], inf["corner_im"])
    drawn = np.sqrt(mean / m.mean())   # the corner before it followed the drawn mean mass
    if spec.get("rectangular"):
        cx, cy = corner.real * drawn, corner.imag * drawn
        assert np.all(np.abs(st[:, 0]) <= cx * (1 + 1e-12)) and np.all(np.abs(st[:, 1]) <= cy * (1 + 1e-12))
        assert stats.kstest(st[:, 0] / cx, "uniform", args=(-1, 2)).pvalue > 1e-4
        assert stats.kstest(st[:, 1] / cy, "uniform", args=(-1, 2)).pvalue > 1e-4
    else:
        R = abs(corner) * drawn
        r2 = (st[:, 0] ** 2 + st[:, 1] ** 2) / R ** 2
        assert r2.max() <= 1 + 1e-12 and r2.max() > 0.99
        assert stats.kstest(r2, "uniform").pvalue > 1e-4
        ang = np.mod(np.arctan2(st[:, 1], st[:, 0]), 2 * np.pi) / (2 * np.pi)
        assert stats.kstest(ang, "uniform").pvalue > 1e-4
    # kappa_star_actual is taken before the corner follows the drawn mean
    # mass (ccf.cuh:676-700): kappa_star scaled by drawn / expected mean mass
    assert inf["kappa_star_actual"] == pytest.approx(spec["kappa_star"] * m.mean() / mean, rel=1e-9)
    again, _ = _generated(1234, **spec)
    assert np.array_equal(st, again)
    other, _ = _generated(1235, **spec)
    assert not np.array_equal(st[:, :2], other[:, :2])


def test_random_seed_zero_takes_the_clock():
    """ccf.cuh:619-622: seed 0 is replaced by the clock (never 0), and the
    field is the one that seed gives when it is set explicitly"""
    spec = RS_FIELDS[0]
    st, inf = _generated(0, **spec)
    seed = inf["random_seed"]
    assert seed != 0
    again, inf2 = _generated(seed, **spec)
    assert np.array_equal(st, again)
    import time
    time.sleep(0.01)
    _, inf3 = _generated(0, **spec)
    assert inf3["random_seed"] not in (0, seed)


def test_cli_random_field_matches_oracle(workdir):
    """ccf_gpu -rs: a seeded field generated on the device, solved, and the
    files it writes; the oracle solves the star file that run wrote, and the
    critical curves, caustics and parameter file agree (the parameter file's
    random_seed and mass-function lines are the generated run's own)."""
    import subprocess

    gp = os.path.join(str(workdir), "rs_")
    op = os.path.join(str(workdir), "rs_o_")
    r = subprocess.run([ccf_amd.CLI_PATH, "-k", "0.45", "-y", "0.35", "-ks", "0.45", "-ns", "300", "-mf", "salpeter",
                        "-ml", "0.08", "-mh", "1", "-rs", "4242", "-np", "60", "-nb", "3", "-ws", "1", "-wcc", "1",
                        "-wc", "1", "-o", gp, "-v", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    kw = dict(kappa_tot=0.45, shear=0.35, num_phi=60, num_branches=3)
    ost = oracle_run(gp + "ccf_stars.bin", op, **kw)
    R = ost["num_roots"]
    assert R == 600
    for f in ("ccf_ccs.bin", "ccf_caustics.bin"):
        worst, nbad = compare_arrays(io.read_array(gp + f), io.read_array(op + f), R, 3)
        assert nbad == 0, (f, worst, nbad)
    a = io.read_params(gp + "ccf_parameter_info.txt")
    b = io.read_params(op + "ccf_parameter_info.txt")
    assert a["random_seed"] == 4242 and a["mass_function"] == "salpeter" and "kappa_star_actual" in a
    for k in ("kappa_tot", "shear", "num_stars", "num_roots", "num_phi", "num_branches", "rad", "mean_mass_actual"):
        assert a[k] == pytest.approx(b[k], rel=1e-8), k
