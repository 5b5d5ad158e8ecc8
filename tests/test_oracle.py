"""Pins of the FP64 oracle (CPU). Each check fixes a part of the restated
algorithm independently of the oracle's own code: closed forms, a library
routine at 50 digits, direct sums, finite differences and symmetries."""
import math
import os

import numpy as np
import pytest
from scipy import integrate

import ccf_amd
import oracle
from ccf_amd import io
from helpers import FIELDS, make_field

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "witt_poly_roots.npz")


# ---------------------------------------------------------------- binomials
def test_binomials_exact_and_wrapped():
    # binomial_coefficients.cuh builds Pascal's rows in int32: entries past
    # row 33 are C(n, k) mod 2^32 read as signed (row 34 is the first to wrap)
    for n in range(0, 62):
        for k in range(0, n + 1):
            c = math.comb(n, k)
            assert oracle.binomial(n, k, False) == float(c)
            w = c % 2**32
            w = w - 2**32 if w >= 2**31 else w
            assert oracle.binomial(n, k, True) == float(w)
    assert oracle.binomial(33, 16, True) == math.comb(33, 16)
    assert oracle.binomial(34, 17, True) != math.comb(34, 17)


# ---------------------------------------------------------- mass functions
def _pdf_kroupa(m, ms=1.0):
    # continuous broken power law, slopes -0.3 / -1.3 / -2.3 at 0.08, 0.5 (unnormalised)
    m0, m1 = 0.08 * ms, 0.5 * ms
    if m < m0:
        return m ** -0.3
    if m < m1:
        return m0 ** (-0.3 + 1.3) * m ** -1.3
    return m0 ** (-0.3 + 1.3) * m1 ** (-1.3 + 2.3) * m ** -2.3


@pytest.mark.parametrize("name,ml,mu,pdf", [
    ("salpeter", 0.08, 1.0, lambda m: m ** -2.35),
    ("uniform", 0.1, 2.0, lambda m: 1.0),
    ("kroupa", 0.08, 1.0, _pdf_kroupa),
    ("kroupa", 0.01, 1.0, _pdf_kroupa),
    ("kroupa", 0.1, 0.4, _pdf_kroupa),
    ("kroupa", 0.6, 5.0, _pdf_kroupa),
])
def test_mass_function_moments_and_inverse_cdf(name, ml, mu, pdf):
    pts = [x for x in (0.08, 0.5) if ml < x < mu]
    Z = integrate.quad(pdf, ml, mu, points=pts, limit=200)[0]
    mom = [integrate.quad(lambda m: f(m) * pdf(m), ml, mu, points=pts, limit=200)[0] / Z
           for f in (lambda m: m, lambda m: m * m, lambda m: m * m * math.log(m))]
    for which in (1, 2, 3):
        for impl in (oracle.mass_function, ccf_amd.mass_function):
            v = impl(name, which, 0.0, ml, mu, 1.0)
            assert v == pytest.approx(mom[which - 1], rel=1e-8, abs=1e-12), (impl, which)
    for p in (0.001, 0.2, 0.5, 0.77, 0.999):
        for impl in (oracle.mass_function, ccf_amd.mass_function):
            m = impl(name, 0, p, ml, mu, 1.0)
            cdf = integrate.quad(pdf, ml, m, points=[x for x in pts if x < m], limit=200)[0] / Z
            assert cdf == pytest.approx(p, abs=1e-9)


def test_salpeter_mean_mass_published():
    # <m> of Salpeter on [0.08, 1] quoted for the 2k and 10k fields: 0.18728
    assert oracle.mass_function("salpeter", 1, 0, 0.08, 1.0) == pytest.approx(0.18728, abs=5e-6)


def test_equal_and_optical_depth():
    assert oracle.mass_function("equal", 1, 0, 1, 1) == 1 and oracle.mass_function("equal", 3, 0, 1, 1) == 0
    ml, mu = 0.2, 1.4
    # two masses with equal total mass: number fractions mu/(ml+mu) and ml/(ml+mu)
    fl, fu = mu / (ml + mu), ml / (ml + mu)
    assert fl * ml == pytest.approx(fu * mu)
    for impl in (oracle.mass_function, ccf_amd.mass_function):
        assert impl("optical_depth", 1, 0, ml, mu, 1) == pytest.approx(fl * ml + fu * mu)
        assert impl("optical_depth", 2, 0, ml, mu, 1) == pytest.approx(fl * ml ** 2 + fu * mu ** 2)
        assert impl("optical_depth", 3, 0, ml, mu, 1) == pytest.approx(
            fl * ml ** 2 * math.log(ml) + fu * mu ** 2 * math.log(mu))
        assert impl("optical_depth", 0, 0.3, ml, mu, 1) == ml and impl("optical_depth", 0, 0.9, ml, mu, 1) == mu
    assert math.isnan(oracle.mass_function("nope", 1, 0, 1, 2))


# ------------------------------------------------------ FMM against direct sums
@pytest.mark.parametrize("name,kw,tol", [
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, binomial_exact=1), 2e-7),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35), 2e-6),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, binomial_exact=1), 2e-7),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0, binomial_exact=1), 2e-7),
])
def test_fmm_matches_direct_sums(workdir, name, kw, tol):
    """Tree + P2M/M2M/M2L/L2L/P2L + L2P against summing every star: a wrong
    sign, index or transposed operand in any operator breaks this by O(1)."""
    sf = make_field(workdir, name)
    n, rect, corner, theta, st = io.read_stars(sf)
    rng = np.random.default_rng(3)
    z = st[rng.integers(0, n, 400), 0] + 1j * st[rng.integers(0, n, 400), 1] + 0.7 * np.exp(2j * np.pi * rng.random(400))
    p = oracle.params(starfile=sf, num_phi=10, **kw)
    a, s = oracle.evaluate(p, z, 0.7, use_tree=True)
    b, _ = oracle.evaluate(p, z, 0.7, use_tree=False)
    assert s["tree_levels"] >= 2 and s["expansion_order"] >= 20
    for col in (0, 2, 4):
        d = np.abs((a[:, col] + 1j * a[:, col + 1]) - (b[:, col] + 1j * b[:, col + 1]))
        scale = np.maximum(np.abs(b[:, col] + 1j * b[:, col + 1]), 1.0)
        assert (d / scale).max() < tol, (col, (d / scale).max())


@pytest.mark.parametrize("name,kw", [
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, approx=0)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),
])
def test_f1_is_derivative_of_f0(workdir, name, kw):
    sf = make_field(workdir, name)
    n, rect, corner, theta, st = io.read_stars(sf)
    rng = np.random.default_rng(5)
    z = st[rng.integers(0, n, 100), 0] + 1j * st[rng.integers(0, n, 100), 1] + 0.6 * np.exp(2j * np.pi * rng.random(100))
    p = oracle.params(starfile=sf, num_phi=10, **kw)
    h = 1e-5
    fp, _ = oracle.evaluate(p, z + h, 1.1)
    fm, _ = oracle.evaluate(p, z - h, 1.1)
    f, _ = oracle.evaluate(p, z, 1.1)
    fd = ((fp[:, 0] + 1j * fp[:, 1]) - (fm[:, 0] + 1j * fm[:, 1])) / (2 * h)
    f1 = f[:, 2] + 1j * f[:, 3]
    ok = np.abs(fd - f1) <= 1e-6 * np.maximum(np.abs(f1), 1.0)
    # the exact rectangular sheet is discontinuous on the field's edge
    assert ok.mean() > 0.97 if kw.get("approx", 1) == 0 else ok.all()


def test_rectangular_taylor_series_matches_exact_sheet(workdir):
    """The approximate rectangular sheet (Taylor series of degree
    taylor_smooth) agrees with the exact 4-logarithm form inside the field
    to the error bound the degree was chosen for (ccf.cuh:441-447)."""
    sf = make_field(workdir, "rect200")
    n, rect, corner, theta, st = io.read_stars(sf)
    rng = np.random.default_rng(11)
    # inside |z| < |corner| / safety_scale, where the series' error bound holds
    z = 0.6 * (corner.real * (2 * rng.random(200) - 1) + 1j * corner.imag * (2 * rng.random(200) - 1))
    kw = dict(starfile=sf, num_phi=10, kappa_tot=0.7, shear=0.25, direct=1)
    a, _ = oracle.evaluate(oracle.params(approx=1, **kw), z, 0.3)
    b, _ = oracle.evaluate(oracle.params(approx=0, **kw), z, 0.3)
    # lens maps differ by a constant (the two forms fix alpha_smooth(0) differently)
    dw = (a[:, 4] + 1j * a[:, 5]) - (b[:, 4] + 1j * b[:, 5])
    assert np.abs(dw - dw.mean()).max() < 1e-6
    df0 = np.abs((a[:, 0] + 1j * a[:, 1]) - (b[:, 0] + 1j * b[:, 1]))
    assert df0.max() < 1e-6


# ------------------------------------------------------------ root solves
def _golden_cases():
    d = np.load(GOLDEN)
    for i in range(int(d["ncases"])):
        yield {k: d[f"c{i}_{k}"] for k in ("stars", "rad", "kappa", "kappa_star", "gamma", "num_phi", "roots")}


@pytest.mark.parametrize("case", list(range(4)))
def test_oracle_roots_match_polynomial_roots(workdir, case):
    """Oracle critical curves of tiny fields against the 50-digit roots of
    the numerator polynomial of Witt's function (tests/golden)."""
    c = list(_golden_cases())[case]
    sf = os.path.join(str(workdir), f"golden{case}.bin")
    io.write_stars(sf, c["stars"], 0, complex(float(c["rad"]), 0.0))
    pre = os.path.join(str(workdir), f"gold{case}_")
    npp = int(c["num_phi"])
    s = oracle.run(oracle.params(starfile=sf, outfile_prefix=pre, kappa_tot=float(c["kappa"]), shear=float(c["gamma"]),
                                 num_phi=npp, num_branches=1))
    assert s["kappa_star"] == pytest.approx(float(c["kappa_star"]), rel=1e-12)
    ccs = io.read_array(pre + "ccf_ccs.bin")
    R = s["num_roots"]
    assert R == 2 * len(c["stars"])
    cols = io.columns_by_phi(ccs, R, 1)[0]
    from helpers import match_column
    for k in range(npp + 1):
        d = match_column(cols[k], c["roots"][k])
        assert (d / np.maximum(np.abs(c["roots"][k]), 1)).max() < 1e-9, (k, d.max())


def test_chang_refsdal_closed_form(workdir):
    """One star at the origin: z = +-theta / sqrt((1 - kappa + kappa_star) e^{-i phi} - gamma)."""
    kappa, ks, gamma = 0.5, 0.3, 0.15
    sf = os.path.join(str(workdir), "cr.bin")
    io.write_stars(sf, np.array([[0.0, 0.0, 1.0]]), 0, complex(math.sqrt(1 / ks), 0))
    pre = os.path.join(str(workdir), "cr_")
    s = oracle.run(oracle.params(starfile=sf, outfile_prefix=pre, kappa_tot=kappa, shear=gamma, num_phi=40, num_branches=2))
    ccs = io.read_array(pre + "ccf_ccs.bin")
    cols = io.columns_by_phi(ccs, 2, 2)
    for c in range(2):
        for k in range(cols.shape[1]):
            phi = 2 * np.pi * (c * 20 + k) / 40
            r = 1 / np.sqrt((1 - kappa + ks) * np.exp(-1j * phi) - gamma)
            from helpers import match_column
            assert match_column(cols[c, k], np.array([r, -r])).max() < 1e-10
    assert s["points_converged"] == s["points_solved"]


def _symmetric_field(workdir, kind):
    rng = np.random.default_rng(17)
    half = rng.uniform(-4, 4, (6, 2))
    m = rng.uniform(0.5, 1.5, 6)
    if kind == "mirror":   # z -> conj(z)
        other = np.stack([half[:, 0], -half[:, 1]], axis=1)
    else:                  # z -> -z
        other = -half
    st = np.concatenate([np.column_stack([half, m]), np.column_stack([other, m])])
    sf = os.path.join(str(workdir), f"sym_{kind}.bin")
    io.write_stars(sf, st, 0, complex(math.sqrt(st[:, 2].sum() / 0.35), 0))
    return sf


def test_mirror_and_point_symmetry(workdir):
    """A field symmetric under z -> conj(z) has roots(2 pi - phi) = conj(roots(phi));
    one symmetric under z -> -z has roots in +- pairs."""
    for kind in ("mirror", "point"):
        sf = _symmetric_field(workdir, kind)
        pre = os.path.join(str(workdir), f"sym_{kind}_")
        s = oracle.run(oracle.params(starfile=sf, outfile_prefix=pre, kappa_tot=0.5, shear=0.2, num_phi=40, num_branches=1))
        R = s["num_roots"]
        cols = io.columns_by_phi(io.read_array(pre + "ccf_ccs.bin"), R, 1)[0]
        from helpers import match_column
        for k in range(41):
            if kind == "mirror":
                d = match_column(np.conj(cols[40 - k]), cols[k])
            else:
                d = match_column(-cols[k], cols[k])
            assert d.max() < 1e-8, (kind, k, d.max())


def test_continuation_invariants(workdir):
    """Witt residual at every converged root, and the rows phi = 0 and
    phi = 2 pi (solved from different branch centres) hold the same set."""
    sf = make_field(workdir, "salpeter300")
    pre = os.path.join(str(workdir), "inv_")
    kw = dict(kappa_tot=0.45, shear=0.35, num_phi=60, num_branches=3)
    s = oracle.run(oracle.params(starfile=sf, outfile_prefix=pre, **kw))
    R = s["num_roots"]
    cols = io.columns_by_phi(io.read_array(pre + "ccf_ccs.bin"), R, 3)
    from helpers import match_column
    d = match_column(cols[0, 0], cols[2, -1])
    assert d.max() < 1e-8
    p = oracle.params(starfile=sf, **kw)
    for c, k in [(0, 0), (1, 10), (2, 20)]:
        phi = 2 * np.pi * (c * 20 + k) / 60
        v, _ = oracle.evaluate(p, cols[c, k], phi)
        assert v[:, 7].max() < 1e-8  # 1/mu bound at the roots
    assert s["points_converged"] == s["points_solved"]
