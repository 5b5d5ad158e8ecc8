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


# ------------------------------------- lens map, 1/mu bound, length scale, sheets
# These pin the oracle's lens map w(z) (ccf_functions.cuh:33-45), its 1/mu
# error bound (:449-463), its magnification length scale (:144-166) and the
# smooth-matter sheets (alpha_smooth.cuh:62-301) by what the mathematics fixes:
# finite differences of w against Witt's function, the Jacobian determinant of
# w, the magnification of the two images of a source beside a fold caustic,
# and a quadrature of the uniform disc and rectangle. All of them evaluate the
# direct-sum field (the tree's local expansions are pinned against it above).
FD_H = 1e-5


def _eval(sf, kw, z, phi=0.0):
    v, s = oracle.evaluate(oracle.params(starfile=sf, num_phi=10, **kw), z, phi, use_tree=False)
    return v, s


def _cx(v, col):
    return v[:, col] + 1j * v[:, col + 1]


def _jacobian(sf, kw, z, h=FD_H):
    """dw/dz and dw/dzbar of the oracle's lens map by central differences"""
    z = np.asarray(z, dtype=np.complex128)
    v, _ = _eval(sf, kw, np.concatenate([z + h, z - h, z + 1j * h, z - 1j * h]))
    w = _cx(v, 4).reshape(4, -1)
    wx = (w[0] - w[1]) / (2 * h)
    wy = (w[2] - w[3]) / (2 * h)
    return 0.5 * (wx - 1j * wy), 0.5 * (wx + 1j * wy)


def _field_kappa_star(sf):
    """kappa_star of a star file from its own contents: theta^2 sum(m) / |c|^2
    for the disc, pi theta^2 sum(m) / (4 c1 c2) for the rectangle"""
    n, rect, c, theta, st = io.read_stars(sf)
    if rect:
        return np.pi * theta ** 2 * st[:, 2].sum() / (4 * c.real * c.imag)
    return theta ** 2 * st[:, 2].sum() / abs(c) ** 2


def _off_star_points(sf, rng, count, spread=0.6, min_dist=0.25, edge_gap=0.02):
    """random points near the stars, at least min_dist from every star and,
    for the disc / rectangle edge where the exact sheets jump, edge_gap away"""
    n, rect, c, theta, st = io.read_stars(sf)
    s = st[:, 0] + 1j * st[:, 1]
    z = s[rng.integers(0, n, 4 * count)] + spread * np.exp(2j * np.pi * rng.random(4 * count)) * (0.5 + rng.random(4 * count))
    far = 1.4 * max(abs(c.real), abs(c.imag), abs(c)) * np.exp(2j * np.pi * rng.random(count // 4))
    z = np.concatenate([z, far])
    ok = np.abs(z[:, None] - s[None, :]).min(axis=1) > min_dist
    if rect:
        ok &= (np.abs(np.abs(z.real) - abs(c.real)) > edge_gap) & (np.abs(np.abs(z.imag) - abs(c.imag)) > edge_gap)
    else:
        ok &= np.abs(np.abs(z) - abs(c)) > edge_gap
    return z[ok][:count]


def _inside(sf, z):
    n, rect, c, theta, st = io.read_stars(sf)
    if rect:
        return (np.abs(z.real) <= abs(c.real)) & (np.abs(z.imag) <= abs(c.imag))
    return np.abs(z) <= abs(c)


def _dzbar_extra(sf, kw, z):
    """What the reference's d(alpha_smooth)/dzbar holds beyond the derivative
    of its own alpha_smooth. The exact circular sheet's derivative is
    kappa_star R^2 / zbar^2 on both sides of the edge (alpha_smooth.cuh:226-233)
    while alpha_smooth is -kappa_star z inside the disc (:113-116), so inside
    the disc f0 and f1 carry a kappa_star R^2 / z^2 term that w does not (the
    pole at z = 0 the Aberth sums add for this sheet, ccf_functions.cuh:223-226,
    is that term's). Zero for every other sheet."""
    n, rect, c, theta, st = io.read_stars(sf)
    if rect or kw.get("approx", 1):
        return np.zeros(np.shape(z), complex)
    return np.where(np.abs(z) <= abs(c), _field_kappa_star(sf) * abs(c) ** 2 / np.conj(z) ** 2, 0)


LENS_CASES = [
    ("cfg1", dict(kappa_tot=0.3, shear=0.1)),                          # circular, approximate sheet
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),           # circular, exact sheet
    ("smooth150", dict(kappa_tot=0.6, shear=0.2)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25)),                      # rectangular, Taylor series
    ("rect200", dict(kappa_tot=0.7, shear=0.25, approx=0)),            # rectangular, exact sheet
]


@pytest.mark.parametrize("name,kw", LENS_CASES)
def test_lens_map_jacobian_matches_witt_function(workdir, name, kw):
    """Central differences of the lens map w give
      dw/dz = 1 - kappa - d(alpha_smooth)/dz, a real number: 1 - kappa +
        kappa_star where the sheet covers z (everywhere for the approximate
        sheets) and 1 - kappa outside the exact ones;
      conj(dw/dzbar) = f0(z; phi) + (dw/dz) e^{-i phi} for every phi, where
        f0 is Witt's function (pinned by the polynomial roots above), plus
        inside the exact disc the reference's extra sheet term (_dzbar_extra).
    A sign or conjugation slip in any deflection term of w, or a smooth term
    missing from w or f0, breaks one of the two by order one."""
    sf = make_field(workdir, name)
    rng = np.random.default_rng(23)
    z = _off_star_points(sf, rng, 120)
    a, b = _jacobian(sf, kw, z)
    ks = _field_kappa_star(sf)
    cover = np.ones(z.size, bool) if kw.get("approx", 1) else _inside(sf, z)
    expect_a = 1 - kw["kappa_tot"] + ks * cover
    assert np.abs(a - expect_a).max() < 1e-7, np.abs(a - expect_a).max()
    for phi in (0.0, 0.9, 2.5, 4.4):
        v, s = _eval(sf, kw, z, phi)
        assert s["kappa_star"] == pytest.approx(ks, rel=1e-12)
        rhs = _cx(v, 0) + expect_a * np.exp(-1j * phi) + np.conj(_dzbar_extra(sf, kw, z))
        assert np.abs(np.conj(b) - rhs).max() < 1e-7, (phi, np.abs(np.conj(b) - rhs).max())


def _sheet_quadrature(sf, z):
    """(1/pi) integral over the disc or rectangle of dA / conj(z - s), by
    rays from z: with s = z + r e^{i t}, dA / conj(z - s) = -e^{i t} dr dt,
    so the integral is -(1/pi) int e^{i t} (r_out - r_in) dt over the chord
    [r_in, r_out] (clipped to r >= 0) that each ray cuts from the convex set"""
    from scipy.integrate import quad

    n, rect, c, theta, st = io.read_stars(sf)

    def chord(t, z0):
        e = np.exp(1j * t)
        if rect:
            lo, hi = -np.inf, np.inf
            for p0, d, half in ((z0.real, e.real, abs(c.real)), (z0.imag, e.imag, abs(c.imag))):
                if abs(d) < 1e-300:
                    if abs(p0) > half:
                        return 0.0
                    continue
                r1, r2 = sorted(((-half - p0) / d, (half - p0) / d))
                lo, hi = max(lo, r1), min(hi, r2)
            if hi <= lo:
                return 0.0
        else:
            R = abs(c)
            bq = (np.conj(z0) * e).real
            disc = bq * bq - (abs(z0) ** 2 - R * R)
            if disc <= 0:
                return 0.0
            lo, hi = -bq - np.sqrt(disc), -bq + np.sqrt(disc)
        return max(hi, 0.0) - max(lo, 0.0)

    out = []
    for z0 in z:
        if rect:
            brk = sorted({float(np.angle(complex(sx * abs(c.real), sy * abs(c.imag)) - z0)) % (2 * np.pi)
                          for sx in (-1, 1) for sy in (-1, 1)})
        elif abs(z0) > abs(c):
            # the chord length has square-root ends where the ray grazes the disc
            t0, dt = float(np.angle(-z0)), float(np.arcsin(abs(c) / abs(z0)))
            brk = sorted({(t0 - dt) % (2 * np.pi), (t0 + dt) % (2 * np.pi)})
        else:
            brk = []
        re = quad(lambda t: -np.cos(t) * chord(t, z0), 0, 2 * np.pi, points=brk or None, limit=400, epsabs=1e-12, epsrel=1e-11)[0]
        im = quad(lambda t: -np.sin(t) * chord(t, z0), 0, 2 * np.pi, points=brk or None, limit=400, epsabs=1e-12, epsrel=1e-11)[0]
        out.append(complex(re, im) / np.pi)
    return np.array(out)


@pytest.mark.parametrize("name,kw", [
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),   # circular exact: alpha_smooth.cuh:107-125
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, approx=0)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2)),             # circular approximate: infinite sheet
    ("rect200", dict(kappa_tot=0.7, shear=0.25, approx=0)),    # rectangular exact: :96-105
])
def test_smooth_sheet_is_the_uniform_sheet_deflection(workdir, name, kw):
    """The oracle's smooth deflection, alpha_smooth = (1 - kappa) z + gamma
    zbar - alpha_star - w with alpha_star = theta^2 sum m / conj(z - s) summed
    here, against the deflection of a sheet of convergence -kappa_star:
    -kappa_star (1/pi) int dA / conj(z - s) over the disc (exact circular:
    -kappa_star z inside, -kappa_star R^2 / zbar outside) or the rectangle,
    by quadrature; the approximate circular sheet is the infinite one,
    -kappa_star z everywhere. Points inside and outside the field."""
    sf = make_field(workdir, name)
    n, rect, c, theta, st = io.read_stars(sf)
    rng = np.random.default_rng(29)
    z = _off_star_points(sf, rng, 40, spread=2.0)
    z = np.concatenate([z, np.array([0.3 * c, 0.9 * c + 0.1j, 1.3 * abs(c) * np.exp(0.4j), 2.5 * abs(c) * np.exp(2.1j)])])
    if rect:
        z = np.concatenate([z, np.array([complex(0.5 * c.real, 0.97 * c.imag), complex(1.2 * c.real, 0.2 * c.imag),
                                         complex(-0.2 * c.real, -1.5 * c.imag)])])
    v, s = _eval(sf, kw, z)
    ks = _field_kappa_star(sf)
    s_ = st[:, 0] + 1j * st[:, 1]
    a_star = theta ** 2 * (st[None, :, 2] / np.conj(z[:, None] - s_[None, :])).sum(axis=1)
    k, g = kw["kappa_tot"], kw["shear"]
    a_smooth = (1 - k) * z + g * np.conj(z) - a_star - _cx(v, 4)
    if kw.get("approx", 1) and not rect:
        expect = -ks * z
    else:
        expect = -ks * _sheet_quadrature(sf, z)
    err = np.abs(a_smooth - expect) / np.maximum(np.abs(expect), 1.0)
    assert err.max() < 1e-9, (err.max(), z[np.argmax(err)])


@pytest.mark.parametrize("name,kw", LENS_CASES)
def test_error_bound_bounds_inverse_magnification(workdir, name, kw):
    """1/mu = |dw/dz|^2 - |dw/dzbar|^2 (the Jacobian determinant of w, from
    central differences) never exceeds the oracle's bound
    max |f0| | |f0| +- 2 (1 - kappa - d alpha_smooth/dz) | at any phi, and
    equals it where f0 is parallel to (dw/dz) e^{-i phi} with
    |dw/dzbar| > dw/dz (phi = -arg conj(dw/dzbar)): the bound is the extreme
    of 1/mu over the unknown direction of f0."""
    sf = make_field(workdir, name)
    rng = np.random.default_rng(31)
    z = _off_star_points(sf, rng, 120)
    a, b = _jacobian(sf, kw, z)
    b = b - _dzbar_extra(sf, kw, z)          # the reference's d w / d zbar (equal to b but inside the exact disc)
    inv_mu = np.abs(a) ** 2 - np.abs(b) ** 2
    for phi in (0.3, 1.7, 3.9):
        v, _ = _eval(sf, kw, z, phi)
        assert np.all(np.abs(inv_mu) <= v[:, 7] * (1 + 1e-7) + 1e-7)
        assert np.any(np.abs(inv_mu) < 0.9 * v[:, 7])   # not attained at a random phi
    sel = np.abs(b) > np.abs(a) + 1e-3
    assert sel.sum() >= 5
    for zi, ai, bi, im in zip(z[sel], a[sel], b[sel], inv_mu[sel]):
        phi = -np.angle(np.conj(bi))
        v, _ = _eval(sf, kw, np.array([zi]), phi)
        assert v[0, 7] == pytest.approx(abs(im), rel=1e-6, abs=1e-8)


def _critical_points(sf, kw, phis, per_phi):
    """points of the oracle's critical curve at the angles phis: roots of an
    oracle run, polished by Newton steps z - f0/f1 of the direct-sum field"""
    pre = sf + "_cp_"
    s = oracle.run(oracle.params(starfile=sf, outfile_prefix=pre, num_phi=8, num_branches=1, direct=1,
                                 write_stars=0, write_caustics=0, **kw))
    cols = io.columns_by_phi(io.read_array(pre + "ccf_ccs.bin"), s["num_roots"], 1)[0]
    pts = []
    for k in phis:
        phi = 2 * np.pi * k / 8
        z = cols[k][:: max(1, cols[k].size // per_phi)][:per_phi]
        for _ in range(3):
            v, _ = _eval(sf, kw, z, phi)
            z = z - _cx(v, 0) / _cx(v, 2)
        pts += [(zi, phi) for zi in z]
    return pts


@pytest.mark.parametrize("name,kw", [
    ("cfg1", dict(kappa_tot=0.3, shear=0.1)),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25)),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, approx=0)),
])
def test_mu_length_scale_is_the_fold_magnification(workdir, name, kw):
    """The length scale d0 is defined by mu = sqrt(d0 / d) for an image of a
    source a distance d from a fold caustic. Here, for points z_c on the
    oracle's critical curve: the source is put at w(z_c) + d N with N the
    caustic's normal (the range of the lens Jacobian at z_c is the tangent
    line), the two images beside z_c are found by Newton's method on w, and
    each image's |1/det J| (central differences) must be sqrt(d0 / d) to
    the fold approximation's order sqrt(d / d0), whose first-order term has
    opposite signs for the two images (their mean is held to 1e-2; near a
    cusp the approximation holds only very close to the caustic, hence
    d = 1e-10 d0). A factor in d0, or a wrong derivative or conjugate in the
    tangents it is formed from, fails this."""
    sf = make_field(workdir, name)
    n, rect, c, theta, st = io.read_stars(sf)
    s_ = st[:, 0] + 1j * st[:, 1]
    pts = _critical_points(sf, kw, (1, 3, 6), 3)
    checked = 0
    for zc, phi in pts:
        if np.abs(zc - s_).min() < 0.2:
            continue
        if not kw.get("approx", 1) and rect and min(abs(abs(zc.real) - c.real), abs(abs(zc.imag) - c.imag)) < 0.05:
            continue
        v, _ = _eval(sf, kw, np.array([zc]), phi)
        d0 = v[0, 6]
        a, b = _jacobian(sf, kw, np.array([zc]))
        a, b = float(a[0].real), b[0]
        assert abs(abs(b) - abs(a)) < 1e-6                 # on the critical curve: det J = 0
        N = 1j * np.exp(0.5j * np.angle(b))                # normal to the range of J
        tdir = 1j * np.exp(0.5j * np.angle(b))             # kernel of J (image-plane direction)
        wc = _cx(_eval(sf, kw, np.array([zc]))[0], 4)[0]
        d = 1e-10 * d0
        vscale = 4 * abs(a) * np.sqrt(d * d0)              # |v| = sqrt(d / K), K = 1 / (16 a^2 d0)
        found = None
        for side in (1, -1):
            y = wc + side * d * N
            imgs = []
            for sgn in (1, -1):
                zz = zc + sgn * vscale * tdir
                for _ in range(40):
                    a2, b2 = _jacobian(sf, kw, np.array([zz]))
                    r = _cx(_eval(sf, kw, np.array([zz]))[0], 4)[0] - y
                    # solve a2 dz + b2 conj(dz) = -r  (2x2 real system)
                    A = np.array([[(a2[0] + b2[0]).real, (-a2[0] + b2[0]).imag],
                                  [(a2[0] + b2[0]).imag, (a2[0] - b2[0]).real]])
                    dz = np.linalg.solve(A, [-r.real, -r.imag])
                    zz = zz + complex(dz[0], dz[1])
                    if abs(complex(dz[0], dz[1])) < 1e-15 * max(1.0, abs(zz)):
                        break
                if abs(_cx(_eval(sf, kw, np.array([zz]))[0], 4)[0] - y) < 1e-13 * max(1, abs(y)) and abs(zz - zc) < 10 * vscale:
                    imgs.append(zz)
            if len(imgs) == 2 and abs(imgs[0] - imgs[1]) > 0.5 * vscale:
                found = imgs
                break
        assert found is not None, (zc, phi)
        ratios = []
        for zi in found:
            a2, b2 = _jacobian(sf, kw, np.array([zi]), h=1e-6)
            mu = 1 / (abs(a2[0]) ** 2 - abs(b2[0]) ** 2)
            ratios.append(abs(mu) * np.sqrt(d / d0))
        assert np.all(np.abs(np.array(ratios) - 1) < 0.1), (zc, phi, ratios, d0)
        assert np.mean(ratios) == pytest.approx(1.0, abs=1e-2), (zc, phi, ratios, d0)
        checked += 1
    assert checked >= 5


@pytest.mark.parametrize("plant,check,name,kw", [
    (oracle.PLANT_W_STAR_UNCONJUGATED, "jacobian", "salpeter300", dict(kappa_tot=0.45, shear=0.35)),
    (oracle.PLANT_W_STAR_UNCONJUGATED, "sheet", "cfg1", dict(kappa_tot=0.3, shear=0.1, approx=0)),
    (oracle.PLANT_BOUND_SMOOTH_SIGN, "bound", "salpeter300", dict(kappa_tot=0.45, shear=0.35)),
    (oracle.PLANT_MU_SCALE_FACTOR, "mu", "cfg1", dict(kappa_tot=0.3, shear=0.1)),
    (oracle.PLANT_DISC_OUTSIDE_Z, "sheet", "smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),
    (oracle.PLANT_DISC_DZBAR_INSIDE, "jacobian", "smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0)),
])
def test_pins_catch_planted_mistakes(workdir, plant, check, name, kw):
    """Each pin above fails when the oracle carries the slip it guards
    against (planted in the oracle itself through oracle_params.plant)."""
    fn = {"jacobian": test_lens_map_jacobian_matches_witt_function,
          "sheet": test_smooth_sheet_is_the_uniform_sheet_deflection,
          "bound": test_error_bound_bounds_inverse_magnification,
          "mu": test_mu_length_scale_is_the_fold_magnification}[check]
    fn(workdir, name, kw)  # passes without the mistake
    with pytest.raises(AssertionError):
        fn(workdir, name, dict(kw, plant=plant))
