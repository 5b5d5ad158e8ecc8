"""ccf_aberth_sweep_noalloc, the sweep entry point that allocates nothing,
against ccf_aberth_sweep (the same kernels over buffers from the context's
pool): bit-identical updates on two calls in a row, the second with a
smaller row than the first (the fixed scratch array keeps the first call's
contents), in the FP64 and mixed forms, with direct star sums, the star
FMM and the root-sum FMM; a row too large for the array is refused."""
import numpy as np
import pytest

import ccf_amd
from ccf_amd import io
from helpers import make_field

pytestmark = pytest.mark.gpu


def _rows(sf, extra, seed):
    n, rect, corner, theta, st = io.read_stars(sf)
    rng = np.random.default_rng(seed)
    ext = np.array([abs(corner.real), abs(corner.imag)]) if rect else np.array([abs(corner)] * 2)
    roots = np.concatenate([st[:, 0] + 1 + 1j * st[:, 1], st[:, 0] - 1 + 1j * st[:, 1],
                            ext[0] * rng.uniform(-1, 1, extra) + 1j * ext[1] * rng.uniform(-1, 1, extra)])
    return roots, roots[rng.permutation(roots.size)[: roots.size // 3]]


@pytest.mark.parametrize("name,kw,extra", [
    ("cfg1", dict(kappa_tot=0.3, shear=0.1, direct=1), 0),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35), 200),
    ("rect200", dict(kappa_tot=0.7, shear=0.25, precision=1), 300),
    ("salpeter300", dict(kappa_tot=0.45, shear=0.35, root_fmm=1), 2400),
    ("smooth150", dict(kappa_tot=0.6, shear=0.2, approx=0, precision=1, root_fmm=1), 1500),
])
def test_noalloc_sweep_matches_the_allocating_sweep(workdir, name, kw, extra):
    sf = make_field(workdir, name)
    big, small = _rows(sf, extra, 3)
    phi = 2.2
    print(f"# noalloc sweep {name} {kw}: rows of {big.size} and {small.size} roots", flush=True)
    with ccf_amd.Context(starfile=sf, outfile_prefix=str(workdir) + "/", num_phi=10, **kw) as c:
        c.prepare()
        got = [c.aberth_sweep_noalloc(big, phi), c.aberth_sweep_noalloc(small, phi)]
        ref = [c.aberth_sweep(big, phi), c.aberth_sweep(small, phi)]
        again = c.aberth_sweep_noalloc(big, phi)
    for g, r, z in zip(got, ref, (big, small)):
        assert g.shape == z.shape and np.all(np.isfinite(g))
        assert np.array_equal(g, r), np.abs(g - r).max()
        assert np.any(g != z)
    assert np.array_equal(again, ref[0])


def test_noalloc_sweep_refuses_a_row_beyond_its_scratch(workdir):
    sf = make_field(workdir, "salpeter300")
    with ccf_amd.Context(starfile=sf, outfile_prefix=str(workdir) + "/", num_phi=10, kappa_tot=0.45, shear=0.35) as c:
        c.prepare()
        z = np.linspace(-5, 5, 200000) + 0.5j
        with pytest.raises(RuntimeError, match="bytes of scratch, more than the 33554432 it has"):
            c.aberth_sweep_noalloc(z, 1.0)
        # the context still works afterwards
        small = np.array([1.0 + 0.5j, -2.0 + 1.0j, 0.3 - 0.7j])
        assert np.array_equal(c.aberth_sweep_noalloc(small, 1.0), c.aberth_sweep(small, 1.0))
