"""The batched form (ccf_batch_*): several fields of uneven size solved in
shared persistent launches, checked field by field against the FP64 oracle
run on the same star file at the batch's expansion order, and against the
single-field path of the product."""
import os

import numpy as np
import pytest

import ccf_amd
from ccf_amd import io
from helpers import compare_arrays, newton_step, oracle_run

pytestmark = pytest.mark.gpu

# uneven fields in the spirit of the 256-field workload: equal masses,
# kappa = kappa_star, shears and sizes drawn per field (one rectangular)
FIELDS = [
    dict(num_stars=60, kappa_tot=0.3, kappa_star=0.3, shear=0.1, seed=100),
    dict(num_stars=180, kappa_tot=0.5, kappa_star=0.5, shear=0.35, seed=101),
    dict(num_stars=110, kappa_tot=0.2, kappa_star=0.2, shear=0.45, seed=102),
    dict(num_stars=240, kappa_tot=0.45, kappa_star=0.45, shear=0.05, seed=103),
    dict(num_stars=90, kappa_tot=0.6, kappa_star=0.3, shear=0.25, rectangular=1, mass_function="salpeter",
         m_lower=0.1, m_upper=1.0, seed=104),
]
RUN = dict(num_phi=24, num_branches=2)


def _star_files(workdir):
    out = []
    for i, spec in enumerate(FIELDS):
        spec = dict(spec)
        seed = spec.pop("seed")
        path = os.path.join(str(workdir), f"batch{i}.bin")
        if not os.path.exists(path):
            ccf_amd.write_synthetic_stars(path, seed, **spec)
        out.append((path, spec))
    return out


def _field_kw(spec):
    return dict(kappa_tot=spec["kappa_tot"], shear=spec["shear"])


def _run_batch(workdir, tag, **extra):
    files = _star_files(workdir)
    b = ccf_amd.Batch()
    for i, (sf, spec) in enumerate(files):
        b.add(starfile=sf, outfile_prefix=os.path.join(str(workdir), f"{tag}{i}_"), write_mu_length_scales=1,
              **_field_kw(spec), **RUN, **extra)
    assert b.run(0), b.last_error()
    return b, files


def test_batch_matches_oracle_per_field(workdir):
    b, files = _run_batch(workdir, "bg")
    tot = b.info(-1)
    assert tot["points_converged"] == tot["points_total"] > 0
    assert b.workgroup() in (256, 512, 1024)
    for i, (sf, spec) in enumerate(files):
        gi = b.info(i)
        order = gi["expansion_order"]
        op = os.path.join(str(workdir), f"bo{i}_")
        ost = oracle_run(sf, op, fmm_order=order, **_field_kw(spec), **RUN)
        R, nb = ost["num_roots"], RUN["num_branches"]
        assert gi["num_roots"] == R and gi["tree_levels"] == ost["tree_levels"]
        g = b.array(i, ccf_amd.ARRAY_CCS)
        o = io.read_array(op + "ccf_ccs.bin")
        assert g.shape == o.shape
        worst, nbad = compare_arrays(g, o, R, nb)
        assert nbad == 0, (i, worst, nbad)
        gc = b.array(i, ccf_amd.ARRAY_CAUSTICS)
        oc = io.read_array(op + "ccf_caustics.bin")
        worst_c, nbad_c = compare_arrays(gc, oc, R, nb)
        assert nbad_c == 0, (i, worst_c, nbad_c)
        assert gi["points_converged"] == ost["points_converged"] == R * (RUN["num_phi"] + nb)
        assert gi["max_error"] < 1e-8


def test_batch_files_are_the_reference_layout(workdir):
    b, files = _run_batch(workdir, "bf")
    for i, (sf, spec) in enumerate(files):
        assert b.save(i, 0), b.last_error()
        pre = os.path.join(str(workdir), f"bf{i}_")
        par = io.read_params(pre + "ccf_parameter_info.txt")
        R = par["num_roots"]
        ccs = io.read_array(pre + "ccf_ccs.bin")
        assert ccs.shape == (R * RUN["num_branches"], RUN["num_phi"] // RUN["num_branches"] + 1)
        mu = io.read_array(pre + "ccf_mu_length_scales.bin", complex_values=False)
        assert mu.shape == ccs.shape and np.all(mu > 0)
        _, _, _, _, st = io.read_stars(pre + "ccf_stars.bin")
        _, _, _, _, st0 = io.read_stars(sf)
        assert np.array_equal(np.sort(st.view("f8,f8,f8"), axis=0), np.sort(st0.view("f8,f8,f8"), axis=0))


def test_batch_matches_single_field_path(workdir):
    b, files = _run_batch(workdir, "bs")
    for i, (sf, spec) in enumerate(files):
        order = b.info(i)["expansion_order"]
        with ccf_amd.Context(starfile=sf, outfile_prefix=str(workdir) + "/", fmm_order=order, **_field_kw(spec), **RUN) as c:
            assert c.run(0), c.last_error()
            s = c.array(ccf_amd.ARRAY_CCS)
            R = c.info()["num_roots"]
        worst, nbad = compare_arrays(b.array(i, ccf_amd.ARRAY_CCS), s, R, RUN["num_branches"])
        assert nbad == 0, (i, worst, nbad)


@pytest.mark.parametrize("predictor", [2, 4])
def test_mixed_batch_with_predictor_within_band(workdir, predictor):
    """precision 1 and a continuation predictor change the path to each root:
    roots sit within the oracle's convergence band, caustics are the lens
    map of the product's own critical points"""
    b, files = _run_batch(workdir, f"bm{predictor}", precision=1, predictor=predictor)
    for i, (sf, spec) in enumerate(files):
        gi = b.info(i)
        op = os.path.join(str(workdir), f"bmo{predictor}_{i}_")
        ost = oracle_run(sf, op, fmm_order=gi["expansion_order"], **_field_kw(spec), **RUN)
        R, nb = ost["num_roots"], RUN["num_branches"]
        g = b.array(i, ccf_amd.ARRAY_CCS)
        o = io.read_array(op + "ccf_ccs.bin")
        worst, nbad, bad = compare_arrays(g, o, R, nb, details=True)
        allowed = 2 * newton_step(sf, bad, RUN["num_phi"], nb, fmm_order=gi["expansion_order"], **_field_kw(spec))
        over = [(x, a) for x, a in zip(bad, allowed) if x[3] > 1e-9 * max(abs(x[2]), 1) + a]
        assert not over, over[:5]
        assert nbad <= max(5, 1e-3 * R * (RUN["num_phi"] + nb)), nbad
        assert gi["points_converged"] == R * (RUN["num_phi"] + nb)


def test_batch_rebuild_with_more_fields_is_exact(workdir):
    """A batch solved, grown by more fields and solved again (a larger result
    arena, a new forest and graph, buffers from the pool) gives its first
    fields bit-identical results, and the new fields those of a batch of
    their own (one expansion order for all, so that it does not change)."""
    files = _star_files(workdir)
    kw = dict(RUN, fmm_order=20, precision=1, predictor=4)

    def add(b, group, tag):
        for i, (sf, spec) in group:
            b.add(starfile=sf, outfile_prefix=os.path.join(str(workdir), f"{tag}{i}_"), **_field_kw(spec), **kw)

    first = list(enumerate(files))[:3]
    rest = list(enumerate(files))[3:]
    b = ccf_amd.Batch()
    add(b, first, "rb")
    assert b.run(0), b.last_error()
    before = [b.array(k, ccf_amd.ARRAY_CCS) for k in range(len(first))]
    add(b, rest, "rb")
    assert b.run(0), b.last_error()
    for k in range(len(first)):
        assert np.array_equal(b.array(k, ccf_amd.ARRAY_CCS), before[k])
    alone = ccf_amd.Batch()
    add(alone, rest, "ra")
    assert alone.run(0), alone.last_error()
    for k in range(len(rest)):
        assert np.array_equal(b.array(len(first) + k, ccf_amd.ARRAY_CCS), alone.array(k, ccf_amd.ARRAY_CCS))
    b.close()
    alone.close()
