#!/usr/bin/env python3
"""Throughput of the critical-curve hot path on one MI355X.

Workload (default): 10,000 stars, Salpeter IMF (slope -2.35) on [0.08, 1],
uniform disc for kappa_star = kappa_tot = 0.5, shear 0.2, num_phi 2000,
num_branches 20, FMM order 24, star-file seed 2 (synthetic, written by
ccf_write_synthetic_stars). One step = tree + FMM build, the Aberth solve at
the 20 branch centres, the phi continuation, the 1/mu error pass, the
transpose and the caustic map, with the stars already in HBM. The metric is
converged critical-curve points per second.

Prints one JSON line. The Aberth kernel's time comes from HIP events
recorded on the stream it is launched on; its algorithmic flops from the
per-update counts (10 per Cauchy term, 36 per near-field star, 16 per FMM
order). The CPU baseline is the FP64 oracle with OpenMP on the process's
host cores, timed on whole branches (median of three runs) and scaled to the
whole job; that leg also takes the accuracy gate against the oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import ccf_amd  # noqa: E402
from ccf_amd import io  # noqa: E402

METRIC = "converged cc points/s at 10k stars, mixed vs FP64 on 1 MI355X; max root error"
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak
PEAK_FP32_TFLOPS = 157.3

WORKLOADS = {
    "cc10k": dict(stars=dict(num_stars=10000, kappa_tot=0.5, kappa_star=0.5, shear=0.2, mass_function="salpeter",
                             m_lower=0.08, m_upper=1.0), seed=2,
                  run=dict(kappa_tot=0.5, shear=0.2, num_phi=2000, num_branches=20, fmm_order=24)),
    "cc2k": dict(stars=dict(num_stars=2000, kappa_tot=0.45, kappa_star=0.45, shear=0.35, mass_function="salpeter",
                            m_lower=0.08, m_upper=1.0), seed=1,
                 run=dict(kappa_tot=0.45, shear=0.35, num_phi=1000, num_branches=10, fmm_order=20)),
    # 50,000 stars, Kroupa IMF (slopes -1.3 / -2.3, break 0.5) on [0.08, 1], rectangular field for
    # kappa_star 0.35 with kappa_smooth 0.35, shear 0.25, num_phi 2000, num_branches 100, FMM order 28
    "cc50k": dict(stars=dict(num_stars=50000, kappa_tot=0.7, kappa_star=0.35, shear=0.25, mass_function="kroupa",
                             m_lower=0.08, m_upper=1.0, rectangular=1), seed=3,
                  run=dict(kappa_tot=0.7, shear=0.25, num_phi=2000, num_branches=100, fmm_order=28)),
}


def run_gpu(starfile: str, run_kw: dict, precision: int, steps: int, warmup: int, predictor: int = 0,
            root_fmm: int = 0, fmm_kw: dict | None = None):
    times, infos = [], []
    ctx = ccf_amd.Context(starfile=starfile, outfile_prefix="./", write_stars=0, write_critical_curves=0,
                          write_caustics=1, precision=precision, timing=0, predictor=predictor, root_fmm=root_fmm,
                          **(fmm_kw or {}), **run_kw)
    tune_s = 0.0
    probe = None
    for i in range(warmup + steps + 1):
        # one step after the timed ones records HIP events around every
        # sweep's regions (the roofline's launch durations); the events cost
        # about 4 ms of that step, which is therefore not a timed step
        events = i == warmup + steps
        if events:
            ctx.set(timing=1)
        ctx.stage("load")
        t0 = time.perf_counter()
        for st in ("build_tree_fmm", "find_initial_roots", "find_ccs", "find_caustics"):
            ctx.stage(st)
        dt = time.perf_counter() - t0
        kind = "warmup" if i < warmup else ("event-timed, not timed" if events else "timed")
        print(f"# precision={precision} predictor={predictor} root_fmm={root_fmm} step {i} ({kind}): "
              f"{dt:.3f} s", file=sys.stderr, flush=True)
        inf = ctx.info()
        tune_s += inf["t_autotune"]
        if events:
            probe = inf
        elif i >= warmup:
            times.append(dt)
            infos.append(inf)
    ccs = ctx.array(ccf_amd.ARRAY_CCS)
    ctx.close()
    # (params.autotune: the first warmup step tunes; the timed steps reuse it)
    infos[-1]["t_autotune_first_step"] = tune_s
    return times, infos, ccs, probe


def set_error(a: np.ndarray, b: np.ndarray, R: int, nb: int, theta: float = 1.0):
    """max over phi columns of the matched |a - b| / theta_E (index match, then
    Hungarian assignment over the entries that disagree), and the matched
    pairs (branch, column, a_i, b_i) that differ by more than 1e-9 theta_E."""
    from scipy.optimize import linear_sum_assignment
    from scipy.spatial import cKDTree

    ca = io.columns_by_phi(a, R, nb)
    cb = io.columns_by_phi(b, R, nb)
    d = np.abs(ca - cb)
    worst = 0.0
    pairs = []
    bad = np.argwhere(d.max(axis=2) > 1e-9 * theta)
    for n, (c, k) in enumerate(bad):
        if n % 200 == 0:
            print(f"# matching root sets: column {n} of {len(bad)}", file=sys.stderr, flush=True)
        x, y = ca[c, k], cb[c, k]
        idx = np.nonzero(np.abs(x - y) > 1e-9 * theta)[0]
        if idx.size <= 400:
            cost = np.abs(x[idx][:, None] - y[idx][None, :])
            r, s = linear_sum_assignment(cost)
            dist = cost[r, s]
        else:
            # large differing sets: nearest neighbour in the other set (roots
            # are distinct points, far apart compared with the differences)
            yy = y[idx]
            dist, s = cKDTree(np.column_stack([yy.real, yy.imag])).query(np.column_stack([x[idx].real, x[idx].imag]))
            r = np.arange(idx.size)
        dd = np.abs(x - y)
        dd[idx[r]] = dist
        worst = max(worst, float(dd.max()))
        for i, j, dv in zip(r, s, dist):
            if dv > 1e-9 * theta:
                pairs.append((int(c), int(k), x[idx[i]], y[idx[j]]))
    mask = np.ones(d.shape[:2], bool)
    if bad.size:
        mask[bad[:, 0], bad[:, 1]] = False
    if mask.any():
        worst = max(worst, float(d[mask].max()))
    return worst / theta, pairs


def relative_error(pairs, theta: float = 1.0) -> float:
    """max |a - b| / max(|b|, theta_E) over the differing pairs (the parity
    bar's relative form; pairs within 1e-9 theta_E absolute are within it)"""
    return max((abs(x - y) / max(abs(y), theta) for _, _, x, y in pairs), default=0.0)


def gate_summary(pairs_x, pairs_y, exact):
    """distances of two forms' roots to the exact roots at the given points"""
    x, y = np.asarray(pairs_x), np.asarray(pairs_y)
    return dict(points=int(x.size), form_max_dist_to_exact_thetaE=float(np.abs(x - exact).max(initial=0.0)),
                fp64_max_dist_to_exact_thetaE=float(np.abs(y - exact).max(initial=0.0)))


def cpu_threads() -> dict:
    """the host threads the CPU baseline uses: the process's CPU share
    (OMP_NUM_THREADS, else the CPUs this process may run on), and what the
    host reports besides"""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return dict(threads=omp or affinity, omp_num_threads=omp or None, sched_affinity=affinity, os_cpu_count=os.cpu_count())


def cpu_baseline(starfile: str, run_kw: dict, branches: int, threads: int, repeats: int, gpu_ccs: dict, R: int,
                 job_pairs: dict):
    """The CPU leg: the oracle (FP64, OpenMP) on the first `branches`
    branches, every phi step (initial solve + the whole continuation of
    each), timed `repeats` times (median) and scaled to all branches. As the
    checker it also compares the GPU rows of those branches with the
    oracle's, and gives the exact roots (the oracle's roots polished by FP64
    Newton steps of the oracle's field) at which the accuracy gate is taken:
      - every GPU form's roots on the sampled branches against the oracle's;
      - the points of the whole job where the mixed and FP64 GPU forms differ
        by more than 1e-9 theta_E (job_pairs)."""
    import oracle
    from scipy.optimize import linear_sum_assignment

    nb, npp = run_kw["num_branches"], run_kw["num_phi"]
    branches = min(branches, nb)
    walls, stats = [], []
    with tempfile.TemporaryDirectory() as d:
        for i in range(repeats):
            p = oracle.params(starfile=starfile, outfile_prefix=d + "/", branch_limit=branches, compact_output=1,
                              num_threads=threads, write_stars=0, write_caustics=0, **run_kw)
            t0 = time.perf_counter()
            st = oracle.run(p)
            walls.append(time.perf_counter() - t0)
            stats.append(st)
            print(f"# cpu baseline run {i}: {walls[-1]:.1f} s", file=sys.stderr, flush=True)
        occs = io.read_array(d + "/ccf_ccs.bin")
    order = np.argsort([s_["t_init_roots"] + s_["t_ccs"] for s_ in stats])
    s = stats[order[len(order) // 2]]
    t_job = (s["t_init_roots"] + s["t_ccs"]) * nb / branches
    points = R * (npp + nb)
    res = dict(value=points / t_job, unit="points/s", cores=threads, kind="port",
               sample=f"FP64 oracle, branches 0..{branches - 1} of {nb}, every phi step ({s['points_solved']} points), "
                      f"median of {repeats} runs ({', '.join(f'{w:.1f}' for w in walls)} s wall); "
                      f"job time = {nb}/{branches} x (t_init + t_ccs)",
               sample_seconds=float(np.median(walls)), t_init=s["t_init_roots"], t_ccs=s["t_ccs"], repeats=repeats)
    okw = {k: v for k, v in run_kw.items() if k not in ("num_phi", "num_branches")}
    p = oracle.params(starfile=starfile, **okw)
    cols = npp // nb + 1

    def phi(c, k):
        return 2 * np.pi * (c * npp // nb + k) / npp

    gate = {}
    for name, arr in gpu_ccs.items():
        ca = io.columns_by_phi(arr, R, nb)[:branches]
        cb = io.columns_by_phi(occs, R, branches)
        worst, xs, ys, phs = 0.0, [], [], []
        for c in range(branches):
            for k in range(cols):
                dd = np.abs(ca[c, k] - cb[c, k])
                # absolute 1e-9 theta_E: the gate's unit (the parity bar's
                # relative form is the smaller test)
                idx = np.nonzero(dd > 1e-9)[0]
                if idx.size:
                    cost = np.abs(ca[c, k][idx][:, None] - cb[c, k][idx][None, :])
                    r, cc = linear_sum_assignment(cost)
                    dd[idx[r]] = cost[r, cc]
                    for i, j in zip(r, cc):
                        if cost[i, j] > 1e-9:
                            xs.append(ca[c, k][idx[i]])
                            ys.append(cb[c, k][idx[j]])
                            phs.append(phi(c, k))
                worst = max(worst, float((dd / np.maximum(np.abs(cb[c, k]), 1)).max()))
        g = dict(max_rel_diff_vs_oracle=worst, points_differing_over_1e9_thetaE=len(xs))
        if xs:
            exact = oracle.newton_polish(p, ys, phs)
            g.update(gate_summary(xs, ys, exact))
        gate[name] = g
    res["oracle_max_rel_diff"] = gate.get("fp64", {}).get("max_rel_diff_vs_oracle")
    res["sampled_branches_vs_oracle"] = gate
    if job_pairs:
        # whole job: where the mixed form differs from the FP64 form, both
        # against the exact root
        xs = [x for _, _, x, _ in job_pairs["pairs"]]
        ys = [y for _, _, _, y in job_pairs["pairs"]]
        phs = [phi(c, k) for c, k, _, _ in job_pairs["pairs"]]
        res["job_mixed_vs_fp64_differing_points"] = gate_summary(xs, ys, oracle.newton_polish(p, ys, phs)) if xs else \
            dict(points=0, form_max_dist_to_exact_thetaE=0.0, fp64_max_dist_to_exact_thetaE=0.0)
    return res


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE, scaled by a copy
    of known size); None when there is none for this workload."""
    path = os.path.join(ROOT, "profiles", f"{workload}_sums_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return json.loads(line.split(":", 1)[1]).get(kernel)
    except (OSError, ValueError):
        pass
    return None


def summarize(times, infos, probe):
    """times and infos: the timed steps; probe: the info of the event-timed
    step after them (launch durations of the roofline's regions)"""
    inf = infos[-1]
    t = float(np.mean(times))
    conv = inf["points_converged"]
    k_ms = probe["aberth_kernel_ms"]
    launches = max(1, probe["aberth_launches"])
    achieved = probe["flops_sums"] / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    w_ms = probe["witt_kernel_ms"]
    w_launches = max(1, probe["witt_launches"])
    return dict(points_per_s=conv / t, s_per_step=t, points_converged=conv, points_total=inf["points_total"],
                t_tree=inf["t_tree"], t_init_roots=inf["t_init_roots"], t_ccs=inf["t_ccs"], t_caustics=inf["t_caustics"],
                updates_full=inf["updates_full"], updates_early=inf["updates_early"], updates_newton=inf["updates_newton"],
                sweeps=inf["sweeps"],
                root_fmm_updates=inf["root_fmm_updates"], root_fmm_rows=inf["root_fmm_rows"],
                root_fmm_order=inf["root_fmm_order"], root_fmm_leaf_level=inf["root_fmm_leaf_level"],
                cluster_chunks=inf["cluster_chunks"],
                aberth_ms_total=k_ms, aberth_launches=launches, aberth_ms_avg=k_ms / launches,
                sums_tflops=achieved, sums_flops=probe["flops_sums"], update_flops=inf["flops_update"],
                witt_ms_total=w_ms, witt_launches=w_launches, witt_ms_avg=w_ms / w_launches,
                witt_tflops=probe["flops_witt"] / (w_ms * 1e-3) / 1e12 if w_ms > 0 else 0.0, witt_flops=probe["flops_witt"],
                tail_launches=probe["tail_launches"], tail_sweeps=probe["tail_sweeps"], tail_ms_total=probe["tail_kernel_ms"],
                tail_flops=probe["flops_tail"],
                tail_tflops=(probe["flops_tail"] / (probe["tail_kernel_ms"] * 1e-3) / 1e12
                             if probe["tail_kernel_ms"] > 0 else 0.0),
                expansion_order=inf["expansion_order"], tree_levels=inf["tree_levels"], num_roots=inf["num_roots"],
                max_error_1_over_mu=inf["max_error"], max_witt_residual=inf["max_witt_residual"],
                points_outside_tree=inf["points_outside_tree"],
                roots_per_thread=inf["roots_per_thread"], root_fmm_leaf=inf["root_fmm_leaf"],
                autotune_s=inf.get("t_autotune_first_step", 0.0))


# 256 independent fields: N uniform in [50, 1000], equal masses, kappa_star =
# kappa_tot uniform in [0.1, 0.6], shear uniform in [0, 0.5], num_phi 504
# (500 breaks the rule num_phi % (2 num_branches) == 0), num_branches 4,
# field k drawn and written with seed 100 + k
BATCH_FIELDS, BATCH_RUN = 256, dict(num_phi=504, num_branches=4)


def batch_fields(tmp: str):
    fields = []
    for k in range(BATCH_FIELDS):
        seed = 100 + k
        r = np.random.default_rng(seed)
        n = int(r.integers(50, 1001))
        ks = float(r.uniform(0.1, 0.6))
        g = float(r.uniform(0.0, 0.5))
        path = os.path.join(tmp, f"field{k}.bin")
        ccf_amd.write_synthetic_stars(path, seed, num_stars=n, kappa_tot=ks, kappa_star=ks, shear=g, mass_function="equal")
        fields.append((path, dict(kappa_tot=ks, shear=g), n))
    return fields


def run_batch(fields, precision: int, predictor: int, steps: int, warmup: int):
    b = ccf_amd.Batch()
    for path, kw, _ in fields:
        b.add(starfile=path, outfile_prefix="./", write_stars=0, write_critical_curves=0, write_caustics=1,
              precision=precision, predictor=predictor, **kw, **BATCH_RUN)
    times = []
    for i in range(warmup + steps):
        b.load()
        t0 = time.perf_counter()
        b.build()
        b.solve()
        dt = time.perf_counter() - t0
        print(f"# batch precision={precision} predictor={predictor} step {i} ({'warmup' if i < warmup else 'timed'}): "
              f"{dt:.3f} s (workgroup {b.workgroup()})", file=sys.stderr, flush=True)
        if i >= warmup:
            times.append(dt)
    inf = b.info(-1)
    orders = sorted({b.info(f)["expansion_order"] for f in range(len(fields))})
    ccs = [b.array(f, ccf_amd.ARRAY_CCS) for f in range(len(fields))]
    wg = b.workgroup()
    b.close()
    t = float(np.mean(times))
    return dict(points_per_s=inf["points_converged"] / t, s_per_step=t, points_converged=inf["points_converged"],
                points_total=inf["points_total"], sweeps=inf["sweeps"], updates_full=inf["updates_full"],
                updates_early=inf["updates_early"], updates_newton=inf["updates_newton"], max_error_1_over_mu=inf["max_error"],
                max_witt_residual=inf["max_witt_residual"], workgroup=wg,
                expansion_orders=orders, t_tree=inf["t_tree"], t_solve=inf["t_ccs"]), ccs


def run_batched_workload(W: dict, tmp: str, nfields: int, predictor: int, steps: int, warmup: int):
    """the batched form (one persistent launch per phase, HIP graph) over
    nfields fields of the workload's size and shape, seeds seed, seed + 1, ...
    (mixed precision, the continuation predictor): whole-batch points/s"""
    fields = []
    for k in range(nfields):
        path = os.path.join(tmp, f"batched_{k}.bin")
        ccf_amd.write_synthetic_stars(path, W["seed"] + k, **W["stars"])
        fields.append((path, W["run"], None))
    b = ccf_amd.Batch()
    for path, kw, _ in fields:
        b.add(starfile=path, outfile_prefix="./", write_stars=0, write_critical_curves=0, write_caustics=1,
              precision=1, predictor=predictor, **kw)
    times = []
    for i in range(warmup + steps):
        b.load()
        t0 = time.perf_counter()
        b.build()
        b.solve()
        dt = time.perf_counter() - t0
        print(f"# batched {nfields} fields step {i} ({'warmup' if i < warmup else 'timed'}): {dt:.3f} s "
              f"(workgroup {b.workgroup()})", file=sys.stderr, flush=True)
        if i >= warmup:
            times.append(dt)
    inf = b.info(-1)
    wg = b.workgroup()
    b.close()
    t = float(np.mean(times))
    return dict(fields=nfields, points_per_s=inf["points_converged"] / t, s_per_step=t,
                points_converged=inf["points_converged"], points_total=inf["points_total"], sweeps=inf["sweeps"],
                workgroup=wg, t_tree=inf["t_tree"], t_solve=inf["t_ccs"])


def run_sequential(fields, nfields: int):
    """the plain FP64 single-field path, one field after another (stars in
    HBM before each timed field, as for the batch)"""
    tot = conv = pts = 0.0
    for path, kw, _ in fields[:nfields]:
        with ccf_amd.Context(starfile=path, outfile_prefix="./", write_stars=0, write_critical_curves=0, write_caustics=1,
                             **kw, **BATCH_RUN) as c:
            c.stage("load")
            t0 = time.perf_counter()
            for st in ("build_tree_fmm", "find_initial_roots", "find_ccs", "find_caustics"):
                c.stage(st)
            tot += time.perf_counter() - t0
            inf = c.info()
            conv += inf["points_converged"]
            pts += inf["points_total"]
    return dict(points_per_s=conv / tot, seconds=tot, fields=nfields, points_converged=conv, points_total=pts)


def bench_batch(args):
    tmp = tempfile.mkdtemp(prefix="ccfbatch_")
    fields = batch_fields(tmp)
    fp64, ccs64 = run_batch(fields, 0, 0, args.steps, args.warmup)
    mixed, ccsm = run_batch(fields, 1, args.predictor, args.steps, args.warmup)
    # the mixed batch against the FP64 batch, field by field (the gate
    # against the oracle's exact roots is tests/test_gpu_configs.py's)
    worst = rel = 0.0
    ndiff = 0
    for (path, kw, n), a, b in zip(fields, ccsm, ccs64):
        err, pairs = set_error(a, b, 2 * n, BATCH_RUN["num_branches"])
        worst = max(worst, err)
        rel = max(rel, relative_error(pairs))
        ndiff += len(pairs)
    mixed["max_root_err_vs_fp64_thetaE"] = worst
    mixed["max_root_err_vs_fp64_rel"] = rel
    mixed["points_differing_over_1e-9"] = ndiff
    seq = run_sequential(fields, min(args.seq_fields, len(fields)))
    line = {
        "metric": "converged cc points/s over 256 uneven fields (batched form) on 1 MI355X",
        "value": mixed["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["s_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "batch256: 256 fields, N ~ U{50..1000} equal masses, kappa_star = kappa_tot ~ U(0.1, 0.6), "
                               "shear ~ U(0, 0.5), num_phi 504, num_branches 4, seeds 100..355",
                   "points": mixed["points_total"]},
        "batched_mixed": mixed,
        "batched_fp64": fp64,
        "single_field_fp64_sequential": seq,
        "speedup_batched_fp64_vs_sequential": fp64["points_per_s"] / seq["points_per_s"],
        "speedup_batched_mixed_vs_sequential": mixed["points_per_s"] / seq["points_per_s"],
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cc10k", choices=sorted(WORKLOADS) + ["batch256"])
    ap.add_argument("--seq-fields", type=int, default=64,
                    help="batch256: fields of the sample timed one by one on the single-field FP64 path")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--cpu-branches", type=int, default=2, help="branches in the CPU sample, all phi steps (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=None, help="0: skip the CPU sample (same as --cpu-branches 0)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the process's CPU share (OMP_NUM_THREADS or affinity)")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="oracle runs of the CPU sample (the median is reported)")
    ap.add_argument("--ablation-steps", type=int, default=2,
                    help="timed steps of the default ablation (FP64 + predictor, mixed with direct root sums; 0: none)")
    ap.add_argument("--predictor", type=int, default=4, help="continuation predictor of the mixed form (0..4)")
    ap.add_argument("--root-fmm", type=int, default=1, help="root sums of dense sweeps by the FMM in the mixed form")
    ap.add_argument("--root-fmm-order", type=int, default=0, help="its expansion order (0: the library default)")
    ap.add_argument("--root-fmm-leaf", type=int, default=0, help="its mean roots per leaf cell (0: the library default)")
    ap.add_argument("--roots-per-thread", type=int, default=0, help="the sums' roots per thread (0: automatic)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: the first warmup step times the automatic choices (sums tile, root-sum FMM order and "
                         "cells) and the timed steps keep the fastest")
    ap.add_argument("--ablation", action="store_true", help="also time the mixed form without the predictor")
    ap.add_argument("--batched-fields", type=int, default=8,
                    help="the batched mixed form over this many fields of the workload's size (0: skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device: the known-size kernel that scales the HBM byte counters")
    args = ap.parse_args()
    if args.cpu_steps == 0:
        args.cpu_branches = 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1 (the path would shard by branch across GPUs, "
              "which is not part of this build)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy > 0:
        import torch

        a = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        print(f"# copied {args.calibrate_copy} MiB device to device", file=sys.stderr, flush=True)

    if args.workload == "batch256":
        bench_batch(args)
        return

    W = WORKLOADS[args.workload]
    tmp = tempfile.mkdtemp(prefix="ccfbench_")
    sf = os.path.join(tmp, "stars.bin")
    ccf_amd.write_synthetic_stars(sf, W["seed"], **W["stars"])
    run_kw = W["run"]

    res = {}
    ccs = {}
    # plain FP64 path = the reference's method; mixed form = FMM root sums of
    # the dense sweeps (FP64 near field), FP32 direct sums otherwise, FMM
    # pole sum, continuation predictor. --ablation takes the parts away.
    mixed = ("mixed", 1, args.predictor, args.root_fmm)
    modes = {"fp64": [("fp64", 0, 0, 0)], "mixed": [mixed],
             "both": [("fp64", 0, 0, 0), mixed]}[args.precision]
    # the ablation: what the FP64 form gains from the predictor alone, and
    # what the mixed form loses without the root-sum FMM (short runs)
    ablation = []
    if args.ablation_steps > 0 and args.precision == "both":
        ablation = [("fp64_predictor", 0, args.predictor, 0), ("mixed_direct_root_sums", 1, args.predictor, 0)]
        if args.ablation:
            ablation.append(("mixed_no_predictor", 1, 0, args.root_fmm))
    for name, prec, pred, rfmm in modes + ablation:
        fmm_kw = dict(root_fmm_order=args.root_fmm_order, root_fmm_leaf=args.root_fmm_leaf, autotune=args.autotune,
                      roots_per_thread=args.roots_per_thread)
        steps, warm = (args.steps, args.warmup) if (name, prec, pred, rfmm) in modes else (args.ablation_steps, 1)
        times, infos, ccs[name], probe = run_gpu(sf, run_kw, prec, steps, warm, pred, rfmm, fmm_kw)
        res[name] = summarize(times, infos, probe)
        res[name]["predictor"] = pred
        res[name]["root_fmm"] = rfmm
    R = res[next(iter(res))]["num_roots"]
    job_pairs = None
    if "fp64" in ccs:
        for name in res:
            if name == "fp64":
                continue
            err, pairs = set_error(ccs[name], ccs["fp64"], R, run_kw["num_branches"])
            res[name]["max_root_err_vs_fp64_thetaE"] = err
            res[name]["max_root_err_vs_fp64_rel"] = relative_error(pairs)
            res[name]["points_differing_over_1e-9"] = len(pairs)
            if name == "mixed":
                job_pairs = dict(pairs=pairs)
            res[name]["speedup_vs_fp64"] = res[name]["points_per_s"] / res["fp64"]["points_per_s"]

    batched = None
    if args.batched_fields > 0 and args.precision != "fp64":
        batched = run_batched_workload(W, tmp, args.batched_fields, args.predictor, 2, 1)
        if "fp64" in res:
            batched["vs_fp64_single_field"] = batched["points_per_s"] / res["fp64"]["points_per_s"]

    head = res["mixed"] if "mixed" in res else res["fp64"]
    mixed_head = "mixed" in res

    def region(r, which):
        # the two event-timed regions of a sweep: the root sums (FMM pass and
        # direct sums) and the Witt evaluation (k_witt, which also plans the sums)
        if which == "sums":
            name = ("root sums: FMM pass (k_r*) + k_sums32" if r.get("root_fmm") else
                    ("k_sums32" if r is res.get("mixed") else "k_sums64"))
            return dict(kernel=name, achieved=r["sums_tflops"], total_ms=r["aberth_ms_total"],
                        avg_launch_ms=r["aberth_ms_avg"], launches=r["aberth_launches"])
        if which == "tail":
            n = max(1, r.get("tail_launches", 0))
            return dict(kernel="k_tail (phase tails: Witt evaluations and Cauchy sums)", achieved=r.get("tail_tflops", 0.0),
                        total_ms=r.get("tail_ms_total", 0.0), avg_launch_ms=r.get("tail_ms_total", 0.0) / n,
                        launches=r.get("tail_launches", 0))
        return dict(kernel="k_witt", achieved=r["witt_tflops"], total_ms=r["witt_ms_total"],
                    avg_launch_ms=r["witt_ms_avg"], launches=r["witt_launches"])

    regs = [region(head, "sums"), region(head, "witt")]
    if head.get("tail_launches"):
        regs.append(region(head, "tail"))
    dom = max(regs, key=lambda g: g["total_ms"])
    pk = PEAK_FP64_TFLOPS
    roof = {"bound": "valu",
            "pipe": "FP64 VALU (v_fma_f64, v_rcp_f64): dense reciprocal sums and tree gathers with no contraction, so "
                    "no MFMA work; MI355X FP64 vector peak = FP64 matrix peak; the mixed form's FP32 parts "
                    "(sparse-sweep root sums, far pole field) issue at the same non-packed rate",
            "kernel": dom["kernel"], "achieved": dom["achieved"], "peak": pk, "unit": "TFLOP/s",
            "frac": dom["achieved"] / pk,
            "traffic": measured_traffic(args.workload, dom["kernel"]),
            "traffic_source": f"profiles/{args.workload}_sums_hbm_traffic.md",
            "avg_launch_ms": dom["avg_launch_ms"], "launches": dom["launches"],
            "timed_with": "HIP events on the context's stream around each sweep's regions, during one step run after the timed steps",
            "regions": [dict(g, frac=g["achieved"] / pk) for g in regs]}
    if "fp64" in res:
        roof["fp64_kernel"] = {"kernel": "k_sums64", "achieved": res["fp64"]["sums_tflops"], "peak": PEAK_FP64_TFLOPS,
                               "frac": res["fp64"]["sums_tflops"] / PEAK_FP64_TFLOPS,
                               "avg_launch_ms": res["fp64"]["aberth_ms_avg"],
                               "traffic": measured_traffic(args.workload, "k_sums64")}

    cpu = None
    gate = None
    if args.cpu_branches > 0:
        host = cpu_threads()
        threads = args.cpu_threads or host["threads"]
        cpu = cpu_baseline(sf, run_kw, args.cpu_branches, threads, args.cpu_repeats,
                           {k: v for k, v in ccs.items() if k in ("fp64", "mixed")}, R, job_pairs)
        cpu["host"] = host
        m = cpu["sampled_branches_vs_oracle"].get("mixed")
        if m is not None:
            # north-star accuracy gate: max root error <= 1e-9 theta_E against
            # the FP64 oracle's exact roots, on the sampled branches and at
            # every point of the job where the mixed and FP64 forms differ
            jm = cpu.get("job_mixed_vs_fp64_differing_points") or {}
            err = max(m.get("form_max_dist_to_exact_thetaE", 0.0), jm.get("form_max_dist_to_exact_thetaE", 0.0))
            gate = {"gate_max_root_err_le_1e-9_thetaE": bool(err <= 1e-9), "max_root_err_vs_exact_thetaE": err,
                    "sampled_branches": args.cpu_branches,
                    "mixed_points_differing_from_oracle_over_1e-9": m["points_differing_over_1e9_thetaE"],
                    "mixed_max_rel_diff_vs_oracle": m["max_rel_diff_vs_oracle"],
                    "oracle_max_dist_to_exact_thetaE": m.get("fp64_max_dist_to_exact_thetaE", 0.0)}

    line = {
        "metric": METRIC,
        "value": head["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["s_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if "mixed" not in res else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {W['stars']['num_stars']} stars {W['stars']['mass_function']} "
                               f"[{W['stars']['m_lower']},{W['stars']['m_upper']}], kappa={run_kw['kappa_tot']}, "
                               f"shear={run_kw['shear']}, num_phi={run_kw['num_phi']}, "
                               f"num_branches={run_kw['num_branches']}, fmm_order={run_kw['fmm_order']}, seed={W['seed']}",
                   "num_roots": R, "form": "mixed" if "mixed" in res else "fp64"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "accuracy_gate": gate,
        "fp64": res.get("fp64"),
        "mixed": res.get("mixed"),
        "ablation": {k: v for k, v in res.items() if k not in ("fp64", "mixed")} or None,
        "batched": batched,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
