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
order). The CPU baseline is the FP64 oracle with OpenMP on the host cores,
timed on one branch over a few phi steps and scaled to the whole job.
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
    for i in range(warmup + steps):
        if i == warmup + steps - 1:
            # the last timed step records HIP events around every sweep's
            # regions (the roofline's launch durations); the events cost
            # about 4 ms of this one step
            ctx.set(timing=1)
        ctx.stage("load")
        t0 = time.perf_counter()
        for st in ("build_tree_fmm", "find_initial_roots", "find_ccs", "find_caustics"):
            ctx.stage(st)
        dt = time.perf_counter() - t0
        print(f"# precision={precision} predictor={predictor} root_fmm={root_fmm} step {i} ({'warmup' if i < warmup else 'timed'}): "
              f"{dt:.3f} s", file=sys.stderr, flush=True)
        if i >= warmup:
            times.append(dt)
            infos.append(ctx.info())
    ccs = ctx.array(ccf_amd.ARRAY_CCS)
    ctx.close()
    return times, infos, ccs


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


def polished_error(starfile: str, run_kw: dict, pairs):
    """Over the differing pairs (a, b), with p = two FP64 Newton steps
    z - f0/f1 on the GPU (the exact root of Witt's equation to rounding):
    max |p(a) - p(b)| (the two forms converge to the same root), max |a - p(a)|
    and max |b - p(b)| (how far each form stopped from that root; the
    reference's 1/mu early exit bounds |f0|, not the distance)."""
    if not pairs:
        return 0.0, 0.0, 0.0
    npp, nb = run_kw["num_phi"], run_kw["num_branches"]
    worst = da = db = 0.0
    with ccf_amd.Context(starfile=starfile, outfile_prefix="./", **run_kw) as ctx:
        ctx.prepare()
        by_col = {}
        for c, k, x, y in pairs:
            by_col.setdefault((c, k), []).append((x, y))
        for (c, k), zz in by_col.items():
            phi = 2 * np.pi * (c * npp // nb + k) / npp
            z0 = np.array([p for pr in zz for p in pr], dtype=np.complex128)
            z = z0
            for _ in range(2):
                v = ctx.eval_points(z, phi)
                z = z - (v[:, 0] + 1j * v[:, 1]) / (v[:, 2] + 1j * v[:, 3])
            worst = max(worst, float(np.abs(z[0::2] - z[1::2]).max()))
            da = max(da, float(np.abs(z0[0::2] - z[0::2]).max()))
            db = max(db, float(np.abs(z0[1::2] - z[1::2]).max()))
    return worst, da, db


def cpu_baseline(starfile: str, run_kw: dict, branches: int, threads: int, gpu_ccs: np.ndarray | None, R: int):
    """Oracle (FP64, OpenMP) on the first `branches` branches, every phi step
    (initial solve + the whole continuation of each), scaled to all branches;
    also checks those rows against the GPU FP64 result."""
    import oracle

    nb, npp = run_kw["num_branches"], run_kw["num_phi"]
    branches = min(branches, nb)
    with tempfile.TemporaryDirectory() as d:
        p = oracle.params(starfile=starfile, outfile_prefix=d + "/", branch_limit=branches,
                          num_threads=threads, write_stars=0, write_caustics=0, **run_kw)
        t0 = time.perf_counter()
        s = oracle.run(p)
        wall = time.perf_counter() - t0
        occs = io.read_array(d + "/ccf_ccs.bin")
    t_job = (s["t_init_roots"] + s["t_ccs"]) * nb / branches
    points = R * (npp + nb)
    res = dict(value=points / t_job, unit="points/s", cores=threads, kind="port",
               sample=f"FP64 oracle, branches 0..{branches - 1} of {nb}, every phi step ({s['points_solved']} points, "
                      f"{wall:.1f} s wall); job time = {nb}/{branches} x (t_init + t_ccs)",
               sample_seconds=wall, t_init=s["t_init_roots"], t_ccs=s["t_ccs"])
    if gpu_ccs is not None:
        ca = io.columns_by_phi(gpu_ccs, R, nb)[:branches]
        cb = io.columns_by_phi(occs, R, nb)[:branches]
        worst = 0.0
        from scipy.optimize import linear_sum_assignment
        for c in range(ca.shape[0]):
            for k in range(ca.shape[1]):
                dd = np.abs(ca[c, k] - cb[c, k])
                idx = np.nonzero(dd > 1e-9 * np.maximum(np.abs(cb[c, k]), 1))[0]
                if idx.size:
                    cost = np.abs(ca[c, k][idx][:, None] - cb[c, k][idx][None, :])
                    r, cc = linear_sum_assignment(cost)
                    dd[idx[r]] = cost[r, cc]
                worst = max(worst, float((dd / np.maximum(np.abs(cb[c, k]), 1)).max()))
        res["oracle_max_rel_diff"] = worst
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


def summarize(times, infos):
    inf = infos[-1]
    t = float(np.mean(times))
    conv = inf["points_converged"]
    k_ms = inf["aberth_kernel_ms"]
    launches = max(1, inf["aberth_launches"])
    achieved = inf["flops_sums"] / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    w_ms = inf["witt_kernel_ms"]
    w_launches = max(1, inf["witt_launches"])
    return dict(points_per_s=conv / t, s_per_step=t, points_converged=conv, points_total=inf["points_total"],
                t_tree=inf["t_tree"], t_init_roots=inf["t_init_roots"], t_ccs=inf["t_ccs"], t_caustics=inf["t_caustics"],
                updates_full=inf["updates_full"], updates_early=inf["updates_early"], updates_newton=inf["updates_newton"],
                sweeps=inf["sweeps"],
                root_fmm_updates=inf["root_fmm_updates"], root_fmm_rows=inf["root_fmm_rows"],
                root_fmm_order=inf["root_fmm_order"], root_fmm_leaf_level=inf["root_fmm_leaf_level"],
                cluster_chunks=inf["cluster_chunks"],
                aberth_ms_total=k_ms, aberth_launches=launches, aberth_ms_avg=k_ms / launches,
                sums_tflops=achieved, sums_flops=inf["flops_sums"], update_flops=inf["flops_update"],
                witt_ms_total=w_ms, witt_launches=w_launches, witt_ms_avg=w_ms / w_launches,
                witt_tflops=inf["flops_witt"] / (w_ms * 1e-3) / 1e12 if w_ms > 0 else 0.0, witt_flops=inf["flops_witt"],
                tail_launches=inf["tail_launches"], tail_sweeps=inf["tail_sweeps"], tail_ms_total=inf["tail_kernel_ms"],
                tail_flops=inf["flops_tail"],
                tail_tflops=inf["flops_tail"] / (inf["tail_kernel_ms"] * 1e-3) / 1e12 if inf["tail_kernel_ms"] > 0 else 0.0,
                expansion_order=inf["expansion_order"], tree_levels=inf["tree_levels"], num_roots=inf["num_roots"],
                max_error_1_over_mu=inf["max_error"], max_witt_residual=inf["max_witt_residual"])


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
    # the mixed batch against the FP64 batch, field by field; where they
    # differ, each form's distance to the exact root (two FP64 Newton steps
    # on the field at the batch's expansion order)
    worst = rel = pol = da = db = 0.0
    ndiff = 0
    order = max(mixed["expansion_orders"])
    for (path, kw, n), a, b in zip(fields, ccsm, ccs64):
        err, pairs = set_error(a, b, 2 * n, BATCH_RUN["num_branches"])
        worst = max(worst, err)
        rel = max(rel, relative_error(pairs))
        ndiff += len(pairs)
        if pairs:
            p, x, y = polished_error(path, dict(kw, fmm_order=order, **BATCH_RUN), pairs[:200])
            pol, da, db = max(pol, p), max(da, x), max(db, y)
    mixed["max_root_err_vs_fp64_thetaE"] = worst
    mixed["max_root_err_vs_fp64_rel"] = rel
    mixed["points_differing_over_1e-9"] = ndiff
    mixed["max_root_err_vs_fp64_after_newton_polish_thetaE"] = pol
    mixed["differing_points_max_dist_to_exact_root_thetaE"] = da
    mixed["differing_points_fp64_max_dist_to_exact_root_thetaE"] = db
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
    ap.add_argument("--cpu-branches", type=int, default=6, help="branches in the CPU sample, all phi steps (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=None, help="0: skip the CPU sample (same as --cpu-branches 0)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--predictor", type=int, default=4, help="continuation predictor of the mixed form (0..4)")
    ap.add_argument("--root-fmm", type=int, default=1, help="root sums of dense sweeps by the FMM in the mixed form")
    ap.add_argument("--root-fmm-order", type=int, default=0, help="its expansion order (0: the library default)")
    ap.add_argument("--root-fmm-leaf", type=int, default=0, help="its mean roots per leaf cell (0: the library default)")
    ap.add_argument("--ablation", action="store_true",
                    help="also time FP64 + predictor, mixed without predictor, mixed with direct root sums")
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
    if args.ablation:
        modes += [("fp64_predictor", 0, args.predictor, 0), ("mixed_no_predictor", 1, 0, args.root_fmm),
                  ("mixed_direct_root_sums", 1, args.predictor, 0)]
    for name, prec, pred, rfmm in modes:
        fmm_kw = dict(root_fmm_order=args.root_fmm_order, root_fmm_leaf=args.root_fmm_leaf)
        times, infos, ccs[name] = run_gpu(sf, run_kw, prec, args.steps, args.warmup, pred, rfmm, fmm_kw)
        res[name] = summarize(times, infos)
        res[name]["predictor"] = pred
        res[name]["root_fmm"] = rfmm
    R = res[next(iter(res))]["num_roots"]
    if "fp64" in ccs:
        for name in res:
            if name == "fp64":
                continue
            err, pairs = set_error(ccs[name], ccs["fp64"], R, run_kw["num_branches"])
            res[name]["max_root_err_vs_fp64_thetaE"] = err
            res[name]["max_root_err_vs_fp64_rel"] = relative_error(pairs)
            res[name]["points_differing_over_1e-9"] = len(pairs)
            pol, da, db = polished_error(sf, run_kw, pairs[:2000])
            res[name]["max_root_err_vs_fp64_after_newton_polish_thetaE"] = pol
            # over the same pairs: each form's distance to the exact root
            res[name]["differing_points_max_dist_to_exact_root_thetaE"] = da
            res[name]["differing_points_fp64_max_dist_to_exact_root_thetaE"] = db
            res[name]["speedup_vs_fp64"] = res[name]["points_per_s"] / res["fp64"]["points_per_s"]

    head = res["mixed"] if "mixed" in res else res["fp64"]
    mixed_head = "mixed" in res

    def region(r, which):
        # the two event-timed regions of a sweep: the root sums (FMM pass and
        # direct sums) and the Witt evaluation (k_witt + k_witt_far)
        if which == "sums":
            name = ("root sums: FMM pass (k_r*) + k_sums32 + k_sums32_sparse" if r.get("root_fmm") else
                    ("k_sums32 + k_sums32_sparse" if r is res.get("mixed") else "k_sums64"))
            return dict(kernel=name, achieved=r["sums_tflops"], total_ms=r["aberth_ms_total"],
                        avg_launch_ms=r["aberth_ms_avg"], launches=r["aberth_launches"])
        if which == "tail":
            n = max(1, r.get("tail_launches", 0))
            return dict(kernel="k_tail (phase tails: Witt evaluations and Cauchy sums)", achieved=r.get("tail_tflops", 0.0),
                        total_ms=r.get("tail_ms_total", 0.0), avg_launch_ms=r.get("tail_ms_total", 0.0) / n,
                        launches=r.get("tail_launches", 0))
        return dict(kernel="k_witt + k_witt_far", achieved=r["witt_tflops"], total_ms=r["witt_ms_total"],
                    avg_launch_ms=r["witt_ms_avg"], launches=r["witt_launches"])

    regs = [region(head, "sums"), region(head, "witt")]
    if head.get("tail_launches"):
        regs.append(region(head, "tail"))
    dom = max(regs, key=lambda g: g["total_ms"])
    pk = PEAK_FP64_TFLOPS
    roof = {"bound": "mfma",
            "pipe": "FP64 VALU (v_fma_f64, v_rcp_f64): MI355X FP64 vector peak = FP64 matrix peak; the mixed form's "
                    "FP32 parts (sparse-sweep root sums, far pole field) issue at the same non-packed rate",
            "kernel": dom["kernel"], "achieved": dom["achieved"], "peak": pk, "unit": "TFLOP/s",
            "frac": dom["achieved"] / pk,
            "traffic": measured_traffic(args.workload, dom["kernel"]),
            "traffic_source": f"profiles/{args.workload}_sums_hbm_traffic.md",
            "avg_launch_ms": dom["avg_launch_ms"], "launches": dom["launches"],
            "timed_with": "HIP events on the context's stream around each sweep's regions, during the last timed step",
            "regions": [dict(g, frac=g["achieved"] / pk) for g in regs]}
    if "fp64" in res:
        roof["fp64_kernel"] = {"kernel": "k_sums64", "achieved": res["fp64"]["sums_tflops"], "peak": PEAK_FP64_TFLOPS,
                               "frac": res["fp64"]["sums_tflops"] / PEAK_FP64_TFLOPS,
                               "avg_launch_ms": res["fp64"]["aberth_ms_avg"],
                               "traffic": measured_traffic(args.workload, "k_sums64")}

    cpu = None
    if args.cpu_branches > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(sf, run_kw, args.cpu_branches, threads, ccs.get("fp64"), R)

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
        "fp64": res.get("fp64"),
        "mixed": res.get("mixed"),
        "ablation": {k: v for k, v in res.items() if k not in ("fp64", "mixed")} or None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
