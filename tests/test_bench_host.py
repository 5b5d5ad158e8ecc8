"""Host logic of bench.py (no GPU): what its JSON line takes from the timed
steps and what from the event-timed step run after them."""
import bench


def _info(**kw):
    base = dict(points_converged=1000, points_total=1000, aberth_kernel_ms=0.0, aberth_launches=0, flops_sums=0.0,
                witt_kernel_ms=0.0, witt_launches=0, flops_witt=0.0, tail_launches=0, tail_sweeps=0,
                tail_kernel_ms=0.0, flops_tail=0.0, t_tree=0.001, t_init_roots=0.002, t_ccs=0.003, t_caustics=0.0,
                updates_full=10, updates_early=20, updates_newton=30, sweeps=40, root_fmm_updates=0, root_fmm_rows=0,
                root_fmm_order=6, root_fmm_leaf_level=6, cluster_chunks=0, flops_update=0.0, expansion_order=24,
                tree_levels=7, num_roots=20000, max_error=1e-8, max_witt_residual=1e-9, points_outside_tree=0,
                roots_per_thread=1, root_fmm_leaf=6)
    base.update(kw)
    return base


def test_summary_rate_comes_from_the_timed_steps_only():
    # the event-timed step is slower (its events open gaps); it must not
    # enter the rate
    infos = [_info(), _info(t_ccs=0.004)]
    probe = _info(aberth_kernel_ms=8.0, aberth_launches=4, flops_sums=4e10, witt_kernel_ms=2.0, witt_launches=4,
                  flops_witt=2e10, tail_launches=2, tail_sweeps=6, tail_kernel_ms=1.0, flops_tail=1e8)
    r = bench.summarize([0.020, 0.022], infos, probe)
    assert abs(r["s_per_step"] - 0.021) < 1e-15
    assert abs(r["points_per_s"] - 1000 / 0.021) < 1e-6
    # stage times and counts from the last timed step
    assert r["t_ccs"] == 0.004
    assert r["updates_full"] == 10


def test_summary_launch_durations_come_from_the_event_step():
    infos = [_info(aberth_kernel_ms=0.0, witt_kernel_ms=0.0)]
    probe = _info(aberth_kernel_ms=8.0, aberth_launches=4, flops_sums=4e10, witt_kernel_ms=2.0, witt_launches=4,
                  flops_witt=2e10, tail_launches=2, tail_sweeps=6, tail_kernel_ms=1.0, flops_tail=1e8)
    r = bench.summarize([0.02], infos, probe)
    assert r["aberth_ms_total"] == 8.0 and r["aberth_ms_avg"] == 2.0
    # 4e10 flop in 8 ms: 5 TFLOP/s; 2e10 in 2 ms: 10 TFLOP/s; 1e8 in 1 ms: 0.1
    assert abs(r["sums_tflops"] - 5.0) < 1e-12
    assert abs(r["witt_tflops"] - 10.0) < 1e-12
    assert abs(r["tail_tflops"] - 0.1) < 1e-12
    assert r["tail_launches"] == 2
