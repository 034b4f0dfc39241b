"""Benchmark of the per-step particle pipeline on one MI355X.

Workload: the dam break with mDBC walls (3.22 x 1.8 m tank, 1.2 x 0.6 m column,
dx = 1 mm, 740,478 particles, fluid positions jittered by +-0.01 dx from a fixed
seed), one symplectic step = mDBC wall density, EOS, predictor, cell list, EOS,
pair loop, corrector, removal, cell list, EOS, pair loop (the stage order of
cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:59-116).

Measured in one run:
  * mixed path (FP32 pair arithmetic, FP64 integration)  -> `value`
  * plain FP64 path                                        -> `fp64` block
  * the mixed path's error against the FP64 path after 1000 steps (rho, p, v, r)
  * the oracle (the C++ restatement of the reference OpenMP path) on the host
    cores, on a bounded sample of the same workload       -> `cpu_baseline`
  * the other workloads, each timed over a short run     -> `workloads`:
    still water DBC (5.7k particles, FP64), the ALE-SPH MUSCL dam break (dx 2 mm,
    minmod, mixed and FP64), the 64-tank batched ensemble (7.8M particles, one
    launch per stage, autotuned, mixed and FP64) and the 4M-particle inflow
    channel with buffer zones (mixed)

`value` = ordered pair interactions (|r_ij| < 2h, i active) per second over the
timed steps, inputs resident in HBM.  Pair-kernel time comes from HIP events
around every pair launch on the library's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tinysph_old_amd as T  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic bytes (per pair: neighbour record; per active particle: own record + outputs)
B_PAIR = {T.PREC_FP64: 50, T.PREC_MIXED: 26}
B_PART = {T.PREC_FP64: 74, T.PREC_MIXED: 50}  # mixed keeps FP64 outputs (a 16 B + dRho 8 B)
SEED = 20241202
KERNEL = "k_tile32<0>"  # the mixed WCSPH pair kernel (mDBC case: fluid-only interactions)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_dam_break_mdbc_740k.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (lines `traffic_bytes_per_launch <kernel> <bytes>`), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for ln in f:
                t = ln.split()
                if len(t) == 3 and t[0] == "traffic_bytes_per_launch" and t[1] == kernel:
                    return float(t[2])
    except OSError:
        pass
    return None


def copy_probe(mib: int = 512, reps: int = 3):
    """A plain device copy of known size (k_copy_probe in libtsph_copy_probe.so),
    run so that rocprofv3's FETCH_SIZE and WRITE_SIZE per dispatch can be scaled
    against it: each dispatch reads and writes exactly `mib` MiB."""
    import ctypes as C

    lib = C.CDLL(os.path.join(ROOT, "tinysph_old_amd", "libtsph_copy_probe.so"))
    lib.tsph_copy_probe.restype = C.c_int
    lib.tsph_copy_probe.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    ms = C.c_double()
    rc = lib.tsph_copy_probe(mib, reps, C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"tsph_copy_probe failed ({rc})")
    gbs = 2 * mib * 2**20 / (ms.value * 1e-3) / 1e9
    print(f"[bench] copy probe: {reps} x ({mib} MiB read + {mib} MiB written), "
          f"{ms.value:.4f} ms per copy = {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    return ms.value


def run_path(case, precision: int, steps: int, warmup: int, total_steps: int = 0):
    ps = T.system_for_case(case, precision=precision)
    ps.particle_to_cell()
    ps.density_to_pressure()
    ps.compute_forces(T.forces_variant(case.wall))
    # warmup (graph capture + replay)
    ps.step(case.wall, case.dt, warmup)
    ps.synchronize()
    # timed region A: graph replay, whole-step throughput
    ps.reset_counters()
    ps.synchronize()
    t0 = time.perf_counter()
    ps.step(case.wall, case.dt, steps)
    ps.synchronize()
    t_graph = time.perf_counter() - t0
    cnt = ps.counters()
    # timed region B: same steps launched eagerly with HIP events on every pair kernel
    ps.pair_timing(2 * steps + 8)
    ps.synchronize()
    t0 = time.perf_counter()
    ps.step(case.wall, case.dt, steps)
    ps.synchronize()
    t_eager = time.perf_counter() - t0
    pair_ms, launches = ps.pair_timing_read()
    ps.pair_timing(0)
    cnt_b = ps.counters()
    done = warmup + 2 * steps
    if total_steps > done:  # on to the step count the error is quoted at
        ps.step(case.wall, case.dt, total_steps - done)
        done = total_steps
    state = ps.download(("rho", "r", "v", "p"))
    ps.close()
    pairs_a = cnt["pairs"]
    pairs_b = cnt_b["pairs"] - cnt["pairs"]
    act_b = cnt_b["active"] - cnt["active"]
    return dict(t_graph=t_graph, t_eager=t_eager, pairs=pairs_a, cand=cnt["candidates"],
                passes=cnt["pair_passes"], pair_ms=pair_ms, launches=launches,
                pairs_b=pairs_b, active_b=act_b, state=state, counters=cnt_b, steps_done=done)


def timed(ps, fn, steps: int, warmup: int):
    """(seconds, counters) of `steps` calls-worth of fn after `warmup`."""
    fn(warmup)
    ps.synchronize()
    ps.reset_counters()
    t0 = time.perf_counter()
    fn(steps)
    ps.synchronize()
    return time.perf_counter() - t0, ps.counters()


def leg_line(t, cnt, steps, particles, **extra):
    out = {"pair_interactions_per_s": cnt["pairs"] / t, "steps_per_s": steps / t,
           "ms_per_step": t / steps * 1e3, "particles": particles, "steps": steps}
    out.update(extra)
    return out


def leg_still_water(steps):
    """Still water with DBC walls, 5,650 particles, plain FP64 (CPU-runnable size)."""
    case = T.build_case(T.CASE_STILL_WATER_DBC)
    with T.system_for_case(case, precision=T.PREC_FP64) as ps:
        ps.particle_to_cell()
        t, cnt = timed(ps, lambda n: ps.step(case.wall, case.dt, n), steps, 10)
    return {"fp64": leg_line(t, cnt, steps, case.np)}


def leg_ale(steps):
    """ALE-SPH dam break, BT walls, dx = 2 mm, minmod-limited MUSCL."""
    case = T.build_case(T.CASE_ALE_DAM_BREAK)
    out = {}
    for name, prec in (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64)):
        with T.system_for_case(case, precision=prec) as ps:
            ps.ale_enable()
            ps.step_ale(case.dt, 1, 1, T.MUSCL_MINMOD)
            state = {"s": 2}

            def run(n):
                ps.step_ale(case.dt, state["s"], n, T.MUSCL_MINMOD)
                state["s"] += n

            t, cnt = timed(ps, run, steps, 5)
            out[name] = leg_line(t, cnt, steps, case.np)
    out["mixed_over_fp64"] = (out["mixed"]["pair_interactions_per_s"] /
                              out["fp64"]["pair_interactions_per_s"])
    return out


def leg_ensemble(steps):
    """64 independent tanks (7.8M particles) advanced together: one launch per
    stage, a segmented cell list, graph-replayed steps, autotuned pair loop."""
    cases = T.build_ensemble(64, SEED)
    n = sum(c.np for c in cases)
    out = {"tanks": len(cases), "particles": n,
           "fluid": sum(c.n_fluid for c in cases)}
    for name, prec in (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64)):
        with T.batch_for_cases(cases, prec) as bs:
            bs.particle_to_cell()
            rows, waves, bal, _ = T.autotune(bs, T.FORCES_WCSPH)
            t, cnt = timed(bs, lambda k: bs.step(T.WALL_MDBC, k), steps, 3)
            out[name] = leg_line(t, cnt, steps, n, tile_rows=rows, waves_per_cu=waves,
                                 balanced=bal)
    out["mixed_over_fp64"] = (out["mixed"]["pair_interactions_per_s"] /
                              out["fp64"]["pair_interactions_per_s"])
    return out


def leg_inflow(steps):
    """8 x 0.5 m channel with inlet/outlet buffer zones, dx = 0.5 mm (4M particles)."""
    case = T.build_case(T.CASE_INFLOW_CHANNEL)
    with T.system_for_case(case, precision=T.PREC_MIXED) as ps:
        ps.particle_to_cell()
        t, cnt = timed(ps, lambda k: ps.step_inflow(case.wall, case.dt, k), steps, 3)
    return {"mixed": leg_line(t, cnt, steps, case.np, appended=cnt["appended"],
                             removed=cnt["removed"], refused=cnt["zone_rejected"])}


def cpu_baseline(case, budget_s: float = 20.0, threads: int = 0, max_steps: int = 50):
    """Oracle (C++ restatement of the reference path, OpenMP) on the host cores."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_init, oracle_step

    threads = threads or min(16, os.cpu_count() or 1)
    O = po.oracle_case(case.oracle_params(), case.arrays, omp_threads=threads)
    oracle_init(O)
    O.call("density_to_pressure")
    O.call("compute_forces", 0)
    # pair count of one pass is taken from the GPU counters of the same case
    t0 = time.perf_counter()
    n = 0
    while True:
        oracle_step(O, case.wall, case.dt)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= max_steps:
            break
    return n, time.perf_counter() - t0, threads


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.001)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 512 MiB device copy (calibration for HBM counters)")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="step count at which the mixed path's error vs FP64 is quoted")
    ap.add_argument("--workloads", default="all",
                    help="comma list of still_water,ale,ensemble,inflow; 'none' or 'all'")
    ap.add_argument("--workloads-only", action="store_true",
                    help="run only the --workloads legs (for profiling them alone)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): the pipeline runs on a single MI355X; "
              "sharding across GPUs is designed, not built", file=sys.stderr)
        return 2

    legs = {"still_water": (leg_still_water, 200), "ale": (leg_ale, 50),
            "ensemble": (leg_ensemble, 20), "inflow": (leg_inflow, 50)}
    want = legs if args.workloads == "all" else \
        ([] if args.workloads == "none" else args.workloads.split(","))

    def run_workloads():
        out = {}
        for name in want:
            fn, steps = legs[name]
            out[name] = fn(steps)
            print(f"[bench] workload {name} done", file=sys.stderr, flush=True)
        return out

    if args.workloads_only:
        print(json.dumps({"workloads": run_workloads()}), flush=True)
        return 0

    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=args.dx, jitter=0.01, seed=SEED)
    print(f"[bench] case dam_break_mdbc np={case.np} fluid={case.n_fluid} wall={case.n_wall} "
          f"dt={case.dt:.4g}", file=sys.stderr, flush=True)
    fp64 = run_path(case, T.PREC_FP64, args.steps, args.warmup, args.error_steps)
    print(f"[bench] fp64 done {fp64['t_graph']:.3f}s", file=sys.stderr, flush=True)
    mixed = run_path(case, T.PREC_MIXED, args.steps, args.warmup, args.error_steps)
    print(f"[bench] mixed done {mixed['t_graph']:.3f}s", file=sys.stderr, flush=True)

    def rate(r):
        return r["pairs"] / r["t_graph"]

    k = args.steps

    def rel_l2(f):
        a, b = mixed["state"][f], fp64["state"][f]
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    # error of the mixed path vs the FP64 path (the FP64 path matches the
    # oracle within 1e-9) after the same steps from the same state
    assert mixed["steps_done"] == fp64["steps_done"]
    err = {"rho_rel_l2": rel_l2("rho"), "p_rel_l2": rel_l2("p"), "v_rel_l2": rel_l2("v"),
           "r_rel_l2": rel_l2("r"), "after_steps": mixed["steps_done"]}

    def roof(r, prec):
        avg_ms = r["pair_ms"] / max(r["launches"], 1)
        pairs_per_launch = r["pairs_b"] / max(r["launches"], 1)
        act_per_launch = r["active_b"] / max(r["launches"], 1)
        bytes_per_launch = pairs_per_launch * B_PAIR[prec] + act_per_launch * B_PART[prec]
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        return avg_ms, achieved, bytes_per_launch, pairs_per_launch

    m_avg, m_ach, m_bytes, m_ppl = roof(mixed, T.PREC_MIXED)
    traffic = measured_traffic(KERNEL)
    d_avg, d_ach, d_bytes, d_ppl = roof(fp64, T.PREC_FP64)

    line = {
        "metric": "pair-interactions/s & steps/s, mixed vs FP64 GPU, at stated error",
        "value": rate(mixed),
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": mixed["t_graph"] / k * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairs / f64 integration",
        "data": "synthetic (dam-break lattice, seeded jitter)",
        "config": {"workload": "dam break mDBC 3.22x1.8 m tank, 1.2x0.6 m column, dx=1 mm",
                   "particles": case.np, "fluid": case.n_fluid, "wall": case.n_wall,
                   "dt": case.dt, "pair_passes_per_step": 2, "path": "mixed, hipGraph step"},
        "steps_per_s": k / mixed["t_graph"],
        "pairs_per_step": mixed["pairs"] / k,
        # cell-list scan blocks that fell back to summing the counts themselves
        # (both paths, timed runs; expected 0)
        "scan_stalls": mixed["counters"]["scan_stalls"] + fp64["counters"]["scan_stalls"],
        "fp64": {"pair_interactions_per_s": rate(fp64), "steps_per_s": k / fp64["t_graph"],
                 "ms_per_step": fp64["t_graph"] / k * 1e3, "pair_kernel_avg_ms": d_avg,
                 "roofline_frac": d_ach / HBM_PEAK_GBS},
        "mixed_over_fp64": rate(mixed) / rate(fp64),
        "error_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": m_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m_ach / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_gbs": traffic / (m_avg * 1e-3) / 1e9 if traffic else None,
                     "kernel": KERNEL, "avg_launch_ms": m_avg,
                     "algorithmic_bytes_per_launch": m_bytes, "pairs_per_launch": m_ppl},
    }
    if want:
        line["workloads"] = run_workloads()
    if not args.no_cpu:
        n, t, threads = cpu_baseline(case, args.cpu_budget)
        pairs_per_step = fp64["pairs"] / k
        line["cpu_baseline"] = {"value": n * pairs_per_step / t, "unit": "pair-interactions/s",
                                "cores": threads, "kind": "port",
                                "sample": f"{n} full symplectic steps of the same 740k-particle "
                                          f"dam break (oracle, OpenMP, {threads} threads), "
                                          f"{t:.1f} s",
                                "steps_per_s": n / t}
        n1, t1, _ = cpu_baseline(case, args.cpu_budget / 3, threads=1, max_steps=5)
        line["cpu_baseline"]["single_thread"] = {
            "value": n1 * pairs_per_step / t1, "unit": "pair-interactions/s", "cores": 1,
            "sample": f"{n1} full symplectic steps, 1 thread, {t1:.1f} s"}
    if args.copy_probe:
        copy_probe()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
