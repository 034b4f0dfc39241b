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
# the symplectic corrector fused into the first mixed pair pass of a step, per
# particle: reads v_o, r_o (16 B each), rho_o, rho (8 B each), writes v, r
# (16 B each), rho (8 B) and the removal state (1 B)
B_CORR = 89
SEED = 20241202
KERNEL = "k_tile32<0>"  # the mixed WCSPH pair kernel (mDBC case: fluid-only interactions)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_dam_break_mdbc_740k.md")


SQ_PROFILE = os.path.join(ROOT, "profiles", "sq_counters_dam_break_mdbc_740k.md")
ALE_SQ_PROFILE = os.path.join(ROOT, "profiles", "sq_counters_ale_muscl_183k.md")
# VALU issue: one wave64 instruction per 4 cycles per SIMD, 256 CUs x 4 SIMDs, 2.4 GHz
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4


def profile_value(path: str, key: str, kernel: str):
    """The value of a machine-readable line `<key> <kernel> <value>` of a
    committed rocprofv3 summary, or None."""
    try:
        with open(path) as f:
            for ln in f:
                t = ln.split()
                if len(t) == 3 and t[0] == key and t[1] == kernel:
                    return float(t[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (lines `traffic_bytes_per_launch <kernel> <bytes>`), or None."""
    return profile_value(TRAFFIC_PROFILE, "traffic_bytes_per_launch", kernel)


def valu_ceiling(kernel: str, avg_ms: float, path: str = SQ_PROFILE) -> dict:
    """The VALU-issue ceiling of `kernel` from a committed rocprofv3 summary:
    SQ_INSTS_VALU per launch at one wave instruction per 4 cycles on every
    SIMD (valu_floor_ms, valu_frac), and SQ_ACTIVE_INST_VALU, the quad-cycles
    those instructions occupy the SIMDs (transcendentals take four), at the
    same rate (issue_floor_ms, issue_frac); the fractions are the floor over
    the measured launch time."""
    insts = profile_value(path, "valu_insts_per_launch", kernel)
    quads = profile_value(path, "valu_quad_cycles_per_launch", kernel)
    out = {"valu_insts_per_launch": insts, "valu_floor_ms": None, "valu_frac": None,
           "issue_floor_ms": None, "issue_frac": None}
    if insts is not None:
        out["valu_floor_ms"] = insts / VALU_ISSUE_PER_S * 1e3
        out["valu_frac"] = out["valu_floor_ms"] / avg_ms if avg_ms > 0 else None
    if quads is not None:
        out["issue_floor_ms"] = quads / VALU_ISSUE_PER_S * 1e3
        out["issue_frac"] = out["issue_floor_ms"] / avg_ms if avg_ms > 0 else None
    return out


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


# ALE-SPH MUSCL pair passes with reconstruction (minmod): algorithmic bytes per
# pair (the staged neighbour slot: r, v, {rho, omega}, n, type, and for the
# density and acceleration passes the three gradients) and per active particle
# (own slot + outputs: d omega, d omega rho; the three gradients; d omega v)
ALE_PASS_BYTES = {
    T.PREC_MIXED: {"density": (57, 57 + 16), "gradient": (33, 33 + 48), "acceleration": (57, 57 + 16)},
    T.PREC_FP64: {"density": (113, 113 + 16), "gradient": (65, 65 + 48),
                  "acceleration": (113, 113 + 16)},
}


def ale_pass_roofline(ps, prec, reps: int = 10) -> dict:
    """Each ALE pair pass timed alone on the current state (HIP events around
    each launch), with its pair and active-particle counts: achieved
    algorithmic GB/s and the fraction of the HBM peak."""
    calls = {"density": lambda: ps.ale_compute_density(limited=True),
             "gradient": ps.ale_kernel_gradient,
             "acceleration": lambda: ps.ale_compute_acceleration(limited=True)}
    out = {}
    for name, fn in calls.items():
        fn()
        ps.synchronize()
        c0 = ps.counters()
        ps.pair_timing(reps + 2)
        for _ in range(reps):
            fn()
        ms, n = ps.pair_timing_read()
        ps.pair_timing(0)
        c1 = ps.counters()
        pairs = (c1["pairs"] - c0["pairs"]) / max(n, 1)
        act = (c1["active"] - c0["active"]) / max(n, 1)
        bp, ba = ALE_PASS_BYTES[prec][name]
        nbytes = pairs * bp + act * ba
        avg = ms / max(n, 1)
        gbs = nbytes / (avg * 1e-3) / 1e9
        key = f"ale{'32' if prec == T.PREC_MIXED else '64'}:{name}"
        out[name] = {"avg_launch_ms": avg, "pairs_per_launch": pairs, "algorithmic_bytes_per_launch":
                     nbytes, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                     **valu_ceiling(key, avg, ALE_SQ_PROFILE)}
    return out


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
            out[name]["roofline_passes"] = ale_pass_roofline(ps, prec)
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
            rows, waves, bal, _, ppl = T.autotune(bs, T.FORCES_WCSPH, True)
            t, cnt = timed(bs, lambda k: bs.step(T.WALL_MDBC, k), steps, 3)
            out[name] = leg_line(t, cnt, steps, n, tile_rows=rows, waves_per_cu=waves,
                                 balanced=bal, particles_per_lane=round(ppl, 3))
    out["mixed_over_fp64"] = (out["mixed"]["pair_interactions_per_s"] /
                              out["fp64"]["pair_interactions_per_s"])
    return out


def leg_inflow(steps):
    """8 x 0.5 m channel with inlet/outlet buffer zones, dx = 0.5 mm (4M particles),
    FP32 pairs with Kahan-compensated FP64 integration."""
    case = T.build_case(T.CASE_INFLOW_CHANNEL)
    with T.system_for_case(case, precision=T.PREC_MIXED) as ps:
        ps.set_compensated(True)
        ps.particle_to_cell()
        t, cnt = timed(ps, lambda k: ps.step_inflow(case.wall, case.dt, k), steps, 3)
    return {"mixed": leg_line(t, cnt, steps, case.np, appended=cnt["appended"],
                             removed=cnt["removed"], refused=cnt["zone_rejected"],
                             integration="fp64, Kahan-compensated r and rho")}


def host_threads() -> int:
    """Threads for the oracle: the host cores this process may use, i.e. the CPU
    affinity mask, or OMP_NUM_THREADS where a machine that shares its host
    between GPUs sets a smaller share there."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(int(env), aff) if env.isdigit() and int(env) > 0 else aff)


def host_info() -> dict:
    """The host CPU as lscpu and the kernel report it, and the share this
    process may use (affinity, cgroup quota, OMP_NUM_THREADS)."""
    import subprocess
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for ln in txt.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in keys and keys[k.strip()] not in info:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    if isinstance(info.get("sockets"), int) and isinstance(info.get("cores_per_socket"), int):
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        info["cgroup_cpus"] = None if quota == "max" else float(quota) / float(period)
    except (OSError, ValueError):
        pass
    info["omp_num_threads"] = os.environ.get("OMP_NUM_THREADS")
    return info


def cpu_baseline(case, budget_s: float = 20.0, threads: int = 0, max_steps: int = 50):
    """Oracle (C++ restatement of the reference path, OpenMP) on the host cores.
    Returns (steps, seconds, threads, oracle) so that the same run can go on as
    the checker of the error legs."""
    from oracle import pyoracle as po
    from tests.helpers import oracle_init, oracle_step

    threads = threads or host_threads()
    O = po.oracle_case(case.oracle_params(), case.arrays, omp_threads=threads)
    # the initial cell list only, as the GPU runs it is compared with start
    # (tsph_particle_to_cell, then steps); the pair count of one pass is taken
    # from the GPU counters of the same case
    oracle_init(O)
    t0 = time.perf_counter()
    n = 0
    while True:
        oracle_step(O, case.wall, case.dt)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= max_steps:
            break
    return n, time.perf_counter() - t0, threads, O


# ---------------------------------------------------------------------------
# error of every reported field against the FP64 oracle, per workload.  The
# oracle runs here only as the checker (the timed GPU legs never call it).
# ---------------------------------------------------------------------------
def field_errors(got: dict, ref: dict, rho0: float = 1000.0) -> dict:
    """Relative L2 error of rho - rho0, p, v and r (and max |d| / max |ref|)."""
    def l2(a, b):
        a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def mx(a, b):
        a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    if len(got["type"]) != len(ref["type"]) or not np.array_equal(got["type"], ref["type"]):
        return {"types_equal": False, "np": [int(len(got["type"])), int(len(ref["type"]))]}
    return {"rho-rho0_rel_l2": l2(got["rho"] - rho0, ref["rho"] - rho0),
            "p_rel_l2": l2(got["p"], ref["p"]), "v_rel_l2": l2(got["v"], ref["v"]),
            "r_rel_l2": l2(got["r"], ref["r"]), "rho_rel_max": mx(got["rho"], ref["rho"]),
            "r_rel_max": mx(got["r"], ref["r"])}


def oracle_fields(O) -> dict:
    from oracle import pyoracle as po
    return {"type": O.types(), "r": O.get(po.F_R), "v": O.get(po.F_V), "p": O.get(po.F_P),
            "rho": O.get(po.F_RHO)}


def gpu_fields(case, precision, steps, kind="symplectic", muscl=0, compensated=False):
    """The GPU state after `steps` driver steps from the case's initial state."""
    with T.system_for_case(case, precision=precision) as ps:
        if compensated:
            ps.set_compensated(True)
        if kind == "ale":
            ps.ale_enable()
            ps.step_ale(case.dt, 1, steps, muscl)
        else:
            ps.particle_to_cell()
            if kind == "inflow":
                ps.step_inflow(case.wall, case.dt, steps)
            else:
                ps.step(case.wall, case.dt, steps)
        return ps.download(("type", "r", "v", "p", "rho")), ps.counters()


def errors_vs_oracle(threads: int, O_full, full_case, full_steps: int, gpu_full: dict) -> dict:
    """Each workload's mixed (and FP64) GPU state against the oracle after the
    workload's step count, at the sizes stated in each entry."""
    from tests.helpers import (oracle_add_zones, oracle_ale, oracle_ale_step, oracle_for,
                               oracle_inflow_step, oracle_init, oracle_step)

    out = {}
    precs = (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64))

    def log(name, t0):
        print(f"[bench] error leg {name} done {time.perf_counter() - t0:.1f}s", file=sys.stderr,
              flush=True)

    # dam break at full size (740k): the cpu_baseline run continued
    t0 = time.perf_counter()
    ref = oracle_fields(O_full)
    out["dam_break_740k"] = {"particles": full_case.np, "after_steps": full_steps,
                             **{k: field_errors(g, ref) for k, g in gpu_full.items()}}
    log("dam_break_740k", t0)

    # still water, the first workload at its configured size (dx 10 mm, 2000 steps)
    t0 = time.perf_counter()
    case = T.build_case(T.CASE_STILL_WATER_DBC, dx=0.01)
    O = oracle_for(case, threads)
    oracle_init(O)
    for _ in range(case.steps):
        oracle_step(O, case.wall, case.dt)
    ref = oracle_fields(O)
    out["still_water"] = {"particles": case.np, "after_steps": case.steps, "dx": 0.01,
                          **{n: field_errors(gpu_fields(case, p, case.steps)[0], ref) for n, p in precs}}
    log("still_water", t0)

    # dam break, 1000 steps at dx = 4 mm (the 740k case at 1000 steps:
    # profiles/error_vs_oracle_dam_break_740k.md)
    t0 = time.perf_counter()
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.004, jitter=0.01, seed=SEED)
    O = oracle_for(case, threads)
    oracle_init(O)
    for _ in range(1000):
        oracle_step(O, case.wall, case.dt)
    ref = oracle_fields(O)
    out["dam_break_dx4mm"] = {"particles": case.np, "after_steps": 1000, "dx": 0.004,
                              **{n: field_errors(gpu_fields(case, p, 1000)[0], ref) for n, p in precs}}
    log("dam_break_dx4mm", t0)

    # ALE-SPH MUSCL (minmod) dam break, 1000 steps at dx = 8 mm
    t0 = time.perf_counter()
    case = T.build_case(T.CASE_ALE_DAM_BREAK, dx=0.008)
    O = oracle_ale(case, threads)
    for s in range(1, 1001):
        oracle_ale_step(O, s, case.dt, T.MUSCL_MINMOD)
    ref = oracle_fields(O)
    out["ale_muscl_dx8mm"] = {
        "particles": case.np, "after_steps": 1000, "dx": 0.008, "muscl": "minmod",
        **{n: field_errors(gpu_fields(case, p, 1000, "ale", T.MUSCL_MINMOD)[0], ref)
           for n, p in precs}}
    log("ale_muscl_dx8mm", t0)

    # inflow channel (8 x 0.5 m), 1000 steps at dx = 5 mm: fields and the mass
    # fluxes through the inlet (appended) and the outlet plane (fluid turned
    # outlet, src/SPH_buffer_zones.h:164-171), times m over the run time
    t0 = time.perf_counter()
    case = T.build_case(T.CASE_INFLOW_CHANNEL, dx=0.005)
    O = oracle_for(case, threads)
    oracle_add_zones(O, case.zones)
    oracle_init(O)
    steps = 1000
    for _ in range(steps):
        oracle_inflow_step(O, case.wall, case.dt)
    ref = oracle_fields(O)
    m = 1000.0 * case.params["dp"] ** 2
    t_run = steps * case.dt
    out_t = case.zones[1]["btype"]

    def fluxes(types, removed):
        n0 = len(case.arrays["type"])
        conv = int((types == out_t).sum()) - int((case.arrays["type"] == out_t).sum()) + removed
        app = len(types) - n0 + removed
        return app * m / t_run, conv * m / t_run

    ref_in, ref_out = fluxes(ref["type"], O.call("np_out"))
    entry = {"particles": case.np, "after_steps": steps, "dx": 0.005,
             "oracle_mass_flux_in_kg_per_s_m": ref_in, "oracle_mass_flux_out_kg_per_s_m": ref_out}
    entry["mixed_integration"] = "fp64, Kahan-compensated r and rho"
    for n, p in precs:
        g, cnt = gpu_fields(case, p, steps, "inflow", compensated=p == T.PREC_MIXED)
        f_in, f_out = fluxes(g["type"], cnt["removed"])
        entry[n] = {**field_errors(g, ref),
                    "mass_flux_in_rel_err": abs(f_in - ref_in) / max(abs(ref_in), 1e-300),
                    "mass_flux_out_rel_err": abs(f_out - ref_out) / max(abs(ref_out), 1e-300)}
    out["inflow_dx5mm"] = entry
    log("inflow_dx5mm", t0)

    # batched ensemble: its three smallest tanks, batched together on the GPU,
    # each against its own oracle run (500 steps)
    t0 = time.perf_counter()
    tanks = sorted(T.build_ensemble(64, SEED), key=lambda c: c.np)[:3]
    refs = []
    for c in tanks:
        O = oracle_for(c, threads)
        oracle_init(O)
        for _ in range(500):
            oracle_step(O, c.wall, c.dt)
        refs.append(oracle_fields(O))
    entry = {"tanks": [c.np for c in tanks], "after_steps": 500}
    for n, p in precs:
        with T.batch_for_cases(tanks, p) as bs:
            bs.particle_to_cell()
            bs.step(T.WALL_MDBC, 500)
            got = [bs.download_case(q, ("type", "r", "v", "p", "rho")) for q in range(len(tanks))]
        entry[n] = [field_errors(g, r) for g, r in zip(got, refs)]
    out["ensemble_smallest_3"] = entry
    log("ensemble_smallest_3", t0)
    return out


def developed_run(dx: float = 0.004, t_end: float = 0.6) -> dict:
    """The mixed path against the FP64 path (both on the GPU) over a physically
    developed dam break: the 3.22 m tank at dx = 4 mm to t = 0.6 s, past the
    front's impact on the right wall (about t = 0.55 s)."""
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=dx, jitter=0.01, seed=SEED)
    steps = int(round(t_end / case.dt))
    st = {n: gpu_fields(case, p, steps)[0] for n, p in (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64))}
    front = {n: float(s["r"][s["type"] == 0, 0].max()) for n, s in st.items()}
    return {"particles": case.np, "dx": dx, "steps": steps, "t_end_s": steps * case.dt,
            "front_x_m": front, **field_errors(st["mixed"], st["fp64"])}


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
    ap.add_argument("--oracle-steps", type=int, default=100,
                    help="steps of the 740k-particle case at which its error against the "
                         "oracle is quoted (0: no error legs against the oracle)")
    ap.add_argument("--no-developed", dest="developed", action="store_false",
                    help="skip the mixed-vs-FP64 run to the dam-break front's wall impact")
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
        if prec == T.PREC_MIXED:  # one of the step's two launches also corrects every particle
            bytes_per_launch += 0.5 * B_CORR * case.np
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
                 "roofline_frac": d_ach / HBM_PEAK_GBS, **valu_ceiling("k_tile64<0>", d_avg)},
        "mixed_over_fp64": rate(mixed) / rate(fp64),
        "error_vs_fp64": err,
        # `frac` is the algorithmic-byte roofline (neighbour records gathered per
        # pair, as if none were reused); the staged tiles serve most of them from
        # LDS, so the kernel's own ceiling is VALU issue (`valu_frac`, from the
        # SQ_INSTS_VALU count) and its measured HBM rate is `traffic_gbs`
        "roofline": {"bound": "hbm", "achieved": m_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m_ach / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_gbs": traffic / (m_avg * 1e-3) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (m_avg * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "binding": "valu issue", **valu_ceiling(KERNEL, m_avg),
                     "kernel": KERNEL, "avg_launch_ms": m_avg,
                     "algorithmic_bytes_per_launch": m_bytes, "pairs_per_launch": m_ppl},
    }
    if want:
        line["workloads"] = run_workloads()
    if args.developed:
        line["error_mixed_vs_fp64_developed"] = developed_run()
        print("[bench] developed run done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        n, t, threads, O = cpu_baseline(case, args.cpu_budget)
        pairs_per_step = fp64["pairs"] / k
        line["cpu_baseline"] = {"value": n * pairs_per_step / t, "unit": "pair-interactions/s",
                                "cores": threads, "kind": "port",
                                "sample": f"{n} full symplectic steps of the same 740k-particle "
                                          f"dam break (oracle, OpenMP, {threads} threads), "
                                          f"{t:.1f} s",
                                "steps_per_s": n / t, "host": host_info()}
        print(f"[bench] cpu baseline done ({threads} threads)", file=sys.stderr, flush=True)
        if args.oracle_steps > 0:
            # the same oracle run goes on, as the checker, to the step count the
            # full-size error is quoted at
            for _ in range(max(args.oracle_steps - n, 0)):
                from tests.helpers import oracle_step
                oracle_step(O, case.wall, case.dt)
            steps_full = max(args.oracle_steps, n)
            gpu_full = {nm: gpu_fields(case, p, steps_full)[0]
                        for nm, p in (("mixed", T.PREC_MIXED), ("fp64", T.PREC_FP64))}
            line["error_vs_oracle"] = errors_vs_oracle(threads, O, case, steps_full, gpu_full)
        del O
        n1, t1, _, _ = cpu_baseline(case, args.cpu_budget / 3, threads=1, max_steps=5)
        line["cpu_baseline"]["single_thread"] = {
            "value": n1 * pairs_per_step / t1, "unit": "pair-interactions/s", "cores": 1,
            "sample": f"{n1} full symplectic steps, 1 thread, {t1:.1f} s"}
    if args.copy_probe:
        copy_probe()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
