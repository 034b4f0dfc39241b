"""GPU parity: the HIP pipeline against the oracle on the same inputs.

Integer results (cell indices, cell lists, neighbour counts) are compared
bit-exactly.  FP64 fields carry a tolerance because the device pow() (Fourtakas
hydrostatic term, Tait EOS) is not glibc's: a few ulp per call, amplified by
the cancellation in p = B((rho/rho0)^7 - 1).  The stated bounds are:
  FP64 pair loop, one evaluation  : rel. error <= 1e-9 on a and dRho
  FP64, 20 symplectic steps       : rel. error <= 1e-9 on r, v, rho; <= 1e-6 on p
  mixed (FP32 pairs), one pass    : rel. L2 <= 1e-4 on a and dRho
  mixed, 1000 symplectic steps    : rel. L2 <= 1e-4 on rho (FP64, same run: <= 1e-8)
"""
import math

import numpy as np
import pytest

import tinysph_old_amd as T
from tinysph_old_amd import _lib as L
from oracle import pyoracle as po
from tests.helpers import (brute_force_pairs, oracle_fields, oracle_for, oracle_init, oracle_step,
                           oracle_variant, rel_err, rel_l2)

pytestmark = pytest.mark.gpu

CASES = {
    "still_water_dbc": dict(preset=T.CASE_STILL_WATER_DBC, dx=0.02),
    "dam_break_mdbc": dict(preset=T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01),
    "dam_break_bt": dict(preset=T.CASE_DAM_BREAK_BT, dx=0.01),
}


def make(name):
    print(f"[case] {name} {CASES[name]}", flush=True)
    return T.build_case(**CASES[name])


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def active_mask(types, wall):
    if wall == T.WALL_DBC:
        return (types != 6) & (types != 11)
    return types == 0


@pytest.mark.parametrize("name", list(CASES))
def test_cell_lists_bit_exact(name):
    case = make(name)
    O = oracle_for(case)
    oracle_init(O)
    start_o, items_o = O.cell_csr()
    with T.system_for_case(case) as ps:
        assert ps.grid() == O.grid()
        ps.particle_to_cell()
        start_g, items_g = ps.cell_lists()
    assert np.array_equal(start_g, start_o)
    assert np.array_equal(items_g, items_o)


def test_cell_lists_after_motion():
    """Jitter every particle by up to 0.4 cell, rebuild: still identical."""
    case = make("dam_break_mdbc")
    rng = np.random.default_rng(7)
    kh = case.params["kap"] * case.params["h"]
    case.arrays["r"] = case.arrays["r"] + rng.uniform(-0.4 * kh, 0.4 * kh, case.arrays["r"].shape)
    O = oracle_for(case)
    oracle_init(O)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.particle_to_cell()  # second build starts from the sorted order
        s, it = ps.cell_lists()
    s_o, it_o = O.cell_csr()
    assert np.array_equal(s, s_o) and np.array_equal(it, it_o)


@pytest.mark.parametrize("mode", [1, 2], ids=["ranked", "separate"])
def test_cell_order_modes(mode):
    """Both forms of the per-cell order (ranked by the pass that moves the
    particles, or a pass of its own) give the oracle's cell lists after motion
    and its state after 20 FP64 symplectic steps (pooled pair passes, whose
    chunks the builds cut)."""
    print(f"[case] dam_break_mdbc dx=0.01 cell order {mode}", flush=True)
    case = make("dam_break_mdbc")
    moved = T.build_case(**CASES["dam_break_mdbc"])
    rng = np.random.default_rng(7)
    kh = moved.params["kap"] * moved.params["h"]
    moved.arrays["r"] = moved.arrays["r"] + rng.uniform(-0.4 * kh, 0.4 * kh, moved.arrays["r"].shape)
    O = oracle_for(moved)
    oracle_init(O)
    with T.system_for_case(moved) as ps:
        ps.set_cell_order(mode)
        ps.particle_to_cell()
        ps.particle_to_cell()
        s, it = ps.cell_lists()
    s_o, it_o = O.cell_csr()
    assert np.array_equal(s, s_o) and np.array_equal(it, it_o)
    nsteps = 20
    O = oracle_for(case)
    oracle_init(O)
    for _ in range(nsteps):
        oracle_step(O, case.wall, case.dt)
    ref = oracle_fields(O)
    with T.system_for_case(case) as ps:
        ps.set_cell_order(mode)
        ps.set_balanced(True)
        ps.particle_to_cell()
        ps.step(case.wall, case.dt, nsteps)
        got = ps.download()
    assert np.array_equal(got["type"], ref["type"])
    assert rel_err(got["r"], ref["r"]) <= 1e-9
    assert rel_err(got["v"], ref["v"]) <= 1e-9
    assert rel_err(got["rho"], ref["rho"]) <= 1e-9
    assert rel_err(got["p"], ref["p"]) <= 1e-6


@pytest.mark.parametrize("name", list(CASES))
def test_forces_fp64(name):
    case = make(name)
    wall = case.wall
    O = oracle_for(case)
    oracle_init(O)
    if wall == T.WALL_MDBC:
        O.call("mdbc_density")
    elif wall == T.WALL_BT:
        O.call("bt_density")
    O.call("density_to_pressure")
    O.call("compute_forces", oracle_variant(wall))
    ref = oracle_fields(O)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        if wall == T.WALL_MDBC:
            ps.mdbc_density()
        elif wall == T.WALL_BT:
            ps.bt_density()
        ps.density_to_pressure()
        ps.compute_forces(T.forces_variant(wall))
        got = ps.download()
        cnt = ps.counters()
    act = active_mask(ref["type"], wall)
    assert rel_err(got["rho"], ref["rho"]) <= 1e-12
    assert rel_err(got["p"], ref["p"]) <= 1e-6
    assert rel_err(got["a"][act], ref["a"][act]) <= 1e-9
    assert rel_err(got["drho"][act], ref["drho"][act]) <= 1e-9
    # pair count against brute force
    assert cnt["active"] == int(act.sum())
    assert cnt["pairs"] == brute_force_pairs(case.arrays["r"], act, case.params["h"]) or \
        wall == T.WALL_DBC  # DBC skips wall-wall pairs; checked separately below


def test_pair_count_dbc_skips_wall_wall():
    case = make("still_water_dbc")
    r, t, h = case.arrays["r"], case.arrays["type"], case.params["h"]
    expect = 0
    for i in range(len(r)):
        d = np.sqrt(((r - r[i]) ** 2).sum(axis=1))
        m = d / h < 2.0
        m[i] = False
        if t[i] == 1:
            m &= t != 1
        expect += int(m.sum())
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.density_to_pressure()
        ps.compute_forces(T.FORCES_DBC)
        assert ps.counters()["pairs"] == expect


@pytest.mark.parametrize("name", list(CASES))
def test_steps_fp64(name):
    case = make(name)
    nsteps = 20
    O = oracle_for(case)
    oracle_init(O)
    for _ in range(nsteps):
        oracle_step(O, case.wall, case.dt)
    ref = oracle_fields(O)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.step(case.wall, case.dt, nsteps)
        got = ps.download()
    assert len(got["rho"]) == len(ref["rho"])
    assert np.array_equal(got["type"], ref["type"])
    assert rel_err(got["r"], ref["r"]) <= 1e-9
    assert rel_err(got["v"], ref["v"]) <= 1e-9
    assert rel_err(got["rho"], ref["rho"]) <= 1e-9
    assert rel_err(got["p"], ref["p"]) <= 1e-6


def stage_step(ps, wall, dt):
    """One symplectic step launched stage by stage, like the drivers
    (cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:59-116)."""
    variant = T.forces_variant(wall)
    if wall == T.WALL_MDBC:
        ps.mdbc_density(); ps.density_to_pressure()
    elif wall == T.WALL_BT:
        ps.bt_density(); ps.density_to_pressure()
    ps.predictor(dt)
    ps.clear_cells(); ps.particle_to_cell(); ps.density_to_pressure()
    ps.compute_forces(variant); ps.corrector(dt); ps.remove_out_of_domain()
    ps.clear_cells(); ps.particle_to_cell(); ps.density_to_pressure()
    ps.compute_forces(variant)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED], ids=["fp64", "mixed"])
def test_graph_replay_equals_eager(name, precision):
    """The captured step (wall densities in workgroups of the first cell
    build's hash launch; on the mixed path the corrector inside the first pair
    pass) equals the stages launched one by one, bit for bit."""
    case = make(name)
    with T.system_for_case(case, precision) as a, T.system_for_case(case, precision) as b:
        a.particle_to_cell()
        b.particle_to_cell()
        a.step(case.wall, case.dt, 5)  # graph
        for _ in range(5):
            stage_step(b, case.wall, case.dt)
        ga, gb = a.download(), b.download()
        assert a.counters()["removed"] == b.counters()["removed"]
    for k in ("type", "r", "v", "rho", "p", "a", "drho", "r_o", "v_o", "rho_o"):
        assert np.array_equal(ga[k], gb[k]), k


def test_step_counters_equal_stage_by_stage():
    """The captured step folds the EOS count after the wall densities into the
    wall pass and the removal decision into the cell build: the out-of-range
    density and removal counts equal those of the stages run one by one."""
    case = make("dam_break_mdbc")
    rho = case.arrays["rho"].copy()
    fluid = np.nonzero(case.arrays["type"] == 0)[0]
    rho[fluid[::97]] = 1600.0  # out of range for the EOS warning (src/SPH_pressure.h:76-80)
    case.arrays["rho"] = rho
    with T.system_for_case(case) as a, T.system_for_case(case) as b:
        a.particle_to_cell()
        b.particle_to_cell()
        a.reset_counters()
        b.reset_counters()
        a.step(case.wall, case.dt, 2)
        for _ in range(2):
            b.mdbc_density(); b.density_to_pressure(); b.predictor(case.dt)
            b.clear_cells(); b.particle_to_cell(); b.density_to_pressure()
            b.compute_forces(T.FORCES_WCSPH); b.corrector(case.dt); b.remove_out_of_domain()
            b.clear_cells(); b.particle_to_cell(); b.density_to_pressure()
            b.compute_forces(T.FORCES_WCSPH)
        ca, cb = a.counters(), b.counters()
        ga, gb = a.download(), b.download()
    assert ca["density_out_of_range"] > 0
    for key in ("density_out_of_range", "removed", "pairs", "active"):
        assert ca[key] == cb[key], key
    for k in ("r", "v", "rho", "p", "a", "drho"):
        assert np.array_equal(ga[k], gb[k]), k


@pytest.mark.parametrize("name", list(CASES))
def test_forces_mixed(name):
    case = make(name)
    wall = case.wall
    O = oracle_for(case)
    oracle_init(O)
    if wall == T.WALL_MDBC:
        O.call("mdbc_density")
    elif wall == T.WALL_BT:
        O.call("bt_density")
    O.call("density_to_pressure")
    O.call("compute_forces", oracle_variant(wall))
    ref = oracle_fields(O)
    with T.system_for_case(case, precision=T.PREC_MIXED) as ps:
        ps.particle_to_cell()
        if wall == T.WALL_MDBC:
            ps.mdbc_density()
        elif wall == T.WALL_BT:
            ps.bt_density()
        ps.density_to_pressure()
        ps.compute_forces(T.forces_variant(wall))
        got = ps.download()
    act = active_mask(ref["type"], wall)
    assert rel_l2(got["a"][act], ref["a"][act]) <= 1e-4
    # BT on the hydrostatic start: dRho is a small remainder of the diffusion
    # term's (rho_j - rho_i) - drho_H cancellation (|dRho| ~ 1e-2 against terms
    # of ~1e-1 per pair), so its FP32 relative error is ~10x the others'
    drho_tol = 1e-3 if wall == T.WALL_BT else 1e-4
    assert rel_l2(got["drho"][act], ref["drho"][act]) <= drho_tol


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED])
def test_tile_size_does_not_change_results(precision):
    """The per-particle summation order does not depend on the tile shape, so
    every tile size (including the LDS-overflow fallback of 24-row tiles in
    dense columns) gives identical bits."""
    case = make("dam_break_mdbc")
    outs = []
    for rows in (1, 3, 7, 24):
        with T.system_for_case(case, precision=precision) as ps:
            ps.set_tile_rows(rows)
            ps.particle_to_cell()
            ps.mdbc_density()
            ps.density_to_pressure()
            ps.compute_forces(T.FORCES_WCSPH)
            outs.append(ps.download(("a", "drho")))
    for o in outs[1:]:
        if precision == T.PREC_FP64:
            assert np.array_equal(o["a"], outs[0]["a"])
            assert np.array_equal(o["drho"], outs[0]["drho"])
        else:
            # the mixed path stages positions relative to the tile, so the
            # FP32 rounding of dr (not the summation order) moves with it
            assert rel_l2(o["a"], outs[0]["a"]) <= 1e-5
            assert rel_l2(o["drho"], outs[0]["drho"]) <= 1e-5


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED])
def test_tile_walks_agree(precision):
    """The static and the pooled tile walk (and the waves per CU) only change
    which wavefront computes a tile, never the tile frame or the summation
    order: identical bits and counts.  dx = 4 mm gives 6,808 tiles: at 1 wave
    per CU (26 tiles per wave) the pooled walk is taken, at 2 (13) it strides."""
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.004, jitter=0.01, seed=11)
    print(f"[case] dam_break_mdbc dx=0.004 precision={precision}", flush=True)
    outs = []
    with T.system_for_case(case, precision=precision) as ps:
        ps.particle_to_cell()
        ps.mdbc_density()
        ps.density_to_pressure()
        for bal, waves in ((0, 8), (1, 1), (1, 2), (0, 2)):
            ps.set_balanced(bal)
            ps.set_waves_per_cu(waves)
            # first with the list re-cut for the new walk by the tile-list
            # pass, then with the list and chunks the cell-list build cuts
            for rebuild in (False, True):
                if rebuild:
                    ps.particle_to_cell()
                    ps.density_to_pressure()
                ps.reset_counters()
                ps.compute_forces(T.FORCES_WCSPH)
                out = ps.download(("a", "drho"))
                out["counters"] = ps.counters()
                outs.append(out)
    for o in outs[1:]:
        assert np.array_equal(o["a"], outs[0]["a"])
        assert np.array_equal(o["drho"], outs[0]["drho"])
        for key in ("pairs", "candidates", "active", "pair_passes"):
            assert o["counters"][key] == outs[0]["counters"][key], key
        assert o["counters"]["scan_stalls"] == 0


def test_water_level_and_probe():
    case = make("dam_break_mdbc")
    O = oracle_for(case)
    oracle_init(O)
    O.call("density_to_pressure")
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.density_to_pressure()
        for x in (0.3, 0.865):
            wl_g = ps.water_elevation(x, 0.0)
            wl_o = O.call("water_elevation", x, 0.0)
            assert abs(wl_g - wl_o) <= 1e-12, (x, wl_g, wl_o)
            wli_g = ps.water_elevation(x, 0.9, inverse=True)
            wli_o = O.call("water_elevation_inverse", x, 0.9)
            assert abs(wli_g - wli_o) <= 1e-12, (x, wli_g, wli_o)
        for pt in ((0.5, 0.1), (1.0, 0.05), (0.01, 0.3)):
            g = ps.interpolate_point(*pt)
            o = O.interpolate_point(*pt)
            assert np.allclose(g, o, rtol=1e-9, atol=1e-9), (pt, g, o)


def test_remove_out_of_domain_alternation():
    """Three consecutive fluid particles outside the domain: the reference's
    in-loop erase removes the 1st and 3rd and skips the 2nd."""
    case = make("still_water_dbc")
    r = case.arrays["r"].copy()
    t = case.arrays["type"]
    fl = np.nonzero(t == 0)[0]
    for k in (fl[10], fl[11], fl[12], fl[40]):
        r[k] = (case.params["x_m"] + 0.0, 0.2)  # x > x_m - eps
    case.arrays["r"] = r
    O = oracle_for(case)
    oracle_init(O)
    removed_o = O.call("remove_out_of_domain")
    assert O.call("particle_to_cell") == 0
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.remove_out_of_domain()
        ps.particle_to_cell()
        got = ps.download()
        cnt = ps.counters()
    assert removed_o == 3
    assert cnt["removed"] == 3
    assert ps is not None
    assert np.array_equal(got["r"], O.get(po.F_R))
    s_o, it_o = O.cell_csr()
    assert len(got["r"]) == O.np


# 1000 oracle steps on the host take about two and a half minutes
@pytest.mark.timeout(600)
def test_mixed_density_error_after_1000_steps():
    """The mixed path's density after 1000 dam-break steps against the oracle
    (dx = 4 mm, 46k particles, jittered lattice): relative L2 <= 1e-4."""
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.004, jitter=0.01, seed=11)
    nsteps = 1000
    print(f"[case] dam break dx=4mm np={case.np}, {nsteps} steps", flush=True)
    O = oracle_for(case)
    oracle_init(O)
    for _ in range(nsteps):
        oracle_step(O, case.wall, case.dt)
    ref = oracle_fields(O)
    got = {}
    for prec in (T.PREC_FP64, T.PREC_MIXED):
        with T.system_for_case(case, precision=prec) as ps:
            ps.particle_to_cell()
            ps.step(case.wall, case.dt, nsteps)
            got[prec] = ps.download()
    errs = {}
    for prec, g in got.items():
        assert np.array_equal(g["type"], ref["type"])
        errs[prec] = {k: rel_l2(g[k], ref[k]) for k in ("rho", "p", "v", "r")}
    print(f"[errors] fp64 {errs[T.PREC_FP64]} mixed {errs[T.PREC_MIXED]}", flush=True)
    assert errs[T.PREC_FP64]["rho"] <= 1e-8
    assert errs[T.PREC_MIXED]["rho"] <= 1e-4


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED], ids=["fp64", "mixed"])
def test_scan_fallback_gives_same_cell_lists(precision):
    """A cell-list scan block that outwaits its bounded poll for the sums of
    the blocks before it forms its prefix from the cell counts instead.  Forced
    for every block (scan wait -1), the cell lists, the pooled chunks (through
    the forces and pair counts) and whole graph-replayed steps are bit-identical
    to the polled scan's, and the stalls are counted."""
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.004, jitter=0.01, seed=11)
    print(f"[case] dam_break_mdbc dx=0.004 np={case.np} precision={precision}", flush=True)
    outs = []
    for wait in (1 << 20, -1):
        with T.system_for_case(case, precision=precision) as ps:
            ps.set_scan_wait(wait)
            ps.set_waves_per_cu(1)  # enough tiles per wave for the pooled walk
            ps.set_balanced(True)
            ps.particle_to_cell()
            cells = ps.cell_lists()
            ps.mdbc_density()
            ps.density_to_pressure()
            ps.reset_counters()
            ps.compute_forces(T.FORCES_WCSPH)
            first = ps.counters()
            forces = ps.download(("a", "drho"))
            ps.step(case.wall, case.dt, 5)
            state = ps.download(("r", "v", "rho", "p", "a", "drho"))
            outs.append((cells, first, forces, state, ps.counters()))
    (c0, f0, a0, s0, n0), (c1, f1, a1, s1, n1) = outs
    assert np.array_equal(c0[0], c1[0]) and np.array_equal(c0[1], c1[1])
    for key in ("pairs", "candidates", "active"):
        assert f0[key] == f1[key], key
    for k in a0:
        assert np.array_equal(a0[k], a1[k]), k
    for k in s0:
        assert np.array_equal(s0[k], s1[k]), k
    assert n0["scan_stalls"] == 0
    assert n1["scan_stalls"] > 0


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED], ids=["fp64", "mixed"])
def test_empty_system(precision):
    """No particles: every stage and whole graph-replayed steps run, count
    nothing and leave an empty system."""
    case = make("dam_break_mdbc")
    empty = {k: np.asarray(v)[:0] for k, v in case.arrays.items() if v is not None}
    with T.ParticleSystem(case.params, precision=precision, capacity=64) as ps:
        ps.upload(empty)
        ps.particle_to_cell()
        ps.mdbc_density()
        ps.density_to_pressure()
        ps.compute_forces(T.FORCES_WCSPH)
        ps.step(case.wall, case.dt, 3)
        assert ps.np == 0
        cnt = ps.counters()
        for key in ("pairs", "candidates", "active", "outside_grid", "removed", "scan_stalls"):
            assert cnt[key] == 0, key
        start, items = ps.cell_lists()
        assert not start.any() and len(items) == 0
        assert len(ps.download(("r", "rho"))["rho"]) == 0


def test_particle_outside_grid_is_counted_and_binned():
    """A particle beyond the cell grid (undefined behaviour in the reference,
    src/SPH_find_pairs.h:64-69) is counted, binned in the nearest edge cell and
    paired like any other; every particle appears in the cell lists once."""
    case = make("dam_break_mdbc")
    arrays = {k: (np.array(v, copy=True) if v is not None else None) for k, v in case.arrays.items()}
    fluid = np.flatnonzero(np.asarray(arrays["type"]) == T.FLUID)
    r = arrays["r"].reshape(-1, 2)
    r[fluid[0]] = (case.params["x_m"] + 1.0, case.params["y_m"] + 1.0)
    with T.ParticleSystem(case.params, precision=T.PREC_FP64, capacity=case.np) as ps:
        ps.upload(arrays)
        ps.particle_to_cell()
        cnt = ps.counters()
        assert cnt["outside_grid"] == 1
        start, items = ps.cell_lists()
        assert start[-1] == case.np
        assert np.array_equal(np.sort(items), np.arange(case.np))
        ps.density_to_pressure()
        ps.compute_forces(T.FORCES_WCSPH)
        out = ps.download(("a", "drho"))
        assert np.isfinite(out["a"]).all() and np.isfinite(out["drho"]).all()
