"""GPU parity of the batched ensemble (tsph_batch_*) and of the autotuner.

A batch of uneven cases advanced together must give every case exactly the
state a single-case context gives it: the same kernels visit the same
neighbours in the same order, so the comparison is bit-exact, in FP64 and in
the mixed path.  FP64 batch cases also match the oracle within 1e-9 relative
(1e-6 on p), like the single-case FP64 parity tests.  Removal runs must not
continue across case boundaries (each case keeps the reference's in-loop erase
rule on its own).  The 64-tank workload runs at full size with its particle
bookkeeping checked and two of its tanks compared with single-case runs.
"""
import ctypes as C

import numpy as np
import pytest

import tinysph_old_amd as T
from tinysph_old_amd import _lib as L
from oracle import pyoracle as po
from tests.helpers import oracle_for, oracle_init, oracle_step, rel_err

pytestmark = pytest.mark.gpu

SEED = 20241202
FIELDS = ("type", "index", "p", "rho", "drho", "r", "v", "a", "r_o", "v_o", "rho_o")


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def run_single(case, precision, steps, prep=None):
    with T.system_for_case(case, precision=precision) as ps:
        ps.particle_to_cell()
        if prep:
            prep(ps)
        if steps:
            ps.step(case.wall, case.dt, steps)
        return ps.download(FIELDS), ps.counters()


def assert_same(got, ref, what):
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), (what, k)


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED], ids=["fp64", "mixed"])
def test_batch_equals_single_cases(precision):
    cases = T.build_ensemble(6, SEED, 2000, 12000)
    steps = 25
    print(f"batch of {len(cases)}: {[c.np for c in cases]}", flush=True)
    with T.batch_for_cases(cases, precision) as bs:
        bs.particle_to_cell()
        bs.step(T.WALL_MDBC, steps)
        npc = bs.np_per_case()
        got = [bs.download_case(q, FIELDS) for q in range(len(cases))]
        cnt = bs.counters()
    pairs = 0
    for q, case in enumerate(cases):
        ref, c1 = run_single(case, precision, steps)
        assert npc[q] == len(ref["type"])
        assert_same(got[q], ref, f"case {q}")
        pairs += c1["pairs"]
    assert cnt["pairs"] == pairs
    assert cnt["outside_grid"] == 0


def test_batch_fp64_matches_oracle():
    cases = T.build_ensemble(3, SEED + 1, 2000, 6000)
    steps = 6
    with T.batch_for_cases(cases, T.PREC_FP64) as bs:
        bs.particle_to_cell()
        bs.step(T.WALL_MDBC, steps)
        got = [bs.download_case(q, FIELDS) for q in range(len(cases))]
    for q, case in enumerate(cases):
        O = oracle_for(case)
        oracle_init(O)
        for _ in range(steps):
            oracle_step(O, case.wall, case.dt)
        g = got[q]
        assert np.array_equal(g["type"], O.types())
        assert rel_err(g["r"], O.get(po.F_R)) <= 1e-9
        assert rel_err(g["v"], O.get(po.F_V)) <= 1e-9
        assert rel_err(g["rho"], O.get(po.F_RHO)) <= 1e-9
        assert rel_err(g["p"], O.get(po.F_P)) <= 1e-6


@pytest.mark.parametrize("precision", [T.PREC_FP64, T.PREC_MIXED], ids=["fp64", "mixed"])
def test_batch_cell_order_modes_agree(precision):
    """A batch ordered by a pass of its own (the default above 2M particles)
    and one ranked by the reorder give identical states."""
    cases = T.build_ensemble(3, SEED + 3, 2000, 6000)
    outs = []
    for mode in (1, 2):
        print(f"[case] batch of 3, cell order {mode}, precision {precision}", flush=True)
        with T.batch_for_cases(cases, precision) as bs:
            bs.set_cell_order(mode)
            bs.particle_to_cell()
            bs.step(T.WALL_MDBC, 6)
            outs.append([bs.download_case(q, FIELDS) for q in range(len(cases))])
    for q in range(len(cases)):
        assert_same(outs[1][q], outs[0][q], f"case {q}")


def test_batch_graph_replay_equals_eager_launches():
    cases = T.build_ensemble(4, SEED + 2, 2000, 8000)
    out = []
    for eager in (False, True):
        with T.batch_for_cases(cases, T.PREC_MIXED) as bs:
            bs.particle_to_cell()
            if eager:
                bs.pair_timing(64)  # eager launches, each pair pass timed
            bs.step(T.WALL_MDBC, 7)
            out.append([bs.download_case(q, FIELDS) for q in range(len(cases))])
    for q in range(len(cases)):
        assert_same(out[0][q], out[1][q], f"case {q}")


def _escape(arrays, slots, x):
    a = {k: (v.copy() if v is not None else None) for k, v in arrays.items()}
    for s in slots:
        a["type"][s] = T.FLUID
        a["r"][s] = (x, a["r"][s][1])
    return a


def test_removal_runs_stop_at_case_boundaries():
    cases = T.build_ensemble(2, SEED + 3, 2000, 3000)
    A, B = cases
    xa = A.params["x_m"] - 0.005   # inside the grid, past x_m - eps (src/SPH_btdebug.h:17-22)
    xb = B.params["x_m"] - 0.005
    # a run of three candidates at the end of case A and at the start of case B
    A.arrays = _escape(A.arrays, [A.np - 3, A.np - 2, A.np - 1], xa)
    B.arrays = _escape(B.arrays, [0, 1, 2], xb)

    def prep(ps):
        ps.remove_out_of_domain()
        ps.particle_to_cell()

    refs = [run_single(c, T.PREC_FP64, 0, prep)[0] for c in cases]
    for c, ref in zip(cases, refs):
        assert len(ref["type"]) == c.np - 2   # the 1st and 3rd of the run are erased
    with T.batch_for_cases(cases, T.PREC_FP64) as bs:
        bs.particle_to_cell()
        prep(bs)
        assert list(bs.np_per_case()) == [c.np - 2 for c in cases]
        for q in range(2):
            assert_same(bs.download_case(q, FIELDS), refs[q], f"case {q}")
        assert bs.counters()["removed"] == 4


def test_batch_refuses_single_case_entry_points():
    cases = T.build_ensemble(2, SEED, 2000, 3000)
    lib = L.lib()
    with T.batch_for_cases(cases) as bs:
        h, _keep = T.system.host_particles(cases[0].arrays)
        assert lib.tsph_upload(bs.ctx, C.byref(h)) == L.ESTATE
        assert lib.tsph_step_symplectic(bs.ctx, T.WALL_MDBC, 1e-5, 1) == L.ESTATE
        assert lib.tsph_symplectic_predictor(bs.ctx, 1e-5) == L.ESTATE
        assert lib.tsph_ale_enable(bs.ctx) == L.ESTATE
        out = (C.c_double * 4)()
        assert lib.tsph_interpolate_point(bs.ctx, 0.1, 0.1, out) == L.ESTATE
        assert lib.tsph_batch_download(bs.ctx, 2, C.byref(h)) == L.EINVAL
        n = C.c_int()
        assert lib.tsph_batch_cases(bs.ctx, C.byref(n)) == 0 and n.value == 2
    with T.system_for_case(cases[0]) as ps:
        assert lib.tsph_batch_step(ps.ctx, T.WALL_MDBC, 1) == L.ESTATE


@pytest.mark.parametrize("batch", [False, True], ids=["single", "batch"])
def test_autotune_keeps_results(batch):
    """The tuned tile depth and wave count change only the launch shape: FP64
    results stay bit-identical to the default setting's."""
    if batch:
        cases = T.build_ensemble(4, SEED + 4, 2000, 20000)
    else:
        cases = [T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.004, jitter=0.01, seed=3)]
    res = []
    # both runs start from the a, drho of one pair pass (the tuner leaves them)
    for tune in (False, True):
        mk = (lambda: T.batch_for_cases(cases, T.PREC_FP64)) if batch else \
            (lambda: T.system_for_case(cases[0], precision=T.PREC_FP64))
        with mk() as ps:
            ps.particle_to_cell()
            if tune:
                rows, waves, bal, ms, ppl = T.autotune(ps, T.FORCES_WCSPH, True)
                assert rows in (5, 6, 7, 8, 10, 12, 14, 16, 20, 24)
                assert waves in (4, 8, 12, 16) and bal in (0, 1)
                assert ms > 0
                # particles per lane of the chosen depth: a full cell holds
                # about 8-12 particles here, so a tile of `rows` cells about
                # 0.13-0.19 per lane and row (less where columns are partly
                # filled)
                print(f"tuned: {rows} rows, {waves} waves per CU, balanced {bal}, "
                      f"{ppl:.2f} particles per lane", flush=True)
                assert 0.02 * rows < ppl < 0.3 * rows
                assert ps.tile_rows == rows
            else:
                ps.compute_forces(T.FORCES_WCSPH)
            if batch:
                ps.step(T.WALL_MDBC, 5)
                res.append([ps.download_case(q, FIELDS) for q in range(len(cases))])
            else:
                ps.step(cases[0].wall, cases[0].dt, 5)
                res.append([ps.download(FIELDS)])
    for q in range(len(cases)):
        assert_same(res[0][q], res[1][q], f"case {q}")


def test_ensemble_64_tanks_full_size():
    """The 64-tank workload at full size (7.8M particles): bookkeeping over a
    few mixed-precision steps and two tanks against single-case runs."""
    cases = T.build_ensemble(64, SEED)
    steps = 4
    n0 = np.array([c.np for c in cases])
    print(f"64 tanks, {int(n0.sum())} particles", flush=True)
    with T.batch_for_cases(cases, T.PREC_MIXED) as bs:
        bs.particle_to_cell()
        rows, waves, bal, _ = T.autotune(bs, T.FORCES_WCSPH)
        print(f"tuned: {rows} cells per tile, {waves} waves per CU, balanced {bal}", flush=True)
        bs.reset_counters()
        bs.step(T.WALL_MDBC, steps)
        cnt = bs.counters()
        npc = bs.np_per_case()
        small = [int(q) for q in np.argsort(n0)[:2]]
        got = {q: bs.download_case(q, FIELDS) for q in small}
    assert np.array_equal(npc, n0)
    assert cnt["outside_grid"] == 0 and cnt["removed"] == 0
    assert cnt["pair_passes"] == 2 * steps
    fluid = sum(c.n_fluid for c in cases)
    assert cnt["active"] == 2 * steps * fluid
    assert 15 * cnt["active"] < cnt["pairs"] < 40 * cnt["active"]
    def prep(ps):
        # the same tile frame as the tuned batch (the mixed path rounds in it)
        # and a, drho from one pair pass, as the tuner leaves them
        ps.set_tile_rows(rows)
        ps.compute_forces(T.FORCES_WCSPH)

    for q in small:
        ref, _ = run_single(cases[q], T.PREC_MIXED, steps, prep)
        assert_same(got[q], ref, f"tank {q}")


def test_batch_scan_fallback_gives_same_state():
    """The batched grid's scan with every block forming its prefix from the
    cell counts (scan wait -1) gives the same steps as the polled scan."""
    cases = T.build_ensemble(4, SEED + 5, 2000, 12000)
    got = []
    for wait in (1 << 20, -1):
        with T.batch_for_cases(cases, T.PREC_MIXED) as bs:
            bs.set_scan_wait(wait)
            bs.set_waves_per_cu(1)
            bs.set_balanced(True)
            bs.particle_to_cell()
            bs.step(T.WALL_MDBC, 5)
            got.append(([bs.download_case(q, FIELDS) for q in range(len(cases))], bs.counters()))
    (d0, n0), (d1, n1) = got
    for q in range(len(cases)):
        assert_same(d1[q], d0[q], f"case {q}")
    assert n0["pairs"] == n1["pairs"]
    assert n0["scan_stalls"] == 0 and n1["scan_stalls"] > 0
