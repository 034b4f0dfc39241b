"""GPU parity of the other WCSPH integrators and the density reinitialisation
against the oracle (which is pinned bit for bit to the reference's functions
in tests/test_oracle_integrators.py):

* single stages on the same state: Integrate_Verlet, Integrate_Euler and the
  leap-frog halves are HBM streams of IEEE operations without pow(), so they
  are compared bit for bit; the reinitialisation (sum + overwrite) too;
* whole driver steps (Verlet driver with its Euler steps, leap-frog driver,
  symplectic steps with the reinitialisation): FP64 within 1e-9 relative on
  r, v, rho and the history fields (1e-6 on p), the bounds of the other FP64
  step tests; graph replay bit-identical to eager launches;
* the mixed path on the Verlet driver: relative L2 <= 1e-4 on rho.
"""
import numpy as np
import pytest

import tinysph_old_amd as T
from oracle import pyoracle as po
from tests.helpers import (oracle_for, oracle_init, oracle_leapfrog_step, oracle_reinit, oracle_step,
                           oracle_verlet_step, rel_err, rel_l2)

pytestmark = pytest.mark.gpu

CASES = {
    "still_water_dbc": dict(preset=T.CASE_STILL_WATER_DBC, dx=0.02),
    "dam_break_mdbc": dict(preset=T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01),
    "dam_break_bt": dict(preset=T.CASE_DAM_BREAK_BT, dx=0.01),
}


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def make(name):
    print(f"[case] {name} {CASES[name]}", flush=True)
    return T.build_case(**CASES[name])


def compare(ps, O, tol=1e-9, tol_p=1e-6, history=True):
    g = ps.download(("type", "index", "r", "v", "rho", "p"))
    assert np.array_equal(g["type"], O.types())
    assert np.array_equal(g["index"], O.index())
    assert rel_err(g["r"], O.get(po.F_R)) <= tol
    assert rel_err(g["v"], O.get(po.F_V)) <= tol
    assert rel_err(g["rho"], O.get(po.F_RHO)) <= tol
    assert rel_err(g["p"], O.get(po.F_P)) <= tol_p
    if history:
        h = ps.history_download()
        assert rel_err(h["v_oo"], O.get(po.F_VOO)) <= tol
        assert rel_err(h["rho_oo"], O.get(po.F_RHOOO)) <= tol
        assert rel_err(h["rho_temp"], O.get(po.F_RHOTMP)) <= tol


def exact(ps, O):
    g = ps.download(("r", "v", "rho", "v_o", "rho_o"))
    h = ps.history_download()
    for k, f in (("r", po.F_R), ("v", po.F_V), ("rho", po.F_RHO), ("v_o", po.F_VO),
                 ("rho_o", po.F_RHOO)):
        assert np.array_equal(g[k], O.get(f)), k
    for k, f in (("v_oo", po.F_VOO), ("rho_oo", po.F_RHOOO), ("rho_temp", po.F_RHOTMP)):
        assert np.array_equal(h[k], O.get(f)), k


def test_integrator_stages_bit_exact():
    """Each integrator stage on the same state (a, drho from one pair pass)."""
    case = make("dam_break_mdbc")
    dt = case.dt
    O = oracle_for(case)
    oracle_init(O)
    O.call("density_to_pressure")
    O.call("compute_forces", 0)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.history_enable()
        ps.density_to_pressure()
        ps.compute_forces(T.FORCES_WCSPH)
        # a and drho must agree before the stages are compared bit for bit
        g = ps.download(("a", "drho"))
        O.set(po.F_A, g["a"])
        O.set(po.F_DRHO, g["drho"])
        for name, args in (("euler", (dt,)), ("verlet", (dt,)), ("verlet", (dt,)),
                           ("leapfrog_one", (dt,)), ("leapfrog_two", (dt, 1)),
                           ("leapfrog_one", (dt,)), ("leapfrog_two", (dt, 2))):
            getattr(ps, name)(*args)
            O.call(name, *args)
            exact(ps, O)


def test_density_reinit_bit_exact():
    case = make("dam_break_mdbc")
    O = oracle_for(case)
    oracle_init(O)
    for _ in range(5):
        oracle_step(O, case.wall, case.dt)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.history_enable()
        ps.step(case.wall, case.dt, 5)
        # the same state on both sides, then one reinitialisation
        g = ps.download(("r", "rho"))
        O.set(po.F_R, g["r"])
        O.set(po.F_RHO, g["rho"])
        O.call("clear_cells")
        assert O.call("particle_to_cell") == 0
        ps.reset_counters()
        bad = oracle_reinit(O)
        ps.density_reinit()
        h = ps.history_download()
        assert np.array_equal(h["rho_temp"], O.get(po.F_RHOTMP))
        assert np.array_equal(ps.download(("rho",))["rho"], O.get(po.F_RHO))
        assert ps.counters()["reinit_out_of_range"] == bad


@pytest.mark.parametrize("name", list(CASES))
def test_verlet_driver_steps(name):
    case = make(name)
    O = oracle_for(case)
    oracle_init(O)
    nsteps = 22  # Euler on steps 1 and 20
    for step in range(1, nsteps + 1):
        oracle_verlet_step(O, case.wall, case.dt, step)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.history_enable()
        ps.step_verlet(case.wall, case.dt, 1, nsteps)
        compare(ps, O)


@pytest.mark.parametrize("name", list(CASES))
def test_leapfrog_driver_steps(name):
    case = make(name)
    O = oracle_for(case)
    oracle_init(O)
    nsteps = 15
    for step in range(1, nsteps + 1):
        oracle_leapfrog_step(O, case.wall, case.dt, step)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.step_leapfrog(case.wall, case.dt, 1, 5)
        ps.step_leapfrog(case.wall, case.dt, 6, nsteps - 5)
        compare(ps, O, history=False)


def test_symplectic_reinit_steps():
    case = make("dam_break_mdbc")
    O = oracle_for(case)
    oracle_init(O)
    O.call("density_to_pressure")
    O.call("compute_forces", 0)
    for step in range(1, 13):
        oracle_step(O, case.wall, case.dt)
        if step % 4 == 0:
            oracle_reinit(O)
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.history_enable()
        ps.density_to_pressure()
        ps.compute_forces(T.FORCES_WCSPH)
        ps.step_symplectic_reinit(case.wall, case.dt, 1, 12, interval=4)
        compare(ps, O)


def test_verlet_graph_equals_eager():
    case = make("dam_break_mdbc")
    out = []
    for eager in (False, True):
        with T.system_for_case(case) as ps:
            ps.particle_to_cell()
            ps.history_enable()
            if eager:
                ps.pair_timing(64)  # eager launches
            ps.step_verlet(case.wall, case.dt, 1, 22)
            out.append((ps.download(), ps.history_download()))
    for k in ("r", "v", "rho", "p", "a", "drho"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    for k in ("v_oo", "rho_oo"):
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


def test_verlet_driver_mixed():
    case = make("dam_break_mdbc")
    O = oracle_for(case)
    oracle_init(O)
    for step in range(1, 23):
        oracle_verlet_step(O, case.wall, case.dt, step)
    with T.system_for_case(case, precision=T.PREC_MIXED) as ps:
        ps.particle_to_cell()
        ps.history_enable()
        ps.step_verlet(case.wall, case.dt, 1, 22)
        g = ps.download(("type", "rho", "r"))
    assert np.array_equal(g["type"], O.types())
    assert rel_l2(g["rho"], O.get(po.F_RHO)) <= 1e-4
    assert rel_l2(g["r"], O.get(po.F_R)) <= 1e-6


def isolated_movers(n=256, seed=5):
    """n fluid particles more than 2h apart (no pair terms: a = -g, dRho = 0),
    moving at vx = 1 m/s, at x in [1, 8) m inside a tall, wide domain."""
    rng = np.random.default_rng(seed)
    dp = 0.01
    h = 1.4 * dp
    xs = 1.0 + 7.0 * rng.random(n)
    ys = 40.0 + 0.1 * np.arange(n)  # 0.1 m > 2h apart in y
    arrays = dict(type=np.zeros(n, np.int32), p=np.zeros(n), rho=np.full(n, 1000.0),
                  r=np.stack([xs, ys], 1), v=np.tile([1.0, 0.0], (n, 1)))
    params = dict(h=h, kap=2.0, cs=20.0, rho0=1000., avisc=0.01, delta=0.1, dp=dp,
                  x_0=0.0, x_m=12.0, y_0=-5.0, y_m=70.0)
    return params, arrays


def test_compensated_corrector_keeps_position_at_rounding_level():
    """Kahan-compensated corrector (tsph_set_compensated) against the exact
    sum: particles moving at a constant vx = 1 m/s for 20000 symplectic steps
    advance by exactly (v + v_o) dt / 2 = dt per step, so x_n = x_0 + n dt in
    exact arithmetic (evaluated here with fractions).  The plain FP64 corrector
    (src/SPH_integration.h:287-309) rounds every step's sum and drifts; the
    compensated one stays within one ulp of the exact value."""
    from fractions import Fraction
    params, arrays = isolated_movers()
    dt, steps = 1.234567e-4, 20000
    out = {}
    for comp in (False, True):
        with T.ParticleSystem(params, precision=T.PREC_FP64, capacity=len(arrays["type"])) as ps:
            ps.upload(arrays)
            if comp:
                ps.set_compensated(True)
            ps.particle_to_cell()
            ps.step(T.WALL_DBC, dt, steps)
            got = ps.download(("type", "r", "v"))
        assert np.array_equal(got["type"], arrays["type"])
        assert np.all(got["v"][:, 0] == 1.0)
        out[comp] = got["r"][:, 0]
    fdt = Fraction(dt)
    exact = np.array([float(Fraction(float(x0)) + steps * fdt) for x0 in arrays["r"][:, 0]])
    ulp = np.spacing(exact)
    err_plain = np.abs(out[False] - exact) / ulp
    err_comp = np.abs(out[True] - exact) / ulp
    print(f"[kahan] {steps} steps: plain mean {err_plain.mean():.1f} ulp (max {err_plain.max():.0f}), "
          f"compensated mean {err_comp.mean():.2f} ulp (max {err_comp.max():.0f})", flush=True)
    assert err_comp.max() <= 1.0
    assert err_plain.mean() >= 10.0 * max(err_comp.mean(), 0.05)


def test_compensated_inflow_matches_oracle():
    """The compensated corrector on the inflow slab (appends, conversions,
    erases) stays within the FP64 bounds of the oracle's uncompensated run."""
    from tests.helpers import oracle_add_zones, oracle_inflow_step
    case = T.build_case(T.CASE_INFLOW_SLAB, dx=0.01)
    steps = 150
    print(f"[case] inflow slab dx=0.01 np={case.np} steps={steps} compensated", flush=True)
    O = oracle_for(case)
    oracle_add_zones(O, case.zones)
    oracle_init(O)
    for _ in range(steps):
        oracle_inflow_step(O, case.wall, case.dt)
    with T.system_for_case(case) as ps:
        ps.set_compensated(True)
        ps.particle_to_cell()
        ps.step_inflow(case.wall, case.dt, steps)
        g = ps.download(("type", "r", "v", "rho", "p"))
        assert ps.counters()["zone_rejected"] == 0
    assert np.array_equal(g["type"], O.types())
    assert rel_err(g["r"], O.get(po.F_R)) <= 1e-9
    assert rel_err(g["v"], O.get(po.F_V)) <= 1e-9
    assert rel_err(g["rho"], O.get(po.F_RHO)) <= 1e-9
    assert rel_err(g["p"], O.get(po.F_P)) <= 1e-6
