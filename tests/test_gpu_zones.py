"""GPU parity of the buffer zones (BZ_Integrate_inlet / BZ_Integrate_outlet,
src/SPH_buffer_zones.h) against the oracle on the same inputs.

Integer results (types, reference indices, particle counts, the order of
appended particles) are compared exactly; FP64 fields within 1e-9 relative
(1e-6 on p), as for the other FP64 parity tests.  The inflow channel at its
full size (4M particles) is checked through size-independent properties:
particle bookkeeping (live = initial + appended - removed), no refused or
chained appends, and the inflow rate against its closed form.
"""
import math

import numpy as np
import pytest

import tinysph_old_amd as T
from oracle import pyoracle as po
from tests.helpers import (oracle_add_zones, oracle_for, oracle_inflow_step, oracle_init, rel_err)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def compare(ps, O, tol=1e-9, tol_p=1e-6):
    g = ps.download(("type", "index", "r", "v", "rho", "p"))
    assert np.array_equal(g["type"], O.types())
    assert np.array_equal(g["index"], O.index())
    assert rel_err(g["r"], O.get(po.F_R)) <= tol
    assert rel_err(g["v"], O.get(po.F_V)) <= tol
    assert rel_err(g["rho"], O.get(po.F_RHO)) <= tol
    assert rel_err(g["p"], O.get(po.F_P)) <= tol_p


@pytest.mark.parametrize("preset,dx,steps", [(T.CASE_INFLOW_SLAB, 0.01, 150),
                                             (T.CASE_INFLOW_CHANNEL, 0.01, 120)],
                         ids=["slab", "channel"])
def test_inflow_steps_fp64(preset, dx, steps):
    """Whole inflow-driver steps: appends in parent order, conversions, erases."""
    print(f"[case] preset={preset} dx={dx} steps={steps}", flush=True)
    case = T.build_case(preset, dx=dx)
    O = oracle_for(case)
    oracle_add_zones(O, case.zones)
    oracle_init(O)
    for _ in range(steps):
        oracle_inflow_step(O, case.wall, case.dt)
    assert O.np > case.np  # the inlet appended particles
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.step_inflow(case.wall, case.dt, steps)
        c = ps.counters()
        assert c["zone_rejected"] == 0 and c["zone_chained"] == 0
        assert ps.np == O.np
        compare(ps, O)


def outlet_scene(orientation, dp=0.01):
    """Zone particles in runs of 1, 2 and 3 removal candidates, each followed by
    a fluid particle past the zone start, plus candidates below dp/2."""
    s = orientation
    beyond, inside, fluid_past = 1.0 + s * 0.1, 1.0 + s * 0.02, 1.0 + s * 0.01
    rows = [(12, beyond, 0.2), (0, fluid_past, 0.2),
            (12, beyond, 0.2), (12, beyond, 0.25), (0, fluid_past, 0.25),
            (12, beyond, 0.3), (12, inside, 0.001), (12, beyond, 0.35), (0, fluid_past, 0.3),
            (12, inside, 0.2), (0, fluid_past, 0.35), (0, 1.0 - s * 0.2, 0.2),
            (12, inside, 0.004), (0, fluid_past, 0.4)]
    n = len(rows)
    arrays = dict(type=np.array([t for t, _, _ in rows], np.int32), p=np.zeros(n),
                  rho=np.full(n, 1000.0), r=np.array([[x, y] for _, x, y in rows]),
                  v=np.tile([0.3 * s, 0.0], (n, 1)))
    params = dict(h=dp * 1.4, kap=2.1, cs=40.0, rho0=1000., avisc=0.1, delta=0.1, dp=dp,
                  x_0=-0.5, x_m=2.5, y_0=-0.5, y_m=1.5)
    zone = dict(x0=1.0, xm=1.0 + 4 * dp * s, y0=dp, ym=0.5, vx=0.5, vy=0.0, ox=s, oy=0, btype=12)
    return params, arrays, zone


@pytest.mark.parametrize("orientation", [+1, -1])
def test_outlet_in_loop_erase(orientation):
    params, arrays, zone = outlet_scene(orientation)
    q = [params[k] for k in ("h", "kap", "cs", "rho0", "avisc", "delta", "dp", "x_0", "x_m", "y_0", "y_m")]
    O = po.oracle_case(q, arrays)
    oracle_add_zones(O, [zone])
    removed = O.call("bz_integrate_outlet", 0, 1e-4)
    assert removed == 5
    with T.ParticleSystem(params, capacity=len(arrays["type"])) as ps:
        ps.upload(arrays)
        ps.bz_add_zone(zone)
        ps.bz_integrate_outlet(0, 1e-4)
        ps.particle_to_cell()  # compacts the erased particles
        assert ps.counters()["removed"] == 5
        compare(ps, O)


def inlet_scene(axis, orientation, dp=0.01, shift=0.0):
    """Four layers of zone particles as DRAW_CreateInletY / DRAW_CreateInletX
    lay them out (src/SPH_create.h:439-503), moving out of the zone."""
    o = orientation
    n_along = int(math.ceil(0.1 / dp)) + 1
    pts = []
    for l in range(4):
        for h in range(n_along):
            if axis == "y":
                pts.append((1.0 + dp * l * o + shift, 0.2 + h * dp))
            else:
                pts.append((0.2 + h * dp, 1.0 + dp * l * o + shift))
    n = len(pts)
    vel = (-0.9 * o, 0.0) if axis == "y" else (0.0, -0.9 * o)
    arrays = dict(type=np.full(n, 7, np.int32), p=np.zeros(n), rho=np.full(n, 1000.0),
                  r=np.array(pts), v=np.tile(vel, (n, 1)))
    params = dict(h=dp * 1.4, kap=2.1, cs=40.0, rho0=1000., avisc=0.1, delta=0.1, dp=dp,
                  x_0=-0.5, x_m=2.5, y_0=-0.5, y_m=1.5)
    if axis == "y":
        zone = dict(x0=1.0, xm=1.0 + 4 * dp * o, y0=0.2, ym=0.3, vx=vel[0], vy=0.0, ox=o, oy=0, btype=7)
    else:
        zone = dict(x0=0.2, xm=0.3, y0=1.0, ym=1.0 + 4 * dp * o, vx=0.0, vy=vel[1], ox=0, oy=o, btype=7)
    return params, arrays, zone


@pytest.mark.parametrize("axis,orientation", [("y", -1), ("y", +1), ("x", -1), ("x", +1)])
def test_inlet_branches(axis, orientation):
    """All four inlet branches, three passes each (conversions, appends in
    parent order, the in-loop advance of the new particles)."""
    params, arrays, zone = inlet_scene(axis, orientation,
                                       shift=0.015 if (axis, orientation) == ("y", -1) else 0.0)
    q = [params[k] for k in ("h", "kap", "cs", "rho0", "avisc", "delta", "dp", "x_0", "x_m", "y_0", "y_m")]
    O = po.oracle_case(q, arrays)
    oracle_add_zones(O, [zone])
    n0 = len(arrays["type"])
    with T.ParticleSystem(params, capacity=4 * n0) as ps:
        ps.upload(arrays)
        ps.bz_add_zone(zone)
        for dt in (0.004, 0.008, 0.02):
            O.call("bz_integrate_inlet", 0, dt)
            ps.bz_integrate_inlet(0, dt)
            compare(ps, O, tol=0.0, tol_p=0.0)
        assert ps.np > n0 and ps.counters()["appended"] == ps.np - n0


def test_inlet_refuses_beyond_capacity():
    """Appends that do not fit the context are refused as a whole and counted."""
    params, arrays, zone = inlet_scene("y", +1)
    n0 = len(arrays["type"])
    with T.ParticleSystem(params, capacity=n0) as ps:
        ps.upload(arrays)
        ps.bz_add_zone(zone)
        ps.bz_integrate_inlet(0, 0.004)
        c = ps.counters()
        assert ps.np == n0 and c["appended"] == 0 and c["zone_rejected"] > 0


def test_inflow_channel_full_size():
    """The 8 x 0.5 m channel at dx = 0.5 mm (about 4M particles), mixed
    precision, 300 graph-replayed inflow steps: every appended particle is
    accounted for, nothing is refused, and the particles fed through the inlet
    match the closed form U t H / dx^2 to within one buffer column."""
    case = T.build_case(T.CASE_INFLOW_CHANNEL)
    print(f"[case] inflow channel np={case.np}", flush=True)
    assert 3.9e6 < case.np < 4.1e6
    steps = 300
    with T.system_for_case(case, precision=T.PREC_MIXED) as ps:
        ps.particle_to_cell()
        ps.step_inflow(case.wall, case.dt, steps)
        c = ps.counters()
        np_end = ps.np
    print(f"[channel] {c}", flush=True)
    assert c["zone_rejected"] == 0 and c["zone_chained"] == 0
    assert np_end == case.np + c["appended"] - c["removed"]
    dx = case.params["dp"]
    column = int(math.ceil((0.125 - dx) / dx)) + 1       # zone particles per layer
    columns_fed = 1.0 * steps * case.dt / dx              # U t / dx
    assert abs(c["appended"] / column - columns_fed) <= 1.0 + 1e-9


def test_renumbered_positions_leave_inflow_run_unchanged():
    """Relabelling the reference positions mid-run (what the library does
    before an inlet would run out of them) changes no result: a slab run with
    two relabels equals the run without, bit for bit, and the oracle."""
    case = T.build_case(T.CASE_INFLOW_SLAB, dx=0.01)
    steps = 150
    print(f"[case] inflow slab dx=0.01 np={case.np} steps={steps}, relabelled at 50 and 100", flush=True)
    O = oracle_for(case)
    oracle_add_zones(O, case.zones)
    oracle_init(O)
    for _ in range(steps):
        oracle_inflow_step(O, case.wall, case.dt)
    outs = []
    for relabel in (False, True):
        with T.system_for_case(case) as ps:
            ps.particle_to_cell()
            for _ in range(3):
                ps.step_inflow(case.wall, case.dt, steps // 3)
                if relabel:
                    ps.renumber()
            compare(ps, O)
            outs.append(ps.download(("type", "index", "r", "v", "rho", "p")))
            assert ps.counters()["zone_rejected"] == 0
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
