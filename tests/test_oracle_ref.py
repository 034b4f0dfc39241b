"""Pins the oracle to the reference: the reference's own stage functions,
compiled from its headers in place (oracle/_ref, built by `make -C oracle ref`),
against the oracle restatement on identical inputs, bit for bit.

The reference code is single-threaded here (built without -fopenmp: its kernel
returns a static buffer, src/SPH_kernel.h:54).  Cases are built by the
reference's own geometry routines (src/SPH_create.h, src/SPH_create_bt.h)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import oracle_step

pytestmark = pytest.mark.skipif(
    not (po.ref_available("dbc") and po.ref_available("wcsph") and po.ref_available("bt")),
    reason="oracle/_ref not built (needs the reference tree; `make -C oracle ref`)")

FIELDS = (po.F_R, po.F_V, po.F_A, po.F_P, po.F_RHO, po.F_DRHO, po.F_RO, po.F_VO, po.F_RHOO)


def still_water(variant="dbc", dp=0.02, W=1.0):
    """cases/stillWater_DBC/geometry.h with a W-wide tank"""
    h, kap = 1.5 * dp, 2.0
    cs = 10 * math.sqrt(9.81 * 0.5)
    nvl = math.ceil(kap * h / dp)
    q = [h, kap, cs, 1000., 0.01, 0.1, dp, -nvl * dp, W + 0.2 + nvl * dp, -nvl * dp, 1.0 + nvl * dp]
    R = po.Case(po.ref_lib(variant), "ref_", q)
    R.call("create_wall_x_mdbc", 0. + dp, W - dp, 0., -1, dp / 2)
    R.call("create_wall_y_mdbc", 0., 0 + dp, 1., -1, dp / 2)
    R.call("create_wall_y_mdbc", W, 0 + dp, 1., +1, W - dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0. + dp / 2, 0. + dp / 2)
    R.call("create_corner_mdbc", +1, -1, W - dp / 2, 0. + dp / 2)
    R.call("create_block_of_fluid_wp", 0 + dp, 0 + dp, W - dp, 0.5)
    return R, q, 0.2 * h / cs


def dam_break_mdbc(dp=0.02, L=1.61):
    """cases/damBreakLong_mDBC/geometry.h"""
    h, kap, cs = math.sqrt(2 * dp * dp), 2.1, 34.3
    nvl = math.ceil(kap * h / dp)
    q = [h, kap, cs, 1000., 0.02, 0.1, dp, -nvl * dp, 2. + nvl * dp, -nvl * dp, 1. + nvl * dp]
    R = po.Case(po.ref_lib("wcsph"), "ref_", q)
    R.call("create_wall_x_mdbc", 0. + dp, L, 0., -1, dp / 2)
    R.call("create_wall_y_mdbc", 0., 0 + dp, 1., -1, dp / 2)
    R.call("create_wall_y_mdbc", L, 0 + dp, 1., +1, L - dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0. + dp / 2, 0. + dp / 2)
    R.call("create_corner_mdbc", +1, -1, L - dp / 2, 0. + dp / 2)
    R.call("create_block_of_fluid", 0 + dp, 0 + dp, 0.6, 0.3)
    return R, q, 0.00005 * dp / 0.005


def dam_break_bt(dp=0.02):
    """cases/damBreakLong_BT/geometry.h"""
    h, kap, cs = math.sqrt(2 * dp * dp), 2.1, 34.3
    nvl = math.ceil(kap * h / dp)
    q = [h, kap, cs, 1000., 0.02, 0.1, dp, -nvl * dp, 2. + nvl * dp, -nvl * dp, 1. + nvl * dp]
    R = po.Case(po.ref_lib("bt"), "ref_", q)
    R.call("create_wall_x_bt", 0. + dp, 1.61 - dp, 0., -1, 0., 1.)
    R.call("create_wall_y_bt", 0., 0. + dp, 1., -1, 1., 0.)
    R.call("create_wall_y_bt", 1.61, 0. + dp, 1., +1, -1., 0.)
    R.call("create_corner_bt", -1, -1, 0., 0., math.sqrt(2), math.sqrt(2))
    R.call("create_corner_bt", +1, -1, 1.61, 0., -math.sqrt(2), math.sqrt(2))
    R.call("create_block_of_fluid_wp", 0.0 + dp, 0 + dp, 0.6, 0.3)
    return R, q, 0.00004 * dp / 0.005


def oracle_from_ref(R, q, extra=None):
    """Oracle case holding exactly the reference's particles."""
    parts = dict(type=R.types(), p=R.get(po.F_P), rho=R.get(po.F_RHO), r=R.get(po.F_R),
                 v=R.get(po.F_V))
    n = len(parts["type"])
    if extra is not None:
        e = R.get(extra)
        full = np.zeros((n, 2))
        full[: len(e)] = e  # the reference stores wall extras for the walls, created first
        parts["gnr" if extra == po.F_GNR else "n"] = full
    return po.oracle_case(q, parts)


def ref_step(R, wall, dt):
    """One step in the driver order, through the reference's functions."""
    if wall in ("mdbc", "bt"):
        R.call("wall_density")
        R.call("density_to_pressure")
    R.call("predictor", dt)
    R.call("clear_cells"); R.call("particle_to_cell"); R.call("density_to_pressure")
    R.call("compute_forces")
    R.call("corrector", dt)
    R.call("remove_out_of_domain")
    R.call("clear_cells"); R.call("particle_to_cell"); R.call("density_to_pressure")
    R.call("compute_forces")


def assert_same(R, O, fields=FIELDS):
    assert R.np == O.np
    assert np.array_equal(R.types(), O.types())
    for f in fields:
        a, b = R.get(f), O.get(f)
        assert a.shape == b.shape, f
        assert np.array_equal(a, b), (f, float(np.max(np.abs(a - b))))


@pytest.mark.parametrize("builder,wall,extra,ovar", [
    (still_water, "dbc", None, 1),
    (dam_break_mdbc, "mdbc", po.F_GNR, 0),
    (dam_break_bt, "bt", po.F_N, 2),
])
def test_steps_bit_exact(builder, wall, extra, ovar):
    R, q, dt = builder()
    O = oracle_from_ref(R, q, extra)
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
    assert R.grid() == O.grid()
    s1, i1 = R.cell_csr()
    s2, i2 = O.cell_csr()
    assert np.array_equal(s1, s2) and np.array_equal(i1, i2)
    # initial forces (as after INIT), then 15 steps
    R.call("density_to_pressure"); R.call("compute_forces")
    O.call("density_to_pressure"); O.call("compute_forces", ovar)
    assert_same(R, O)
    wmap = {"dbc": 0, "mdbc": 1, "bt": 2}
    for _ in range(15):
        ref_step(R, wall, dt)
        oracle_step(O, wmap[wall], dt)
    assert_same(R, O)


def dam_break_rale(dp=0.02):
    """cases/ALESPH_dambreakLong/geometry.h with the ALE creation routines
    (src/ALESPH_create.h: ALESPH_Add_particle sets omega = dp^2, ...)"""
    h, kap, cs = math.sqrt(2 * dp * dp), 2.1, 34.3
    nvl = math.ceil(kap * h / dp)
    q = [h, kap, cs, 1000., 0.02, 0.1, dp, -nvl * dp, 2. + nvl * dp, -nvl * dp, 1. + nvl * dp]
    R = po.Case(po.ref_lib("rale"), "ref_", q)
    R.call("create_wall_x_bt", 0. + dp, 1.61 - dp, 0., -1, 0., 1.)
    R.call("create_wall_y_bt", 0., 0. + dp, 1., -1, 1., 0.)
    R.call("create_wall_y_bt", 1.61, 0. + dp, 1., +1, -1., 0.)
    R.call("create_corner_bt", -1, -1, 0., 0., math.sqrt(2), math.sqrt(2))
    R.call("create_corner_bt", +1, -1, 1.61, 0., -math.sqrt(2), math.sqrt(2))
    R.call("create_block_of_fluid", 0.0 + dp, 0 + dp, 0.6, 0.3)
    return R, q, 0.00004 * dp / 0.005


def rale_step(C, step, dt, oracle):
    """cases/ALESPH_dambreakLong/SPH_simulationINT.h:84-132"""
    if step > 1:
        C.call("rale_leapfrog_one", dt)
    C.call("clear_cells")
    C.call("particle_to_cell")
    C.call("density_to_pressure")
    C.call("rale_forces" if oracle else "compute_forces")
    C.call("rale_leapfrog_two", dt, step)
    C.call("remove_out_of_domain")


@pytest.mark.skipif(not po.ref_available("rale"), reason="oracle/_ref/ref_rale.so not built")
def test_rale_steps_bit_exact():
    R, q, dt = dam_break_rale()
    O = oracle_from_ref(R, q, po.F_N)
    O.call("init_ale")
    for f in po.ALE_FIELDS:
        assert np.array_equal(R.get(f), O.get(f)), f
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
    for step in range(1, 16):
        rale_step(R, step, dt, False)
        rale_step(O, step, dt, True)
    assert_same(R, O, (po.F_R, po.F_V, po.F_P, po.F_RHO) + po.ALE_FIELDS)


def ale_step(C, step, dt, oracle, gradients=False):
    """cases/ALESPH_dambreakLong/SPH_simulation.h:78-125; with gradients=True the
    commented-out Kernel_gradient_approx call at :103 (after Compute_Density) is
    made."""
    if step > 1:
        C.call("ale_leapfrog_one", dt)
    C.call("clear_cells")
    C.call("particle_to_cell")
    C.call("ale_compute_density")
    if gradients:
        C.call("ale_kernel_gradient")
    C.call("ale_density_pressure", dt)
    C.call("ale_compute_acceleration")
    C.call("ale_leapfrog_two", dt, step)
    C.call("remove_out_of_domain")


ALE_ALL = (po.F_R, po.F_V, po.F_P, po.F_RHO) + po.ALE_FIELDS + po.GRAD_FIELDS


@pytest.mark.skipif(not po.ref_available("rale"), reason="oracle/_ref/ref_rale.so not built")
@pytest.mark.parametrize("gradients", [False, True])
def test_ale_muscl_steps_bit_exact(gradients):
    """ALE-SPH (Compute_Density / Compute_Acceleration_BT, leap-frog halves and
    Integrate_density_compute_pressure), first order as the live driver runs
    it, and with the MUSCL gradients of Kernel_gradient_approx."""
    R, q, dt = dam_break_rale()
    O = oracle_from_ref(R, q, po.F_N)
    O.call("init_ale")
    for f in ALE_ALL:
        assert np.array_equal(R.get(f), O.get(f)), f
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
    for step in range(1, 16):
        ale_step(R, step, dt, False, gradients)
        ale_step(O, step, dt, True, gradients)
    assert_same(R, O, ALE_ALL)
    if gradients:
        assert np.abs(O.get(po.F_GRADRHO)).max() > 0


@pytest.mark.skipif(not po.ref_available("rale"), reason="oracle/_ref/ref_rale.so not built")
def test_ale_density_pass_fluid_volume_flux_is_zero():
    """Compute_Density reads a never-assigned `dvdW` for fluid neighbours
    (src/ALESPH_density_and_omega_bt.h:135,237).  The reference build evaluates
    that term as zero: with random velocities, interior fluid particles get
    domega == 0 while the RALE pass (which assigns it, src/RALESPH_interaction.h:135)
    does not.  The oracle reproduces the reference build."""
    R, q, dt = dam_break_rale()
    rng = np.random.default_rng(1)
    v = R.get(po.F_V)
    v[:] = rng.normal(0, 0.3, v.shape)
    R.set(po.F_V, v)
    O = oracle_from_ref(R, q, po.F_N)
    O.call("init_ale")
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
        c.call("density_to_pressure")
    R.call("ale_compute_density")
    O.call("ale_compute_density")
    assert np.array_equal(R.get(po.F_DOMEGA), O.get(po.F_DOMEGA))
    assert np.array_equal(R.get(po.F_DOMEGARHO), O.get(po.F_DOMEGARHO))
    r = R.get(po.F_R)
    interior = (R.types() == 0) & (r[:, 0] > 0.2) & (r[:, 0] < 0.4) & (r[:, 1] > 0.1) & (r[:, 1] < 0.2)
    assert interior.sum() > 10
    assert np.all(R.get(po.F_DOMEGA)[interior] == 0)
    R.call("compute_forces")  # RALE pass assigns dv.dW
    assert np.any(R.get(po.F_DOMEGA)[interior] != 0)


@pytest.mark.skipif(not po.ref_available("rale"), reason="oracle/_ref/ref_rale.so not built")
def test_rale_isolated_particle_branch():
    """The reference never resets gamma_sum between the particles of a cell
    (src/RALESPH_interaction.h:48,257,277-280): an isolated fluid particle gets
    -g*omega*rho if an earlier fluid particle of its cell had neighbours, and
    zero if it is the first of its cell.  Both situations, against the reference."""
    dp = 0.02
    h, kap = math.sqrt(2 * dp * dp), 2.1
    kh = kap * h
    q = [h, kap, 34.3, 1000., 0.02, 0.1, dp, 0.0, 1.0, 0.0, 1.0]
    of = 0.273 * h
    # cell (5, 5): lower-left corner x0 = 5 kh - of
    cx0, cy0 = 5 * kh - of, 5 * kh - of
    pts = [
        (cx0 + 0.05 * kh, cy0 + 0.05 * kh, 0.1, 0.0),    # has a neighbour (next entry)
        (cx0 - 0.3 * h, cy0 + 0.05 * kh, 0.0, 0.2),      # neighbour in the cell to the left
        (cx0 + 0.97 * kh, cy0 + 0.97 * kh, 0.3, -0.1),   # isolated, later in the same cell
        (cx0 + 3.5 * kh, cy0 + 3.5 * kh, -0.2, 0.1),     # isolated, first of its cell
    ]
    R = po.Case(po.ref_lib("rale"), "ref_", q)
    for x, y, vx, vy in pts:
        R.call("add_ale", 0, 0.0, 1001.0, x, y, vx, vy)
    O = po.oracle_case(q, dict(type=R.types(), p=R.get(po.F_P), rho=R.get(po.F_RHO),
                               r=R.get(po.F_R), v=R.get(po.F_V)))
    O.call("init_ale")
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
        c.call("density_to_pressure")
    R.call("compute_forces")
    O.call("rale_forces")
    dov = R.get(po.F_DOMEGAV)
    assert np.array_equal(dov, O.get(po.F_DOMEGAV))
    w = dp * dp
    assert dov[2][1] == -9.81 * w * 1001.0  # isolated, after a particle with neighbours
    assert dov[3][1] == 0.0                  # isolated, first of its cell


def test_probes_bit_exact():
    R, q, dt = dam_break_mdbc(dp=0.01)
    O = oracle_from_ref(R, q, po.F_GNR)
    for c in (R, O):
        c.call("divide_to_cells"); c.call("particle_to_cell"); c.call("density_to_pressure")
    for x in (0.1, 0.3, 0.45):
        assert R.call("water_elevation", x, 0.0) == O.call("water_elevation", x, 0.0)
        assert R.call("water_elevation_inverse", x, 0.5) == O.call("water_elevation_inverse", x, 0.5)
        for y in (0.02, 0.1, 0.25):
            assert R.call("sum_kernel_function", x, y) == O.call("sum_kernel_function", x, y)
            assert np.array_equal(R.interpolate_point(x, y), O.interpolate_point(x, y))


def test_removal_bit_exact():
    """A run of consecutive out-of-domain fluid particles: the reference's
    in-loop erase removes every other one."""
    R, q, dt = still_water(dp=0.05)
    r = R.get(po.F_R)
    t = R.types()
    fl = np.nonzero(t == 0)[0]
    for k in (fl[3], fl[4], fl[5], fl[6], fl[9]):
        r[k] = (q[8] + 0.1, 0.3)
    R.set(po.F_R, r)
    O = oracle_from_ref(R, q)
    for c in (R, O):
        c.call("divide_to_cells")
    R.call("remove_out_of_domain")
    removed = O.call("remove_out_of_domain")
    assert removed == 3  # fl[3], fl[5], fl[9]; fl[4] and fl[6] are skipped
    assert R.np == O.np
    assert np.array_equal(R.get(po.F_R), O.get(po.F_R))
    assert np.array_equal(R.types(), O.types())


# ---------------------------------------------------------------------------
# Buffer zones (src/SPH_buffer_zones.h), reference build REF_WCSPH
# ---------------------------------------------------------------------------
INLET1, OUTLET1 = 7, 12  # cases/inflowSlab_mDBC/geometry.h:15-17 (btype values)


def copy_zones(R, O):
    out = np.zeros(12)
    for i in range(R.lib.ref_zone_count(R.h)):
        R.lib.ref_zone_get(R.h, i, out.ctypes.data_as(po._dp))
        x0, xm, y0, ym, _nbl, bl, bh, bt, vx, vy, ox, oy = out
        k = O.lib.orc_add_zone(O.h, x0, xm, y0, ym, vx, vy, ox, oy, int(bt))
        assert k == i


def inflow_slab(dp=0.01):
    """cases/inflowSlab_mDBC/geometry.h with the driver's constants
    (cases/inflowSlab_mDBC/constants.h)"""
    h, kap, cs = math.sqrt(2 * dp * dp), 2.1, 42.48
    nvl = math.ceil(kap * h / dp)
    q = [h, kap, cs, 1000., 0.1, 0.1, dp, -nvl * dp, 2. + nvl * dp, -nvl * dp, 1. + nvl * dp]
    R = po.Case(po.ref_lib("wcsph"), "ref_", q)
    R.call("create_wall_x_mdbc", 0.0, 0.4 - 3 * dp, 0.1, -1, 0.1 + dp / 2)
    R.call("create_wall_y_mdbc", 0.4, 0 + dp, 0.1 - 3 * dp, -1, 0.4 + dp / 2)
    R.call("create_wall_x_mdbc", 0.4 + dp, 1.2, 0., -1, 0 + dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0.4 + dp / 2, 0. + dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0.4 + dp / 2, 0.1 + dp / 2)
    R.call("draw_create_inlet_y", 0.1, 0.1 + dp, 0.2, -1, 4, 0.75, 0., INLET1)
    R.call("draw_create_outlet_y", 1.0, 0.0 + dp, 0.5, +1, 4, 0.5, 0., OUTLET1, 1)
    return R, q, 0.00005 * dp / 0.005


def inflow_step(C, dt, oracle):
    """cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:57-123"""
    C.call("wall_density" if not oracle else "mdbc_density")
    C.call("density_to_pressure")
    C.call("predictor", dt)
    C.call("bz_integrate_inlet", 0, dt)
    C.call("bz_integrate_outlet", 1, dt)
    C.call("clear_cells"); C.call("particle_to_cell"); C.call("density_to_pressure")
    C.call("compute_forces", 0) if oracle else C.call("compute_forces")
    C.call("corrector", dt)
    C.call("remove_out_of_domain")
    C.call("clear_cells"); C.call("particle_to_cell"); C.call("density_to_pressure")
    C.call("compute_forces", 0) if oracle else C.call("compute_forces")


def test_inflow_slab_steps_bit_exact():
    """Inlet conversions and appends (in loop order) inside whole driver steps."""
    R, q, dt = inflow_slab()
    O = oracle_from_ref(R, q, po.F_GNR)
    copy_zones(R, O)
    n0 = R.np
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
    for _ in range(12):
        inflow_step(R, dt, False)
        inflow_step(O, dt, True)
    assert R.np > n0  # the inlet appended particles
    assert_same(R, O, FIELDS)
    assert np.array_equal(R.index(), O.index())


def outlet_scene(orientation):
    """Zone particles in runs of 1, 2 and 3 removal candidates, each followed by
    a fluid particle beyond the zone start (converted only if examined), plus
    candidates below dp/2 and kept zone particles."""
    dp = 0.01
    q = [dp * 1.4, 2.1, 40.0, 1000., 0.1, 0.1, dp, -0.5, 2.5, -0.5, 1.5]
    R = po.Case(po.ref_lib("wcsph"), "ref_", q)
    R.call("draw_create_outlet_y", 1.0, dp, 0.5, orientation, 4, 0.5, 0., OUTLET1, 1)
    s = orientation
    beyond = 1.0 + s * 0.1          # past x_b + b_len (b_len = 0.04)
    inside = 1.0 + s * 0.02         # inside the zone
    fluid_past = 1.0 + s * 0.01     # fluid beyond x_b: converts when examined
    seq = [
        (OUTLET1, beyond, 0.2), (0, fluid_past, 0.2),                           # run of 1
        (OUTLET1, beyond, 0.2), (OUTLET1, beyond, 0.25), (0, fluid_past, 0.25),  # run of 2
        (OUTLET1, beyond, 0.3), (OUTLET1, inside, 0.001), (OUTLET1, beyond, 0.35),
        (0, fluid_past, 0.3),                                                   # run of 3
        (OUTLET1, inside, 0.2), (0, fluid_past, 0.35), (0, 1.0 - s * 0.2, 0.2),
        (OUTLET1, inside, 0.004), (0, fluid_past, 0.4),
    ]
    for t, x, y in seq:
        R.call("add", t, 0.0, 1000.0, x, y, 0.3 * s, 0.0)
    return R, q


@pytest.mark.parametrize("orientation", [+1, -1])
def test_outlet_in_loop_erase_bit_exact(orientation):
    R, q = outlet_scene(orientation)
    O = oracle_from_ref(R, q)
    copy_zones(R, O)
    R.call("bz_integrate_outlet", 0, 1e-4)
    O.call("bz_integrate_outlet", 0, 1e-4)
    assert R.np < 14 and R.np == O.np
    assert_same(R, O, FIELDS)
    assert np.array_equal(R.index(), O.index())


@pytest.mark.parametrize("axis,orientation", [("y", -1), ("y", +1), ("x", -1), ("x", +1)])
def test_inlet_orientations_bit_exact(axis, orientation):
    """All four inlet branches (three of them advance the zone particles inside
    the loop), with particles appended during the pass visited by it."""
    dp = 0.01
    q = [dp * 1.4, 2.1, 40.0, 1000., 0.1, 0.1, dp, -0.5, 2.5, -0.5, 1.5]
    R = po.Case(po.ref_lib("wcsph"), "ref_", q)
    if axis == "y":   # zone normal along x (DRAW_CreateInletY)
        R.call("draw_create_inlet_y", 1.0, 0.2, 0.3, orientation, 4, -0.9 * orientation, 0., INLET1)
    else:             # zone normal along y (DRAW_CreateInletX)
        R.call("draw_create_inlet_x", 1.0, 0.2, 0.3, orientation, 4, 0., -0.9 * orientation, INLET1)
    if (axis, orientation) == ("y", -1):
        # this branch does not advance the zone particles: move them across
        r = R.get(po.F_R)
        r[:, 0] += 0.015
        R.set(po.F_R, r)
        R.set(po.F_RO, r)
    O = oracle_from_ref(R, q)
    copy_zones(R, O)
    n0 = R.np
    for dt in (0.004, 0.008, 0.02):
        R.call("bz_integrate_inlet", 0, dt)
        O.call("bz_integrate_inlet", 0, dt)
        assert_same(R, O, FIELDS)
        assert np.array_equal(R.index(), O.index())
    assert R.np > n0 and (R.types() == 0).sum() == R.np - n0
