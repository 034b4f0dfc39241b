"""The oracle's other WCSPH integrators and density reinitialisation, pinned:

* against the reference's own functions compiled in place (oracle/_ref), bit for
  bit, over whole driver steps: the Verlet driver (Euler on steps 1 and 20), the
  leap-frog driver, and symplectic steps with the reinitialisation;
* against closed forms: constant-acceleration kinematics of the Verlet/Euler
  pair (exact for that scheme) and of the staggered leap-frog, and the
  reinitialisation of a uniform-density lattice (rhoTemp = rho) and of an
  isolated particle (rhoTemp kept)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import (oracle_init, oracle_leapfrog_step, oracle_reinit, oracle_step,
                           oracle_verlet_step)
from tests.test_oracle_ref import (FIELDS, assert_same, dam_break_bt, dam_break_mdbc,
                                   oracle_from_ref, still_water)

needs_ref = pytest.mark.skipif(
    not (po.ref_available("dbc") and po.ref_available("wcsph") and po.ref_available("bt")),
    reason="oracle/_ref not built (needs the reference tree; `make -C oracle ref`)")

HIST = FIELDS + po.HISTORY_FIELDS
WMAP = {"dbc": 0, "mdbc": 1, "bt": 2}
BUILDERS = [(still_water, "dbc", None), (dam_break_mdbc, "mdbc", po.F_GNR), (dam_break_bt, "bt", po.F_N)]


def ref_verlet_step(R, wall, dt, step):
    """cases/*/SPH_simulationVerlet.h, through the reference's functions"""
    if wall == "mdbc":
        R.call("wall_density")
    elif wall == "bt":
        R.call("wall_density"); R.call("density_to_pressure")
    R.call("clear_cells"); R.call("particle_to_cell"); R.call("density_to_pressure")
    R.call("compute_forces")
    R.call("verlet" if step > 1 and step % 20 != 0 else "euler", dt)
    R.call("remove_out_of_domain")


def ref_leapfrog_step(R, wall, dt, step):
    """cases/*/SPH_simulationLeapFrog.h, through the reference's functions"""
    if step > 1:
        R.call("leapfrog_one", dt)
    R.call("clear_cells"); R.call("particle_to_cell")
    if wall in ("mdbc", "bt"):
        R.call("wall_density")
    R.call("density_to_pressure")
    R.call("compute_forces")
    R.call("leapfrog_two", dt, step)
    R.call("remove_out_of_domain")


def start(builder, extra):
    R, q, dt = builder()
    O = oracle_from_ref(R, q, extra)
    for c in (R, O):
        c.call("divide_to_cells")
        c.call("particle_to_cell")
    return R, O, dt


@needs_ref
@pytest.mark.parametrize("builder,wall,extra", BUILDERS, ids=[b[1] for b in BUILDERS])
def test_verlet_driver_bit_exact(builder, wall, extra):
    R, O, dt = start(builder, extra)
    assert_same(R, O, HIST)
    for step in range(1, 23):  # Euler at steps 1 and 20
        ref_verlet_step(R, wall, dt, step)
        oracle_verlet_step(O, WMAP[wall], dt, step)
    assert_same(R, O, HIST)


@needs_ref
@pytest.mark.parametrize("builder,wall,extra", BUILDERS, ids=[b[1] for b in BUILDERS])
def test_leapfrog_driver_bit_exact(builder, wall, extra):
    R, O, dt = start(builder, extra)
    for step in range(1, 16):
        ref_leapfrog_step(R, wall, dt, step)
        oracle_leapfrog_step(O, WMAP[wall], dt, step)
    assert_same(R, O, HIST)


@needs_ref
def test_symplectic_with_reinit_bit_exact():
    R, O, dt = start(dam_break_mdbc, po.F_GNR)
    R.call("density_to_pressure"); R.call("compute_forces")
    O.call("density_to_pressure"); O.call("compute_forces", 0)
    from tests.test_oracle_ref import ref_step
    for step in range(1, 13):
        ref_step(R, "mdbc", dt)
        oracle_step(O, WMAP["mdbc"], dt)
        if step % 4 == 0:  # the driver's interval is 20; 4 exercises it three times here
            R.call("reinit_sum"); R.call("reinit_overwrite")
            oracle_reinit(O)
    assert_same(R, O, HIST)


def free_particles(n=5, seed=3):
    """n fluid particles far apart (no pair interactions) in a large domain"""
    rng = np.random.default_rng(seed)
    q = [0.01, 2.0, 10.0, 1000., 0.01, 0.1, 0.005, -10., 10., -10., 10.]
    r = np.stack([np.linspace(-5, 5, n), np.zeros(n)], axis=1)
    v = rng.uniform(-1, 1, (n, 2))
    O = po.oracle_case(q, dict(type=np.zeros(n, np.int32), p=np.zeros(n), rho=np.full(n, 1000.),
                               r=r, v=v))
    a = rng.uniform(-3, 3, (n, 2))
    drho = rng.uniform(-5, 5, n)
    O.set(po.F_A, a)
    O.set(po.F_DRHO, drho)
    return O, r, v, a, drho


def test_verlet_constant_acceleration_is_exact():
    """Euler then Verlet under a constant a is exact kinematics:
    r_n = r0 + v0 t + a t^2 / 2, v_n = v0 + a t, rho_n = rho0 + drho t (t = n dt)."""
    O, r0, v0, a, drho = free_particles()
    dt = 1e-3
    for step in range(1, 26):
        O.call("verlet" if step > 1 and step % 20 != 0 else "euler", dt)
        t = step * dt
        np.testing.assert_allclose(O.get(po.F_V), v0 + a * t, rtol=0, atol=1e-12)
        np.testing.assert_allclose(O.get(po.F_R), r0 + v0 * t + 0.5 * a * t * t, rtol=0, atol=1e-12)
        np.testing.assert_allclose(O.get(po.F_RHO), 1000. + drho * t, rtol=0, atol=1e-10)


def test_leapfrog_constant_acceleration_closed_form():
    """The drivers' leap-frog with a constant a: after step n,
    v_n = v0 + a dt (n - 1/2), r_n = r0 + v0 dt (n - 1/2) + a dt^2 (n(n-1)/2 + 1/4)
    (the first step advances v by half a step and r by a quarter-step drift)."""
    O, r0, v0, a, drho = free_particles()
    dt = 1e-3
    for n in range(1, 16):
        if n > 1:
            O.call("leapfrog_one", dt)
        O.call("leapfrog_two", dt, n)
        np.testing.assert_allclose(O.get(po.F_V), v0 + a * dt * (n - 0.5), rtol=0, atol=1e-12)
        np.testing.assert_allclose(O.get(po.F_R), r0 + v0 * dt * (n - 0.5) +
                                   a * dt * dt * (n * (n - 1) / 2 + 0.25), rtol=0, atol=1e-12)


def test_reinit_uniform_lattice_and_isolated_particle():
    """On a lattice of equal density, rhoTemp = rho (the Shepard ratio of a
    constant); an isolated fluid particle keeps its previous rhoTemp."""
    dp, h = 0.01, 0.015
    q = [h, 2.0, 20.0, 1000., 0.01, 0.1, dp, -0.1, 1.2, -0.1, 1.2]
    xs, ys = np.meshgrid(np.arange(0, 0.3, dp), np.arange(0, 0.3, dp))
    r = np.concatenate([np.stack([xs.ravel(), ys.ravel()], 1), [[1.0, 1.0]]])
    n = len(r)
    rho = np.full(n, 1003.25)
    rho[-1] = 1017.0  # the isolated one (out of [990, 1010]: counted)
    O = po.oracle_case(q, dict(type=np.zeros(n, np.int32), p=np.zeros(n), rho=rho, r=r,
                               v=np.zeros((n, 2))))
    oracle_init(O)
    rt0 = O.get(po.F_RHOTMP).copy()
    np.testing.assert_array_equal(rt0, rho)  # Add_particle: rhoTemp = rho
    O.set(po.F_RHOTMP, np.where(np.arange(n) == n - 1, 1001.5, rho))
    bad = oracle_reinit(O)
    out = O.get(po.F_RHO)
    np.testing.assert_allclose(out[:-1], 1003.25, rtol=1e-13)
    assert out[-1] == 1001.5
    assert bad == 1
