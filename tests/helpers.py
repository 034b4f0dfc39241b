"""Shared test helpers: drive the oracle through the same stage sequence as the
reference case drivers, and compare device results with it."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po

WALL_DBC, WALL_MDBC, WALL_BT = 0, 1, 2


def oracle_for(case, threads: int = 0):
    return po.oracle_case(case.oracle_params(), case.arrays, omp_threads=threads)


def oracle_variant(wall: int) -> int:
    return {WALL_DBC: 1, WALL_MDBC: 0, WALL_BT: 2}[wall]


def oracle_init(O) -> None:
    """INIT: Divide_To_Cells + Particle_To_Cell (cases/*/SPH_simulationSymplectic.h INIT)."""
    O.call("divide_to_cells")
    rc = O.call("particle_to_cell")
    assert rc == 0, "a particle lies outside the cell grid"


def oracle_step(O, wall: int, dt: float) -> None:
    """One symplectic step in the drivers' order
    (cases/stillWater_DBC/SPH_simulationSymplectic.h:141-198,
    cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:59-116,
    cases/damBreakLong_BT/SPH_simulationSymplectic.h:66-122)."""
    var = oracle_variant(wall)
    if wall == WALL_MDBC:
        O.call("mdbc_density")
        O.call("density_to_pressure")
    elif wall == WALL_BT:
        O.call("bt_density")
        O.call("density_to_pressure")
    O.call("predictor", dt)
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("compute_forces", var)
    O.call("corrector", dt)
    O.call("remove_out_of_domain")
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("compute_forces", var)


def oracle_fields(O) -> dict:
    return {
        "type": O.types(),
        "r": O.get(po.F_R), "v": O.get(po.F_V), "a": O.get(po.F_A),
        "p": O.get(po.F_P), "rho": O.get(po.F_RHO), "drho": O.get(po.F_DRHO),
    }


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max |b| (scale-aware elementwise error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def brute_force_pairs(r: np.ndarray, active: np.ndarray, h: float) -> int:
    """Ordered pairs (i active, j != i) with |r_i - r_j| < 2h, by brute force."""
    n = len(r)
    total = 0
    for i in np.nonzero(active)[0]:
        d = np.sqrt(((r - r[i]) ** 2).sum(axis=1))
        m = d / h < 2.0
        m[i] = False
        total += int(m.sum())
    return total


def oracle_add_zones(O, zones) -> None:
    """Register buffer zones (dicts as CaseData.zones) with an oracle case."""
    for i, z in enumerate(zones):
        k = O.lib.orc_add_zone(O.h, z["x0"], z["xm"], z["y0"], z["ym"], z["vx"], z["vy"],
                               float(z["ox"]), float(z["oy"]), int(z["btype"]))
        assert k == i


def oracle_inflow_step(O, wall: int, dt: float, inlet: int = 0, outlet: int = 1) -> None:
    """One step of the inflow driver (cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:57-123)."""
    var = oracle_variant(wall)
    if wall == WALL_MDBC:
        O.call("mdbc_density")
        O.call("density_to_pressure")
    elif wall == WALL_BT:
        O.call("bt_density")
        O.call("density_to_pressure")
    O.call("predictor", dt)
    if inlet >= 0:
        O.call("bz_integrate_inlet", inlet, dt)
    if outlet >= 0:
        O.call("bz_integrate_outlet", outlet, dt)
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("compute_forces", var)
    O.call("corrector", dt)
    O.call("remove_out_of_domain")
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("compute_forces", var)


def oracle_verlet_step(O, wall: int, dt: float, step: int) -> None:
    """One Verlet-driver step (cases/damBreakLong_mDBC/SPH_simulationVerlet.h:64-116;
    the BT variant adds an EOS after the wall density, the DBC variant has none):
    Euler on step 1 and every 20th step, Verlet otherwise."""
    if wall == WALL_MDBC:
        O.call("mdbc_density")
    elif wall == WALL_BT:
        O.call("bt_density")
        O.call("density_to_pressure")
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("compute_forces", oracle_variant(wall))
    O.call("verlet" if step > 1 and step % 20 != 0 else "euler", dt)
    O.call("remove_out_of_domain")


def oracle_leapfrog_step(O, wall: int, dt: float, step: int) -> None:
    """One leap-frog-driver step (cases/damBreakLong_mDBC/SPH_simulationLeapFrog.h,
    cases/stillWater_DBC/SPH_simulationLeapFrog.h)."""
    if step > 1:
        O.call("leapfrog_one", dt)
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    if wall == WALL_MDBC:
        O.call("mdbc_density")
    elif wall == WALL_BT:
        O.call("bt_density")
    O.call("density_to_pressure")
    O.call("compute_forces", oracle_variant(wall))
    O.call("leapfrog_two", dt, step)
    O.call("remove_out_of_domain")


def oracle_reinit(O) -> int:
    """Density_reinitialization_sumation + _overwrite; returns the out-of-range count."""
    O.call("reinit_sum")
    return O.call("reinit_overwrite")


def oracle_ale(case, threads: int = 0):
    """Oracle ALE case: ALESPH_Add_particle's volume variables, the grid and the
    first cell list (cases/ALESPH_dambreakLong/SPH_simulation.h INIT)."""
    O = oracle_for(case, threads)
    O.call("init_ale")
    O.call("divide_to_cells")
    assert O.call("particle_to_cell") == 0
    return O


def oracle_ale_step(O, step, dt, muscl):
    """cases/ALESPH_dambreakLong/SPH_simulation.h:78-125 (gradients after
    Compute_Density where the driver has the commented call, :103).
    muscl: 0 none, 1 linear, 2 minmod (TSPH_MUSCL_*)."""
    sfx = "_limited" if muscl == 2 else ""
    if step > 1:
        O.call("ale_leapfrog_one", dt)
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("ale_compute_density" + sfx)
    if muscl:
        O.call("ale_kernel_gradient")
    O.call("ale_density_pressure", dt)
    O.call("ale_compute_acceleration" + sfx)
    O.call("ale_leapfrog_two", dt, step)
    O.call("remove_out_of_domain")
