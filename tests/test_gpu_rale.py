"""GPU parity of the Riemann-ALE-SPH path (src/RALESPH_interaction.h,
src/RALESPH_integration.h) against the oracle on the same inputs.

Bounds (FP64): the device pow() in the Tait EOS differs from glibc's by a few
ulp, so FP64 fields carry rel. error <= 1e-9 (1e-6 on p, where
p = B((rho/rho0)^7 - 1) cancels).  The mixed path evaluates the pair fluxes in
FP32 and is held to rel. L2 <= 2e-4 on the flux derivatives of one pass.
"""
import math

import numpy as np
import pytest

import tinysph_old_amd as T
from oracle import pyoracle as po
from tests.helpers import oracle_for, rel_err, rel_l2

pytestmark = pytest.mark.gpu

DX = 0.02


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def ale_case(dx=DX):
    print(f"[case] ale_dam_break dx={dx}", flush=True)
    return T.build_case(T.CASE_ALE_DAM_BREAK, dx=dx)


def oracle_rale(case):
    O = oracle_for(case)
    O.call("init_ale")
    O.call("divide_to_cells")
    assert O.call("particle_to_cell") == 0
    return O


def oracle_rale_step(O, step, dt):
    """cases/ALESPH_dambreakLong/SPH_simulationINT.h:84-132"""
    if step > 1:
        O.call("rale_leapfrog_one", dt)
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("rale_forces")
    O.call("rale_leapfrog_two", dt, step)
    O.call("remove_out_of_domain")


def compare(ps, O, tol=1e-9, tol_p=1e-6):
    g = ps.download(("type", "r", "v", "p", "rho"))
    ga = ps.ale_download()
    assert np.array_equal(g["type"], O.types())
    assert rel_err(g["r"], O.get(po.F_R)) <= tol
    assert rel_err(g["v"], O.get(po.F_V)) <= tol
    assert rel_err(g["rho"], O.get(po.F_RHO)) <= tol
    assert rel_err(g["p"], O.get(po.F_P)) <= tol_p
    names = dict(zip(T._lib.ALE_NAMES, po.ALE_FIELDS))
    for k, f in names.items():
        err = rel_err(ga[k], O.get(f))
        assert err <= tol, (k, err)


def test_rale_init_matches_oracle():
    case = ale_case()
    O = oracle_rale(case)
    with T.system_for_case(case) as ps:
        ps.ale_enable()
        ga = ps.ale_download()
    for k, f in zip(T._lib.ALE_NAMES[:6], po.ALE_FIELDS[:6]):
        assert np.array_equal(ga[k], O.get(f)), k


def test_rale_forces_fp64():
    case = ale_case()
    O = oracle_rale(case)
    O.call("density_to_pressure")
    O.call("rale_forces")
    with T.system_for_case(case) as ps:
        ps.ale_enable()
        ps.particle_to_cell()
        ps.density_to_pressure()
        ps.rale_forces()
        ga = ps.ale_download()
    for k, f in (("domega", po.F_DOMEGA), ("domegarho", po.F_DOMEGARHO), ("domegav", po.F_DOMEGAV)):
        err = rel_err(ga[k], O.get(f))
        assert err <= 1e-9, (k, err)


def test_rale_steps_fp64():
    """20 leap-frog steps (the first is the special step == 1 update)."""
    case = ale_case()
    O = oracle_rale(case)
    for s in range(1, 21):
        oracle_rale_step(O, s, case.dt)
    with T.system_for_case(case) as ps:
        ps.ale_enable()
        ps.step_rale(case.dt, 1, 20)
        compare(ps, O)


def test_rale_graph_equals_stages():
    """Graph-replayed steps equal the same steps launched stage by stage."""
    case = ale_case()
    with T.system_for_case(case) as a, T.system_for_case(case) as b:
        a.ale_enable()
        b.ale_enable()
        a.step_rale(case.dt, 1, 8)
        for s in range(1, 9):
            if s > 1:
                b.rale_leapfrog_one(case.dt)
            b.particle_to_cell()
            b.density_to_pressure()
            b.rale_forces()
            b.rale_leapfrog_two(case.dt, s)
            b.remove_out_of_domain()
        fa, fb = a.download(("r", "v", "rho", "p")), b.download(("r", "v", "rho", "p"))
        aa, ab = a.ale_download(), b.ale_download()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    for k in aa:
        assert np.array_equal(aa[k], ab[k]), k


def test_rale_forces_mixed():
    case = ale_case(0.01)
    O = oracle_rale(case)
    O.call("density_to_pressure")
    O.call("rale_forces")
    with T.system_for_case(case, precision=T.PREC_MIXED) as ps:
        ps.ale_enable()
        ps.particle_to_cell()
        ps.density_to_pressure()
        ps.rale_forces()
        ga = ps.ale_download()
    for k, f in (("domega", po.F_DOMEGA), ("domegarho", po.F_DOMEGARHO), ("domegav", po.F_DOMEGAV)):
        err = rel_l2(ga[k], O.get(f))
        print(f"[mixed] {k} rel L2 {err:.3e}", flush=True)
        assert err <= 2e-4, (k, err)


def test_rale_isolated_particle_branch():
    """gamma_sum is not reset between the particles of a cell
    (src/RALESPH_interaction.h:48,257,277-280): an isolated fluid particle gets
    -g*omega*rho when an earlier particle of its cell had neighbours, and zero
    when it is the first of its cell."""
    dp = 0.02
    h, kap = math.sqrt(2 * dp * dp), 2.1
    kh = kap * h
    prm = dict(h=h, kap=kap, cs=34.3, rho0=1000., avisc=0.02, delta=0.1, dp=dp,
               x_0=0.0, x_m=1.0, y_0=0.0, y_m=1.0)
    of = 0.273 * h
    cx0, cy0 = 5 * kh - of, 5 * kh - of
    pts = np.array([
        (cx0 + 0.05 * kh, cy0 + 0.05 * kh, 0.1, 0.0),
        (cx0 - 0.3 * h, cy0 + 0.05 * kh, 0.0, 0.2),
        (cx0 + 0.97 * kh, cy0 + 0.97 * kh, 0.3, -0.1),
        (cx0 + 3.5 * kh, cy0 + 3.5 * kh, -0.2, 0.1),
    ])
    arrays = dict(type=np.zeros(4, np.int32), p=np.zeros(4), rho=np.full(4, 1001.0),
                  r=pts[:, :2].copy(), v=pts[:, 2:].copy())
    q = [prm[k] for k in ("h", "kap", "cs", "rho0", "avisc", "delta", "dp", "x_0", "x_m", "y_0", "y_m")]
    O = po.oracle_case(q, arrays)
    O.call("init_ale")
    O.call("divide_to_cells")
    assert O.call("particle_to_cell") == 0
    O.call("density_to_pressure")
    O.call("rale_forces")
    for prec in (T.PREC_FP64, T.PREC_MIXED):
        with T.ParticleSystem(prm, precision=prec, capacity=4) as ps:
            ps.upload(arrays)
            ps.ale_enable()
            ps.particle_to_cell()
            ps.density_to_pressure()
            ps.rale_forces()
            dov = ps.ale_download()["domegav"]
        w = dp * dp
        assert dov[2][1] == -9.81 * w * 1001.0
        assert dov[3][1] == 0.0
        if prec == T.PREC_FP64:
            assert rel_err(dov, O.get(po.F_DOMEGAV)) <= 1e-9
