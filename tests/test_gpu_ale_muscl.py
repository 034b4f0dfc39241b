"""GPU parity of the ALE-SPH MUSCL path (Compute_Density, Kernel_gradient_approx,
Integrate_density_compute_pressure, Compute_Acceleration_BT and the ALESPH
leap-frog halves) against the oracle on the same inputs.

Bounds: FP64 fields <= 1e-9 relative after 20 steps (the device sqrt in the
pair sound speed against glibc's pow(x, 0.5) is the only arithmetic that may
differ); the mixed path (FP32 fluxes) <= 1e-5 relative L2 on the flux
derivatives and gradients of one pass.
"""
import numpy as np
import pytest

import tinysph_old_amd as T
from oracle import pyoracle as po
from tests.helpers import oracle_ale, oracle_ale_step, oracle_for, rel_err, rel_l2

pytestmark = pytest.mark.gpu

DX = 0.02
MODES = {T.MUSCL_NONE: "none", T.MUSCL_LINEAR: "linear", T.MUSCL_MINMOD: "minmod"}


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def ale_case(dx=DX):
    print(f"[case] ale_dam_break dx={dx}", flush=True)
    return T.build_case(T.CASE_ALE_DAM_BREAK, dx=dx)


ALE_CMP = {"omega": po.F_OMEGA, "omegarho": po.F_OMEGARHO, "omegav": po.F_OMEGAV,
           "domega": po.F_DOMEGA, "domegarho": po.F_DOMEGARHO, "domegav": po.F_DOMEGAV,
           "gradrho": po.F_GRADRHO, "gradvx": po.F_GRADVX, "gradvy": po.F_GRADVY}


def compare(ps, O, tol=1e-9):
    g = ps.download(("type", "r", "v", "p", "rho"))
    ga = ps.ale_download()
    assert np.array_equal(g["type"], O.types())
    for k, f in (("r", po.F_R), ("v", po.F_V), ("rho", po.F_RHO), ("p", po.F_P)):
        err = rel_err(g[k], O.get(f))
        assert err <= tol, (k, err)
    for k, f in ALE_CMP.items():
        err = rel_err(ga[k], O.get(f))
        assert err <= tol, (k, err)


@pytest.mark.parametrize("muscl", list(MODES), ids=list(MODES.values()))
def test_ale_steps_fp64(muscl):
    case = ale_case()
    O = oracle_ale(case)
    for s in range(1, 21):
        oracle_ale_step(O, s, case.dt, muscl)
    with T.system_for_case(case) as ps:
        ps.ale_enable()
        ps.step_ale(case.dt, 1, 20, muscl)
        compare(ps, O)


def test_ale_graph_equals_stages():
    """Graph-replayed MUSCL steps equal the same steps launched stage by stage."""
    case = ale_case()
    with T.system_for_case(case) as a, T.system_for_case(case) as b:
        a.ale_enable()
        b.ale_enable()
        a.step_ale(case.dt, 1, 8, T.MUSCL_MINMOD)
        for s in range(1, 9):
            if s > 1:
                b.ale_leapfrog_one(case.dt)
            b.particle_to_cell()
            b.ale_compute_density(limited=True)
            b.ale_kernel_gradient()
            b.ale_density_pressure(case.dt)
            b.ale_compute_acceleration(limited=True)
            b.ale_leapfrog_two(case.dt, s)
            b.remove_out_of_domain()
        fa, fb = a.download(("r", "v", "rho", "p")), b.download(("r", "v", "rho", "p"))
        aa, ab = a.ale_download(), b.ale_download()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    for k in aa:
        assert np.array_equal(aa[k], ab[k]), k


def passes_state(dx):
    """A mid-run state (gradients non-zero) from 30 oracle MUSCL steps."""
    case = ale_case(dx)
    O = oracle_ale(case)
    for s in range(1, 31):
        oracle_ale_step(O, s, case.dt, T.MUSCL_LINEAR)
    arrays = dict(type=O.types(), p=O.get(po.F_P), rho=O.get(po.F_RHO), r=O.get(po.F_R),
                  v=O.get(po.F_V), n=O.get(po.F_N))
    ale = {k: O.get(f) for k, f in ALE_CMP.items()}
    return case, O, arrays, ale


@pytest.mark.parametrize("precision,tol", [(T.PREC_FP64, 1e-9), (T.PREC_MIXED, 1e-5)],
                         ids=["fp64", "mixed"])
def test_ale_passes_on_a_running_state(precision, tol):
    """Each pass on the same mid-run state (non-zero gradients, moving fluid)."""
    case, O, arrays, ale = passes_state(0.01)
    O.call("clear_cells")
    assert O.call("particle_to_cell") == 0
    with T.ParticleSystem(case.params, precision=precision, capacity=len(arrays["type"])) as ps:
        ps.upload(arrays)
        ps.ale_upload(ale)
        ps.particle_to_cell()
        for limited in (False, True):
            sfx = "_limited" if limited else ""
            O.call("ale_compute_density" + sfx)
            O.call("ale_compute_acceleration" + sfx)
            ps.ale_compute_density(limited)
            ps.ale_compute_acceleration(limited)
            ga = ps.ale_download()
            for k in ("domega", "domegarho", "domegav"):
                ref = O.get(ALE_CMP[k])
                err = rel_err(ga[k], ref) if precision == T.PREC_FP64 else rel_l2(ga[k], ref)
                print(f"[pass] limited={limited} {k} err {err:.3e}", flush=True)
                assert err <= tol, (limited, k, err)
        O.call("ale_kernel_gradient")
        ps.ale_kernel_gradient()
        ga = ps.ale_download()
        for k in ("gradrho", "gradvx", "gradvy"):
            ref = O.get(ALE_CMP[k])
            err = rel_err(ga[k], ref) if precision == T.PREC_FP64 else rel_l2(ga[k], ref)
            print(f"[pass] {k} err {err:.3e}", flush=True)
            assert err <= tol, (k, err)
