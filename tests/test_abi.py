"""The C-ABI boundary without a GPU: the library loads, exports every entry
point that include/tinysph.h declares, builds the case presets exactly as the
reference's geometry routines do, and the C++ drop-in header compiles."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import tinysph_old_amd as T
from tinysph_old_amd import _lib as L
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tinysph.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tsph_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = C.CDLL(L.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) <= set(L.EXPORTS)


def test_abi_version():
    assert L.lib().tsph_abi_version() == 1


def test_invalid_parameters_rejected_without_device_work():
    prm = L.Params()
    prm.h = -1.0
    prm.capacity = 10
    ctx = C.c_void_p()
    assert L.lib().tsph_create(C.byref(prm), C.byref(ctx)) == L.EINVAL
    assert not ctx
    assert b"invalid" in L.lib().tsph_last_error()


def _has_gpu() -> bool:
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_context_creation_fails_loudly_without_device():
    case = T.build_case(T.CASE_STILL_WATER_DBC, dx=0.05)
    with pytest.raises(T.TsphError):
        T.system_for_case(case)


def _ref_case(variant, q):
    return po.Case(po.ref_lib(variant), "ref_", q)


needs_ref = pytest.mark.skipif(not po.ref_available("dbc"), reason="oracle/_ref not built")


@needs_ref
def test_still_water_preset_matches_reference_geometry():
    case = T.build_case(T.CASE_STILL_WATER_DBC)
    q = case.oracle_params()
    dp, W = case.params["dp"], 1.0
    R = _ref_case("dbc", q)
    R.call("create_wall_x_mdbc", 0. + dp, W - dp, 0., -1, dp / 2)
    R.call("create_wall_y_mdbc", 0., 0 + dp, 1., -1, dp / 2)
    R.call("create_wall_y_mdbc", W, 0 + dp, 1., +1, W - dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0. + dp / 2, 0. + dp / 2)
    R.call("create_corner_mdbc", +1, -1, W - dp / 2, 0. + dp / 2)
    R.call("create_block_of_fluid_wp", 0 + dp, 0 + dp, W - dp, 0.5)
    assert case.np == R.np == 5865
    assert (case.n_fluid, case.n_wall) == (4950, 915)
    assert np.array_equal(case.arrays["type"], R.types())
    for f, k in ((po.F_R, "r"), (po.F_RHO, "rho"), (po.F_P, "p"), (po.F_V, "v")):
        assert np.array_equal(case.arrays[k], R.get(f)), k
    g = R.get(po.F_GNR)
    assert np.array_equal(case.arrays["gnr"][: len(g)], g)
    # the reference mass m = rho0 / (1/dp * 1/dp)
    assert R.call("mass") == po.oracle_lib().orc_mass(po.oracle_case(q, case.arrays).h)


@needs_ref
def test_dam_break_preset_matches_reference_geometry():
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.005, jitter=0.0)
    q = case.oracle_params()
    dp, L_, Ht = case.params["dp"], 3.22, 1.8
    R = _ref_case("wcsph", q)
    R.call("create_wall_x_mdbc", 0. + dp, L_, 0., -1, dp / 2)
    R.call("create_wall_y_mdbc", 0., 0 + dp, Ht, -1, dp / 2)
    R.call("create_wall_y_mdbc", L_, 0 + dp, Ht, +1, L_ - dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0. + dp / 2, 0. + dp / 2)
    R.call("create_corner_mdbc", +1, -1, L_ - dp / 2, 0. + dp / 2)
    R.call("create_block_of_fluid", 0 + dp, 0 + dp, 1.2, 0.6)
    assert np.array_equal(case.arrays["type"], R.types())
    assert np.array_equal(case.arrays["r"], R.get(po.F_R))
    assert np.array_equal(case.arrays["rho"], R.get(po.F_RHO))
    g = R.get(po.F_GNR)
    assert np.array_equal(case.arrays["gnr"][: len(g)], g)


@needs_ref
def test_bt_preset_matches_reference_geometry():
    case = T.build_case(T.CASE_DAM_BREAK_BT)
    q = case.oracle_params()
    dp = case.params["dp"]
    R = _ref_case("bt", q)
    s2 = float(np.sqrt(2))
    R.call("create_wall_x_bt", 0. + dp, 1.61 - dp, 0., -1, 0., 1.)
    R.call("create_wall_y_bt", 0., 0. + dp, 1., -1, 1., 0.)
    R.call("create_wall_y_bt", 1.61, 0. + dp, 1., +1, -1., 0.)
    R.call("create_corner_bt", -1, -1, 0., 0., s2, s2)
    R.call("create_corner_bt", +1, -1, 1.61, 0., -s2, s2)
    R.call("create_block_of_fluid_wp", 0.0 + dp, 0 + dp, 0.6, 0.3)
    assert np.array_equal(case.arrays["r"], R.get(po.F_R))
    assert np.array_equal(case.arrays["rho"], R.get(po.F_RHO))
    nrm = R.get(po.F_N)
    assert np.array_equal(case.arrays["n"][: len(nrm)], nrm)


def test_dam_break_jitter_is_seeded_and_bounded():
    a = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01, seed=5)
    b = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01, seed=5)
    c = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.0)
    assert np.array_equal(a.arrays["r"], b.arrays["r"])
    fl = a.arrays["type"] == 0
    d = np.abs(a.arrays["r"][fl] - c.arrays["r"][fl])
    assert 0 < d.max() <= 0.01 * 0.01
    assert np.array_equal(a.arrays["r"][~fl], c.arrays["r"][~fl])


def test_dropin_header_compiles(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text('''
#define TSPH_DROPIN_FORCES TSPH_FORCES_WCSPH
#include "tinysph_dropin.hpp"
struct SPH_simulation {
  const double dp = 0.02, hh = 1.0 * sqrt(2 * dp * dp), kap = 2.1, dt = 5e-5;
  Particle_system particles{hh, kap, 34.3, 1000., 0.02, 0.1, dp};
  Simulation_data simulation_data{0., 2., 0., 1.};
  void RUN() {
    mDBC_compute_density_mdbcGeo(particles, simulation_data);
    Density_to_pressure(particles);
    Integrate_SymplecticPredictor(particles, dt);
    Clear_cells(particles);
    Particle_To_Cell(particles, simulation_data);
    Density_to_pressure(particles);
    Compute_Forces(particles);
    Integrate_SymplecticCorrector(particles, dt);
    RemoveParticlesOutOfDomain(particles, simulation_data);
    (void)Water_Elevation(particles, simulation_data, 0.3, 0.);
    (void)Pressure_in_point(particles, simulation_data, {1.6, 0.01});
  }
};
int main() { SPH_simulation s; Create_Block_of_Fluid(s.particles, 0.02, 0.02, 0.6, 0.3); return 0; }
''')
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_dropin_header_compiles_rale(tmp_path):
    """A Riemann-ALE driver in the shape of cases/ALESPH_dambreakLong/SPH_simulationINT.h
    compiles against the drop-in header."""
    src = tmp_path / "driver_rale.cpp"
    src.write_text('''
#define TSPH_DROPIN_RALE
#include "tinysph_dropin.hpp"
struct SPH_simulation {
  const double dp = 0.02, hh = 1.0 * sqrt(2 * dp * dp), kap = 2.1, dt = 1.6e-4;
  Particle_system particles{hh, kap, 34.3, 1000., 0.02, 0.1, dp};
  Simulation_data simulation_data{-0.1, 1.7, -0.1, 1.1};
  void RUN(int steps) {
    Divide_To_Cells(particles, simulation_data);
    for (int step = 1; step < steps; step++) {
      if (step > 1) Integrate_LeapFrog_partOne(particles, dt);
      Clear_cells(particles);
      Particle_To_Cell(particles, simulation_data);
      Density_to_pressure(particles);
      Compute_Forces(particles);
      Integrate_LeapFrog_partTwo(particles, dt, step);
      RemoveParticlesOutOfDomain(particles, simulation_data);
    }
    tsph_sync_host(particles);
  }
};
int main() {
  SPH_simulation s;
  Create_Wall_x_bt(s.particles, 0. + 0.02, 1.61 - 0.02, 0., -1, {0., 1.});
  Create_Block_of_Fluid(s.particles, 0.02, 0.02, 0.6, 0.3);
  return s.particles.special.omega.size() == s.particles.np ? 0 : 1;
}
''')
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


# header struct <-> ctypes mirror: every field at the same offset, same size
STRUCTS = {"tsph_params": "Params", "tsph_particles": "Particles", "tsph_counters": "Counters",
           "tsph_history": "History", "tsph_zone": "Zone", "tsph_case": "CaseC",
           "tsph_ale_fields": "AleFields"}


def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "tinysph.h"',
             "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(L, pyname)
        lines.append(f'  printf("{pyname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{pyname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            cls_name, key, value = line.split()
            got[(cls_name, key)] = int(value)
    for pyname in STRUCTS.values():
        cls = getattr(L, pyname)
        assert got[(pyname, "size")] == C.sizeof(cls), pyname
        for field, _ in cls._fields_:
            assert got[(pyname, field)] == getattr(cls, field).offset, (pyname, field)
