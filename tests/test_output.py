"""The output writers' text against the reference's own writers (compiled in
place, oracle/_ref), byte for byte, without a GPU:

* tsph_format_vtk vs write_to_ASCII_VTK (src/SPH_data_to_vtk.h:4-51) on a dam
  break built by the reference's geometry routines, after a few oracle steps so
  that every field carries non-trivial values;
* tsph_format_interpolated_vtk vs write_INTERPOLATED_mesh_to_ASCII_VTK
  (src/SPH_grid_to_vtk.h:139-243) on GenerateInterpol's grid
  (src/SPH_dataInterpolation.h:13-91) evaluated by the oracle's
  Pressure_in_point interpolation, itself pinned to the reference in
  tests/test_oracle_ref.py."""
import ctypes as C
import os

import numpy as np
import pytest

import tinysph_old_amd as T
from tinysph_old_amd import _lib as L
from oracle import pyoracle as po
from tests.test_oracle_ref import dam_break_mdbc, ref_step

needs_ref = pytest.mark.skipif(not po.ref_available("wcsph"), reason="oracle/_ref not built")


def host_particles(R):
    n = R.np
    arr = dict(type=R.types().astype(np.int32), r=np.ascontiguousarray(R.get(po.F_R)),
               v=np.ascontiguousarray(R.get(po.F_V)), p=np.ascontiguousarray(R.get(po.F_P)),
               rho=np.ascontiguousarray(R.get(po.F_RHO)))
    h = L.Particles()
    h.np = n
    h.type = L.iptr(arr["type"])
    for k in ("r", "v", "p", "rho"):
        setattr(h, k, L.dptr(arr[k]))
    return h, arr


@needs_ref
def test_particle_vtk_matches_reference_writer(tmp_path):
    R, q, dt = dam_break_mdbc(dp=0.02)
    R.call("divide_to_cells"); R.call("particle_to_cell")
    R.call("density_to_pressure"); R.call("compute_forces")
    for _ in range(3):
        ref_step(R, "mdbc", dt)
    ref_file, our_file = tmp_path / "ref.vtk", tmp_path / "ours.vtk"
    lib = R.lib
    lib.ref_write_vtk.argtypes = [C.c_void_p, C.c_char_p]
    lib.ref_write_vtk(R.h, str(ref_file).encode())
    h, keep = host_particles(R)
    assert L.lib().tsph_format_vtk(C.byref(h), str(our_file).encode()) == 0
    a, b = ref_file.read_bytes(), our_file.read_bytes()
    assert len(a) > 10000
    assert a == b


@needs_ref
def test_interpolated_grid_vtk_matches_reference_writer(tmp_path):
    R, q, dt = dam_break_mdbc(dp=0.02)
    O = po.oracle_case(q, dict(type=R.types(), p=R.get(po.F_P), rho=R.get(po.F_RHO),
                               r=R.get(po.F_R), v=R.get(po.F_V)))
    O.call("divide_to_cells"); O.call("particle_to_cell"); O.call("density_to_pressure")
    h = q[0]
    x0, y0, xm, ym = 0.0, 0.0, 0.8, 0.4
    ncx, ncy = int((xm - x0) / h) * 4, int((ym - y0) / h) * 4
    dh = h / 4
    rho, p, v = np.zeros(ncx * ncy), np.zeros(ncx * ncy), np.zeros(2 * ncx * ncy)
    for i in range(ncx):
        for j in range(ncy):
            out = O.interpolate_point(x0 + i * dh, y0 + j * dh)
            t = j + i * ncy
            rho[t], p[t], v[2 * t], v[2 * t + 1] = out
    assert np.count_nonzero(rho) > 100
    ref_file, our_file = tmp_path / "ref.vtk", tmp_path / "ours.vtk"
    lib = R.lib
    lib.ref_write_grid_vtk.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       *([C.POINTER(C.c_double)] * 3), C.c_char_p]
    lib.ref_write_grid_vtk(ncx, ncy, dh, x0, y0, L.dptr(v), L.dptr(rho), L.dptr(p),
                           str(ref_file).encode())
    assert L.lib().tsph_format_interpolated_vtk(ncx, ncy, dh, x0, y0, L.dptr(rho), L.dptr(p),
                                                L.dptr(v), str(our_file).encode()) == 0
    assert ref_file.read_bytes() == our_file.read_bytes()


@needs_ref
def test_grid_vtk_matches_reference_writer(tmp_path):
    """tsph_format_grid_vtk vs write_mesh_to_ASCII_VTK (src/SPH_grid_to_vtk.h:6-67)
    on the reference's own cell counts after a few steps."""
    R, q, dt = dam_break_mdbc(dp=0.02)
    R.call("divide_to_cells"); R.call("particle_to_cell")
    R.call("density_to_pressure"); R.call("compute_forces")
    for _ in range(3):
        ref_step(R, "mdbc", dt)
    ncx, ncy = R.grid()
    start, _ = R.cell_csr()
    counts = np.ascontiguousarray(np.diff(start)[: ncx * ncy], dtype=np.int32)
    assert counts.sum() == R.np
    ref_file, our_file = tmp_path / "ref.vtk", tmp_path / "ours.vtk"
    R.lib.ref_write_mesh_vtk.argtypes = [C.c_void_p, C.c_char_p]
    R.lib.ref_write_mesh_vtk(R.h, str(ref_file).encode())
    h, kap, x0, y0 = q[0], q[1], q[7], q[9]
    assert L.lib().tsph_format_grid_vtk(ncx, ncy, h * kap, h * 0.273, x0, y0,
                                        counts.ctypes.data_as(C.POINTER(C.c_int32)),
                                        str(our_file).encode()) == 0
    a, b = ref_file.read_bytes(), our_file.read_bytes()
    assert len(a) > 10000
    assert a == b


@needs_ref
def test_hydrostatic_pressure_file_matches_reference_writer(tmp_path):
    """tsph_format_hydrostatic_pressure vs Output_hydrostatic_pressure
    (src/SPH_measure_pressure.h:17-34)."""
    R, q, dt = dam_break_mdbc(dp=0.02)
    R.call("divide_to_cells"); R.call("particle_to_cell")
    R.call("density_to_pressure"); R.call("compute_forces")
    for _ in range(3):
        ref_step(R, "mdbc", dt)
    ref_file, our_file = tmp_path / "ref.dat", tmp_path / "ours.dat"
    R.lib.ref_output_hydrostatic_pressure.argtypes = [C.c_void_p, C.c_char_p]
    R.lib.ref_output_hydrostatic_pressure(R.h, str(ref_file).encode())
    h, keep = host_particles(R)
    assert L.lib().tsph_format_hydrostatic_pressure(C.byref(h), str(our_file).encode()) == 0
    a, b = ref_file.read_bytes(), our_file.read_bytes()
    assert a.count(b"\n") == int((R.types() == 0).sum()) > 100
    assert a == b


DROPIN_TEXT = r"""
#include "tinysph_dropin.hpp"
int main(int argc, char **argv) {
  const double h = %r, kap = %r, cs = %r, rho0 = %r, visco = %r, delta = %r, dp = %r;
  Particle_system ps(h, kap, cs, rho0, visco, delta, dp);
  Simulation_data sd(%r, %r, %r, %r);
  sd.hh = h;
  sd.kh = h * kap;
  sd.nvl = %d;
  Output_file(ps.data_const, sd, argv[1]);
  for (int k = 0; k < 5; ++k) {
    Write_water_level(k * 0.0123456789, 0.3 - k * 1e-7, argv[2]);
    Write_pressure(k * 0.0123456789, 1234.56789 * k, argv[3]);
  }
  return 0;
}
"""


@needs_ref
def test_driver_text_files_match_reference_writers(tmp_path):
    """Output_file, Write_water_level and Write_pressure of the drop-in header
    vs src/SPH_output_info.h:8-76, byte for byte."""
    import subprocess
    R, q, dt = dam_break_mdbc(dp=0.005)
    nvl = 3
    src = tmp_path / "text.cpp"
    src.write_text(DROPIN_TEXT % (*q[:7], q[7], q[8], q[9], q[10], nvl))
    exe = tmp_path / "text"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.join(root, "tinysph_old_amd")
    res = subprocess.run(["g++", "-std=c++17", f"-I{os.path.join(root, 'include')}", str(src), "-o",
                          str(exe), f"-L{pkg}", "-ltinysph_hip", f"-Wl,-rpath,{pkg}",
                          "-Wl,-rpath-link,/opt/rocm/lib"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    ours = [tmp_path / n for n in ("info.txt", "wl.dat", "p.dat")]
    res = subprocess.run([str(exe), *map(str, ours)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    assert "Info file generated." in res.stdout
    ref = [tmp_path / n for n in ("ref_info.txt", "ref_wl.dat", "ref_p.dat")]
    lib = R.lib
    lib.ref_output_file.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
    lib.ref_write_water_level.argtypes = [C.c_double, C.c_double, C.c_char_p]
    lib.ref_write_pressure.argtypes = [C.c_double, C.c_double, C.c_char_p]
    lib.ref_output_file(R.h, nvl, str(ref[0]).encode())
    for k in range(5):
        lib.ref_write_water_level(k * 0.0123456789, 0.3 - k * 1e-7, str(ref[1]).encode())
        lib.ref_write_pressure(k * 0.0123456789, 1234.56789 * k, str(ref[2]).encode())
    for a, b in zip(ref, ours):
        assert a.read_bytes() == b.read_bytes(), a.name
