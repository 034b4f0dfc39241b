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
