"""GPU side of the output path and of snapshot/resume:

* tsph_write_vtk queues its copies at the call and writes while later steps
  run: the file equals the reference-format text (tsph_format_vtk, pinned to
  write_to_ASCII_VTK in tests/test_output.py) of the state at the call;
* tsph_interpolate_grid (GenerateInterpol's grid on the device) against the
  oracle's Pressure_in_point at every grid point, within 1e-9 relative (1e-6 on
  p), the probes' bound; tsph_write_interpolated_vtk writes those values;
* a run saved with tsph_save_state and resumed with tsph_load_state continues
  bit-identically (symplectic steps, and Verlet-driver steps with the history
  fields)."""
import ctypes as C

import numpy as np
import pytest

import tinysph_old_amd as T
from tinysph_old_amd import _lib as L
from oracle import pyoracle as po
from tests.helpers import oracle_for, oracle_init, oracle_step, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def case():
    return T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01, seed=7)


def format_state(ps, path):
    g = ps.download(("type", "r", "v", "p", "rho"))
    h = L.Particles()
    h.np = len(g["type"])
    t = np.ascontiguousarray(g["type"], dtype=np.int32)
    arr = {k: np.ascontiguousarray(g[k]).reshape(-1) for k in ("r", "v", "p", "rho")}
    h.type = L.iptr(t)
    for k, a in arr.items():
        setattr(h, k, L.dptr(a))
    assert L.lib().tsph_format_vtk(C.byref(h), str(path).encode()) == 0


def test_async_vtk_is_the_state_at_the_call(tmp_path):
    c = case()
    with T.system_for_case(c) as ps:
        ps.particle_to_cell()
        ps.step(c.wall, c.dt, 5)
        format_state(ps, tmp_path / "sync.vtk")
        ps.write_vtk(tmp_path / "async.vtk")
        ps.step(c.wall, c.dt, 5)  # runs while the file is written
        ps.io_wait()
        format_state(ps, tmp_path / "later.vtk")
    a = (tmp_path / "sync.vtk").read_bytes()
    assert a == (tmp_path / "async.vtk").read_bytes()
    assert len(a) > 100000
    assert a != (tmp_path / "later.vtk").read_bytes()  # the state moved on after the call


def test_interpolated_grid_matches_oracle(tmp_path):
    c = case()
    O = oracle_for(c)
    oracle_init(O)
    for _ in range(4):
        oracle_step(O, c.wall, c.dt)
    x0, y0, xm, ym = 0.05, 0.0, 0.9, 0.5
    with T.system_for_case(c) as ps:
        ps.particle_to_cell()
        ps.step(c.wall, c.dt, 4)
        rho, p, v = ps.interpolate_grid(x0, y0, xm, ym)
        ps.write_interpolated_vtk(x0, y0, xm, ym, tmp_path / "grid.vtk")
    h = c.params["h"]
    ncx, ncy = int((xm - x0) / h) * 4, int((ym - y0) / h) * 4
    assert rho.shape == (ncx, ncy)
    dh = h / 4
    ref = np.zeros((ncx, ncy, 4))
    for i in range(ncx):
        for j in range(ncy):
            ref[i, j] = O.interpolate_point(x0 + i * dh, y0 + j * dh)
    assert np.count_nonzero(ref[..., 0]) > 100
    assert rel_err(rho, ref[..., 0]) <= 1e-9
    assert rel_err(p, ref[..., 1]) <= 1e-6
    assert rel_err(v, ref[..., 2:]) <= 1e-9
    # the file holds the same values, in the reference writer's format
    ours = tmp_path / "fmt.vtk"
    assert L.lib().tsph_format_interpolated_vtk(
        ncx, ncy, dh, x0, y0, L.dptr(np.ascontiguousarray(rho).reshape(-1)),
        L.dptr(np.ascontiguousarray(p).reshape(-1)), L.dptr(np.ascontiguousarray(v).reshape(-1)),
        str(ours).encode()) == 0
    assert ours.read_bytes() == (tmp_path / "grid.vtk").read_bytes()


FIELDS = ("type", "index", "p", "rho", "drho", "r", "v", "a", "r_o", "v_o", "rho_o")


@pytest.mark.parametrize("driver", ["symplectic", "verlet"])
def test_snapshot_resume_is_bit_identical(tmp_path, driver):
    c = case()
    snap = tmp_path / "state.bin"

    def run(ps, first, n):
        if driver == "symplectic":
            ps.step(c.wall, c.dt, n)
        else:
            ps.step_verlet(c.wall, c.dt, first, n)

    with T.system_for_case(c) as ps:
        ps.particle_to_cell()
        if driver == "verlet":
            ps.history_enable()
        run(ps, 1, 10)
        ps.save_state(snap)
        run(ps, 11, 12)
        straight = ps.download(FIELDS)
        hist = ps.history_download() if driver == "verlet" else None
    with T.system_for_case(c) as ps:
        ps.load_state(snap)
        ps.particle_to_cell()
        run(ps, 11, 12)
        resumed = ps.download(FIELDS)
        for k in FIELDS:
            assert np.array_equal(straight[k], resumed[k]), k
        if hist is not None:
            h2 = ps.history_download()
            for k in hist:
                assert np.array_equal(hist[k], h2[k]), k
