"""tsph_interpolate_grid_noalloc (no device allocation; grid values staged in
a per-GPU device array) against tsph_interpolate_grid on the same state: two
calls in a row, the second on a smaller grid, bit-identical values; a grid
beyond TSPH_GRID_SCRATCH_POINTS is refused with TSPH_ECAPACITY."""
import ctypes as C

import numpy as np
import pytest

import tinysph_old_amd as T
from tinysph_old_amd import _lib as L

pytestmark = pytest.mark.gpu


def grid(fn, ctx, box, sizes_only=False):
    nx, ny = C.c_int(-7), C.c_int(-7)
    rc = fn(ctx, *box, C.byref(nx), C.byref(ny), None, None, None)
    if rc != 0 or sizes_only:
        return rc, nx.value, ny.value, None
    n = nx.value * ny.value
    rho, p, v = np.full(n, np.nan), np.full(n, np.nan), np.full(2 * n, np.nan)
    rc = fn(ctx, *box, C.byref(nx), C.byref(ny), L.dptr(rho), L.dptr(p), L.dptr(v))
    return rc, nx.value, ny.value, (rho, p, v)


def test_noalloc_grid_equals_allocating_grid_twice():
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01, seed=3)
    print(f"[case] dam break dx=0.01 np={case.np}", flush=True)
    lib = L.lib()
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        ps.step(case.wall, case.dt, 5)
        # a large grid first, then a smaller one: the second call must not see
        # values the first left in the shared staging array
        for box in ((0.0, 0.0, 1.3, 0.7), (0.2, 0.1, 0.6, 0.4)):
            rc_a, nx_a, ny_a, want = grid(lib.tsph_interpolate_grid, ps.ctx, box)
            rc_b, nx_b, ny_b, got = grid(lib.tsph_interpolate_grid_noalloc, ps.ctx, box)
            print(f"[grid] box={box} {nx_b}x{ny_b}", flush=True)
            assert rc_a == 0 and rc_b == 0, L.lib().tsph_last_error()
            assert (nx_a, ny_a) == (nx_b, ny_b) and nx_b * ny_b > 100
            for a, b in zip(want, got):
                assert np.array_equal(a, b)
            assert np.count_nonzero(got[0]) > 10  # some points lie in the water


def test_noalloc_grid_refuses_oversized_grid():
    case = T.build_case(T.CASE_DAM_BREAK_MDBC, dx=0.01, jitter=0.01, seed=3)
    lib = L.lib()
    with T.system_for_case(case) as ps:
        ps.particle_to_cell()
        h = case.params["h"]
        # (xm - x0) / h * 4 * (ym - y0) / h * 4 points, well beyond the limit
        side = h * (int(np.sqrt(L.GRID_SCRATCH_POINTS) / 4) + 8)
        rc, nx, ny, _ = grid(lib.tsph_interpolate_grid_noalloc, ps.ctx, (0.0, 0.0, side, side),
                             sizes_only=True)
        assert rc == L.ECAPACITY
        assert (nx, ny) == (-7, -7)
        assert b"TSPH_GRID_SCRATCH_POINTS" in lib.tsph_last_error()
