"""The cell-list scan on a grid of more than 4.2M cells, where its blocks
(4096 cells each above 1M cells) outnumber what one look-back round covers
(1024 blocks) and the scan chains inclusive prefixes from block to block
(tsph_cells.hip, k_scan).

The still-water tank (dx = 10 mm, 5,865 particles) is placed in a 64 x 64 m
domain: 2,134 x 2,134 cells of 30 mm, 1,112 scan blocks, almost all cells
empty.  Against the oracle: the cell lists bit-exact, FP64 forces and three
graph-replayed symplectic steps within 1e-9 (the pair loops walk the tile
list the scan writes, so a tile missing from it shows as missing forces);
with the bounded poll (scan wait 2^20) no block stalls, with scan wait -1
every block past the first forms its prefix from the cell counts instead,
with the same results.
"""
import dataclasses

import numpy as np
import pytest

import tinysph_old_amd as T
from tests.helpers import (oracle_fields, oracle_for, oracle_init, oracle_step, oracle_variant,
                           rel_err)

pytestmark = pytest.mark.gpu

CELLS_PER_SCAN_BLOCK = 256 * 16
ONE_ROUND_BLOCKS = 4 * 256


@pytest.fixture(scope="module", autouse=True)
def _oracle(oracle_built):
    return oracle_built


def sparse_case():
    case = T.build_case(T.CASE_STILL_WATER_DBC, dx=0.01)
    prm = dict(case.params)
    prm["x_m"] = prm["x_0"] + 64.0
    prm["y_m"] = prm["y_0"] + 64.0
    return dataclasses.replace(case, params=prm)


@pytest.mark.parametrize("wait", [1 << 20, -1], ids=["polled", "counts"])
def test_chained_scan_on_a_sparse_large_grid(wait):
    case = sparse_case()
    O = oracle_for(case)
    oracle_init(O)
    start_o, items_o = O.cell_csr()
    with T.system_for_case(case) as ps:
        ncx, ncy = ps.grid()
        assert (ncx, ncy) == O.grid()
        blocks = (ncx * ncy + 1 + CELLS_PER_SCAN_BLOCK - 1) // CELLS_PER_SCAN_BLOCK
        print(f"[case] still water in 64 x 64 m: np={case.np} grid {ncx}x{ncy} = {ncx * ncy} "
              f"cells, {blocks} scan blocks, scan wait {wait}", flush=True)
        assert blocks > ONE_ROUND_BLOCKS
        ps.set_scan_wait(wait)
        ps.particle_to_cell()
        start_g, items_g = ps.cell_lists()
        assert np.array_equal(start_g, start_o)
        assert np.array_equal(items_g, items_o)
        # forces on the first cell list, then whole steps
        ps.density_to_pressure()
        ps.compute_forces(T.forces_variant(case.wall))
        O.call("density_to_pressure")
        O.call("compute_forces", oracle_variant(case.wall))
        got = ps.download(("a", "drho"))
        ref = oracle_fields(O)
        assert rel_err(got["a"], ref["a"]) <= 1e-9
        assert rel_err(got["drho"], ref["drho"]) <= 1e-9
        ps.step(case.wall, case.dt, 3)
        for _ in range(3):
            oracle_step(O, case.wall, case.dt)
        got = ps.download(("type", "r", "v", "rho", "p"))
        cnt = ps.counters()
    ref = oracle_fields(O)
    assert np.array_equal(got["type"], ref["type"])
    for k in ("r", "v", "rho"):
        assert rel_err(got[k], ref[k]) <= 1e-9, k
    assert rel_err(got["p"], ref["p"]) <= 1e-6
    if wait < 0:
        assert cnt["scan_stalls"] > 0
    else:
        assert cnt["scan_stalls"] == 0
