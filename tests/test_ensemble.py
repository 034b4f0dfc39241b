"""The batched-ensemble workload (tsph_ensemble_build) without a GPU: seeded and
reproducible, each tank built with the reference's geometry routines
(src/SPH_create.h, as cases/damBreakLong_mDBC/geometry.h uses them), per-tank
constants as stated in include/tinysph.h, and the realised size of the
64-tank workload recorded."""
import numpy as np
import pytest

import tinysph_old_amd as T
from oracle import pyoracle as po

SEED = 20241202


def test_ensemble_is_seeded():
    a = T.build_ensemble(6, 7, 2000, 20000)
    b = T.build_ensemble(6, 7, 2000, 20000)
    c = T.build_ensemble(6, 8, 2000, 20000)
    for x, y in zip(a, b):
        assert np.array_equal(x.arrays["r"], y.arrays["r"])
        assert x.dt == y.dt
    assert any(x.np != y.np for x, y in zip(a, c))


def test_ensemble_tank_constants_and_bounds():
    cases = T.build_ensemble(16, SEED, 2000, 30000)
    for c in cases:
        p = c.params
        assert c.wall == T.WALL_MDBC and c.steps == 500
        assert p["h"] == pytest.approx(np.sqrt(2) * p["dp"], rel=1e-15)
        assert p["kap"] == 2.1 and p["avisc"] == 0.02 and p["delta"] == 0.1
        assert c.dt == pytest.approx(0.2 * p["h"] / p["cs"], rel=1e-15)
        # the lattice count is estimated from the block area: within a few percent
        assert 0.9 * 2000 <= c.n_fluid <= 1.05 * 30000
        r = c.arrays["r"]
        assert r[:, 0].min() > p["x_0"] and r[:, 0].max() < p["x_m"]
        assert r[:, 1].min() > p["y_0"] and r[:, 1].max() < p["y_m"]
        fl = c.arrays["type"] == T.FLUID
        # water depth from c0 = 10 sqrt(g depth)
        depth = (p["cs"] / 10.0) ** 2 / 9.81
        assert r[fl, 1].max() <= depth + 0.02 * p["dp"]


def test_ensemble_workload_size():
    """The 64-tank workload: realised totals for the benchmark seed."""
    cases = T.build_ensemble(64, SEED)
    n = np.array([c.np for c in cases])
    nf = np.array([c.n_fluid for c in cases])
    assert len(cases) == 64
    assert int(n.sum()) == 7835545
    assert int(nf.sum()) == 7517719
    assert nf.min() >= 0.9 * 2000 and nf.max() <= 1.05 * 400000
    kinds = [bool(np.ptp(c.arrays["r"][c.arrays["type"] == T.FLUID, 0]) <
                  0.6 * np.ptp(c.arrays["r"][:, 0])) for c in cases]
    assert 20 <= sum(kinds) <= 44   # dam-break columns vs still water, p = 1/2 each


def splitmix64(seed):
    """SplitMix64 (Steele, Lea, Flood 2014), as the builder draws from it."""
    M = (1 << 64) - 1
    s = seed & M
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield z ^ (z >> 31)


def tank_draws(ncases, seed, nmin, nmax):
    """(W, depth, dam, dx) per tank, from the documented draw order."""
    g = splitmix64(seed)
    u = lambda: (next(g) >> 11) * (1.0 / 9007199254740992.0)  # noqa: E731
    out = []
    for _ in range(ncases):
        W = 0.5 + 2.5 * u()
        fill = 0.3 + 0.5 * u()
        dx = 0.002 + 0.008 * u()
        dam = u() < 0.5
        next(g)  # jitter seed
        depth = fill * 2.0 * W
        area = (0.5 * W if dam else W - dx) * depth
        est = area / (dx * dx)
        if est > nmax:
            dx = np.sqrt(area / nmax)
        elif est < nmin:
            dx = np.sqrt(area / nmin)
        out.append((W, depth, dam, dx))
    return out


def test_ensemble_draws_follow_the_stated_distribution():
    draws = tank_draws(64, SEED, 2000, 400000)
    cases = T.build_ensemble(64, SEED)
    for (W, depth, dam, dx), c in zip(draws, cases):
        assert c.params["dp"] == dx
        assert c.params["cs"] == pytest.approx(10.0 * np.sqrt(9.81 * depth), rel=1e-15)


needs_ref = pytest.mark.skipif(not po.ref_available("wcsph"), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("q", [0, 1, 2, 3])
def test_ensemble_tank_matches_reference_geometry(q):
    case = T.build_ensemble(4, SEED, 2000, 20000)[q]
    W, depth, dam, dx = tank_draws(4, SEED, 2000, 20000)[q]
    dp = case.params["dp"]
    assert dp == dx
    Ht = 2.0 * W
    R = po.Case(po.ref_lib("wcsph"), "ref_", case.oracle_params())
    R.call("create_wall_x_mdbc", 0. + dp, W, 0., -1, dp / 2)
    R.call("create_wall_y_mdbc", 0., 0 + dp, Ht, -1, dp / 2)
    R.call("create_wall_y_mdbc", W, 0 + dp, Ht, +1, W - dp / 2)
    R.call("create_corner_mdbc", -1, -1, 0. + dp / 2, 0. + dp / 2)
    R.call("create_corner_mdbc", +1, -1, W - dp / 2, 0. + dp / 2)
    R.call("create_block_of_fluid", 0 + dp, 0 + dp, 0.5 * W if dam else W - dp, depth)
    r = case.arrays["r"]
    fl = case.arrays["type"] == T.FLUID
    rr = R.get(po.F_R)
    assert np.array_equal(case.arrays["type"], R.types())
    assert np.array_equal(r[~fl], rr[~fl])
    # fluid positions carry the seeded +-0.01 dx jitter
    d = np.abs(r[fl] - rr[fl])
    assert 0 < d.max() <= 0.01 * dp
    g = R.get(po.F_GNR)
    assert np.array_equal(case.arrays["gnr"][: len(g)], g)
