"""The example drivers written against the reference's stage API
(include/tinysph_dropin.hpp, tinysph_old_amd/csrc/dropin_driver.cpp): their
particle set-up equals the reference geometry, and on the GPU their whole run
equals the oracle driven through the same stage sequence."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import oracle_init, oracle_reinit, oracle_step, oracle_verlet_step, rel_err

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tinysph_old_amd")
DRIVERS = {k: os.path.join(PKG, f"tsph_dropin_{k}") for k in ("mdbc", "rale", "verlet", "inflow")}
DP = 0.02


def load(path, ale):
    raw = open(path, "rb").read()
    np_, ne, np_out = np.frombuffer(raw[:12], dtype=np.int32)
    off = 12
    out = {"np_out": int(np_out)}

    def take(name, count, dtype, shape=None):
        nonlocal off
        size = np.dtype(dtype).itemsize * count
        a = np.frombuffer(raw[off:off + size], dtype=dtype).copy()
        off += size
        out[name] = a.reshape(shape) if shape else a

    take("type", np_, np.int32)
    take("r", 2 * np_, np.float64, (np_, 2))
    take("v", 2 * np_, np.float64, (np_, 2))
    take("rho", np_, np.float64)
    take("p", np_, np.float64)
    take("extra", 2 * ne, np.float64, (ne, 2))
    if ale:
        take("omega", np_, np.float64)
        take("omegarho", np_, np.float64)
        take("omegav", 2 * np_, np.float64, (np_, 2))
    assert off == len(raw)
    return out


def params(dp=DP):
    h = math.sqrt(2 * dp * dp)
    nvl = math.ceil(2.1 * h / dp)
    return [h, 2.1, 34.3, 1000., 0.02, 0.1, dp, -nvl * dp, 2. + nvl * dp, -nvl * dp, 1. + nvl * dp]


def run_driver(kind, steps, tmp_path, expect_ok=True, outdir=None):
    exe = DRIVERS[kind]
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run `python -m tinysph_old_amd.build`")
    init, final = tmp_path / f"{kind}_init.bin", tmp_path / f"{kind}_final.bin"
    print(f"[driver] {kind} dp={DP} steps={steps}", flush=True)
    extra = [str(outdir)] if outdir is not None else []
    res = subprocess.run([exe, str(DP), str(steps), str(init), str(final), *extra],
                         capture_output=True, text=True, timeout=120)
    if expect_ok:
        assert res.returncode == 0, res.stderr
    return res, init, final


def with_extra(d, kind):
    """Per-particle gnr / n from the wall-only list (walls come first)."""
    arrays = dict(type=d["type"], p=d["p"], rho=d["rho"], r=d["r"], v=d["v"])
    ext = np.zeros_like(d["r"])
    ext[: len(d["extra"])] = d["extra"]
    arrays["gnr" if kind == "mdbc" else "n"] = ext
    return arrays


@pytest.mark.parametrize("kind", ["mdbc", "rale"])
@pytest.mark.skipif(not (po.ref_available("wcsph") and po.ref_available("bt")),
                    reason="oracle/_ref not built")
def test_driver_geometry_matches_reference(kind, tmp_path):
    """The set-up the driver dumps before INIT equals the reference's geometry
    routines on the same arguments (cases/damBreakLong_mDBC/geometry.h,
    cases/ALESPH_dambreakLong/geometry.h).  Without a GPU the driver then stops
    at Divide_To_Cells with an error."""
    res, init, _ = run_driver(kind, 1, tmp_path, expect_ok=False)
    d = load(init, kind == "rale")
    dp, q = DP, params()
    if kind == "mdbc":
        R = po.Case(po.ref_lib("wcsph"), "ref_", q)
        R.call("create_wall_x_mdbc", 0. + dp, 1.61, 0., -1, dp / 2)
        R.call("create_wall_y_mdbc", 0., 0 + dp, 1., -1, dp / 2)
        R.call("create_wall_y_mdbc", 1.61, 0 + dp, 1., +1, 1.61 - dp / 2)
        R.call("create_corner_mdbc", -1, -1, 0. + dp / 2, 0. + dp / 2)
        R.call("create_corner_mdbc", +1, -1, 1.61 - dp / 2, 0. + dp / 2)
        R.call("create_block_of_fluid", 0 + dp, 0 + dp, 0.6, 0.3)
        extra = R.get(po.F_GNR)
    else:
        R = po.Case(po.ref_lib("rale"), "ref_", q)
        R.call("create_wall_x_bt", 0. + dp, 1.61 - dp, 0., -1, 0., 1.)
        R.call("create_wall_y_bt", 0., 0. + dp, 1., -1, 1., 0.)
        R.call("create_wall_y_bt", 1.61, 0. + dp, 1., +1, -1., 0.)
        R.call("create_corner_bt", -1, -1, 0., 0., math.sqrt(2), math.sqrt(2))
        R.call("create_corner_bt", +1, -1, 1.61, 0., -math.sqrt(2), math.sqrt(2))
        R.call("create_block_of_fluid", 0.0 + dp, 0 + dp, 0.6, 0.3)
        extra = R.get(po.F_N)
        assert np.array_equal(d["omega"], R.get(po.F_OMEGA))
        assert np.array_equal(d["omegarho"], R.get(po.F_OMEGARHO))
        assert np.array_equal(d["omegav"], R.get(po.F_OMEGAV))
    assert np.array_equal(d["type"], R.types())
    for k, f in (("r", po.F_R), ("v", po.F_V), ("rho", po.F_RHO), ("p", po.F_P)):
        assert np.array_equal(d[k], R.get(f)), k
    nw = len(d["extra"])
    assert np.array_equal(d["extra"], extra[:nw])


@pytest.mark.gpu
def test_driver_mdbc_matches_oracle(tmp_path):
    """20 steps of the unchanged reference driver sequence on the GPU (FP64)
    against the oracle on the dumped initial state."""
    _, init, final = run_driver("mdbc", 20, tmp_path)
    d0, d1 = load(init, False), load(final, False)
    O = po.oracle_case(params(), with_extra(d0, "mdbc"))
    oracle_init(O)
    dt = 0.2 * math.sqrt(2 * DP * DP) / 34.3
    for _ in range(20):
        oracle_step(O, 1, dt)
    assert np.array_equal(d1["type"], O.types())
    assert rel_err(d1["r"], O.get(po.F_R)) <= 1e-9
    assert rel_err(d1["v"], O.get(po.F_V)) <= 1e-9
    assert rel_err(d1["rho"], O.get(po.F_RHO)) <= 1e-9
    assert rel_err(d1["p"], O.get(po.F_P)) <= 1e-6


@pytest.mark.gpu
def test_driver_verlet_matches_oracle(tmp_path):
    """22 steps of the Verlet driver with density reinitialisation (Euler on
    steps 1 and 20, reinitialisation after step 20) against the oracle."""
    _, init, final = run_driver("verlet", 22, tmp_path)
    d0, d1 = load(init, False), load(final, False)
    O = po.oracle_case(params(), with_extra(d0, "mdbc"))
    oracle_init(O)
    dt = 0.2 * math.sqrt(2 * DP * DP) / 34.3
    for step in range(1, 23):
        oracle_verlet_step(O, 1, dt, step)
        if step % 20 == 0:
            oracle_reinit(O)
    assert np.array_equal(d1["type"], O.types())
    assert rel_err(d1["r"], O.get(po.F_R)) <= 1e-9
    assert rel_err(d1["v"], O.get(po.F_V)) <= 1e-9
    assert rel_err(d1["rho"], O.get(po.F_RHO)) <= 1e-9
    assert rel_err(d1["p"], O.get(po.F_P)) <= 1e-6


@pytest.mark.gpu
def test_driver_rale_matches_oracle(tmp_path):
    _, init, final = run_driver("rale", 20, tmp_path)
    d0, d1 = load(init, True), load(final, True)
    O = po.oracle_case(params(), with_extra(d0, "rale"))
    O.call("init_ale")
    oracle_init(O)
    dt = 0.2 * math.sqrt(2 * DP * DP) / 34.3
    for step in range(1, 21):
        if step > 1:
            O.call("rale_leapfrog_one", dt)
        O.call("clear_cells")
        assert O.call("particle_to_cell") == 0
        O.call("density_to_pressure")
        O.call("rale_forces")
        O.call("rale_leapfrog_two", dt, step)
        O.call("remove_out_of_domain")
    assert np.array_equal(d1["type"], O.types())
    for k, f in (("r", po.F_R), ("v", po.F_V), ("rho", po.F_RHO), ("omega", po.F_OMEGA),
                 ("omegarho", po.F_OMEGARHO), ("omegav", po.F_OMEGAV)):
        assert rel_err(d1[k], O.get(f)) <= 1e-9, k
    assert rel_err(d1["p"], O.get(po.F_P)) <= 1e-6


def inflow_params(dp=DP):
    """cases/inflowSlab_mDBC/constants.h with the domain padding of PREP_SIMULATION_DATA."""
    h = math.sqrt(2 * dp * dp)
    nvl = math.ceil(2.1 * h / dp)
    return [h, 2.1, 42.48, 1000., 0.1, 0.1, dp, -nvl * dp, 2. + nvl * dp, -nvl * dp, 1. + nvl * dp]


@pytest.mark.gpu
def test_driver_inflow_matches_oracle(tmp_path):
    """60 steps of the unchanged inflow driver sequence
    (cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:57-123) through the
    drop-in header: the inlet appends, RemoveParticlesOutOfDomain erases the
    column placed past the domain bound (every other particle per pass), and
    np_out counts the removals only (src/SPH_particle_system.h:115), as the
    oracle on the dumped initial state does."""
    from tests.helpers import oracle_add_zones, oracle_inflow_step
    steps = 60
    _, init, final = run_driver("inflow", steps, tmp_path)
    d0, d1 = load(init, False), load(final, False)
    O = po.oracle_case(inflow_params(), with_extra(d0, "mdbc"))
    dp = DP
    zones = [dict(x0=0.1, xm=0.1 - 4 * dp, y0=0.1 + dp, ym=0.2, vx=0.75, vy=0.0, ox=-1, oy=0,
                  btype=7),
             dict(x0=1.0, xm=1.0 + 4 * dp, y0=0.0 + dp, ym=0.5, vx=0.5, vy=0.0, ox=+1, oy=0,
                  btype=12)]
    oracle_add_zones(O, zones)
    oracle_init(O)
    dt = 0.00005 * DP / 0.005
    for _ in range(steps):
        oracle_inflow_step(O, 1, dt)
    removed = O.call("np_out")
    appended = O.np - len(d0["type"]) + removed
    assert removed > 0 and appended > 0
    assert d1["np_out"] == removed
    assert len(d1["type"]) == O.np
    assert np.array_equal(d1["type"], O.types())
    assert rel_err(d1["r"], O.get(po.F_R)) <= 1e-9
    assert rel_err(d1["v"], O.get(po.F_V)) <= 1e-9
    assert rel_err(d1["rho"], O.get(po.F_RHO)) <= 1e-9
    assert rel_err(d1["p"], O.get(po.F_P)) <= 1e-6


@pytest.mark.gpu
def test_driver_mdbc_measurements_and_text_output(tmp_path):
    """The mDBC driver with its measurement block
    (cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:156-201) and INIT's
    text output: the gauge and sensor files hold one line per sample with the
    oracle's values, the grid file is the cell grid of the oracle's first
    Particle_To_Cell, and the info file carries the case constants."""
    import ctypes as C

    from tinysph_old_amd import _lib as L
    steps = 100
    out = tmp_path / "out"
    out.mkdir()
    _, init, final = run_driver("mdbc", steps, tmp_path, outdir=out)
    d0 = load(init, False)
    q = params()
    O = po.oracle_case(q, with_extra(d0, "mdbc"))
    oracle_init(O)
    ncx, ncy = O.grid()
    start, _ = O.cell_csr()
    counts = np.ascontiguousarray(np.diff(start)[: ncx * ncy], dtype=np.int32)
    want_grid = tmp_path / "grid.vtk"
    assert L.lib().tsph_format_grid_vtk(ncx, ncy, q[0] * q[1], q[0] * 0.273, q[7], q[9],
                                        counts.ctypes.data_as(C.POINTER(C.c_int32)),
                                        str(want_grid).encode()) == 0
    assert (out / "grid.vtk").read_bytes() == want_grid.read_bytes()
    info = (out / "info.txt").read_text()
    assert "Dp: 0.02\n" in info and "Cs: 34.3\n" in info and "nvl: 3\n" in info
    dt = 0.2 * math.sqrt(2 * DP * DP) / 34.3
    gauges = [(0.3, 0.0, False), (0.3, 0.45, True), (0.865, 0.0, False), (0.865, 0.45, True),
              (1.114, 0.0, False), (1.114, 0.45, True), (1.3625, 0.0, False), (1.3625, 0.45, True)]
    names = ["h1", "h1_inv", "h2", "h2_inv", "h3", "h3_inv", "h4", "h4_inv"]
    wl = {n: [] for n in names}
    pr = {k: [] for k in range(1, 5)}
    for step in range(1, steps + 1):
        oracle_step(O, 1, dt)
        if step % 50 == 0:
            for n, (x, y0, inv) in zip(names, gauges):
                fn = "water_elevation_inverse" if inv else "water_elevation"
                wl[n].append((step * dt * math.sqrt(9.81 / 0.3), O.call(fn, x, y0) / 0.3))
        if step % 10 == 0:
            for k, y in zip(range(1, 5), (0.003, 0.015, 0.03, 0.08)):
                pr[k].append((step * dt * math.sqrt(9.81 / 0.3),
                              O.interpolate_point(1.61, y)[1] / (1000. * 9.81 * 0.3)))
    for n in names:
        got = np.loadtxt(out / f"waterLevel_{n}.dat", ndmin=2)
        want = np.array(wl[n])
        assert got.shape == want.shape == (steps // 50, 2), n
        # six significant digits as written; levels agree to a marching step
        assert np.allclose(got[:, 0], want[:, 0], rtol=1e-5), n
        assert np.allclose(got[:, 1], want[:, 1], rtol=1e-5, atol=2e-3), n
    for k in pr:
        got = np.loadtxt(out / f"pressure_p{k}.dat", ndmin=2)
        want = np.array(pr[k])
        assert got.shape == want.shape == (steps // 10, 2), k
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5), k
