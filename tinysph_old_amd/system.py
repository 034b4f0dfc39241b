"""Python host layer over the C ABI: case presets and a device-resident
particle system whose methods are the reference stage functions
(src/SPH_find_pairs.h, src/SPH_pressure.h, src/SPH_interaction*.h,
src/SPH_mDBC.h, src/SPH_BT.h, src/SPH_integration.h, src/SPH_btdebug.h,
src/SPH_water_level.h, src/SPH_measure_pressure.h).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


@dataclass
class CaseData:
    """A case preset: constants, padded domain, time step and particles in
    reference order (vectors as (np, 2) arrays)."""

    params: dict
    dt: float
    wall: int
    steps: int
    n_fluid: int
    n_wall: int
    arrays: dict = field(default_factory=dict)
    scheme: int = 0
    zones: list = field(default_factory=list)

    @property
    def np(self) -> int:
        return len(self.arrays["type"])

    def oracle_params(self) -> list:
        p = self.params
        return [p["h"], p["kap"], p["cs"], p["rho0"], p["avisc"], p["delta"], p["dp"],
                p["x_0"], p["x_m"], p["y_0"], p["y_m"]]


def _case_from_c(c) -> CaseData:
    prm = {n: getattr(c.prm, n) for n, _ in L.Params._fields_}
    h = c.particles
    n = h.np

    def arr(ptr, count, dtype):
        if not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    arrays = {
        "type": arr(h.type, n, np.int32),
        "index": arr(h.index, n, np.int32),
        "p": arr(h.p, n, np.float64),
        "rho": arr(h.rho, n, np.float64),
        "r": arr(h.r, 2 * n, np.float64).reshape(n, 2),
        "v": arr(h.v, 2 * n, np.float64).reshape(n, 2),
    }
    if h.gnr:
        arrays["gnr"] = arr(h.gnr, 2 * n, np.float64).reshape(n, 2)
    if h.n:
        arrays["n"] = arr(h.n, 2 * n, np.float64).reshape(n, 2)
    zones = [{n: getattr(c.zones[i], n) for n, _ in L.Zone._fields_} for i in range(c.n_zones)]
    return CaseData(prm, c.dt, c.wall, c.steps, c.n_fluid, c.n_wall, arrays, c.scheme, zones)


def build_case(preset: int, dx: float = 0.0, jitter: float = -1.0, seed: int = 12345) -> CaseData:
    lib = L.lib()
    out = C.POINTER(L.CaseC)()
    L.check(lib.tsph_case_build(preset, float(dx), float(jitter), int(seed), C.byref(out)),
            "tsph_case_build")
    try:
        return _case_from_c(out.contents)
    finally:
        lib.tsph_case_free(out)


def build_ensemble(ncases: int, seed: int, nmin: int = 2000, nmax: int = 400000) -> list:
    """The batched-ensemble workload (tsph_ensemble_build): ncases tanks drawn
    from seed, each holding nmin..nmax fluid particles."""
    lib = L.lib()
    arr = (C.POINTER(L.CaseC) * ncases)()
    L.check(lib.tsph_ensemble_build(int(ncases), int(seed), int(nmin), int(nmax), arr),
            "tsph_ensemble_build")
    try:
        return [_case_from_c(arr[q].contents) for q in range(ncases)]
    finally:
        for q in range(ncases):
            if arr[q]:
                lib.tsph_case_free(arr[q])


VEC = ("r", "v", "a", "r_o", "v_o", "gnr", "n")
SCAL = ("p", "rho", "drho", "rho_o")


DOWNLOAD_FIELDS = ("type", "index", "p", "rho", "drho", "r", "v", "a", "r_o", "v_o", "rho_o")


def host_particles(arrays: dict):
    """tsph_particles over the numpy arrays of a case (plus the arrays it keeps alive)."""
    n = len(arrays["type"])
    keep = {}
    h = L.Particles()
    h.np = n
    t = np.ascontiguousarray(arrays["type"], dtype=np.int32)
    keep["type"] = t
    h.type = L.iptr(t)
    if arrays.get("index") is not None:
        ix = np.ascontiguousarray(arrays["index"], dtype=np.int32)
        keep["index"] = ix
        h.index = L.iptr(ix)
    for k in VEC + SCAL:
        a = arrays.get(k)
        if a is None:
            continue
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        keep[k] = a
        setattr(h, k, L.dptr(a))
    return h, keep


def download_buffers(n: int, fields):
    out = {}
    h = L.Particles()
    for k in fields:
        if k in ("type", "index"):
            a = np.zeros(max(n, 1), dtype=np.int32)
            setattr(h, k, L.iptr(a))
        elif k in VEC:
            a = np.zeros(max(2 * n, 1), dtype=np.float64)
            setattr(h, k, L.dptr(a))
        else:
            a = np.zeros(max(n, 1), dtype=np.float64)
            setattr(h, k, L.dptr(a))
        out[k] = a
    return h, out


def trim_download(h, out: dict) -> dict:
    m = h.np
    for k in list(out):
        out[k] = out[k][: 2 * m].reshape(m, 2) if k in VEC else out[k][:m]
    return out


class ParticleSystem:
    """Particle_system + Simulation_data on one GPU (src/SPH_particle_system.h:7-213)."""

    def __init__(self, params: dict, precision: int = L.PREC_FP64, capacity: int | None = None):
        self.lib = L.lib()
        prm = L.Params()
        for n, _ in L.Params._fields_:
            if n in params and n not in ("precision", "capacity"):
                setattr(prm, n, float(params[n]))
        prm.precision = precision
        prm.capacity = int(capacity if capacity is not None else params.get("capacity", 1))
        self.capacity = prm.capacity
        self.precision = precision
        self.ctx = C.c_void_p()
        L.check(self.lib.tsph_create(C.byref(prm), C.byref(self.ctx)), "tsph_create")

    def close(self):
        if self.ctx:
            self.lib.tsph_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- transfer ----
    def upload(self, arrays: dict) -> None:
        h, keep = host_particles(arrays)
        L.check(self.lib.tsph_upload(self.ctx, C.byref(h)), "tsph_upload")

    def download(self, fields=DOWNLOAD_FIELDS) -> dict:
        h, out = download_buffers(self.np_upper(), fields)
        L.check(self.lib.tsph_download(self.ctx, C.byref(h)), "tsph_download")
        return trim_download(h, out)

    def np_upper(self) -> int:
        return self.np

    @property
    def np(self) -> int:
        n = C.c_int()
        L.check(self.lib.tsph_get_np(self.ctx, C.byref(n)), "tsph_get_np")
        return n.value

    def grid(self):
        x, y = C.c_int(), C.c_int()
        L.check(self.lib.tsph_grid(self.ctx, C.byref(x), C.byref(y)), "tsph_grid")
        return x.value, y.value

    @property
    def mass(self) -> float:
        m = C.c_double()
        L.check(self.lib.tsph_mass(self.ctx, C.byref(m)), "tsph_mass")
        return m.value

    # ---- stages ----
    def clear_cells(self):
        L.check(self.lib.tsph_clear_cells(self.ctx), "Clear_cells")

    def particle_to_cell(self):
        L.check(self.lib.tsph_particle_to_cell(self.ctx), "Particle_To_Cell")

    def cell_lists(self):
        ncx, ncy = self.grid()
        n = self.np
        start = np.zeros(ncx * ncy + 1, dtype=np.int32)
        items = np.zeros(max(n, 1), dtype=np.int32)
        L.check(self.lib.tsph_get_cell_lists(self.ctx, L.iptr(start), L.iptr(items)),
                "tsph_get_cell_lists")
        return start, items[:n]

    def density_to_pressure(self):
        L.check(self.lib.tsph_density_to_pressure(self.ctx), "Density_to_pressure")

    def compute_forces(self, variant: int):
        L.check(self.lib.tsph_compute_forces(self.ctx, variant), "Compute_Forces")

    def mdbc_density(self):
        L.check(self.lib.tsph_mdbc_density(self.ctx), "mDBC_compute_density_mdbcGeo")

    def bt_density(self):
        L.check(self.lib.tsph_bt_density(self.ctx), "BT_compute_density_BTGeo")

    def predictor(self, dt: float):
        L.check(self.lib.tsph_symplectic_predictor(self.ctx, dt), "Integrate_SymplecticPredictor")

    def corrector(self, dt: float):
        L.check(self.lib.tsph_symplectic_corrector(self.ctx, dt), "Integrate_SymplecticCorrector")

    def remove_out_of_domain(self):
        L.check(self.lib.tsph_remove_out_of_domain(self.ctx), "RemoveParticlesOutOfDomain")

    def step(self, wall: int, dt: float, nsteps: int = 1):
        L.check(self.lib.tsph_step_symplectic(self.ctx, wall, dt, nsteps), "tsph_step_symplectic")

    def synchronize(self):
        L.check(self.lib.tsph_synchronize(self.ctx), "tsph_synchronize")

    def water_elevation(self, x: float, y0: float, inverse: bool = False) -> float:
        out = C.c_double()
        L.check(self.lib.tsph_water_elevation(self.ctx, x, y0, int(inverse), C.byref(out)),
                "Water_Elevation")
        return out.value

    def interpolate_point(self, x: float, y: float) -> np.ndarray:
        out = np.zeros(4)
        L.check(self.lib.tsph_interpolate_point(self.ctx, x, y, L.dptr(out)), "Pressure_in_point")
        return out

    def counters(self) -> dict:
        c = L.Counters()
        L.check(self.lib.tsph_get_counters(self.ctx, C.byref(c)), "tsph_get_counters")
        return {n: int(getattr(c, n)) for n, _ in L.Counters._fields_}

    def reset_counters(self):
        L.check(self.lib.tsph_reset_counters(self.ctx), "tsph_reset_counters")

    def set_tile_rows(self, rows: int):
        L.check(self.lib.tsph_set_tile_rows(self.ctx, rows), "tsph_set_tile_rows")

    # ---- Riemann-ALE-SPH (src/RALESPH_interaction.h, src/RALESPH_integration.h) ----
    def ale_enable(self):
        L.check(self.lib.tsph_ale_enable(self.ctx), "tsph_ale_enable")

    def ale_upload(self, fields: dict):
        """Overwrite ALE variables (reference order; missing keys keep the
        ALESPH_Add_particle values)."""
        h = L.AleFields()
        h.np = self.np
        keep = []
        for k in L.ALE_NAMES:
            if fields.get(k) is None:
                continue
            a = np.ascontiguousarray(fields[k], dtype=np.float64).reshape(-1)
            keep.append(a)
            setattr(h, k, L.dptr(a))
            h.np = len(a) // 2 if k in L.ALE_VEC else len(a)
        L.check(self.lib.tsph_ale_upload(self.ctx, C.byref(h)), "tsph_ale_upload")

    def ale_download(self) -> dict:
        n = self.np
        h = L.AleFields()
        out = {}
        for k in L.ALE_NAMES:
            a = np.zeros(max(2 * n if k in L.ALE_VEC else n, 1))
            setattr(h, k, L.dptr(a))
            out[k] = a
        L.check(self.lib.tsph_ale_download(self.ctx, C.byref(h)), "tsph_ale_download")
        m = h.np
        return {k: (v[: 2 * m].reshape(m, 2) if k in L.ALE_VEC else v[:m]) for k, v in out.items()}

    def rale_forces(self):
        L.check(self.lib.tsph_rale_forces(self.ctx), "Compute_Forces (RALE)")

    def rale_leapfrog_one(self, dt: float):
        L.check(self.lib.tsph_rale_leapfrog_one(self.ctx, dt), "Integrate_LeapFrog_partOne")

    def rale_leapfrog_two(self, dt: float, step: int):
        L.check(self.lib.tsph_rale_leapfrog_two(self.ctx, dt, step), "Integrate_LeapFrog_partTwo")

    def step_rale(self, dt: float, step: int, nsteps: int = 1):
        L.check(self.lib.tsph_step_rale(self.ctx, dt, step, nsteps), "tsph_step_rale")

    # ---- buffer zones (src/SPH_buffer_zones.h) ----
    def bz_add_zone(self, zone: dict) -> int:
        z = L.Zone()
        for n, _ in L.Zone._fields_:
            setattr(z, n, zone[n])
        idx = C.c_int()
        L.check(self.lib.tsph_bz_add_zone(self.ctx, C.byref(z), C.byref(idx)), "tsph_bz_add_zone")
        return idx.value

    def bz_integrate_inlet(self, zone: int, dt: float):
        L.check(self.lib.tsph_bz_integrate_inlet(self.ctx, zone, dt), "BZ_Integrate_inlet")

    def bz_integrate_outlet(self, zone: int, dt: float):
        L.check(self.lib.tsph_bz_integrate_outlet(self.ctx, zone, dt), "BZ_Integrate_outlet")

    def set_compensated(self, on: bool = True):
        """Kahan-compensated position and density updates (tsph_set_compensated)."""
        L.check(self.lib.tsph_set_compensated(self.ctx, int(on)), "tsph_set_compensated")

    def renumber(self):
        """Relabel the live particles' reference positions 0..np-1 (tsph_renumber)."""
        L.check(self.lib.tsph_renumber(self.ctx), "tsph_renumber")

    def step_inflow(self, wall: int, dt: float, nsteps: int = 1, inlet: int = 0, outlet: int = 1):
        L.check(self.lib.tsph_step_inflow(self.ctx, wall, dt, nsteps, inlet, outlet), "tsph_step_inflow")

    # ---- ALE-SPH with MUSCL reconstruction (src/ALESPH_*.h) ----
    def ale_compute_density(self, limited: bool = False):
        L.check(self.lib.tsph_ale_compute_density(self.ctx, int(limited)), "Compute_Density")

    def ale_kernel_gradient(self):
        L.check(self.lib.tsph_ale_kernel_gradient(self.ctx), "Kernel_gradient_approx")

    def ale_density_pressure(self, dt: float):
        L.check(self.lib.tsph_ale_density_pressure(self.ctx, dt), "Integrate_density_compute_pressure")

    def ale_compute_acceleration(self, limited: bool = False):
        L.check(self.lib.tsph_ale_compute_acceleration(self.ctx, int(limited)), "Compute_Acceleration_BT")

    def ale_leapfrog_one(self, dt: float):
        L.check(self.lib.tsph_ale_leapfrog_one(self.ctx, dt), "Integrate_LeapFrog_partOne_ALESPH")

    def ale_leapfrog_two(self, dt: float, step: int):
        L.check(self.lib.tsph_ale_leapfrog_two(self.ctx, dt, step), "Integrate_LeapFrog_partTwo_ALESPH")

    def step_ale(self, dt: float, step: int, nsteps: int = 1, muscl: int = 0):
        L.check(self.lib.tsph_step_ale(self.ctx, dt, step, nsteps, muscl), "tsph_step_ale")

    # ---- output, snapshot, resume (src/SPH_data_to_vtk.h, src/SPH_dataInterpolation.h) ----
    def write_vtk(self, path: str):
        L.check(self.lib.tsph_write_vtk(self.ctx, str(path).encode()), "write_to_ASCII_VTK")

    def write_grid_vtk(self, x_0: float, y_0: float, path: str):
        """write_mesh_to_ASCII_VTK (src/SPH_grid_to_vtk.h:6-67)."""
        L.check(self.lib.tsph_write_grid_vtk(self.ctx, float(x_0), float(y_0), str(path).encode()),
                "write_mesh_to_ASCII_VTK")

    def write_hydrostatic_pressure(self, path: str):
        """Output_hydrostatic_pressure (src/SPH_measure_pressure.h:17-34)."""
        L.check(self.lib.tsph_write_hydrostatic_pressure(self.ctx, str(path).encode()),
                "Output_hydrostatic_pressure")

    def io_wait(self):
        L.check(self.lib.tsph_io_wait(self.ctx), "tsph_io_wait")

    def interpolate_grid(self, x0: float, y0: float, xm: float, ym: float):
        """GenerateInterpol's grid: (rho, p, v) as (ncx, ncy) / (ncx, ncy, 2) arrays."""
        nx, ny = C.c_int(), C.c_int()
        L.check(self.lib.tsph_interpolate_grid(self.ctx, x0, y0, xm, ym, C.byref(nx), C.byref(ny),
                                               None, None, None), "tsph_interpolate_grid")
        n = nx.value * ny.value
        rho, p, v = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(2 * n, 1))
        L.check(self.lib.tsph_interpolate_grid(self.ctx, x0, y0, xm, ym, C.byref(nx), C.byref(ny),
                                               L.dptr(rho), L.dptr(p), L.dptr(v)),
                "tsph_interpolate_grid")
        shp = (nx.value, ny.value)
        return rho[:n].reshape(shp), p[:n].reshape(shp), v[:2 * n].reshape(shp + (2,))

    def write_interpolated_vtk(self, x0: float, y0: float, xm: float, ym: float, path: str):
        L.check(self.lib.tsph_write_interpolated_vtk(self.ctx, x0, y0, xm, ym, str(path).encode()),
                "GenerateInterpol")

    def save_state(self, path: str):
        L.check(self.lib.tsph_save_state(self.ctx, str(path).encode()), "tsph_save_state")

    def load_state(self, path: str):
        L.check(self.lib.tsph_load_state(self.ctx, str(path).encode()), "tsph_load_state")

    # ---- other WCSPH integrators and density reinitialisation
    # (src/SPH_integration.h:103-230, src/SPH_density_reinitialization.h) ----
    def history_enable(self):
        L.check(self.lib.tsph_history_enable(self.ctx), "tsph_history_enable")

    def history_download(self) -> dict:
        n = self.np
        voo, roo, rt = np.zeros(max(2 * n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
        h = L.History(n, L.dptr(voo), L.dptr(roo), L.dptr(rt))
        L.check(self.lib.tsph_history_download(self.ctx, C.byref(h)), "tsph_history_download")
        m = h.np
        return {"v_oo": voo[: 2 * m].reshape(m, 2), "rho_oo": roo[:m], "rho_temp": rt[:m]}

    def verlet(self, dt: float):
        L.check(self.lib.tsph_verlet(self.ctx, dt), "Integrate_Verlet")

    def euler(self, dt: float):
        L.check(self.lib.tsph_euler(self.ctx, dt), "Integrate_Euler")

    def leapfrog_one(self, dt: float):
        L.check(self.lib.tsph_leapfrog_one(self.ctx, dt), "Integrate_LeapFrog_partOne_withDensity")

    def leapfrog_two(self, dt: float, step: int):
        L.check(self.lib.tsph_leapfrog_two(self.ctx, dt, step),
                "Integrate_LeapFrog_partTwo_withDensity")

    def density_reinit(self):
        L.check(self.lib.tsph_density_reinit_sum(self.ctx), "Density_reinitialization_sumation")
        L.check(self.lib.tsph_density_reinit_overwrite(self.ctx),
                "Density_reinitialization_overwrite")

    def step_verlet(self, wall: int, dt: float, step: int, nsteps: int = 1):
        L.check(self.lib.tsph_step_verlet(self.ctx, wall, dt, step, nsteps), "tsph_step_verlet")

    def step_leapfrog(self, wall: int, dt: float, step: int, nsteps: int = 1):
        L.check(self.lib.tsph_step_leapfrog(self.ctx, wall, dt, step, nsteps), "tsph_step_leapfrog")

    def step_symplectic_reinit(self, wall: int, dt: float, step: int, nsteps: int = 1,
                               interval: int = 20):
        L.check(self.lib.tsph_step_symplectic_reinit(self.ctx, wall, dt, step, nsteps, interval),
                "tsph_step_symplectic_reinit")

    def set_balanced(self, on: bool):
        L.check(self.lib.tsph_set_balanced(self.ctx, int(bool(on))), "tsph_set_balanced")

    def set_cell_order(self, mode: int):
        """0: by size (default), 1: ranked by the reorder, 2: separate ordering pass."""
        L.check(self.lib.tsph_set_cell_order(self.ctx, mode), "tsph_set_cell_order")

    def set_waves_per_cu(self, waves: int):
        L.check(self.lib.tsph_set_waves_per_cu(self.ctx, waves), "tsph_set_waves_per_cu")

    def set_scan_wait(self, polls: int):
        L.check(self.lib.tsph_set_scan_wait(self.ctx, polls), "tsph_set_scan_wait")

    @property
    def tile_rows(self) -> int:
        r = C.c_int()
        L.check(self.lib.tsph_get_tile_rows(self.ctx, C.byref(r)), "tsph_get_tile_rows")
        return r.value

    def pair_timing(self, max_launches: int):
        L.check(self.lib.tsph_pair_timing(self.ctx, max_launches), "tsph_pair_timing")

    def pair_timing_read(self):
        ms, n = C.c_double(), C.c_int()
        L.check(self.lib.tsph_pair_timing_read(self.ctx, C.byref(ms), C.byref(n)),
                "tsph_pair_timing_read")
        return ms.value, n.value


def system_for_case(case: CaseData, precision: int = L.PREC_FP64, capacity: int | None = None
                    ) -> ParticleSystem:
    if capacity is None:
        capacity = max(case.np, int(case.params.get("capacity", 0)))
    ps = ParticleSystem(case.params, precision=precision, capacity=capacity)
    ps.upload(case.arrays)
    for z in case.zones:
        ps.bz_add_zone(z)
    return ps


def forces_variant(wall: int) -> int:
    return {L.WALL_DBC: L.FORCES_DBC, L.WALL_MDBC: L.FORCES_WCSPH, L.WALL_BT: L.FORCES_BT}[wall]


class BatchSystem(ParticleSystem):
    """A batched ensemble: many independent cases advanced together, one launch
    per stage (tsph_batch_*).  The stage methods of ParticleSystem that act on
    all cases (particle_to_cell, density_to_pressure, compute_forces,
    mdbc_density, bt_density, remove_out_of_domain, counters, tuning) apply."""

    def __init__(self, cases: list, precision: int = L.PREC_FP64, capacity: list | None = None):
        self.lib = L.lib()
        n = len(cases)
        prms = (L.Params * n)()
        for q, case in enumerate(cases):
            for name, _ in L.Params._fields_:
                if name in case.params and name not in ("precision", "capacity"):
                    setattr(prms[q], name, float(case.params[name]))
            prms[q].precision = precision
            prms[q].capacity = int(capacity[q] if capacity is not None else case.np)
        dts = np.ascontiguousarray([c.dt for c in cases], dtype=np.float64)
        self.ncases = n
        self.capacity = sum(int(prms[q].capacity) for q in range(n))
        self.precision = precision
        self.ctx = C.c_void_p()
        L.check(self.lib.tsph_batch_create(prms, L.dptr(dts), n, C.byref(self.ctx)),
                "tsph_batch_create")

    def upload(self, arrays_per_case: list) -> None:
        hs = (L.Particles * self.ncases)()
        keep = []
        for q, arrays in enumerate(arrays_per_case):
            h, k = host_particles(arrays)
            hs[q] = h
            keep.append(k)
        L.check(self.lib.tsph_batch_upload(self.ctx, hs), "tsph_batch_upload")

    def np_per_case(self) -> np.ndarray:
        out = np.zeros(self.ncases, dtype=np.int32)
        L.check(self.lib.tsph_batch_get_np(self.ctx, L.iptr(out)), "tsph_batch_get_np")
        return out

    def download_case(self, q: int, fields=DOWNLOAD_FIELDS) -> dict:
        h, out = download_buffers(int(self.np_per_case()[q]), fields)
        L.check(self.lib.tsph_batch_download(self.ctx, int(q), C.byref(h)), "tsph_batch_download")
        return trim_download(h, out)

    def grid_of(self, q: int):
        x, y, m = C.c_int(), C.c_int(), C.c_double()
        L.check(self.lib.tsph_batch_grid(self.ctx, int(q), C.byref(x), C.byref(y), C.byref(m)),
                "tsph_batch_grid")
        return x.value, y.value, m.value

    def step(self, wall: int, nsteps: int = 1):
        L.check(self.lib.tsph_batch_step(self.ctx, wall, int(nsteps)), "tsph_batch_step")


def batch_for_cases(cases: list, precision: int = L.PREC_FP64) -> BatchSystem:
    bs = BatchSystem(cases, precision=precision)
    bs.upload([c.arrays for c in cases])
    return bs


def autotune(ps: ParticleSystem, variant: int, with_particles_per_lane: bool = False):
    """tsph_autotune: (tile_rows, waves_per_cu, balanced, ms per pair pass) of the
    fastest setting, and its mean particles per lane when asked for."""
    r, w, b, ms, ppl = C.c_int(), C.c_int(), C.c_int(), C.c_double(), C.c_double()
    L.check(ps.lib.tsph_autotune(ps.ctx, variant, C.byref(r), C.byref(w), C.byref(b),
                                 C.byref(ms), C.byref(ppl)), "tsph_autotune")
    if with_particles_per_lane:
        return r.value, w.value, b.value, ms.value, ppl.value
    return r.value, w.value, b.value, ms.value
