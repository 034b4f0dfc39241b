"""ctypes wrappers around the oracle (liboracle.so) and the reference harness
(oracle/_ref/ref_*.so).  Test infrastructure only: imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the product.

Both libraries share one calling convention, so a test can drive the oracle and
the reference through the same ``Case`` interface and compare field by field.
Field ids: 0 r, 1 v, 2 a, 3 p, 4 rho, 5 drho, 6 r_o, 7 v_o, 8 rho_o, 9 gnr, 10 n.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

F_R, F_V, F_A, F_P, F_RHO, F_DRHO, F_RO, F_VO, F_RHOO, F_GNR, F_N = range(11)
# ALE volume variables
(F_OMEGA, F_OMEGARHO, F_OMEGAV, F_OMEGA_O, F_OMEGARHO_O, F_OMEGAV_O, F_DOMEGA, F_DOMEGARHO,
 F_DOMEGAV) = range(11, 20)
F_GRADRHO, F_GRADVX, F_GRADVY = 20, 21, 22
GRAD_FIELDS = (F_GRADRHO, F_GRADVX, F_GRADVY)
# Verlet second time layer and the density-reinitialisation buffer
F_VOO, F_RHOOO, F_RHOTMP = 23, 24, 25
HISTORY_FIELDS = (F_VOO, F_RHOOO, F_RHOTMP)
VEC_FIELDS = {F_R, F_V, F_A, F_RO, F_VO, F_GNR, F_N, F_OMEGAV, F_OMEGAV_O, F_DOMEGAV, F_GRADRHO,
              F_GRADVX, F_GRADVY, F_VOO}
ALE_FIELDS = tuple(range(11, 20))

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref: bool = False) -> None:
    """Compile the oracle (and, when the reference tree is present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load(path: str) -> C.CDLL:
    return C.CDLL(path)


def oracle_lib(omp: bool = False) -> C.CDLL:
    name = "liboracle_omp.so" if omp else "liboracle.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = _load(path)
    _declare(lib, "orc_")
    return lib


def ref_available(variant: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"ref_{variant}.so"))


def ref_lib(variant: str) -> C.CDLL:
    """variant in {"wcsph", "dbc", "bt", "rale"}"""
    lib = _load(os.path.join(REF_DIR, f"ref_{variant}.so"))
    _declare(lib, "ref_")
    return lib


def _declare(lib: C.CDLL, pfx: str) -> None:
    def f(name, res, *args):
        fn = getattr(lib, pfx + name, None)
        if fn is None:
            return
        fn.restype = res
        fn.argtypes = list(args)

    f("new", C.c_void_p, _dp)
    f("free", None, C.c_void_p)
    f("np", C.c_int, C.c_void_p)
    f("np_out", C.c_int, C.c_void_p)
    f("mass", C.c_double, C.c_void_p)
    f("get", C.c_int, C.c_void_p, C.c_int, _dp)
    f("set", C.c_int, C.c_void_p, C.c_int, _dp)
    f("get_type", None, C.c_void_p, _ip)
    f("get_index", None, C.c_void_p, _ip)
    f("divide_to_cells", None, C.c_void_p)
    f("clear_cells", None, C.c_void_p)
    f("particle_to_cell", C.c_int if pfx == "orc_" else None, C.c_void_p)
    f("grid", None, C.c_void_p, _ip, _ip)
    f("cell_csr", None, C.c_void_p, _ip, _ip)
    f("density_to_pressure", C.c_int if pfx == "orc_" else None, C.c_void_p)
    f("predictor", None, C.c_void_p, C.c_double)
    f("corrector", None, C.c_void_p, C.c_double)
    f("remove_out_of_domain", C.c_int if pfx == "orc_" else None, C.c_void_p)
    f("water_elevation", C.c_double, C.c_void_p, C.c_double, C.c_double)
    f("water_elevation_inverse", C.c_double, C.c_void_p, C.c_double, C.c_double)
    f("sum_kernel_function", C.c_double, C.c_void_p, C.c_double, C.c_double)
    f("interpolate_point", None, C.c_void_p, C.c_double, C.c_double, _dp)
    f("rale_leapfrog_one", None, C.c_void_p, C.c_double)
    f("rale_leapfrog_two", None, C.c_void_p, C.c_double, C.c_int)
    # ALE-SPH with MUSCL reconstruction
    f("ale_compute_density", None, C.c_void_p)
    f("ale_compute_acceleration", None, C.c_void_p)
    f("ale_kernel_gradient", None, C.c_void_p)
    f("ale_density_pressure", None, C.c_void_p, C.c_double)
    f("ale_leapfrog_one", None, C.c_void_p, C.c_double)
    f("ale_leapfrog_two", None, C.c_void_p, C.c_double, C.c_int)
    # alternate WCSPH integrators and density reinitialisation
    f("verlet", None, C.c_void_p, C.c_double)
    f("euler", None, C.c_void_p, C.c_double)
    f("leapfrog_one", None, C.c_void_p, C.c_double)
    f("leapfrog_two", None, C.c_void_p, C.c_double, C.c_int)
    f("reinit_sum", None, C.c_void_p)
    f("reinit_overwrite", C.c_int if pfx == "orc_" else None, C.c_void_p)
    if pfx == "orc_":
        f("add_zone", C.c_int, C.c_void_p, *([C.c_double] * 8), C.c_int)
        f("bz_integrate_inlet", C.c_int, C.c_void_p, C.c_int, C.c_double)
        f("bz_integrate_outlet", C.c_int, C.c_void_p, C.c_int, C.c_double)
        f("ale_density_pressure", C.c_int, C.c_void_p, C.c_double)
        f("ale_compute_density_limited", None, C.c_void_p)
        f("ale_compute_acceleration_limited", None, C.c_void_p)
        f("minmod_phi", C.c_double, *([C.c_double] * 6))
        f("muscl_states", None, _dp, C.c_int, _dp)
        f("init_ale", None, C.c_void_p)
        f("rale_forces", None, C.c_void_p)
        f("set_threads", None, C.c_void_p, C.c_int)
        f("load", None, C.c_void_p, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp)
        f("compute_forces", None, C.c_void_p, C.c_int)
        f("mdbc_density", None, C.c_void_p)
        f("bt_density", None, C.c_void_p)
        f("wendland_w", C.c_double, C.c_double, C.c_double)
        f("wendland_f", C.c_double, C.c_double, C.c_double)
        f("tait", C.c_double, C.c_double, C.c_double, C.c_double)
        f("pressure_to_density", C.c_double, C.c_double, C.c_double, C.c_double)
        f("artificial_viscosity", C.c_double, *([C.c_double] * 6))
        f("diffusion_fourtakas", C.c_double, *([C.c_double] * 11))
    else:
        f("add", None, C.c_void_p, C.c_int, *([C.c_double] * 6))
        f("draw_create_inlet_y", None, C.c_void_p, *([C.c_double] * 3), C.c_int, C.c_int,
          C.c_double, C.c_double, C.c_int)
        f("draw_create_inlet_x", None, C.c_void_p, *([C.c_double] * 3), C.c_int, C.c_int,
          C.c_double, C.c_double, C.c_int)
        f("draw_create_outlet_y", None, C.c_void_p, *([C.c_double] * 3), C.c_int, C.c_int,
          C.c_double, C.c_double, C.c_int, C.c_int)
        f("zone_count", C.c_int, C.c_void_p)
        f("zone_get", None, C.c_void_p, C.c_int, _dp)
        f("bz_integrate_inlet", None, C.c_void_p, C.c_int, C.c_double)
        f("bz_integrate_outlet", None, C.c_void_p, C.c_int, C.c_double)
        f("add_ale", None, C.c_void_p, C.c_int, *([C.c_double] * 6))
        f("push_gnr", None, C.c_void_p, C.c_double, C.c_double)
        f("push_n", None, C.c_void_p, C.c_double, C.c_double)
        f("compute_forces", None, C.c_void_p)
        f("wall_density", None, C.c_void_p)
        f("create_wall_x_mdbc", None, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double)
        f("create_wall_y_mdbc", None, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double)
        f("create_corner_mdbc", None, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double)
        f("create_block_of_fluid", None, C.c_void_p, *([C.c_double] * 4))
        f("create_block_of_fluid_wp", None, C.c_void_p, *([C.c_double] * 4))
        f("create_wall_x_bt", None, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double)
        f("create_wall_y_bt", None, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double)
        f("create_corner_bt", None, C.c_void_p, C.c_int, C.c_int, *([C.c_double] * 4))


def _ptr(a: np.ndarray, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Case:
    """One particle system held by the oracle ("orc_") or the reference ("ref_")."""

    def __init__(self, lib: C.CDLL, pfx: str, q):
        self.lib, self.pfx = lib, pfx
        qa = np.ascontiguousarray(q, dtype=np.float64)
        self.h = getattr(lib, pfx + "new")(_ptr(qa))

    def __del__(self):
        try:
            getattr(self.lib, self.pfx + "free")(self.h)
        except Exception:
            pass

    def call(self, name, *args):
        return getattr(self.lib, self.pfx + name)(self.h, *args)

    @property
    def np(self) -> int:
        return self.call("np")

    def get(self, f: int) -> np.ndarray:
        n = self.np
        gn = getattr(self.lib, self.pfx + "field_len", None)
        size = 2 * n if f in VEC_FIELDS else n
        if self.pfx == "ref_" and gn is not None:
            gn.restype = C.c_int
            gn.argtypes = [C.c_void_p, C.c_int]
            size = gn(self.h, f)
        out = np.zeros(max(size, 1), dtype=np.float64)
        k = self.call("get", f, _ptr(out))
        out = out[:k]
        return out.reshape(-1, 2) if f in VEC_FIELDS else out

    def set(self, f: int, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self.call("set", f, _ptr(a))

    def types(self) -> np.ndarray:
        out = np.zeros(max(self.np, 1), dtype=np.int32)
        self.call("get_type", _ptr(out, C.c_int))
        return out[: self.np]

    def index(self) -> np.ndarray:
        out = np.zeros(max(self.np, 1), dtype=np.int32)
        self.call("get_index", _ptr(out, C.c_int))
        return out[: self.np]

    def grid(self):
        x, y = C.c_int(), C.c_int()
        self.call("grid", C.byref(x), C.byref(y))
        return x.value, y.value

    def cell_csr(self):
        ncx, ncy = self.grid()
        start = np.zeros(ncx * ncy + 1, dtype=np.int32)
        items = np.zeros(max(self.np, 1), dtype=np.int32)
        self.call("cell_csr", _ptr(start, C.c_int), _ptr(items, C.c_int))
        return start, items[: self.np]

    def interpolate_point(self, x, y):
        out = np.zeros(4)
        self.call("interpolate_point", C.c_double(x), C.c_double(y), _ptr(out))
        return out


def params_vector(h, kap, cs, rho0, avisc, delta, dp, x0, xm, y0, ym):
    return [h, kap, cs, rho0, avisc, delta, dp, x0, xm, y0, ym]


def oracle_case(q, particles: dict, omp_threads: int = 0) -> Case:
    """Oracle case loaded from arrays in reference order (dict of numpy arrays
    with keys type, p, rho, r, v and optionally gnr, n)."""
    lib = oracle_lib(omp=omp_threads > 1)
    c = Case(lib, "orc_", q)
    if omp_threads > 1:
        c.call("set_threads", omp_threads)
    t = np.ascontiguousarray(particles["type"], dtype=np.int32)
    n = len(t)
    arrs = [np.ascontiguousarray(particles[k], dtype=np.float64) for k in ("p", "rho", "r", "v")]
    gnr = particles.get("gnr")
    nn = particles.get("n")
    gnr = None if gnr is None else np.ascontiguousarray(gnr, dtype=np.float64)
    nn = None if nn is None else np.ascontiguousarray(nn, dtype=np.float64)
    c.call("load", n, _ptr(t, C.c_int), *[_ptr(a) for a in arrs],
           _ptr(gnr) if gnr is not None else None, _ptr(nn) if nn is not None else None)
    c._keep = (t, arrs, gnr, nn)
    return c
