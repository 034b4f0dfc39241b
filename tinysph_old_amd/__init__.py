"""tinysph_old_amd: MI355X-native per-step particle pipeline of the tinySPH
2-D WCSPH solver (cell list, pair loops, integration) as HIP kernels behind a
C ABI (include/tinysph.h).  This package is the Python host layer."""
from ._lib import (  # noqa: F401
    CASE_ALE_DAM_BREAK, CASE_DAM_BREAK_BT, CASE_DAM_BREAK_MDBC_REF, CASE_INFLOW_CHANNEL, CASE_INFLOW_SLAB, INLET1, OUTLET1, CASE_DAM_BREAK_MDBC, CASE_STILL_WATER_DBC, FLUID, FORCES_BT, FORCES_DBC,
    FORCES_WCSPH, PREC_FP64, PREC_MIXED, WALL, WALL_BT, WALL_DBC, WALL_MDBC, SCHEME_RALE,
    SCHEME_SYMPLECTIC, MUSCL_LINEAR, MUSCL_MINMOD, MUSCL_NONE, TsphError, lib,
)
from .system import (  # noqa: F401
    BatchSystem, CaseData, ParticleSystem, autotune, batch_for_cases, build_case, build_ensemble,
    forces_variant, system_for_case,
)
