"""Build libtinysph_hip.so in-tree with hipcc for gfx950.

    python -m tinysph_old_amd.build

Every HIP source is compiled with -ffp-contract=off (the FP64 path reproduces
the reference's unfused IEEE operations; the mixed-precision kernels opt into
contraction locally) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libtinysph_hip.so")
# HBM-counter calibration kernel (a plain copy of known size), its own library
PROBE_LIB = os.path.join(PKG, "libtsph_copy_probe.so")
PROBE_SOURCE = "tsph_copy_probe.hip"
OBJDIR = os.path.join(PKG, "_build")

HIP_SOURCES = ["tsph_api.hip", "tsph_cells.hip", "tsph_pairs.hip", "tsph_stages.hip", "tsph_ale.hip",
               "tsph_alemuscl.hip", "tsph_zones.hip", "tsph_integrators.hip", "tsph_io.hip"]
HOST_SOURCES = ["tsph_cases.cpp"]

ARCH = os.environ.get("TSPH_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [hipcc(), *COMMON, *HIP_FLAGS, "-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = ["g++", *COMMON, "-I/opt/rocm/include", "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


# example drivers written against include/tinysph_dropin.hpp (reference stage API)
DRIVER_SOURCE = "dropin_driver.cpp"
DRIVERS = {
    os.path.join(PKG, "tsph_dropin_mdbc"): [],
    os.path.join(PKG, "tsph_dropin_rale"): ["-DTSPH_DROPIN_RALE"],
    os.path.join(PKG, "tsph_dropin_verlet"): ["-DTSPH_DROPIN_VERLET"],
    os.path.join(PKG, "tsph_dropin_inflow"): ["-DTSPH_DROPIN_INFLOW"],
}


def _driver(exe: str, defs: list) -> None:
    inc = os.path.join(os.path.dirname(PKG), "include")
    tmp = exe + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", *defs, f"-I{inc}", os.path.join(CSRC, DRIVER_SOURCE),
           "-o", tmp, f"-L{PKG}", "-ltinysph_hip", f"-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, exe)


def _link(out: str, objs: list) -> None:
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = HIP_SOURCES + HOST_SOURCES + [PROBE_SOURCE]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    _link(LIB, objs[:-1])
    _link(PROBE_LIB, objs[-1:])
    for exe, defs in DRIVERS.items():
        _driver(exe, defs)
    if verbose:
        print("built", LIB, PROBE_LIB, *DRIVERS)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
