"""The reference's own case drivers compile unchanged against the drop-in
header (include/tinysph_dropin.hpp) and link against libtinysph_hip.so.

Each driver (a struct SPH_simulation with INIT/RUN, e.g.
cases/damBreakLong_mDBC/SPH_simulationSymplectic.h) includes the reference's
src/*.h headers by name and its case's constants.h / geometry.h.  A shim
directory maps every src/*.h name to the drop-in header, the case directory
supplies constants.h and geometry.h, and the pair-loop variant is chosen on the
command line with -DTSPH_DROPIN_FORCES / -DTSPH_DROPIN_RALE (the reference chooses
it by #include).
The drivers are compiled and linked here only; they are never run, and nothing
of the reference leaves this machine.  Skipped where /root/reference is absent."""
import concurrent.futures as cf
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "tinysph_old_amd")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")),
                                reason="reference tree not present")

# the drivers of the hot-path cases: compiled to objects and linked
DRIVERS = [
    "damBreakLong_mDBC/SPH_simulationSymplectic.h",
    "damBreakLong_BT/SPH_simulationSymplectic.h",
    "inflowSlab_mDBC/SPH_simulationSymplectic.h",
    "ALESPH_dambreakLong/SPH_simulationINT.h",
    "ALESPH_dambreakLong/SPH_simulation.h",
]
# the other drivers whose headers the reference itself provides (syntax only)
OTHER_DRIVERS = [
    "ALESPH_dambreakLong/SPH_simulationSymplectic.h",
    "ALESPH_stillWater/SPH_simulation.h",
    "ALESPH_stillWater/SPH_simulationINT.h",
    "ALESPH_stillWater/SPH_simulationSymplectic.h",
    "damBreakLong_BT/SPH_simulationLeapFrog.h",
    "damBreakLong_BT/SPH_simulationVerlet.h",
    "damBreakLong_DBC/SPH_simulationLeapFrog.h",
    "damBreakLong_DBC/SPH_simulationSymplectic.h",
    "damBreakLong_DBC/SPH_simulationVerlet.h",
    "damBreakLong_mDBC/SPH_simulationLeapFrog.h",
    "damBreakLong_mDBC/SPH_simulationSymplectic_reinit.h",
    "damBreakLong_mDBC/SPH_simulationVerlet.h",
    "damBreakLong_mDBC/SPH_simulationVerlet_reinit.h",
    "flowDrop_BT/SPH_simulationLeapFrog.h",
    "flowDrop_BT/SPH_simulationSymplectic.h",
    "flowDrop_mDBC/SPH_simulationSymplectic.h",
    "inflowSlab_mDBC/SPH_simulationVerlet.h",
    "inflowTankBottom_mDBC/SPH_simulationLeapFrog.h",
    "inflowTankBottom_mDBC/SPH_simulationSymplectic.h",
    "stillWater_BT/SPH_simulationLeapFrog.h",
    "stillWater_BT/SPH_simulationSymplectic.h",
    "stillWater_BT/SPH_simulationVerlet.h",
    "stillWater_DBC/SPH_simulationLeapFrog.h",
    "stillWater_DBC/SPH_simulationVerlet.h",
    "stillWater_mDBC/SPH_simulationLeapFrog.h",
]
MAIN = """#include "{driver}"
int main() {{
  SPH_simulation SIMULATION;
  SIMULATION.PREP_SIMULATION_DATA();
  SIMULATION.INIT();
  SIMULATION.RUN();
  return 0;
}}
"""


def variant_flags(driver: str) -> list:
    """The pair-loop selection the driver's #include list makes in the reference."""
    case = driver.split("/")[0]
    if case.startswith("ALESPH"):
        return ["-DTSPH_DROPIN_RALE"]
    if case.endswith("_BT"):
        return ["-DTSPH_DROPIN_FORCES=TSPH_FORCES_BT"]
    if case.endswith("_DBC"):
        return ["-DTSPH_DROPIN_FORCES=TSPH_FORCES_DBC"]
    return []


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("shim")
    for name in sorted(os.listdir(os.path.join(REF, "src"))):
        if name.endswith(".h"):
            (d / name).write_text('#include "tinysph_dropin.hpp"\n')
    return str(d)


def compile_driver(driver: str, shim: str, out_dir: str, link: bool):
    case_dir = os.path.join(REF, "cases", os.path.dirname(driver))
    tag = driver.replace("/", "_").replace(".h", "")
    src = os.path.join(out_dir, tag + ".cpp")
    with open(src, "w") as f:
        f.write(MAIN.format(driver=os.path.join(REF, "cases", driver)))
    cmd = ["g++", "-std=c++17", *variant_flags(driver), f"-I{shim}", f"-I{INC}", f"-I{case_dir}"]
    if link:
        exe = os.path.join(out_dir, tag)
        cmd += ["-O1", src, "-o", exe, f"-L{PKG}", "-ltinysph_hip", "-Wl,-rpath-link,/opt/rocm/lib"]
    else:
        cmd += ["-fsyntax-only", src]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    return driver, res.returncode, res.stderr[-2000:]


def test_shim_covers_every_reference_header(shim):
    names = {n for n in os.listdir(os.path.join(REF, "src")) if n.endswith(".h")}
    assert names and names <= set(os.listdir(shim))


def test_hot_path_drivers_compile_and_link_unchanged(shim, tmp_path):
    if not os.path.exists(os.path.join(PKG, "libtinysph_hip.so")):
        pytest.fail("libtinysph_hip.so is missing: run `python -m tinysph_old_amd.build`")
    with cf.ThreadPoolExecutor(max_workers=5) as ex:
        results = list(ex.map(lambda d: compile_driver(d, shim, str(tmp_path), True), DRIVERS))
    bad = [(d, err) for d, rc, err in results if rc != 0]
    assert not bad, bad
    for d in DRIVERS:
        assert os.path.exists(tmp_path / d.replace("/", "_").replace(".h", ""))


def test_other_reference_drivers_compile_unchanged(shim, tmp_path):
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        results = list(ex.map(lambda d: compile_driver(d, shim, str(tmp_path), False), OTHER_DRIVERS))
    bad = [(d, err) for d, rc, err in results if rc != 0]
    assert not bad, bad
