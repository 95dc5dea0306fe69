"""The reference's own host driver builds against this library (CPU only, compile and link, not run).

SurfaceGrowthConsole/SurfaceGrowth.cpp and SurfaceGrowthProto.h are read from the reference tree at
test time, edited in a temporary directory with the edits CPP_EDITS and PROTO_EDITS list below (the
driver's CUDA device queries, its Windows entry point and key wait, and its header include), compiled
with g++ against include/surfacegrowth_driver.h, and linked against libsurfacegrowth.so.  Nothing of
the reference is written into the repository and the linked program is never run.  Skipped where
the reference tree is absent.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/SurfaceGrowthConsole"
LIBDIR = os.path.join(ROOT, "surfacegrowthconsole_amd", "lib")

# (regex, replacement): the driver edits, in SurfaceGrowth.cpp order.
CPP_EDITS = [
    # .cpp:12 - the CUDA device-properties global goes
    (r"cudaDeviceProp g_hDeviceProp;[^\n]*\n", ""),
    # .cpp:109 - device count from the library
    (r"cudaGetDeviceCount\(&deviceCount\);", "deviceCount = sgDeviceCount();"),
    # .cpp:114-122 - the CUDA compute-capability test goes
    (r"\n[ \t]*// Get cuda device properties\.\n[ \t]*cudaGetDeviceProperties\(&g_hDeviceProp, 0\);"
     r".*?return 1;[^\n]*\n[ \t]*\}", ""),
    # .cpp:317 - the grid limit is the engine's, not one CUDA block per cell
    (r"\(uint\)g_hDeviceProp\.maxGridSize\[0\]", "sgMaxCellGrid()"),
    # .cpp:104, 159 - the Windows console entry point and key wait
    (r"int _tmain\(int argc, _TCHAR\* argv\[\]\)", "int main(int argc, char* argv[])"),
    (r"_getch\(\);", "getchar();"),
]
PROTO_EDITS = [
    # SurfaceGrowthProto.h:11 - the library's driver header in place of the CUDA one
    (r'#include "SurfaceGrowth.h"', '#include "surfacegrowth_driver.h"'),
]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present on this machine")


def _edit(path, edits):
    text = open(path).read()
    for pat, rep in edits:
        text, n = re.subn(pat, rep, text, flags=re.S)
        assert n == 1, f"edit {pat!r} matched {n} times in {os.path.basename(path)}"
    return text


@pytest.fixture(scope="module")
def driver_obj(tmp_path_factory):
    d = tmp_path_factory.mktemp("driver")
    (d / "SurfaceGrowth.cpp").write_text(_edit(os.path.join(REF, "SurfaceGrowth.cpp"), CPP_EDITS))
    (d / "SurfaceGrowthProto.h").write_text(_edit(os.path.join(REF, "SurfaceGrowthProto.h"), PROTO_EDITS))
    obj = d / "SurfaceGrowth.o"
    r = subprocess.run(["g++", "-std=gnu++17", "-O1", "-c", f"-I{d}", f"-I{os.path.join(ROOT, 'include')}",
                        str(d / "SurfaceGrowth.cpp"), "-o", str(obj)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    errors = [ln for ln in r.stderr.splitlines() if "error" in ln]
    assert not errors, errors
    return d, obj


def test_reference_driver_compiles_with_the_listed_edits(driver_obj):
    d, obj = driver_obj
    assert obj.exists()
    # no CUDA name is left in the edited driver
    text = (d / "SurfaceGrowth.cpp").read_text()
    assert "cuda" not in re.sub(r"//[^\n]*|/\*.*?\*/|\"[^\"\n]*\"", "", text, flags=re.S).replace("CudaInitW", "")


def test_reference_driver_links_against_the_library(driver_obj):
    """Every symbol the driver leaves undefined - the four wrappers, AccumProps with C++ linkage and
    the two device queries - resolves in libsurfacegrowth.so (the program is linked, not run)."""
    d, obj = driver_obj
    so = os.path.join(LIBDIR, "libsurfacegrowth.so")
    assert os.path.exists(so)
    nm = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout
    undefined = {ln.split()[-1] for ln in nm.splitlines() if ln.strip()}
    for sym in ("CudaInitW", "SetParametersW", "InitCoordsW", "DoComputationsW", "sgDeviceCount",
                "sgMaxCellGrid", "_Z10AccumPropsiP9SimParams"):
        assert sym in undefined, sym
    exe = d / "surfacegrowth_reference_driver"
    r = subprocess.run(["g++", str(obj), f"-L{LIBDIR}", "-lsurfacegrowth", f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exe.unlink()
