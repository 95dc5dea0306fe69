"""The C-ABI library loads on a machine without a GPU, exports every symbol the public headers
declare, and its struct layouts match the reference SimParams layout and the Python mirrors."""
import ctypes as C
import os
import re
import subprocess

import pytest

from surfacegrowthconsole_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("surfacegrowth.h", "sgk.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            name = m.group(1)
            if name not in ("sizeof",):
                names.add(name)
    return names


def test_library_loads_without_gpu():
    lib = N.lib()
    assert lib.sgVersion().decode().startswith("surfacegrowth-mi355x")
    assert lib.sgk_device_count() >= 0


# Functions the headers declare with C++ linkage, as the reference has them: their exported names.
CXX_LINKAGE = {"AccumProps": "_Z10AccumPropsiP9SimParams"}


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    exported |= {k for k, v in CXX_LINKAGE.items() if v in exported}
    declared = declared_functions()
    assert declared, "header parse found nothing"
    assert declared <= exported, sorted(declared - exported)
    assert set(N.EXPORTED_SYMBOLS) == declared


def test_struct_layouts():
    N.check_layouts(N.lib())
    # Reference layout (SurfaceGrowth.h:69-186): 4-byte fields, 12-byte float3/uint3/int3/Prop,
    # szNameMe[3] plus one pad byte, szRdfPath[264], szBckup0/1[20], szDiffusePath[264]:
    # 36+20+24+16+8+24+24+48+104+4+24+20+16+16+20+36+36+28+264+40+20+20+264 = 1112 bytes.
    assert C.sizeof(N.Float3) == 12 and C.sizeof(N.Prop) == 12
    assert C.alignment(N.SimParams) == 4
    assert N.SimParams.szNameMe.offset % 4 == 0
    assert N.SimParams.temperatureU.offset == N.SimParams.szNameMe.offset + 4
    assert C.sizeof(N.SimParams) == 1112


def test_create_without_device_fails_cleanly():
    lib = N.lib()
    if lib.sgk_device_count() > 0:
        pytest.skip("a device is present")
    p = N.SimParams()
    h = C.c_void_p()
    rc = lib.sgk_create(C.byref(p), None, C.byref(h))
    assert rc != 0 and not h.value
    assert lib.sgk_last_error(None)
