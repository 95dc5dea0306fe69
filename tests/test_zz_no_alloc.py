"""sgk_forces_noalloc (the force evaluation that allocates nothing) against the allocating path it
stands for: sgk_create + upload + sgk_bin + sgk_build_nebr + sgk_compute_forces on a fresh context.
Same kernels in the same order, so accelerations, per-atom energies and friction must be equal bit
for bit.  Two calls in a row share the GPU's scratch array; the second has a smaller input, so it
must not see the first call's cells, lists or state.  Device buffers come from the HIP runtime the
library itself uses (hipMalloc / hipMemcpy through ctypes): test plumbing, not the call under test."""
import ctypes as C
import os

import numpy as np
import pytest

import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N

pytestmark = pytest.mark.gpu


def _systems():
    big = sg.config_system("small", "reference")  # 960 C + 201 Ni
    small = sg.make_system(regime=2, ux=3, uy=2, n_metal=13, temperature=600.0, geometry="sphere", seed=4)
    assert small.params.nMol < big.params.nMol
    return [("960 C + 201 Ni", big), ("192 C + 13 Ni", small)]


def _reference(s, prec):
    with sg.Engine(s.params, precision=prec) as eng:
        eng.upload(s.r.astype(np.float64) if prec == "mixed" else s.r)
        eng.bin()
        eng.build_nebr()
        eng.compute_forces()
        return eng.forces()


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libsurfacegrowth.so links (/opt/rocm/lib)."""

    def __init__(self):
        self.h = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def to_host(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _noalloc(lib, s, prec, hip):
    p = s.params
    cfg = N.SgkConfig()
    lib.sgk_default_config(C.byref(cfg))
    cfg.precision = {"mixed": N.SG_PREC_MIXED, "fp32": N.SG_PREC_FP32}[prec]
    n, nme = p.nMol, p.nMolMe
    dt = np.float64 if prec == "mixed" else np.float32
    r = np.ascontiguousarray(s.r.astype(np.float64).astype(dt))
    acc0 = np.full(3 * n, np.nan, dt)
    fric0 = np.full(max(nme, 1), np.nan, np.float32)
    flags0 = np.full(1, -1, np.int32)
    pos, acc, fric, flags = hip.to_device(r), hip.to_device(acc0), hip.to_device(fric0), hip.to_device(flags0)
    need = lib.sgk_noalloc_scratch_bytes(C.byref(p), C.byref(cfg))
    assert 0 < need <= N.SGK_NOALLOC_SCRATCH_BYTES
    rc = lib.sgk_forces_noalloc(C.byref(p), C.byref(cfg), pos, None, acc, fric, flags, None)
    assert rc == 0, lib.sgk_noalloc_last_error().decode()
    hip.sync()
    assert int(hip.to_host(flags, flags0)[0]) == 0
    a = hip.to_host(acc, acc0).astype(np.float64).reshape(3, n).T
    e = hip.to_host(pos, r)[:, 3].astype(np.float64)
    f = hip.to_host(fric, fric0)[:nme].astype(np.float64)
    return a, e, f


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_noalloc_equals_the_allocating_path_on_two_calls(gpu, prec):
    lib = N.lib()
    hip = Hip()
    try:
        for label, s in _systems():
            print(f"[noalloc] {prec} {label}: allocating path", flush=True)
            ra, re_, rf = _reference(s, prec)
            print(f"[noalloc] {prec} {label}: sgk_forces_noalloc", flush=True)
            a, e, f = _noalloc(lib, s, prec, hip)
            assert np.array_equal(a, ra), np.abs(a - ra).max()
            assert np.array_equal(e, re_), np.abs(e - re_).max()
            assert np.array_equal(f, rf), np.abs(f - rf).max()
    finally:
        hip.free()


def test_noalloc_rejects_a_system_beyond_the_scratch(gpu):
    """The medium workload (44,353 atoms, 60 x 54 x 11 cells) needs more than the scratch holds: the
    call returns SGK_E_UNSUPPORTED with the reason and enqueues nothing."""
    lib = N.lib()
    s = sg.config_system("medium", "config")
    p = s.params
    cfg = N.SgkConfig()
    lib.sgk_default_config(C.byref(cfg))
    cfg.precision = N.SG_PREC_MIXED
    assert lib.sgk_noalloc_scratch_bytes(C.byref(p), C.byref(cfg)) > N.SGK_NOALLOC_SCRATCH_BYTES
    hip = Hip()
    try:
        buf = hip.to_device(np.zeros(8))
        rc = lib.sgk_forces_noalloc(C.byref(p), C.byref(cfg), buf, None, buf, None, None, None)
    finally:
        hip.free()
    assert rc == N.SGK_E_UNSUPPORTED
    assert b"limit without allocation" in lib.sgk_noalloc_last_error()
