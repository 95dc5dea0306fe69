"""ctypes bindings of libsurfacegrowth.so (include/surfacegrowth.h and include/sgk.h).

The structs below mirror the C layouts field by field; ``check_layouts()`` compares their sizes
with the sizes the library reports, so a drifted mirror fails loudly instead of corrupting memory.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsurfacegrowth.so")
MAX_PATH = 264


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class UInt3(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Prop(C.Structure):
    _fields_ = [("val", C.c_float), ("sum", C.c_float), ("sum2", C.c_float)]


_f = C.c_float
_i = C.c_int


class SimParams(C.Structure):
    """Mirror of SimParams (reference SurfaceGrowthConsole/SurfaceGrowth.h:69-186)."""

    _fields_ = [
        ("region", Float3), ("vSum", Float3), ("initUcell", UInt3),
        ("deltaT", _f), ("rCutC", _f), ("rrCutC", _f), ("temperature", _f), ("velMag", _f),
        ("moreCycles", _i), ("stepAvg", _i), ("stepEquil", _i), ("stepCount", _i), ("stepPdb", _i),
        ("stepLimit", _i), ("bPdb", _i), ("bResult", _i), ("nMol", _i), ("nMolMe", _i),
        ("gridSize", C.c_uint), ("blockSize", C.c_uint), ("cells", Int3), ("invWidth", Float3),
        ("rNebrShell", _f), ("rrNebr", _f), ("iNebrMax", _i), ("maxMol", _i), ("nebrNow", _i),
        ("dispHi", _f), ("kinEnergy", Prop), ("totEnergy", Prop), ("potEnergy", Prop), ("oneStep", Prop),
        ("a", _f), ("re", _f), ("fe", _f), ("rhoe", _f), ("alpha", _f), ("beta", _f), ("A", _f), ("B", _f),
        ("kappa", _f), ("lambda_", _f), ("Fn", _f * 4), ("F", _f * 4), ("eta", _f), ("Fe", _f),
        ("massMe", _f), ("density", _f), ("rCutEam", _f), ("rrCutEam", _f), ("rei", _f), ("rhoei", _f),
        ("szNameMe", C.c_char * 3),
        ("temperatureU", _f), ("kB", _f), ("enU", _f), ("lengthU", _f), ("forceU", _f), ("massU", _f),
        ("mu_r", _f), ("theta_0", _f), ("mu_theta", _f), ("mu_p", _f), ("z_0", _f),
        ("stepThermostat", _i), ("gammaBerendsen", _f), ("stepCool", _i), ("iRegime", _i),
        ("sigmaLJ", _f), ("epsLJ", _f), ("rCutLJ", _f), ("rrCutLJ", _f),
        ("velMagDepos", _f), ("randSeedP", _i), ("stepDeposit", _i), ("nMolDeposited", _i),
        ("nMolToDeposit", _i), ("centerOfMass", Prop), ("cmVel", Prop), ("frictForce", Prop),
        ("deltaF", _f), ("particleSize", Float3), ("cellShiftZ", _i), ("shear", _f), ("totalShear", _f),
        ("initUcellMeX", _i), ("initUcellMeY", _i),
        ("rangeRdf", _f), ("limitRdf", _i), ("sizeHistRdf", _i), ("stepRdf", _i), ("countRdf", _i),
        ("intervalRdf", _f), ("bRdf", _i), ("szRdfPath", C.c_char * MAX_PATH),
        ("szBckup0", C.c_char * 20), ("szBckup1", C.c_char * 20), ("bBckup", _i), ("bStartBckup", _i),
        ("stepBckup", _i), ("totalTime", _f), ("iMaterial", _i),
        ("nValDiffuse", _i), ("nBuffDiffuse", _i), ("stepDiffuse", _i), ("limitDiffuseAv", _i),
        ("countDiffuseAv", _i), ("szDiffusePath", C.c_char * MAX_PATH),
    ]


class SgOptions(C.Structure):
    _fields_ = [
        ("rNebrShell", _f), ("iNebrMax", _i), ("blockSize", _i), ("stepThermostat", _i),
        ("gammaBerendsen", _f), ("limitRdf", _i), ("sizeHistRdf", _i), ("nValDiffuse", _i),
        ("nBuffDiffuse", _i), ("stepDiffuse", _i), ("limitDiffuseAv", _i), ("stepBckup", _i),
        ("seed", _i), ("geometry", _i), ("gapAngstrom", _f), ("velocityRule", _i), ("precision", _i),
        ("rebuildEvery", _i), ("rangeRdfAngstrom", _f), ("stepsPerSync", _i), ("useGraphs", _i),
        ("metalThermostatAll", _i), ("tauPs", _f),
    ]


class SgInput(C.Structure):
    _fields_ = [
        ("iRegime", _i), ("iMaterial", _i), ("ux", _i), ("uy", _i), ("uz", _i), ("temperature", _f),
        ("nMolMe", _i), ("epsilon", _f), ("sigma", _f), ("stepLimit", _i), ("deltaT", _f),
        ("stepAvg", _i), ("stepEquil", _i), ("deltaF", _f), ("stepCool", _i), ("deposEnergy", _f),
        ("nMolToDeposit", _i), ("stepDeposit", _i), ("stepPdb", _i), ("stepRdf", _i), ("bBckup", _i),
        ("bStartBckup", _i),
    ]


class TBuf(C.Structure):
    _fields_ = [("orgR", Float3), ("rTrue", Float3), ("rrDiffuse", C.POINTER(C.c_float)), ("count", _i)]


class Material(C.Structure):
    _fields_ = [("szName", C.c_char_p)] + [(n, _f) for n in (
        "re", "fe", "rhoe", "alpha", "beta", "A", "B", "kappa", "lambda_")] + [
        ("Fn", _f * 4), ("F", _f * 4), ("eta", _f), ("Fe", _f), ("massMe", _f), ("density", _f)]


class SgkTuneEntry(C.Structure):
    _fields_ = [("metal_lanes", C.c_int), ("xcd_map", C.c_int), ("us", C.c_double)]


class SgkBatchTuneEntry(C.Structure):
    _fields_ = [("metal_lo", C.c_int), ("metal_hi", C.c_int), ("replicas", C.c_int), ("atoms", C.c_int),
                ("metal_lanes", C.c_int), ("us", C.c_double)]


class SgkConfig(C.Structure):
    _fields_ = [
        ("precision", _i), ("nl_capacity", _i), ("cell_capacity", _i), ("lj_capacity", _i),
        ("close_capacity", _i), ("close_margin", _f), ("metal_lanes", _i), ("block_size", _i),
        ("keep_full_nl", _i), ("rebuild_every", _i), ("records_capacity", _i), ("use_graphs", _i),
        ("metal_thermostat_all", _i), ("pair_fp64", _i), ("xcd_map", _i), ("bin_group", _i),
        ("nl_scheme", _i), ("rdf_complete", _i), ("sync_rebuilds", _i),
    ]


_d = C.c_double


class SgkRecord(C.Structure):
    _fields_ = [
        ("step", _i), ("rebuilt", _i), ("nebrNow", _i), ("errors", _i),
        ("kinEnergy", _d), ("potEnergy", _d), ("totEnergy", _d), ("vSum", _d * 3), ("vvSum", _d),
        ("vvMax", _d), ("uSum", _d), ("comSum", _d * 3), ("cmVelSum", _d * 3), ("frict", _d),
        ("bboxMin", _d * 3), ("bboxMax", _d * 3), ("dispHi", _d), ("shear", _d), ("shearCount", _d),
        ("nMolDeposited", _i), ("pad", _i), ("virial", _d),
    ]


class SgkKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", _i), ("ms", _d)]


class SgkAdjointInfo(C.Structure):
    _fields_ = [("nsteps", _i), ("checkpoint_every", _i), ("segments", _i), ("list_copies", _i),
                ("mismatches", _i), ("pad", _i), ("bytes", C.c_longlong), ("forward_ms", _d),
                ("backward_ms", _d), ("obs", _d * 2)]


SGK_PARAMS = ("epsilon", "sigma", "embedding", "density")  # SGK_PAR_*: scale factors differentiated
SGK_OBSERVABLES = ("adhesion", "height")                    # SGK_OBS_*


SG_PREC_FP32, SG_PREC_MIXED = 0, 1
SGK_E_ARG, SGK_E_HIP, SGK_E_NOMEM, SGK_E_OVERFLOW, SGK_E_UNSUPPORTED, SGK_E_STATE, SGK_E_DISPLACEMENT = range(-1, -8, -1)
SGK_NOALLOC_SCRATCH_BYTES = 64 << 20  # sgk.h
SG_GEOM_REFERENCE, SG_GEOM_SPHERE, SG_GEOM_ICOSAHEDRON = 0, 1, 2
SG_VEL_REFERENCE, SG_VEL_MAXWELL = 0, 1
SGK_FLAGS = {1: "cell overflow", 2: "nl overflow", 4: "lj overflow", 8: "close overflow",
             16: "unbinned", 32: "rdf edge", 64: "displacement"}

# Every symbol the two public headers declare.
EXPORTED_SYMBOLS = [
    "CudaInitW", "SetParametersW", "InitCoordsW", "DoComputationsW",
    "sgDefaultOptions", "sgSetOptions", "sgGetOptions", "sgDefaultInput", "sgReadInputFile",
    "sgSetParams", "sgInitRand", "sgRandR", "sgVRand", "sgInitVels", "sgOpenResultFile",
    "sgWriteBackup", "sgReadBackup", "sgVersion", "sgStructSizes",
    "sgDiffusionCreate", "sgDiffusionEval", "sgDiffusionDestroy", "sgDeviceCount", "sgMaxCellGrid",
    "AccumProps",
    "sgk_default_config", "sgk_create", "sgk_destroy", "sgk_last_error", "sgk_device_count",
    "sgk_upload", "sgk_download", "sgk_upload_f64", "sgk_download_f64", "sgk_set_run_state",
    "sgk_set_shear_state",
    "sgk_bin", "sgk_get_cells", "sgk_build_nebr", "sgk_build_nl_order", "sgk_get_nebr", "sgk_get_carbon_lists",
    "sgk_compute_rho", "sgk_compute_forces", "sgk_get_rho", "sgk_get_forces", "sgk_leapfrog",
    "sgk_thermostat", "sgk_eval_props", "sgk_rdf_clear", "sgk_rdf_accumulate", "sgk_rdf_get",
    "sgk_run", "sgk_last_run_ms", "sgk_set_profiling", "sgk_kernel_stats", "sgk_device_bytes",
    "sgk_struct_sizes", "sgk_metal_lanes", "sgk_autotune", "sgk_set_deposited",
    "sgk_batch_create", "sgk_batch_destroy", "sgk_batch_last_error", "sgk_batch_upload_f64",
    "sgk_batch_download_f64", "sgk_batch_set_run_state", "sgk_batch_run", "sgk_batch_last_run_ms",
    "sgk_batch_rdf_get", "sgk_batch_rdf_clear", "sgk_batch_device_bytes", "sgk_batch_metal_lanes",
    "sgk_batch_atoms", "sgk_batch_replica_lanes", "sgk_batch_set_lanes", "sgk_batch_autotune",
    "sgk_adjoint_forward", "sgk_adjoint_backward", "sgk_adjoint_get_info", "sgk_force_vjp",
    "sgk_adhesion_energy", "sgk_prepare_run", "sgk_diffusion_enable", "sgk_diffusion_get", "sgk_forces_noalloc", "sgk_noalloc_scratch_bytes", "sgk_noalloc_last_error",
]

_lib = None


def _proto(lib):
    P, vp, dp, ip = C.POINTER, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    spec = {
        "sgDefaultOptions": (None, [P(SgOptions)]),
        "sgSetOptions": (None, [P(SgOptions)]),
        "sgGetOptions": (None, [P(SgOptions)]),
        "sgDefaultInput": (None, [P(SgInput)]),
        "sgReadInputFile": (_i, [C.c_char_p, P(SgInput), P(SgOptions)]),
        "sgOpenResultFile": (vp, [P(SimParams), _f, C.c_char_p]),
        "sgDiffusionCreate": (vp, [P(SimParams)]),
        "sgDiffusionEval": (_i, [vp, P(SimParams), P(Float3), vp]),
        "sgDiffusionDestroy": (None, [vp]),
        "sgSetParams": (_i, [P(SgInput), P(SimParams), C.c_longlong]),
        "sgInitRand": (None, [_i, P(SimParams)]),
        "sgRandR": (_f, [P(SimParams)]),
        "sgVRand": (None, [P(Float3), P(SimParams)]),
        "sgInitVels": (None, [P(SimParams), vp]),
        "sgWriteBackup": (_i, [C.c_char_p, P(SimParams), vp, vp, vp]),
        "sgReadBackup": (_i, [C.c_char_p, P(SimParams), vp, vp, vp, _i]),
        "sgVersion": (C.c_char_p, []),
        "sgStructSizes": (_i, [ip, _i]),
        "InitCoordsW": (C.c_char_p, [vp, vp, P(SimParams)]),
        "DoComputationsW": (C.c_char_p, [vp, vp, vp, P(SimParams), vp, C.c_char_p]),
        "SetParametersW": (None, [P(SimParams)]),
        "sgk_default_config": (None, [P(SgkConfig)]),
        "sgk_create": (_i, [P(SimParams), P(SgkConfig), P(vp)]),
        "sgk_diffusion_enable": (_i, [vp, _i]),
        "sgk_prepare_run": (_i, [vp, _i, _i]),
        "sgk_diffusion_get": (_i, [vp, vp, _i, P(_i)]),
        "sgk_forces_noalloc": (_i, [P(SimParams), P(SgkConfig), vp, vp, vp, vp, vp, vp]),
        "sgk_noalloc_scratch_bytes": (C.c_longlong, [P(SimParams), P(SgkConfig)]),
        "sgk_noalloc_last_error": (C.c_char_p, []),
        "sgk_destroy": (_i, [vp]),
        "sgk_last_error": (C.c_char_p, [vp]),
        "sgk_device_count": (_i, []),
        "sgk_upload": (_i, [vp, vp, vp, vp]),
        "sgk_download": (_i, [vp, vp, vp, vp]),
        "sgk_upload_f64": (_i, [vp, vp, vp, vp]),
        "sgk_download_f64": (_i, [vp, vp, vp, vp]),
        "sgk_set_run_state": (_i, [vp, _i, _i, C.c_double]),
        "sgk_set_shear_state": (_i, [vp, C.c_double, C.c_double, C.c_double]),
        "sgk_bin": (_i, [vp]),
        "sgk_get_cells": (_i, [vp, vp, vp, vp]),
        "sgk_build_nebr": (_i, [vp]),
        "sgk_build_nl_order": (_i, [vp]),
        "sgk_get_nebr": (_i, [vp, vp, vp]),
        "sgk_get_carbon_lists": (_i, [vp, vp, vp, vp, vp, vp, vp]),
        "sgk_compute_rho": (_i, [vp]),
        "sgk_compute_forces": (_i, [vp]),
        "sgk_get_rho": (_i, [vp, vp]),
        "sgk_get_forces": (_i, [vp, vp, vp, vp]),
        "sgk_leapfrog": (_i, [vp, _i]),
        "sgk_thermostat": (_i, [vp, _i, C.c_double]),
        "sgk_eval_props": (_i, [vp, P(SgkRecord)]),
        "sgk_rdf_clear": (_i, [vp]),
        "sgk_rdf_accumulate": (_i, [vp]),
        "sgk_rdf_get": (_i, [vp, vp]),
        "sgk_run": (_i, [vp, _i, _i, vp]),
        "sgk_last_run_ms": (C.c_double, [vp]),
        "sgk_set_profiling": (_i, [vp, _i]),
        "sgk_kernel_stats": (_i, [vp, P(SgkKernelStat), _i]),
        "sgk_device_bytes": (C.c_longlong, [vp]),
        "sgk_struct_sizes": (_i, [ip, _i]),
        "sgk_metal_lanes": (_i, [vp]),
        "sgk_autotune": (_i, [vp, _i, vp, _i, vp]),
        "sgk_set_deposited": (_i, [vp, _i]),
        "sgk_batch_create": (_i, [_i, vp, P(SgkConfig), P(vp)]),
        "sgk_batch_destroy": (_i, [vp]),
        "sgk_batch_last_error": (C.c_char_p, [vp]),
        "sgk_batch_upload_f64": (_i, [vp, _i, vp, vp, vp]),
        "sgk_batch_download_f64": (_i, [vp, _i, vp, vp, vp]),
        "sgk_batch_set_run_state": (_i, [vp, _i, _i, _i, C.c_double]),
        "sgk_batch_run": (_i, [vp, _i, _i, vp]),
        "sgk_batch_last_run_ms": (C.c_double, [vp]),
        "sgk_batch_rdf_get": (_i, [vp, _i, vp]),
        "sgk_batch_rdf_clear": (_i, [vp]),
        "sgk_batch_device_bytes": (C.c_longlong, [vp]),
        "sgk_batch_metal_lanes": (_i, [vp]),
        "sgk_batch_atoms": (_i, [vp]),
        "sgk_batch_replica_lanes": (_i, [vp, _i]),
        "sgk_batch_set_lanes": (_i, [vp, vp]),
        "sgk_batch_autotune": (_i, [vp, _i, P(SgkBatchTuneEntry), _i, P(_i)]),
        "sgk_adjoint_forward": (_i, [vp, _i, _i, vp, vp]),
        "sgk_adjoint_backward": (_i, [vp, vp, vp, vp, vp]),
        "sgk_adjoint_get_info": (_i, [vp, P(SgkAdjointInfo)]),
        "sgk_force_vjp": (_i, [vp, vp, C.c_double, vp, vp]),
        "sgk_adhesion_energy": (_i, [vp, dp]),
    }
    for name, (res, args) in spec.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def library_built():
    return os.path.exists(LIB_PATH)


def lib():
    """Load libsurfacegrowth.so (built in-tree by __graft_entry__.build()); raises if missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _proto(_lib)
        check_layouts(_lib)
    return _lib


def check_layouts(l):
    out = (C.c_int * 8)()
    n = l.sgStructSizes(out, 8)
    mine = [C.sizeof(SimParams), C.sizeof(SgOptions), C.sizeof(SgInput), C.sizeof(TBuf), C.sizeof(Material)]
    if list(out[:n]) != mine[:n]:
        raise RuntimeError(f"ABI struct size mismatch: library {list(out[:n])} vs bindings {mine[:n]}")
    n = l.sgk_struct_sizes(out, 8)
    mine = [C.sizeof(SgkConfig), C.sizeof(SgkRecord), C.sizeof(SgkKernelStat), C.sizeof(SgkAdjointInfo)]
    if list(out[:n]) != mine[:n]:
        raise RuntimeError(f"sgk struct size mismatch: library {list(out[:n])} vs bindings {mine[:n]}")
