"""ctypes loader of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to check its result and by
bench.py's cpu_baseline leg.  The product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")


class OrcParams(C.Structure):
    _f, _i = C.c_float, C.c_int
    _fields_ = [
        ("nMol", _i), ("nMolMe", _i), ("iRegime", _i), ("region", _f * 3), ("cells", _i * 3),
        ("invWidth", _f * 3), ("rrNebr", _f), ("iNebrMax", _i), ("cellCap", _i),
        ("re", _f), ("fe", _f), ("rhoe", _f), ("alpha", _f), ("beta", _f), ("A", _f), ("B", _f),
        ("kappa", _f), ("lambda_", _f), ("Fn", _f * 4), ("F", _f * 4), ("eta", _f), ("Fe", _f),
        ("massMe", _f), ("rCutEam", _f), ("rrCutEam", _f), ("rei", _f), ("rhoei", _f),
        ("mu_r", _f), ("theta_0", _f), ("mu_theta", _f), ("mu_p", _f), ("z_0", _f), ("rCutC", _f),
        ("rrCutC", _f), ("sigmaLJ", _f), ("epsLJ", _f), ("rrCutLJ", _f),
        ("deltaT", _f), ("temperature", _f), ("kB", _f), ("gammaBerendsen", _f), ("rNebrShell", _f),
        ("stepThermostat", _i), ("stepEquil", _i), ("stepCool", _i),
        ("rangeRdf", _f), ("intervalRdf", _f), ("sizeHistRdf", _i), ("deltaF", _f),
        ("velMagDepos", _f), ("stepDeposit", _i), ("nMolToDeposit", _i), ("parScale", C.c_double * 4),
    ]


ND = 4  # directions per tangent call (ORC_ND)
PARAMS = ("eps", "sigma", "embed", "density")  # scale factors of the differentiated parameters


class OrcProps(C.Structure):
    _d = C.c_double
    _fields_ = [("vSum", _d * 3), ("vvSum", _d), ("vvMax", _d), ("uSum", _d), ("comSum", _d * 3),
                ("cmVelSum", _d * 3), ("frict", _d), ("bboxMin", _d * 3), ("bboxMax", _d * 3),
                ("kinEnergy", _d), ("potEnergy", _d), ("totEnergy", _d)]


class OrcRecord(C.Structure):
    _d = C.c_double
    _fields_ = [("step", C.c_int), ("rebuilt", C.c_int), ("nebrNow", C.c_int), ("nMolDeposited", C.c_int),
                ("kinEnergy", _d), ("potEnergy", _d), ("totEnergy", _d), ("vSum", _d * 3), ("vvSum", _d),
                ("vvMax", _d), ("uSum", _d), ("comSum", _d * 3), ("cmVelSum", _d * 3), ("frict", _d),
                ("bboxMin", _d * 3), ("bboxMax", _d * 3), ("dispHi", _d)]


OMP_LIB = os.path.join(HERE, "_build", "liboracle_omp.so")  # CPU-baseline timing build (FP32, OpenMP)
OMP64_LIB = os.path.join(HERE, "_build", "liboracle_omp64.so")  # FP64 oracle on OpenMP threads, bitwise lib()'s

_lib = None
_omp = None
_omp64 = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def timing_lib():
    """The FP32 restatement built for speed (-O3 -march=x86-64-v3, OpenMP per-atom loops): the CPU
    baseline's timing library.  Its *_f32 functions have the signatures of lib()'s."""
    global _omp
    if _omp is None:
        if not os.path.exists(OMP_LIB):
            build()
        _omp = C.CDLL(OMP_LIB)
        _declare(_omp, ("f32",))
    return _omp


def threaded_f64_lib():
    """The FP64 oracle with OpenMP per-atom loops, built with liboracle's flags: its *_f64 results
    are bitwise those of lib() (each atom's sums keep their order); for long FP64 runs."""
    global _omp64
    if _omp64 is None:
        if not os.path.exists(OMP64_LIB):
            build()
        _omp64 = C.CDLL(OMP64_LIB)
        _declare(_omp64, ("f64",))
        _omp64.orc_freeze_f64.restype = C.c_longlong
        _omp64.orc_freeze_f64.argtypes = [C.c_int]
        _omp64.orc_freeze_ok_f64.restype = C.c_int
    return _omp64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _declare(_lib, ("f32", "f64"))
        vp = C.c_void_p
        _lib.orc_diffusion_create.restype = vp
        _lib.orc_diffusion_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, vp]
        _lib.orc_diffusion_destroy.argtypes = [vp]
        _lib.orc_freeze_f64.restype = C.c_longlong
        _lib.orc_freeze_f64.argtypes = [C.c_int]
        _lib.orc_freeze_ok_f64.restype = C.c_int
        _lib.orc_tangent_window.restype = C.c_int
        _lib.orc_tangent_window.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        _lib.orc_tangent_forces.restype = None
        _lib.orc_tangent_forces.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    return _lib


def _declare(L, suffixes):
    """ctypes signatures of the precision-suffixed functions present in library L."""
    vp = C.c_void_p
    for sfx in suffixes:
        R = C.c_float if sfx == "f32" else C.c_double
        spec = {
            "orc_bin": (C.c_int, [vp, vp, vp, vp, vp]),
            "orc_nebr": (C.c_int, [vp, vp, vp, vp, vp, vp]),
            "orc_rho": (None, [vp, vp, vp, vp, vp]),
            "orc_forces": (None, [vp, vp, vp, vp, vp, vp, vp]),
            "orc_forces_virial": (None, [vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_embed": (R, [vp, R, vp]),
            "orc_pair": (None, [vp, R, vp]),
            "orc_leapfrog": (None, [vp, C.c_int, vp, vp, vp]),
            "orc_thermostat": (None, [vp, vp, C.c_double, C.c_int]),
            "orc_props": (None, [vp, vp, vp, vp, vp]),
            "orc_rdf": (C.c_int, [vp, vp, vp, vp, vp, vp]),
            "orc_diffusion_sample": (C.c_int, [vp, vp, vp]),
            "orc_run": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, vp, vp]),
            "orc_adhesion": (R, [vp, vp, vp, vp]),
            "orc_window": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp]),
        }
        for name, (res, args) in spec.items():
            fn = getattr(L, f"{name}_{sfx}", None)
            if fn is not None:
                fn.restype, fn.argtypes = res, args


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def params_from(sp, cell_cap=64):
    """orc_params with the float values of a SimParams (ctypes mirror)."""
    o = OrcParams()
    o.nMol, o.nMolMe, o.iRegime = sp.nMol, sp.nMolMe, sp.iRegime
    o.region[:] = [sp.region.x, sp.region.y, sp.region.z]
    o.cells[:] = [sp.cells.x, sp.cells.y, sp.cells.z]
    o.invWidth[:] = [sp.invWidth.x, sp.invWidth.y, sp.invWidth.z]
    o.rrNebr, o.iNebrMax, o.cellCap = sp.rrNebr, sp.iNebrMax, cell_cap
    for k in ("re", "fe", "rhoe", "alpha", "beta", "A", "B", "kappa", "eta", "Fe", "massMe", "rCutEam",
              "rrCutEam", "rei", "rhoei", "mu_r", "theta_0", "mu_theta", "mu_p", "z_0", "rCutC", "rrCutC",
              "sigmaLJ", "epsLJ", "rrCutLJ", "deltaT", "temperature", "kB", "gammaBerendsen", "rNebrShell",
              "stepThermostat", "stepEquil", "stepCool", "rangeRdf", "intervalRdf", "sizeHistRdf", "deltaF",
              "velMagDepos", "stepDeposit", "nMolToDeposit"):
        setattr(o, k, getattr(sp, k))
    o.lambda_ = sp.lambda_
    o.Fn[:] = list(sp.Fn)
    o.F[:] = list(sp.F)
    o.parScale[:] = [1.0, 1.0, 1.0, 1.0]
    return o


class Oracle:
    """One precision of the oracle ("f32" restatement or "f64" reference-of-record)."""

    def __init__(self, sp, prec="f64", cell_cap=64, library=None):
        self.l = library if library is not None else lib()
        self.sfx = prec
        self.dt = np.float32 if prec == "f32" else np.float64
        self.p = params_from(sp, cell_cap)
        self.n, self.n_metal = sp.nMol, sp.nMolMe
        self.ncell = sp.cells.x * sp.cells.y * sp.cells.z
        self.cap = cell_cap
        self.nbmax = sp.iNebrMax
        self.deposited = sp.nMolDeposited

    def f(self, name):
        return getattr(self.l, f"{name}_{self.sfx}")

    def arr(self, x):
        return np.ascontiguousarray(x, self.dt)

    def bin(self, r):
        r = self.arr(r)
        cell_of = np.zeros(self.n, np.int32)
        cnt = np.zeros(self.ncell, np.int32)
        slots = np.full(self.ncell * self.cap, -1, np.int32)
        err = self.f("orc_bin")(C.byref(self.p), _p(r), _p(cell_of), _p(cnt), _p(slots))
        return err, cell_of, cnt, slots.reshape(self.ncell, self.cap)

    def nebr(self, r, cnt, slots):
        r = self.arr(r)
        nn = np.zeros(self.n, np.int32)
        nbl = np.zeros(self.n * self.nbmax, np.int32)
        err = self.f("orc_nebr")(C.byref(self.p), _p(r), _p(np.ascontiguousarray(cnt, np.int32)),
                                 _p(np.ascontiguousarray(slots, np.int32)), _p(nn), _p(nbl))
        return err, nn, nbl

    def neighbour_list(self, r):
        e1, cell_of, cnt, slots = self.bin(r)
        e2, nn, nbl = self.nebr(r, cnt, slots)
        return e1 | e2, nn, nbl

    def rho(self, r, nn, nbl):
        r = self.arr(r)
        out = np.zeros(max(self.n_metal, 1), self.dt)
        self.f("orc_rho")(C.byref(self.p), _p(r), _p(nn), _p(nbl), _p(out))
        return out[: self.n_metal]

    def forces(self, r, nn, nbl, rho=None):
        """Returns (acc (N,3), energy (N,), friction (nMe,)) for positions r (N,4)."""
        r = self.arr(r).copy()
        if rho is None:
            rho = self.rho(r, nn, nbl)
        rho = np.ascontiguousarray(np.resize(rho, max(self.n_metal, 1)), self.dt)
        acc = np.zeros((self.n, 3), self.dt)
        fr = np.zeros(max(self.n_metal, 1), self.dt)
        self.f("orc_forces")(C.byref(self.p), _p(r), _p(nn), _p(nbl), _p(rho), _p(acc), _p(fr))
        return acc, r[:, 3].copy(), fr[: self.n_metal]

    def virial(self, r, nn, nbl, rho=None):
        """Per-atom pair virial 1/2 sum_k f_ik |d_ik|^2 of the EAM and LJ terms (N,)."""
        r = self.arr(r).copy()
        if rho is None:
            rho = self.rho(r, nn, nbl)
        rho = np.ascontiguousarray(np.resize(rho, max(self.n_metal, 1)), self.dt)
        acc = np.zeros((self.n, 3), self.dt)
        fr = np.zeros(max(self.n_metal, 1), self.dt)
        vir = np.zeros(self.n, self.dt)
        self.f("orc_forces_virial")(C.byref(self.p), _p(r), _p(nn), _p(nbl), _p(rho), _p(acc), _p(fr), _p(vir))
        return vir

    def embed(self, rho):
        d = (C.c_float if self.sfx == "f32" else C.c_double)()
        F = self.f("orc_embed")(C.byref(self.p), rho, C.byref(d))
        return F, d.value

    def pair(self, r):
        out = np.zeros(4, self.dt)
        self.f("orc_pair")(C.byref(self.p), r, _p(out))
        return out

    def leapfrog(self, part, r, v, a):
        r, v, a = self.arr(r).copy(), self.arr(v).copy(), self.arr(a)
        self.f("orc_leapfrog")(C.byref(self.p), part, _p(r), _p(v), _p(a))
        return r, v

    def thermostat(self, v, vv_sum, step):
        v = self.arr(v).copy()
        self.f("orc_thermostat")(C.byref(self.p), _p(v), vv_sum, step)
        return v

    def props(self, r, v, fric=None):
        out = OrcProps()
        fr = self.arr(np.resize(fric, max(self.n_metal, 1))) if fric is not None else None
        self.f("orc_props")(C.byref(self.p), _p(self.arr(r)), _p(self.arr(v)), _p(fr), C.byref(out))
        return {k: (list(getattr(out, k)) if hasattr(getattr(out, k), "__len__") else getattr(out, k))
                for k, _ in OrcProps._fields_}

    def rdf(self, r, cell_of, cnt, slots, hist=None):
        if hist is None:
            hist = np.zeros(self.p.sizeHistRdf, np.uint32)
        edge = self.f("orc_rdf")(C.byref(self.p), _p(self.arr(r)), _p(cell_of), _p(cnt),
                                 _p(np.ascontiguousarray(slots, np.int32)), _p(hist))
        return hist, edge

    def run(self, r, v, a, nsteps, step0=0, nebr_now=1, disp_hi=0.0, rebuild_every=0, deposited=None):
        """deposited: regime-1 deposition counter at the start (default: the SimParams value)."""
        r, v, a = self.arr(r).copy(), self.arr(v).copy(), self.arr(a).copy()
        recs = (OrcRecord * nsteps)()
        nb = C.c_int(nebr_now)
        dh = C.c_double(disp_hi)
        dep = C.c_int(self.deposited if deposited is None else deposited)
        err = self.f("orc_run")(C.byref(self.p), _p(r), _p(v), _p(a), nsteps, step0, C.byref(nb),
                                C.byref(dh), rebuild_every, recs, C.byref(dep))
        out = [{k: (list(getattr(x, k)) if hasattr(getattr(x, k), "__len__") else getattr(x, k))
                for k, _ in OrcRecord._fields_} for x in recs]
        return err, r, v, a, out

    # ---- adjoint observables and the reference gradient ------------------------------------------
    def set_scale(self, scale):
        """Scale factors of (LJ epsilon, LJ sigma, embedding F, density f); (1,1,1,1) = SimParams."""
        self.p.parScale[:] = [float(x) for x in scale]

    def adhesion(self, r, nn, nbl):
        return float(self.f("orc_adhesion")(C.byref(self.p), _p(self.arr(r)), _p(nn), _p(nbl)))

    def window(self, r, v, nsteps, step0=0, rebuild_every=0):
        """Adjoint window (orc_window): returns (J[2], r, v, a, err); J = (mean adhesion energy,
        final metal centre-of-mass height above z_0)."""
        r, v = self.arr(r).copy(), self.arr(v).copy()
        a = np.zeros((self.n, 3), self.dt)
        J = np.zeros(2, self.dt)
        err = self.f("orc_window")(C.byref(self.p), _p(r), _p(v), _p(a), nsteps, step0, rebuild_every, _p(J))
        return J.astype(np.float64), r, v, a, err

    def freeze(self, mode):
        """FP64 decision tape: "record" (clear, then record the next runs' rebuild steps, lists and
        F(rho) branches), "replay" (take them from the tape, from the start) or "off".  Returns the
        tape length."""
        return int(self.l.orc_freeze_f64({"off": 0, "record": 1, "replay": 2}[mode]))

    def freeze_ok(self):
        """True when the replay since freeze("replay") took exactly the recorded decisions."""
        return bool(self.l.orc_freeze_ok_f64())

    def tangent_window(self, r, v, nsteps, step0=0, rebuild_every=0, pdir=None, rdir=None, vdir=None):
        """FP64 window with forward-mode derivatives along up to ND directions: pdir (ND,4) seeds the
        parameter scale factors, rdir/vdir (ND,N,3) the initial positions/velocities.  Returns
        (J[2], dJ[2,ND], r, v, a, err)."""
        r = np.ascontiguousarray(r, np.float64).copy()
        v = np.ascontiguousarray(v, np.float64).copy()
        a = np.zeros((self.n, 3))
        J, dJ = np.zeros(2), np.zeros((2, ND))
        d = [None if x is None else np.ascontiguousarray(x, np.float64) for x in (pdir, rdir, vdir)]
        for x, shape in zip(d, ((ND, 4), (ND, self.n, 3), (ND, self.n, 3))):
            if x is not None:
                assert x.shape == shape, (x.shape, shape)
        err = self.l.orc_tangent_window(C.byref(self.p), _p(r), _p(v), _p(a), nsteps, step0, rebuild_every,
                                        _p(d[0]), _p(d[1]), _p(d[2]), _p(J), _p(dJ))
        return J, dJ, r, v, a, err

    def tangent_forces(self, r, nn, nbl, pdir=None, rdir=None):
        """Forces with forward-mode derivatives: (acc (N,3), dacc (ND,N,3), adhesion, dadh (ND,))."""
        r = np.ascontiguousarray(r, np.float64)
        acc = np.zeros((self.n, 3))
        dacc = np.zeros((ND, self.n, 3))
        adh, dadh = C.c_double(), np.zeros(ND)
        d = [None if x is None else np.ascontiguousarray(x, np.float64) for x in (pdir, rdir)]
        self.l.orc_tangent_forces(C.byref(self.p), _p(r), _p(np.ascontiguousarray(nn, np.int32)),
                                  _p(np.ascontiguousarray(nbl, np.int32)), _p(d[0]), _p(d[1]), _p(acc),
                                  _p(dacc), C.byref(adh), _p(dadh))
        return acc, dacc, adh.value, dadh
