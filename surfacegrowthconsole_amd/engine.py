"""Python front end of the MI355X engine: builds systems through the host C-ABI and drives the
sgk_* stages on the GPU.  Every call goes through libsurfacegrowth.so; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _native as N

PRECISIONS = {"fp32": N.SG_PREC_FP32, "mixed": N.SG_PREC_MIXED}
GEOMETRIES = {"reference": N.SG_GEOM_REFERENCE, "sphere": N.SG_GEOM_SPHERE, "icosahedron": N.SG_GEOM_ICOSAHEDRON}
VELOCITIES = {"reference": N.SG_VEL_REFERENCE, "maxwell": N.SG_VEL_MAXWELL}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class System:
    """Parameters plus initial state built by the library's host code (SetParams, InitCoordsW,
    InitVels of the reference driver, SurfaceGrowth.cpp:189-621)."""

    def __init__(self, params, r, v, a, options):
        self.params = params
        self.r = r  # float32 (N,4): x, y, z, E
        self.v = v  # float32 (N,3)
        self.a = a  # float32 (N,3)
        self.options = options

    @property
    def n(self):
        return self.params.nMol

    @property
    def n_metal(self):
        return self.params.nMolMe


def make_options(**kw):
    lib = N.lib()
    o = N.SgOptions()
    lib.sgDefaultOptions(C.byref(o))
    for k, v in kw.items():
        if v is None:
            continue
        if k == "geometry":
            v = GEOMETRIES[v]
        elif k == "velocityRule":
            v = VELOCITIES[v]
        elif k == "precision":
            v = PRECISIONS[v]
        setattr(o, k, v)
    return o


def make_system(regime=2, material=3, ux=5, uy=6, uz=0, n_metal=201, temperature=300.0, epsilon=0.023,
                sigma=2.85, step_limit=1000, dt=0.005, step_avg=100, step_equil=0, delta_f=0.0, step_cool=0,
                step_pdb=100000, step_rdf=50, seed=1, geometry="sphere", gap=2.5, velocities="reference",
                depos_energy=None, n_to_deposit=None, step_deposit=None, **options):
    """Build a system with the reference's parameter arithmetic; extension options by keyword
    (SgOptions field names, e.g. tauPs=0.1, rangeRdfAngstrom=10.0, sizeHistRdf=200)."""
    lib = N.lib()
    inp = N.SgInput()
    lib.sgDefaultInput(C.byref(inp))
    inp.iRegime, inp.iMaterial, inp.ux, inp.uy, inp.uz = regime, material, ux, uy, uz
    inp.temperature, inp.nMolMe, inp.epsilon, inp.sigma = temperature, n_metal, epsilon, sigma
    inp.stepLimit, inp.deltaT, inp.stepAvg, inp.stepEquil = step_limit, dt, step_avg, step_equil
    inp.deltaF, inp.stepCool, inp.stepPdb, inp.stepRdf = delta_f, step_cool, step_pdb, step_rdf
    if depos_energy is not None:  # regime 1 rows 16-18
        inp.deposEnergy = depos_energy
    if n_to_deposit is not None:
        inp.nMolToDeposit = n_to_deposit
    if step_deposit is not None:
        inp.stepDeposit = step_deposit
    opts = make_options(seed=seed, geometry=geometry, gapAngstrom=gap, velocityRule=velocities, **options)
    lib.sgSetOptions(C.byref(opts))
    p = N.SimParams()
    if not lib.sgSetParams(C.byref(inp), C.byref(p), 0):
        raise RuntimeError("sgSetParams rejected the system")
    n = p.nMol
    r = np.zeros((n, 4), np.float32)
    v = np.zeros((n, 3), np.float32)
    a = np.zeros((n, 3), np.float32)
    err = lib.InitCoordsW(None, _ptr(r), C.byref(p))
    if err:
        raise RuntimeError(err.decode())
    lib.sgInitVels(C.byref(p), _ptr(v))
    p.nebrNow = 1
    return System(p, r, v, a, opts)


class Engine:
    """One sgk_ctx: device buffers for one system."""

    def __init__(self, params, precision="fp32", keep_full_nl=False, rebuild_every=0, metal_lanes=0,
                 records_capacity=1024, use_graphs=True, metal_thermostat_all=False, **cfg):
        self.lib = N.lib()
        self.params = params
        c = N.SgkConfig()
        self.lib.sgk_default_config(C.byref(c))
        c.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        c.keep_full_nl = int(keep_full_nl)
        c.rebuild_every = rebuild_every
        c.metal_lanes = metal_lanes
        c.records_capacity = records_capacity
        c.use_graphs = int(use_graphs)
        c.metal_thermostat_all = int(metal_thermostat_all)
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        h = C.c_void_p()
        rc = self.lib.sgk_create(C.byref(params), C.byref(c), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"sgk_create failed ({rc}): {self.lib.sgk_last_error(None).decode()}")
        self.h = h
        self.n = params.nMol
        self.n_metal = params.nMolMe
        self.ncell = params.cells.x * params.cells.y * params.cells.z

    # -- helpers ---------------------------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.sgk_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sgk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def metal_lanes(self):
        return self.lib.sgk_metal_lanes(self.h)

    # -- state -----------------------------------------------------------------------------
    def upload(self, r, v=None, a=None):
        """float32 arrays (N,4), (N,3), (N,3) in the reference layout, or float64 (full state)."""
        if r.dtype == np.float64:
            args = [np.ascontiguousarray(x, np.float64) if x is not None else None for x in (r, v, a)]
            self._check(self.lib.sgk_upload_f64(self.h, *[_ptr(x) for x in args]), "sgk_upload_f64")
        else:
            args = [np.ascontiguousarray(x, np.float32) if x is not None else None for x in (r, v, a)]
            self._check(self.lib.sgk_upload(self.h, *[_ptr(x) for x in args]), "sgk_upload")

    def download(self):
        r = np.zeros((self.n, 4))
        v = np.zeros((self.n, 3))
        a = np.zeros((self.n, 3))
        self._check(self.lib.sgk_download_f64(self.h, _ptr(r), _ptr(v), _ptr(a)), "sgk_download_f64")
        return r, v, a

    def set_run_state(self, step=0, nebr_now=1, disp_hi=0.0):
        self._check(self.lib.sgk_set_run_state(self.h, step, nebr_now, disp_hi), "sgk_set_run_state")

    # -- stages ----------------------------------------------------------------------------
    def bin(self):
        self._check(self.lib.sgk_bin(self.h), "sgk_bin")

    def get_cells(self, cap=64):
        cell_of = np.zeros(self.n, np.int32)
        cnt = np.zeros(self.ncell, np.int32)
        slots = np.zeros(self.ncell * self.cfg.cell_capacity, np.int32)
        self._check(self.lib.sgk_get_cells(self.h, _ptr(cell_of), _ptr(cnt), _ptr(slots)), "sgk_get_cells")
        return cell_of, cnt, slots.reshape(self.ncell, self.cfg.cell_capacity)

    def build_nebr(self):
        self._check(self.lib.sgk_build_nebr(self.h), "sgk_build_nebr")

    def build_nl_order(self):
        """nl_scheme 3: rebuild the list-build order from the current positions now."""
        self._check(self.lib.sgk_build_nl_order(self.h), "sgk_build_nl_order")

    def get_nebr(self):
        cap = self.cfg.nl_capacity or self.params.iNebrMax
        nn = np.zeros(self.n, np.int32)
        nbl = np.zeros(cap * self.n, np.int32)
        self._check(self.lib.sgk_get_nebr(self.h, _ptr(nn), _ptr(nbl)), "sgk_get_nebr")
        return nn, nbl

    def carbon_lists(self):
        nc = self.n - self.n_metal
        lc, cc = self.cfg.lj_capacity, self.cfg.close_capacity
        out = dict(n_lj=np.zeros(nc, np.int32), lj_idx=np.zeros(nc * lc, np.int32),
                   lj_pos=np.zeros(nc * lc, np.int32), n_close=np.zeros(nc, np.int32),
                   close_idx=np.zeros(nc * cc, np.int32), close_pos=np.zeros(nc * cc, np.int32))
        self._check(self.lib.sgk_get_carbon_lists(self.h, *[_ptr(out[k]) for k in (
            "n_lj", "lj_idx", "lj_pos", "n_close", "close_idx", "close_pos")]), "sgk_get_carbon_lists")
        for k, w in (("lj_idx", lc), ("lj_pos", lc), ("close_idx", cc), ("close_pos", cc)):
            out[k] = out[k].reshape(nc, w)
        return out

    def compute_forces(self):
        self._check(self.lib.sgk_compute_forces(self.h), "sgk_compute_forces")

    def rho(self):
        out = np.zeros(max(self.n_metal, 1))
        self._check(self.lib.sgk_get_rho(self.h, _ptr(out)), "sgk_get_rho")
        return out[: self.n_metal]

    def forces(self):
        acc = np.zeros((self.n, 3))
        e = np.zeros(self.n)
        fr = np.zeros(max(self.n_metal, 1))
        self._check(self.lib.sgk_get_forces(self.h, _ptr(acc), _ptr(e), _ptr(fr)), "sgk_get_forces")
        return acc, e, fr[: self.n_metal]

    def leapfrog(self, part):
        self._check(self.lib.sgk_leapfrog(self.h, part), "sgk_leapfrog")

    def thermostat(self, step, vv_sum):
        self._check(self.lib.sgk_thermostat(self.h, step, vv_sum), "sgk_thermostat")

    def props(self):
        rec = N.SgkRecord()
        self._check(self.lib.sgk_eval_props(self.h, C.byref(rec)), "sgk_eval_props")
        return record_dict(rec)

    def rdf_clear(self):
        self._check(self.lib.sgk_rdf_clear(self.h), "sgk_rdf_clear")

    def rdf_accumulate(self):
        self._check(self.lib.sgk_rdf_accumulate(self.h), "sgk_rdf_accumulate")

    def rdf(self):
        h = np.zeros(self.params.sizeHistRdf, np.uint32)
        self._check(self.lib.sgk_rdf_get(self.h, _ptr(h)), "sgk_rdf_get")
        return h

    def diffusion_enable(self, max_tables=64):
        """Accumulate the centre-of-mass diffusion on the device in FP64 during run()."""
        self._check(self.lib.sgk_diffusion_enable(self.h, max_tables), "sgk_diffusion_enable")

    def diffusion_tables(self, max_tables=64):
        """(completed tables so far, array (k, nValDiffuse) of the most recent k <= max_tables)."""
        nv = self.params.nValDiffuse
        out = np.zeros((max_tables, nv))
        n = C.c_int()
        self._check(self.lib.sgk_diffusion_get(self.h, _ptr(out), max_tables, C.byref(n)), "sgk_diffusion_get")
        return n.value, out[: min(n.value, max_tables)]

    # -- device-resident loop ------------------------------------------------------------------
    def run(self, nsteps, rdf_every=0, records=True):
        recs = (N.SgkRecord * nsteps)() if records else None
        self._check(self.lib.sgk_run(self.h, nsteps, rdf_every, recs), "sgk_run")
        return [record_dict(r) for r in recs] if records else None

    def prepare_run(self, nsteps, rdf_every=0):
        """Capture the graph of the next run(nsteps, rdf_every) now, so that run does not pay for it."""
        self._check(self.lib.sgk_prepare_run(self.h, nsteps, rdf_every), "sgk_prepare_run")

    def last_run_ms(self):
        return self.lib.sgk_last_run_ms(self.h)

    def set_profiling(self, on):
        self._check(self.lib.sgk_set_profiling(self.h, int(on)), "sgk_set_profiling")

    def autotune(self, reps=20):
        """Pick the fastest launch shape of the density and force passes on the current state;
        returns [(metal_lanes, xcd_map, us), ...] for every candidate."""
        buf = (N.SgkTuneEntry * 32)()
        n = C.c_int()
        self._check(self.lib.sgk_autotune(self.h, reps, buf, 32, C.byref(n)), "sgk_autotune")
        return [(e.metal_lanes, e.xcd_map, e.us) for e in buf[: n.value]]

    def kernel_stats(self):
        n = self.lib.sgk_kernel_stats(self.h, None, 0)
        arr = (N.SgkKernelStat * max(n, 1))()
        n = self.lib.sgk_kernel_stats(self.h, arr, n)
        return {arr[k].name.decode(): (arr[k].launches, arr[k].ms) for k in range(n)}

    def device_bytes(self):
        return self.lib.sgk_device_bytes(self.h)

    # -- adjoint (mixed precision) -------------------------------------------------------------
    def adjoint_forward(self, nsteps, checkpoint_every=20, records=False):
        """Run a differentiation window from the current state (sgk_adjoint_forward): returns
        {"adhesion": mean Ni-C adhesion energy, "height": final metal centre-of-mass height above
        the sheet plane} in reduced units, plus the step records if asked."""
        obs = np.zeros(2)
        recs = (N.SgkRecord * nsteps)() if records else None
        self._check(self.lib.sgk_adjoint_forward(self.h, nsteps, checkpoint_every, _ptr(obs), recs),
                    "sgk_adjoint_forward")
        out = dict(zip(N.SGK_OBSERVABLES, obs.tolist()))
        if records:
            out["records"] = [record_dict(r) for r in recs]
        return out

    def adjoint_backward(self, weights, initial_state=False):
        """Gradient of weights . (adhesion, height) with respect to the scale factors of
        (LJ epsilon, LJ sigma, embedding F, density f); with initial_state also the gradients with
        respect to the window's initial positions and velocities, (N,3) each."""
        w = np.ascontiguousarray(weights, np.float64)
        g = np.zeros(4)
        dr0 = np.zeros((self.n, 3)) if initial_state else None
        dv0 = np.zeros((self.n, 3)) if initial_state else None
        self._check(self.lib.sgk_adjoint_backward(self.h, _ptr(w), _ptr(g), _ptr(dr0), _ptr(dv0)),
                    "sgk_adjoint_backward")
        return (g, dr0, dv0) if initial_state else g

    def adjoint_info(self):
        info = N.SgkAdjointInfo()
        self._check(self.lib.sgk_adjoint_get_info(self.h, C.byref(info)), "sgk_adjoint_get_info")
        return {k: (list(getattr(info, k)) if k == "obs" else getattr(info, k)) for k, _ in info._fields_
                if k != "pad"}

    def force_vjp(self, abar, w_adhesion=0.0):
        """(d/dr (N,3), d/d scale (4,)) of abar . a(r) + w_adhesion E_adh(r) at the current state."""
        ab = np.ascontiguousarray(abar, np.float64)
        rbar = np.zeros((self.n, 3))
        pbar = np.zeros(4)
        self._check(self.lib.sgk_force_vjp(self.h, _ptr(ab), float(w_adhesion), _ptr(rbar), _ptr(pbar)),
                    "sgk_force_vjp")
        return rbar, pbar

    def adhesion_energy(self):
        e = C.c_double()
        self._check(self.lib.sgk_adhesion_energy(self.h, C.byref(e)), "sgk_adhesion_energy")
        return e.value


class Batch:
    """Many replicas advanced together (sgk_batch): one launch per stage for all of them.  Replicas
    share the potentials, time step and schedules; geometry, atom counts and temperature differ."""

    def __init__(self, params_list, precision="mixed", rebuild_every=0, metal_lanes=0, records_capacity=1024,
                 use_graphs=True, **cfg):
        self.lib = N.lib()
        self.params = list(params_list)
        c = N.SgkConfig()
        self.lib.sgk_default_config(C.byref(c))
        c.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        c.rebuild_every = rebuild_every
        c.metal_lanes = metal_lanes
        c.records_capacity = records_capacity
        c.use_graphs = int(use_graphs)
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        arr = (N.SimParams * len(self.params))(*self.params)
        h = C.c_void_p()
        rc = self.lib.sgk_batch_create(len(self.params), C.cast(arr, C.c_void_p), C.byref(c), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"sgk_batch_create failed ({rc}): {self.lib.sgk_batch_last_error(None).decode()}")
        self.h = h
        self.nrep = len(self.params)

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.sgk_batch_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sgk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def metal_lanes(self):
        return self.lib.sgk_batch_metal_lanes(self.h)

    @property
    def atoms(self):
        return self.lib.sgk_batch_atoms(self.h)

    def replica_lanes(self):
        """Lanes per metal atom of every replica."""
        return [self.lib.sgk_batch_replica_lanes(self.h, r) for r in range(self.nrep)]

    def set_lanes(self, lanes):
        arr = np.ascontiguousarray(lanes, np.int32)
        assert arr.shape == (self.nrep,)
        self._check(self.lib.sgk_batch_set_lanes(self.h, _ptr(arr)), "sgk_batch_set_lanes")

    def autotune(self, reps=20):
        """Per replica-size bucket, the fastest lanes per metal atom on the current state; returns
        [(metal_lo, metal_hi, replicas, atoms, lanes, us), ...] for every (bucket, lanes) tried."""
        buf = (N.SgkBatchTuneEntry * 64)()
        n = C.c_int()
        self._check(self.lib.sgk_batch_autotune(self.h, reps, buf, 64, C.byref(n)), "sgk_batch_autotune")
        return [(e.metal_lo, e.metal_hi, e.replicas, e.atoms, e.metal_lanes, e.us) for e in buf[: n.value]]

    def upload(self, rep, r, v=None, a=None):
        n = self.params[rep].nMol
        r4 = np.zeros((n, 4))
        r = np.asarray(r, np.float64)
        r4[:, : r.shape[1]] = r
        args = [np.ascontiguousarray(x, np.float64) if x is not None else None for x in (r4, v, a)]
        self._check(self.lib.sgk_batch_upload_f64(self.h, rep, *[_ptr(x) if x is not None else None for x in args]),
                    "sgk_batch_upload_f64")

    def download(self, rep):
        n = self.params[rep].nMol
        r, v, a = np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self.lib.sgk_batch_download_f64(self.h, rep, _ptr(r), _ptr(v), _ptr(a)), "sgk_batch_download_f64")
        return r, v, a

    def set_run_state(self, rep, step=0, nebr_now=1, disp_hi=0.0):
        self._check(self.lib.sgk_batch_set_run_state(self.h, rep, step, nebr_now, disp_hi), "sgk_batch_set_run_state")

    def run(self, nsteps, rdf_every=0, records=True):
        """nsteps steps of every replica; returns one list of record dicts per replica."""
        if not records:
            self._check(self.lib.sgk_batch_run(self.h, nsteps, rdf_every, None), "sgk_batch_run")
            return None
        buf = (N.SgkRecord * (nsteps * self.nrep))()
        self._check(self.lib.sgk_batch_run(self.h, nsteps, rdf_every, buf), "sgk_batch_run")
        return [[record_dict(buf[r * nsteps + k]) for k in range(nsteps)] for r in range(self.nrep)]

    def last_run_ms(self):
        return self.lib.sgk_batch_last_run_ms(self.h)

    def rdf(self, rep):
        h = np.zeros(self.params[rep].sizeHistRdf, np.uint32)
        self._check(self.lib.sgk_batch_rdf_get(self.h, rep, _ptr(h)), "sgk_batch_rdf_get")
        return h


def record_dict(r):
    d = {}
    for name, _ in N.SgkRecord._fields_:
        val = getattr(r, name)
        d[name] = list(val) if hasattr(val, "__len__") else val
    return d
