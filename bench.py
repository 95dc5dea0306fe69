"""Throughput of the per-step MD engine on one MI355X.

A step is one full MD step of the hot path (leapfrog + boundary + binning when due, neighbour list
when due, EAM density, fused EAM/LJ/graphene force pass with leapfrog part 2 and the step
reductions, thermostat on its steps) over the whole system, with all state resident in HBM, plus
the samples the large workload takes on the device: the radial distribution function (400 bins to
12 A, every metal pair within range) every stepRdf steps (50, the reference input's row 20) and the
centre-of-mass MSD buffers every stepDiffuse steps (100); --rdf-every / --msd-every override them.
The step graph of the timed pattern is captured before the timed region.

Default workload: the large system - 1,003,520-atom periodic graphene sheet + 100,000-atom Ni fcc
sphere 2.5 A above it, 1,000 K Maxwell-Boltzmann velocities, dt = 1 fs, Berendsen tau = 0.1 ps,
neighbour list rebuilt every 10 steps (or sooner by the skin rule).  `value` is the mixed-precision
form's atom-steps/s; the plain FP32 form is timed on the same workload and reported beside it.

Roofline: the dominant kernel is the force pass; its algorithmic bytes per launch are the
per-atom-step figure of the state traffic (plain 48 + 20 f_Me B, mixed 96 + 20 f_Me B) times the
atoms, divided by its mean launch duration from HIP events on the engine stream.  `bound` is the
larger of the pass's measured HBM-traffic fraction and VALU-issue fraction (committed rocprofv3
counter summaries); `step_frac` is the same algorithmic bytes over the whole step time.

cpu_baseline: the FP32 CPU restatement of the same step (oracle/), built for speed (-O3
-march=x86-64-v3, OpenMP over the per-atom loops), on one core and on the host's cores, each on a
bounded number of steps of the same workload.  The FP64 oracle gives the force and
energy error of both GPU forms, on the initial state and on the state the timed mixed run ends in.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_if_needed():
    lib = os.path.join(ROOT, "surfacegrowthconsole_amd", "lib", "libsurfacegrowth.so")
    orc = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not (os.path.exists(lib) and os.path.exists(orc)):
        import __graft_entry__

        __graft_entry__.build()


def calibrate(gib=1.0, reps=5):
    """A plain elementwise kernel of known traffic (out = -in, float32): reads and writes `gib` GiB per
    launch.  Profiled with FETCH_SIZE / WRITE_SIZE it gives the factor that converts those counters
    to bytes on this machine."""
    import torch

    n = int(gib * (1 << 30)) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "torch.neg float32", "bytes_read_per_launch": 4 * n,
                      "bytes_written_per_launch": 4 * n, "launches": reps}))


def profile_figure(name, key):
    """A figure recorded in a committed rocprofv3 summary (profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def measured_traffic(workload):
    """HBM bytes per k_force launch from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or None."""
    return profile_figure(f"k_force_{workload}_traffic.md", "traffic_bytes_per_launch")


def valu_issue_us(workload):
    """Lower bound on the SIMD time of one k_force launch (mixed form): its VALU wave-instructions
    (SQ_INSTS_VALU, committed rocprofv3 summary) x 4 cycles / (1,024 SIMDs x 2.4 GHz), or None."""
    n = profile_figure(f"k_force_{workload}_valu.md", "valu_instructions_per_launch")
    return None if n is None else n * 4.0 / (1024 * 2.4e9) * 1e6


def nve_drift(config, steps, oracle_steps):
    """Total-energy drift without thermostat (eV/atom/ns, slope of a least-squares line) of each GPU
    form over `steps` steps, and of the FP64 oracle over `oracle_steps` steps of the same start.
    The reference's graphene energy is not the potential of its forces, so the drift is not zero
    even in exact arithmetic: what a form adds to the FP64 oracle's drift is its precision cost."""
    import numpy as np
    import surfacegrowthconsole_amd as sg

    s = sg.config_system(config, "config", stepThermostat=1 << 30)
    p = s.params
    dt_ns = p.deltaT * 0.2e-3  # time unit 0.2 ps
    out = {"workload": config, "atoms": p.nMol, "steps": steps}

    def slope(tot):
        t = np.arange(1, len(tot) + 1) * dt_ns
        return float(np.polyfit(t, np.asarray(tot) * p.enU, 1)[0])

    for form in ("mixed", "mixed_fp32pairs", "fp32"):
        with sg.Engine(p, rebuild_every=10, records_capacity=1000, **FORMS[form]) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            tot = []
            while len(tot) < steps:
                tot += [x["totEnergy"] for x in eng.run(min(1000, steps - len(tot)))]
        out[form] = slope(tot)
        out[form + "_first_" + str(oracle_steps)] = slope(tot[:oracle_steps])
        log(f"[bench] NVE {form}: {out[form]:.4e} eV/atom/ns")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    # the FP64 oracle on OpenMP threads (bitwise the serial one), 500 steps per call
    o = oracle.Oracle(p, "f64", library=oracle.threaded_f64_lib())
    r, v, a = s.r.astype(np.float64), s.v.astype(np.float64), s.a.astype(np.float64)
    tot, nebr, disp = [], 1, 0.0
    t0 = time.perf_counter()
    while len(tot) < oracle_steps:
        n = min(500, oracle_steps - len(tot))
        err, r, v, a, recs = o.run(r, v, a, n, len(tot), nebr, disp, 10)
        assert err == 0
        nebr, disp = recs[-1]["nebrNow"], recs[-1]["dispHi"]
        tot += [x["totEnergy"] for x in recs]
        log(f"[bench] NVE FP64 oracle: {len(tot)} steps, {time.perf_counter() - t0:.0f} s")
    out["fp64_oracle_first_" + str(oracle_steps)] = slope(tot)
    log(f"[bench] NVE FP64 oracle ({oracle_steps} steps): {out['fp64_oracle_first_' + str(oracle_steps)]:.4e}")
    return out


def batch_workload(nrep=64, seed=2024):
    """The uneven-replica ensemble: Ni fcc spheres of log-uniform size in [55, 2000] atoms, each on a
    graphene sheet whose sides are at least 3 cluster diameters (2,016-16,128 C, 32 ux uy atoms),
    temperatures uniform in [300, 1200] K; config variant otherwise."""
    import numpy as np
    import surfacegrowthconsole_amd as sg

    rng = np.random.default_rng(seed)
    a = 3.5196  # Ni lattice constant, angstrom
    out = []
    for k in range(nrep):
        n = int(round(math.exp(rng.uniform(math.log(55), math.log(2000)))))
        d = 2.0 * a * (3.0 * n / (16.0 * math.pi)) ** (1.0 / 3.0)
        ux = max(2, math.ceil(3.0 * d / (8 * math.cos(math.pi / 6) * 1.42)))
        uy = max(2, math.ceil(3.0 * d / (6 * 1.42)))
        # sheet area log-uniform between the smallest allowed and 16,128 atoms (ux*uy <= 504)
        cells = math.exp(rng.uniform(math.log(max(63, ux * uy)), math.log(504)))
        while ux * uy < cells and (ux + 1) * uy <= 504:
            ux += 1
            if ux * uy < cells and ux * (uy + 1) <= 504:
                uy += 1
        t = float(rng.uniform(300.0, 1200.0))
        out.append(sg.config_system("small", "config", ux=ux, uy=uy, n_metal=n, temperature=t, seed=100 + k))
    return out


def bench_batch(steps, warmup, nrep, tune=True, sequential=True):
    """Atom-steps/s of the whole ensemble batched (one launch per stage, list rebuilds synchronised
    across replicas), with lanes per metal atom tuned per replica-size bucket after the warm-up (the
    table is reported), and, for comparison, of the same replicas advanced one after another, each in
    its own context."""
    import surfacegrowthconsole_amd as sg

    systems = batch_workload(nrep)
    natoms = sum(s.params.nMol for s in systems)
    log(f"[bench] batch: {nrep} replicas, {natoms} atoms")
    res = {"workload": f"{nrep} uneven replicas (55-2000 Ni, 2016-16128 C, 300-1200 K)", "replicas": nrep,
           "atoms": natoms, "steps": steps}
    res["sync_rebuilds"] = 1
    for prec in ("mixed", "fp32"):
        # sync_rebuilds: a step that rebuilds one replica's lists rebuilds all (the list kernels then
        # run on about one step in five instead of on nearly every step)
        with sg.Batch([s.params for s in systems], precision=prec, rebuild_every=10, records_capacity=500,
                      sync_rebuilds=1) as bt:
            for k, s in enumerate(systems):
                bt.upload(k, s.r, s.v, s.a)
                bt.set_run_state(k, 0, 1, 0.0)
            bt.run(warmup, records=False)
            if tune:
                table = bt.autotune(10)
                res[f"tune_{prec}"] = [{"metal_atoms": [e[0], e[1]], "replicas": e[2], "atoms": e[3],
                                        "lanes_per_metal_atom": e[4], "us": round(e[5], 2)} for e in table]
                res[f"lanes_{prec}"] = bt.replica_lanes()
                bt.run(warmup, records=False)  # capture the tuned step graphs
            t0 = time.perf_counter()
            done = 0
            while done < steps:
                n = min(500, steps - done)
                bt.run(n, records=False)
                done += n
            wall = time.perf_counter() - t0
        res[f"batched_{prec}"] = natoms * steps / wall
        log(f"[bench] batch {prec}: {wall / steps * 1e3:.3f} ms/step, {natoms * steps / wall:.4e} atom-steps/s")
        if not sequential:
            continue
        # one context per replica, advanced in turn
        engines = []
        for s in systems:
            e = sg.Engine(s.params, precision=prec, rebuild_every=10, records_capacity=500)
            e.upload(s.r, s.v, s.a)
            e.set_run_state(0, 1, 0.0)
            e.run(warmup, records=False)
            engines.append(e)
        seq_steps = max(50, steps // 10)
        t0 = time.perf_counter()
        for e in engines:
            e.run(seq_steps, records=False)
        wall = time.perf_counter() - t0
        for e in engines:
            e.close()
        res[f"sequential_{prec}"] = natoms * seq_steps / wall
        log(f"[bench] one context per replica {prec}: {natoms * seq_steps / wall:.4e} atom-steps/s")
    return res


def b_alg(precision, f_me):
    return (96.0 if precision == "mixed" else 48.0) + 20.0 * f_me


def time_form(sys_, precision, steps, warmup, rebuild, chunk, sample_every, msd):
    """Warm-up, then `steps` timed device-resident steps (complete-range RDF samples every
    `sample_every` steps, MSD samples every stepDiffuse steps if `msd`), then the same number again
    with HIP events around every launch for the per-kernel durations.  The graph of the timed chunk
    is captured after the warm-up."""
    import surfacegrowthconsole_amd as sg

    p = sys_.params
    eng = sg.Engine(p, precision=precision, rebuild_every=rebuild, records_capacity=chunk, rdf_complete=1)
    eng.upload(sys_.r, sys_.v, sys_.a)
    eng.set_run_state(0, 1, 0.0)
    if msd:
        eng.diffusion_enable(4)
    eng.rdf_clear()
    done = 0
    while done < warmup:
        n = min(chunk, warmup - done)
        eng.run(n, sample_every, records=False)
        done += n
    eng.prepare_run(min(chunk, steps), sample_every)
    # timed region: `steps` device-resident steps; sgk_run synchronises the stream on return
    t0 = time.perf_counter()
    dev_ms = 0.0
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        eng.run(n, sample_every, records=False)
        dev_ms += eng.last_run_ms()
        done += n
    wall = time.perf_counter() - t0
    # per-kernel durations: HIP events around each launch on the engine stream, same step count
    eng.set_profiling(True)
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        eng.run(n, sample_every, records=False)
        done += n
    stats = eng.kernel_stats()
    eng.set_profiling(False)
    r, v, a = eng.download()
    eng.close()
    return wall, dev_ms, stats, r


# Precision forms: plain FP32; mixed with FP64 pair arithmetic (default); mixed with FP32 pairs.
FORMS = {"mixed": dict(precision="mixed"), "fp32": dict(precision="fp32"),
         "mixed_fp32pairs": dict(precision="mixed", pair_fp64=0)}


def gpu_forces(params, r, form):
    import surfacegrowthconsole_amd as sg

    with sg.Engine(params, **FORMS[form]) as eng:
        eng.upload(r)
        eng.bin()
        eng.build_nebr()
        eng.compute_forces()
        acc, e, _ = eng.forces()
    return acc, e


def force_errors(params, states):
    """Per state and form, against the FP64 oracle on the same positions: e_F = max_i |F_i - F_i^64| /
    rms_j |F_j^64|, the global ||dF||_2 / ||F^64||_2 and the potential-energy error |E - E^64| / |E^64|
    (E = sum of the per-atom energies).  The plain form is evaluated at the float-rounded positions it
    stores, its reference too."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    o64 = oracle.Oracle(params, "f64")
    out = {}
    for label, r in states.items():
        for prec in ("mixed", "mixed_fp32pairs", "fp32"):
            rin = np.asarray(r, np.float64) if prec != "fp32" else np.asarray(r, np.float32).astype(np.float64)
            _, nn, nbl = o64.neighbour_list(rin)
            a64, e64, _ = o64.forces(rin, nn, nbl)
            acc, e = gpu_forces(params, rin, prec)
            rms = float(np.sqrt((a64 ** 2).sum(1).mean()))
            key = f"{prec}_{label}"
            out[key] = float(np.abs(acc - a64).max() / rms)
            out[key + "_l2"] = float(np.sqrt(((acc - a64) ** 2).sum()) / np.sqrt((a64 ** 2).sum()))
            out[key + "_energy"] = float(abs(e.sum() - e64.sum()) / abs(e64.sum()))
            log(f"[bench] {prec} {label}: e_F {out[key]:.3e}, l2 {out[key + '_l2']:.3e}, "
                f"energy {out[key + '_energy']:.3e}")
    return out


def host_cpu():
    """Model name and core counts of the host this runs on."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def cpu_baseline(sys_, rebuild, seconds):
    """The FP32 CPU restatement of the step built for speed (oracle/_build/liboracle_omp.so: -O3
    -march=x86-64-v3, OpenMP per-atom loops), timed on the host: one thread, then all the threads
    OpenMP is given (OMP_NUM_THREADS, else the cores this process may use), each on about `seconds`
    of steps of the same workload after one untimed step.  Rebuild cadence as the GPU run."""
    import ctypes

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    p = sys_.params
    lib = oracle.timing_lib()
    gomp = ctypes.CDLL("libgomp.so.1")
    model, ncpu, nallowed = host_cpu()
    threads_all = int(os.environ.get("OMP_NUM_THREADS") or nallowed)
    res = {}
    for threads in (1, threads_all):
        gomp.omp_set_num_threads(threads)
        o = oracle.Oracle(p, "f32", library=lib)
        t0 = time.perf_counter()
        err, r, v, a, _ = o.run(sys_.r, sys_.v, sys_.a, 1, 0, 1, 0.0, rebuild)
        t1 = time.perf_counter() - t0
        n = max(2, min(200, int(seconds / max(t1, 1e-6))))
        t0 = time.perf_counter()
        err, r, v, a, _ = o.run(r, v, a, n, 1, 0, 0.0, rebuild)
        dt = time.perf_counter() - t0
        res[threads] = (p.nMol * n / dt, n)
        log(f"[bench] cpu {threads} thread(s): {res[threads][0]:.4e} atom-steps/s ({n} steps)")
    one, allc = res[1], res[threads_all]
    return {"value": one[0], "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"{one[1]} steps (after 1 untimed) of the same workload, FP32 restatement built -O3 "
                      f"-march=x86-64-v3, 1 thread",
            "all_cores": {"value": allc[0], "cores": threads_all, "sample": f"{allc[1]} steps, OpenMP"},
            "host": {"cpu": model, "nproc": ncpu, "cores_allowed": nallowed}}


def bench_adjoint(equil, window, every, fd_steps):
    """Differentiation workload: 8,064 C graphene + 561-atom Ni Mackay icosahedron 3 A above it,
    600 K, equilibrated `equil` steps on the GPU, then a `window`-step window checkpointed every
    `every` steps.  Gradients of the window-mean adhesion energy and the final cluster height with
    respect to the scale factors of LJ epsilon, LJ sigma, embedding F and density f from the GPU
    adjoint (mixed form), against the tangent build of the FP64 oracle on the same initial state
    and against central differences of the FP64 oracle at each relative step in fd_steps, free and
    with the unperturbed run's discrete decisions frozen."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import surfacegrowthconsole_amd as sg
    from surfacegrowthconsole_amd.configs import rebuild_every

    s = sg.config_system("icosa", "config")
    p = s.params
    rebuild = rebuild_every("config")
    eng = sg.Engine(p, precision="mixed", rebuild_every=rebuild)
    eng.upload(s.r.astype(np.float64), s.v.astype(np.float64))
    log(f"[adjoint] equilibrating {equil} steps")
    left = equil
    while left > 0:
        n = min(500, left)
        eng.run(n, records=False)
        left -= n
    r0, v0, _ = eng.download()
    runs = []
    for rep in range(2):  # the first pass captures the forward graphs; the second is timed
        eng.upload(r0, v0)
        eng.set_run_state(equil, 1, 0.0)
        t0 = time.perf_counter()
        obs = eng.adjoint_forward(window, every)
        t1 = time.perf_counter()
        g_adh = eng.adjoint_backward([1.0, 0.0])
        t2 = time.perf_counter()
        g_h = eng.adjoint_backward([0.0, 1.0])
        t3 = time.perf_counter()
        runs.append(dict(obs=obs, grad=np.array([g_adh, g_h]), info=eng.adjoint_info(),
                         wall_ms=[(t1 - t0) * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3]))
        log(f"[adjoint] pass {rep}: forward {runs[-1]['wall_ms'][0]:.1f} ms, backward "
            f"{runs[-1]['wall_ms'][1]:.1f} + {runs[-1]['wall_ms'][2]:.1f} ms")
    run = runs[1]
    grad = run["grad"]
    assert np.array_equal(grad, runs[0]["grad"]) or np.allclose(grad, runs[0]["grad"], rtol=1e-12)
    o = oracle.Oracle(p, "f64")
    log("[adjoint] tangent oracle")
    t = time.perf_counter()
    J, dJ, *_ = o.tangent_window(r0, v0, window, step0=equil, rebuild_every=rebuild, pdir=np.eye(4))
    t_tan = time.perf_counter() - t
    out = {
        "workload": f"icosa: {p.nMol - p.nMolMe} C graphene + {p.nMolMe} Ni icosahedron, 600 K, "
                    f"equilibrated {equil} steps, window {window} steps, checkpoint every {every}",
        "atoms": p.nMol,
        "observables": {"gpu": [run["obs"]["adhesion"], run["obs"]["height"]], "oracle_fp64": J.tolist()},
        "params": list(oracle.PARAMS),
        "grad_gpu": grad.tolist(),
        "grad_tangent_fp64": dJ.tolist(),
        "rel_err_vs_tangent": (np.abs(grad - dJ) / np.abs(dJ)).tolist(),
        "gpu_ms": {"forward": run["info"]["forward_ms"], "backward_adhesion": run["wall_ms"][1],
                   "backward_height": run["wall_ms"][2], "forward_wall": run["wall_ms"][0]},
        "adjoint_bytes": run["info"]["bytes"],
        "segments": run["info"]["segments"],
        "list_copies": run["info"]["list_copies"],
        "recompute_mismatches": run["info"]["mismatches"],
        "oracle_tangent_s": t_tan,
        "fd": {},
    }
    # Central differences free (each perturbed run takes its own decisions) and frozen (the decision
    # tape of the unperturbed run replayed: rebuild steps, lists, cutoff bits, F(rho) branches), on
    # the threaded build of the FP64 oracle (bitwise the serial one).
    o = oracle.Oracle(p, "f64", library=oracle.threaded_f64_lib())
    out["fd_frozen"] = {}
    if fd_steps:
        o.freeze("record")
        o.window(r0, v0, window, step0=equil, rebuild_every=rebuild)
        o.freeze("off")
    for h in fd_steps:
        for frozen in (False, True):
            log(f"[adjoint] central differences, h = {h:g}, {'frozen' if frozen else 'free'}")
            fd = np.zeros((2, 4))
            for k in range(4):
                Js = []
                for sgn in (1, -1):
                    sc = np.ones(4)
                    sc[k] = 1 + sgn * h
                    o.set_scale(sc)
                    o.freeze("replay" if frozen else "off")
                    Js.append(o.window(r0, v0, window, step0=equil, rebuild_every=rebuild)[0])
                    assert not frozen or o.freeze_ok()
                fd[:, k] = (Js[0] - Js[1]) / (2 * h)
            o.set_scale(np.ones(4))
            o.freeze("off")
            out["fd_frozen" if frozen else "fd"][f"{h:g}"] = {
                "grad": fd.tolist(), "rel_err_gpu": (np.abs(grad - fd) / np.abs(fd)).tolist(),
                "rel_err_tangent": (np.abs(dJ - fd) / np.abs(fd)).tolist()}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", default="large",
                    choices=["small", "medium", "large", "icosa", "shard8_sheet", "shard8"])
    ap.add_argument("--variant", default="config", choices=["config", "reference"])
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--rdf-every", type=int, default=-1,
                    help="RDF sample interval in the timed steps (-1: the workload's stepRdf; 0: none)")
    ap.add_argument("--msd-every", type=int, default=-1,
                    help="centre-of-mass MSD sample interval (-1: the workload's stepDiffuse; 0: none)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--calibrate", action="store_true", help="run the copy kernel for counter calibration")
    ap.add_argument("--nve", default=None, choices=["small", "medium", "large", "icosa"],
                    help="measure the NVE energy drift of every form on this workload instead")
    ap.add_argument("--nve-steps", type=int, default=5000)
    ap.add_argument("--batch", type=int, default=0, metavar="NREP",
                    help="time an ensemble of NREP uneven replicas, batched and one context each")
    ap.add_argument("--nve-oracle-steps", type=int, default=1000)
    ap.add_argument("--batch-only", action="store_true", help="with --batch: skip the one-context-per-replica legs")
    ap.add_argument("--no-tune", action="store_true", help="with --batch: keep the default lanes per metal atom")
    ap.add_argument("--adjoint", action="store_true",
                    help="differentiation workload: adjoint gradients against the FP64 tangent and FD")
    ap.add_argument("--adjoint-equil", type=int, default=2000)
    ap.add_argument("--adjoint-window", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--adjoint-fd", default="1e-4,1e-6", help="relative FD steps, comma separated ('' for none)")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    if args.adjoint:
        build_if_needed()
        fd = [float(x) for x in args.adjoint_fd.split(",") if x]
        print(json.dumps({"adjoint": bench_adjoint(args.adjoint_equil, args.adjoint_window, args.adjoint_every, fd)}))
        return
    if args.batch:
        build_if_needed()
        print(json.dumps({"batch": bench_batch(args.steps, args.warmup, args.batch, tune=not args.no_tune,
                                               sequential=not args.batch_only)}))
        return
    if args.nve:
        build_if_needed()
        print(json.dumps({"nve_drift_eV_per_atom_per_ns": nve_drift(args.nve, args.nve_steps, args.nve_oracle_steps)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    build_if_needed()
    import surfacegrowthconsole_amd as sg
    from surfacegrowthconsole_amd.configs import rebuild_every

    log(f"[bench] building {args.config}/{args.variant}")
    extra = {"stepDiffuse": args.msd_every} if args.msd_every > 0 else {}
    s = sg.config_system(args.config, args.variant, **extra)
    p = s.params
    rdf_every = p.stepRdf if args.rdf_every < 0 else args.rdf_every
    msd_every = 0 if args.msd_every == 0 else p.stepDiffuse
    rebuild = rebuild_every(args.variant)
    f_me = p.nMolMe / p.nMol
    res = {}
    for prec in ("mixed", "fp32"):
        log(f"[bench] {prec}: {args.warmup} warm-up + {args.steps} timed steps")
        wall, dev_ms, stats, r_end = time_form(s, prec, args.steps, args.warmup, rebuild, args.chunk,
                                               rdf_every, msd_every > 0)
        res[prec] = dict(wall=wall, dev_ms=dev_ms, stats=stats, r_end=r_end)
        log(f"[bench] {prec}: {wall / args.steps * 1e3:.3f} ms/step, {p.nMol * args.steps / wall:.4e} atom-steps/s")
    m = res["mixed"]
    value = p.nMol * args.steps / m["wall"]
    ms_per_step = m["wall"] / args.steps * 1e3
    fl, fms = m["stats"]["force"]
    force_us = fms / fl * 1e3
    alg_per_atom = b_alg("mixed", f_me)
    alg_bytes = alg_per_atom * p.nMol
    achieved = alg_bytes / (force_us * 1e-6) / 1e9
    traffic = measured_traffic(args.config)
    hbm_frac = traffic / (force_us * 1e-6) / 1e9 / HBM_PEAK_GBS if traffic else None
    valu_us = valu_issue_us(args.config)
    valu_frac = valu_us / force_us if valu_us is not None else None
    bound = "hbm"
    if valu_frac is not None and (hbm_frac is None or valu_frac > hbm_frac):
        bound = "valu"
    kernels = {k: {"launches": v[0], "us_per_launch": round(v[1] / v[0] * 1e3, 3)} for k, v in m["stats"].items()}
    samples = {}
    for k, every in (("rdf", rdf_every), ("diffuse", msd_every)):
        if k in m["stats"]:
            samples[k] = {"us_per_sample": round(m["stats"][k][1] / m["stats"][k][0] * 1e3, 3),
                          "every_steps": every}
    lu = p.lengthU  # angstrom per reduced length unit
    out = {
        "metric": "atom-steps/s on 1 MI355X, mixed vs FP32 at <=1e-5 rel force err vs FP64",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, accelerations and EAM/LJ pair math; f32 graphene force law",
        "data": "synthetic (generated lattices, seeded velocities)",
        "config": {"workload": f"{args.config}: {p.nMol - p.nMolMe} C graphene + {p.nMolMe} Ni fcc sphere, "
                               f"{p.temperature * p.temperatureU:.0f} K, dt 1 fs, variant {args.variant}",
                   "atoms": p.nMol, "metal_fraction": round(f_me, 4), "precision": "mixed",
                   "rebuild_every": rebuild,
                   "eam_cutoff_A": round(p.rCutEam * lu, 3),
                   "lj_cutoff_A": round(min(p.rCutLJ, math.sqrt(p.rrNebr)) * lu, 3),
                   "cutoffs_note": "reference rules: EAM cut at 1.45 a; LJ limited by the neighbour-list radius "
                                   "(rCutEam + skin), not 2.5 sigma",
                   "rdf": f"{p.sizeHistRdf} bins to {p.rangeRdf * lu:.1f} A, every metal pair in range, "
                          f"every {rdf_every} steps" if rdf_every else "none",
                   "msd": f"centre-of-mass buffers on the device (FP64), every {msd_every} steps"
                          if msd_every else "none",
                   "parallelism": "single GPU"},
        "fp32": {"value": p.nMol * args.steps / res["fp32"]["wall"],
                 "ms_per_step": res["fp32"]["wall"] / args.steps * 1e3},
        "roofline": {"bound": bound, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_force (fused metal+carbon force pass)", "us_per_launch": force_us,
                     "alg_bytes_per_atom_step": alg_per_atom,
                     "hbm_traffic_frac": hbm_frac, "valu_issue_frac": valu_frac,
                     "counters_from": [f"profiles/k_force_{args.config}_traffic.md",
                                       f"profiles/k_force_{args.config}_valu.md"] if traffic else None,
                     "step_frac": alg_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "samples": samples,
        "kernels": kernels,
    }
    if not args.no_check:
        log("[bench] force and energy error against the FP64 oracle (initial state, state after the mixed run)")
        out["force_error_vs_fp64"] = force_errors(p, {"initial": s.r, "final": res["mixed"]["r_end"]})
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(s, rebuild, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
