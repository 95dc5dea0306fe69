"""The benchmark workloads at their full sizes on the GPU, against the oracle.

  medium (40,320 C + 4,033 Ni, 900 K): forces and energy after 10,000 device steps, each form against
      the oracle of its precision.
  large (1,003,520 C + 100,000 Ni, 1,000 K): mixed-form forces and energy against the FP64 oracle
      (e_F <= 1e-5, energy <= 1e-7 relative) at the initial state and after 300 device steps; the
      plain form against the FP32 restatement at the initial state (e_F <= 1e-3: see the test).
  ensemble (64 uneven replicas, 55-2,000 Ni on 2,016-16,128 C, 300-1,200 K): every replica of the
      batched engine bitwise equal to the replica run alone; forces of the smallest and the largest
      replica against the FP64 oracle.
  differentiation (8,064 C + 561 Ni icosahedron, 600 K, equilibrated 2,000 steps): gradients of a
      200-step window checkpointed every 20 steps against the FP64 tangent and against central
      differences at h = 1e-4 with the discrete decisions frozen (1e-4 relative, all 8 components).
"""
import numpy as np
import pytest

import bench
import oracle
import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd.configs import rebuild_every

pytestmark = pytest.mark.gpu


def e_f(acc, ref):
    return np.abs(acc - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def engine_forces(p, r, prec, **cfg):
    with sg.Engine(p, precision=prec, **cfg) as eng:
        eng.upload(r if prec == "mixed" else np.asarray(r, np.float32))
        eng.bin()
        eng.build_nebr()
        eng.compute_forces()
        return eng.forces()


@pytest.fixture(scope="module")
def large():
    print("[configs] building the large workload", flush=True)
    return sg.config_system("large", "config")


def test_large_initial_forces_both_forms(gpu, large):
    p = large.params
    r = large.r.astype(np.float64)
    print(f"[configs] large: {p.nMol} atoms, FP64 oracle forces", flush=True)
    o64 = oracle.Oracle(p, "f64")
    err, nn, nbl = o64.neighbour_list(large.r)
    assert err == 0
    a64, e64, _ = o64.forces(r, nn, nbl)
    print("[configs] large: mixed form", flush=True)
    acc, e, _ = engine_forces(p, r, "mixed")
    assert e_f(acc, a64) <= 1e-5
    assert abs(e.sum() - e64.sum()) <= 1e-7 * abs(e64.sum())
    # The plain form forms x_i - x_k before the periodic wrap, as the reference does, so a bond across
    # the edge of this 1,108-unit box carries ~ulp(L)/2 = 3e-5 of error in FP32: its angle deviation
    # is ~3.5e-5 where the lattice has 0.  One ulp of acosf (ROCm's against glibc's) is then ~0.7 % of
    # that term, ~5e-5 of force, against an rms force of 0.13 on the near-perfect initial lattice:
    # e_F ~3.5e-4 between two correct FP32 evaluations.  Hence 1e-3 here, 2e-4 on smaller boxes.
    print("[configs] large: plain form against the FP32 restatement", flush=True)
    o32 = oracle.Oracle(p, "f32")
    a32, _, _ = o32.forces(large.r, nn, nbl)
    acc32, _, _ = engine_forces(p, large.r, "fp32")
    assert e_f(acc32, a32.astype(np.float64)) <= 1e-3


def test_large_forces_after_300_steps(gpu, large):
    p = large.params
    print("[configs] large: 300 mixed steps on the device", flush=True)
    with sg.Engine(p, precision="mixed", rebuild_every=rebuild_every("config"), records_capacity=300) as eng:
        eng.upload(large.r, large.v, large.a)
        eng.set_run_state(0, 1, 0.0)
        recs = eng.run(300)
        r, _, _ = eng.download()
    assert recs[-1]["step"] == 300 and recs[-1]["errors"] == 0
    o64 = oracle.Oracle(p, "f64")
    err, nn, nbl = o64.neighbour_list(r.astype(np.float32))
    assert err == 0
    a64, e64, _ = o64.forces(r, nn, nbl)
    acc, e, _ = engine_forces(p, r, "mixed")
    print(f"[configs] large after 300 steps: e_F {e_f(acc, a64):.3e}", flush=True)
    assert e_f(acc, a64) <= 1e-5
    assert abs(e.sum() - e64.sum()) <= 1e-7 * abs(e64.sum())


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_medium_forces_at_step_10000(gpu, prec):
    """medium (40,320 C + 4,033 Ni, 900 K): forces and energy of the state after 10,000 device steps
    against the oracle at that state - the mixed form against FP64 (e_F <= 1e-5, energy <= 1e-7),
    the plain form against the FP32 restatement (e_F <= 2e-4, energy <= 1e-6)."""
    s = sg.config_system("medium", "config")
    p = s.params
    print(f"[configs] medium: 10,000 {prec} steps", flush=True)
    with sg.Engine(p, precision=prec, rebuild_every=rebuild_every("config"), records_capacity=1000) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        for _ in range(10):
            recs = eng.run(1000)
        r, _, _ = eng.download()
    assert recs[-1]["step"] == 10000 and recs[-1]["errors"] == 0
    o = oracle.Oracle(p, "f64" if prec == "mixed" else "f32")
    rin = r if prec == "mixed" else r.astype(np.float32)
    err, nn, nbl = o.neighbour_list(rin.astype(np.float32) if prec == "mixed" else rin)
    assert err == 0
    aref, eref, _ = o.forces(rin, nn, nbl)
    aref, eref = aref.astype(np.float64), eref.astype(np.float64)
    acc, e, _ = engine_forces(p, rin, prec)
    ef = e_f(acc, aref)
    print(f"[configs] medium {prec} at step 10,000: e_F {ef:.3e}", flush=True)
    assert ef <= (1e-5 if prec == "mixed" else 2e-4)
    assert abs(e.sum() - eref.sum()) <= (1e-7 if prec == "mixed" else 1e-6) * abs(eref.sum())


def test_ensemble_of_64_replicas(gpu):
    systems = bench.batch_workload(64)
    sizes = [s.params.nMol for s in systems]
    assert len(systems) == 64 and min(s.params.nMolMe for s in systems) >= 55
    steps = 30
    print(f"[configs] ensemble: 64 replicas, {sum(sizes)} atoms, batched", flush=True)
    with sg.Batch([s.params for s in systems], precision="mixed", rebuild_every=10, metal_lanes=4, xcd_map=0,
                  records_capacity=steps) as bt:
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        recs = bt.run(steps)
        out = [bt.download(k) for k in range(64)]
    for k, s in enumerate(systems):  # one context at a time
        with sg.Engine(s.params, precision="mixed", rebuild_every=10, metal_lanes=4, xcd_map=0,
                       records_capacity=steps) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            ref = eng.run(steps)
            r, v, _ = eng.download()
        assert np.array_equal(out[k][0][:, :3], r[:, :3]), k
        assert np.array_equal(out[k][1], v), k
        assert [x["totEnergy"] for x in recs[k]] == [x["totEnergy"] for x in ref], k
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        s = systems[k]
        p = s.params
        print(f"[configs] ensemble replica {k}: {p.nMol} atoms against the FP64 oracle", flush=True)
        o64 = oracle.Oracle(p, "f64")
        err, nn, nbl = o64.neighbour_list(s.r)
        a64, _, _ = o64.forces(s.r.astype(np.float64), nn, nbl)
        acc, _, _ = engine_forces(p, s.r.astype(np.float64), "mixed", metal_lanes=4, xcd_map=0)
        assert e_f(acc, a64) <= 1e-5, k


def test_differentiation_window_against_fp64_tangent(gpu):
    s = sg.config_system("icosa", "config")
    p = s.params
    rebuild = rebuild_every("config")
    equil, window, every = 2000, 200, 20
    print(f"[configs] icosahedron: {p.nMol} atoms, {equil} equilibration steps", flush=True)
    with sg.Engine(p, precision="mixed", rebuild_every=rebuild, records_capacity=500) as eng:
        eng.upload(s.r.astype(np.float64), s.v.astype(np.float64))
        for _ in range(equil // 500):
            eng.run(500, records=False)
        r0, v0, _ = eng.download()
        eng.upload(r0, v0)
        eng.set_run_state(equil, 1, 0.0)
        obs = eng.adjoint_forward(window, every)
        g = np.array([eng.adjoint_backward([1.0, 0.0]), eng.adjoint_backward([0.0, 1.0])])
        info = eng.adjoint_info()
    assert info["mismatches"] == 0 and info["segments"] == window // every
    print("[configs] icosahedron: FP64 tangent of the window", flush=True)
    o = oracle.Oracle(p, "f64")
    J, dJ, *_ = o.tangent_window(r0, v0, window, step0=equil, rebuild_every=rebuild, pdir=np.eye(4))
    assert obs["adhesion"] == pytest.approx(J[0], rel=1e-4)
    assert obs["height"] == pytest.approx(J[1], rel=1e-4)
    rel = np.abs(g - dJ) / np.abs(dJ)
    print(f"[configs] icosahedron: gradient rel. error {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-4, (g, dJ)
    # Central differences at h = 1e-4 of the FP64 window with the unperturbed run's decisions held
    # (rebuild steps, lists, cutoff bits, F(rho) branches): the map the adjoint differentiates.  Free
    # differences at this step miss by up to ~0.6 (pairs cross the LJ cutoff inside the window).
    print("[configs] icosahedron: frozen central differences, h = 1e-4", flush=True)
    h = 1e-4
    o.freeze("record")
    J0 = o.window(r0, v0, window, step0=equil, rebuild_every=rebuild)[0]
    assert np.array_equal(J0, J)
    fd = np.zeros((2, 4))
    for k in range(4):
        Js = []
        for sgn in (1, -1):
            sc = np.ones(4)
            sc[k] = 1 + sgn * h
            o.set_scale(sc)
            o.freeze("replay")
            Js.append(o.window(r0, v0, window, step0=equil, rebuild_every=rebuild)[0])
            assert o.freeze_ok(), k
        fd[:, k] = (Js[0] - Js[1]) / (2 * h)
    o.set_scale(np.ones(4))
    o.freeze("off")
    rel_fd = np.abs(g - fd) / np.abs(fd)
    print(f"[configs] icosahedron: GPU gradient vs frozen differences, rel. error {rel_fd.max():.2e}", flush=True)
    assert rel_fd.max() <= 1e-4, (g, fd)
