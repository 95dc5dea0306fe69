"""Batched replicas (sgk_batch): each replica's trajectory equals the same replica run alone.

The batched kernels run the single-system per-atom code over each replica's own ranges with the
same workgroup decomposition, so with xcd_map = 0 and the same lanes per metal atom the results are
bitwise those of a separate sgk_ctx per replica.
"""
import ctypes as C

import numpy as np
import pytest

import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N

# (ux, uy, metal atoms, temperature K): uneven sheets, clusters and temperatures
SPECS = [(4, 4, 55, 300.0), (5, 6, 201, 600.0), (6, 5, 147, 900.0), (7, 7, 309, 1200.0)]


def replicas():
    return [sg.config_system("small", "config", ux=ux, uy=uy, n_metal=n, temperature=t, seed=11 + k)
            for k, (ux, uy, n, t) in enumerate(SPECS)]


def test_batch_rejects_replicas_with_different_potentials():
    a = sg.config_system("small", "config")
    b = sg.config_system("small", "config", epsilon=0.05)
    arr = (N.SimParams * 2)(a.params, b.params)
    lib = N.lib()
    cfg = N.SgkConfig()
    lib.sgk_default_config(C.byref(cfg))
    h = C.c_void_p()
    rc = lib.sgk_batch_create(2, C.cast(arr, C.c_void_p), C.byref(cfg), C.byref(h))
    assert rc != 0 and not h.value
    assert b"LJ" in lib.sgk_batch_last_error(None)


def test_batch_rejects_workgroups_other_than_256_threads():
    """The kernels are written for 4-wave workgroups: block_size other than 0 or 256 is an error,
    raised before any device work."""
    a = sg.config_system("small", "config")
    arr = (N.SimParams * 1)(a.params)
    lib = N.lib()
    for bs in (128, 512):
        cfg = N.SgkConfig()
        lib.sgk_default_config(C.byref(cfg))
        cfg.block_size = bs
        h = C.c_void_p()
        rc = lib.sgk_batch_create(1, C.cast(arr, C.c_void_p), C.byref(cfg), C.byref(h))
        assert rc != 0 and not h.value, bs
        assert b"block size" in lib.sgk_batch_last_error(None)


@pytest.mark.gpu
def test_engine_rejects_workgroups_other_than_256_threads(gpu):
    s = sg.config_system("small", "config")
    with pytest.raises(RuntimeError, match="block size must be 256"):
        sg.Engine(s.params, block_size=128)
    with sg.Engine(s.params, block_size=0) as eng:  # 0 takes the default, 256
        assert eng.cfg.block_size == 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mixed", "fp32"])
@pytest.mark.parametrize("nl_scheme", [3, 2, 0])
def test_batch_equals_replicas_run_alone(gpu, prec, nl_scheme):
    systems = replicas()
    steps = 40
    print(f"[batch] {prec}, list scheme {nl_scheme}: {len(systems)} replicas", flush=True)
    with sg.Batch([s.params for s in systems], precision=prec, rebuild_every=10, metal_lanes=4, xcd_map=0,
                  nl_scheme=nl_scheme) as bt:
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        recs = bt.run(steps)
        out = [bt.download(k) for k in range(len(systems))]
    for k, s in enumerate(systems):
        with sg.Engine(s.params, precision=prec, rebuild_every=10, metal_lanes=4, xcd_map=0, nl_scheme=nl_scheme) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            ref = eng.run(steps)
            r, v, a = eng.download()
        assert np.array_equal(out[k][0][:, :3], r[:, :3]), k
        assert np.array_equal(out[k][1], v), k
        assert [x["totEnergy"] for x in recs[k]] == [x["totEnergy"] for x in ref], k
        assert [x["rebuilt"] for x in recs[k]] == [x["rebuilt"] for x in ref], k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_batch_fused_drift_bitwise_equals_the_leapfrog_kernel(gpu, prec):
    """In 40-step chunks the batched force pass drifts the positions of every pre-kicked step but at
    most one (kb_lf1 then only bins); in chunks of four steps the first pre-kicked step is unfused
    and the next two fused; in chunks of two steps nothing is fused and kb_lf1 drifts.  With
    synchronised rebuilds the batches agree bit for bit with the last: every replica's positions
    with their energies, velocities, accelerations and records."""
    systems = replicas()
    out = []
    for chunk in (40, 4, 2):
        with sg.Batch([s.params for s in systems], precision=prec, rebuild_every=10, metal_lanes=4, xcd_map=0,
                      sync_rebuilds=1) as bt:
            for k, s in enumerate(systems):
                bt.upload(k, s.r, s.v, s.a)
                bt.set_run_state(k, 0, 1, 0.0)
            recs = [[] for _ in systems]
            for _ in range(80 // chunk):
                for k, x in enumerate(bt.run(chunk)):
                    recs[k] += x
            out.append(([bt.download(k) for k in range(len(systems))], recs))
    d1, r1 = out[-1]
    assert any(x["rebuilt"] for x in r1[0])
    for d0, r0 in out[:-1]:
        for k in range(len(systems)):
            for a, b in zip(d0[k], d1[k]):
                assert np.array_equal(a, b), k
            assert r0[k] == r1[k], k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_batch_with_a_metal_free_and_a_one_atom_replica(gpu, prec):
    """A replica without metal (no metal workgroups in any pass) and one with a single metal atom,
    beside an ordinary one, with synchronised rebuilds: each equals the replica run alone with the
    same rebuild steps, bit for bit."""
    systems = [sg.config_system("small", "config", ux=4, uy=4, n_metal=0, temperature=600.0, seed=3),
               sg.config_system("small", "config", ux=4, uy=5, n_metal=1, temperature=900.0, seed=4),
               sg.config_system("small", "config", ux=5, uy=5, n_metal=147, temperature=300.0, seed=5)]
    assert [s.params.nMolMe for s in systems] == [0, 1, 147]
    steps = 40
    print(f"[batch] {prec}: replicas with 0, 1 and 147 metal atoms", flush=True)
    with sg.Batch([s.params for s in systems], precision=prec, rebuild_every=0, metal_lanes=4, xcd_map=0,
                  sync_rebuilds=1) as bt:
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        recs = bt.run(steps)
        out = [bt.download(k) for k in range(len(systems))]
    rebuilt = [x["rebuilt"] for x in recs[0]]
    for k, s in enumerate(systems):
        with sg.Engine(s.params, precision=prec, rebuild_every=0, metal_lanes=4, xcd_map=0,
                       records_capacity=1) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            ref = []
            for step in range(steps):
                if step > 0 and rebuilt[step]:
                    eng.set_run_state(step, 1, ref[-1]["dispHi"])
                ref += eng.run(1)
            r, v, _ = eng.download()
        assert [x["rebuilt"] for x in ref] == rebuilt, k
        assert np.array_equal(out[k][0][:, :3], r[:, :3]), k
        assert np.array_equal(out[k][1], v), k
        assert [x["totEnergy"] for x in recs[k]] == [x["totEnergy"] for x in ref], k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_batch_with_lanes_per_replica_equals_replicas_run_alone(gpu, prec):
    """Replicas with different lanes per metal atom in one batch (one density/force launch per lane
    count): each equals the replica run alone with its own lanes, bit for bit."""
    systems = replicas()
    lanes = [16, 2, 8, 4]
    steps = 30
    print(f"[batch] {prec}: lanes per replica {lanes}", flush=True)
    with sg.Batch([s.params for s in systems], precision=prec, rebuild_every=10, xcd_map=0) as bt:
        bt.set_lanes(lanes)
        assert bt.replica_lanes() == lanes and bt.metal_lanes == 0
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        recs = bt.run(steps)
        out = [bt.download(k) for k in range(len(systems))]
    for k, s in enumerate(systems):
        with sg.Engine(s.params, precision=prec, rebuild_every=10, metal_lanes=lanes[k], xcd_map=0) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            ref = eng.run(steps)
            r, v, a = eng.download()
        assert np.array_equal(out[k][0][:, :3], r[:, :3]), k
        assert np.array_equal(out[k][1], v), k
        assert [x["totEnergy"] for x in recs[k]] == [x["totEnergy"] for x in ref], k


@pytest.mark.gpu
def test_batch_autotune_per_size_bucket(gpu):
    """The autotuner tries every lane count for each replica-size bucket and keeps the fastest; the
    forces it leaves equal the untuned ones to FP64 summation order (the lanes change only the order
    in which a metal atom's pair terms are added)."""
    systems = replicas()
    with sg.Batch([s.params for s in systems], precision="mixed", rebuild_every=10, xcd_map=0) as bt:
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        bt.run(10, records=False)
        before = [bt.download(k)[2] for k in range(len(systems))]
        table = bt.autotune(5)
        after = [bt.download(k)[2] for k in range(len(systems))]
        lanes = bt.replica_lanes()
        recs = bt.run(20)
    buckets = {(e[0], e[1]) for e in table}
    metal = [s.params.nMolMe for s in systems]
    assert len(buckets) == len({next(b for b in buckets if b[0] <= m < b[1]) for m in metal})
    assert len(table) == 5 * len(buckets)  # every lane count per bucket
    for lo, hi in buckets:  # the chosen lanes are the fastest of the bucket's entries
        best = min((e for e in table if (e[0], e[1]) == (lo, hi)), key=lambda e: e[5])[4]
        assert all(lanes[k] == best for k, m in enumerate(metal) if lo <= m < hi)
    for b, a in zip(before, after):
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-12 * scale
    assert all(np.isfinite(x["totEnergy"]) and x["errors"] == 0 for rr in recs for x in rr)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_batch_sync_rebuilds_equals_replicas_with_the_same_rebuild_steps(gpu, prec):
    """sync_rebuilds: a step rebuilds every replica's lists when any replica's own rule asked for it
    (the records keep each replica's own decision).  Each replica then equals the replica run alone
    with its list rebuilt at those same steps, bit for bit."""
    systems = replicas()
    steps = 60
    print(f"[batch] {prec}: sync_rebuilds", flush=True)
    with sg.Batch([s.params for s in systems], precision=prec, rebuild_every=0, metal_lanes=4, xcd_map=0,
                  sync_rebuilds=1) as bt:
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        recs = bt.run(steps)
        out = [bt.download(k) for k in range(len(systems))]
    rebuilt = np.array([[x["rebuilt"] for x in r] for r in recs])
    own = np.array([[x["nebrNow"] for x in r] for r in recs])
    assert (rebuilt == rebuilt[0]).all()
    assert rebuilt[0, 0] == 1 and np.array_equal(rebuilt[0, 1:], own[:, :-1].any(0))
    assert 1 < rebuilt[0].sum() and (own[:, :-1].sum(1) < rebuilt[0, 1:].sum()).any()  # schedules differ
    for k, s in enumerate(systems):
        with sg.Engine(s.params, precision=prec, rebuild_every=0, metal_lanes=4, xcd_map=0,
                       records_capacity=1) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            ref = []
            for step in range(steps):
                if step > 0 and rebuilt[0, step]:
                    eng.set_run_state(step, 1, ref[-1]["dispHi"])
                ref += eng.run(1)
            r, v, _ = eng.download()
        assert [x["rebuilt"] for x in ref] == rebuilt[0].tolist(), k
        assert np.array_equal(out[k][0][:, :3], r[:, :3]), k
        assert np.array_equal(out[k][1], v), k
        assert [x["totEnergy"] for x in recs[k]] == [x["totEnergy"] for x in ref], k


@pytest.mark.gpu
def test_batch_sync_rebuilds_over_more_replicas_than_a_workgroup_has_lanes(gpu):
    """The OR of the rebuild flags runs in the last replica workgroup of the closing kernel, 256
    threads striding over the replicas.  With 300 replicas - the first 256 cold, the last 44 hot -
    steps occur where only replicas past the first stride asked for a rebuild, and every replica must
    still rebuild on exactly the steps after one where any replica asked."""
    n_cold, n_hot = 256, 44
    systems = [sg.config_system("small", "config", ux=3, uy=3, n_metal=13,
                                temperature=300.0 if k < n_cold else 1500.0, seed=100 + k)
               for k in range(n_cold + n_hot)]
    steps = 40
    print(f"[batch] sync_rebuilds over {len(systems)} replicas", flush=True)
    with sg.Batch([s.params for s in systems], precision="mixed", rebuild_every=0, metal_lanes=4, xcd_map=0,
                  sync_rebuilds=1) as bt:
        for k, s in enumerate(systems):
            bt.upload(k, s.r, s.v, s.a)
            bt.set_run_state(k, 0, 1, 0.0)
        recs = bt.run(steps)
        second = bt.run(steps)  # the counter is back at 0 after every launch
    for rr in (recs, second):
        rebuilt = np.array([[x["rebuilt"] for x in r] for r in rr])
        own = np.array([[x["nebrNow"] for x in r] for r in rr])
        assert (rebuilt == rebuilt[0]).all()
        assert np.array_equal(rebuilt[0, 1:], own[:, :-1].any(0))
        assert all(x["errors"] == 0 for r in rr for x in r)
    own = np.array([[x["nebrNow"] for x in r] for r in recs])
    # steps that only the hot replicas (indices 256..299) asked for
    assert (own[n_cold:, :-1].any(0) & ~own[:n_cold, :-1].any(0)).any()
