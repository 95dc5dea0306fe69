"""Samples taken on the device inside the step loop: the centre-of-mass diffusion buffers in FP64.

The device keeps the reference's staggered buffers (EvalDiffusion / AccumDiffusion,
SurfaceGrowth.cu:2898-2952) itself; the FP64 oracle restatement of the same buffers, fed the centre
of mass of the same steps (from the step records), must give the same tables bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N

pytestmark = pytest.mark.gpu


def _small_diffusion_params(nval=20, nbuff=4, step=2, limit=3):
    s = sg.config_system("small", "config")
    q = N.SimParams.from_buffer_copy(s.params)
    q.nValDiffuse, q.nBuffDiffuse, q.stepDiffuse, q.limitDiffuseAv = nval, nbuff, step, limit
    return s, q


def _oracle_tables(q, coms):
    lo = oracle.lib()
    L = np.array([q.region.x, q.region.y, q.region.z], np.float32)
    od = lo.orc_diffusion_create(q.nValDiffuse, q.nBuffDiffuse, q.stepDiffuse, q.limitDiffuseAv,
                                 float(np.float32(q.deltaT)), L.ctypes.data_as(C.c_void_p))
    out = np.zeros(q.nValDiffuse)
    tabs = []
    for c in coms:
        c = np.ascontiguousarray(c, np.float64)
        if lo.orc_diffusion_sample_f64(od, c.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)):
            tabs.append(out.copy())
    lo.orc_diffusion_destroy(od)
    return tabs


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_device_diffusion_equals_fp64_oracle_on_the_same_samples(gpu, prec):
    s, q = _small_diffusion_params()
    nsteps = 400
    print(f"[samples] device diffusion, {prec}, {nsteps} steps", flush=True)
    with sg.Engine(q, precision=prec, rebuild_every=10, records_capacity=200) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        eng.diffusion_enable(64)
        recs = eng.run(200) + eng.run(200)
        ntab, tabs = eng.diffusion_tables(64)
    coms = [np.array(r["comSum"]) / q.nMolMe for r in recs if r["step"] % q.stepDiffuse == 0]
    assert len(coms) == nsteps // q.stepDiffuse
    otabs = _oracle_tables(q, coms)
    assert ntab == len(otabs) >= 10
    for k, (t, o) in enumerate(zip(tabs, otabs)):
        assert np.array_equal(t, o), (k, np.abs(t - o).max())
    # D_j is an MSD over a time j*stepDiffuse*dt: non-negative and not all zero (the cluster moves)
    assert (tabs[:, 1:] >= 0).all() and tabs[:, 1:].max() > 0


def test_device_diffusion_off_by_default_and_restart(gpu):
    """Without sgk_diffusion_enable no sample is taken; enabling again restarts the buffers."""
    s, q = _small_diffusion_params()
    with sg.Engine(q, precision="mixed", records_capacity=200) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        with pytest.raises(RuntimeError):
            eng.diffusion_tables()
        eng.run(100, records=False)
        eng.diffusion_enable(8)
        assert eng.diffusion_tables(8)[0] == 0
        eng.run(150, records=False)
        n1 = eng.diffusion_tables(8)[0]
        assert n1 > 0
        eng.diffusion_enable(8)
        assert eng.diffusion_tables(8)[0] == 0


def test_prepared_graph_gives_the_same_run(gpu):
    """sgk_prepare_run captures the next run's graph without running it: the run that follows
    equals a run of an identical engine that captured on the fly, bit for bit."""
    s, q = _small_diffusion_params()
    out = []
    for prepare in (False, True):
        with sg.Engine(q, precision="mixed", rebuild_every=10, records_capacity=100) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            eng.run(7, records=False)
            if prepare:
                eng.prepare_run(60, 10)
                eng.prepare_run(60, 10)  # idempotent
            recs = eng.run(60, 10)
            r, v, a = eng.download()
            out.append((recs, r, v))
    assert [x["totEnergy"] for x in out[0][0]] == [x["totEnergy"] for x in out[1][0]]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
