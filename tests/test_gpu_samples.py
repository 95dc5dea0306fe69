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


# --------------------------------------------------------------------------------------------
# complete-range RDF (extension)

def brute_force_rdf(p, r):
    """Every ordered metal-metal pair within rangeRdf: FP32 displacement with the single-image wrap,
    r^2 = (dx^2 + dy^2) + dz^2, bin = int(sqrt(r^2) * intervalRdf) - the extension's definition."""
    n = p.nMolMe
    x = np.asarray(r[:n, :3], np.float32)
    L = np.array([p.region.x, p.region.y, p.region.z], np.float32)
    d = x[:, None, :] - x[None, :, :]
    half = np.float32(0.5) * L
    d = np.where(d >= half, d - L, np.where(d < -half, d + L, d)).astype(np.float32)
    rr = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    np.fill_diagonal(rr, np.inf)
    range2 = np.float32(np.float32(p.rangeRdf) * np.float32(p.rangeRdf))
    sel = rr < range2
    b = (np.sqrt(rr[sel]) * np.float32(p.intervalRdf)).astype(np.int64)
    return np.bincount(b[b < p.sizeHistRdf], minlength=p.sizeHistRdf).astype(np.int64)


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
@pytest.mark.parametrize("name", ["small", "icosa"])
def test_complete_range_rdf_equals_brute_force(gpu, prec, name):
    """rdf_complete: the histogram of every metal pair within rangeRdf (10 A for the small workload,
    its 1.5 rc for the icosahedron) equals the FP32 brute force over all pairs, bin for bin; below one
    cell width it also equals the reference stencil's histogram (which is complete there)."""
    s = sg.config_system(name, "config")
    p = s.params
    o = oracle.Oracle(p, "f64")
    err, r, v, a, _ = o.run(s.r, s.v, s.a, 150)
    assert err == 0
    rin = r if prec == "mixed" else r.astype(np.float32)
    hists = {}
    for complete in (1, 0):
        print(f"[samples] RDF {name} {prec} complete={complete}", flush=True)
        with sg.Engine(p, precision=prec, rdf_complete=complete) as eng:
            eng.upload(rin)
            eng.bin()
            eng.build_nebr()
            eng.rdf_clear()
            eng.rdf_accumulate()
            eng.rdf_accumulate()
            hists[complete] = eng.rdf().astype(np.int64)
    bf = brute_force_rdf(p, np.asarray(rin, np.float64).astype(np.float32))
    assert np.array_equal(hists[1], 2 * bf)
    width = min(p.region.x / p.cells.x, p.region.y / p.cells.y, p.region.z / p.cells.z)
    inner = int(width * p.intervalRdf)  # bins entirely below one cell width
    assert inner > 10
    assert np.array_equal(hists[0][:inner], hists[1][:inner])
    assert hists[1][inner:].sum() > hists[0][inner:].sum()  # the stencil misses pairs beyond a cell
