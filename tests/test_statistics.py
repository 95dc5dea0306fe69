"""Statistical agreement over runs: g(r) and the diffusion tables of GPU runs against the FP64 oracle's
runs.

Small workload (960 C + 201 Ni, 300 K, config variant), eight seeds (independent Maxwell velocities),
5,000 steps each.  Every 10 steps the metal-metal pair histogram is sampled (every pair within
rangeRdf: the engine's complete-range RDF, the brute force over the oracle's positions; below one
cell width this is also the reference stencil's histogram) and the metal centre of mass is fed to
the diffusion buffers (SurfaceGrowth.cu:2898-2952; the device's own buffers, the oracle's
restatement fed from its step records).  Per seed: the mean histogram per sample and the mean of
the diffusion tables; across seeds: means and standard errors; per bin z = difference / combined
standard error.
  - Independent samples: the GPU runs of seeds 1-8 against the oracle's runs of seeds 9-16 must
    agree within statistical error (|z| < 5, mean z^2 < 2.5; the oracle split into two halves of
    its own seeds gives max |z| ~2, mean z^2 ~0.6).
  - Paired runs: over 5,000 steps at 300 K the GPU run of a seed and the oracle's run of the same
    seed stay close (FP64 rounding grows by ~e^5), so their difference is far inside the error.
  - Power: the same statistic separates the GPU runs from the oracle's at 450 K (g(r) only; the
    diffusion tables are dominated by each seed's centre-of-mass drift).
"""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N
from surfacegrowthconsole_amd.configs import rebuild_every

pytestmark = pytest.mark.gpu

SEEDS = range(1, 9)
STEPS, EVERY = 5000, 10
DIFF = dict(nValDiffuse=20, nBuffDiffuse=4, stepDiffuse=EVERY, limitDiffuseAv=5)


def system(seed, temperature=300.0):
    s = sg.config_system("small", "config", seed=seed, temperature=temperature)
    q = N.SimParams.from_buffer_copy(s.params)
    for k, v in DIFF.items():
        setattr(q, k, v)
    return s, q


def pair_histogram(q, r):
    """Every ordered metal pair within rangeRdf: FP32 displacement with the single-image wrap,
    bin = int(sqrt(r^2) * intervalRdf) (the complete-range RDF's definition)."""
    n = q.nMolMe
    x = np.asarray(r[:n, :3], np.float32)
    L = np.array([q.region.x, q.region.y, q.region.z], np.float32)
    d = x[:, None, :] - x[None, :, :]
    half = np.float32(0.5) * L
    d = np.where(d >= half, d - L, np.where(d < -half, d + L, d)).astype(np.float32)
    rr = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    np.fill_diagonal(rr, np.inf)
    sel = rr < np.float32(np.float32(q.rangeRdf) * np.float32(q.rangeRdf))
    b = (np.sqrt(rr[sel]) * np.float32(q.intervalRdf)).astype(np.int64)
    return np.bincount(b[b < q.sizeHistRdf], minlength=q.sizeHistRdf)


def diffusion_tables(q, coms):
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
    return np.array(tabs)


def gpu_run(s, q, prec):
    with sg.Engine(q, precision=prec, rebuild_every=rebuild_every("config"), records_capacity=500,
                   rdf_complete=1) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        eng.diffusion_enable(64)
        eng.rdf_clear()
        for _ in range(STEPS // 500):
            eng.run(500, EVERY, records=False)
        hist = eng.rdf().astype(np.float64)
        ntab, tabs = eng.diffusion_tables(64)
    assert ntab == len(tabs)
    return hist / (STEPS // EVERY), tabs.mean(axis=0)


def oracle_run(s, q):
    o = oracle.Oracle(q, "f64")
    r, v, a = s.r.astype(np.float64), s.v.astype(np.float64), s.a.astype(np.float64)
    hist = np.zeros(q.sizeHistRdf)
    coms = []
    nebr, disp = 1, 0.0
    for c in range(STEPS // EVERY):
        err, r, v, a, recs = o.run(r, v, a, EVERY, step0=c * EVERY, nebr_now=nebr, disp_hi=disp,
                                   rebuild_every=rebuild_every("config"))
        assert err == 0
        nebr, disp = recs[-1]["nebrNow"], recs[-1]["dispHi"]
        hist += pair_histogram(q, r)
        coms.append(np.array(recs[-1]["comSum"]) / q.nMolMe)
    return hist / (STEPS // EVERY), diffusion_tables(q, coms).mean(axis=0)


def oracle_runs(temperature, seeds):
    """The oracle's run of each seed, on threads (the oracle releases the GIL and keeps no shared
    state apart from read-only parameter scales)."""
    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(lambda k: oracle_run(*system(k, temperature)), seeds))


def z_scores(a, b, floor=0.5):
    """Per component: (mean a - mean b) / combined standard error over seeds, where both means
    exceed `floor` (histogram bins: half a count per sample); plus the set of components used."""
    a, b = np.asarray(a), np.asarray(b)
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    use = (np.minimum(ma, mb) > floor) & (se > 0)
    return (ma - mb)[use] / se[use], use


@pytest.fixture(scope="module")
def oracle_300():
    print("[stats] FP64 oracle runs, 300 K, seeds 1-16", flush=True)
    return oracle_runs(300.0, range(1, 17))


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_rdf_and_diffusion_match_the_oracle_within_statistical_error(gpu, prec, oracle_300):
    print(f"[stats] GPU runs ({prec}), {len(SEEDS)} seeds x {STEPS} steps", flush=True)
    g = [gpu_run(*system(k), prec) for k in SEEDS]
    gh, gd = np.array([x[0] for x in g]), np.array([x[1] for x in g])
    oh, od = np.array([x[0] for x in oracle_300]), np.array([x[1] for x in oracle_300])
    n = len(SEEDS)
    for what, other in (("independent seeds", slice(n, 2 * n)), ("same seeds", slice(0, n))):
        z, use = z_scores(gh, oh[other])
        print(f"[stats] {what}: g(r) {use.sum()} bins, max |z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z * z):.2f}",
              flush=True)
        assert use.sum() >= 50
        assert np.abs(z).max() < 5 and np.mean(z * z) < 2.5
        # D_j = MSD(j) / (6 j stepDiffuse dt), j >= 1.  The lags of one seed move together (its centre
        # of mass drifts), so they are one degree of freedom: z of the longest lag and of the mean.
        zd, used = z_scores(np.c_[gd[:, -1], gd[:, 1:].mean(1)], np.c_[od[other, -1], od[other, 1:].mean(1)],
                            floor=0.0)
        print(f"[stats] {what}: D(t) at the longest lag and averaged over lags, z {np.round(zd, 2)}", flush=True)
        assert used.all() and np.abs(zd).max() < 4
    if prec == "mixed":
        # power: the same statistic separates the 300 K runs from the oracle at 450 K
        print("[stats] FP64 oracle runs, 450 K", flush=True)
        hot = np.array([x[0] for x in oracle_runs(450.0, SEEDS)])
        zh, _ = z_scores(gh, hot)
        print(f"[stats] g(r) against 450 K: max |z| {np.abs(zh).max():.2f}, mean z^2 {np.mean(zh * zh):.2f}",
              flush=True)
        assert np.abs(zh).max() > 5 and np.mean(zh * zh) > 2.5
