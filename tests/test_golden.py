"""Committed golden vectors (tests/golden/, written by make_golden.py from the FP64 oracle).

CPU: the generators reproduce the fixture's initial state bit for bit, and the oracle reproduces
its own recorded outputs (a regression pin for both).  GPU: both engine forms against the vectors,
with the force-error metric e_F = max_i |F_i - F_i^ref| / rms_j |F_j^64|.
"""
import os

import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_generators_reproduce_the_fixture_state():
    g = load("small_forces.npz")
    s = sg.config_system("small", "reference")
    assert np.array_equal(s.r.astype(np.float64), g["r_0"])
    assert np.array_equal(s.v.astype(np.float64), g["v_0"])


@pytest.mark.parametrize("tag", ["0", "100"])
def test_oracle_reproduces_fixture(tag):
    g = load("small_forces.npz")
    s = sg.config_system("small", "reference")
    o = oracle.Oracle(s.params, "f64")
    r = g[f"r_{tag}"]
    err, nn, nbl = o.neighbour_list(r)
    assert err == 0
    rho = o.rho(r, nn, nbl)
    acc, e, fr = o.forces(r, nn, nbl, rho)
    scale = np.abs(g[f"acc_{tag}"]).max()
    assert np.abs(acc - g[f"acc_{tag}"]).max() <= 1e-12 * scale
    assert np.abs(e - g[f"energy_{tag}"]).max() <= 1e-12 * np.abs(g[f"energy_{tag}"]).max()
    assert np.abs(rho - g[f"rho_{tag}"]).max() <= 1e-12 * np.abs(g[f"rho_{tag}"]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["0", "100"])
def test_engine_forces_against_golden_vectors(gpu, tag):
    g = load("small_forces.npz")
    s = sg.config_system("small", "reference")
    ref = g[f"acc_{tag}"]
    rms = np.sqrt((ref ** 2).sum(1).mean())
    print(f"[golden] forces state {tag}", flush=True)
    for prec in ("mixed", "fp32"):
        r = g[f"r_{tag}"] if prec == "mixed" else g[f"r_{tag}"].astype(np.float32)
        with sg.Engine(s.params, precision=prec) as eng:
            eng.upload(r)
            eng.bin()
            eng.build_nebr()
            eng.compute_forces()
            acc, e, fr = eng.forces()
        if prec == "mixed":
            assert np.abs(acc - ref).max() / rms <= 1e-5
            assert abs(e.sum() - g[f"energy_{tag}"].sum()) <= 1e-7 * np.abs(g[f"energy_{tag}"]).sum()
        elif tag == "100":  # the plain form's reference is the FP32 restatement at float positions
            assert np.abs(acc - g[f"acc32_{tag}"]).max() / rms <= 2e-4


@pytest.mark.gpu
def test_engine_run_against_golden_trajectory(gpu):
    g = load("small_run.npz")
    s = sg.config_system("small", "reference")
    print("[golden] 40-step run", flush=True)
    with sg.Engine(s.params, precision="mixed") as eng:
        eng.upload(g["r0"], g["v0"], g["a0"])
        eng.set_run_state(0, 1, 0.0)
        recs = eng.run(40)
        r, v, _ = eng.download()
    assert [x["rebuilt"] for x in recs] == g["rebuilt"].tolist()
    for k in ("kin", "pot", "tot"):
        key = {"kin": "kinEnergy", "pot": "potEnergy", "tot": "totEnergy"}[k]
        a = np.array([x[key] for x in recs])
        assert np.abs(a - g[k]).max() <= 1e-6 * np.abs(g[k]).max(), k
    assert np.abs(r[:, :3] - g["r"][:, :3]).max() < 1e-6
