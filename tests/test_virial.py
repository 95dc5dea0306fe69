"""The pair virial W = 1/2 sum_i sum_k f_ik |d_ik|^2 of the EAM and metal-carbon LJ forces.

The reference computes no virial; it is pinned by the mathematics.  For forces that derive from a
pair-decomposable energy E (EAM: phi(r) and F(rho) with rho a sum over pairs; LJ), scaling every
position and the box by lambda gives dE/dlambda = -W at lambda = 1, i.e. W = -3 V dE/dV.  The CPU
tests check the FP64 oracle's virial against central differences of its own energy (neighbour list
held fixed), one axis at a time so that each scaling is exact in the float box, W = sum_a W_aa, on a
perfect and a strained fcc lattice (Ni, Au) and on a perturbed Ni cluster with scattered carbons
(every carbon term LJ, no graphene bonds); and the two-atom LJ closed form.  The GPU test checks the
virial the engine records for a step (FP64 in the mixed form) against the oracle's at the positions
the step's force pass saw.
"""
import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N

MATERIALS = {"Cu": 0, "Au": 2, "Ni": 3}


def scaled_params(sp, lam):
    """SimParams with the box scaled by lam = (lx, ly, lz), rounded to float as the box is stored."""
    q = N.SimParams.from_buffer_copy(sp)
    q.region.x, q.region.y, q.region.z = [float(np.float32(x * l)) for x, l in
                                          zip((sp.region.x, sp.region.y, sp.region.z), lam)]
    return q


def energy_at(sp, r, nn, nbl, lam):
    """Total energy with the box scaled by lam (float-rounded) and positions by the same exact
    per-axis factors (list fixed).  Returns the energy and the factors."""
    q = scaled_params(sp, lam)
    f = np.array([q.region.x / sp.region.x, q.region.y / sp.region.y, q.region.z / sp.region.z])
    rs = r.copy()
    rs[:, :3] = r[:, :3] * f
    o = oracle.Oracle(q, "f64")
    _, e, _ = o.forces(rs, nn, nbl)
    return e.sum(), f


def virial_check(sp, r, h=1e-5):
    """(oracle W, -sum_a dE/dlambda_a): each axis scaled on its own by its exact float factor, so
    the diagonal components W_aa = -dE/dlambda_a are central differences of a true scaling."""
    o = oracle.Oracle(sp, "f64")
    err, nn, nbl = o.neighbour_list(r.astype(np.float32))
    assert err == 0
    W = o.virial(r, nn, nbl).sum()
    ref = 0.0
    for ax in range(3):
        up, dn = np.ones(3), np.ones(3)
        up[ax] += h
        dn[ax] -= h
        ep, fp = energy_at(sp, r, nn, nbl, up)
        em, fm = energy_at(sp, r, nn, nbl, dn)
        ref -= (ep - em) / (fp[ax] - fm[ax])
    return W, ref


@pytest.mark.parametrize("mat", ["Ni", "Au"])
@pytest.mark.parametrize("strain", [1.0, 1.02, 0.98])
def test_fcc_lattice_virial_equals_minus_dE_dlambda(mat, strain):
    s = sg.make_system(regime=0, material=MATERIALS[mat], ux=5, uy=5, uz=5, temperature=300.0, geometry="reference")
    sp = s.params
    q = scaled_params(sp, (strain, strain, strain))
    r = s.r.astype(np.float64)
    r[:, :3] *= np.array([q.region.x / sp.region.x, q.region.y / sp.region.y, q.region.z / sp.region.z])
    W, ref = virial_check(q, r)
    n = sp.nMol
    # central differences at h = 1e-5: truncation ~h^2 relative to the third derivative
    assert abs(W - ref) < 1e-6 * n + 1e-6 * abs(ref), (W, ref)
    if strain != 1.0:
        assert abs(W) > 1.0 * n  # a strained lattice carries a real stress
        assert (W < 0) == (strain > 1.0)  # stretched: tension (negative virial); compressed: positive


def test_perturbed_cluster_with_carbons_virial():
    """A Ni cluster with randomly displaced atoms and carbons scattered around it (none within the
    bond cutoff of another carbon, so every carbon term is LJ)."""
    s = sg.make_system(regime=2, ux=6, uy=6, n_metal=201, temperature=300.0, geometry="sphere", gap=2.5, seed=3)
    sp = s.params
    rng = np.random.default_rng(11)
    r = s.r.astype(np.float64)
    nMe = sp.nMolMe
    r[:nMe, :3] += rng.uniform(-0.08, 0.08, (nMe, 3))
    # keep a sparse subset of carbons: every 7th atom of the sheet, lifted to the cluster's base
    keep = np.concatenate([np.arange(nMe), nMe + np.arange(0, sp.nMol - nMe, 7)])
    r = r[keep]
    zmin = r[:nMe, 2].min()
    r[nMe:, 2] = zmin - 1.6 + rng.uniform(-0.2, 0.2, len(r) - nMe)
    q = N.SimParams.from_buffer_copy(sp)
    q.nMol = len(r)
    rc = r[nMe:, :3]
    d = rc[:, None, :] - rc[None, :, :]
    dd = np.sqrt((d ** 2).sum(-1)) + np.eye(len(rc)) * 10
    assert dd.min() > q.rCutC  # no carbon bonds: the carbon energies are LJ halves only
    W, ref = virial_check(q, r)
    assert abs(W) > 1.0
    assert abs(W - ref) < 1e-7 * q.nMol + 1e-6 * abs(ref), (W, ref)


def test_virial_sign_and_closed_form_for_two_lj_atoms():
    """One Ni and one C: W = f r^2 with f = 48 eps s^6 (s^6 - 1/2) / r^2 (force factor), the closed
    form of -r dU/dr for U = 4 eps (s^12 - s^6)."""
    s = sg.make_system(regime=2, ux=4, uy=4, n_metal=13, temperature=300.0)
    sp = s.params
    q = N.SimParams.from_buffer_copy(sp)
    q.nMol, q.nMolMe = 2, 1
    o = oracle.Oracle(q, "f64")
    sig, eps = float(q.sigmaLJ), float(q.epsLJ)
    for rr in (0.95 * sig, 1.2 * sig, 1.6 * sig):
        x = np.array([[0.0, 0.0, 0.0, 0.0], [rr * 0.6, rr * 0.8, 0.0, 0.0]])
        err, nn, nbl = o.neighbour_list(x)
        vir = o.virial(x, nn, nbl)
        inside = rr * rr < min(q.rrCutLJ, q.rrNebr)
        sr6 = (sig / rr) ** 6
        W = 24 * eps * sr6 * (2 * sr6 - 1) if inside else 0.0   # -r dU/dr
        assert abs(vir.sum() - W) < 1e-10 * max(1.0, abs(W)), (rr, vir, W)
        assert abs(vir[0] - vir[1]) < 1e-12 * max(1.0, abs(W))  # each side holds half


# --------------------------------------------------------------------------------------------
# engine

def _thermalised(name, steps):
    s = sg.config_system(name, "config")
    o = oracle.Oracle(s.params, "f64")
    err, r, v, a, _ = o.run(s.r, s.v, s.a, steps)
    assert err == 0
    return s, r, v


@pytest.mark.gpu
# mixed: FP64 pair terms and sum, but F'(rho) is held in float (its rounding, ~6e-8 of the embedding
# part of each EAM term, is the same in the force); plain: FP32 throughout
@pytest.mark.parametrize("prec,tol", [("mixed", 1e-6), ("fp32", 2e-4)])
@pytest.mark.parametrize("name", ["small", "icosa"])
def test_engine_virial_matches_oracle(gpu, prec, tol, name):
    """One device step from a thermalised state: the step's force pass runs at the drifted positions
    on a list rebuilt from them, so the recorded virial is the oracle's at the downloaded positions."""
    s, r, v = _thermalised(name, 100)
    p = s.params
    print(f"[virial] {name} {prec}", flush=True)
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(r if prec == "mixed" else r.astype(np.float32), v if prec == "mixed" else v.astype(np.float32))
        eng.set_run_state(100, 1, 0.0)
        rec = eng.run(1)[0]
        r1, _, _ = eng.download()
    o = oracle.Oracle(p, "f64")
    err, nn, nbl = o.neighbour_list(r1.astype(np.float32))
    assert err == 0
    W = o.virial(r1, nn, nbl).sum()
    print(f"[virial] {name} {prec}: gpu {rec['virial']:.12e} oracle {W:.12e}", flush=True)
    assert abs(rec["virial"] - W) <= tol * abs(W)
