"""The adjoint on the GPU (sgk_adjoint_*, sgk_force_vjp) against the tangent build of the FP64
oracle (forward-mode derivatives of the same algorithm, pinned by central differences in
tests/test_oracle_adjoint.py).

Tolerances:
  - force-pass vector-Jacobian product at identical FP64 positions and identical lists: both sides
    are FP64 derivatives of the same expressions, so 1e-9 relative (summation order only);
  - window observables and gradients: the GPU integrates the mixed-precision form, whose graphene
    force law is FP32 (relative force error ~1e-7 against FP64), so its trajectory - and with it the
    observables and their derivatives - departs slightly from the FP64 oracle's; the bar is 1e-4
    relative, the adjoint's accuracy target, on every gradient component.
"""
import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny():
    s = sg.make_system(regime=2, material=3, ux=2, uy=2, n_metal=13, temperature=600.0, geometry="icosahedron",
                       gap=3.0, velocities="maxwell", seed=3)
    return s, oracle.Oracle(s.params, "f64")


def contact_state(s, o):
    """FP64 oracle state after 60 steps: the cluster sits on the sheet, the sheet is distorted."""
    _, r, v, a, err = o.window(s.r.astype(np.float64), s.v.astype(np.float64), 60)
    assert err == 0
    return r, v


def test_force_vjp_matches_tangent(gpu, tiny):
    s, o = tiny
    p = s.params
    r, v = contact_state(s, o)
    e, nn, nbl = o.neighbour_list(r)
    assert e == 0
    rng = np.random.default_rng(5)
    abar = rng.standard_normal((p.nMol, 3))
    w_adh = 0.7
    print(f"[adjoint] force VJP, {p.nMol} atoms", flush=True)
    with sg.Engine(p, precision="mixed") as eng:
        eng.upload(r, v)
        eng.bin()
        eng.build_nebr()
        e_gpu = eng.adhesion_energy()
        rbar, pbar = eng.force_vjp(abar, w_adh)
    assert e_gpu == pytest.approx(o.adhesion(r, nn, nbl), rel=1e-12)
    # parameters: pbar_k = abar . da/ds_k + w dE/ds_k
    _, dacc, _, dadh = o.tangent_forces(r, nn, nbl, pdir=np.eye(4))
    want = np.einsum("nx,knx->k", abar, dacc) + w_adh * dadh
    assert np.allclose(pbar, want, rtol=1e-9, atol=1e-12 * np.abs(want).max()), (pbar, want)
    # positions, every component: (J^T abar)_j = abar . J e_j, four unit directions per tangent call
    n3 = 3 * p.nMol
    full = np.zeros(n3)
    for j0 in range(0, n3, oracle.ND):
        dirs = np.zeros((oracle.ND, p.nMol, 3))
        for k in range(oracle.ND):
            if j0 + k < n3:
                dirs[k].flat[j0 + k] = 1.0
        _, dacc, _, dadh = o.tangent_forces(r, nn, nbl, rdir=dirs)
        for k in range(oracle.ND):
            if j0 + k < n3:
                full[j0 + k] = np.sum(abar * dacc[k]) + w_adh * dadh[k]
    full = full.reshape(p.nMol, 3)
    assert np.abs(rbar - full).max() < 1e-9 * np.abs(full).max()


def window_case(s, o, nsteps, every, rebuild_every=0):
    p = s.params
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    eng = sg.Engine(p, precision="mixed", rebuild_every=rebuild_every)
    eng.upload(r0, v0)
    obs = eng.adjoint_forward(nsteps, every)
    end = eng.download()
    g_adh = eng.adjoint_backward([1.0, 0.0])
    g_h, dr0, dv0 = eng.adjoint_backward([0.0, 1.0], initial_state=True)
    info = eng.adjoint_info()
    after = eng.download()
    return eng, obs, end, np.array([g_adh, g_h]), dr0, dv0, info, after


def test_window_gradient_matches_tangent(gpu, tiny):
    """30 steps in 3 checkpointed segments (a thermostat step, rebuilds inside segments)."""
    s, o = tiny
    p = s.params
    print(f"[adjoint] window 30 steps / 10, {p.nMol} atoms", flush=True)
    eng, obs, end, grad, dr0, dv0, info, after = window_case(s, o, 30, 10)
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    J, dJ, *_ = o.tangent_window(r0, v0, 30, pdir=np.eye(4))
    assert info["mismatches"] == 0 and info["segments"] == 3
    for x, y in zip(end, after):  # the context is left where the forward ended, bitwise
        assert np.array_equal(x, y)
    assert obs["adhesion"] == pytest.approx(J[0], rel=1e-5)
    assert obs["height"] == pytest.approx(J[1], rel=1e-5)
    rel = np.abs(grad - dJ) / np.abs(dJ).max(axis=1, keepdims=True)
    print("[adjoint] gradient rel err", rel.max(), flush=True)
    assert rel.max() < 1e-4, (grad, dJ)
    # initial-state gradients of the height: tangent along metal-position and velocity directions
    rng = np.random.default_rng(9)
    u = rng.standard_normal((oracle.ND, p.nMol, 3))
    u[:, p.nMolMe:] = 0  # carbons start on a cell boundary: see test_oracle_adjoint
    w = rng.standard_normal((oracle.ND, p.nMol, 3))
    _, dJr, *_ = o.tangent_window(r0, v0, 30, rdir=u)
    _, dJv, *_ = o.tangent_window(r0, v0, 30, vdir=w)
    got_r = np.einsum("nx,knx->k", dr0, u)
    got_v = np.einsum("nx,knx->k", dv0, w)
    assert np.allclose(got_r, dJr[1], rtol=1e-4, atol=1e-4 * np.abs(dJr[1]).max()), (got_r, dJr[1])
    assert np.allclose(got_v, dJv[1], rtol=1e-4, atol=1e-4 * np.abs(dJv[1]).max()), (got_v, dJv[1])
    eng.close()


def test_window_gradient_fixed_cadence_and_repeat(gpu, tiny):
    """Rebuild every 4 steps (several list copies per segment), a window that does not divide into
    the checkpoint interval, and a second forward/backward on the same context."""
    s, o = tiny
    p = s.params
    print(f"[adjoint] window 27 steps / 8, rebuild every 4", flush=True)
    eng, obs, end, grad, _, _, info, _ = window_case(s, o, 27, 8, rebuild_every=4)
    assert info["segments"] == 4 and info["mismatches"] == 0 and info["list_copies"] >= 2
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    J, dJ, *_ = o.tangent_window(r0, v0, 27, rebuild_every=4, pdir=np.eye(4))
    rel = np.abs(grad - dJ) / np.abs(dJ).max(axis=1, keepdims=True)
    assert rel.max() < 1e-4, (grad, dJ)
    # continue from the end state: a second window on the same context
    o2 = oracle.Oracle(p, "f64")
    r1, v1, _ = eng.download()
    obs2 = eng.adjoint_forward(12, 5)
    g2 = eng.adjoint_backward([1.0, 0.0])
    J2, dJ2, *_ = o2.tangent_window(r1, v1, 12, step0=27, rebuild_every=4, pdir=np.eye(4))
    assert obs2["adhesion"] == pytest.approx(J2[0], rel=1e-5)
    assert np.abs(g2 - dJ2[0]).max() < 1e-4 * np.abs(dJ2[0]).max(), (g2, dJ2[0])
    eng.close()


def test_adjoint_rejects_what_it_does_not_differentiate(gpu, tiny):
    s, _ = tiny
    with sg.Engine(s.params, precision="fp32") as eng:
        eng.upload(s.r, s.v)
        with pytest.raises(RuntimeError, match="mixed-precision"):
            eng.adjoint_forward(10, 5)
    with sg.Engine(s.params, precision="mixed") as eng:
        eng.upload(s.r.astype(np.float64), s.v.astype(np.float64))
        with pytest.raises(RuntimeError, match="no forward window"):
            eng.adjoint_backward([1.0, 0.0])


def test_bulk_metal_height_gradients_closed_form(gpu):
    """Bulk Ni (regime 0, no thermostat): forces sum to zero pair by pair, so the metal centre of mass
    moves ballistically and J_h = mean z_0 + W dt mean vz_0.  Its gradients are closed forms:
    d J_h/d v0_z = W dt / nMe, d J_h/d r0_z = 1 / nMe, zero for x, y and for every parameter - which
    checks the transpose of the leapfrog and of the momentum-conserving EAM pass without an oracle."""
    s = sg.make_system(regime=0, material=3, ux=4, uy=4, uz=4, temperature=600.0, velocities="maxwell", seed=2)
    p = s.params
    W = 30
    print(f"[adjoint] bulk {p.nMol} atoms, window {W}", flush=True)
    with sg.Engine(p, precision="mixed") as eng:
        eng.upload(s.r.astype(np.float64), s.v.astype(np.float64))
        obs = eng.adjoint_forward(W, 10)
        g, dr0, dv0 = eng.adjoint_backward([0.0, 1.0], initial_state=True)
        info = eng.adjoint_info()
    assert info["mismatches"] == 0
    assert obs["adhesion"] == 0.0
    n = p.nMolMe
    dt = float(np.float32(p.deltaT))
    assert np.abs(g).max() < 1e-12
    assert np.allclose(dv0[:, 2], W * dt / n, rtol=1e-9, atol=0)
    assert np.allclose(dr0[:, 2], 1.0 / n, rtol=1e-9, atol=0)
    assert np.abs(dv0[:, :2]).max() < 1e-12 * W * dt / n and np.abs(dr0[:, :2]).max() < 1e-12 / n
