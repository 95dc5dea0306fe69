"""Pinning the reference gradient: the tangent (forward-mode) build of the FP64 oracle.

The adjoint's reference gradient is oracle/tangent.cpp: the FP64 oracle's own source evaluated on
dual numbers.  These tests pin it on the CPU:
  - its values are bitwise the FP64 oracle's (same operations, same decisions);
  - its parameter derivatives of the window observables equal central finite differences of the
    FP64 oracle with a step small enough (1e-6 relative) that no cutoff or branch event falls
    inside it, to the O(h^2) + round-off bound;
  - its position and initial-velocity derivatives equal central differences the same way, for the
    forces (every term of the force law: EAM, LJ, bonds, plane, both angle families) and for the
    window (leapfrog, thermostat, neighbour-list rebuilds on the way).
A wrong sign, a dropped term or a wrong factor in any derivative fails one of these.
"""
import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg


@pytest.fixture(scope="module")
def tiny():
    # 128 C + 13-atom Mackay icosahedron 3 A above the sheet, 600 K: the cluster lands on the sheet
    # inside the window, so metal-carbon LJ terms switch on and change.
    s = sg.make_system(regime=2, material=3, ux=2, uy=2, n_metal=13, temperature=600.0, geometry="icosahedron",
                       gap=3.0, velocities="maxwell", seed=3)
    return s, oracle.Oracle(s.params, "f64")


def window_fd(o, r0, v0, nsteps, h):
    fd = np.zeros((2, 4))
    for k in range(4):
        Js = []
        for sgn in (1, -1):
            sc = np.ones(4)
            sc[k] = 1 + sgn * h
            o.set_scale(sc)
            Js.append(o.window(r0, v0, nsteps)[0])
        fd[:, k] = (Js[0] - Js[1]) / (2 * h)
    o.set_scale(np.ones(4))
    return fd


def test_tangent_values_are_the_fp64_oracle(tiny):
    s, o = tiny
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    J, r, v, a, err = o.window(r0, v0, 30)
    Jt, dJ, rt, vt, at, errt = o.tangent_window(r0, v0, 30, pdir=np.eye(4))
    assert err == 0 and errt == 0
    assert np.array_equal(J, Jt) and np.array_equal(r, rt) and np.array_equal(v, vt) and np.array_equal(a, at)
    assert np.all(np.isfinite(dJ)) and np.all(dJ != 0)


def test_tangent_parameter_derivatives_match_central_differences(tiny):
    """30 steps (a thermostat step at 25, rebuilds on the way): d J / d scale_k for the adhesion
    energy and the cluster height, all four parameters."""
    s, o = tiny
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    _, dJ, *_ = o.tangent_window(r0, v0, 30, pdir=np.eye(4))
    fd = window_fd(o, r0, v0, 30, 1e-6)
    scale = np.abs(dJ).max(axis=1, keepdims=True)
    assert np.abs(fd - dJ).max() / scale.max() < 1e-6
    assert (np.abs(fd - dJ) / scale).max() < 1e-5, (fd, dJ)


def test_tangent_initial_state_derivatives_match_central_differences(tiny):
    s, o = tiny
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    rng = np.random.default_rng(7)
    # Position directions on the metal atoms only: the sheet lies exactly on a cell boundary in z, so
    # moving a carbon of the initial state by any amount changes its cell and with it the list order
    # its bond search depends on - a discrete event that no derivative sees.
    u = rng.standard_normal((oracle.ND, s.params.nMol, 3))
    u[:, s.params.nMolMe:] = 0
    w = rng.standard_normal((oracle.ND, s.params.nMol, 3))
    _, dJr, *_ = o.tangent_window(r0, v0, 30, rdir=u)
    _, dJv, *_ = o.tangent_window(r0, v0, 30, vdir=w)
    h = 1e-7
    for k in range(oracle.ND):
        rp, rm = r0.copy(), r0.copy()
        rp[:, :3] += h * u[k]
        rm[:, :3] -= h * u[k]
        fd_r = (o.window(rp, v0, 30)[0] - o.window(rm, v0, 30)[0]) / (2 * h)
        fd_v = (o.window(r0, v0 + h * w[k], 30)[0] - o.window(r0, v0 - h * w[k], 30)[0]) / (2 * h)
        assert np.allclose(fd_r, dJr[:, k], rtol=1e-5, atol=1e-7 * np.abs(dJr).max()), (k, fd_r, dJr[:, k])
        assert np.allclose(fd_v, dJv[:, k], rtol=1e-5, atol=1e-7 * np.abs(dJv).max()), (k, fd_v, dJv[:, k])


def test_tangent_force_jacobian_matches_central_differences(tiny):
    """Force Jacobian-vector products with a fixed list, on a state where the cluster touches the
    sheet (LJ active) and the sheet is thermally distorted (all graphene terms non-zero)."""
    s, o = tiny
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    _, r, v, a, err = o.window(r0, v0, 60)
    assert err == 0
    e, nn, nbl = o.neighbour_list(r)
    assert e == 0
    rng = np.random.default_rng(11)
    u = rng.standard_normal((oracle.ND, s.params.nMol, 3))
    acc, dacc, adh, dadh = o.tangent_forces(r, nn, nbl, rdir=u)
    acc64 = o.forces(r, nn, nbl)[0]
    assert np.array_equal(acc, acc64)
    assert adh == o.adhesion(r, nn, nbl) and adh < 0
    h = 1e-7
    for k in range(oracle.ND):
        rp, rm = r.copy(), r.copy()
        rp[:, :3] += h * u[k]
        rm[:, :3] -= h * u[k]
        fd = (o.forces(rp, nn, nbl)[0] - o.forces(rm, nn, nbl)[0]) / (2 * h)
        assert np.abs(fd - dacc[k]).max() < 1e-6 * np.abs(dacc[k]).max(), k
        fde = (o.adhesion(rp, nn, nbl) - o.adhesion(rm, nn, nbl)) / (2 * h)
        assert abs(fde - dadh[k]) < 1e-6 * abs(dadh[k]), (k, fde, dadh[k])
    # parameter directions of the forces and the adhesion energy
    acc, daccp, adh, dadhp = o.tangent_forces(r, nn, nbl, pdir=np.eye(4))
    for k in range(4):
        Fs, Es = [], []
        for sgn in (1, -1):
            sc = np.ones(4)
            sc[k] = 1 + sgn * 1e-6
            o.set_scale(sc)
            Fs.append(o.forces(r, nn, nbl)[0])
            Es.append(o.adhesion(r, nn, nbl))
        o.set_scale(np.ones(4))
        fd = (Fs[0] - Fs[1]) / 2e-6
        assert np.abs(fd - daccp[k]).max() < 1e-6 * np.abs(daccp[k]).max(), k
        fde = (Es[0] - Es[1]) / 2e-6
        assert abs(fde - dadhp[k]) <= 1e-6 * abs(dadhp[k]) + 1e-12, (k, fde, dadhp[k])
    # the adhesion energy is linear in epsilon and carries no EAM parameter
    assert dadhp[0] == pytest.approx(adh, rel=1e-12)
    assert dadhp[2] == 0 and dadhp[3] == 0


def frozen_window_fd(o, r0, v0, nsteps, h, **kw):
    """Central differences of the window with every discrete decision of the unperturbed run held
    (orc_freeze_f64): rebuild steps, lists, cutoff bits, F(rho) branches."""
    o.freeze("record")
    J0 = o.window(r0, v0, nsteps, **kw)[0]
    fd = np.zeros((2, 4))
    for k in range(4):
        Js = []
        for sgn in (1, -1):
            sc = np.ones(4)
            sc[k] = 1 + sgn * h
            o.set_scale(sc)
            o.freeze("replay")
            Js.append(o.window(r0, v0, nsteps, **kw)[0])
            assert o.freeze_ok(), k
        fd[:, k] = (Js[0] - Js[1]) / (2 * h)
    o.set_scale(np.ones(4))
    o.freeze("off")
    return J0, fd


def test_decision_tape_replays_the_recorded_run(tiny):
    """Replaying the tape at the recorded parameters reproduces the run bit for bit; a replay that
    runs longer than the recording reports it."""
    s, o = tiny
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    n = o.freeze("record")
    J, r, v, a, err = o.window(r0, v0, 60)
    assert o.freeze("replay") > n == 0
    Jr, rr, vr, ar, errr = o.window(r0, v0, 60)
    assert o.freeze_ok()
    assert np.array_equal(J, Jr) and np.array_equal(r, rr) and np.array_equal(v, vr) and np.array_equal(a, ar)
    o.freeze("replay")
    o.window(r0, v0, 61)
    assert not o.freeze_ok()
    o.freeze("replay")
    o.window(r0, v0, 30)
    assert not o.freeze_ok()  # a shorter replay leaves part of the tape unread
    o.freeze("off")
    assert np.array_equal(o.window(r0, v0, 60)[0], J)


def test_frozen_differences_at_a_large_step_match_the_tangent(tiny):
    """Over 200 steps at h = 1e-3 pairs cross the LJ cutoff and lists change between the perturbed
    runs, so free central differences miss the tangent by ~5e-2; with the decisions frozen they match
    it to O(h^2)."""
    s, o = tiny
    r0, v0 = s.r.astype(np.float64), s.v.astype(np.float64)
    n = 200
    _, dJ, *_ = o.tangent_window(r0, v0, n, pdir=np.eye(4))
    J0, fd = frozen_window_fd(o, r0, v0, n, 1e-3)
    assert np.array_equal(J0, o.window(r0, v0, n)[0])
    scale = np.abs(dJ).max(axis=1, keepdims=True)
    # O(h^2): the sigma direction enters as sigma^12, whose third derivative gives ~h^2 12^2/6 ~ 2e-5
    assert (np.abs(fd - dJ) / scale).max() < 1e-4, (fd, dJ)
    free = window_fd(o, r0, v0, n, 1e-3)
    assert (np.abs(free - dJ) / scale).max() > 1e-2
