"""Pins the oracle's graphene force law in the sheet plane and under finite deformation (CPU only).

The reference's carbon branch (SurfaceGrowth.cu:919-1061) is, term by term, the negative gradient
with respect to the atom's own position of a local energy of that atom, its three bonded
neighbours k and their other bonded neighbours m, all held fixed:

    U_i(x) = sum_k  1/2 mu_r (|x - r_k| - 1)^2                                 stretch (.cu:989-994)
           + sum_(k<k') 1/2 mu_theta (theta_kk'(x) - theta_0)^2                central angles (.cu:1013-1058)
           + sum_k sum_m 1/2 mu_theta (acos c~_km(x) - theta_0)^2               neighbour angles (.cu:950-966)

with c~_km(x) = (x - r_k).(r_m - r_k) / (|x - r_k| |r_m - r_k|^2), the reference's division by
|r_km|^2 where |r_km| is meant (.cu:957, 962), plus two terms linear in z that are no gradient:
-(2/3) mu_p (2 dz_i - sum_k dz_k) (.cu:1001-1005) and -mu_p dz_m / 9 per neighbour angle (.cu:975).
Its reported energy is 1/2 [sum_k mu_r (r_k - 1)^2 + mu_p (dz_i - sum_k dz_k / 3)^2
+ sum_(k<k') mu_theta (theta_kk' - theta_0)^2] (.cu:996-997, 1007-1010, 1032, 1045, 1058, 1087).

The checks below come from these statements, not from the oracle's code:
  * one atom displaced in the plane by d: to first order the force on it is
    -(3/2 mu_r + 9/2 mu_theta + 3 mu_theta) d (three unit bonds at 120 degrees: sum e e^T = 3/2 I;
    each central angle has |grad theta| = sqrt(3); each of the six neighbour angles |grad theta| = 1);
    each bonded neighbour feels +mu_r (e.d) e from its stretched bond plus angle terms;
  * a uniformly dilated sheet (bond length 1 + eps): energy 3/2 mu_r eps^2 per carbon, zero force
    (every site has three-fold symmetry);
  * a sheet with every atom displaced at random (finite deformation, in and out of the plane):
    every carbon's force equals -grad U_i by central differences plus the linear z terms, and its
    energy equals the stated sum.
A mutation test builds copies of the oracle with one planted mistake each (flipped stretch sign,
swapped central-angle coefficients, a dropped or sign-flipped neighbour-angle term, the true cosine
in place of the /r^2 quirk, wrong energy factors) and shows that these checks reject every one.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N

ORACLE_DIR = os.path.dirname(os.path.abspath(oracle.__file__))


def sheet(ux=4, uy=4):
    return sg.make_system(regime=2, ux=ux, uy=uy, n_metal=0, temperature=300.0)


def box(sp):
    return np.array([sp.region.x, sp.region.y, sp.region.z], np.float64)


def min_image(d, L):
    """Single-image wrap as SurfaceGrowth.cu:838-849 (one shift per axis)."""
    d = d.copy()
    for ax in range(3):
        h = 0.5 * L[ax]
        d[..., ax] = np.where(d[..., ax] >= h, d[..., ax] - L[ax], np.where(d[..., ax] < -h, d[..., ax] + L[ax], d[..., ax]))
    return d


def bond_table(x, L, rc):
    """For every atom the three atoms within rc (geometry only: on the sheets used here each atom has
    exactly three, so the reference's 'first three in list order' picks the same set)."""
    n = len(x)
    d = min_image(x[:, None, :] - x[None, :, :], L)
    rr = (d ** 2).sum(-1)
    np.fill_diagonal(rr, np.inf)
    nb = [np.where(rr[i] < rc * rc)[0] for i in range(n)]
    assert all(len(b) == 3 for b in nb), "every carbon must have exactly three bonds"
    B = np.array(nb)
    # for each bond (i, k) the two other bonded neighbours m of k
    M = np.array([[[m for m in B[k] if m != i] for k in B[i]] for i in range(n)])
    assert M.shape == (n, 3, 2)
    return B, M


def local_energy(xi, x, B, M, L, sp):
    """U_i(xi) of the module docstring for every atom i at once (xi: (N,3) trial positions of i,
    x: the fixed positions of all other atoms)."""
    mu_r, mu_t, th0 = float(sp.mu_r), float(sp.mu_theta), float(sp.theta_0)
    u = min_image(xi[:, None, :] - x[B], L)                 # (N,3,3): x_i - x_k
    ru = np.sqrt((u ** 2).sum(-1))                          # (N,3)
    U = 0.5 * mu_r * ((ru - 1.0) ** 2).sum(-1)
    for a, b in ((0, 1), (1, 2), (2, 0)):
        c = (u[:, a] * u[:, b]).sum(-1) / (ru[:, a] * ru[:, b])
        U += 0.5 * mu_t * (np.arccos(c) - th0) ** 2
    w = min_image(x[M] - x[B][:, :, None, :], L)            # (N,3,2,3): x_m - x_k
    rw2 = (w ** 2).sum(-1)
    c = (u[:, :, None, :] * w).sum(-1) / (ru[:, :, None] * rw2)
    U += (0.5 * mu_t * (np.arccos(c) - th0) ** 2).sum((-1, -2))
    return U


# The reference writes 2/3 and 1/3 as the float literals 0.666666667f and 0.33333333f (.cu:1004,
# 1008); rounded to float they differ from 2/3 and 1/3 by 1e-8 relative, which shows at the FP64
# tolerances used here.
TWO_THIRDS = float(np.float32(0.666666667))
ONE_THIRD = float(np.float32(0.33333333))


def linear_z_force(x, B, M, sp):
    """The two z terms that are not gradients (.cu:1001-1005 and .cu:975)."""
    z0, mu_p = float(sp.z_0), float(sp.mu_p)
    dz = x[:, 2] - z0
    return -TWO_THIRDS * mu_p * (2.0 * dz - dz[B].sum(-1)) - (mu_p / 9.0) * dz[M].sum((-1, -2))


def stated_energy(x, B, L, sp):
    mu_r, mu_t, mu_p, th0, z0 = (float(sp.mu_r), float(sp.mu_theta), float(sp.mu_p), float(sp.theta_0),
                                 float(sp.z_0))
    u = min_image(x[:, None, :] - x[B], L)
    ru = np.sqrt((u ** 2).sum(-1))
    E = mu_r * ((ru - 1.0) ** 2).sum(-1)
    dz = x[:, 2] - z0
    E += mu_p * (dz - ONE_THIRD * dz[B].sum(-1)) ** 2
    for a, b in ((0, 1), (1, 2), (2, 0)):
        c = (u[:, a] * u[:, b]).sum(-1) / (ru[:, a] * ru[:, b])
        E += mu_t * (np.arccos(c) - th0) ** 2
    return 0.5 * E


def fd_gradient(x, B, M, L, sp, h=1e-5):
    g = np.zeros_like(x)
    for ax in range(3):
        xp, xm = x.copy(), x.copy()
        xp[:, ax] += h
        xm[:, ax] -= h
        g[:, ax] = (local_energy(xp, x, B, M, L, sp) - local_energy(xm, x, B, M, L, sp)) / (2 * h)
    return g


# --------------------------------------------------------------------------------------------
# the checks; each returns a dimensionless error (pass: < 1)

def check_displaced_atom(o, s):
    """First-order response to one in-plane displacement (tolerance 1e-4 of the expected force)."""
    sp = s.params
    r0 = s.r.astype(np.float64)
    err, nn, nbl = o.neighbour_list(r0)
    assert err == 0
    i = len(r0) // 2 + 3
    k_self = 1.5 * sp.mu_r + 7.5 * sp.mu_theta
    worst = 0.0
    for d in (np.array([1.0, 0.0, 0.0]), np.array([0.6, -0.8, 0.0])):
        h = 1e-4
        rp, rm = r0.copy(), r0.copy()
        rp[i, :3] += h * d
        rm[i, :3] -= h * d
        ap, _, _ = o.forces(rp, nn, nbl)
        am, _, _ = o.forces(rm, nn, nbl)
        dF = (ap - am) / (2.0 * h)               # d F / d h, second order cancelled
        worst = max(worst, np.abs(dF[i] + k_self * d).max() / (1e-4 * k_self))
        # every z stays zero in the plane
        worst = max(worst, np.abs(dF[:, 2]).max() / (1e-6 * k_self))
    return worst


def check_dilation(o, s, eps=0.05):
    """Uniform in-plane dilation: energy 3/2 mu_r eps^2 per carbon, zero force."""
    sp = s.params
    q = N.SimParams.from_buffer_copy(sp)
    Lx = float(np.float32(sp.region.x * (1 + eps)))
    Ly = float(np.float32(sp.region.y * (1 + eps)))
    fx, fy = Lx / sp.region.x, Ly / sp.region.y           # per-axis factors that keep the box periodic
    q.region.x, q.region.y = Lx, Ly
    w = q.rCutEam + q.rNebrShell
    q.cells.x, q.cells.y = int((1.0 / w) * Lx), int((1.0 / w) * Ly)
    q.invWidth.x = float(np.float32(np.float32(q.cells.x) / np.float32(Lx)))
    q.invWidth.y = float(np.float32(np.float32(q.cells.y) / np.float32(Ly)))
    od = oracle.Oracle(q, "f64")
    od.l = o.l
    r = s.r.astype(np.float64)
    r[:, 0] *= fx
    r[:, 1] *= fy
    err, nn, nbl = od.neighbour_list(r)
    assert err == 0
    acc, e, _ = od.forces(r, nn, nbl)
    e_expect = 1.5 * sp.mu_r * eps * eps
    # bond lengths of the FP32 generator are 1 to ~1e-6; relative energy error ~ 2e-6 / eps
    err_e = np.abs(e / e_expect - 1.0).max() / 2e-3
    err_f = np.abs(acc).max() / (2e-4 * sp.mu_r * eps)
    return max(err_e, err_f)


def deformed(s, seed=3, amp_xy=0.06, amp_z=0.04):
    rng = np.random.default_rng(seed)
    r = s.r.astype(np.float64)
    r[:, :2] += rng.uniform(-amp_xy, amp_xy, (len(r), 2))
    r[:, 2] += rng.uniform(-amp_z, amp_z, len(r))
    return r


def check_local_gradient(o, s, seed=3):
    """Finite deformation: force = -grad U_i + linear z terms; energy = the stated sum."""
    sp = s.params
    L = box(sp)
    r = deformed(s, seed)
    err, nn, nbl = o.neighbour_list(r)
    assert err == 0
    acc, e, _ = o.forces(r, nn, nbl)
    x = r[:, :3]
    B, M = bond_table(x, L, float(sp.rCutC))
    F = -fd_gradient(x, B, M, L, sp)
    F[:, 2] += linear_z_force(x, B, M, sp)
    scale = np.sqrt((F ** 2).sum(1).mean())
    err_f = np.abs(acc - F).max() / (1e-7 * scale)
    E = stated_energy(x, B, L, sp)
    err_e = np.abs(e - E).max() / (1e-9 * np.abs(E).max())
    return max(err_f, err_e)


CHECKS = (check_displaced_atom, check_dilation, check_local_gradient)


# --------------------------------------------------------------------------------------------
# the oracle as built

def test_graphene_in_plane_first_order_response():
    s = sheet()
    assert check_displaced_atom(oracle.Oracle(s.params, "f64"), s) < 1.0


def test_graphene_bonded_neighbour_stretch_first_order():
    """Displace atom i by h along its bond e (unit, k -> i).  The direction k -> i does not change,
    so no angle at k changes.  Of k's neighbour angles only the two at i through the bond k-i move,
    each by d theta = -h / sqrt(3) (the /r^2 cosine at unit bonds: d c~ = h / 2); their forces on k
    are -(2/3) mu_theta h (f_j - e/2) with f_j the unit vectors i -> m_j, and f_1 + f_2 = e, so they
    cancel.  What is left on k is the pure stretch: dF_k / dh = mu_r e."""
    s = sheet()
    sp = s.params
    o = oracle.Oracle(sp, "f64")
    r0 = s.r.astype(np.float64)
    err, nn, nbl = o.neighbour_list(r0)
    L = box(sp)
    i = len(r0) // 2 + 3
    B, _ = bond_table(r0[:, :3], L, float(sp.rCutC))
    k = B[i][0]
    e = min_image(r0[i, :3] - r0[k, :3], L)
    e /= np.linalg.norm(e)
    h = 1e-4
    rp, rm = r0.copy(), r0.copy()
    rp[i, :3] += h * e
    rm[i, :3] -= h * e
    ap, _, _ = o.forces(rp, nn, nbl)
    am, _, _ = o.forces(rm, nn, nbl)
    dF = (ap - am) / (2 * h)
    assert np.allclose(dF[k], sp.mu_r * e, rtol=0, atol=1e-4 * sp.mu_r)


def test_graphene_uniform_dilation_energy_and_force():
    s = sheet()
    assert check_dilation(oracle.Oracle(s.params, "f64"), s) < 1.0


@pytest.mark.parametrize("seed", [3, 11])
def test_graphene_forces_are_local_energy_gradients(seed):
    s = sheet()
    assert check_local_gradient(oracle.Oracle(s.params, "f64"), s, seed) < 1.0


def test_graphene_checks_on_the_fp32_restatement():
    """The FP32 restatement (the CPU baseline and the plain form's reference) obeys the same law to
    FP32 rounding: force error against -grad U_i below 1e-4 of the rms force."""
    s = sheet()
    sp = s.params
    o = oracle.Oracle(sp, "f32")
    L = box(sp)
    r = deformed(s).astype(np.float32).astype(np.float64)
    err, nn, nbl = o.neighbour_list(r)
    acc, e, _ = o.forces(r, nn, nbl)
    x = r[:, :3]
    B, M = bond_table(x, L, float(sp.rCutC))
    F = -fd_gradient(x, B, M, L, sp)
    F[:, 2] += linear_z_force(x, B, M, sp)
    scale = np.sqrt((F ** 2).sum(1).mean())
    assert np.abs(acc - F).max() < 1e-4 * scale


# --------------------------------------------------------------------------------------------
# mutation test: each planted mistake fails at least one check

SRC = os.path.join(ORACLE_DIR, "oracle_impl.c")

# (name, [(text, replacement, occurrence)]) - occurrence counts from 0 among identical lines
MUTANTS = [
    ("stretch sign", [("REAL fc = -mu_r * (absRnn[b] - K(1.)) / absRnn[b];",
                       "REAL fc = mu_r * (absRnn[b] - K(1.)) / absRnn[b];", 0)]),
    ("swapped central-angle coefficients fca/fcb",
     [("REAL fcb = tmp * (K(1.) - absRnn[a] * c / absRnn[b])", "REAL fcb = tmp * (K(1.) - absRnn[b] * c / absRnn[a])", 0),
      ("REAL fca = tmp * (K(1.) - absRnn[b] * c / absRnn[a])", "REAL fca = tmp * (K(1.) - absRnn[a] * c / absRnn[b])", 0)]),
    ("central-angle vectors swapped", [("fx += fcb * rnn[b][0];", "fx += fcb * rnn[a][0];", 0),
                                       ("fy += fcb * rnn[b][1];", "fy += fcb * rnn[a][1];", 0),
                                       ("fx += fca * rnn[a][0];", "fx += fca * rnn[b][0];", 0),
                                       ("fy += fca * rnn[a][1];", "fy += fca * rnn[b][1];", 0)]),
    ("central-angle sign", [("REAL tmp = mu_t * dth / SQRT(K(1.) - c * c);",
                             "REAL tmp = -mu_t * dth / SQRT(K(1.) - c * c);", 1)]),
    ("neighbour-angle sign", [("REAL tmp = mu_t * dth / SQRT(K(1.) - c * c);",
                               "REAL tmp = -mu_t * dth / SQRT(K(1.) - c * c);", 0)]),
    ("neighbour-angle second term dropped", [("fc = -tmp * c / (absRnn[counti] * absRnn[counti]);", "fc = 0;", 0)]),
    ("true cosine instead of the /r^2 quirk",
     [("(absRnn[counti] * rr2);\n                        REAL dth", "(absRnn[counti] * SQRT(rr2));\n                        REAL dth", 0),
      ("REAL fc = tmp / (absRnn[counti] * rr2);", "REAL fc = tmp / (absRnn[counti] * SQRT(rr2));", 0)]),
    ("stretch energy halved", [("uSum += mu_r * (absRnn[b] - K(1.)) * (absRnn[b] - K(1.));",
                                "uSum += K(0.5) * mu_r * (absRnn[b] - K(1.)) * (absRnn[b] - K(1.));", 0)]),
    ("central-angle energy dropped", [("uSum += mu_t * dth * dth;", "", 0)]),
    ("stretch on two bonds only", [("for (int b = 0; b < 3; ++b) {", "for (int b = 0; b < 2; ++b) {", 0)]),
]


def _replace_nth(text, old, new, n):
    idx = -1
    for _ in range(n + 1):
        idx = text.find(old, idx + 1)
        assert idx >= 0, f"mutation anchor not found: {old!r}"
    return text[:idx] + new + text[idx + len(old):]


def _mutant(tmp_path, edits):
    text = open(SRC).read()
    for old, new, n in edits:
        text = _replace_nth(text, old, new, n)
    src = tmp_path / "oracle_impl.c"
    src.write_text(text)
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-DORC_F64",
                    f"-I{ORACLE_DIR}", str(src), "-o", str(so), "-lm"], check=True, capture_output=True)
    lib = C.CDLL(str(so))
    vp = C.c_void_p
    for name, (res, args) in {"orc_bin": (C.c_int, [vp] * 5), "orc_nebr": (C.c_int, [vp] * 6),
                              "orc_rho": (None, [vp] * 5), "orc_forces": (None, [vp] * 7)}.items():
        fn = getattr(lib, f"{name}_f64")
        fn.restype, fn.argtypes = res, args
    return lib


@pytest.mark.parametrize("name,edits", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_checks_reject_planted_mistakes(tmp_path, name, edits):
    s = sheet()
    lib = _mutant(tmp_path, edits)
    o = oracle.Oracle(s.params, "f64")
    o.l = lib
    errs = {}
    for chk in CHECKS:
        try:
            errs[chk.__name__] = chk(o, s)
        except AssertionError:    # a mutant may break the sheet itself
            errs[chk.__name__] = float("inf")
    assert max(errs.values()) > 10.0, (name, errs)


def test_unmutated_copy_passes(tmp_path):
    """The mutation harness itself: an unmodified copy built the same way passes every check."""
    s = sheet()
    o = oracle.Oracle(s.params, "f64")
    o.l = _mutant(tmp_path, [])
    for chk in CHECKS:
        assert chk(o, s) < 1.0, chk.__name__


# --------------------------------------------------------------------------------------------
# the engine against the same statements

@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("mixed", 1e-5), ("fp32", 2e-4)])
@pytest.mark.parametrize("seed", [3, 11])
def test_engine_graphene_law_under_finite_deformation(gpu, prec, tol, seed):
    """The engine's carbon forces on a randomly deformed sheet against -grad U_i plus the linear z
    terms, and its per-atom energies against the stated sum: the GPU law checked against the
    mathematics directly, not only against the oracle.  Tolerance: e_F <= 1e-5 (mixed: FP64
    deviations, FP32 law) and 2e-4 (plain FP32) of the rms force."""
    s = sheet()
    sp = s.params
    L = box(sp)
    r = deformed(s, seed)
    if prec == "fp32":
        r = r.astype(np.float32).astype(np.float64)
    x = r[:, :3]
    B, M = bond_table(x, L, float(sp.rCutC))
    F = -fd_gradient(x, B, M, L, sp)
    F[:, 2] += linear_z_force(x, B, M, sp)
    E = stated_energy(x, B, L, sp)
    print(f"[graphene] engine {prec} seed {seed}", flush=True)
    with sg.Engine(sp, precision=prec) as eng:
        eng.upload(r if prec == "mixed" else r.astype(np.float32))
        eng.bin()
        eng.build_nebr()
        eng.compute_forces()
        acc, e, _ = eng.forces()
    scale = np.sqrt((F ** 2).sum(1).mean())
    ef = np.abs(acc - F).max() / scale
    ee = np.abs(e - E).max() / np.abs(E).max()
    print(f"[graphene] {prec}: e_F {ef:.2e}, energy {ee:.2e}", flush=True)
    assert ef < tol
    assert ee < tol
