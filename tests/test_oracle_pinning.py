"""Pins the CPU oracle against what the mathematics and the published numbers fix, independently of
the oracle's own formulas (CPU only).

  EAM        fcc shell sums: perfect-lattice energy per atom of Ni, Au and Cu at a = re*sqrt(2) with
             the 54 neighbours inside 1.45 a, and rho(Ni); zero force on the perfect lattice;
             phi', f', F' against central differences of phi, f, F; F continuous at its branch
             points; forces = -grad of the total EAM + LJ energy (both are conservative).
  known answer  bulk Au 298 K (paper Fig. 8): total energy -3.8914285 eV/atom, conserved in NVE.
  LJ         two atoms against 4 eps ((s/r)^12 - (s/r)^6) and its derivative; cutoff.
  graphene   zero force and energy on the ideal sheet; first-order response to one out-of-plane
             displacement (plane and neighbour-angle terms, SurfaceGrowth.cu:1003-1011, 975).
  lists      neighbour lists against brute force O(N^2) with the reference order (27 offsets in
             table order, ascending index within a cell); RDF histogram against brute force for
             r below one cell width; perfect-lattice RDF shell counts 12/6/24/12.
  integrator leapfrog with constant acceleration (closed form), a harmonic dimer (energy error
             O(dt^2), no drift), frozen carbon perimeter, periodic wrap.
  thermostat Berendsen lambda = sqrt(1 + gamma (T0/T - 1)), T = 2 KE / (3 kB).
  reductions hand-computed sums with the reference's quirks.
  diffusion  rigid translation: D(t) = |u|^2 t / 6 per unit time, across periodic jumps.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg
from surfacegrowthconsole_amd import _native as N

MATERIALS = {"Cu": 0, "Ag": 1, "Au": 2, "Ni": 3, "Al": 4, "Pb": 5}


def bulk(material, k=6, temperature=298.0):
    return sg.make_system(regime=0, material=MATERIALS[material], ux=k, uy=k, uz=k, temperature=temperature,
                          geometry="reference")


def with_box(sp, L, n, n_metal):
    """Copy of SimParams with another periodic box and atom count; cells as SurfaceGrowth.cpp:276-285."""
    q = N.SimParams.from_buffer_copy(sp)
    L = [float(np.float32(x)) for x in L]
    q.region.x, q.region.y, q.region.z = L
    w = q.rCutEam + q.rNebrShell
    cells = [int((1.0 / w) * x) for x in L]
    q.cells.x, q.cells.y, q.cells.z = cells
    q.invWidth.x, q.invWidth.y, q.invWidth.z = [float(np.float32(np.float32(c) / np.float32(x)))
                                                for c, x in zip(cells, L)]
    q.nMol, q.nMolMe = n, n_metal
    return q


def fcc(a, k):
    """k^3 fcc unit cells of edge a, centred in [-L/2, L/2)."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.array([[i, j, l] for i in range(k) for j in range(k) for l in range(k)], float)
    x = (g[:, None, :] + basis[None]).reshape(-1, 3) * a - 0.5 * k * a + 0.25 * a
    return x


def pos4(x):
    r = np.zeros((len(x), 4))
    r[:, :3] = x
    return r


# --------------------------------------------------------------------------------------------
# EAM


# Perfect-lattice energies from the fcc shell sums of the Zhou et al. functions with the
# reference parameters (shells at a/sqrt2, a, a sqrt(3/2), a sqrt2: 12 + 6 + 24 + 12 = 54).
SHELL_SUM = {"Ni": -4.449961, "Au": -3.929978, "Cu": -3.539741}


@pytest.mark.parametrize("mat", ["Ni", "Au", "Cu"])
def test_eam_fcc_shell_sum_energy(mat):
    s = bulk(mat)
    sp = s.params
    a = float(np.float32(sp.re * math.sqrt(2.0)))  # float lattice edge so that L = 8 a is exact
    k = 8
    x = fcc(a, k)
    q = with_box(sp, [k * a] * 3, len(x), len(x))
    o = oracle.Oracle(q, "f64")
    r = pos4(x)
    err, nn, nbl = o.neighbour_list(r)
    assert err == 0
    rho = o.rho(r, nn, nbl)
    acc, e, _ = o.forces(r, nn, nbl, rho)
    E = e.mean() * sp.enU
    assert abs(E - SHELL_SUM[mat]) < 2e-6 * abs(SHELL_SUM[mat]), E
    assert np.ptp(e) < 1e-12 * abs(e.mean())
    assert np.abs(acc).max() < 1e-11
    if mat == "Ni":
        assert abs(rho.mean() - 27.98446) < 2e-6 * 27.98446
    # 54 neighbours inside the EAM cutoff
    nbl = nbl.reshape(-1, len(x))
    i = 0
    d = x[nbl[: nn[i], i]] - x[i]
    d -= k * a * np.round(d / (k * a))
    assert int((np.sqrt((d ** 2).sum(1)) < sp.rCutEam).sum()) == 54


def test_bulk_au_known_answer_and_nve():
    """Paper Fig. 8: bulk Au at 298 K, totEn = -3.8914285 eV/atom, constant in NVE.  Here 6^3 cells
    instead of 24^3; the kinetic term differs by O(1/N)."""
    s = bulk("Au", 6, 298.0)
    sp = s.params
    o = oracle.Oracle(sp, "f64")
    r = s.r.astype(np.float64)
    err, nn, nbl = o.neighbour_list(r)
    acc, e, _ = o.forces(r, nn, nbl)
    pr = o.props(r, s.v.astype(np.float64))
    tot0 = (e.mean() + pr["kinEnergy"]) * sp.enU
    assert abs(tot0 - (-3.8914285)) < 5e-5 * 3.8914285, tot0
    err, r2, v2, a2, recs = o.run(s.r, s.v, acc, 200)
    assert err == 0
    tot = np.array([x["totEnergy"] for x in recs]) * sp.enU
    assert np.abs(tot - tot0).max() < 2e-5 * abs(tot0)
    # no thermostat in the bulk regime: the momentum the FP32 initial velocities carry is conserved
    assert np.abs(np.array(recs[-1]["vSum"]) - np.array(pr["vSum"])).max() < 1e-10


@pytest.mark.parametrize("mat", ["Ni", "Au"])
def test_eam_derivatives_are_central_differences(mat):
    sp = bulk(mat).params
    o = oracle.Oracle(sp, "f64")
    h = 1e-6
    for r in np.linspace(0.75 * sp.re, sp.rCutEam - 0.01, 25):
        phi, dphi, f, df = o.pair(float(r))
        p1, p0 = o.pair(float(r + h)), o.pair(float(r - h))
        assert abs(dphi - (p1[0] - p0[0]) / (2 * h)) < 1e-6 * max(1.0, abs(dphi))
        assert abs(df - (p1[2] - p0[2]) / (2 * h)) < 1e-6 * max(1.0, abs(df))
    rhoe = sp.rhoe
    for rho in np.concatenate([np.linspace(0.3, 0.84, 6), np.linspace(0.86, 1.14, 6), np.linspace(1.16, 1.6, 6)]):
        rho = float(rho * rhoe)
        F, dF = o.embed(rho)
        Fp, _ = o.embed(rho + 1e-5)
        Fm, _ = o.embed(rho - 1e-5)
        assert abs(dF - (Fp - Fm) / 2e-5) < 1e-6 * max(1.0, abs(dF))
    # F is continuous where its branches meet (0.85 rho_e and 1.15 rho_e), to the table's 6 digits
    for x in (0.85, 1.15):
        lo, _ = o.embed(x * rhoe * (1 - 1e-12))
        hi, _ = o.embed(x * rhoe * (1 + 1e-12))
        assert abs(lo - hi) < 1e-5 * abs(lo)


def cluster_with_carbons(seed=3):
    """A perturbed Ni fcc cluster plus isolated carbon atoms (no two within rCutC, so no bonded
    terms) inside the LJ range, in a box large enough that no image interacts."""
    s = bulk("Ni", 6)
    sp = s.params
    rng = np.random.default_rng(seed)
    a = sp.re * math.sqrt(2.0)
    x = fcc(a, 3)
    x = x[np.sqrt((x ** 2).sum(1)) < 1.3 * a]
    x = x + rng.uniform(-0.06, 0.06, x.shape) * a
    top = x[:, 2].max()
    cs = []
    for gx in (-1.5, 0.0, 1.5):
        for gy in (-1.5, 1.5):
            cs.append([gx + rng.uniform(-0.2, 0.2), gy + rng.uniform(-0.2, 0.2), top + 1.5 + rng.uniform(0, 0.4)])
    c = np.array(cs)
    L = 40.0
    q = with_box(sp, [L, L, L], len(x) + len(c), len(x))
    return q, pos4(np.concatenate([x, c]))


def test_forces_are_the_gradient_of_eam_plus_lj_energy():
    q, r = cluster_with_carbons()
    o = oracle.Oracle(q, "f64")
    err, nn, nbl = o.neighbour_list(r)
    assert err == 0
    acc, e, _ = o.forces(r, nn, nbl)
    mass = np.where(np.arange(q.nMol) < q.nMolMe, q.massMe, 1.0)[:, None]
    F = acc * mass
    assert np.abs(F).max() > 1.0  # the check is not vacuous
    h = 1e-6
    worst = 0.0
    for i in list(range(0, q.nMolMe, 7)) + list(range(q.nMolMe, q.nMol)):
        for d in range(3):
            rp, rm = r.copy(), r.copy()
            rp[i, d] += h
            rm[i, d] -= h
            Ep = o.forces(rp, nn, nbl)[1].sum()
            Em = o.forces(rm, nn, nbl)[1].sum()
            worst = max(worst, abs(-(Ep - Em) / (2 * h) - F[i, d]))
    assert worst < 1e-6 * np.abs(F).max(), worst
    # Newton's third law for the pair forces: zero net force on the whole system
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()


def test_lennard_jones_two_atoms():
    sp = bulk("Ni").params
    s = sp.sigmaLJ
    eps = sp.epsLJ
    q = with_box(sp, [30.0, 30.0, 30.0], 2, 1)
    o = oracle.Oracle(q, "f64")
    F0, _ = o.embed(0.0)  # a lone metal atom keeps its embedding energy at rho = 0
    for r in np.linspace(0.9 * s, 2.45 * s, 12):
        x = np.array([[0.0, 0.0, 0.0, 0.0], [r * 0.6, r * 0.8, 0.0, 0.0]])
        err, nn, nbl = o.neighbour_list(x)
        acc, e, fr = o.forces(x, nn, nbl)
        inside = r * r < min(sp.rrCutLJ, sp.rrNebr)  # 2.5 sigma, limited by the list radius
        U = 4 * eps * ((s / r) ** 12 - (s / r) ** 6) if inside else 0.0
        dU = 4 * eps * (-12 * s ** 12 / r ** 13 + 6 * s ** 6 / r ** 7) if inside else 0.0
        assert abs((e[0] - F0) + e[1] - U) < 1e-12 + 1e-10 * abs(U)
        fc = -dU * np.array([0.6, 0.8, 0.0])  # force on the carbon, along +r
        assert np.allclose(acc[1], fc, rtol=1e-10, atol=1e-12)
        assert np.allclose(acc[0] * q.massMe, -fc, rtol=1e-10, atol=1e-12)
    # zero force at the minimum 2^(1/6) sigma
    r = 2 ** (1 / 6) * s
    x = np.array([[0.0, 0.0, 0.0, 0.0], [r, 0.0, 0.0, 0.0]])
    err, nn, nbl = o.neighbour_list(x)
    acc, e, _ = o.forces(x, nn, nbl)
    assert abs(acc[1, 0]) < 1e-10 * 24 * eps / s


# --------------------------------------------------------------------------------------------
# graphene


def sheet(ux=4, uy=4):
    s = sg.make_system(regime=2, ux=ux, uy=uy, n_metal=0, temperature=300.0)
    return s


def test_ideal_graphene_sheet_has_zero_force_and_energy():
    s = sheet()
    sp = s.params
    o = oracle.Oracle(sp, "f64")
    r = s.r.astype(np.float64)
    err, nn, nbl = o.neighbour_list(r)
    assert err == 0
    acc, e, _ = o.forces(r, nn, nbl)
    # positions are the FP32 generator's: bond lengths are 1 to ~1e-7
    assert np.abs(acc).max() < 2e-6 * sp.mu_r
    assert np.abs(e).max() < 1e-9 * sp.mu_r


def test_graphene_out_of_plane_response():
    """Displace one carbon by dz: to first order the plane term gives -(4/3) mu_p dz on it and
    +(2/3) mu_p dz on each of its three bonded neighbours (SurfaceGrowth.cu:1003-1005), and the
    neighbour-angle terms give -mu_p dz / 9 on each of its six second neighbours (.cu:975)."""
    s = sheet()
    sp = s.params
    o = oracle.Oracle(sp, "f64")
    r0 = s.r.astype(np.float64)
    err, nn, nbl = o.neighbour_list(r0)
    acc0, _, _ = o.forces(r0, nn, nbl)
    i = len(r0) // 2 + 3
    dz = 1e-5
    r = r0.copy()
    r[i, 2] += dz
    acc, _, _ = o.forces(r, nn, nbl)
    dF = (acc - acc0)[:, 2] / (sp.mu_p * dz)
    d = r0[:, :3] - r0[i, :3]
    L = np.array([sp.region.x, sp.region.y, sp.region.z])
    d -= L * np.round(d / L)
    dist = np.sqrt((d ** 2).sum(1))
    first = np.where(np.abs(dist - 1.0) < 1e-3)[0]
    second = np.where(np.abs(dist - math.sqrt(3.0)) < 1e-3)[0]
    assert len(first) == 3 and len(second) == 6
    assert abs(dF[i] + 4.0 / 3.0) < 1e-4
    assert np.allclose(dF[first], 2.0 / 3.0, atol=1e-4)
    assert np.allclose(dF[second], -1.0 / 9.0, atol=1e-4)
    rest = np.setdiff1d(np.arange(len(r0)), np.concatenate([[i], first, second]))
    assert np.abs(dF[rest]).max() < 1e-4
    # in-plane components are second order in dz
    assert np.abs(acc - acc0)[:, :2].max() < 1e-3 * sp.mu_p * dz


# --------------------------------------------------------------------------------------------
# neighbour lists and RDF

OFFS = [(a, b, c) for a in (0, 1, -1) for b in (0, 1, -1) for c in (0, 1, -1)]  # ComputeDefs.h:221-234


def small_hot():
    s = sg.config_system("small", "reference")
    rng = np.random.default_rng(7)
    r = s.r.copy()
    r[:, :3] += rng.uniform(-0.15, 0.15, (len(r), 3)).astype(np.float32)
    return s, r


def test_neighbour_list_matches_brute_force_and_reference_order():
    s, r = small_hot()
    p = s.params
    o = oracle.Oracle(p, "f32")
    err, cell_of, cnt, slots = o.bin(r)
    assert err == 0
    err, nn, nbl = o.nebr(r, cnt, slots)
    assert err == 0
    n = p.nMol
    nbl = nbl.reshape(-1, n)
    x = r[:, :3].astype(np.float32)
    L = np.array([p.region.x, p.region.y, p.region.z], np.float32)
    cells = np.array([p.cells.x, p.cells.y, p.cells.z])
    # cell coordinates by the reference formula, in FP32
    cc = ((x + np.float32(0.5) * L) * np.array([p.invWidth.x, p.invWidth.y, p.invWidth.z], np.float32)).astype(int)
    assert np.array_equal((cc[:, 2] * cells[1] + cc[:, 1]) * cells[0] + cc[:, 0], cell_of)
    for i in range(0, n, 5):
        d = x[i] - x
        d = np.where(d >= np.float32(0.5) * L, d - L, np.where(d < np.float32(-0.5) * L, d + L, d)).astype(np.float32)
        rr = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        brute = set(np.where((rr < np.float32(p.rrNebr)) & (np.arange(n) != i))[0].tolist())
        lst = nbl[: nn[i], i].tolist()
        assert set(lst) == brute and len(lst) == len(brute), i
        # order: offset index in table order, then ascending atom index
        keys = []
        for k in lst:
            delta = (cc[k] - cc[i] + 1) % cells - 1
            keys.append((OFFS.index(tuple(int(v) for v in delta)), k))
        assert keys == sorted(keys), i


def test_rdf_matches_brute_force_below_one_cell_width():
    s, r = small_hot()
    p = s.params
    o = oracle.Oracle(p, "f32")
    err, cell_of, cnt, slots = o.bin(r)
    hist, edge = o.rdf(r, cell_of, cnt, slots)
    nMe = p.nMolMe
    x = r[:nMe, :3].astype(np.float32)
    L = np.array([p.region.x, p.region.y, p.region.z], np.float32)
    ref = np.zeros_like(hist)
    for i in range(nMe):
        d = x[i] - x
        d = np.where(d >= np.float32(0.5) * L, d - L, np.where(d < np.float32(-0.5) * L, d + L, d)).astype(np.float32)
        rr = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        ok = (rr < np.float32(p.rangeRdf) ** 2) & (np.arange(nMe) != i)
        b = (np.sqrt(rr[ok]) * np.float32(p.intervalRdf)).astype(int)
        np.add.at(ref, b[b < len(ref)], 1)
    width = min(p.region.x / p.cells.x, p.region.y / p.cells.y, p.region.z / p.cells.z)
    nb = int(width * p.intervalRdf)
    assert hist.sum() > 0
    assert np.array_equal(hist[:nb], ref[:nb])


def test_rdf_perfect_lattice_shell_counts():
    sp = bulk("Ni", 6).params
    a = float(np.float32(sp.re * math.sqrt(2.0)))
    k = 8
    x = fcc(a, k)
    q = with_box(sp, [k * a] * 3, len(x), len(x))
    q.rangeRdf = float(np.float32(1.5 * a))
    q.sizeHistRdf = 300
    q.intervalRdf = float(np.float32(300 / q.rangeRdf))
    o = oracle.Oracle(q, "f32")
    r = pos4(x).astype(np.float32)
    err, cell_of, cnt, slots = o.bin(r)
    hist, edge = o.rdf(r, cell_of, cnt, slots)
    n = len(x)
    for shell, count in ((a / math.sqrt(2), 12), (a, 6), (a * math.sqrt(1.5), 24), (a * math.sqrt(2), 12)):
        b = int(shell * q.intervalRdf)
        assert hist[b - 1: b + 2].sum() == count * n, (shell, hist[b - 1: b + 2])
    assert hist.sum() == 54 * n


# --------------------------------------------------------------------------------------------
# integrator, thermostat, reductions, diffusion


def test_leapfrog_constant_acceleration_and_wrap():
    s = sheet(3, 3)
    p = s.params
    o = oracle.Oracle(p, "f64")
    rng = np.random.default_rng(1)
    n = p.nMol
    r = s.r.astype(np.float64)
    r[:, 2] += 0.3
    v = rng.normal(size=(n, 3))
    a = rng.normal(size=(n, 3))
    dt = p.deltaT
    r1, v1 = o.leapfrog(1, r, v, a)
    r2, v2 = o.leapfrog(2, r1, v1, a)
    Lx = p.region.x
    cos30 = float(np.float32(math.cos(math.pi / 6)))
    xl, xr = np.float32(0.5 * (cos30 - Lx)), np.float32(0.5 * (Lx - cos30))
    yb, yt = np.float32(0.5 * (1 - p.region.y)), np.float32(0.5 * (p.region.y - 1))
    frozen = ~((r[:, 0] != xl) & (r[:, 0] < xr) & (r[:, 1] != yt) & (r[:, 1] != yb))
    assert frozen.sum() > 0 and (~frozen).sum() > 0
    mv = ~frozen
    # velocity Verlet: r(dt) = r + v dt + a dt^2/2, v(dt) = v + a dt (constant a), modulo the box
    L = np.array([p.region.x, p.region.y, p.region.z])
    want = r[mv, :3] + v[mv] * dt + 0.5 * a[mv] * dt * dt
    got = r2[mv, :3]
    dd = got - want
    assert np.abs(dd - L * np.round(dd / L)).max() < 1e-12
    assert np.all((got >= -0.5 * L) & (got < 0.5 * L))
    assert np.allclose(v2[mv], v[mv] + a[mv] * dt, rtol=0, atol=1e-12)
    assert np.array_equal(r2[frozen, :3], r[frozen, :3]) and np.array_equal(v2[frozen], v[frozen])


def test_leapfrog_harmonic_dimer_energy_error_is_second_order():
    s = bulk("Ni", 2)
    p = s.params
    q = with_box(p, [50.0, 50.0, 50.0], 2, 2)
    o = oracle.Oracle(q, "f64")
    w = 2.0
    k = w * w

    def acc(r):
        d = r[0, :3] - r[1, :3]
        return np.array([-0.5 * k * d, 0.5 * k * d])

    def energy(r, v):
        d = r[0, :3] - r[1, :3]
        return 0.5 * (v ** 2).sum() + 0.25 * k * (d ** 2).sum()

    errs = []
    for dt in (0.02, 0.01):
        q.deltaT = dt
        o = oracle.Oracle(q, "f64")
        r = np.array([[0.5, 0, 0, 0], [-0.5, 0, 0, 0]], float)
        v = np.array([[0.0, 0.3, 0], [0.0, -0.3, 0]])
        a = acc(r)
        E0 = energy(r, v)
        worst = 0.0
        for _ in range(int(20.0 / dt)):
            r, v = o.leapfrog(1, r, v, a)
            a = acc(r)
            r, v = o.leapfrog(2, r, v, a)
            worst = max(worst, abs(energy(r, v) - E0))
        errs.append(worst / E0)
    assert errs[0] < 1e-3
    assert 3.5 < errs[0] / errs[1] < 4.5  # O(dt^2)


def test_berendsen_lambda():
    s = sg.config_system("small", "reference")
    p = s.params
    o = oracle.Oracle(p, "f64")
    rng = np.random.default_rng(2)
    v = rng.normal(size=(p.nMol, 3)) * 0.1
    vv = float(1.7 * (v ** 2).sum())
    ke = 0.5 * vv / p.nMol
    T = 2.0 * ke / (3.0 * p.kB)
    lam = math.sqrt(1.0 + p.gammaBerendsen * (p.temperature / T - 1.0))
    v2 = o.thermostat(v, vv, p.stepEquil + 1)
    C_ = np.arange(p.nMol) >= p.nMolMe
    assert np.allclose(v2[C_], lam * v[C_], rtol=1e-12)
    assert np.array_equal(v2[~C_], v[~C_])  # metal only inside the cooling window


def test_reductions_with_reference_quirks():
    s = sg.config_system("small", "reference")
    p = s.params
    o = oracle.Oracle(p, "f64")
    n, nMe, m = p.nMol, p.nMolMe, p.massMe
    rng = np.random.default_rng(4)
    r = s.r.astype(np.float64)
    r[:, 3] = rng.normal(size=n)
    v = rng.normal(size=(n, 3))
    fr = rng.normal(size=nMe)
    pr = o.props(r, v, fr)
    me = np.arange(n) < nMe
    assert np.isclose(pr["vSum"][0], m * v[me, 0].sum() + v[~me, 0].sum(), rtol=1e-12)
    assert np.allclose(pr["vSum"][1:], v[:, 1:].sum(0), rtol=1e-12)
    assert np.isclose(pr["vvMax"], (v ** 2).sum(1).max(), rtol=1e-14)
    assert np.isclose(pr["uSum"], r[:, 3].sum(), rtol=1e-12)
    assert np.allclose(pr["comSum"], r[me, :3].sum(0), rtol=1e-12)
    assert np.allclose(pr["cmVelSum"], v[me].sum(0), rtol=1e-12)
    assert np.isclose(pr["frict"], fr.sum(), rtol=1e-12)
    # bounding box includes the zero padding of the reference reduction
    assert np.allclose(pr["bboxMin"], np.minimum(r[me, :3].min(0), 0.0))
    assert np.allclose(pr["bboxMax"], np.maximum(r[me, :3].max(0), 0.0))
    # regime 2: metal x velocity relative to the cluster's mean x velocity
    vx = v[:, 0].copy()
    vx[me] -= v[me, 0].mean()
    vv = m * ((vx[me] ** 2) + (v[me, 1:] ** 2).sum(1)).sum() + (vx[~me] ** 2 + (v[~me, 1:] ** 2).sum(1)).sum()
    assert np.isclose(pr["vvSum"], vv, rtol=1e-12)
    assert np.isclose(pr["kinEnergy"], 0.5 * vv / n, rtol=1e-12)


def test_diffusion_rigid_translation():
    L = np.array([20.0, 30.0, 40.0], np.float32)
    nVal, nBuff, stepD, lim, dt = 10, 5, 4, 3, 0.005
    lib = oracle.lib()
    d = lib.orc_diffusion_create(nVal, nBuff, stepD, lim, dt, L.ctypes.data_as(C.c_void_p))
    u = np.array([0.37, -0.21, 0.05])
    out = np.zeros(nVal)
    tables = []
    x0 = np.array([9.0, -14.0, 1.0])
    for t in range(200):
        x = x0 + u * t
        x = x - L * np.floor((x + 0.5 * L) / L)  # the cluster crosses the box edges
        if lib.orc_diffusion_sample_f64(d, x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)):
            tables.append(out.copy())
    lib.orc_diffusion_destroy(d)
    assert tables
    j = np.arange(1, nVal)
    want = (u @ u) * (j * j) / (6.0 * j * stepD * np.float64(np.float32(dt)))
    for tab in tables:
        assert tab[0] == 0.0
        assert np.allclose(tab[1:], want, rtol=1e-9)


def test_deposition_schedule_and_parking():
    """Regime 1 (InsertAtomsK, SurfaceGrowth.cu:488-514, schedule .cu:2148-2168): before the end of
    equilibration every metal atom waits above the box at rest; afterwards a group of
    nMolToDeposit atoms enters at z = 0.49 Lz with vz = -velMagDepos every stepDeposit steps."""
    s = sg.make_system(regime=1, material=3, ux=4, uy=4, n_metal=12, temperature=300.0, step_equil=3,
                       depos_energy=0.03, n_to_deposit=5, step_deposit=4, dt=0.005)
    p = s.params
    assert p.nMolDeposited == 0
    o = oracle.Oracle(p, "f64")
    r, v, a = s.r.astype(np.float64), s.v.astype(np.float64), s.a.astype(np.float64)
    dep, nebr, dh = 0, 1, 0.0
    Lz = np.float32(p.region.z)
    for step in range(1, 16):
        err, r, v, a, rec = o.run(r, v, a, 1, step - 1, nebr, dh, 0, dep)
        assert err == 0
        nebr, dh = rec[0]["nebrNow"], rec[0]["dispHi"]
        want = dep + (5 if step > 3 and dep < 12 and (step - 3) % 4 == 0 else 0)
        assert rec[0]["nMolDeposited"] == want, step
        if want > dep:  # the new group entered this step at rest force (parked: no neighbours):
            g = np.arange(dep, min(want, 12))  # z = 0.49 Lz - dt velMagDepos after the drift
            z = float(np.float32(0.49) * Lz) - p.deltaT * float(np.float32(p.velMagDepos))
            assert np.allclose(r[g, 2], z, atol=1e-9)
        dep = want
        for n in range(dep, 12):  # parked: one periodic wrap of the parked height, at rest
            park = float(np.float32(1.5) * Lz + np.float32(0.5) * np.float32(n + 1) * Lz)
            assert r[n, 2] == pytest.approx(park - float(Lz), rel=1e-6)
            assert np.all(v[n] == 0.0)
    assert dep == 15  # groups at steps 7, 11, 15 (the last group exceeds nMolMe, as in the reference)


def test_threaded_fp64_oracle_is_bitwise_the_serial_one():
    """liboracle_omp64 (per-atom loops on OpenMP threads, the checker's flags) gives the serial FP64
    oracle's run bit for bit: the long FP64 reference runs may use it."""
    s = sg.config_system("small", "config")
    o = oracle.Oracle(s.params, "f64")
    ot = oracle.Oracle(s.params, "f64", library=oracle.threaded_f64_lib())
    a = o.run(s.r, s.v, s.a, 40, 0, 1, 0.0, 10)
    b = ot.run(s.r, s.v, s.a, 40, 0, 1, 0.0, 10)
    assert a[0] == b[0] == 0
    assert all(np.array_equal(x, y) for x, y in zip(a[1:4], b[1:4]))
    assert a[4] == b[4]
