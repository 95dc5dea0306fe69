"""GPU parity: every stage of the HIP engine against the CPU oracle on the same inputs.

Integer results (cell index, cell lists, neighbour list and its order, carbon lists, RDF histogram)
must be bit-exact.  Floating point uses the force-error metric
e_F = max_i |F_i - F_i^ref| / rms_j |F_j^64| on the same positions and neighbour list: the mixed form
against the FP64 oracle (<= 1e-5), the plain FP32 form against the FP32 restatement (<= 2e-4).
"""
import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg

pytestmark = pytest.mark.gpu

_CACHE = {}


def system(name):
    if name not in _CACHE:
        cfg, variant = {"small": ("small", "reference"), "icosa": ("icosa", "config"),
                        "medium": ("medium", "config")}[name]
        print(f"[parity] building {name}", flush=True)
        _CACHE[name] = sg.config_system(cfg, variant)
    return _CACHE[name]


def thermalised(name, steps=200):
    """State after `steps` FP64-oracle steps, so forces are away from the perfect lattice."""
    key = (name, "hot", steps)
    if key not in _CACHE:
        s = system(name)
        o = oracle.Oracle(s.params, "f64")
        err, r, v, a, _ = o.run(s.r, s.v, s.a, steps)
        assert err == 0
        _CACHE[key] = (r, v, a)
    return _CACHE[key]


def force_error(acc, ref):
    rms = np.sqrt((ref ** 2).sum(1).mean())
    return np.abs(acc - ref).max() / rms


def check_lists(p, prec, scheme, r, ordered=False, xcd_map=None):
    o = oracle.Oracle(p, "f32")
    opts = {} if xcd_map is None else {"xcd_map": xcd_map}
    with sg.Engine(p, precision=prec, keep_full_nl=True, nl_scheme=scheme, **opts) as eng:
        eng.upload(r)
        if ordered:
            eng.build_nl_order()
        eng.bin()
        cell_of, cnt, slots = eng.get_cells()
        err, ocell, ocnt, oslots = o.bin(r)
        assert err == 0
        assert np.array_equal(cell_of, ocell)
        assert np.array_equal(cnt, ocnt)
        assert np.array_equal(slots, oslots)
        eng.build_nebr()
        nn, nbl = eng.get_nebr()
        err, onn, onbl = o.nebr(r, ocnt, oslots)
        assert err == 0
        assert np.array_equal(nn, onn)
        assert np.array_equal(nbl, onbl)
        if p.nMol > p.nMolMe:
            check_carbon_lists(eng, p, np.asarray(r, np.float32), onn, onbl)


@pytest.mark.parametrize("name", ["small", "icosa", "medium"])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("scheme", [3, 2, 1, 0])
def test_cells_and_neighbour_list_bit_exact(gpu, name, prec, scheme):
    """Every list scheme (lanes per metal atom and a cell-ordered walk, LDS-staged per workgroup, one
    wave per cell, one thread per atom) gives the reference list."""
    s = system(name)
    print(f"[parity] cells/NL {name} {prec} scheme {scheme}", flush=True)
    check_lists(s.params, prec, scheme, s.r)


@pytest.mark.parametrize("name", ["small", "icosa", "medium"])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_cell_ordered_neighbour_list_bit_exact(gpu, name, prec):
    """nl_scheme 3 with its list-build order made from the same positions."""
    s = system(name)
    print(f"[parity] cells/NL {name} {prec} scheme 3, cell order", flush=True)
    check_lists(s.params, prec, 3, s.r, ordered=True)


@pytest.mark.parametrize("name", ["icosa", "medium"])
@pytest.mark.parametrize("scheme", [3, 2, 0])
def test_neighbour_list_bit_exact_thermalised(gpu, name, scheme):
    """On a thermalised state the sheet straddles a cell boundary in z and the cluster has moved:
    more occupied cells per workgroup box, and the metal workgroups exceed the LDS budget, so the
    staged scheme runs both its LDS path and its fallback."""
    s = system(name)
    r, _, _ = thermalised(name)
    print(f"[parity] NL thermalised {name} scheme {scheme}", flush=True)
    check_lists(s.params, "mixed", scheme, r, ordered=scheme == 3)


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_cell_ordered_lists_with_the_xcd_remap(gpu, prec):
    """xcd_map bit 0: the cell-ordered walk's workgroups, remapped so that each XCD takes a contiguous
    eighth of the metal and of the carbon ranges, build the same lists."""
    s = system("medium")
    r, _, _ = thermalised("medium")
    print(f"[parity] NL cell-ordered with XCD remap {prec}", flush=True)
    check_lists(s.params, prec, 3, r, ordered=True, xcd_map=3)


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("scheme", [3, 0])
@pytest.mark.parametrize("dims,grid", [((2, 2, 2), (1, 1, 1)), ((6, 6, 2), (3, 3, 1)), ((2, 6, 6), (1, 3, 3)),
                                       ((4, 4, 4), (2, 2, 2))])
def test_neighbour_list_bit_exact_wrapped_grid(gpu, prec, scheme, dims, grid):
    """Bulk Ni in cell grids with an axis shorter than 3 cells.  With 1 x 1 x 1 cells all 27 offsets
    wrap onto the atom's own cell with a shift, so every offset holds the atom itself and the other
    atoms' periodic images; with one z layer the three cells of every column are one cell seen with
    z shifts 0, +Lz and -Lz; with 2 cells an offset of +1 and -1 reach the same cell."""
    s = sg.make_system(regime=0, material=3, ux=dims[0], uy=dims[1], uz=dims[2], temperature=600.0,
                       velocities="maxwell", seed=2)
    p = s.params
    assert (p.cells.x, p.cells.y, p.cells.z) == grid
    print(f"[parity] NL wrapped grid {grid} {prec} scheme {scheme}", flush=True)
    check_lists(p, prec, scheme, s.r)


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("shape,grid,n_metal", [((4, 4), (6, 6, 2), 0), ((2, 2), (3, 3, 2), 0),
                                                ((3, 2), (5, 3, 2), 0)])
def test_neighbour_list_bit_exact_graphene_short_z_grid(gpu, prec, ordered, shape, grid, n_metal):
    """Graphene in a box two cells tall: the sheet lies on the z cell boundary, so its atoms fall in
    both z layers, and the z offsets -1 and +1 of every column reach the same cell, once directly and
    once through the periodic face with a z shift.  This drives the carbon column walk through
    repeated cells within one column and nonzero per-candidate z shifts; the (2, 2) sheet also has
    only three cells along x and y.  Every atom is displaced at random so the sheet straddles the
    boundary in both directions."""
    s = sg.make_system(regime=2, ux=shape[0], uy=shape[1], n_metal=n_metal, temperature=300.0)
    p = s.params
    assert (p.cells.x, p.cells.y, p.cells.z) == grid
    rng = np.random.default_rng(5)
    r = s.r.copy()
    r[:, :3] += rng.uniform(-0.08, 0.08, (len(r), 3)).astype(np.float32)
    assert (r[:, 2] < 0).any() and (r[:, 2] >= 0).any()
    print(f"[parity] NL graphene short-z grid {grid} {prec} ordered={ordered}", flush=True)
    check_lists(p, prec, 3, r, ordered=ordered)


@pytest.mark.parametrize("name", ["icosa", "medium"])
def test_cell_ordered_lists_after_run(gpu, name):
    """nl_scheme 3 after a run: the carbon lists are built in the cell order made at the start of the
    run (from positions that have since moved), and still equal the reference list."""
    s = system(name)
    p = s.params
    print(f"[parity] cell-ordered NL after 40 steps, {name}", flush=True)
    with sg.Engine(p, precision="mixed", keep_full_nl=True, rebuild_every=10, nl_scheme=3) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        eng.run(40, records=False)
        r, _, _ = eng.download()
        eng.bin()
        eng.build_nebr()
        nn, nbl = eng.get_nebr()
        o = oracle.Oracle(p, "f32")
        err, ocell, ocnt, oslots = o.bin(r)
        assert err == 0
        err, onn, onbl = o.nebr(r, ocnt, oslots)
        assert err == 0
        assert np.array_equal(nn, onn)
        assert np.array_equal(nbl, onbl)
        check_carbon_lists(eng, p, np.asarray(r, np.float32), onn, onbl)


@pytest.mark.parametrize("name", ["icosa", "medium"])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_list_schemes_same_trajectory(gpu, name, prec):
    """The cell-ordered scheme inside the device-resident step loop (lists built from the new cell
    buffer before the commit, skin-rule and fixed-cadence rebuilds, the order built at the start of
    the run) gives the per-atom scheme's trajectory bit for bit."""
    s = system(name)
    print(f"[parity] list schemes, 300 steps, {name} {prec}", flush=True)
    out = []
    for scheme in (0, 3):
        with sg.Engine(s.params, precision=prec, rebuild_every=10, nl_scheme=scheme) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            eng.run(300, records=False)
            out.append(eng.download())
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def check_carbon_lists(eng, p, r, nn, nbl, lj_carbons_positions_only=False):
    """Carbon lists = the metal entries of the reference list, and the entries closer than the close
    radius, each with its list position, in list order.  lj_carbons_positions_only: compare the close
    entries' list positions only for carbons with LJ partners (the only place positions are read);
    then every carbon with LJ partners is checked, beside 400 drawn at random."""
    cl = eng.carbon_lists()
    N, nMe = p.nMol, p.nMolMe
    nbl = nbl.reshape(-1, N)
    margin = eng.cfg.close_margin
    close_r = np.float32(p.rCutC + 2.0 * (0.5 * p.rNebrShell + margin))
    close2 = np.float32(close_r * close_r)
    L = np.array([p.region.x, p.region.y, p.region.z], np.float32)
    rng = np.random.default_rng(0)
    picks = rng.choice(np.arange(nMe, N), size=min(400, N - nMe), replace=False)
    if lj_carbons_positions_only:
        picks = np.union1d(picks, nMe + np.nonzero(cl["n_lj"])[0])
    for i in picks:
        ic = i - nMe
        ks = nbl[: nn[i], i]
        pos = np.arange(nn[i])
        metal = ks < nMe
        assert np.array_equal(cl["lj_idx"][ic, : cl["n_lj"][ic]], ks[metal])
        assert np.array_equal(cl["lj_pos"][ic, : cl["n_lj"][ic]], pos[metal])
        d = r[i, :3] - r[ks, :3]
        d = np.where(d >= np.float32(0.5) * L, d - L, np.where(d < -np.float32(0.5) * L, d + L, d))
        rr = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        close = rr < close2
        assert np.array_equal(cl["close_idx"][ic, : cl["n_close"][ic]], ks[close])
        if not lj_carbons_positions_only or cl["n_lj"][ic] > 0:
            assert np.array_equal(cl["close_pos"][ic, : cl["n_close"][ic]], pos[close])


@pytest.mark.parametrize("name", ["small", "icosa", "medium"])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_carbon_lists_without_the_reference_layout(gpu, name, prec):
    """The engines of the step loop keep no reference-layout list (keep_full_nl off): their carbon
    lists, built in the list-build order, equal the reference list's entries in order with their
    list positions, and their forces equal those of an engine that also keeps the reference layout,
    bit for bit."""
    r, _, _ = thermalised(name)
    s = system(name)
    p = s.params
    rin = r if prec == "mixed" else np.asarray(r, np.float32)
    o = oracle.Oracle(p, "f32")
    r32 = np.asarray(r, np.float32)
    err, ocell, ocnt, oslots = o.bin(r32)
    err2, onn, onbl = o.nebr(r32, ocnt, oslots)
    assert err == 0 and err2 == 0
    out = []
    for keep in (False, True):
        print(f"[parity] carbon lists {name} {prec} keep_full_nl={keep}", flush=True)
        with sg.Engine(p, precision=prec, keep_full_nl=keep, nl_scheme=3) as eng:
            eng.upload(rin)
            eng.build_nl_order()
            eng.bin()
            eng.build_nebr()
            check_carbon_lists(eng, p, r32, onn, onbl)
            eng.compute_forces()
            out.append(eng.forces())
    for x, y in zip(*out):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name", ["small", "icosa", "medium"])
@pytest.mark.parametrize("state", ["initial", "thermalised"])
def test_forces_vs_oracle(gpu, name, state):
    """Mixed form against the FP64 oracle (e_F <= 1e-5).  Plain form against the FP32 restatement
    of the same arithmetic (e_F <= 2e-4: FP32 acosf of two libms may differ by 1 ulp, which the angle
    force amplifies by mu_theta/sin(theta) ~ 110).  The reference forms x_i - x_k before the periodic wrap, so
    in FP32 a bond across the boundary of a box of edge L carries an error ~ulp(L) times the bond
    stiffness - a property of the reference algorithm, reported here against FP64, not a defect."""
    s = system(name)
    p = s.params
    if state == "initial":
        r = s.r.astype(np.float64)
    else:
        r = thermalised(name, 100 if name == "medium" else 200)[0]
    o64 = oracle.Oracle(p, "f64")
    o32 = oracle.Oracle(p, "f32")
    for prec in ("fp32", "mixed"):
        # the plain form stores float positions: its references are evaluated at those positions
        rin = r if prec == "mixed" else r.astype(np.float32)
        err, nn, nbl = o64.neighbour_list(rin)
        assert err == 0
        rho64 = o64.rho(rin, nn, nbl)
        a64, e64, f64 = o64.forces(rin, nn, nbl, rho64)
        print(f"[parity] forces {name} {state} {prec}", flush=True)
        with sg.Engine(p, precision=prec) as eng:
            eng.upload(rin)
            eng.bin()
            eng.build_nebr()
            eng.compute_forces()
            rho = eng.rho()
            acc, e, fr = eng.forces()
        ef64 = force_error(acc, a64)
        if prec == "mixed":
            print(f"[parity]   e_F vs FP64 = {ef64:.3e}", flush=True)
            assert ef64 < 1e-5, (name, state, ef64)
            if p.nMolMe:
                assert np.abs(rho - rho64).max() / np.abs(rho64).max() < 1e-6
            assert abs(e.sum() - e64.sum()) <= 1e-7 * np.abs(e64).sum()
            if p.nMolMe:
                assert np.abs(fr - f64).max() <= 1e-6 * max(np.abs(f64).max(), 1e-3)
        else:
            a32, e32, f32 = o32.forces(rin, nn, nbl)
            ef32 = np.abs(acc - a32).max() / np.sqrt((a64 ** 2).sum(1).mean())
            print(f"[parity]   e_F vs FP32 restatement = {ef32:.3e}; vs FP64 = {ef64:.3e}", flush=True)
            assert ef32 < 2e-4, (name, state, ef32)
            if p.nMolMe:
                rho32 = o32.rho(rin, nn, nbl)
                assert np.abs(rho - rho32).max() / np.abs(rho32).max() < 1e-5
            assert abs(e.sum() - e32.sum()) <= 1e-5 * np.abs(e32).sum()


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_bulk_metal_forces_across_periodic_faces(gpu, prec):
    """Periodic bulk Ni (5x5x5 fcc cells, 500 atoms, a box 3 list radii wide) with every atom
    displaced at random: nearly every atom lies within one list radius of a face, so its density and
    force pairs take the single-image correction, which the pair displacement applies only when a component reaches half
    the box.  Density and forces against the oracle at the same positions: the mixed form against
    FP64 (e_F <= 1e-5, rho within 1e-6), the plain form against the FP32 restatement (rho within
    1e-5; e_F <= 2e-5, since no angle term amplifies its rounding here)."""
    s = sg.make_system(regime=0, material=3, ux=5, uy=5, uz=5, temperature=300.0, geometry="reference")
    p = s.params
    rng = np.random.default_rng(11)
    r = s.r.astype(np.float64)
    r[:, :3] += rng.uniform(-0.04, 0.04, (p.nMol, 3))
    half = 0.5 * np.array([p.region.x, p.region.y, p.region.z])
    near_face = (np.abs(r[:, :3]) > half - np.sqrt(p.rrNebr)).any(1)
    assert near_face.mean() > 0.5
    rin = r if prec == "mixed" else r.astype(np.float32)
    o64 = oracle.Oracle(p, "f64")
    err, nn, nbl = o64.neighbour_list(rin)
    assert err == 0
    rho64 = o64.rho(rin, nn, nbl)
    a64, e64, _ = o64.forces(rin, nn, nbl, rho64)
    print(f"[parity] bulk Ni {p.nMol} atoms {prec}", flush=True)
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(rin)
        eng.bin()
        eng.build_nebr()
        eng.compute_forces()
        rho = eng.rho()
        acc, e, _ = eng.forces()
    if prec == "mixed":
        ef = force_error(acc, a64)
        assert np.abs(rho - rho64).max() / np.abs(rho64).max() < 1e-6
        assert abs(e.sum() - e64.sum()) <= 1e-7 * np.abs(e64).sum()
    else:
        o32 = oracle.Oracle(p, "f32")
        rho32 = o32.rho(rin, nn, nbl)
        a32, e32, _ = o32.forces(rin, nn, nbl)
        ef = np.abs(acc - a32).max() / np.sqrt((a64 ** 2).sum(1).mean())
        assert np.abs(rho - rho32).max() / np.abs(rho32).max() < 1e-5
        assert abs(e.sum() - e32.sum()) <= 1e-5 * np.abs(e32).sum()
    print(f"[parity]   e_F = {ef:.3e}", flush=True)
    assert ef < (1e-5 if prec == "mixed" else 2e-5), ef


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_integrator_thermostat_props(gpu, prec):
    s = system("small")
    p = s.params
    r, v, a = thermalised("small", 100)
    orc = oracle.Oracle(p, "f32" if prec == "fp32" else "f64")
    if prec == "fp32":
        r, v, a = r.astype(np.float32), v.astype(np.float32), a.astype(np.float32)
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(r, v, a)
        eng.leapfrog(1)
        gr, gv, _ = eng.download()
        orr, ov = orc.leapfrog(1, r, v, a)
        # leapfrog and the boundary wrap are restated operation for operation: bit-exact
        assert np.array_equal(gr[:, :3], orr[:, :3].astype(np.float64))
        assert np.array_equal(gv, ov.astype(np.float64))
        eng.leapfrog(2)
        gr, gv, _ = eng.download()
        _, ov2 = orc.leapfrog(2, orr, ov, a)
        assert np.array_equal(gv, ov2.astype(np.float64))
        # reductions with the reference semantics
        rec = eng.props()
        opr = orc.props(orr, ov2)
        for k in ("vvSum", "uSum"):
            assert abs(rec[k] - opr[k]) <= 1e-5 * max(abs(opr[k]), 1.0), k
        assert abs(rec["vvMax"] - opr["vvMax"]) <= 1e-6 * opr["vvMax"]
        for k in ("vSum", "comSum", "cmVelSum"):
            assert np.allclose(rec[k], opr[k], rtol=1e-5, atol=1e-3), k
        for k in ("bboxMin", "bboxMax"):
            assert np.array_equal(np.float32(rec[k]), np.float32(opr[k])), k
        # thermostat with the same vvSum
        vv = opr["vvSum"]
        eng.thermostat(30, vv)
        _, gv3, _ = eng.download()
        ov3 = orc.thermostat(ov2, vv, 30)
        assert np.allclose(gv3, ov3, rtol=1e-6, atol=0)


@pytest.mark.parametrize("name", ["small", "medium"])
def test_rdf_histogram_bit_exact(gpu, name):
    s = system(name)
    p = s.params
    r = thermalised(name, 50)[0].astype(np.float32)
    o = oracle.Oracle(p, "f32")
    err, cell_of, cnt, slots = o.bin(r)
    hist, edge = o.rdf(r, cell_of, cnt, slots)
    with sg.Engine(p) as eng:
        eng.upload(r)
        eng.bin()
        eng.rdf_clear()
        eng.rdf_accumulate()
        eng.rdf_accumulate()
        g = eng.rdf()
    assert hist.sum() > 0
    assert np.array_equal(g, 2 * hist)


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_device_loop_vs_oracle_run(gpu, prec):
    """sgk_run (device-resident, graph captured) against the oracle's DoComputationsW loop."""
    s = system("small")
    p = s.params
    n = 60
    o = oracle.Oracle(p, "f64" if prec == "mixed" else "f32")
    err, r_o, v_o, a_o, orecs = o.run(s.r, s.v, s.a, n)
    assert err == 0
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        recs = eng.run(n)
        r_g, v_g, _ = eng.download()
    assert [x["step"] for x in recs] == list(range(1, n + 1))
    assert [x["rebuilt"] for x in recs] == [x["rebuilt"] for x in orecs]
    tol = 1e-6 if prec == "mixed" else 2e-4
    for k in ("totEnergy", "potEnergy", "kinEnergy"):
        a = np.array([x[k] for x in recs])
        b = np.array([x[k] for x in orecs])
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), k
    dr = np.abs(r_g[:, :3] - r_o[:, :3]).max()
    assert dr < (1e-6 if prec == "mixed" else 1e-3), dr


def test_graph_and_eager_bitwise_identical(gpu):
    s = system("small")
    out = []
    for graphs in (True, False, True):
        with sg.Engine(s.params, precision="fp32", use_graphs=graphs) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            eng.run(40)
            eng.run(40)
            out.append(eng.download())
    for x in out[1:]:
        for a, b in zip(out[0], x):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_fused_drift_bitwise_equals_the_leapfrog_kernel(gpu, prec):
    """In a chunk of 40 steps the force pass of every pre-kicked step but at most one also drifts
    the positions into the second position buffer.  A chunk of five steps holds four pre-kicked
    steps, all fused; a chunk of four holds three, the first unfused and the next two fused; a chunk
    of two holds one, never fused, so k_lf1 drifts every step there.  All runs agree bit for bit
    with the last: positions with their per-atom energies, velocities, accelerations and every
    record, through the list rebuilds of 80 steps.  The RDF sampled every 10 steps (on fused steps
    too, where it must read the step's positions, not the drifted ones) agrees bin for bin."""
    s = system("small")
    out = []
    for chunk in (40, 5, 4, 2):
        with sg.Engine(s.params, precision=prec) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.set_run_state(0, 1, 0.0)
            eng.rdf_clear()
            recs = []
            for _ in range(80 // chunk):
                recs += eng.run(chunk, rdf_every=10)
            out.append((eng.download(), recs, eng.rdf()))
    d1, r1, h1 = out[-1]
    assert any(x["rebuilt"] for x in r1)
    assert h1.sum() > 0
    for d0, r0, h0 in out[:-1]:
        for a, b in zip(d0, d1):
            assert np.array_equal(a, b)
        assert r0 == r1
        assert np.array_equal(h0, h1)


def test_autotune_keeps_results(gpu):
    """Every launch shape the autotuner tries gives the default shape's forces up to the FP
    summation order of the metal neighbour loop; the tuned engine keeps stepping correctly."""
    s = system("medium")
    p = s.params
    for prec, tol in (("mixed", 1e-12), ("fp32", 2e-6)):
        print(f"[parity] autotune {prec}", flush=True)
        with sg.Engine(p, precision=prec) as eng:
            eng.upload(s.r, s.v, s.a)
            eng.bin()
            eng.build_nebr()
            eng.compute_forces()
            a0, e0, _ = eng.forces()
            table = eng.autotune(reps=3)
            assert len(table) == 20 and all(us > 0 for _, _, us in table)  # 5 lane counts x 4 maps
            a1, e1, _ = eng.forces()
            scale = np.abs(a0).max()
            assert np.abs(a1 - a0).max() <= tol * scale
            assert abs(e1.sum() - e0.sum()) <= tol * np.abs(e0).sum()
            lanes = eng.metal_lanes
            assert lanes in (1, 2, 4, 8, 16)
            eng.upload(s.r, s.v, s.a)  # the run starts from the initial state (zero acceleration)
            eng.set_run_state(0, 1, 0.0)
            recs = eng.run(30)
        with sg.Engine(p, precision=prec, metal_lanes=lanes) as ref:
            ref.upload(s.r, s.v, s.a)
            ref.set_run_state(0, 1, 0.0)
            recs_ref = ref.run(30)
        a = np.array([x["totEnergy"] for x in recs])
        b = np.array([x["totEnergy"] for x in recs_ref])
        # the XCD map regroups the block partials: FP summation order of the energy sums only
        assert np.abs(a - b).max() <= (1e-9 if prec == "mixed" else 1e-6) * np.abs(b).max()


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_device_loop_shear(gpu, prec):
    """Regime 2 past stepEquil + stepCool: ApplyShearK's force on the metal atoms behind the cluster's
    centre of mass, raised by deltaF on every step the cluster's x velocity stays below 0.005
    (SurfaceGrowth.cu:2398-2414), inside the device-resident loop, against the oracle's loop.  On
    these steps the force pass must leave the next half-kick to k_lf1 (the shear changes the
    accelerations after it)."""
    s = sg.make_system(regime=2, material=3, ux=5, uy=6, n_metal=40, temperature=300.0, step_equil=5,
                       step_cool=10, delta_f=0.001, dt=0.005)
    p = s.params
    n = 60
    o = oracle.Oracle(p, "f64" if prec == "mixed" else "f32")
    err, r_o, v_o, a_o, orecs = o.run(s.r, s.v, s.a, n)
    assert err == 0
    print(f"[parity] shear {prec}", flush=True)
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        recs = eng.run(n)
        r_g, v_g, _ = eng.download()
    # the shear sequence the oracle's records imply
    want, sh = [], 0.0
    for x in orecs:
        if x["step"] > p.stepEquil + p.stepCool and x["cmVelSum"][0] / p.nMolMe < 0.005:
            sh += float(np.float32(p.deltaF))
        want.append(sh)
    got = [x["shear"] for x in recs]
    assert want[-1] > 0
    # the mixed form accumulates in FP64, as this sum; the plain form in FP32 (45 additions)
    assert np.allclose(got, want, rtol=1e-12 if prec == "mixed" else 1e-5, atol=0), (got, want)
    assert [x["rebuilt"] for x in recs] == [x["rebuilt"] for x in orecs]
    tol = 1e-6 if prec == "mixed" else 2e-4
    for k in ("totEnergy", "potEnergy", "kinEnergy"):
        a = np.array([x[k] for x in recs])
        b = np.array([x[k] for x in orecs])
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), k
    assert np.abs(r_g[:, :3] - r_o[:, :3]).max() < (1e-6 if prec == "mixed" else 1e-3)
    assert np.abs(v_g - v_o).max() < (1e-6 if prec == "mixed" else 1e-3) * np.abs(v_o).max()


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_device_loop_deposition(gpu, prec):
    """Regime 1: metal atoms wait above the box and enter in groups (InsertAtomsK and its schedule)
    inside the device-resident loop, against the oracle's loop."""
    s = sg.make_system(regime=1, material=3, ux=5, uy=6, n_metal=40, temperature=300.0, step_equil=10,
                       depos_energy=0.03, n_to_deposit=8, step_deposit=15, dt=0.005)
    p = s.params
    n = 80
    o = oracle.Oracle(p, "f64" if prec == "mixed" else "f32")
    err, r_o, v_o, a_o, orecs = o.run(s.r, s.v, s.a, n)
    assert err == 0
    print(f"[parity] deposition {prec}", flush=True)
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        recs = eng.run(n)
        r_g, v_g, _ = eng.download()
    assert [x["nMolDeposited"] for x in recs] == [x["nMolDeposited"] for x in orecs]
    assert recs[-1]["nMolDeposited"] == 32  # groups of 8 at steps 25, 40, 55, 70
    tol = 1e-6 if prec == "mixed" else 2e-4
    for k in ("totEnergy", "potEnergy", "kinEnergy"):
        a = np.array([x[k] for x in recs])
        b = np.array([x[k] for x in orecs])
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), k
    assert np.abs(r_g[:, :3] - r_o[:, :3]).max() < (1e-6 if prec == "mixed" else 1e-3)
