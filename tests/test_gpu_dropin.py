"""The drop-in path end to end on the GPU: the console driver (lib/surfacegrowth, the reference's
_tmain flow over CudaInitW / SetParametersW / InitCoordsW / DoComputationsW) reads a sugr_in.txt in
the reference's format and writes sg_*.txt, sugr_*.pdb and rdf_*.txt.  The files are checked
against the oracle's run of the same system: the averaged rows of sg_*.txt (SurfaceGrowth.cu:
2659-2731), the PDB coordinates (2734-2810) and the RDF table layout (2812-2844).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "surfacegrowthconsole_amd", "lib", "surfacegrowth")

ROWS = ["regime, 0 - bulk, 1 - surface growth, 2 - shear", "metal Cu(0), Ag(1), Au(2), Ni(3), Al(4), Pb(5)",
        "unit cells in x", "unit cells in y", "unit cells in z", "temperature (K)",
        "number of metal atoms (if regime is not bulk)", "epsilon for metal-carbon LJ interaction",
        "sigma for metal-carbon LJ interaction", "step limit", "time step", "averaging", "equilibration",
        "increment of shear", "cool interval", "energy of deposited atoms",
        "deposit by groups, containing this number of atoms", "deposit after this number of time steps",
        "create pdb file after this number of steps", "compute rdf after this number of steps",
        "create backup, 0 - no, 1 - yes", "start from backup"]


def write_input(path, values, extra):
    lines = [f"{v:<12}# {c}\t{k + 1}" for k, (v, c) in enumerate(zip(values, ROWS))]
    lines.insert(9, "\n# time parameters\n")
    open(path, "w").write("\n".join(lines + extra) + "\n")


def run_driver(tmp_path, values, extra):
    write_input(tmp_path / "sugr_in.txt", values, extra)
    print(f"[dropin] driver in {tmp_path}", flush=True)
    out = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=150)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def rows_of(path):
    lines = open(path).read().splitlines()
    data = [l for l in lines[1:] if l.strip() and not l.startswith("Duration")]
    return lines, np.array([[float(v) for v in l.split()] for l in data])


def block_means(recs, key, avg):
    x = np.array([np.float32(r[key]) for r in recs], np.float64)
    return x.reshape(-1, avg).mean(1)


def test_console_driver_bulk(gpu, tmp_path):
    """Bulk Ag, the reference input's material and LJ rows, 6x6x6 cells, 200 steps, plain FP32."""
    values = [0, 1, 6, 6, 6, 330, 25000, 0.0087381, 2.4945, 200, 0.001, 50, 10, 0.0001, 120000, 0.03, 40, 1100,
              100, 10, 0, 0]
    stdout = run_driver(tmp_path, values, ["seed=1", "rdf_limit=10"])
    assert "Performing computations" in stdout
    s = sg.make_system(regime=0, material=1, ux=6, uy=6, uz=6, temperature=330.0, epsilon=0.0087381, sigma=2.4945,
                       step_limit=200, dt=0.001, step_avg=50, step_equil=10, delta_f=0.0001, step_cool=120000,
                       step_pdb=100, step_rdf=10, seed=1, geometry="reference", velocities="reference", limitRdf=10)
    p = s.params
    o = oracle.Oracle(p, "f32")
    err, r1, v1, a1, recs1 = o.run(s.r, s.v, s.a, 100)
    err, r2, v2, a2, recs2 = o.run(r1, v1, a1, 100, 100, recs1[-1]["nebrNow"], recs1[-1]["dispHi"])
    recs = recs1 + recs2
    name = tmp_path / "sg_blk_x6_y6_z6_Me864_Av50_Pdb100_T330Ag.txt"
    lines, rows = rows_of(name)
    assert lines[0].startswith("stepCnt\t impulse\t totEn(eV)")
    assert lines[-1].startswith("Duration of the simulation = ")
    assert rows[:, 0].tolist() == [50, 100, 150, 200]
    tot = block_means(recs, "totEnergy", 50) * p.enU
    pot = block_means(recs, "potEnergy", 50) * p.enU
    temp = 2 * block_means(recs, "kinEnergy", 50) * p.temperatureU / (3 * p.kB)
    assert np.allclose(rows[:, 2], tot, rtol=2e-5, atol=2e-7)
    assert np.allclose(rows[:, 4], pot, rtol=2e-5, atol=2e-7)
    assert np.allclose(rows[:, 6], temp, rtol=1e-3)
    # PDB snapshots at steps 0, 100, 200: one ATOM line per atom, coordinates in angstrom (%-8.2lf)
    for k, ref in ((0, s.r), (1, r1), (2, r2)):
        pdb = [l for l in open(tmp_path / f"sugr_{k}.pdb").read().splitlines() if l.startswith("ATOM")]
        assert len(pdb) == p.nMol
        # columns (CreatePdbFile): "ATOM  ", %-7i, "%2s ", " ", " %2s  ", " ", "    ", %-4i, then x y z
        xyz = np.array([[float(l[31 + 8 * d: 39 + 8 * d]) for d in range(3)] for l in pdb])
        assert np.abs(xyz - 1.42 * ref[:, :3]).max() < 0.011, k
        assert pdb[0][13:15] == "Ag"
    # RDF printed after 10 samples, at steps 100 and 200: 200 rows "%8.4f\t %8.4f"
    for step in (100, 200):
        tab = np.loadtxt(tmp_path / f"rdf_stepCount_{step}_Ag.txt")
        assert tab.shape == (p.sizeHistRdf, 2)
        peak = tab[np.argmax(tab[:, 1]), 0]
        assert abs(peak - p.re * 1.42) < 0.15  # first shell of fcc Ag at r_e = 2.89 A


def test_console_driver_cluster_on_graphene_mixed(gpu, tmp_path):
    """201 Ni atoms on a 960-atom sheet (sphere geometry), mixed precision, 150 steps: the regime-2
    rows (energies, cluster centre of mass, size) against the FP64 oracle."""
    values = [2, 3, 5, 6, 0, 300, 201, 0.023, 2.85, 150, 0.005, 50, 0, 0.0, 0, 0.03, 40, 1100, 100000, 50, 0, 0]
    extra = ["seed=1", "geometry=sphere", "gap=2.5", "precision=mixed"]
    run_driver(tmp_path, values, extra)
    s = sg.make_system(regime=2, material=3, ux=5, uy=6, n_metal=201, temperature=300.0, epsilon=0.023, sigma=2.85,
                       step_limit=150, dt=0.005, step_avg=50, step_equil=0, delta_f=0.0, step_cool=0,
                       step_pdb=100000, step_rdf=50, seed=1, geometry="sphere", gap=2.5, velocities="reference")
    p = s.params
    o = oracle.Oracle(p, "f64")
    err, r, v, a, recs = o.run(s.r, s.v, s.a, 150)
    assert err == 0
    name = tmp_path / "sg_sh_x5_y6_Me201_Eq0_C0_Av50_Pdb100000_T300Ni.txt"
    lines, rows = rows_of(name)
    assert rows[:, 0].tolist() == [50, 100, 150]
    tot = block_means(recs, "totEnergy", 50) * p.enU
    assert np.allclose(rows[:, 2], tot, rtol=1e-6, atol=2e-7)
    cm = np.array([np.float32(x["comSum"][0] / p.nMolMe) for x in recs]).reshape(-1, 50).mean(1) * p.lengthU
    assert np.allclose(rows[:, 10], cm, atol=2e-6)
    size_z = np.array([recs[k]["bboxMax"][2] - recs[k]["bboxMin"][2] for k in (49, 99, 149)]) * p.lengthU
    assert np.allclose(rows[:, 14], size_z, atol=2e-6)
    assert os.path.exists(tmp_path / "Diffuse.txt")


def test_console_driver_deposition(gpu, tmp_path):
    """Regime 1 through the console driver: the surface-growth file name (SurfaceGrowth.cpp:432-437)
    and the PDB convention for atoms not yet deposited (z written as 2 Lz, .cu:2787-2790)."""
    values = [1, 3, 5, 6, 0, 300, 40, 0.023, 2.85, 60, 0.005, 20, 10, 0.0, 0, 0.03, 8, 15, 30, 1000, 0, 0]
    run_driver(tmp_path, values, ["seed=1"])
    lines, rows = rows_of(tmp_path / "sg_x5_y6_Me40_Eq10_Dep15_TD8_Pdb30_T300Ni.txt")
    assert rows[:, 0].tolist() == [20, 40, 60]
    assert "deposit energy = 0.030000 eV" in lines[0]
    s = sg.make_system(regime=1, material=3, ux=5, uy=6, n_metal=40, temperature=300.0)
    lz2 = 1.42 * 2 * s.params.region.z
    pdb = [l for l in open(tmp_path / "sugr_1.pdb").read().splitlines() if l.startswith("ATOM")]
    z = np.array([float(l[47:55]) for l in pdb[:40]])
    # at step 30 the groups of steps 25 (8 atoms) are in: the other 32 print at 2 Lz
    assert np.sum(np.abs(z - lz2) < 0.01) == 32


def test_console_driver_restart_from_backup(gpu, tmp_path):
    """Row 22 (start from backup): a run to step 200 with backups every 100 steps, then the same input
    with step limit 300 and row 22 set, continues from the newer backup (step 200) and gives the
    continuous 300-step run's sg_*.txt rows and final backup, bit for bit (plain FP32, list rebuilt
    every 10 steps, so both runs rebuild at step 201)."""
    base = [2, 3, 5, 6, 0, 300, 201, 0.023, 2.85, 300, 0.005, 50, 0, 0.0, 0, 0.03, 40, 1100, 100000, 50, 1, 0]
    extra = ["seed=1", "geometry=sphere", "gap=2.5", "backup_every=100", "rebuild=10"]
    name = "sg_sh_x5_y6_Me201_Eq0_C0_Av50_Pdb100000_T300Ni.txt"
    cont = tmp_path / "continuous"
    cont.mkdir()
    run_driver(cont, base, extra)
    part = tmp_path / "restart"
    part.mkdir()
    first = list(base)
    first[9] = 200
    run_driver(part, first, extra)
    _, rows200 = rows_of(part / name)
    assert rows200[:, 0].tolist() == [50, 100, 150, 200]
    again = list(base)
    again[21] = 1
    run_driver(part, again, extra)
    _, rows = rows_of(part / name)
    _, ref = rows_of(cont / name)
    assert rows[:, 0].tolist() == [250, 300]
    keep = [c for c in range(ref.shape[1]) if c != 8]  # column 8 is the wall time per step
    assert ref[:, 0].tolist() == [50, 100, 150, 200, 250, 300]
    assert np.array_equal(rows[:, keep], ref[4:, keep])
    assert np.array_equal(rows200[:, keep], ref[:4, keep])
    # the step-300 backups (bckup0.sugr in both): SimParams header, then r, v, a byte for byte
    n = 1161
    hdr = os.path.getsize(cont / "bckup0.sugr") - n * (16 + 12 + 12)
    a, b = open(cont / "bckup0.sugr", "rb").read(), open(part / "bckup0.sugr", "rb").read()
    assert len(a) == len(b) and a[hdr:] == b[hdr:]
    # a backup of another system is refused with the reference's error convention
    other = list(again)
    other[6] = 55
    write_input(part / "sugr_in.txt", other, extra)
    out = subprocess.run([DRIVER], cwd=part, capture_output=True, text=True, timeout=150)
    assert out.returncode != 0 and "does not match" in out.stdout
