"""GPU edge cases: a sheet with no metal atom, a cluster of one metal atom, and capacities too small
for the system.

The two edge systems run through the device loop and the force pass against the oracle.  The mixed
form is compared with the FP64 oracle, the plain form with the FP32 restatement (energies within
1e-6 / 2e-4 relative, e_F within 1e-5 / 2e-4).  A capacity too small for the system is reported by
the run as an error; the kernels clamp every count to its capacity, so nothing is written past it.
"""
import numpy as np
import pytest

import oracle
import surfacegrowthconsole_amd as sg

pytestmark = pytest.mark.gpu


def e_f(acc, ref):
    return np.abs(acc - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def loop_and_forces(s, prec, n=60):
    p = s.params
    o = oracle.Oracle(p, "f64" if prec == "mixed" else "f32")
    err, r_o, _, _, orecs = o.run(s.r, s.v, s.a, n)
    assert err == 0
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        recs = eng.run(n)
    assert [x["rebuilt"] for x in recs] == [x["rebuilt"] for x in orecs]
    assert all(x["errors"] == 0 for x in recs)
    tol = 1e-6 if prec == "mixed" else 2e-4
    for k in ("totEnergy", "potEnergy", "kinEnergy"):
        a = np.array([x[k] for x in recs])
        b = np.array([x[k] for x in orecs])
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), k
    # the force pass at the state the oracle's run ends in
    rin = np.asarray(r_o, np.float64 if prec == "mixed" else np.float32)
    err, nn, nbl = o.neighbour_list(np.asarray(rin, np.float32))
    assert err == 0
    aref, eref, _ = o.forces(rin, nn, nbl)
    with sg.Engine(p, precision=prec) as eng:
        eng.upload(rin)
        eng.bin()
        eng.build_nebr()
        eng.compute_forces()
        acc, e, _ = eng.forces()
    ef = e_f(acc, np.asarray(aref, np.float64))
    assert ef <= (1e-5 if prec == "mixed" else 2e-4), ef
    assert abs(e.sum() - eref.sum()) <= (1e-7 if prec == "mixed" else 1e-6) * abs(eref.sum())
    return recs


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_sheet_without_metal(gpu, prec):
    """nMolMe = 0: no metal workgroups in any pass, no density, and the metal sums of the step
    reduction (centre of mass, bounding box, friction) empty."""
    s = sg.config_system("small", "config", n_metal=0)
    assert s.params.nMolMe == 0
    print(f"[edges] sheet without metal, {prec}", flush=True)
    recs = loop_and_forces(s, prec)
    assert all(x["comSum"] == [0.0, 0.0, 0.0] and x["frict"] == 0.0 for x in recs)


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_single_metal_atom(gpu, prec):
    """A cluster of one atom: its density is that of its carbon-free neighbourhood (zero), so the
    embedding function is evaluated at rho = 0, in its lowest branch; every G-lane group but one is
    idle."""
    s = sg.config_system("small", "config", n_metal=1)
    assert s.params.nMolMe == 1
    print(f"[edges] single metal atom, {prec}", flush=True)
    loop_and_forces(s, prec)


def bulk():
    return sg.make_system(regime=0, material=3, ux=4, uy=4, uz=4, temperature=600.0, velocities="maxwell",
                          seed=2)


def test_cell_capacity_exceeded_is_an_error(gpu):
    """Bulk Ni in 2 x 2 x 2 cells holds 32 atoms per cell: with 8 slots per cell the run fails with
    the capacity named, and the same system runs with the default capacity."""
    s = bulk()
    print("[edges] cell capacity 8 on 32 atoms per cell", flush=True)
    with sg.Engine(s.params, precision="mixed", cell_capacity=8) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        with pytest.raises(RuntimeError, match="cell capacity exceeded"):
            eng.run(5)
    with sg.Engine(s.params, precision="mixed") as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        assert all(x["errors"] == 0 for x in eng.run(5))


def test_neighbour_list_capacity_exceeded_is_an_error(gpu):
    """About 78 neighbours per atom within the list radius of bulk Ni: a capacity of 16 fails with the
    capacity named."""
    s = bulk()
    print("[edges] neighbour-list capacity 16", flush=True)
    with sg.Engine(s.params, precision="fp32", nl_capacity=16) as eng:
        eng.upload(s.r, s.v, s.a)
        eng.set_run_state(0, 1, 0.0)
        with pytest.raises(RuntimeError, match="neighbour-list capacity"):
            eng.run(5)
