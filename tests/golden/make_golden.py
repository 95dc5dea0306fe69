"""Writes the golden fixtures of this directory from the FP64 oracle (run from the repository root:
python tests/golden/make_golden.py).

small_forces.npz  the small workload (960 C + 201 Ni sphere, reference variant) at step 0 and after
                  100 oracle steps: inputs r (N,4), v (N,3) and the FP64 outputs acc (N,3),
                  energy (N,), rho (nMe,), friction (nMe,); the FP32 restatement's acc32 beside them.
small_run.npz     40 steps of the reference step loop from step 0: per-step kinetic, potential and
                  total energy and the rebuild flags, with the final r and v.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
import surfacegrowthconsole_amd as sg  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    s = sg.config_system("small", "reference")
    p = s.params
    o64 = oracle.Oracle(p, "f64")
    o32 = oracle.Oracle(p, "f32")
    err, r1, v1, a1, _ = o64.run(s.r, s.v, s.a, 100)
    assert err == 0
    out = {}
    for tag, r, v in (("0", s.r.astype(np.float64), s.v.astype(np.float64)), ("100", r1, v1)):
        err, nn, nbl = o64.neighbour_list(r)
        assert err == 0
        rho = o64.rho(r, nn, nbl)
        acc, e, fr = o64.forces(r, nn, nbl, rho)
        r32 = r.astype(np.float32)
        _, nn32, nbl32 = o32.neighbour_list(r32)
        acc32, _, _ = o32.forces(r32, nn32, nbl32)
        out.update({f"r_{tag}": r, f"v_{tag}": v, f"acc_{tag}": acc, f"energy_{tag}": e, f"rho_{tag}": rho,
                    f"friction_{tag}": fr, f"acc32_{tag}": acc32.astype(np.float64)})
    np.savez_compressed(os.path.join(HERE, "small_forces.npz"), **out)
    err, r, v, a, recs = o64.run(s.r, s.v, s.a, 40)
    assert err == 0
    np.savez_compressed(os.path.join(HERE, "small_run.npz"), r0=s.r, v0=s.v, a0=s.a, r=r, v=v,
                        kin=np.array([x["kinEnergy"] for x in recs]),
                        pot=np.array([x["potEnergy"] for x in recs]),
                        tot=np.array([x["totEnergy"] for x in recs]),
                        rebuilt=np.array([x["rebuilt"] for x in recs], np.int32))


if __name__ == "__main__":
    main()
