"""The benchmark workloads, restated as reference input rows plus extension options.

Units: dt 0.005 (reduced, time unit 0.2 ps) = 1 fs; epsilon in eV, sigma and gap in angstrom.
The graphene generator makes 32*ux*uy atoms, so a 1,008-atom sheet is not representable: the small
workload uses 5x6 = 960 atoms (4.9 x 5.1 nm).

Two variants per workload:
  "reference": the reference's own rules - fixed-magnitude random-direction velocities, Berendsen
               gamma 0.005 every 25 steps on carbon, adaptive neighbour-list rebuild.
  "config":    Maxwell-Boltzmann velocities, Berendsen tau = 0.1 ps (gamma 0.25 every 25 steps),
               neighbour list rebuilt every 10 steps.
Cutoffs are the reference's in both (EAM 1.45 a, LJ limited by the neighbour-list radius).
"""
from .engine import make_system

_COMMON = dict(regime=2, material=3, epsilon=0.023, sigma=2.85, dt=0.005, step_equil=0, step_cool=0,
               delta_f=0.0, geometry="sphere", gap=2.5)

CONFIGS = {
    # small: 960 C + 201 Ni sphere, 300 K, 1,000 steps, RDF 200 bins to 10 A
    "small": dict(ux=5, uy=6, n_metal=201, temperature=300.0, step_limit=1000, step_rdf=50,
                  rangeRdfAngstrom=10.0, sizeHistRdf=200),
    # medium: 40,320 C + 4,033 Ni sphere (r ~ 2.2 nm), 900 K, 10,000 steps, MSD every 100 steps
    "medium": dict(ux=35, uy=36, n_metal=4033, temperature=900.0, step_limit=10000, step_rdf=50,
                   stepDiffuse=100),
    # differentiation window: 8,064 C + 561-atom Mackay icosahedron 3.0 A above the sheet, 600 K
    "icosa": dict(ux=18, uy=14, n_metal=561, temperature=600.0, step_limit=2000, geometry="icosahedron",
                  gap=3.0),
    # large: 1,003,520 C + 100,000 Ni sphere (r ~ 6.4 nm), 1,000 K, RDF 400 bins to 12 A
    "large": dict(ux=160, uy=196, n_metal=100000, temperature=1000.0, step_limit=2000, step_rdf=50,
                  rangeRdfAngstrom=12.0, sizeHistRdf=400),
    # one x-slab of the large workload cut 8 ways (20 of its 160 sheet columns, 125,440 C), alone and
    # with an eighth of its metal as a 12,500-atom cluster (137,940 atoms, an eighth of the whole):
    # the shards a multi-GPU run of the large workload would hand one GPU, timed on one GPU (no halo)
    "shard8_sheet": dict(ux=20, uy=196, n_metal=0, temperature=1000.0, step_limit=2000, step_rdf=50,
                         rangeRdfAngstrom=12.0, sizeHistRdf=400),
    "shard8": dict(ux=20, uy=196, n_metal=12500, temperature=1000.0, step_limit=2000, step_rdf=50,
                   rangeRdfAngstrom=12.0, sizeHistRdf=400),
}

_VARIANTS = {
    "reference": dict(velocities="reference"),
    "config": dict(velocities="maxwell", tauPs=0.1, rebuildEvery=10),
}


def config_system(name, variant="config", seed=1, **overrides):
    kw = dict(_COMMON)
    kw.update(CONFIGS[name])
    kw.update(_VARIANTS[variant])
    kw.update(overrides)
    kw["seed"] = seed
    return make_system(**kw)


def rebuild_every(variant):
    return _VARIANTS[variant].get("rebuildEvery", 0)
