/*
 * sgk.h - kernel-level C-ABI of the MI355X per-step MD engine (library libsurfacegrowth.so).
 *
 * The reference has no such layer: its wrapper DoComputationsW (SurfaceGrowth.cu:2013-2529)
 * launches the kernels directly.  This layer exposes each stage of the reference step so a host
 * (the drop-in wrappers in surfacegrowth.h, tests, benchmarks) can drive it stage by stage or as a
 * device-resident loop.  Each stage cites the reference kernel whose results it reproduces.
 *
 * All functions return SGK_OK (0) or a negative SGK_E* code; sgk_last_error() gives the message.
 * Host arrays are plain pointers in the reference layouts: positions float4 (x,y,z,E_pot),
 * velocities and accelerations float3, atom order metal [0,nMolMe) then carbon [nMolMe,nMol).
 */
#ifndef SGK_MI355X_H
#define SGK_MI355X_H

#include "surfacegrowth.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SGK_OK = 0,
    SGK_E_ARG = -1,        /* invalid argument */
    SGK_E_HIP = -2,        /* HIP runtime error */
    SGK_E_NOMEM = -3,      /* device allocation failed */
    SGK_E_OVERFLOW = -4,   /* a cell or neighbour list exceeded its capacity */
    SGK_E_UNSUPPORTED = -5,/* input outside what the engine implements */
    SGK_E_STATE = -6,      /* call made in the wrong order (e.g. no neighbour list yet) */
    SGK_E_DISPLACEMENT = -7/* an atom moved further than the neighbour-list skin allows */
};

/* Device error flags accumulated by the kernels (bitmask in sgk_record.errors). */
enum {
    SGK_FLAG_CELL_OVERFLOW = 1,   /* more than cell_capacity atoms in one cell */
    SGK_FLAG_NL_OVERFLOW = 2,     /* more than iNebrMax neighbours (reference: silent corruption) */
    SGK_FLAG_LJ_OVERFLOW = 4,     /* carbon: more metal neighbours than lj_capacity */
    SGK_FLAG_CLOSE_OVERFLOW = 8,  /* carbon: more close neighbours than close_capacity */
    SGK_FLAG_UNBINNED = 16,       /* an atom's cell index fell outside the grid (reference drops it) */
    SGK_FLAG_RDF_EDGE = 32,       /* a pair rounded into bin sizeHistRdf (reference writes past the end) */
    SGK_FLAG_DISPLACEMENT = 64    /* skin bound for the carbon close lists exceeded */
};

typedef struct sgk_ctx sgk_ctx;

typedef struct sgk_config {
    int precision;        /* SG_PREC_FP32: FP32 state and sums, as the reference.
                             SG_PREC_MIXED: FP64 positions/velocities and FP64 sums, FP32 pair math */
    int nl_capacity;      /* neighbours per atom; 0 -> SimParams.iNebrMax (reference 128) */
    int cell_capacity;    /* atoms per cell; 0 -> 64 (reference BLOCK_SIZE) */
    int lj_capacity;      /* metal neighbours kept per carbon atom; 0 -> 64 */
    int close_capacity;   /* close neighbours kept per carbon atom; 0 -> 16 */
    float close_margin;   /* close radius = rCutC + 2*(0.5*rNebrShell + margin); 0 -> 0.2 */
    int metal_lanes;      /* lanes per metal atom in the rho/force kernels (1..16); 0 -> auto */
    int block_size;       /* threads per workgroup of the per-atom kernels: 0 or 256 (the kernels, their
                             launch bounds and reductions are written for 4-wave workgroups; any other
                             value is rejected with SGK_E_UNSUPPORTED) */
    int keep_full_nl;     /* also store the neighbour list in the reference layout (tests) */
    int rebuild_every;    /* 0: adaptive rule dispHi > rNebrShell/2 only; k>0: also every k steps */
    int records_capacity; /* step records buffered on the device; 0 -> 1024 */
    int use_graphs;       /* capture the device loop in hipGraphs */
    int metal_thermostat_all; /* thermostat metal after stepEquil too (reference: only while
                                 stepEquil < step < stepEquil + stepCool) */
    int pair_fp64;        /* mixed form only: EAM and LJ pair arithmetic in FP64 (default) or FP32 */
    int xcd_map;          /* launch shape: bit 0 neighbour list, bit 1 density/near lists, bit 2 force
                             pass - give each of the 8 XCDs a contiguous 1/8 of the atoms (default 2;
                             bits 0 and 1 leave every result unchanged) */
    int bin_group;        /* binning: one atomicAdd per distinct cell of a wave instead of per atom */
    int nl_scheme;        /* neighbour list: 0 one thread per atom, 1 one wave per occupied cell,
                             2 LDS-staged per workgroup, 3 (default) lanes per metal atom, atoms
                             visited in cell order (sgk_build_nl_order); every scheme gives the same
                             lists */
    int rdf_complete;     /* RDF: 0 the reference's 27-cell stencil of the last rebuilt cells (complete
                             only to one cell width); 1 every metal-metal pair within rangeRdf at the
                             current positions (minimum image; rangeRdf must stay below half the box),
                             over a grid of cells rangeRdf/2 wide; 2-6 the same over cells rangeRdf/k
                             wide, cells beyond range culled (same histogram; on the 100,000-atom
                             cluster 2 is the fastest).  No counterpart in the reference. */
    int sync_rebuilds;    /* batched engine: when the skin rule or the cadence of any replica calls for a
                             list rebuild, every replica rebuilds at that step, so the rebuild launches
                             run on fewer steps.  A replica then equals the replica run alone with the
                             same rebuild steps.  0 (default): each replica decides alone.  No
                             counterpart in the reference. */
} sgk_config;

/* One record per step, written on the device by the reduction stage.  Sums are over atoms with
 * the reference's semantics (SurfaceGrowth.cu:1101-1891); see sgk_eval_props. */
typedef struct sgk_record {
    int step;             /* stepCount after the step */
    int rebuilt;          /* the neighbour list was rebuilt in this step */
    int nebrNow;          /* rebuild requested for the next step */
    int errors;           /* SGK_FLAG_* raised so far */
    double kinEnergy;     /* 0.5*vvSum/nMol */
    double potEnergy;     /* uSum/nMol */
    double totEnergy;
    double vSum[3];       /* sum v, metal x component times massMe (ComputeVSumK) */
    double vvSum;         /* sum m|v - vcm_x|^2 (ComputeVvSumK) */
    double vvMax;         /* max |v|^2 (ComputeVvMaxK) */
    double uSum;          /* sum E_pot (ComputePotEnergyK) */
    double comSum[3];     /* sum r over metal (ComputeCenterOfMassK) */
    double cmVelSum[3];   /* sum v over metal (ComputeCmVelK) */
    double frict;         /* sum friction x over metal (ComputeNetForceK) */
    double bboxMin[3];    /* min of metal r including the reference's zero padding */
    double bboxMax[3];
    double dispHi;        /* after this step */
    double shear;         /* current shear increment state (regime 2) */
    double shearCount;    /* atoms sheared in this step */
    int nMolDeposited;    /* metal atoms deposited so far (regime 1) */
    int pad;
    double virial;        /* pair virial 1/2 sum_i sum_k f_ik |d_ik|^2 of the step's EAM and
                             metal-carbon LJ forces (every LJ term the carbon side counts; not the
                             graphene law, which has no pair reaction); W = -3V dE/dV for those
                             terms.  No counterpart in the reference.  0 from sgk_eval_props. */
} sgk_record;

void sgk_default_config(sgk_config *cfg);

/* Create an engine for the parameters in *p (copied; the struct is the reference SimParams). */
int sgk_create(const SimParams *p, const sgk_config *cfg, sgk_ctx **out);
int sgk_destroy(sgk_ctx *ctx);
/* Message for the last failing call on ctx (ctx may be NULL for creation errors). */
const char *sgk_last_error(const sgk_ctx *ctx);
/* Number of HIP devices (0 when no runtime/GPU); does not initialise a context. */
int sgk_device_count(void);

/* State transfer (host reference layout <-> device).  In mixed precision the float inputs are
 * widened; the _f64 variants move full FP64 state (r: 4 doubles x,y,z,E per atom; v,a: 3). */
int sgk_upload(sgk_ctx *ctx, const sg_float4 *r, const sg_float3 *v, const sg_float3 *a);
int sgk_download(sgk_ctx *ctx, sg_float4 *r, sg_float3 *v, sg_float3 *a);
int sgk_upload_f64(sgk_ctx *ctx, const double *r, const double *v, const double *a);
int sgk_download_f64(sgk_ctx *ctx, double *r, double *v, double *a);
/* Step counter, rebuild flag and displacement accumulator of the run state. */
int sgk_set_run_state(sgk_ctx *ctx, int stepCount, int nebrNow, double dispHi);
/* Shear state of the run (regime 2): the shear force per metal atom, and the metal centre-of-mass x
 * and centre-of-mass velocity x of the last step, which the next step's shear reads (ApplyShearK
 * and the increment rule, SurfaceGrowth.cu:517-534, 2398-2404).  For restarts; a new context starts
 * at 0. */
int sgk_set_shear_state(sgk_ctx *ctx, double shear, double comX, double cmVelX);

/* --- stages ---------------------------------------------------------------------------------- */
/* Cell binning (BinAtomsIntoCellsK, SurfaceGrowth.cu:547-609): cell index per atom with the
 * reference FP32 formula; cells listed in ascending atom order. */
int sgk_bin(sgk_ctx *ctx);
/* cellOf[nMol] (-1 = outside the grid), cellCount[ncell], cellSlots[ncell*cell_capacity]. */
int sgk_get_cells(sgk_ctx *ctx, int *cellOf, int *cellCount, int *cellSlots);
/* Neighbour list (BuildNebrListK, SurfaceGrowth.cu:612-711) from the current cells. */
int sgk_build_nebr(sgk_ctx *ctx);
/* nl_scheme 3: rebuild now the order in which the list build visits atoms (by the cell of their
 * current position).  sgk_run rebuilds it when it is missing or 100 steps old; the order
 * changes which lane builds which list, never the lists.  No counterpart in the reference (its list
 * build has one thread per atom in index order).  A no-op for the other schemes. */
int sgk_build_nl_order(sgk_ctx *ctx);
/* Reference layout NN[nMol], NBL[iNebrMax*nMol] column-major, -1 filled; needs keep_full_nl. */
int sgk_get_nebr(sgk_ctx *ctx, int *NN, int *NBL);
/* Per-carbon lists the force kernel uses: metal neighbours and close neighbours, each entry
 * (atom index, position in the reference list).  Arrays sized nC*capacity, row per carbon. */
int sgk_get_carbon_lists(sgk_ctx *ctx, int *nLj, int *ljIdx, int *ljPos, int *nClose,
                         int *closeIdx, int *closePos);
/* Electron density and embedding terms (EamComputeRhoK, SurfaceGrowth.cu:716-773). */
int sgk_compute_rho(sgk_ctx *ctx);
/* Forces, per-atom energy, friction (ComputeForcesK, SurfaceGrowth.cu:776-1091) without
 * integration. */
int sgk_compute_forces(sgk_ctx *ctx);
/* rho[nMolMe]; acc[3*nMol]; energy[nMol] (the .w of r); fric[nMolMe].  NULL pointers skipped. */
int sgk_get_rho(sgk_ctx *ctx, double *rho);
int sgk_get_forces(sgk_ctx *ctx, double *acc, double *energy, double *fric);
/* LeapfrogStepK (SurfaceGrowth.cu:397-434) part 1 or 2; part 1 is followed by
 * ApplyBoundaryCondK (SurfaceGrowth.cu:436-453). */
int sgk_leapfrog(sgk_ctx *ctx, int part);
/* ApplyBerendsenThermostat (SurfaceGrowth.cu:455-485) with the given vvSum. */
int sgk_thermostat(sgk_ctx *ctx, int stepCount, double vvSum);
/* Reductions of the current state with the reference semantics (SurfaceGrowth.cu:1101-1891);
 * cmVelX for the regime-2 kinetic energy is this state's own metal centre-of-mass velocity. */
int sgk_eval_props(sgk_ctx *ctx, sgk_record *rec);
/* RDF histogram (EvalRdfK, SurfaceGrowth.cu:1895-1999) over the cells of the last binning. */
int sgk_rdf_clear(sgk_ctx *ctx);
int sgk_rdf_accumulate(sgk_ctx *ctx);
int sgk_rdf_get(sgk_ctx *ctx, unsigned int *hist);
/* Centre-of-mass diffusion accumulated on the device, in FP64 (EvalDiffusion, AccumDiffusion,
 * PrintDiffusion: SurfaceGrowth.cu:2865-2952).  Once enabled, every step that the reference samples
 * (regime 2, step > stepEquil, (step - stepEquil) % stepDiffuse == 0) of sgk_run adds one sample of
 * the metal centre of mass to nBuffDiffuse staggered buffers of nValDiffuse values; every
 * limitDiffuseAv completed buffers give one table of D_j (j = 0..nValDiffuse-1, reduced units, the
 * reference's normalisation).  Sizes come from the SimParams diffusion fields; the last
 * `maxTables` tables are kept.  Enabling again restarts the buffers. */
int sgk_diffusion_enable(sgk_ctx *ctx, int maxTables);
/* Completed tables so far in *nTables; the most recent min(nTables, maxTables) are copied to
 * tables[k * nValDiffuse + j], oldest first (tables may be NULL). */
int sgk_diffusion_get(sgk_ctx *ctx, double *tables, int maxTables, int *nTables);

/* --- device-resident loop --------------------------------------------------------------------- */
/* Advance nsteps steps in the order of DoComputationsW (SurfaceGrowth.cu:2141-2489) without host
 * synchronisation between steps: leapfrog 1 + boundary, rebuild when flagged on the device,
 * density, forces fused with leapfrog 2 and the reductions, thermostat on its steps, shear,
 * RDF accumulation every rdfEvery steps (0: none).  records[nsteps] receives one sgk_record per
 * step (may be NULL).  nsteps must not exceed records_capacity. */
int sgk_run(sgk_ctx *ctx, int nsteps, int rdfEvery, sgk_record *records);
/* Capture and instantiate, without running anything, the step graph that the next
 * sgk_run(ctx, nsteps, rdfEvery, ...) from the current step will launch, so that call does not pay
 * for the capture.  A no-op without graphs.  No counterpart in the reference. */
int sgk_prepare_run(sgk_ctx *ctx, int nsteps, int rdfEvery);
/* Milliseconds of device time spent in the last sgk_run (HIP events on the engine stream). */
double sgk_last_run_ms(sgk_ctx *ctx);

/* --- introspection for benchmarks ------------------------------------------------------------ */
typedef struct sgk_kernel_stat {
    char name[48];
    int launches;
    double ms;            /* accumulated device time when profiling is on */
} sgk_kernel_stat;
/* Turn per-kernel event timing on/off (adds a sync per kernel; for measurement only). */
int sgk_set_profiling(sgk_ctx *ctx, int on);
/* Copy up to max kernel stats; returns the number of kernels. */
int sgk_kernel_stats(sgk_ctx *ctx, sgk_kernel_stat *stats, int max);
/* Bytes of device memory owned by ctx. */
long long sgk_device_bytes(sgk_ctx *ctx);
/* sizeof of sgk_config, sgk_record, sgk_kernel_stat (for bindings); returns the number written. */
int sgk_struct_sizes(int *out, int max);
/* Metal atoms already deposited (regime 1; the reference's hparams->nMolDeposited).  Creation
 * takes it from SimParams. */
int sgk_set_deposited(sgk_ctx *ctx, int nMolDeposited);

/* Lanes per metal atom in use (chosen at creation, or by sgk_autotune). */
int sgk_metal_lanes(const sgk_ctx *ctx);

/* Autotuner (no reference counterpart: the reference has fixed launch shapes).  Times the density
 * and force passes on the current state for every candidate launch shape - lanes per metal atom
 * 1/2/4/8/16 and the XCD map bits of the density and force passes (4 combinations, 20 shapes) -
 * `reps` launches each, and keeps the fastest.  The passes only rewrite quantities they recompute identically (acceleration, per-atom
 * energy, friction), so the state is unchanged.  Needs a valid neighbour list.  Results differ
 * between shapes only by the FP summation order of the metal neighbour loop. */
typedef struct sgk_tune_entry {
    int metal_lanes;
    int xcd_map;
    double us;            /* mean time of one density + force evaluation */
} sgk_tune_entry;
int sgk_autotune(sgk_ctx *ctx, int reps, sgk_tune_entry *entries, int max_entries, int *n_entries);

/* ---- adjoint (no reference counterpart: the reference integrates forward only) -----------------
 * Gradients of two trajectory observables with respect to four force-field parameters, by reverse-
 * mode differentiation of the discrete step - hand-written FP64 vector-Jacobian kernels of the
 * force pass, the leapfrog and the Berendsen thermostat - with step checkpointing and recomputation.
 * Mixed-precision contexts; regimes 0 and 2 (no deposition; no shear motion inside the window).
 *   window       sgk_adjoint_forward computes a(r0) from a neighbour list built at the current
 *                positions, then runs nsteps steps of the device loop opening with a rebuild, with a
 *                checkpoint (state, lists, run state) every checkpoint_every steps.
 *   observables  obs[SGK_OBS_ADHESION]: mean over the steps of the Ni-C adhesion energy, the
 *                metal-side LJ pair energies within the LJ cutoff, each metal-carbon pair once;
 *                obs[SGK_OBS_HEIGHT]: metal centre-of-mass height above z_0 after the last step.
 *   parameters   scale factors (at 1) of LJ epsilon, LJ sigma, the embedding function F (all three
 *                branches) and the electron density f.
 * Cutoff radii, the F branch points and every discrete decision (lists, bonds, rebuild steps) are
 * held fixed: the gradient is that of the piecewise-smooth step map along the trajectory, in reduced
 * units.  records (may be NULL) receives the window's nsteps step records. */
enum { SGK_PAR_EPSILON = 0, SGK_PAR_SIGMA = 1, SGK_PAR_EMBEDDING = 2, SGK_PAR_DENSITY = 3, SGK_NPAR = 4 };
enum { SGK_OBS_ADHESION = 0, SGK_OBS_HEIGHT = 1, SGK_NOBS = 2 };
int sgk_adjoint_forward(sgk_ctx *ctx, int nsteps, int checkpoint_every, double *obs, sgk_record *records);
/* Reverse sweep for weights w[SGK_NOBS]: grad[SGK_NPAR] = d(w . obs)/d scale; dr0, dv0 (may be NULL)
 * receive d(w . obs)/d r0 and d/d v0, x y z per atom (lists held fixed).  Recomputes each segment
 * from its checkpoint and leaves the context in the state the forward ended in; may be repeated
 * with other weights. */
int sgk_adjoint_backward(sgk_ctx *ctx, const double *w, double *grad, double *dr0, double *dv0);
typedef struct sgk_adjoint_info {
    int nsteps, checkpoint_every, segments;
    int list_copies;      /* neighbour-list copies held for rebuilds inside a segment */
    int mismatches;       /* atoms whose recomputed state differed bitwise from the stored one */
    int pad;
    long long bytes;      /* device memory of the adjoint */
    double forward_ms, backward_ms;
    double obs[2];
} sgk_adjoint_info;
int sgk_adjoint_get_info(sgk_ctx *ctx, sgk_adjoint_info *info);
/* Vector-Jacobian product of the force pass at the current positions and list: for the cotangent
 * abar[3*nMol] of the accelerations and the weight w_adh of the adhesion energy,
 * rbar[3*nMol] = d/dr and pbar[SGK_NPAR] = d/d scale of (abar . a(r) + w_adh E_adh(r)). */
int sgk_force_vjp(sgk_ctx *ctx, const double *abar, double w_adh, double *rbar, double *pbar);
/* Adhesion energy at the current positions and list (reduced units). */
int sgk_adhesion_energy(sgk_ctx *ctx, double *e);

/* ---- batched replicas (no reference counterpart: the reference runs one system per process) ----
 * Many replicas of uneven size (cluster size, sheet size, temperature) advance in lockstep with one
 * launch per stage for all of them, steps captured into hipGraphs.  Replicas share the potentials,
 * time step, neighbour-list shell and capacity, thermostat schedule and RDF bins (checked at
 * creation); geometry, atom counts and temperature are free.  Each replica keeps its own cells,
 * lists, step state, reductions and records, and its results equal those of the same replica run
 * alone in an sgk_ctx with xcd_map = 0 and the same lanes per metal atom.  The list build is the
 * cell-ordered column walk for nl_scheme 3 (the default; one counting sort orders every replica),
 * the LDS-staged one for nl_scheme 2 and the per-atom one otherwise (the lists are the same). */
typedef struct sgk_batch sgk_batch;
int sgk_batch_create(int nrep, const SimParams *params, const sgk_config *cfg, sgk_batch **out);
int sgk_batch_destroy(sgk_batch *b);
const char *sgk_batch_last_error(const sgk_batch *b);
/* State of one replica, reference layouts: r[4*n] (x,y,z,E), v[3*n], a[3*n]; any may be NULL. */
int sgk_batch_upload_f64(sgk_batch *b, int rep, const double *r, const double *v, const double *a);
int sgk_batch_download_f64(sgk_batch *b, int rep, double *r, double *v, double *a);
/* Step counter (replica 0's is the batch's), rebuild flag and displacement bound of one replica. */
int sgk_batch_set_run_state(sgk_batch *b, int rep, int stepCount, int nebrNow, double dispHi);
/* nsteps steps of every replica; records (may be NULL) receives nrep*nsteps records, replica-major. */
int sgk_batch_run(sgk_batch *b, int nsteps, int rdfEvery, sgk_record *records);
double sgk_batch_last_run_ms(const sgk_batch *b);
int sgk_batch_rdf_get(sgk_batch *b, int rep, unsigned int *hist);
int sgk_batch_rdf_clear(sgk_batch *b);
long long sgk_batch_device_bytes(const sgk_batch *b);
/* Lanes per metal atom: the common count of every replica (0 when they differ), or of one replica.
 * At creation every replica gets sgk_config.metal_lanes when it is 1, 2, 4, 8 or 16, else the count
 * the single-system engine would choose for its size. */
int sgk_batch_metal_lanes(const sgk_batch *b);
int sgk_batch_replica_lanes(const sgk_batch *b, int rep);
/* Set the lanes per metal atom of every replica (lanes[nrep], each 1, 2, 4, 8 or 16).  A replica's
 * results equal those of the replica run alone in an sgk_ctx with the same lanes. */
int sgk_batch_set_lanes(sgk_batch *b, const int *lanes);
/* Autotuner by replica size: replicas are bucketed by metal atoms ([0,128), [128,512), [512,2048),
 * [2048,8192), [8192,...)); for each bucket every lane count is timed on the live state (density and
 * force passes of the whole batch, `reps` times) and the fastest kept for the bucket's replicas.
 * Accelerations and energies are left as the chosen lanes compute them.  Needs one run first.
 * One entry per (bucket, lane count) tried.  No counterpart in the reference. */
typedef struct sgk_batch_tune_entry {
    int metal_lo, metal_hi;  /* bucket: replicas with metal_lo <= nMolMe < metal_hi */
    int replicas, atoms;     /* replicas and atoms in the bucket */
    int metal_lanes;         /* lanes per metal atom tried */
    double us;               /* density + force passes of the batch, microseconds per launch pair */
} sgk_batch_tune_entry;
int sgk_batch_autotune(sgk_batch *b, int reps, sgk_batch_tune_entry *entries, int maxEntries, int *nEntries);
int sgk_batch_atoms(const sgk_batch *b);

/* ---- force evaluation without allocation ---------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  It enqueues, on `stream`
 * (a hipStream_t; NULL = the null stream), what sgk_create + sgk_upload + sgk_bin + sgk_build_nebr +
 * sgk_compute_forces do on a fresh context, with the same kernels in the same order, so its results
 * equal theirs bit for bit.  It calls no hipMalloc, hipFree, memset, copy or synchronisation.
 *   pos   device, nMol x {x,y,z,w}: double4 (precision SG_PREC_MIXED) or float4 (SG_PREC_FP32); w
 *         receives the per-atom energy, as the reference's ComputeForcesK writes dr[i].w
 *   vel   device, 3*nMol in the storage precision, SoA (x block, y block, z block), or NULL (zero);
 *         read for the block reductions only
 *   acc   device, 3*nMol in the storage precision, SoA: the accelerations
 *   fric  device, nMolMe floats: metal-carbon friction x, or NULL
 *   flags device, one int: the SGK_FLAG_* bits of the evaluation (capacity overflows), or NULL
 * Intermediate buffers (cells, lists, densities, partials, run state) are carved from one
 * static device array of SGK_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; larger
 * systems are rejected with SGK_E_UNSUPPORTED (sgk_noalloc_scratch_bytes gives a system's need).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU race.  keep_full_nl is
 * not supported.  Errors: the return code, with the reason from sgk_noalloc_last_error(). */
#define SGK_NOALLOC_SCRATCH_BYTES (64ll << 20)
int sgk_forces_noalloc(const SimParams *sp, const sgk_config *cfg, void *pos, const void *vel, void *acc,
                       float *fric, int *flags, void *stream);
/* Scratch bytes sgk_forces_noalloc needs for this system (with fric and vel NULL); -1 if invalid. */
long long sgk_noalloc_scratch_bytes(const SimParams *sp, const sgk_config *cfg);
/* Reason for the last failed sgk_forces_noalloc on this thread. */
const char *sgk_noalloc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SGK_MI355X_H */
