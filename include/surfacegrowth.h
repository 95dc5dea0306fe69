/*
 * surfacegrowth.h - drop-in C-ABI of the SurfaceGrowth per-step MD engine for MI355X (gfx950).
 *
 * This header replaces the boundary the reference host driver binds to:
 *   - struct SimParams            SurfaceGrowthConsole/SurfaceGrowth.h:69-186
 *   - struct TBuf, MATERIAL       SurfaceGrowthConsole/SurfaceGrowth.h:189-215
 *   - the four extern "C" wrappers SurfaceGrowthConsole/SurfaceGrowth.h:217-235
 *     (implemented in the reference at SurfaceGrowthConsole/SurfaceGrowth.cu:2013-2650)
 *
 * The field order, field types and therefore the byte layout of SimParams are those of the
 * reference, so a SimParams written by the reference host (or its backup files, whose header is
 * the raw struct, SurfaceGrowth.cu:2475) is read unchanged.  CUDA's float3/float4/uint3/int3 are
 * replaced by plain PODs of identical size and alignment (12/16/12/12 bytes).
 *
 * No GPU or HIP type appears here: the library behind it is built with hipcc for gfx950 and owns
 * all device memory.
 */
#ifndef SURFACEGROWTH_MI355X_H
#define SURFACEGROWTH_MI355X_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar type of the reference (ComputeDefs.h:22). */
typedef float real;

/* Vector PODs with the layout of CUDA's vector_types.h types. */
typedef struct sg_float3 { float x, y, z; } sg_float3;
typedef struct __attribute__((aligned(16))) sg_float4 { float x, y, z, w; } sg_float4;
typedef struct sg_uint3 { unsigned int x, y, z; } sg_uint3;
typedef struct sg_int3 { int x, y, z; } sg_int3;

/* Property accumulator: value, running sum, running sum of squares (ComputeDefs.h:53-55). */
typedef struct Prop { real val, sum, sum2; } Prop;

#ifndef MAX_PATH
#define MAX_PATH 264 /* SurfaceGrowth.h:53-55 */
#endif
#define SG_BLOCK_SIZE 64 /* BLOCK_SIZE, SurfaceGrowth.h:35: reference atoms-per-cell capacity */

/* Simulation parameters and mutable run state; same fields, order and types as
 * SurfaceGrowth.h:69-186 (TCHAR = char, BOOL = int in the reference's non-Unicode build). */
typedef struct SimParams {
    sg_float3 region;      /* simulation region (reduced units) */
    sg_float3 vSum;        /* total impulse */
    sg_uint3 initUcell;    /* unit cells of the structure */
    real deltaT, rCutC, rrCutC, temperature, velMag;
    int moreCycles, stepAvg, stepEquil, stepCount, stepPdb, stepLimit;
    int bPdb;
    int bResult;
    int nMol;              /* all atoms: metal [0,nMolMe), carbon [nMolMe,nMol) */
    int nMolMe;            /* metal atoms */
    unsigned int gridSize; /* number of cells */
    unsigned int blockSize;/* reference atoms-per-cell capacity (64) */
    sg_int3 cells;         /* cell grid for the neighbour list */
    sg_float3 invWidth;    /* cells / region */
    real rNebrShell;       /* neighbour-list skin */
    real rrNebr;           /* (rCutEam + rNebrShell)^2 */
    int iNebrMax;          /* neighbour-list capacity per atom (128) */
    int maxMol;            /* cells * blockSize */
    int nebrNow;           /* rebuild the neighbour list at the next step */
    real dispHi;           /* accumulated max displacement since the last rebuild */
    Prop kinEnergy, totEnergy, potEnergy, oneStep;
    /* EAM (Zhou et al. 2001) in reduced units */
    real a, re, fe, rhoe, alpha, beta, A, B, kappa, lambda, Fn[4], F[4], eta, Fe, massMe, density;
    real rCutEam, rrCutEam, rei, rhoei;
    char szNameMe[3];
    /* units */
    real temperatureU, kB, enU, lengthU, forceU, massU;
    /* graphene (Sasaki et al.) */
    real mu_r, theta_0, mu_theta, mu_p;
    real z_0;
    /* Berendsen thermostat */
    int stepThermostat;
    real gammaBerendsen;
    int stepCool;
    int iRegime;           /* 0 bulk, 1 surface growth, 2 shear/diffusion */
    /* metal-carbon Lennard-Jones */
    real sigmaLJ, epsLJ, rCutLJ, rrCutLJ;
    /* deposition */
    real velMagDepos;
    int randSeedP;
    int stepDeposit, nMolDeposited, nMolToDeposit;
    /* tribology */
    Prop centerOfMass, cmVel, frictForce;
    real deltaF;
    sg_float3 particleSize;
    int cellShiftZ;
    real shear;
    real totalShear;
    int initUcellMeX;
    int initUcellMeY;
    /* radial distribution function */
    real rangeRdf;
    int limitRdf;
    int sizeHistRdf;
    int stepRdf;
    int countRdf;
    real intervalRdf;
    int bRdf;
    char szRdfPath[MAX_PATH];
    /* backup */
    char szBckup0[20];
    char szBckup1[20];
    int bBckup;
    int bStartBckup;
    int stepBckup;
    real totalTime;
    int iMaterial;
    /* diffusion of the cluster centre of mass */
    int nValDiffuse;
    int nBuffDiffuse;
    int stepDiffuse;
    int limitDiffuseAv;
    int countDiffuseAv;
    char szDiffusePath[MAX_PATH];
} SimParams;

/* Diffusion buffer (SurfaceGrowth.h:189-196). */
typedef struct TBuf {
    sg_float3 orgR, rTrue;
    real *rrDiffuse;
    int count;
} TBuf;

/* EAM material record (SurfaceGrowth.h:199-215); the table itself is SurfaceGrowthProto.h:31-80. */
typedef struct MATERIAL {
    const char *szName;
    real re, fe, rhoe, alpha, beta, A, B, kappa, lambda, Fn[4], F[4], eta, Fe, massMe, density;
} MATERIAL;

/* ---- The reference wrapper API (SurfaceGrowth.h:217-235) ------------------------------------ */

/* Selects device 0 and prints the device count (SurfaceGrowth.cu:2614-2637).
 * Prints and exits with EXIT_FAILURE when the HIP runtime reports an error. */
void CudaInitW(int argc, char **argv);

/* Copies the parameters to the engine (SurfaceGrowth.cu:2639-2650).  Prints and exits with
 * EXIT_FAILURE on error. */
void SetParametersW(SimParams *hostParams);

/* Fills hr[0..nMol) with the initial coordinates for hparams->iRegime (SurfaceGrowth.cu:2532-2609).
 * `dr` is ignored, as in the reference.  Returns 0 on success or a static error string. */
const char *InitCoordsW(sg_float4 *dr, sg_float4 *hr, SimParams *hparams);

/* Runs the whole simulation (SurfaceGrowth.cu:2013-2529): allocates device memory, steps until
 * stepLimit, writes summary rows to fResults, pdb/rdf/Diffuse files, frees everything.  Returns 0
 * on success; on error copies a message into szPdbPath (MAX_PATH bytes) and returns it. */
char *DoComputationsW(sg_float4 *hr, sg_float3 *hv, sg_float3 *ha, SimParams *hparams,
                      FILE *fResults, char *szPdbPath);

/* ---- Device queries the reference driver makes itself (SurfaceGrowth.cpp:108-122, 315-318) ---- */

/* Number of visible gfx950 devices; replaces the driver's cudaGetDeviceCount (SurfaceGrowth.cpp:109).
 * Returns 0 when there is none or the runtime fails. */
int sgDeviceCount(void);
/* Largest cell grid (cells.x * cells.y * cells.z) the engine accepts with the reference's 64-atom
 * cells; replaces the driver's g_hDeviceProp.maxGridSize[0] test (SurfaceGrowth.cpp:317), which
 * limited the reference to one CUDA block per cell.  Here the limit is cells x 64 < 2^31. */
unsigned int sgMaxCellGrid(void);

/* ---- Host-side helpers of the reference driver (SurfaceGrowth.cpp), exported so a host written
 *      against this library can reuse them; behaviour follows the cited lines. ------------------ */

/* Options that the reference hard-codes or takes from its globals, plus extensions.  The
 * reference's values are the defaults returned by sgDefaultOptions(). */
typedef struct SgOptions {
    /* SurfaceGrowth.cpp:49-90 globals */
    real rNebrShell;          /* 0.4 */
    int iNebrMax;             /* 128 */
    int blockSize;            /* 64 */
    int stepThermostat;       /* 25 */
    real gammaBerendsen;      /* 0.005 */
    int limitRdf, sizeHistRdf;/* 100, 200 */
    int nValDiffuse, nBuffDiffuse, stepDiffuse, limitDiffuseAv; /* 25000, 10, 100, 20 */
    int stepBckup;            /* 10000 */
    /* extensions (sugr_in.txt rows >= 23, "key=value") */
    int seed;                 /* 0: time-based seed as the reference; else fixed LCG seed */
    int geometry;             /* SG_GEOM_* ; SG_GEOM_REFERENCE = regime's own generator */
    real gapAngstrom;         /* cluster lowest atom height above the sheet (sphere/icosahedron) */
    int velocityRule;         /* SG_VEL_REFERENCE (fixed |v|, random direction) or SG_VEL_MAXWELL */
    int precision;            /* SG_PREC_FP32 (plain) or SG_PREC_MIXED */
    int rebuildEvery;         /* 0: adaptive rule (SurfaceGrowth.cu:2390-2391); k>0: every k steps */
    real rangeRdfAngstrom;    /* 0: reference 1.5*rCutEam */
    int stepsPerSync;         /* max steps between host synchronisations (records chunk) */
    int useGraphs;            /* capture step chunks in hipGraphs */
    int metalThermostatAll;   /* 0 reference (carbon only, metal in cooling window) */
    real tauPs;               /* >0: gammaBerendsen = stepThermostat*deltaT_ps/tau */
} SgOptions;

enum { SG_GEOM_REFERENCE = 0, SG_GEOM_SPHERE = 1, SG_GEOM_ICOSAHEDRON = 2 };
enum { SG_VEL_REFERENCE = 0, SG_VEL_MAXWELL = 1 };
enum { SG_PREC_FP32 = 0, SG_PREC_MIXED = 1 };

/* Reference defaults (SurfaceGrowth.cpp:49-90) and extension defaults. */
void sgDefaultOptions(SgOptions *opt);
/* Options used by SetParams/InitCoordsW/InitVels/DoComputationsW of this library. */
void sgSetOptions(const SgOptions *opt);
void sgGetOptions(SgOptions *opt);

/* Input values read from sugr_in.txt rows 1-22 (SurfaceGrowth.cpp:689-778). */
typedef struct SgInput {
    int iRegime, iMaterial;
    int ux, uy, uz;
    real temperature;
    int nMolMe;
    real epsilon, sigma;
    int stepLimit;
    real deltaT;
    int stepAvg, stepEquil;
    real deltaF;
    int stepCool;
    real deposEnergy;
    int nMolToDeposit, stepDeposit;
    int stepPdb, stepRdf;
    int bBckup, bStartBckup;
} SgInput;

/* Reference defaults of the globals the input file overrides (SurfaceGrowth.cpp:39-90). */
void sgDefaultInput(SgInput *in);
/* ReadInputFile (SurfaceGrowth.cpp:633-785): positional rows 1-22; rows >= 23 of the form
 * "key=value" set extension options in *opt (may be NULL).  Returns 1 on success, 0 if the file
 * cannot be opened. */
int sgReadInputFile(const char *path, SgInput *in, SgOptions *opt);
/* SetParams + EamInit + GrapheneInit (SurfaceGrowth.cpp:189-396, 480-545).  maxGridSize is the
 * device limit checked at SurfaceGrowth.cpp:317 (<=0: no limit).  Returns 1, or 0 when the cell
 * grid exceeds maxGridSize.  Does not call SetParametersW. */
int sgSetParams(const SgInput *in, SimParams *p, long long maxGridSize);
/* Rapaport LCG (SurfaceGrowth.cpp:548-593). */
void sgInitRand(int seed, SimParams *p);
real sgRandR(SimParams *p);
void sgVRand(sg_float3 *v, SimParams *p);
/* InitVels (SurfaceGrowth.cpp:596-621), or Maxwell-Boltzmann when opt velocityRule says so. */
void sgInitVels(SimParams *p, sg_float3 *hv);
/* Summary-file name and header (SurfaceGrowth.cpp:421-469); returns the opened FILE or NULL. */
FILE *sgOpenResultFile(const SimParams *p, real deposEnergy, char *nameOut);
/* Backup file writer/reader using the reference layout (SurfaceGrowth.cu:2475-2478). */
int sgWriteBackup(const char *path, const SimParams *p, const sg_float4 *hr, const sg_float3 *hv,
                  const sg_float3 *ha);
int sgReadBackup(const char *path, SimParams *p, sg_float4 *hr, sg_float3 *hv, sg_float3 *ha,
                 int capacity);

/* Centre-of-mass diffusion buffers that DoComputationsW keeps (EvalDiffusion, AccumDiffusion,
 * PrintDiffusion: SurfaceGrowth.cu:2865-2952), callable directly.  Each call takes one sample of
 * the metal centre of mass; when an average of limitDiffuseAv buffers completes, a "diffusion"
 * table is appended to `file` (may be NULL).  Returns the number of tables completed so far. */
typedef struct sgDiffusion sgDiffusion;
sgDiffusion *sgDiffusionCreate(const SimParams *p);
int sgDiffusionEval(sgDiffusion *d, SimParams *p, const sg_float3 *com, FILE *file);
void sgDiffusionDestroy(sgDiffusion *d);

/* Library version string. */
const char *sgVersion(void);
/* sizeof of the ABI structs in the order SimParams, SgOptions, SgInput, TBuf, MATERIAL (bindings
 * check their mirrors against these); returns the number written. */
int sgStructSizes(int *out, int max);

#ifdef __cplusplus
}

/* AccumProps (SurfaceGrowth.cu:2659-2686): zero (icode 0), accumulate (1) or average (2) the Prop
 * fields of *hparams.  The reference defines it in its .cu with C++ linkage and its host driver
 * calls it (SurfaceGrowth.cpp:417; prototype SurfaceGrowthProto.h:25), so it keeps that linkage. */
void AccumProps(int icode, SimParams *hparams);
#endif

#endif /* SURFACEGROWTH_MI355X_H */
