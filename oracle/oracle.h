/*
 * oracle.h - CPU restatement of the SurfaceGrowth per-step MD algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness reference for the HIP engine and the CPU baseline
 * leg of bench.py; nothing in the product links, loads or calls it.
 *
 * Every function restates a reference routine, cited by file:line (paths relative to the
 * reference repository; .cu = SurfaceGrowthConsole/SurfaceGrowth.cu).  The arithmetic is
 * restated in a precision-generic form and compiled twice:
 *   *_f32 : float, libm float functions, no FMA contraction - the FP32 restatement
 *   *_f64 : double - the FP64 oracle of the same algorithm (same float parameters, promoted)
 * The integer parts (cell index, neighbour list, RDF bins) of *_f32 are bit-exact definitions.
 *
 * tangent.cpp compiles the same source a third time with forward-mode dual numbers (dual.hpp): the
 * FP64 values plus directional derivatives, the reference gradient of the adjoint.  The FP64 build
 * also keeps a decision tape (orc_freeze_f64): a run that replays it holds the discrete decisions of
 * a recorded run, so central differences of perturbed runs are those of the map the adjoint and the
 * tangent differentiate (pinned against the tangent at a step where free differences fail).
 *
 * Pinning (tests/test_oracle_*.py, tests/test_graphene_pinning.py): EAM by fcc shell sums (closed
 * form), two-atom LJ, central differences of the conservative EAM+LJ energy, the graphene law by the
 * zero-force ideal sheet, first-order in-plane and out-of-plane responses, the uniformly dilated
 * sheet and, under finite deformation, each carbon's force against central differences of its own
 * local energy (with a mutation test that plants mistakes in copies of this source), brute-force
 * O(N^2) neighbour lists and RDF, the Berendsen factor, leapfrog on a harmonic dimer, rigid
 * translation for the diffusion buffers, and the published bulk-Au known answer.
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters the algorithm reads, with the float values of the reference SimParams fields. */
typedef struct orc_params {
    int nMol, nMolMe, iRegime;
    float region[3];
    int cells[3];
    float invWidth[3];
    float rrNebr;
    int iNebrMax;
    int cellCap;             /* atoms per cell capacity (reference 64) */
    /* EAM */
    float re, fe, rhoe, alpha, beta, A, B, kappa, lambda, Fn[4], F[4], eta, Fe, massMe;
    float rCutEam, rrCutEam, rei, rhoei;
    /* graphene */
    float mu_r, theta_0, mu_theta, mu_p, z_0, rCutC, rrCutC;
    /* LJ */
    float sigmaLJ, epsLJ, rrCutLJ;
    /* integration, thermostat */
    float deltaT, temperature, kB, gammaBerendsen, rNebrShell;
    int stepThermostat, stepEquil, stepCool;
    /* RDF */
    float rangeRdf, intervalRdf;
    int sizeHistRdf;
    /* shear */
    float deltaF;
    /* deposition (regime 1) */
    float velMagDepos;
    int stepDeposit, nMolToDeposit;
    /* scale factors of the differentiated parameters (1 = the SimParams values): LJ epsilon, LJ
     * sigma, embedding function F, electron density f */
    double parScale[4];
} orc_params;

/* Reductions of one state with the reference semantics (.cu:1101-1891, 2307-2395). */
typedef struct orc_props {
    double vSum[3], vvSum, vvMax, uSum;
    double comSum[3], cmVelSum[3], frict;
    double bboxMin[3], bboxMax[3];
    double kinEnergy, potEnergy, totEnergy;
} orc_props;

/* Per-step record of orc_run (same meaning as sgk_record). */
typedef struct orc_record {
    int step, rebuilt, nebrNow, nMolDeposited;
    double kinEnergy, potEnergy, totEnergy;
    double vSum[3], vvSum, vvMax, uSum;
    double comSum[3], cmVelSum[3], frict;
    double bboxMin[3], bboxMax[3];
    double dispHi;
} orc_record;

/* Cell lists: count[ncell], slots[ncell*cellCap] ascending atom index, cellOf[nMol] (-1 if the
 * index falls outside [0,ncell)).  Returns 0, or 1 if a cell exceeded cellCap. */
int orc_bin_f32(const orc_params *p, const float *pos4, int *cellOf, int *count, int *slots);
int orc_bin_f64(const orc_params *p, const double *pos4, int *cellOf, int *count, int *slots);
/* Neighbour list from the cells, reference layout NBL[j*nMol+i] (-1 filled), NN[i].
 * Returns 0, or 2 if an atom exceeded iNebrMax neighbours. */
int orc_nebr_f32(const orc_params *p, const float *pos4, const int *count, const int *slots,
                 int *NN, int *NBL);
int orc_nebr_f64(const orc_params *p, const double *pos4, const int *count, const int *slots,
                 int *NN, int *NBL);
/* Electron density (.cu:716-773). */
void orc_rho_f32(const orc_params *p, const float *pos4, const int *NN, const int *NBL, float *rho);
void orc_rho_f64(const orc_params *p, const double *pos4, const int *NN, const int *NBL, double *rho);
/* Forces (.cu:776-1091): acc[3*nMol], pos4[.w] = per-atom energy, fric[nMolMe] (may be NULL). */
void orc_forces_f32(const orc_params *p, float *pos4, const int *NN, const int *NBL,
                    const float *rho, float *acc, float *fric);
void orc_forces_f64(const orc_params *p, double *pos4, const int *NN, const int *NBL,
                    const double *rho, double *acc, double *fric);
/* orc_forces plus each atom's pair virial 1/2 sum_k f_ik |d_ik|^2 of its EAM and LJ terms. */
void orc_forces_virial_f32(const orc_params *p, float *pos4, const int *NN, const int *NBL, const float *rho,
                           float *acc, float *fric, float *vir);
void orc_forces_virial_f64(const orc_params *p, double *pos4, const int *NN, const int *NBL, const double *rho,
                           double *acc, double *fric, double *vir);
/* Embedding function F(rho) and F'(rho) (.cu:105-118, 197-210). */
float orc_embed_f32(const orc_params *p, float rho, float *dF);
double orc_embed_f64(const orc_params *p, double rho, double *dF);
/* Pair functions phi, phi', f, f' at distance r (.cu:46-69, 124-158). */
void orc_pair_f32(const orc_params *p, float r, float *out4);
void orc_pair_f64(const orc_params *p, double r, double *out4);
/* LeapfrogStepK part 1/2 (.cu:397-434); part 1 followed by ApplyBoundaryCondK (.cu:436-453). */
void orc_leapfrog_f32(const orc_params *p, int part, float *pos4, float *vel, const float *acc);
void orc_leapfrog_f64(const orc_params *p, int part, double *pos4, double *vel, const double *acc);
/* ApplyBerendsenThermostat (.cu:455-485). */
void orc_thermostat_f32(const orc_params *p, float *vel, double vvSum, int stepCount);
void orc_thermostat_f64(const orc_params *p, double *vel, double vvSum, int stepCount);
/* Reductions (.cu:1101-1891 and the host arithmetic at .cu:2307-2395). */
void orc_props_f32(const orc_params *p, const float *pos4, const float *vel, const float *fric,
                   orc_props *out);
void orc_props_f64(const orc_params *p, const double *pos4, const double *vel, const double *fric,
                   orc_props *out);
/* RDF histogram accumulation (.cu:1895-1999) over given cells; returns the number of pairs whose
 * bin index equalled sizeHistRdf (the reference writes those past the end; here dropped). */
int orc_rdf_f32(const orc_params *p, const float *pos4, const int *cellOf, const int *count,
                const int *slots, unsigned int *hist);
int orc_rdf_f64(const orc_params *p, const double *pos4, const int *cellOf, const int *count,
                const int *slots, unsigned int *hist);

/* Diffusion of the cluster centre of mass (.cu:2865-2952). */
typedef struct orc_diffusion orc_diffusion;
orc_diffusion *orc_diffusion_create(int nVal, int nBuff, int stepDiffuse, int limitAv, double deltaT,
                                    const float *region);
void orc_diffusion_destroy(orc_diffusion *d);
/* Feed one centre-of-mass sample; returns 1 when a D(t) table was completed into out[nVal]. */
int orc_diffusion_sample_f32(orc_diffusion *d, const float *com, float *out);
int orc_diffusion_sample_f64(orc_diffusion *d, const double *com, double *out);

/* Whole loop of DoComputationsW (.cu:2141-2489) without output files: nsteps steps from
 * stepCount0 with the given rebuild state.  rebuildEvery>0 adds rebuilds to the adaptive rule.
 * nMolDeposited (regime 1; may be NULL) is the deposition counter, updated in place.
 * records[nsteps] may be NULL.  Returns 0 or an error code (1 cell overflow, 2 NL overflow). */
int orc_run_f32(const orc_params *p, float *pos4, float *vel, float *acc, int nsteps,
                int stepCount0, int *nebrNow, double *dispHi, int rebuildEvery, orc_record *records,
                int *nMolDeposited);
int orc_run_f64(const orc_params *p, double *pos4, double *vel, double *acc, int nsteps,
                int stepCount0, int *nebrNow, double *dispHi, int rebuildEvery, orc_record *records,
                int *nMolDeposited);

/* Ni-C adhesion energy: metal-side LJ pair energies within rrCutLJ, each metal-carbon pair once. */
float orc_adhesion_f32(const orc_params *p, const float *pos4, const int *NN, const int *NBL);
double orc_adhesion_f64(const orc_params *p, const double *pos4, const int *NN, const int *NBL);

/* Trajectory window of the adjoint: a(r0) from a list built at r0, then nsteps steps opening with a
 * rebuild.  J[0] = mean adhesion energy of the steps, J[1] = final metal centre-of-mass height
 * above z_0.  pos4/vel/acc are updated in place.  Returns 0 or an orc_run error code. */
int orc_window_f32(const orc_params *p, float *pos4, float *vel, float *acc, int nsteps, int stepCount0,
                   int rebuildEvery, float *J);
int orc_window_f64(const orc_params *p, double *pos4, double *vel, double *acc, int nsteps, int stepCount0,
                   int rebuildEvery, double *J);

/* Tangent (forward-mode) build of the FP64 oracle, tangent.cpp: ORC_ND directions per call.
 * A direction seeds any of: the four parameter scale factors (pdir[ORC_ND][4]), the initial
 * positions (rdir[ORC_ND][3*nMol], x y z per atom) and initial velocities (vdir[ORC_ND][3*nMol]);
 * NULL seeds nothing.  Values are bitwise those of the *_f64 functions. */
#define ORC_ND 4
/* orc_window_f64 with derivatives: J[2], dJ[2][ORC_ND]. */
int orc_tangent_window(const orc_params *p, double *pos4, double *vel, double *acc, int nsteps, int stepCount0,
                       int rebuildEvery, const double *pdir, const double *rdir, const double *vdir, double *J,
                       double *dJ);
/* Density and forces at pos4 with the given list: acc[3*nMol], dacc[ORC_ND][3*nMol], adhesion
 * energy *adh and dadh[ORC_ND]. */
void orc_tangent_forces(const orc_params *p, const double *pos4, const int *NN, const int *NBL, const double *pdir,
                        const double *rdir, double *acc, double *dacc, double *adh, double *dadh);

/* Decision tape of the FP64 step loop (see oracle_impl.c): mode 1 clears it and records the rebuild
 * flag of every step, every list built and the F(rho) branch of every metal atom at every force pass
 * of the runs that follow (orc_run_f64, orc_window_f64); mode 2 replays them from the start instead of
 * deciding; 0 turns it off.  Returns the tape length (ints). */
long long orc_freeze_f64(int mode);
/* 1 when the replay since the last orc_freeze_f64(2) took exactly the recorded decisions. */
int orc_freeze_ok_f64(void);

#ifdef __cplusplus
}
#endif

#endif /* SG_ORACLE_H */
