/*
 * oracle_impl.c - precision-generic CPU restatement of the SurfaceGrowth step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled twice by oracle/Makefile:
 *   -DORC_F32 : REAL = float, float libm functions      -> *_f32 symbols
 *   -DORC_F64 : REAL = double, double libm functions    -> *_f64 symbols
 * with -ffp-contract=off so every + and * rounds separately, as written.
 *
 * Citations: .cu = SurfaceGrowthConsole/SurfaceGrowth.cu, .cpp = SurfaceGrowthConsole/SurfaceGrowth.cpp,
 * Defs.h = SurfaceGrowthConsole/ComputeDefs.h.
 */
#include "oracle.h"

/* OMP_FOR: the per-atom (or per-cell) loops whose iterations write only their own atom's outputs run
 * in parallel in the CPU-baseline timing build (-fopenmp); every other build ignores it. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(dynamic, 256)")
#define OMP_FOR_ERR _Pragma("omp parallel for schedule(dynamic, 256) reduction(| : err)")
#else
#define OMP_FOR
#define OMP_FOR_ERR
#endif

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORC_DUAL) /* tangent build (tangent.cpp): double values, ND derivative directions */
typedef orc::Dual REAL;
#define FN(name) name##_dual
#define EXP orc::exp
#define POW orc::pow
#define LOG orc::log
#define SQRT orc::sqrt
#define ACOS orc::acos
#define CEIL orc::ceil
#elif defined(ORC_F32)
typedef float REAL;
#define FN(name) name##_f32
#define EXP expf
#define POW powf
#define LOG logf
#define SQRT sqrtf
#define ACOS acosf
#define CEIL ceilf
#elif defined(ORC_F64)
typedef double REAL;
#define FN(name) name##_f64
#define EXP exp
#define POW pow
#define LOG log
#define SQRT sqrt
#define ACOS acos
#define CEIL ceil
#else
#error "compile with -DORC_F32, -DORC_F64 or (tangent.cpp) ORC_DUAL"
#endif

/* Float literals of the reference, promoted so the FP64 oracle evaluates the same model. */
#define K(x) ((REAL)(x##f))

/* Parameters the adjoint differentiates, each times its scale factor parScale[k] (1 by default,
 * which leaves every value bitwise unchanged): LJ epsilon and sigma, the embedding function F (all
 * three branches) and the electron density f.  The tangent build seeds the scale factors. */
#if defined(ORC_DUAL)
static REAL g_scale[4];
#define PSCALE(p, k) (g_scale[k])
#else
#define PSCALE(p, k) ((REAL)(p)->parScale[k])
#endif
#define PAR_EPS(p) ((REAL)(p)->epsLJ * PSCALE(p, 0))
#define PAR_SIG(p) ((REAL)(p)->sigmaLJ * PSCALE(p, 1))
#define PAR_FN(p, k) ((REAL)(p)->Fn[k] * PSCALE(p, 2))
#define PAR_F(p, k) ((REAL)(p)->F[k] * PSCALE(p, 2))
#define PAR_FE(p) ((REAL)(p)->Fe * PSCALE(p, 2))
#define PAR_DENS(p) ((REAL)(p)->fe * PSCALE(p, 3))

/* 27-cell offset table (Defs.h:221-234). */
static const int kOffsets[27][3] = {
    {0, 0, 0},   {0, 0, 1},   {0, 0, -1},  {0, 1, 0},   {0, 1, 1},   {0, 1, -1},  {0, -1, 0},
    {0, -1, 1},  {0, -1, -1}, {1, 0, 0},   {1, 0, 1},   {1, 0, -1},  {1, 1, 0},   {1, 1, 1},
    {1, 1, -1},  {1, -1, 0},  {1, -1, 1},  {1, -1, -1}, {-1, 0, 0},  {-1, 0, 1},  {-1, 0, -1},
    {-1, 1, 0},  {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, -1, -1}};

/* Single-image minimum-image correction used by every force kernel (.cu:752-759). */
static inline REAL wrap1(REAL d, REAL L) {
    if (d >= K(0.5) * L)
        d -= L;
    else if (d < -K(0.5) * L)
        d += L;
    return d;
}

/* ------------------------------------------------------------------------------------------ */
/* Cell binning (.cu:547-609): c = VLinear(int((r + L/2) * invWidth)), truncation toward zero.
 * Atoms are listed per cell in ascending index order (the reference's slot order comes from
 * atomicAdd and is not defined; ascending index is the canonical order used throughout). */
int FN(orc_bin)(const orc_params *p, const REAL *pos4, int *cellOf, int *count, int *slots) {
    const int ncell = p->cells[0] * p->cells[1] * p->cells[2];
    int err = 0;
    memset(count, 0, sizeof(int) * (size_t)ncell);
    for (int n = 0; n < p->nMol; ++n) {
        /* The reference computes the cell of every atom in FP32 (.cu:594-600); every precision
         * form uses that FP32 formula on the float-rounded position, so cells, neighbour lists and
         * their order (which the carbon loop depends on) are precision independent. */
        float sx = (float)pos4[4 * n + 0] + 0.5f * p->region[0];
        float sy = (float)pos4[4 * n + 1] + 0.5f * p->region[1];
        float sz = (float)pos4[4 * n + 2] + 0.5f * p->region[2];
        int cx = (int)(sx * p->invWidth[0]);
        int cy = (int)(sy * p->invWidth[1]);
        int cz = (int)(sz * p->invWidth[2]);
        int c = (cz * p->cells[1] + cy) * p->cells[0] + cx;
        if (c < 0 || c >= ncell) { /* no block has this index in the reference: atom not binned */
            cellOf[n] = -1;
            continue;
        }
        cellOf[n] = c;
        if (count[c] >= p->cellCap) {
            err = 1;
            continue;
        }
        slots[(size_t)c * p->cellCap + count[c]] = n;
        ++count[c];
    }
    return err;
}

/* Neighbour list, Anderson et al. algorithm 4 as in BuildNebrListK (.cu:612-711): for each
 * binned atom i, the 27 offset cells in table order with the periodic shift of the cell wrap,
 * atoms of each cell in slot order, r^2 < rrNebr and k != i.  Column-major NBL[j*nMol+i].
 * FP32 on float-rounded positions in every precision form (see orc_bin). */
int FN(orc_nebr)(const orc_params *p, const REAL *pos4, const int *count, const int *slots, int *NN,
                 int *NBL) {
    const int nMol = p->nMol, cap = p->cellCap;
    const int *cells = p->cells;
    int err = 0;
    for (size_t q = 0; q < (size_t)nMol * p->iNebrMax; ++q) NBL[q] = -1;
    for (int i = 0; i < nMol; ++i) NN[i] = 0;
    OMP_FOR_ERR
    for (int c = 0; c < cells[0] * cells[1] * cells[2]; ++c) {
        const int ccz = c / (cells[0] * cells[1]);
        const int ccy = (c - ccz * cells[0] * cells[1]) / cells[0];
        const int ccx = c - (ccz * cells[1] + ccy) * cells[0];
        for (int s = 0; s < count[c]; ++s) {
            const int i = slots[(size_t)c * cap + s];
            const float Ax = (float)pos4[4 * i], Ay = (float)pos4[4 * i + 1], Az = (float)pos4[4 * i + 2];
            int nn = 0;
            for (int off = 0; off < 27; ++off) {
                int mx = ccx + kOffsets[off][0], my = ccy + kOffsets[off][1], mz = ccz + kOffsets[off][2];
                float shx = 0, shy = 0, shz = 0;
                if (mx >= cells[0]) { mx = 0; shx = p->region[0]; }
                else if (mx < 0) { mx = cells[0] - 1; shx = -p->region[0]; }
                if (my >= cells[1]) { my = 0; shy = p->region[1]; }
                else if (my < 0) { my = cells[1] - 1; shy = -p->region[1]; }
                if (mz >= cells[2]) { mz = 0; shz = p->region[2]; }
                else if (mz < 0) { mz = cells[2] - 1; shz = -p->region[2]; }
                const int C = (mz * cells[1] + my) * cells[0] + mx;
                for (int t = 0; t < count[C]; ++t) {
                    const int k = slots[(size_t)C * cap + t];
                    float dx = Ax - (float)pos4[4 * k], dy = Ay - (float)pos4[4 * k + 1], dz = Az - (float)pos4[4 * k + 2];
                    dx = dx - shx;
                    dy = dy - shy;
                    dz = dz - shz;
                    float rr = dx * dx + dy * dy + dz * dz;
                    if (rr < p->rrNebr && i != k) {
                        if (nn < p->iNebrMax)
                            NBL[(size_t)nn * nMol + i] = k;
                        else
                            err = 2;
                        ++nn;
                    }
                }
            }
            NN[i] = nn < p->iNebrMax ? nn : p->iNebrMax;
        }
    }
    return err;
}

/* ------------------------------------------------------------------------------------------ */
/* EAM functions (.cu:46-210), Zhou et al. Acta Mater. 49 (2001) 4005. */
static inline REAL eam_phi(const orc_params *p, REAL r) {
    REAL x = r * (REAL)p->rei;
    return (REAL)p->A * EXP(-(REAL)p->alpha * (x - K(1.))) / (K(1.) + POW(x - (REAL)p->kappa, K(20.))) -
           (REAL)p->B * EXP(-(REAL)p->beta * (x - K(1.))) / (K(1.) + POW(x - (REAL)p->lambda, K(20.)));
}
static inline REAL eam_f(const orc_params *p, REAL r) {
    REAL x = r * (REAL)p->rei;
    return PAR_DENS(p) * EXP(-(REAL)p->beta * (x - K(1.))) / (K(1.) + POW(x - (REAL)p->lambda, K(20.)));
}
static inline REAL eam_dphi(const orc_params *p, REAL r) {
    const REAL rei = p->rei, al = p->alpha, be = p->beta, ka = p->kappa, la = p->lambda;
    REAL x = r * rei;
    REAL dA = K(1.) + POW(x - ka, K(20.));
    REAL dB = K(1.) + POW(x - la, K(20.));
    return (-(REAL)p->A * EXP(-al * (x - K(1.))) * (al + K(20.) * POW(x - ka, K(19.)) / dA) / dA +
            (REAL)p->B * EXP(-be * (x - K(1.))) * (be + K(20.) * POW(x - la, K(19.)) / dB) / dB) *
           rei;
}
static inline REAL eam_df(const orc_params *p, REAL r) {
    const REAL rei = p->rei, be = p->beta, la = p->lambda;
    REAL x = r * rei;
    REAL dB = K(1.) + POW(x - la, K(20.));
    return -PAR_DENS(p) * rei * EXP(-be * (x - K(1.))) * (be + K(20.) * POW(x - la, K(19.)) / dB) / dB;
}
/* Branch of the embedding function at density rho: 0 below 0.85 rho_e, 1 up to 1.15 rho_e, 2 above
 * (the branch points of .cu:72-103 and 160-195). */
static inline int eam_branch(const orc_params *p, REAL rho) {
    const REAL rhoe = p->rhoe, rhon = K(0.85) * rhoe, rhoo = K(1.15) * rhoe;
    return rho < rhon ? 0 : (rho < rhoo ? 1 : 2);
}

/* F(rho) on branch br (br < 0: the branch of rho). */
static inline REAL eam_F_br(const orc_params *p, REAL rho, int br) {
    if (br < 0) br = eam_branch(p, rho);
    if (br == 0) { /* .cu:72-82 */
        REAL rhoni = (REAL)p->rhoei * K(1.176471);
        REAL t = rho * rhoni - K(1.);
        return PAR_FN(p, 0) + PAR_FN(p, 1) * t + PAR_FN(p, 2) * t * t + PAR_FN(p, 3) * t * t * t;
    }
    if (br == 1) { /* .cu:84-94 */
        REAL t = rho * (REAL)p->rhoei - K(1.);
        return PAR_F(p, 0) + PAR_F(p, 1) * t + PAR_F(p, 2) * t * t + PAR_F(p, 3) * t * t * t;
    }
    /* .cu:96-103 */
    REAL x = rho * (REAL)p->rhoei;
    return PAR_FE(p) * (K(1.) - LOG(POW(x, (REAL)p->eta))) * POW(x, (REAL)p->eta);
}
static inline REAL eam_F(const orc_params *p, REAL rho) { return eam_F_br(p, rho, -1); }

/* F'(rho) on branch br (br < 0: the branch of rho). */
static inline REAL eam_dF_br(const orc_params *p, REAL rho, int br) {
    if (br < 0) br = eam_branch(p, rho);
    if (br == 0) { /* .cu:160-171 */
        REAL rhoni = (REAL)p->rhoei * K(1.176471);
        REAL t = rho * rhoni - K(1.);
        return rhoni * (PAR_FN(p, 1) + K(2.) * PAR_FN(p, 2) * t + K(3.) * PAR_FN(p, 3) * t * t);
    }
    if (br == 1) { /* .cu:173-184 */
        REAL x = rho * (REAL)p->rhoei;
        return (REAL)p->rhoei * (PAR_F(p, 1) + K(2.) * PAR_F(p, 2) * (x - K(1.)) +
                                 K(3.) * PAR_F(p, 3) * (x - K(1.)) * (x - K(1.)));
    }
    /* .cu:186-195 */
    REAL x = rho * (REAL)p->rhoei, eta = p->eta;
    return -PAR_FE(p) * eta * (REAL)p->rhoei * POW(x, (eta - K(1.))) * LOG(POW(x, eta));
}
static inline REAL eam_dF(const orc_params *p, REAL rho) { return eam_dF_br(p, rho, -1); }

/* ---- decision tape of the FP64 build (orc_freeze_f64) ------------------------------------------
 * The adjoint differentiates the step map with its discrete decisions held: the neighbour lists
 * (contents and order), the steps that rebuild them, the branch of F(rho) of every metal atom and
 * which list pairs lie inside the EAM and LJ cutoffs.  In record mode the step loop appends each
 * decision in the order it takes them (the rebuild flag of every step, every list built, the cutoff
 * bits of every density, force and adhesion pass, the branches of every force pass); in replay mode it takes them from
 * the tape instead, so a run with perturbed parameters follows the recorded run's decisions and its
 * finite differences are those of the map the adjoint differentiates.  Test infrastructure only. */
/* Allocation for the decision tape and the cutoff bits: test infrastructure, so running out of
 * memory ends the process with a message. */
static inline void *orc_need(void *q) {
    if (!q) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return q;
}

#if defined(ORC_F64)
static struct {
    int mode; /* 0 off, 1 record, 2 replay */
    int *v;
    size_t len, cap, pos;
    int bad; /* a replay ran past the end of the tape */
} g_tape;
static inline int tape_mode(void) { return g_tape.mode; }
static inline void tape_push(const int *x, size_t n) {
    if (g_tape.len + n > g_tape.cap) {
        size_t c = g_tape.cap ? 2 * g_tape.cap : 1 << 20;
        while (c < g_tape.len + n) c *= 2;
        g_tape.v = (int *)orc_need(realloc(g_tape.v, c * sizeof(int)));
        g_tape.cap = c;
    }
    memcpy(g_tape.v + g_tape.len, x, n * sizeof(int));
    g_tape.len += n;
}
static inline void tape_pull(int *x, size_t n) {
    if (g_tape.pos + n > g_tape.len) {
        g_tape.bad = 1;
        memset(x, 0, n * sizeof(int));
        return;
    }
    memcpy(x, g_tape.v + g_tape.pos, n * sizeof(int));
    g_tape.pos += n;
}
long long orc_freeze_f64(int mode) {
    if (mode == 1) g_tape.len = 0;
    g_tape.pos = 0;
    g_tape.bad = 0;
    g_tape.mode = mode;
    return (long long)g_tape.len;
}
/* 1 when the last replay took exactly the recorded decisions, 0 otherwise. */
int orc_freeze_ok_f64(void) { return !g_tape.bad && g_tape.pos == g_tape.len; }
#else
static inline int tape_mode(void) { return 0; }
static inline void tape_push(const int *x, size_t n) { (void)x; (void)n; }
static inline void tape_pull(int *x, size_t n) { (void)x; (void)n; }
#endif

REAL FN(orc_embed)(const orc_params *p, REAL rho, REAL *dF) {
    if (dF) *dF = eam_dF(p, rho);
    return eam_F(p, rho);
}

void FN(orc_pair)(const orc_params *p, REAL r, REAL *out4) {
    out4[0] = eam_phi(p, r);
    out4[1] = eam_dphi(p, r);
    out4[2] = eam_f(p, r);
    out4[3] = eam_df(p, r);
}

/* Cutoff decisions of one pass over list rows [0, nRows), for the decision tape: entry j of row i is
 * 1 when the pair lies inside its cutoff (rrCutEam for two metal atoms, rrCutLJ for metal and carbon;
 * carbon-carbon entries are not taped and read 1).  Recorded from the positions, or replayed; NULL
 * with the tape off.  *off receives the first entry of each row. */
static unsigned char *FN(cut_tape)(const orc_params *p, const REAL *pos4, const int *NN, const int *NBL, int nRows,
                                   size_t **off) {
    if (!tape_mode()) return NULL;
    const int nMe = p->nMolMe;
    const REAL Lx = p->region[0], Ly = p->region[1], Lz = p->region[2];
    size_t *o = (size_t *)orc_need(malloc(sizeof(size_t) * ((size_t)nRows + 1)));
    o[0] = 0;
    for (int i = 0; i < nRows; ++i) o[i + 1] = o[i] + (size_t)NN[i];
    const size_t n = o[nRows], nw = (n + 31) / 32;
    unsigned char *cut = (unsigned char *)orc_need(malloc(n ? n : 1));
    int *w = (int *)orc_need(calloc(nw ? nw : 1, sizeof(int)));
    if (tape_mode() == 2) {
        tape_pull(w, nw);
        for (size_t q = 0; q < n; ++q) cut[q] = (unsigned char)(((unsigned)w[q / 32] >> (q % 32)) & 1u);
    } else {
        for (int i = 0; i < nRows; ++i)
            for (int j = 0; j < NN[i]; ++j) {
                const int k = NBL[(size_t)j * p->nMol + i];
                REAL dx = wrap1(pos4[4 * i] - pos4[4 * k], Lx);
                REAL dy = wrap1(pos4[4 * i + 1] - pos4[4 * k + 1], Ly);
                REAL dz = wrap1(pos4[4 * i + 2] - pos4[4 * k + 2], Lz);
                REAL rr = dx * dx + dy * dy + dz * dz;
                const int mi = i < nMe, mk = k < nMe;
                int in = 1;
                if (mi && mk) in = rr < (REAL)p->rrCutEam;
                else if (mi || mk) in = rr < (REAL)p->rrCutLJ;
                const size_t q = o[i] + (size_t)j;
                cut[q] = (unsigned char)in;
                if (in) w[q / 32] = (int)((unsigned)w[q / 32] | (1u << (q % 32)));
            }
        tape_push(w, nw);
    }
    free(w);
    *off = o;
    return cut;
}

/* EamComputeRhoK (.cu:716-773): rho_i = sum over metal neighbours with r^2 < rrCutEam. */
void FN(orc_rho)(const orc_params *p, const REAL *pos4, const int *NN, const int *NBL, REAL *rho) {
    const REAL Lx = p->region[0], Ly = p->region[1], Lz = p->region[2];
    size_t *off = NULL;
    unsigned char *cut = FN(cut_tape)(p, pos4, NN, NBL, p->nMolMe, &off);
    OMP_FOR
    for (int i = 0; i < p->nMolMe; ++i) {
        REAL sum = 0;
        for (int j = 0; j < NN[i]; ++j) {
            const int k = NBL[(size_t)j * p->nMol + i];
            REAL dx = wrap1(pos4[4 * i] - pos4[4 * k], Lx);
            REAL dy = wrap1(pos4[4 * i + 1] - pos4[4 * k + 1], Ly);
            REAL dz = wrap1(pos4[4 * i + 2] - pos4[4 * k + 2], Lz);
            REAL rr = dx * dx + dy * dy + dz * dz;
            const int in = cut ? cut[off[i] + (size_t)j] : rr < (REAL)p->rrCutEam;
            if (in && k < p->nMolMe) sum += eam_f(p, SQRT(rr));
        }
        rho[i] = sum;
    }
    free(cut);
    free(off);
}

/* Lennard-Jones pair (.cu:880-884): returns the force factor, energy in *u. */
static inline REAL lj_pair(const orc_params *p, REAL rr, REAL *u) {
    const REAL sig = PAR_SIG(p), s2 = sig * sig;
    REAL rri = s2 / rr;
    REAL rri3 = rri * rri * rri;
    *u = K(4.) * PAR_EPS(p) * rri3 * (rri3 - K(1.));
    return K(48.) * PAR_EPS(p) * rri3 * (rri3 - K(0.5)) * rri / s2;
}

/* ComputeForcesK (.cu:776-1091), one atom at a time, neighbours in list order.  vir (may be NULL)
 * receives each atom's pair virial 1/2 sum_k f_ik |d_ik|^2 over the pair-decomposable forces it
 * evaluates - EAM and metal-carbon LJ, every LJ visit of the carbon loop counted - not the graphene
 * law, whose terms have no pair reaction (no counterpart in the reference). */
static void FN(forces_impl)(const orc_params *p, REAL *pos4, const int *NN, const int *NBL, const REAL *rho,
                            REAL *acc, REAL *fric, REAL *vir) {
    const int nMol = p->nMol, nMe = p->nMolMe;
    const REAL Lx = p->region[0], Ly = p->region[1], Lz = p->region[2];
    const REAL z0 = p->z_0, rrC = p->rrCutC, mu_r = p->mu_r, mu_p = p->mu_p, mu_t = p->mu_theta,
               th0 = p->theta_0;
    int *br = NULL; /* F(rho) branches from the decision tape (record or replay), else from rho */
    if (tape_mode() && nMe) {
        br = (int *)orc_need(malloc(sizeof(int) * (size_t)nMe));
        if (tape_mode() == 2) {
            tape_pull(br, (size_t)nMe);
        } else {
            for (int i = 0; i < nMe; ++i) br[i] = eam_branch(p, rho[i]);
            tape_push(br, (size_t)nMe);
        }
    }
    size_t *off = NULL;
    unsigned char *cut = FN(cut_tape)(p, pos4, NN, NBL, nMol, &off);
    OMP_FOR
    for (int i = 0; i < nMol; ++i) {
        const int metal = i < nMe;
        REAL fx = 0, fy = 0, fz = 0, frx = 0, uSum = 0, wSum = 0;
        const REAL Ax = pos4[4 * i], Ay = pos4[4 * i + 1], Az = pos4[4 * i + 2];
        const int nn = NN[i];
        const REAL rhoi = metal ? rho[i] : 0;
        const REAL dFi = metal ? eam_dF_br(p, rhoi, br ? br[i] : -1) : 0;
        REAL rnn[3][3], absRnn[3], deltaZ[3];
        int counti = 0;
        for (int j = 0; j < nn; ++j) {
            const int k = NBL[(size_t)j * nMol + i];
            const REAL Bx = pos4[4 * k], By = pos4[4 * k + 1], Bz = pos4[4 * k + 2];
            REAL dx = wrap1(Ax - Bx, Lx), dy = wrap1(Ay - By, Ly), dz = wrap1(Az - Bz, Lz);
            REAL rr = dx * dx + dy * dy + dz * dz;
            if (metal) {
                if (k < nMe) { /* EAM (.cu:857-877) */
                    const int in = cut ? cut[off[i] + (size_t)j] : rr < (REAL)p->rrCutEam;
                    REAL r = SQRT(rr);
                    REAL fc = -eam_dphi(p, r);
                    REAL u = eam_phi(p, r);
                    fc = fc - eam_df(p, r) * (dFi + eam_dF_br(p, rho[k], br ? br[k] : -1));
                    fc = fc / r;
                    if (!in) fc = 0;
                    fx += fc * dx;
                    fy += fc * dy;
                    fz += fc * dz;
                    wSum += fc * rr;
                    if (in) uSum += u;
                } else { /* LJ with carbon (.cu:878-896) */
                    const int in = cut ? cut[off[i] + (size_t)j] : rr < (REAL)p->rrCutLJ;
                    REAL u, fc = lj_pair(p, rr, &u);
                    if (!in) fc = 0;
                    if (in) uSum += u;
                    fx += fc * dx;
                    fy += fc * dy;
                    fz += fc * dz;
                    frx += fc * dx;
                    wSum += fc * rr;
                }
                continue;
            }
            if (k < nMe) { /* carbon: LJ with metal (.cu:900-917) */
                const int in = cut ? cut[off[i] + (size_t)j] : rr < (REAL)p->rrCutLJ;
                REAL u, fc = lj_pair(p, rr, &u);
                if (!in) fc = 0;
                if (in) uSum += u;
                fx += fc * dx;
                fy += fc * dy;
                fz += fc * dz;
                wSum += fc * rr;
                continue;
            }
            /* carbon-carbon: first three neighbours within rCutC are the bonds (.cu:921-984) */
            if (rr < rrC && counti < 3) {
                absRnn[counti] = SQRT(rr);
                deltaZ[counti] = Bz - z0;
                rnn[counti][0] = dx;
                rnn[counti][1] = dy;
                rnn[counti][2] = dz;
                /* walk the list of k for its (up to two) other bonded neighbours m */
                int countk = 0;
                for (int l = 0; l < NN[k]; ++l) {
                    const int m = NBL[(size_t)l * nMol + k];
                    const REAL mx = pos4[4 * m], my = pos4[4 * m + 1], mz = pos4[4 * m + 2];
                    REAL ex = wrap1(Bx - mx, Lx), ey = wrap1(By - my, Ly), ez = wrap1(Bz - mz, Lz);
                    REAL rr2 = ex * ex + ey * ey + ez * ez;
                    if (rr2 < rrC && m != i) {
                        ex = ex * K(-1.); /* rm - rk */
                        ey = ey * K(-1.);
                        ez = ez * K(-1.);
                        /* the reference divides by |r_km|^2 here (.cu:957, 962) */
                        REAL c = (rnn[counti][0] * ex + rnn[counti][1] * ey + rnn[counti][2] * ez) /
                                 (absRnn[counti] * rr2);
                        REAL dth = ACOS(c) - th0;
                        REAL tmp = mu_t * dth / SQRT(K(1.) - c * c);
                        REAL fc = tmp / (absRnn[counti] * rr2);
                        fx += fc * ex;
                        fy += fc * ey;
                        fz += fc * ez;
                        fc = -tmp * c / (absRnn[counti] * absRnn[counti]);
                        fx += fc * rnn[counti][0];
                        fy += fc * rnn[counti][1];
                        fz += fc * rnn[counti][2];
                        fz += -mu_p * (mz - z0) / K(9.); /* no energy for these terms (.cu:968) */
                        ++countk;
                    }
                    if (countk == 2) break;
                }
                ++counti;
            }
            if (counti == 3) { /* bonded terms of i (.cu:985-1061) */
                for (int b = 0; b < 3; ++b) {
                    REAL fc = -mu_r * (absRnn[b] - K(1.)) / absRnn[b];
                    fx += fc * rnn[b][0];
                    fy += fc * rnn[b][1];
                    fz += fc * rnn[b][2];
                    uSum += mu_r * (absRnn[b] - K(1.)) * (absRnn[b] - K(1.));
                }
                REAL dZ = K(2.) * (Az - z0) - (deltaZ[0] + deltaZ[1] + deltaZ[2]);
                fz += -K(0.666666667) * mu_p * dZ;
                dZ = (Az - z0) - K(0.33333333) * (deltaZ[0] + deltaZ[1] + deltaZ[2]);
                uSum += mu_p * dZ * dZ;
                /* central angles for the pairs (0,1), (1,2), (2,0) */
                static const int pa[3] = {0, 1, 2}, pb[3] = {1, 2, 0};
                for (int q = 0; q < 3; ++q) {
                    const int a = pa[q], b = pb[q];
                    REAL c = (rnn[a][0] * rnn[b][0] + rnn[a][1] * rnn[b][1] + rnn[a][2] * rnn[b][2]) /
                             (absRnn[a] * absRnn[b]);
                    REAL dth = ACOS(c) - th0;
                    REAL tmp = mu_t * dth / SQRT(K(1.) - c * c);
                    REAL fcb = tmp * (K(1.) - absRnn[a] * c / absRnn[b]) / (absRnn[a] * absRnn[b]);
                    fx += fcb * rnn[b][0];
                    fy += fcb * rnn[b][1];
                    fz += fcb * rnn[b][2];
                    REAL fca = tmp * (K(1.) - absRnn[b] * c / absRnn[a]) / (absRnn[a] * absRnn[b]);
                    fx += fca * rnn[a][0];
                    fy += fca * rnn[a][1];
                    fz += fca * rnn[a][2];
                    uSum += mu_t * dth * dth;
                }
                ++counti;
                /* The reference reuses the neighbour-loop counter for the bond loop (.cu:989), so
                 * the neighbour loop resumes at list index 4 after the third bond is found. */
                j = 3;
            }
        }
        if (metal) {
            acc[3 * i] = fx / (REAL)p->massMe;
            acc[3 * i + 1] = fy / (REAL)p->massMe;
            acc[3 * i + 2] = fz / (REAL)p->massMe;
            pos4[4 * i + 3] = K(0.5) * uSum + eam_F_br(p, rhoi, br ? br[i] : -1);
            if (fric) fric[i] = frx;
        } else {
            acc[3 * i] = fx;
            acc[3 * i + 1] = fy;
            acc[3 * i + 2] = fz;
            pos4[4 * i + 3] = K(0.5) * uSum;
        }
        if (vir) vir[i] = K(0.5) * wSum;
    }
    free(br);
    free(cut);
    free(off);
}

void FN(orc_forces)(const orc_params *p, REAL *pos4, const int *NN, const int *NBL, const REAL *rho,
                    REAL *acc, REAL *fric) {
    FN(forces_impl)(p, pos4, NN, NBL, rho, acc, fric, NULL);
}

void FN(orc_forces_virial)(const orc_params *p, REAL *pos4, const int *NN, const int *NBL, const REAL *rho,
                           REAL *acc, REAL *fric, REAL *vir) {
    FN(forces_impl)(p, pos4, NN, NBL, rho, acc, fric, vir);
}

/* ------------------------------------------------------------------------------------------ */
/* LeapfrogStepK (.cu:397-434) and ApplyBoundaryCondK (.cu:436-453). */
void FN(orc_leapfrog)(const orc_params *p, int part, REAL *pos4, REAL *vel, const REAL *acc) {
    const float cos30 = (float)cos(M_PI / 6.0);
    const float xLeft = 0.5f * (cos30 - p->region[0]), xRight = 0.5f * (p->region[0] - cos30);
    const float yBottom = 0.5f * (1.f - p->region[1]), yTop = 0.5f * (p->region[1] - 1.f);
    const REAL hdt = K(0.5) * (REAL)p->deltaT, dt = p->deltaT;
    OMP_FOR
    for (int n = 0; n < p->nMol; ++n) {
        REAL *r = pos4 + 4 * n, *v = vel + 3 * n;
        const REAL *a = acc + 3 * n;
        if (n >= p->nMolMe) { /* carbon perimeter atoms stay fixed (.cu:423-424) */
            if (!(r[0] != (REAL)xLeft && r[0] < (REAL)xRight && r[1] != (REAL)yTop && r[1] != (REAL)yBottom))
                continue;
        }
        v[0] = v[0] + hdt * a[0];
        v[1] = v[1] + hdt * a[1];
        v[2] = v[2] + hdt * a[2];
        if (part == 1) {
            r[0] = r[0] + dt * v[0];
            r[1] = r[1] + dt * v[1];
            r[2] = r[2] + dt * v[2];
        }
    }
    if (part != 1) return;
    for (int n = 0; n < p->nMol; ++n)
        for (int d = 0; d < 3; ++d) {
            REAL L = p->region[d], x = pos4[4 * n + d];
            if (x >= K(0.5) * L)
                x -= L;
            else if (x < -K(0.5) * L)
                x += L;
            pos4[4 * n + d] = x;
        }
}

/* ApplyBerendsenThermostat (.cu:455-485) with the kinetic sum vv of the step. */
static void FN(thermostat_impl)(const orc_params *p, REAL *vel, REAL vv, int stepCount) {
    const REAL ke = vv * K(0.5) / (REAL)p->nMol;
    const REAL bC = SQRT(K(1.) + (REAL)p->gammaBerendsen *
                                     ((REAL)p->temperature * K(1.5) * (REAL)p->kB / ke - K(1.)));
    const REAL bM = SQRT(K(1.) + (REAL)p->gammaBerendsen * ((REAL)p->temperature * (REAL)p->kB / ke - K(1.)));
    const int metalOn = p->iRegime == 2 && stepCount > p->stepEquil && stepCount < p->stepEquil + p->stepCool;
    for (int n = 0; n < p->nMol; ++n) {
        REAL b;
        if (n >= p->nMolMe)
            b = bC;
        else if (metalOn)
            b = bM;
        else
            continue;
        vel[3 * n] = b * vel[3 * n];
        vel[3 * n + 1] = b * vel[3 * n + 1];
        vel[3 * n + 2] = b * vel[3 * n + 2];
    }
}

void FN(orc_thermostat)(const orc_params *p, REAL *vel, double vvSum, int stepCount) {
    FN(thermostat_impl)(p, vel, (REAL)vvSum, stepCount);
}

/* The thermostat's kinetic sum: sum m|v|^2, with the metal centre-of-mass x velocity removed from
 * the metal vx in regime 2 (ComputeVvSumK, .cu:1230-1234; cmVel.val of .cu:2308). */
static REAL FN(vv_sum)(const orc_params *p, const REAL *vel) {
    const int nMol = p->nMol, nMe = p->nMolMe;
    const REAL m = p->massMe;
    REAL cmv0 = 0;
    for (int n = 0; n < nMe; ++n) cmv0 += vel[3 * n];
#if !defined(ORC_F32)
    REAL cmVelXr = nMe ? cmv0 / (REAL)nMe : 0;
#else
    REAL cmVelXr = nMe ? (float)cmv0 / (float)nMe : 0.f; /* cmVel.val, .cu:2308 */
#endif
    REAL vv = 0;
    for (int n = 0; n < nMol; ++n) {
        const REAL *v = vel + 3 * n;
        REAL vx = v[0];
        if (p->iRegime == 2 && n < nMe) vx = vx - cmVelXr;
        REAL s = vx * vx + v[1] * v[1] + v[2] * v[2];
        if (n < nMe) s = m * s;
        vv += s;
    }
    return vv;
}

/* Reductions (.cu:1101-1891) and the host arithmetic that consumes them (.cu:2307-2395).
 * Sums are accumulated sequentially in REAL (the reference's FP32 tree order is not defined by
 * anything observable).  The ComputeNetForceK in-place race (.cu:1828-1858) is not reproduced. */
void FN(orc_props)(const orc_params *p, const REAL *pos4, const REAL *vel, const REAL *fric,
                   orc_props *out) {
    const int nMol = p->nMol, nMe = p->nMolMe;
    const REAL m = p->massMe;
    REAL vs[3] = {0, 0, 0}, vvMax = 0, u = 0, com[3] = {0, 0, 0}, cmv[3] = {0, 0, 0}, fr = 0;
    REAL mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    memset(out, 0, sizeof(*out));
    for (int n = 0; n < nMol; ++n) {
        const REAL *v = vel + 3 * n;
        vs[0] += n < nMe ? m * v[0] : v[0]; /* only x is mass weighted (.cu:1120-1126) */
        vs[1] += v[1];
        vs[2] += v[2];
        REAL vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (vv > vvMax) vvMax = vv;
        u += pos4[4 * n + 3];
    }
    /* metal bounding box: entries past nMolMe enter the reference min/max as zeros
     * (.cu:1671-1682), which is the case for every system below 2^20 metal atoms. */
    const int pad = nMe != 1024 * 1024;
    for (int n = 0; n < nMe; ++n) {
        for (int d = 0; d < 3; ++d) {
            REAL x = pos4[4 * n + d];
            com[d] += x;
            cmv[d] += vel[3 * n + d];
            if ((n == 0 && !pad) || x < mn[d]) mn[d] = x;
            if ((n == 0 && !pad) || x > mx[d]) mx[d] = x;
        }
        if (fric && p->iRegime != 0) fr += fric[n];
    }
    /* regime-2 kinetic energy removes the metal centre-of-mass x velocity (.cu:1230-1234) */
    const REAL vv = FN(vv_sum)(p, vel);
    for (int d = 0; d < 3; ++d) {
        out->vSum[d] = (double)vs[d];
        out->comSum[d] = (double)com[d];
        out->cmVelSum[d] = (double)cmv[d];
        out->bboxMin[d] = nMe ? (double)mn[d] : 0;
        out->bboxMax[d] = nMe ? (double)mx[d] : 0;
    }
    out->vvSum = (double)vv;
    out->vvMax = (double)vvMax;
    out->uSum = (double)u;
    out->frict = (double)fr;
    out->kinEnergy = (double)(K(0.5) * vv / (REAL)nMol);
    out->potEnergy = (double)(u / (REAL)nMol);
    out->totEnergy = (double)((REAL)out->kinEnergy + (REAL)out->potEnergy);
}

/* EvalRdfK (.cu:1895-1999): metal-metal pairs from the 27-cell stencil of the given cells with
 * the cell-wrap shift, r^2 < rangeRdf^2, bin = int(sqrt(r^2)*intervalRdf); FP32 as in the
 * reference, on float-rounded positions in every precision form. */
int FN(orc_rdf)(const orc_params *p, const REAL *pos4, const int *cellOf, const int *count,
                const int *slots, unsigned int *hist) {
    const int *cells = p->cells, cap = p->cellCap, nMe = p->nMolMe;
    const float range2 = p->rangeRdf * p->rangeRdf;
    int edge = 0;
    (void)cellOf;
    for (int c = 0; c < cells[0] * cells[1] * cells[2]; ++c) {
        const int ccz = c / (cells[0] * cells[1]);
        const int ccy = (c - ccz * cells[0] * cells[1]) / cells[0];
        const int ccx = c - (ccz * cells[1] + ccy) * cells[0];
        for (int s = 0; s < count[c]; ++s) {
            const int i = slots[(size_t)c * cap + s];
            if (i >= nMe) continue;
            const float Ax = (float)pos4[4 * i], Ay = (float)pos4[4 * i + 1], Az = (float)pos4[4 * i + 2];
            for (int off = 0; off < 27; ++off) {
                int mx = ccx + kOffsets[off][0], my = ccy + kOffsets[off][1], mz = ccz + kOffsets[off][2];
                float shx = 0, shy = 0, shz = 0;
                if (mx >= cells[0]) { mx = 0; shx = p->region[0]; }
                else if (mx < 0) { mx = cells[0] - 1; shx = -p->region[0]; }
                if (my >= cells[1]) { my = 0; shy = p->region[1]; }
                else if (my < 0) { my = cells[1] - 1; shy = -p->region[1]; }
                if (mz >= cells[2]) { mz = 0; shz = p->region[2]; }
                else if (mz < 0) { mz = cells[2] - 1; shz = -p->region[2]; }
                const int C = (mz * cells[1] + my) * cells[0] + mx;
                for (int t = 0; t < count[C]; ++t) {
                    const int k = slots[(size_t)C * cap + t];
                    if (k >= nMe) continue;
                    float dx = Ax - (float)pos4[4 * k], dy = Ay - (float)pos4[4 * k + 1], dz = Az - (float)pos4[4 * k + 2];
                    dx = dx - shx;
                    dy = dy - shy;
                    dz = dz - shz;
                    float rr = dx * dx + dy * dy + dz * dz;
                    if (rr < range2 && i != k) {
                        int b = (int)(sqrtf(rr) * p->intervalRdf);
                        if (b >= p->sizeHistRdf) {
                            ++edge;
                            continue;
                        }
                        ++hist[b];
                    }
                }
            }
        }
    }
    return edge;
}

/* ------------------------------------------------------------------------------------------ */
/* Diffusion of the centre of mass, Rapaport's staggered multi-origin buffers (.cu:2865-2952). */
struct orc_diffusion {
    int nVal, nBuff, stepDiffuse, limitAv, countAv;
    double deltaT, region[3];
    double *orgR, *rTrue; /* 3 per buffer */
    double *rr;           /* nBuff*nVal */
    int *count;
    double *av;           /* nVal */
};

#if defined(ORC_F64) /* precision-independent lifecycle: defined once */
orc_diffusion *orc_diffusion_create(int nVal, int nBuff, int stepDiffuse, int limitAv, double deltaT,
                                    const float *region) {
    orc_diffusion *d = (orc_diffusion *)calloc(1, sizeof(orc_diffusion));
    d->nVal = nVal;
    d->nBuff = nBuff;
    d->stepDiffuse = stepDiffuse;
    d->limitAv = limitAv;
    d->deltaT = deltaT;
    for (int k = 0; k < 3; ++k) d->region[k] = region[k];
    d->orgR = (double *)calloc((size_t)3 * nBuff, sizeof(double));
    d->rTrue = (double *)calloc((size_t)3 * nBuff, sizeof(double));
    d->rr = (double *)calloc((size_t)nBuff * nVal, sizeof(double));
    d->count = (int *)calloc((size_t)nBuff, sizeof(int));
    d->av = (double *)calloc((size_t)nVal, sizeof(double));
    for (int nb = 0; nb < nBuff; ++nb) d->count[nb] = -nb * nVal / nBuff; /* .cu:2869-2870 */
    return d;
}

void orc_diffusion_destroy(orc_diffusion *d) {
    if (!d) return;
    free(d->orgR);
    free(d->rTrue);
    free(d->rr);
    free(d->count);
    free(d->av);
    free(d);
}
#endif

#if !defined(ORC_DUAL) /* the diffusion buffers hold doubles; not part of the tangent build */
static inline REAL nint_r(REAL x) { /* Nint (Defs.h:64-65); its 0.5 literal makes this double math */
    double xd = (double)x;
    return (REAL)(xd < 0. ? -(int)(0.5 - xd) : (int)(0.5 + xd));
}

int FN(orc_diffusion_sample)(orc_diffusion *d, const REAL *com, REAL *out) {
    int done = 0;
    for (int nb = 0; nb < d->nBuff; ++nb) { /* EvalDiffusion (.cu:2923-2952) */
        double *org = d->orgR + 3 * nb, *tr = d->rTrue + 3 * nb;
        if (d->count[nb] == 0)
            for (int k = 0; k < 3; ++k) org[k] = tr[k] = (REAL)com[k];
        if (d->count[nb] >= 0) {
            const int ni = d->count[nb];
            REAL acc = 0;
            for (int k = 0; k < 3; ++k) {
                REAL L = (REAL)(float)d->region[k];
                REAL dr = (REAL)tr[k] - com[k];
                dr = dr / L;
                dr = nint_r(dr);
                dr = dr * L;
                REAL t = com[k] + dr;
                tr[k] = t;
                REAL e = t - (REAL)org[k];
                acc += e * e;
            }
            d->rr[(size_t)nb * d->nVal + ni] = acc;
        }
        ++d->count[nb];
    }
    for (int nb = 0; nb < d->nBuff; ++nb) { /* AccumDiffusion (.cu:2898-2920) */
        if (d->count[nb] != d->nVal) continue;
        for (int j = 0; j < d->nVal; ++j) d->av[j] = (REAL)d->av[j] + (REAL)d->rr[(size_t)nb * d->nVal + j];
        d->count[nb] = 0;
        if (++d->countAv == d->limitAv) {
            REAL fac = (REAL)(1. / (3 * 2 * d->stepDiffuse * (REAL)(float)d->deltaT * d->limitAv));
            for (int j = 0; j < d->nVal; ++j) out[j] = (j >= 1) ? (REAL)d->av[j] * (fac / (REAL)j) : (REAL)d->av[j];
            done = 1;
            d->countAv = 0;
            for (int j = 0; j < d->nVal; ++j) d->av[j] = 0;
        }
    }
    return done;
}
#endif

/* ------------------------------------------------------------------------------------------ */
/* InsertAtomsK (.cu:488-514): the group [dep, dep + toDeposit) enters at z = 0.49 Lz moving down
 * at velMagDepos; metal atoms not yet due are parked above the box at rest. */
static void FN(orc_insert)(const orc_params *p, REAL *pos4, REAL *vel, int dep, int toDeposit) {
    const float Lz = p->region[2];
    const int interval = dep + toDeposit;
    for (int n = 0; n < p->nMolMe; ++n) {
        if (n >= dep && n < interval) {
            pos4[4 * n + 2] = (REAL)(0.49f * Lz);
            vel[3 * n + 2] = (REAL)(-p->velMagDepos);
        }
        if (n >= interval) {
            pos4[4 * n + 2] = (REAL)(1.5f * Lz + 0.5f * (float)(n + 1) * Lz);
            vel[3 * n] = vel[3 * n + 1] = vel[3 * n + 2] = 0;
        }
    }
}

/* Ni-C adhesion energy: the metal-side Lennard-Jones pair energies of ComputeForcesK (.cu:878-896),
 * each metal-carbon pair within rrCutLJ once (an observable of the adjoint; no reference output). */
REAL FN(orc_adhesion)(const orc_params *p, const REAL *pos4, const int *NN, const int *NBL) {
    const REAL Lx = p->region[0], Ly = p->region[1], Lz = p->region[2];
    size_t *off = NULL;
    unsigned char *cut = FN(cut_tape)(p, pos4, NN, NBL, p->nMolMe, &off);
    REAL sum = 0;
    for (int i = 0; i < p->nMolMe; ++i) {
        for (int j = 0; j < NN[i]; ++j) {
            const int k = NBL[(size_t)j * p->nMol + i];
            if (k < p->nMolMe) continue;
            REAL dx = wrap1(pos4[4 * i] - pos4[4 * k], Lx);
            REAL dy = wrap1(pos4[4 * i + 1] - pos4[4 * k + 1], Ly);
            REAL dz = wrap1(pos4[4 * i + 2] - pos4[4 * k + 2], Lz);
            REAL rr = dx * dx + dy * dy + dz * dz;
            REAL u;
            (void)lj_pair(p, rr, &u);
            if (cut ? cut[off[i] + (size_t)j] : rr < (REAL)p->rrCutLJ) sum += u;
        }
    }
    free(cut);
    free(off);
    return sum;
}

/* A neighbour-list build (bin + list) or, replaying the decision tape, the recorded list. */
static int FN(list_build)(const orc_params *p, REAL *pos4, int *cellOf, int *count, int *slots, int *NN, int *NBL) {
    const int nMol = p->nMol;
    if (tape_mode() == 2) {
        tape_pull(NN, (size_t)nMol);
        for (size_t q = 0; q < (size_t)nMol * p->iNebrMax; ++q) NBL[q] = -1;
        for (int i = 0; i < nMol; ++i)
            for (int j = 0; j < NN[i]; ++j) tape_pull(&NBL[(size_t)j * nMol + i], 1);
        return 0;
    }
    int err = FN(orc_bin)(p, pos4, cellOf, count, slots);
    err |= FN(orc_nebr)(p, pos4, count, slots, NN, NBL);
    if (tape_mode() == 1) {
        tape_push(NN, (size_t)nMol);
        for (int i = 0; i < nMol; ++i)
            for (int j = 0; j < NN[i]; ++j) tape_push(&NBL[(size_t)j * nMol + i], 1);
    }
    return err;
}

/* The step loop of DoComputationsW (.cu:2141-2489) with device work restated serially; adh[s]
 * (optional) receives the adhesion energy of step s. */
static int FN(run_impl)(const orc_params *p, REAL *pos4, REAL *vel, REAL *acc, int nsteps, int stepCount0,
                        int *nebrNow, double *dispHi, int rebuildEvery, orc_record *records, int *nMolDeposited,
                        REAL *adh) {
    const int nMol = p->nMol, nMe = p->nMolMe;
    const int ncell = p->cells[0] * p->cells[1] * p->cells[2];
    int *cellOf = (int *)malloc(sizeof(int) * (size_t)nMol);
    int *count = (int *)malloc(sizeof(int) * (size_t)ncell);
    int *slots = (int *)malloc(sizeof(int) * (size_t)ncell * p->cellCap);
    int *NN = (int *)calloc((size_t)nMol, sizeof(int));
    int *NBL = (int *)malloc(sizeof(int) * (size_t)nMol * p->iNebrMax);
    REAL *rho = (REAL *)calloc((size_t)(nMe ? nMe : 1), sizeof(REAL));
    REAL *fric = (REAL *)calloc((size_t)(nMe ? nMe : 1), sizeof(REAL));
    int err = 0, haveList = 0;
    REAL dh = (REAL)*dispHi, shear = 0;
    for (int s = 0; s < nsteps && !err; ++s) {
        const int step = stepCount0 + s + 1;
        orc_record rec;
        memset(&rec, 0, sizeof(rec));
        if (p->iRegime == 1 && nMe && nMolDeposited) { /* deposition schedule (.cu:2148-2168) */
            if (step <= p->stepEquil) {
                FN(orc_insert)(p, pos4, vel, *nMolDeposited, 0);
            } else if (*nMolDeposited < nMe) {
                int to = 0;
                if (p->stepDeposit > 0 && (step - p->stepEquil) % p->stepDeposit == 0) to = p->nMolToDeposit;
                FN(orc_insert)(p, pos4, vel, *nMolDeposited, to);
                *nMolDeposited += to;
            }
        }
        FN(orc_leapfrog)(p, 1, pos4, vel, acc);
        int rebuild = *nebrNow || !haveList;
        if (tape_mode() == 2)
            tape_pull(&rebuild, 1);
        else if (tape_mode() == 1)
            tape_push(&rebuild, 1);
        if (rebuild) {
            *nebrNow = 0;
            dh = 0;
            err |= FN(list_build)(p, pos4, cellOf, count, slots, NN, NBL);
            haveList = 1;
            rec.rebuilt = 1;
        }
        if (nMe) FN(orc_rho)(p, pos4, NN, NBL, rho);
        FN(orc_forces)(p, pos4, NN, NBL, rho, acc, fric);
        if (adh) adh[s] = FN(orc_adhesion)(p, pos4, NN, NBL);
        FN(orc_leapfrog)(p, 2, pos4, vel, acc);
        orc_props pr;
        FN(orc_props)(p, pos4, vel, (p->iRegime != 0 && nMe) ? fric : NULL, &pr);
        if ((p->iRegime == 2 && step > p->stepEquil && step % p->stepThermostat == 0) ||
            (p->iRegime == 1 && step % p->stepThermostat == 0))
            FN(thermostat_impl)(p, vel, FN(vv_sum)(p, vel), step);
        dh += SQRT((REAL)pr.vvMax) * (REAL)p->deltaT;
        if (dh > K(0.5) * (REAL)p->rNebrShell) *nebrNow = 1; /* .cu:2390-2391 */
        if (rebuildEvery > 0 && step % rebuildEvery == 0) *nebrNow = 1; /* fixed cadence adds rebuilds */
        /* shear (.cu:2398-2414) */
        if (p->iRegime == 2 && nMe && step > p->stepEquil + p->stepCool) {
            REAL cmVelVal = (REAL)pr.cmVelSum[0] / (REAL)nMe;
            if (cmVelVal < K(0.005)) shear += (REAL)p->deltaF;
            REAL comX = (REAL)pr.comSum[0] / (REAL)nMe;
            for (int n = 0; n < nMe; ++n)
                if (comX - pos4[4 * n] > 0) acc[3 * n] += shear;
        }
        if (records) {
            rec.step = step;
            rec.nebrNow = *nebrNow;
            rec.kinEnergy = pr.kinEnergy;
            rec.potEnergy = pr.potEnergy;
            rec.totEnergy = pr.totEnergy;
            for (int d = 0; d < 3; ++d) {
                rec.vSum[d] = pr.vSum[d];
                rec.comSum[d] = pr.comSum[d];
                rec.cmVelSum[d] = pr.cmVelSum[d];
                rec.bboxMin[d] = pr.bboxMin[d];
                rec.bboxMax[d] = pr.bboxMax[d];
            }
            rec.vvSum = pr.vvSum;
            rec.vvMax = pr.vvMax;
            rec.uSum = pr.uSum;
            rec.frict = pr.frict;
            rec.dispHi = (double)dh;
            rec.nMolDeposited = nMolDeposited ? *nMolDeposited : nMe;
            records[s] = rec;
        }
    }
    *dispHi = (double)dh;
    free(cellOf);
    free(count);
    free(slots);
    free(NN);
    free(NBL);
    free(rho);
    free(fric);
    return err;
}

int FN(orc_run)(const orc_params *p, REAL *pos4, REAL *vel, REAL *acc, int nsteps, int stepCount0,
                int *nebrNow, double *dispHi, int rebuildEvery, orc_record *records, int *nMolDeposited) {
    return FN(run_impl)(p, pos4, vel, acc, nsteps, stepCount0, nebrNow, dispHi, rebuildEvery, records,
                        nMolDeposited, NULL);
}

/* A trajectory window for the adjoint: the acceleration of the initial state from a neighbour list
 * built at that state, then nsteps steps opening with a rebuild.  J[0] = mean adhesion energy over
 * the steps, J[1] = metal centre-of-mass height above the sheet plane after the last step. */
int FN(orc_window)(const orc_params *p, REAL *pos4, REAL *vel, REAL *acc, int nsteps, int stepCount0,
                   int rebuildEvery, REAL *J) {
    const int nMol = p->nMol, nMe = p->nMolMe;
    const int ncell = p->cells[0] * p->cells[1] * p->cells[2];
    int *cellOf = (int *)malloc(sizeof(int) * (size_t)nMol);
    int *count = (int *)malloc(sizeof(int) * (size_t)ncell);
    int *slots = (int *)malloc(sizeof(int) * (size_t)ncell * p->cellCap);
    int *NN = (int *)calloc((size_t)nMol, sizeof(int));
    int *NBL = (int *)malloc(sizeof(int) * (size_t)nMol * p->iNebrMax);
    REAL *rho = (REAL *)calloc((size_t)(nMe ? nMe : 1), sizeof(REAL));
    REAL *fric = (REAL *)calloc((size_t)(nMe ? nMe : 1), sizeof(REAL));
    REAL *adh = (REAL *)calloc((size_t)(nsteps > 0 ? nsteps : 1), sizeof(REAL));
    int err = FN(list_build)(p, pos4, cellOf, count, slots, NN, NBL);
    if (nMe) FN(orc_rho)(p, pos4, NN, NBL, rho);
    FN(orc_forces)(p, pos4, NN, NBL, rho, acc, fric);
    int nebrNow = 1;
    double dispHi = 0;
    if (!err)
        err = FN(run_impl)(p, pos4, vel, acc, nsteps, stepCount0, &nebrNow, &dispHi, rebuildEvery, NULL, NULL,
                           adh);
    REAL e = 0, z = 0;
    for (int s = 0; s < nsteps; ++s) e += adh[s];
    for (int n = 0; n < nMe; ++n) z += pos4[4 * n + 2];
    J[0] = nsteps > 0 ? e / (REAL)nsteps : (REAL)0;
    J[1] = nMe ? z / (REAL)nMe - (REAL)p->z_0 : (REAL)0;
    free(cellOf);
    free(count);
    free(slots);
    free(NN);
    free(NBL);
    free(rho);
    free(fric);
    free(adh);
    return err;
}
