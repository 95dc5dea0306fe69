// sgk_adjoint.h - buffers and launchers of the adjoint (reverse-mode) pass of the step.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "sgk.h"
#include "sgk_params.h"

struct sgk_ctx;

namespace sgk {

// Device arrays of the reverse sweep (FP64; per-atom vectors SoA x[n], y[N+n], z[2N+n]).
struct AdjDev {
    double *Rb = nullptr;     // adjoint of the positions r_n
    double *Vb = nullptr;     // adjoint of the velocities v_n
    double *Ab = nullptr;     // adjoint of the accelerations a_n
    double *W = nullptr;      // adjoint of the velocities before the thermostat of step n
    double *Lam = nullptr;    // cotangent of the force pass of step n
    double *G = nullptr;      // vector-Jacobian product of the force pass (d/dr)
    double *rhoBar = nullptr; // adjoint of the electron density [nMolMe]
    double *rho = nullptr, *dF = nullptr, *ddF = nullptr;  // FP64 rho, F'(rho), F''(rho) [nMolMe]
    double *pbar = nullptr;   // [4] d/d scale of the differentiated parameters
    double *therm = nullptr;  // [8] thermostat transpose scalars of the current step
    double *part = nullptr;   // block partials (thermostat sums, adhesion energy)
    double *adh = nullptr;    // adhesion energy per step of the forward window
    int *mismatch = nullptr;  // recomputed-vs-stored state differences
};

// One copy of the neighbour structures the force pass reads.
struct NlSnap {
    int *nlM = nullptr, *nnM = nullptr, *nlj = nullptr, *ncl = nullptr;
    unsigned *lj = nullptr, *cl = nullptr;
    int *fullNN = nullptr, *fullNBL = nullptr;  // only with keep_full_nl
};

struct Checkpoint {
    void *pos = nullptr, *vel = nullptr, *acc = nullptr;
    NlSnap nl;
    DevState st;
    // cell buffers: kept for the final state only (null elsewhere)
    int *cnt[2] = {nullptr, nullptr}, *slots[2] = {nullptr, nullptr}, *cellOf[2] = {nullptr, nullptr};
    int *slotOf = nullptr;
    float4 *cellPos = nullptr;
};

struct Adjoint {
    int nsteps = 0, every = 0, segments = 0, step0 = 0, cap = 0;
    bool ready = false;
    double obs[2] = {0, 0};
    std::vector<Checkpoint> ck;   // segments + 1 (the last is the state the window ends in)
    std::vector<NlSnap> slots;    // lists rebuilt inside the segment being recomputed
    std::vector<int> rebuilt;     // per window step
    std::string thermo;           // per window step: '1' if the thermostat ran
    void *posStore = nullptr;     // [every] positions after each step of a segment
    void *velStore = nullptr;     // [every] velocities before the thermostat of each step
    AdjDev d;
    std::vector<void *> allocs;
    long long bytes = 0;
    int mismatches = 0;
    double forwardMs = 0, backwardMs = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Free every device buffer of the adjoint (sgk_adjoint_api.hip).
void adjoint_free(Adjoint *a);

// Kernels (sgk_adjoint.hip).  Mixed-precision contexts only.
void launch_adh_obs(sgk_ctx *c, double *part, double *out, int recIndex);  // out[st.rec-1] if recIndex
void launch_adj_rho(sgk_ctx *c, const AdjDev &d);
void launch_adj_force(sgk_ctx *c, const AdjDev &d, double wE);
void launch_adj_rho2(sgk_ctx *c, const AdjDev &d);
void launch_adj_therm(sgk_ctx *c, const AdjDev &d, const double *velPre, int step);
void launch_adj_pre(sgk_ctx *c, const AdjDev &d, const double *velPre);
void launch_adj_post(sgk_ctx *c, const AdjDev &d);
void launch_adj_lam0(sgk_ctx *c, const AdjDev &d);
void launch_adj_seed(sgk_ctx *c, const AdjDev &d, double wHeight);
void launch_adj_cmp(sgk_ctx *c, const AdjDev &d, const void *pos, const void *vel, const void *acc);

}  // namespace sgk
