// sgk_params.h - run constants passed by value to every kernel (kernel-argument segment), and the
// device-resident run state.  Plain structs usable from host and device code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgk {

struct DevParams {
    int nMol, nMolMe, nC, iRegime;
    float Lx, Ly, Lz;          // region
    int cx, cy, cz, ncell;     // cell grid
    float iwx, iwy, iwz;       // invWidth
    float rrNebr, rrCutEam, rrCutLJ, rrCutC, closeR2;
    // EAM
    float re, fe, rhoe, alpha, beta, A, B, kappa, lambda, Fn[4], F[4], eta, Fe, massMe, rei, rhoei;
    // graphene
    float mu_r, theta_0, mu_theta, mu_p, z_0;
    // LJ
    float epsLJ, sigma2, inv_sigma2;
    // integration / thermostat
    float deltaT, temperature, kB, gamma, rNebrShell;
    int stepThermostat, stepEquil, stepCool;
    float xLeft, xRight, yBottom, yTop;  // frozen graphene perimeter (SurfaceGrowth.cu:415-424)
    // RDF
    float rangeRdf2, intervalRdf;
    int sizeHistRdf;
    // shear
    float deltaF;
    // deposition (regime 1)
    float velMagDepos;
    int stepDeposit, nMolToDeposit;
    // capacities and policies
    int cellCap, nlCap, ljCap, clCap, keepFull;
    int rebuildEvery;
    float dispBound;           // max dispHi for which the carbon close lists are exact
    int xcdMap, binGroup;      // launch-shape options (sgk_config.xcd_map, bin_group)
    int nlScheme;              // sgk_config.nl_scheme
    float sigmaLJ;             // LJ sigma (the FP64 derivatives of the adjoint square it in FP64)
};

// Two cell buffers; state.cellbuf names the current one (see the rebuild protocol in sgk_nebr.hip).
// cellOf[b][n] is atom n's cell in buffer b (-1: not binned); slotOf[n] its slot in the buffer
// being built; cellPos[c*cap + s] the float position of the atom in slot s of cell c after the sort,
// with the atom's index (its bit pattern) in w, so a list walk reads one 16-byte entry per candidate
// (built with the buffer, read by the neighbour-list pass).
struct CellBufs {
    int *cnt[2];
    int *slots[2];
    int *cellOf[2];
    int *slotOf;
    float4 *cellPos;
    // Buffer k (0 or 1) taken by a select, not by indexing: a kernel that builds the struct in
    // registers (the batched ones, per replica) keeps it there instead of copying it to memory.
    __host__ __device__ int *cntB(int k) const { return k ? cnt[1] : cnt[0]; }
    __host__ __device__ int *slotsB(int k) const { return k ? slots[1] : slots[0]; }
    __host__ __device__ int *cellOfB(int k) const { return k ? cellOf[1] : cellOf[0]; }
};

// Per-block partial reductions written by the force kernels (or the props kernel).
template <class T>
struct Partial {
    T u, fr, com[3], cmv[3], vs[3], vv, vmax, bmin[3], bmax[3];
    T vir;  // pair virial 1/2 sum f_ik |d_ik|^2 of the EAM and LJ terms (force pass only)
};

// Device-resident run state (one per context).
struct DevState {
    int step;              // stepCount after the last finished step
    int nebrNow;           // rebuild at the next step
    int errors;            // SGK_FLAG_* bits
    int cellbuf;           // cell buffer holding the current lists (0/1)
    int rec;               // next record slot
    int rebuilds;          // rebuild counter
    double dispHi;
    double shear;
    double vvSum;          // of the last step (thermostat input)
    double comX;           // metal centre of mass x of the last step (shear input)
    double cmVelX;         // metal centre-of-mass velocity x of the last step
    unsigned int shearCount;
    int nMolDeposited;     // metal atoms deposited so far (regime 1)
};

// Grid of the complete-range RDF: cells at least rangeRdf/2 wide, so a pair within range lies within
// two cells per axis; s* are the stencil half-widths (the whole axis when 2s+1 would wrap onto itself).
struct RdfGrid {
    int nx, ny, nz, ncell;
    float iwx, iwy, iwz;
    float wx, wy, wz;  // cell widths L/n
    float cull2;       // a cell whose box lies at least this far (squared) from the atom's cell is skipped
    int sx, sy, sz;    // stencil half-widths in cells
};

// Device buffers of the centre-of-mass diffusion (sgk_diffusion_enable): nBuff staggered buffers of
// nVal squared displacements each, in FP64.
constexpr int kMaxDiffuseBuffers = 256;
struct DiffusionArgs {
    int *count;        // [nBuff] samples taken per buffer (negative: not started)
    int *state;        // [2] buffers averaged so far, tables completed
    double *org, *tru; // [3 nBuff] origin and unwrapped position of each buffer
    double *rr;        // [nBuff nVal]
    double *av;        // [nVal] running average
    double *out;       // [cap nVal] completed tables (ring)
    int nVal, nBuff, stepDiffuse, limitAv, cap;
    double deltaT;     // the float time step, widened
    double L[3];       // the float box, widened
};

}  // namespace sgk
