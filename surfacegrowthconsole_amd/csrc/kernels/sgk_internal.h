// sgk_internal.h - host-side engine context and the kernel launchers of each translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "sgk.h"
#include "sgk_adjoint.h"
#include "sgk_batch.h"
#include "sgk_params.h"

namespace sgk {

// Buckets per block of the two-pass counting-sort scan (k_ord_blocksum / k_ord_scan); the block-total
// arrays hold one int per kOrdChunk buckets.
constexpr int kOrdChunk = 1024;

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    unsigned long long uses = 0;
};

struct KernelStat {
    int launches = 0;
    double ms = 0;
};

}  // namespace sgk

struct sgk_ctx {
    SimParams sp;
    sgk_config cfg;
    sgk::DevParams dp;
    bool mixed = false;
    int device = 0;
    hipStream_t stream = nullptr;

    // state arrays (positions AoS x,y,z,E; velocity and acceleration SoA, storage precision R)
    void *pos = nullptr;   // float4 or double4 [nMol]
    void *posAlt = nullptr;  // the same: where a force pass with the fused drift writes the next positions
    void *vel = nullptr;   // R [3*nMol]
    void *acc = nullptr;   // R [3*nMol] (FP64 in the mixed form)
    float *rho = nullptr, *embF = nullptr, *embDF = nullptr, *fric = nullptr;  // [nMolMe]
    // cells (two buffers, see the rebuild protocol in sgk_nebr.hip)
    int *cellOf[2] = {nullptr, nullptr};
    int *cnt[2] = {nullptr, nullptr};
    int *slots[2] = {nullptr, nullptr};
    int *slotOf = nullptr;
    float4 *cellPos = nullptr;
    // neighbour structures
    int *nlM = nullptr, *nnM = nullptr;                 // metal: row-major [nMolMe*nlCap]
    unsigned int *lj = nullptr, *cl = nullptr;          // carbon: column-major packed (idx<<8|pos)
    int *nlj = nullptr, *ncl = nullptr;
    uint4 *bondC = nullptr, *bondA = nullptr;          // carbon: per-step near lists (k_rho)
    int *fullNN = nullptr, *fullNBL = nullptr;          // reference layout (keep_full_nl)
    // order in which nl_scheme 3 builds the carbon lists (carbons by cell, then metal atoms) and the
    // counting-sort scratch that makes it (sgk_nebr.hip)
    int *nlOrder = nullptr, *ordCnt = nullptr, *ordKey = nullptr, *ordSlot = nullptr, *ordBlock = nullptr;
    bool orderValid = false;
    int orderAge = 0;                                   // steps run since the order was built
    // reductions and run state
    void *partials = nullptr;
    sgk::DevState *state = nullptr;
    sgk_record *records = nullptr;
    unsigned int *hist = nullptr;
    sgk::DevState *hostState = nullptr;                 // pinned mirror
    // launch geometry
    int block = 256, metalLanes = 8, nbM = 0, nbC = 0, nbRho = 0, nbProps = 0;
    int part2Off = 0;                                   // chunk totals of the step reduction (in partials)
    long long bytes = 0;
    bool haveList = false;
    bool haveCells = false;                             // sgk_bin ran on the current positions
    int hostStep = 0;                                   // host copy of the step counter
    double lastRunMs = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<std::string, sgk::GraphEntry> graphs;
    bool profiling = false;
    std::map<std::string, sgk::KernelStat> stats;
    std::string err;
    sgk::Adjoint *adj = nullptr;                        // adjoint storage (sgk_adjoint_*), on demand
    sgk::DiffusionArgs diff = {};                       // device diffusion (sgk_diffusion_enable), on demand
    void *diffMem = nullptr;
    // complete-range RDF (cfg.rdf_complete): metal atoms sorted by the RDF grid at each sample
    sgk::RdfGrid rdfGrid = {};
    int *rdfCnt = nullptr, *rdfKey = nullptr, *rdfSlot = nullptr, *rdfBlock = nullptr;
    float4 *rdfPos = nullptr;
};

// A batch of replicas advanced together (sgk_batch_*): shared arrays, one range per replica.
namespace sgk {
// Lanes per metal atom of the cell-ordered list build (each takes two of every 2G candidates; 8
// measured no faster on the batch and 1.5 % slower on the 1.1 M-atom system).
constexpr int kNlMetalLanes = 4;
}

// Force-pass workgroups of nMe metal atoms at G lanes each.  With the density or force pass
// remapped to XCDs (xcd_map bit 1 or 2) the count is padded to a multiple of 8, so that the carbon
// range starts on XCD 0; without, the padding workgroups would only idle.
inline int metal_blocks(long long nMe, int G, int xcdMap) {
    const long long nb = (nMe * G + 255) / 256;
    return (int)((xcdMap & 6) ? (nb + 7) / 8 * 8 : nb);
}

struct sgk_batch {
    int nrep = 0;
    std::vector<SimParams> sp;
    sgk_config cfg;
    bool mixed = false;
    int device = 0;
    hipStream_t stream = nullptr;
    int metalLanes = 4;                 // the common lanes per metal atom (0 when replicas differ)
    std::vector<int> repLanes;          // lanes per metal atom of each replica
    // density and force passes: one launch per distinct lane count, its blocks at forceTab[grpOff[g]]
    static constexpr int kLaneKinds = 5;  // 1, 2, 4, 8, 16
    int grpOff[kLaneKinds] = {}, grpN[kLaneKinds] = {};
    std::vector<sgk::BatchRep> reps;  // host copy of the replica table
    sgk::BatchRep *dReps = nullptr;
    int2 *atomTab = nullptr, *forceTab = nullptr, *metalTab = nullptr, *nlTab = nullptr;
    int nAtomBlocks = 0, nForceBlocks = 0, nMetalBlocks = 0, nNlBlocks = 0;
    // cell-ordered list build (nl_scheme 3): the order of every replica from one counting sort
    int *order = nullptr, *ordCnt = nullptr, *ordKey = nullptr, *ordSlot = nullptr, *ordBlock = nullptr;
    long long nOrdBuckets = 0;
    bool orderValid = false;
    int orderAge = 0;
    long long nAtoms = 0, nMetal = 0, nCarbon = 0, nCells = 0;
    void *pos = nullptr, *vel = nullptr, *acc = nullptr;
    void *posAlt = nullptr;  // where a force pass with the fused drift writes the next positions
    float *rho = nullptr, *embF = nullptr, *embDF = nullptr, *fric = nullptr;
    int *cellOf[2] = {nullptr, nullptr}, *cnt[2] = {nullptr, nullptr}, *slots[2] = {nullptr, nullptr};
    int *slotOf = nullptr;
    float4 *cellPos = nullptr;
    int *nlM = nullptr, *nnM = nullptr, *nlj = nullptr, *ncl = nullptr;
    unsigned int *lj = nullptr, *cl = nullptr;
    uint4 *bondC = nullptr, *bondA = nullptr;
    void *partials = nullptr;
    sgk::DevState *state = nullptr;
    sgk_record *records = nullptr;
    unsigned int *hist = nullptr;
    sgk::DevState *hostState = nullptr;  // pinned mirror, nrep entries
    long long bytes = 0;
    int hostStep = 0;
    double lastRunMs = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<std::string, sgk::GraphEntry> graphs;
    std::string err;
    std::vector<void *> allocs;

    sgk::BatchArgs args() const {
        sgk::BatchArgs a;
        a.reps = dReps;
        a.atomTab = atomTab;
        a.forceTab = forceTab;
        a.metalTab = metalTab;
        a.nlTab = nlTab;
        a.order = orderValid ? order : nullptr;
        a.pos = pos;
        a.posOut = nullptr;
        a.vel = vel;
        a.acc = acc;
        a.rho = rho;
        a.embF = embF;
        a.embDF = embDF;
        a.fric = fric;
        for (int k = 0; k < 2; ++k) {
            a.cellOf[k] = cellOf[k];
            a.cnt[k] = cnt[k];
            a.slots[k] = slots[k];
        }
        a.slotOf = slotOf;
        a.cellPos = cellPos;
        a.nlM = nlM;
        a.nnM = nnM;
        a.nlj = nlj;
        a.ncl = ncl;
        a.lj = lj;
        a.cl = cl;
        a.bondC = bondC;
        a.bondA = bondA;
        a.partials = partials;
        a.state = state;
        a.records = records;
        a.recCap = cfg.records_capacity;
        a.hist = hist;
        return a;
    }
};

namespace sgk {

// Device parameters of one system (sgk_api.hip).
void make_dev_params(const SimParams &s, const sgk_config &cfg, DevParams &p);
// Validates the parameters and the configuration (defaults filled in) and sets c->sp, cfg, mixed, dp
// and the launch geometry (lanes per metal atom, block counts, partials layout); allocates nothing.
// Returns SGK_OK or an error code with the reason in err.
int prepare_ctx(const SimParams *sp, const sgk_config *cfg, sgk_ctx *c, std::string &err);

// sgk_nebr.hip
void launch_bin(sgk_ctx *c, bool force);
void launch_batch_sortcells(sgk_batch *b);
void launch_batch_nl(sgk_batch *b);
void build_batch_nl_order(sgk_batch *b);
void launch_batch_rdf(sgk_batch *b);
void launch_sortcells(sgk_ctx *c, bool force);
void launch_nl(sgk_ctx *c, bool force);
void build_nl_order(sgk_ctx *c);
constexpr int kNlOrderSteps = 100;  // age (steps) at which sgk_run rebuilds the nl_scheme 3 list-build order
void launch_commit_rebuild(sgk_ctx *c);
void launch_rdf(sgk_ctx *c);
RdfGrid rdf_grid(const DevParams &p, int S);  // complete-range RDF grid, cells range/S (ncell 0: range too large)
// sgk_force.hip
void launch_rho(sgk_ctx *c);
void launch_batch_rho(sgk_batch *b);
int batch_lane_index(int G);  // position of G in the lane kinds 1, 2, 4, 8, 16
void launch_batch_force(sgk_batch *b, bool prekick = false, void *posOut = nullptr);
void launch_batch_force_eval(sgk_batch *b);  // forces, energies and partials only: velocities untouched
// prekick: the force pass also applies the next step's first half-kick (the caller guarantees that
// nothing touches velocities or accelerations before that step's launch_lf1(kicked = true)).
// posOut (with prekick): also drift the positions of the next step into posOut (sgk_force.hip).
void launch_force(sgk_ctx *c, bool integrate, bool finalize, bool prekick = false, void *posOut = nullptr);
// sgk_step.hip
void launch_lf1(sgk_ctx *c, bool bin, bool kicked = false);
void launch_insert(sgk_ctx *c);
// drifted: the previous force pass drifted the positions (posOut), so only the binning is left.
void launch_batch_lf1(sgk_batch *b, bool kicked = false, bool drifted = false);
void launch_batch_finalize(sgk_batch *b);
void launch_batch_thermostat(sgk_batch *b);
void launch_batch_shear(sgk_batch *b);
void launch_batch_reset_records(sgk_batch *b);
void launch_lf2(sgk_ctx *c);
void launch_props(sgk_ctx *c);
void launch_finalize(sgk_ctx *c, int mode);          // 0 step, 1 evaluate only
void launch_thermostat(sgk_ctx *c, int stepOverride, double vvOverride, int useOverride);
void launch_shear(sgk_ctx *c);
void launch_reset_records(sgk_ctx *c);
void launch_diffuse(sgk_ctx *c);
void launch_diffuse_reset(sgk_ctx *c);

// sgk_api.hip: error reporting, device flags, the step-kernel flags of a run and the device loop.
int set_error(sgk_ctx *c, int code, const std::string &msg);
int check_device_flags(sgk_ctx *c);
// One character per step, 'a' + bits: 1 thermostat, 2 shear, 4 RDF, 8 adhesion energy recorded,
// 16 diffusion sample.
std::string step_flag_string(sgk_ctx *c, int step0, int nsteps, int rdfEvery, bool observe);
// sgk_run with, if adh is not null, the adhesion energy of every step copied to adh[nsteps].
int run_steps(sgk_ctx *c, int nsteps, int rdfEvery, sgk_record *records, double *adh);
// Destroy the cached step graphs (their launches carry grids, parameters and buffer addresses).
void drop_graphs(sgk_ctx *c);

// Per-kernel timing hook used when profiling is on (sgk_api.hip).
void profile_begin(sgk_ctx *c);
void profile_end(sgk_ctx *c, const char *name);

// Partial-record size for the context precision.
size_t partial_bytes(const sgk_ctx *c);
size_t partial_bytes_for(bool mixed);

}  // namespace sgk
