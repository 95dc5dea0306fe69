// sgk_batch_api.hip - the batched engine's C ABI (sgk_batch_*): many replicas of uneven size
// (cluster size, sheet size, temperature) advance in lockstep, one launch per stage for all of
// them, captured step by step into hipGraphs.  Each replica keeps its own cells, lists, step state,
// reductions and records; its numbers are those of the same replica run alone in an sgk_ctx with
// xcd_map = 0 and the same lanes per metal atom (the same per-atom code over the same workgroup
// decomposition).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sgk_device.h"
#include "sgk_internal.h"

using sgk::BatchRep;
using sgk::DevParams;
using sgk::DevState;

namespace {

thread_local std::string g_batch_create_error;

int bfail(sgk_batch *b, int code, const std::string &msg) {
    if (b)
        b->err = msg;
    else
        g_batch_create_error = msg;
    return code;
}

#define BHIP(b, expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return bfail(b, SGK_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int balloc(sgk_batch *b, T **p, size_t count, bool zero) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (hipMalloc((void **)p, bytes) != hipSuccess)
        return bfail(b, SGK_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    b->allocs.push_back(*p);
    b->bytes += (long long)bytes;
    if (zero && hipMemsetAsync(*p, 0, bytes, b->stream) != hipSuccess) return bfail(b, SGK_E_HIP, "hipMemsetAsync");
    return SGK_OK;
}

// Fields every replica must share: the potentials, the time step, the list policy, the step
// schedule and the RDF binning.
const char *mismatch(const SimParams &a, const SimParams &s) {
    if (a.iRegime != s.iRegime) return "iRegime";
    if (a.iMaterial != s.iMaterial || a.re != s.re || a.fe != s.fe || a.A != s.A || a.B != s.B || a.massMe != s.massMe ||
        a.rCutEam != s.rCutEam)
        return "metal (EAM parameters)";
    if (a.epsLJ != s.epsLJ || a.sigmaLJ != s.sigmaLJ || a.rrCutLJ != s.rrCutLJ) return "LJ epsilon/sigma";
    if (a.mu_r != s.mu_r || a.mu_theta != s.mu_theta || a.mu_p != s.mu_p || a.rCutC != s.rCutC) return "graphene constants";
    if (a.deltaT != s.deltaT) return "deltaT";
    if (a.rNebrShell != s.rNebrShell || a.iNebrMax != s.iNebrMax) return "neighbour-list shell/capacity";
    if (a.stepThermostat != s.stepThermostat || a.stepEquil != s.stepEquil || a.stepCool != s.stepCool)
        return "thermostat schedule";
    if (a.sizeHistRdf != s.sizeHistRdf) return "RDF bins";
    return nullptr;
}

int check_batch_flags(sgk_batch *b) {
    BHIP(b, hipMemcpyAsync(b->hostState, b->state, sizeof(DevState) * b->nrep, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    const int fatal = SGK_FLAG_CELL_OVERFLOW | SGK_FLAG_NL_OVERFLOW | SGK_FLAG_LJ_OVERFLOW | SGK_FLAG_CLOSE_OVERFLOW |
                      SGK_FLAG_DISPLACEMENT;
    for (int r = 0; r < b->nrep; ++r) {
        const int f = b->hostState[r].errors & fatal;
        if (f)
            return bfail(b, (f & SGK_FLAG_DISPLACEMENT) ? SGK_E_DISPLACEMENT : SGK_E_OVERFLOW,
                         "replica " + std::to_string(r) + ": device flags " + std::to_string(f));
    }
    return SGK_OK;
}

std::string batch_flags(const sgk_batch *b, int step0, int nsteps, int rdfEvery) {
    const SimParams &s = b->sp[0];
    bool anyShear = false;
    for (const SimParams &q : b->sp) anyShear |= q.iRegime == 2 && q.nMolMe > 0 && q.deltaF != 0.f;
    std::string f(nsteps, 'a');
    for (int k = 0; k < nsteps; ++k) {
        const int step = step0 + k + 1;
        const bool thermo = s.stepThermostat > 0 && step % s.stepThermostat == 0 && s.iRegime == 2 && step > s.stepEquil;
        const bool shear = anyShear && step > s.stepEquil + s.stepCool;
        const bool rdf = rdfEvery > 0 && step % rdfEvery == 0;
        f[k] = (char)('a' + ((thermo ? 1 : 0) | (shear ? 2 : 0) | (rdf ? 4 : 0)));
    }
    return f;
}

void enqueue_batch(sgk_batch *b, const std::string &flags) {
    sgk::launch_batch_reset_records(b);
    // as enqueue_steps: the force pass applies the next step's first half-kick when no thermostat or
    // shear comes in between and the step is not the chunk's last, and it also drifts the positions
    // into posAlt on an even number of those steps (the first stays unfused when the count is odd),
    // the host swapping pos and posAlt after each, so the chunk ends on the buffer it started on
    const size_t n = flags.size();
    std::vector<char> pre(n, 0);
    int nPre = 0;
    for (size_t s = 0; s < n; ++s) {
        pre[s] = !((flags[s] - 'a') & 3) && s + 1 < n;
        nPre += pre[s];
    }
    bool unfuseFirst = (nPre & 1) != 0;
    bool kicked = false, drifted = false;
    for (size_t s = 0; s < n; ++s) {
        const int k = flags[s] - 'a';
        bool fuse = false;
        if (pre[s] && b->posAlt) {
            fuse = !unfuseFirst;
            unfuseFirst = false;
        }
        sgk::launch_batch_lf1(b, kicked, drifted);
        sgk::launch_batch_sortcells(b);
        sgk::launch_batch_nl(b);
        sgk::launch_batch_rho(b);
        sgk::launch_batch_force(b, pre[s], fuse ? b->posAlt : nullptr);
        sgk::launch_batch_finalize(b);
        kicked = pre[s];
        drifted = fuse;
        if (k & 1) sgk::launch_batch_thermostat(b);
        if (k & 2) sgk::launch_batch_shear(b);
        if (k & 4) sgk::launch_batch_rdf(b);  // this step's positions, still in pos
        if (fuse) std::swap(b->pos, b->posAlt);
    }
}

// Lanes per metal atom of a replica with nMe metal atoms when none is configured (the single-system
// rule measured on MI355X: 4 lanes from 65,536 metal atoms, 8 from 2,048, 16 below).
int default_lanes(int nMe) { return nMe >= 65536 ? 4 : nMe >= 2048 ? 8 : 16; }

// Force-pass layout for the current lanes per replica: each replica's metal and carbon blocks (as
// sgk_ctx lays them out), its partials and chunk totals, and the block table grouped by lane count.
std::vector<int2> force_layout(sgk_batch *b) {
    std::vector<std::vector<int2>> groups(sgk_batch::kLaneKinds);
    int partOff = 0;
    for (int r = 0; r < b->nrep; ++r) {
        BatchRep &R = b->reps[r];
        const int G = b->repLanes[r];
        R.nbM = metal_blocks(R.p.nMolMe, G, 0);  // the batch remaps nothing
        R.nbForce = R.nbM + (R.p.nC + 255) / 256;
        R.partOff = partOff;
        R.nChunks = (R.nbForce + 63) / 64;  // sgk::kPartChunk partials per chunk
        partOff += R.nbForce;
        auto &g = groups[sgk::batch_lane_index(G)];
        for (int k = 0; k < R.nbForce; ++k) g.push_back(make_int2(r, k));
    }
    for (BatchRep &R : b->reps) {  // chunk totals after all partials
        R.part2Off = partOff;
        partOff += R.nChunks;
    }
    std::vector<int2> tab;
    for (int q = 0; q < sgk_batch::kLaneKinds; ++q) {
        b->grpOff[q] = (int)tab.size();
        b->grpN[q] = (int)groups[q].size();
        tab.insert(tab.end(), groups[q].begin(), groups[q].end());
    }
    b->nForceBlocks = (int)tab.size();
    int common = b->repLanes.empty() ? 0 : b->repLanes[0];
    for (int G : b->repLanes)
        if (G != common) common = 0;
    b->metalLanes = common;
    return tab;
}

// Largest force-pass layout any lanes give (16 lanes per metal atom): sizes the block table and the
// partials once.
void worst_layout(const sgk_batch *b, size_t &tab, size_t &parts) {
    tab = parts = 0;
    for (const BatchRep &R : b->reps) {
        const size_t nb = (size_t)((((long long)R.p.nMolMe * 16 + 255) / 256 + 7) / 8 * 8) + (R.p.nC + 255) / 256;
        tab += nb;
        parts += nb + (nb + 63) / 64;
    }
}

// Lay out and upload the force pass for the current lanes (tables, replica table); drops the graphs,
// whose launches carry the old grids.
int apply_layout(sgk_batch *b) {
    const std::vector<int2> tab = force_layout(b);
    BHIP(b, hipMemcpyAsync(b->forceTab, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->dReps, b->reps.data(), sizeof(BatchRep) * b->nrep, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    for (auto &g : b->graphs)
        if (g.second.exec) BHIP(b, hipGraphExecDestroy(g.second.exec));
    b->graphs.clear();
    return SGK_OK;
}

}  // namespace

extern "C" {

const char *sgk_batch_last_error(const sgk_batch *b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int sgk_batch_destroy(sgk_batch *b) {
    if (!b) return SGK_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &g : b->graphs)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    for (void *q : b->allocs) (void)hipFree(q);
    if (b->hostState) (void)hipHostFree(b->hostState);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SGK_OK;
}

int sgk_batch_create(int nrep, const SimParams *params, const sgk_config *cfgIn, sgk_batch **out) {
    if (!out || !params || nrep <= 0) return bfail(nullptr, SGK_E_ARG, "need nrep >= 1 replicas");
    *out = nullptr;
    sgk_config cfg, def;
    sgk_default_config(&def);
    cfg = cfgIn ? *cfgIn : def;
    const SimParams &s0 = params[0];
    if (cfg.nl_capacity <= 0) cfg.nl_capacity = s0.iNebrMax > 0 ? s0.iNebrMax : 128;
    if (cfg.cell_capacity <= 0) cfg.cell_capacity = def.cell_capacity;
    if (cfg.lj_capacity <= 0) cfg.lj_capacity = def.lj_capacity;
    if (cfg.close_capacity <= 0) cfg.close_capacity = def.close_capacity;
    if (cfg.close_margin <= 0.f) cfg.close_margin = def.close_margin;
    if (cfg.records_capacity <= 0) cfg.records_capacity = def.records_capacity;
    if (cfg.block_size <= 0) cfg.block_size = def.block_size;
    if (cfg.block_size != 256)
        return bfail(nullptr, SGK_E_UNSUPPORTED, "block size must be 256 (the kernels are written for 4-wave workgroups)");
    if (cfg.nl_capacity > 255) return bfail(nullptr, SGK_E_UNSUPPORTED, "neighbour capacity above 255");
    if (cfg.cell_capacity > 64 || cfg.cell_capacity % 8)
        return bfail(nullptr, SGK_E_UNSUPPORTED, "cell capacity must be a multiple of 8 up to 64");
    if (cfg.close_capacity % 8) return bfail(nullptr, SGK_E_UNSUPPORTED, "close capacity must be a multiple of 8");
    long long sumMe = 0;
    for (int r = 0; r < nrep; ++r) {
        const SimParams &s = params[r];
        if (s.nMol <= 0 || s.nMolMe < 0 || s.nMolMe > s.nMol || s.nMol >= (1 << 24))
            return bfail(nullptr, SGK_E_ARG, "replica " + std::to_string(r) + ": invalid atom counts");
        if (s.iRegime != 0 && s.iRegime != 2)
            return bfail(nullptr, SGK_E_UNSUPPORTED, "batched replicas must be regime 0 or 2");
        if (s.cells.x <= 0 || s.cells.y <= 0 || s.cells.z <= 0)
            return bfail(nullptr, SGK_E_ARG, "replica " + std::to_string(r) + ": empty cell grid");
        if (s.sizeHistRdf <= 0 || s.sizeHistRdf > 8192) return bfail(nullptr, SGK_E_ARG, "RDF bins must be 1..8192");
        if (const char *m = mismatch(s0, s))
            return bfail(nullptr, SGK_E_ARG, "replica " + std::to_string(r) + " differs from replica 0 in " + m);
        sumMe += s.nMolMe;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(nullptr, SGK_E_HIP, "no HIP device");

    sgk_batch *b = new sgk_batch();
    b->nrep = nrep;
    b->sp.assign(params, params + nrep);
    b->cfg = cfg;
    b->mixed = cfg.precision == SG_PREC_MIXED;
    (void)hipGetDevice(&b->device);
    int rc;
#define BTRY(x)                               \
    do {                                      \
        if ((rc = (x)) != SGK_OK) {           \
            g_batch_create_error = b->err;    \
            sgk_batch_destroy(b);             \
            return rc;                        \
        }                                     \
    } while (0)
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return bfail(nullptr, SGK_E_HIP, "hipStreamCreate failed");
    }
    // lanes per metal atom: the configured count for every replica, else per replica by its size
    const int G = cfg.metal_lanes;
    const bool fixedG = G == 1 || G == 2 || G == 4 || G == 8 || G == 16;
    for (int r = 0; r < nrep; ++r) b->repLanes.push_back(fixedG ? G : default_lanes(params[r].nMolMe));
    (void)sumMe;
    // replica table and block tables; per-replica workgroup decomposition as in sgk_ctx
    std::vector<int2> atomTab, metalTab, nlTab, nlTabC;
    long long atomOff = 0, metalOff = 0, carbonOff = 0, cellOff = 0, ordOff = 0;
    for (int r = 0; r < nrep; ++r) {
        const SimParams &s = params[r];
        BatchRep R;
        std::memset(&R, 0, sizeof(R));
        sgk::make_dev_params(s, cfg, R.p);
        R.p.xcdMap = 0;
        R.p.keepFull = 0;  // the reference-layout list is a single-system test aid
        R.atomOff = atomOff;
        R.metalOff = metalOff;
        R.carbonOff = carbonOff;
        R.cellOff = cellOff;
        const int nbAtom = (s.nMol + 255) / 256;
        const int nbMetal = (s.nMolMe + 255) / 256;
        const int nbC = (R.p.nC + 255) / 256;
        for (int k = 0; k < nbAtom; ++k) atomTab.push_back(make_int2(r, k));
        for (int k = 0; k < nbMetal; ++k) metalTab.push_back(make_int2(r, k));
        // cell-ordered list build: kNlMetalLanes lanes per metal atom, then one lane per carbon;
        // every replica's metal workgroups (the long walks) are dispatched before any carbon one
        R.nbNlM = (int)(((long long)s.nMolMe * sgk::kNlMetalLanes + 255) / 256);
        for (int k = 0; k < R.nbNlM; ++k) nlTab.push_back(make_int2(r, k));
        for (int k = 0; k < nbC; ++k) nlTabC.push_back(make_int2(r, R.nbNlM + k));
        R.ordOff = (int)ordOff;
        ordOff += 2LL * R.p.ncell + 2;
        b->reps.push_back(R);
        atomOff += s.nMol;
        metalOff += s.nMolMe;
        carbonOff += R.p.nC;
        cellOff += R.p.ncell;
    }
    nlTab.insert(nlTab.end(), nlTabC.begin(), nlTabC.end());
    const std::vector<int2> forceTab = force_layout(b);
    size_t tabCap = 0, partCap = 0;
    worst_layout(b, tabCap, partCap);
    b->nAtoms = atomOff;
    b->nMetal = metalOff;
    b->nCarbon = carbonOff;
    b->nCells = cellOff;
    b->nAtomBlocks = (int)atomTab.size();
    b->nMetalBlocks = (int)metalTab.size();
    b->nNlBlocks = (int)nlTab.size();
    b->nOrdBuckets = ordOff;
    if (ordOff >= (1LL << 31)) {
        delete b;
        return bfail(nullptr, SGK_E_UNSUPPORTED, "too many cells over all replicas for the list-build order");
    }
    const int cap = cfg.cell_capacity;
    const size_t Rb = b->mixed ? 8 : 4;
    BTRY(balloc(b, &b->dReps, (size_t)nrep, false));
    BTRY(balloc(b, &b->atomTab, atomTab.size(), false));
    BTRY(balloc(b, &b->forceTab, tabCap, false));
    BTRY(balloc(b, &b->metalTab, metalTab.size(), false));
    if (cfg.nl_scheme == 3) {
        BTRY(balloc(b, &b->nlTab, nlTab.size(), false));
        BTRY(balloc(b, &b->order, (size_t)atomOff, false));
        BTRY(balloc(b, &b->ordCnt, (size_t)ordOff, false));
        BTRY(balloc(b, &b->ordKey, (size_t)atomOff, false));
        BTRY(balloc(b, &b->ordSlot, (size_t)atomOff, false));
        BTRY(balloc(b, &b->ordBlock, (size_t)(ordOff + sgk::kOrdChunk - 1) / sgk::kOrdChunk, false));
    }
    BTRY(balloc(b, (char **)&b->pos, (size_t)atomOff * (b->mixed ? 32 : 16), true));
    BTRY(balloc(b, (char **)&b->posAlt, (size_t)atomOff * (b->mixed ? 32 : 16), true));
    BTRY(balloc(b, (char **)&b->vel, (size_t)atomOff * 3 * Rb, true));
    BTRY(balloc(b, (char **)&b->acc, (size_t)atomOff * 3 * Rb, true));
    BTRY(balloc(b, &b->rho, (size_t)metalOff, true));
    BTRY(balloc(b, &b->embF, (size_t)metalOff, true));
    BTRY(balloc(b, &b->embDF, (size_t)metalOff, true));
    BTRY(balloc(b, &b->fric, (size_t)metalOff, true));
    for (int k = 0; k < 2; ++k) {
        BTRY(balloc(b, &b->cellOf[k], (size_t)atomOff, false));
        BTRY(balloc(b, &b->cnt[k], (size_t)cellOff, true));
        BTRY(balloc(b, &b->slots[k], (size_t)cellOff * cap, false));
        BTRY(hipMemsetAsync(b->cellOf[k], 0xff, 4 * (size_t)atomOff, b->stream) == hipSuccess
                 ? SGK_OK
                 : bfail(b, SGK_E_HIP, "hipMemsetAsync of the cell index"));
    }
    BTRY(balloc(b, &b->slotOf, (size_t)atomOff, false));
    BTRY(balloc(b, &b->cellPos, (size_t)cellOff * cap, false));
    BTRY(balloc(b, &b->nlM, (size_t)metalOff * cfg.nl_capacity, false));
    BTRY(balloc(b, &b->nnM, (size_t)metalOff, true));
    BTRY(balloc(b, &b->lj, (size_t)carbonOff * cfg.lj_capacity, false));
    BTRY(balloc(b, &b->nlj, (size_t)carbonOff, true));
    BTRY(balloc(b, &b->cl, (size_t)carbonOff * cfg.close_capacity, false));
    BTRY(balloc(b, &b->ncl, (size_t)carbonOff, true));
    BTRY(balloc(b, &b->bondC, (size_t)carbonOff, true));
    BTRY(balloc(b, &b->bondA, (size_t)carbonOff, true));
    BTRY(balloc(b, (char **)&b->partials, partCap * sgk::partial_bytes_for(b->mixed), true));
    BTRY(balloc(b, &b->state, (size_t)nrep, true));
    BTRY(balloc(b, &b->records, (size_t)nrep * cfg.records_capacity, true));
    BTRY(balloc(b, &b->hist, (size_t)nrep * s0.sizeHistRdf, true));
    BTRY(hipMemcpyAsync(b->dReps, b->reps.data(), sizeof(BatchRep) * nrep, hipMemcpyHostToDevice, b->stream) == hipSuccess
             ? SGK_OK : bfail(b, SGK_E_HIP, "copy of the replica table"));
    BTRY(hipMemcpyAsync(b->atomTab, atomTab.data(), sizeof(int2) * atomTab.size(), hipMemcpyHostToDevice, b->stream) == hipSuccess
             ? SGK_OK : bfail(b, SGK_E_HIP, "copy of the block table"));
    BTRY(hipMemcpyAsync(b->forceTab, forceTab.data(), sizeof(int2) * forceTab.size(), hipMemcpyHostToDevice, b->stream) == hipSuccess
             ? SGK_OK : bfail(b, SGK_E_HIP, "copy of the block table"));
    if (b->nlTab)
        BTRY(hipMemcpyAsync(b->nlTab, nlTab.data(), sizeof(int2) * nlTab.size(), hipMemcpyHostToDevice, b->stream) == hipSuccess
                 ? SGK_OK : bfail(b, SGK_E_HIP, "copy of the block table"));
    if (!metalTab.empty())
        BTRY(hipMemcpyAsync(b->metalTab, metalTab.data(), sizeof(int2) * metalTab.size(), hipMemcpyHostToDevice, b->stream) == hipSuccess
                 ? SGK_OK : bfail(b, SGK_E_HIP, "copy of the block table"));
    if (hipHostMalloc((void **)&b->hostState, sizeof(DevState) * nrep, 0) != hipSuccess) {
        g_batch_create_error = "hipHostMalloc failed";
        sgk_batch_destroy(b);
        return SGK_E_NOMEM;
    }
    std::memset(b->hostState, 0, sizeof(DevState) * nrep);
    for (int r = 0; r < nrep; ++r) b->hostState[r].nebrNow = 1;
    BTRY(hipMemcpyAsync(b->state, b->hostState, sizeof(DevState) * nrep, hipMemcpyHostToDevice, b->stream) == hipSuccess
             ? SGK_OK : bfail(b, SGK_E_HIP, "copy of the run state"));
    BTRY(hipEventCreate(&b->ev0) == hipSuccess ? SGK_OK : bfail(b, SGK_E_HIP, "hipEventCreate"));
    BTRY(hipEventCreate(&b->ev1) == hipSuccess ? SGK_OK : bfail(b, SGK_E_HIP, "hipEventCreate"));
    BTRY(hipStreamSynchronize(b->stream) == hipSuccess ? SGK_OK : bfail(b, SGK_E_HIP, "hipStreamSynchronize"));
#undef BTRY
    *out = b;
    return SGK_OK;
}

int sgk_batch_upload_f64(sgk_batch *b, int rep, const double *r, const double *v, const double *a) {
    if (!b || rep < 0 || rep >= b->nrep) return SGK_E_ARG;
    const BatchRep &R = b->reps[rep];
    const size_t n = (size_t)R.p.nMol;
    if (r) {
        if (b->mixed) {
            BHIP(b, hipMemcpy((double4 *)b->pos + R.atomOff, r, 32 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(4 * n);
            for (size_t i = 0; i < 4 * n; ++i) t[i] = (float)r[i];
            BHIP(b, hipMemcpy((float4 *)b->pos + R.atomOff, t.data(), 16 * n, hipMemcpyHostToDevice));
        }
    }
    const double *src[2] = {v, a};
    void *dst[2] = {b->vel, b->acc};
    for (int k = 0; k < 2; ++k) {
        if (!src[k]) continue;
        if (b->mixed) {
            std::vector<double> t(3 * n);
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) t[d * n + i] = src[k][3 * i + d];
            BHIP(b, hipMemcpy((double *)dst[k] + 3 * R.atomOff, t.data(), 8 * 3 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(3 * n);
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) t[d * n + i] = (float)src[k][3 * i + d];
            BHIP(b, hipMemcpy((float *)dst[k] + 3 * R.atomOff, t.data(), 4 * 3 * n, hipMemcpyHostToDevice));
        }
    }
    return SGK_OK;
}

int sgk_batch_download_f64(sgk_batch *b, int rep, double *r, double *v, double *a) {
    if (!b || rep < 0 || rep >= b->nrep) return SGK_E_ARG;
    BHIP(b, hipStreamSynchronize(b->stream));
    const BatchRep &R = b->reps[rep];
    const size_t n = (size_t)R.p.nMol;
    if (r) {
        if (b->mixed) {
            BHIP(b, hipMemcpy(r, (double4 *)b->pos + R.atomOff, 32 * n, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(4 * n);
            BHIP(b, hipMemcpy(t.data(), (float4 *)b->pos + R.atomOff, 16 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < 4 * n; ++i) r[i] = t[i];
        }
    }
    double *dstv[2] = {v, a};
    void *srcv[2] = {b->vel, b->acc};
    for (int k = 0; k < 2; ++k) {
        if (!dstv[k]) continue;
        if (b->mixed) {
            std::vector<double> t(3 * n);
            BHIP(b, hipMemcpy(t.data(), (double *)srcv[k] + 3 * R.atomOff, 8 * 3 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) dstv[k][3 * i + d] = t[d * n + i];
        } else {
            std::vector<float> t(3 * n);
            BHIP(b, hipMemcpy(t.data(), (float *)srcv[k] + 3 * R.atomOff, 4 * 3 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) dstv[k][3 * i + d] = t[d * n + i];
        }
    }
    return SGK_OK;
}

int sgk_batch_set_run_state(sgk_batch *b, int rep, int stepCount, int nebrNow, double dispHi) {
    if (!b || rep < 0 || rep >= b->nrep) return SGK_E_ARG;
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipMemcpy(b->hostState, b->state, sizeof(DevState) * b->nrep, hipMemcpyDeviceToHost));
    DevState &s = b->hostState[rep];
    s.step = stepCount;
    s.nebrNow = nebrNow ? 1 : 0;
    s.dispHi = dispHi;
    s.errors = 0;
    s.rec = 0;
    BHIP(b, hipMemcpy(b->state + rep, &s, sizeof(DevState), hipMemcpyHostToDevice));
    if (rep == 0) b->hostStep = stepCount;
    return SGK_OK;
}

int sgk_batch_run(sgk_batch *b, int nsteps, int rdfEvery, sgk_record *records) {
    if (!b) return SGK_E_ARG;
    if (nsteps <= 0) return SGK_OK;
    if (nsteps > b->cfg.records_capacity)
        return bfail(b, SGK_E_ARG, "nsteps exceeds records_capacity; call sgk_batch_run in chunks");
    const std::string flags = batch_flags(b, b->hostStep, nsteps, rdfEvery);
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    // the list-build order follows the cells of the current positions; refreshed now and then,
    // outside the graph (the captured launches read it through the same buffer)
    const bool hadOrder = b->orderValid;
    if (b->order && (!b->orderValid || b->orderAge >= sgk::kNlOrderSteps)) sgk::build_batch_nl_order(b);
    if (!hadOrder && b->orderValid) {  // graphs captured before the order existed read none
        for (auto &g : b->graphs)
            if (g.second.exec) BHIP(b, hipGraphExecDestroy(g.second.exec));
        b->graphs.clear();
    }
    if (b->cfg.use_graphs) {
        auto it = b->graphs.find(flags);
        if (it == b->graphs.end()) {
            if (b->graphs.size() >= 8) {
                auto victim = b->graphs.begin();
                for (auto g = b->graphs.begin(); g != b->graphs.end(); ++g)
                    if (g->second.uses < victim->second.uses) victim = g;
                (void)hipGraphExecDestroy(victim->second.exec);
                b->graphs.erase(victim);
            }
            hipGraph_t graph = nullptr;
            BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            enqueue_batch(b, flags);
            hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (e != hipSuccess) return bfail(b, SGK_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            sgk::GraphEntry ge;
            e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return bfail(b, SGK_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            it = b->graphs.emplace(flags, ge).first;
        }
        it->second.uses += 1;
        BHIP(b, hipGraphLaunch(it->second.exec, b->stream));
    } else {
        enqueue_batch(b, flags);
    }
    BHIP(b, hipGetLastError());
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    if (records)  // replica-major: records[r * nsteps + k]
        BHIP(b, hipMemcpy2DAsync(records, sizeof(sgk_record) * (size_t)nsteps, b->records,
                                 sizeof(sgk_record) * (size_t)b->cfg.records_capacity, sizeof(sgk_record) * (size_t)nsteps,
                                 (size_t)b->nrep, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->lastRunMs = ms;
    b->hostStep += nsteps;
    b->orderAge += nsteps;
    return check_batch_flags(b);
}

double sgk_batch_last_run_ms(const sgk_batch *b) { return b ? b->lastRunMs : 0.0; }

int sgk_batch_rdf_get(sgk_batch *b, int rep, unsigned int *hist) {
    if (!b || rep < 0 || rep >= b->nrep || !hist) return SGK_E_ARG;
    const size_t nb = (size_t)b->sp[0].sizeHistRdf;
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipMemcpy(hist, b->hist + rep * nb, 4 * nb, hipMemcpyDeviceToHost));
    return SGK_OK;
}

int sgk_batch_rdf_clear(sgk_batch *b) {
    if (!b) return SGK_E_ARG;
    BHIP(b, hipMemsetAsync(b->hist, 0, 4 * (size_t)b->sp[0].sizeHistRdf * b->nrep, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return SGK_OK;
}

long long sgk_batch_device_bytes(const sgk_batch *b) { return b ? b->bytes : 0; }

int sgk_batch_metal_lanes(const sgk_batch *b) { return b ? b->metalLanes : 0; }

int sgk_batch_replica_lanes(const sgk_batch *b, int rep) {
    return (b && rep >= 0 && rep < b->nrep) ? b->repLanes[rep] : 0;
}

int sgk_batch_set_lanes(sgk_batch *b, const int *lanes) {
    if (!b || !lanes) return SGK_E_ARG;
    for (int r = 0; r < b->nrep; ++r)
        if (lanes[r] != 1 && lanes[r] != 2 && lanes[r] != 4 && lanes[r] != 8 && lanes[r] != 16)
            return bfail(b, SGK_E_ARG, "lanes per metal atom must be 1, 2, 4, 8 or 16");
    BHIP(b, hipStreamSynchronize(b->stream));
    b->repLanes.assign(lanes, lanes + b->nrep);
    return apply_layout(b);
}

int sgk_batch_autotune(sgk_batch *b, int reps, sgk_batch_tune_entry *entries, int maxEntries, int *nEntries) {
    if (!b) return SGK_E_ARG;
    if (b->hostStep == 0) return bfail(b, SGK_E_STATE, "no neighbour lists yet: run at least one step first");
    if (reps <= 0) reps = 20;
    static const int kBounds[] = {0, 128, 512, 2048, 8192, 1 << 30};
    static const int kinds[] = {1, 2, 4, 8, 16};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    int n = 0, rc = SGK_OK;
    for (int k = 0; k + 1 < (int)(sizeof(kBounds) / sizeof(int)) && rc == SGK_OK; ++k) {
        std::vector<int> members;
        for (int r = 0; r < b->nrep; ++r)
            if (b->sp[r].nMolMe >= kBounds[k] && b->sp[r].nMolMe < kBounds[k + 1] && b->sp[r].nMolMe > 0)
                members.push_back(r);
        if (members.empty()) continue;
        int bestG = b->repLanes[members[0]];
        double best = 1e300;
        for (int G : kinds) {
            for (int r : members) b->repLanes[r] = G;
            if ((rc = apply_layout(b)) != SGK_OK) break;
            sgk::launch_batch_rho(b);  // warm-up
            sgk::launch_batch_force_eval(b);
            BHIP(b, hipEventRecord(e0, b->stream));
            for (int q = 0; q < reps; ++q) {
                sgk::launch_batch_rho(b);
                sgk::launch_batch_force_eval(b);
            }
            BHIP(b, hipEventRecord(e1, b->stream));
            BHIP(b, hipGetLastError());
            BHIP(b, hipEventSynchronize(e1));
            float ms = 0.f;
            BHIP(b, hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            if (entries && n < maxEntries) {
                long long atoms = 0;
                for (int r : members) atoms += b->sp[r].nMol;
                entries[n] = sgk_batch_tune_entry{kBounds[k], kBounds[k + 1], (int)members.size(), (int)atoms, G, us};
            }
            ++n;
            if (us < best) {
                best = us;
                bestG = G;
            }
        }
        for (int r : members) b->repLanes[r] = bestG;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SGK_OK) return rc;
    if ((rc = apply_layout(b)) != SGK_OK) return rc;
    // leave accelerations, energies and friction as the chosen lanes compute them
    sgk::launch_batch_rho(b);
    sgk::launch_batch_force_eval(b);
    BHIP(b, hipStreamSynchronize(b->stream));
    if (nEntries) *nEntries = n;
    return check_batch_flags(b);
}

int sgk_batch_atoms(const sgk_batch *b) { return b ? (int)b->nAtoms : 0; }

}  // extern "C"
