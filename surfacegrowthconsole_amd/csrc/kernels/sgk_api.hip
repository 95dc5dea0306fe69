// sgk_api.hip - the sgk_* C-ABI: engine context, state transfer, stages and the device-resident
// step loop (hipGraph-captured chunks of steps).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sgk_device.h"
#include "sgk_internal.h"

using sgk::DevParams;
using sgk::DevState;

namespace {

thread_local std::string g_create_error;

int fail(sgk_ctx *c, int code, const std::string &msg) {
    if (c)
        c->err = msg;
    else
        g_create_error = msg;
    return code;
}

int hip_fail(sgk_ctx *c, hipError_t e, const char *what) {
    return fail(c, SGK_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define SGK_HIP(c, expr)                                   \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(c, e_, #expr); \
    } while (0)

template <class T>
int dalloc(sgk_ctx *c, T **p, size_t count, bool zero = false) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) return fail(c, SGK_E_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed");
    c->bytes += (long long)bytes;
    if (zero) {
        e = hipMemsetAsync(*p, 0, bytes, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "hipMemsetAsync");
    }
    return SGK_OK;
}

// Fatal device flags: capacities exceeded or the carbon close lists no longer exact.
const int kFatal = SGK_FLAG_CELL_OVERFLOW | SGK_FLAG_NL_OVERFLOW | SGK_FLAG_LJ_OVERFLOW | SGK_FLAG_CLOSE_OVERFLOW |
                   SGK_FLAG_DISPLACEMENT;

std::string flag_text(int f) {
    std::string s;
    if (f & SGK_FLAG_CELL_OVERFLOW) s += " cell capacity exceeded;";
    if (f & SGK_FLAG_NL_OVERFLOW) s += " neighbour-list capacity (iNebrMax) exceeded;";
    if (f & SGK_FLAG_LJ_OVERFLOW) s += " carbon metal-neighbour capacity exceeded;";
    if (f & SGK_FLAG_CLOSE_OVERFLOW) s += " carbon close-neighbour capacity exceeded;";
    if (f & SGK_FLAG_DISPLACEMENT) s += " displacement since the last rebuild exceeds the skin bound;";
    return s;
}

int check_flags(sgk_ctx *c) {
    hipError_t e = hipMemcpyAsync(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpyAsync(state)");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamSynchronize");
    const int f = c->hostState->errors & kFatal;
    if (f) return fail(c, (f & SGK_FLAG_DISPLACEMENT) ? SGK_E_DISPLACEMENT : SGK_E_OVERFLOW, "device:" + flag_text(f));
    return SGK_OK;
}

// Per-kernel event timing (profiling mode only; never during graph capture).
struct PendingEvent {
    std::string name;
    hipEvent_t a, b;
};
thread_local std::vector<hipEvent_t> g_event_pool;
thread_local std::vector<PendingEvent> g_pending;
thread_local hipEvent_t g_open = nullptr;
thread_local bool g_capturing = false;

hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

void flush_profile(sgk_ctx *c) {
    if (g_pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto &p : g_pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        auto &s = c->stats[p.name];
        s.launches += 1;
        s.ms += ms;
        g_event_pool.push_back(p.a);
        g_event_pool.push_back(p.b);
    }
    g_pending.clear();
}

}  // namespace

namespace sgk {

void profile_begin(sgk_ctx *c) {
    if (!c->profiling || g_capturing) return;
    g_open = take_event();
    (void)hipEventRecord(g_open, c->stream);
}

void profile_end(sgk_ctx *c, const char *name) {
    if (!c->profiling || g_capturing || !g_open) return;
    hipEvent_t b = take_event();
    (void)hipEventRecord(b, c->stream);
    g_pending.push_back({name, g_open, b});
    g_open = nullptr;
}

}  // namespace sgk

void sgk::make_dev_params(const SimParams &s, const sgk_config &cfg, DevParams &p) {
    std::memset(&p, 0, sizeof(p));
    p.nMol = s.nMol;
    p.nMolMe = s.nMolMe;
    p.nC = s.nMol - s.nMolMe;
    p.iRegime = s.iRegime;
    p.Lx = s.region.x;
    p.Ly = s.region.y;
    p.Lz = s.region.z;
    p.cx = s.cells.x;
    p.cy = s.cells.y;
    p.cz = s.cells.z;
    p.ncell = s.cells.x * s.cells.y * s.cells.z;
    p.iwx = s.invWidth.x;
    p.iwy = s.invWidth.y;
    p.iwz = s.invWidth.z;
    p.rrNebr = s.rrNebr;
    p.rrCutEam = s.rrCutEam;
    p.rrCutLJ = s.rrCutLJ;
    p.rrCutC = s.rrCutC;
    const float margin = cfg.close_margin;
    const float closeR = s.rCutC + 2.f * (0.5f * s.rNebrShell + margin);
    p.closeR2 = closeR * closeR;
    p.dispBound = 0.5f * s.rNebrShell + margin;
    p.re = s.re;
    p.fe = s.fe;
    p.rhoe = s.rhoe;
    p.alpha = s.alpha;
    p.beta = s.beta;
    p.A = s.A;
    p.B = s.B;
    p.kappa = s.kappa;
    p.lambda = s.lambda;
    for (int k = 0; k < 4; ++k) {
        p.Fn[k] = s.Fn[k];
        p.F[k] = s.F[k];
    }
    p.eta = s.eta;
    p.Fe = s.Fe;
    p.massMe = s.massMe;
    p.rei = s.rei;
    p.rhoei = s.rhoei;
    p.mu_r = s.mu_r;
    p.theta_0 = s.theta_0;
    p.mu_theta = s.mu_theta;
    p.mu_p = s.mu_p;
    p.z_0 = s.z_0;
    p.epsLJ = s.epsLJ;
    p.sigma2 = s.sigmaLJ * s.sigmaLJ;
    p.inv_sigma2 = 1.f / p.sigma2;
    p.sigmaLJ = s.sigmaLJ;
    p.deltaT = s.deltaT;
    p.temperature = s.temperature;
    p.kB = s.kB;
    p.gamma = s.gammaBerendsen;
    p.rNebrShell = s.rNebrShell;
    p.stepThermostat = s.stepThermostat;
    p.stepEquil = s.stepEquil;
    p.stepCool = s.stepCool;
    // frozen perimeter bounds (SurfaceGrowth.cu:415-421); cos(M_PI/6.f) is evaluated in double
    const float cos30 = (float)std::cos(M_PI / 6.0);
    p.xLeft = 0.5f * (cos30 - s.region.x);
    p.xRight = 0.5f * (s.region.x - cos30);
    p.yBottom = 0.5f * (1.f - s.region.y);
    p.yTop = 0.5f * (s.region.y - 1.f);
    p.rangeRdf2 = s.rangeRdf * s.rangeRdf;
    p.intervalRdf = s.intervalRdf;
    p.sizeHistRdf = s.sizeHistRdf;
    p.deltaF = s.deltaF;
    p.velMagDepos = s.velMagDepos;
    p.stepDeposit = s.stepDeposit;
    p.nMolToDeposit = s.nMolToDeposit;
    p.cellCap = cfg.cell_capacity;
    p.nlCap = cfg.nl_capacity;
    p.ljCap = cfg.lj_capacity;
    p.clCap = cfg.close_capacity;
    p.keepFull = cfg.keep_full_nl;
    p.rebuildEvery = cfg.rebuild_every;
    p.xcdMap = cfg.xcd_map;
    p.binGroup = cfg.bin_group;
    p.nlScheme = cfg.nl_scheme;
}

int sgk::prepare_ctx(const SimParams *sp, const sgk_config *cfgIn, sgk_ctx *c, std::string &err) {
    sgk_config cfg;
    if (cfgIn)
        cfg = *cfgIn;
    else
        sgk_default_config(&cfg);
    sgk_config def;
    sgk_default_config(&def);
    if (cfg.nl_capacity <= 0) cfg.nl_capacity = sp->iNebrMax > 0 ? sp->iNebrMax : 128;
    if (cfg.cell_capacity <= 0) cfg.cell_capacity = def.cell_capacity;
    if (cfg.lj_capacity <= 0) cfg.lj_capacity = def.lj_capacity;
    if (cfg.close_capacity <= 0) cfg.close_capacity = def.close_capacity;
    if (cfg.close_margin <= 0.f) cfg.close_margin = def.close_margin;
    if (cfg.block_size <= 0) cfg.block_size = def.block_size;
    if (cfg.records_capacity <= 0) cfg.records_capacity = def.records_capacity;
    auto bad = [&err](int code, const char *msg) {
        err = msg;
        return code;
    };
    if (sp->nMol <= 0 || sp->nMolMe < 0 || sp->nMolMe > sp->nMol) return bad(SGK_E_ARG, "invalid atom counts");
    if (sp->nMol >= (1 << 24)) return bad(SGK_E_UNSUPPORTED, "more than 2^24 atoms");
    if (cfg.nl_capacity > 255) return bad(SGK_E_UNSUPPORTED, "neighbour capacity above 255");
    if (cfg.cell_capacity > 64 || cfg.cell_capacity % 8)
        return bad(SGK_E_UNSUPPORTED, "cell capacity must be a multiple of 8 up to 64 (one wave per cell)");
    if (cfg.close_capacity % 8) return bad(SGK_E_UNSUPPORTED, "close capacity must be a multiple of 8");
    if (cfg.block_size != 256)
        return bad(SGK_E_UNSUPPORTED, "block size must be 256 (the kernels are written for 4-wave workgroups)");
    if (sp->cells.x <= 0 || sp->cells.y <= 0 || sp->cells.z <= 0) return bad(SGK_E_ARG, "empty cell grid");
    if ((long long)sp->cells.x * sp->cells.y * sp->cells.z * cfg.cell_capacity >= (1ll << 31))
        return bad(SGK_E_UNSUPPORTED, "cell grid too large: cells x cell capacity must stay below 2^31");
    if (sp->iRegime < 0 || sp->iRegime > 2) return bad(SGK_E_ARG, "regime must be 0, 1 or 2");
    if (sp->sizeHistRdf <= 0 || sp->sizeHistRdf > 8192) return bad(SGK_E_ARG, "RDF bins must be 1..8192");
    c->sp = *sp;
    c->cfg = cfg;
    c->mixed = cfg.precision == SG_PREC_MIXED;
    sgk::make_dev_params(c->sp, c->cfg, c->dp);
    const DevParams &p = c->dp;
    // lanes per metal atom: enough waves to cover the machine for small clusters
    int G = cfg.metal_lanes;
    // measured on MI355X (sgk_autotune): 4 lanes for 100k metal atoms, 8 for 4k, 16 for small clusters
    if (G != 1 && G != 2 && G != 4 && G != 8 && G != 16) G = p.nMolMe >= 65536 ? 4 : p.nMolMe >= 2048 ? 8 : 16;
    c->metalLanes = G;
    c->block = 256;
    c->nbM = metal_blocks(p.nMolMe, G, p.xcdMap);
    c->nbC = (p.nC + 255) / 256;
    c->nbRho = c->nbM;
    c->nbProps = (p.nMol + 255) / 256;
    // partials for the largest force grid any launch shape gives (16 lanes per metal atom), then one
    // chunk total per 64 partials (two-stage step reduction)
    const size_t nbMmax = (size_t)((((long long)p.nMolMe * 16 + 255) / 256 + 7) / 8 * 8);
    const size_t nParts = std::max<size_t>(std::max<size_t>(nbMmax, (size_t)c->nbM) + (size_t)c->nbC,
                                           (size_t)c->nbProps);
    c->part2Off = (int)nParts;
    return SGK_OK;
}

extern "C" {

void sgk_default_config(sgk_config *cfg) {
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->precision = SG_PREC_FP32;
    cfg->cell_capacity = 64;
    cfg->lj_capacity = 64;
    cfg->close_capacity = 16;
    cfg->close_margin = 0.2f;
    cfg->block_size = 256;
    cfg->records_capacity = 1024;
    cfg->use_graphs = 1;
    cfg->pair_fp64 = 1;  // mixed form: FP64 pair arithmetic (measured cost nil, error < 1e-5)
    cfg->xcd_map = 2;    // density pass remapped (57.7 -> 54.2 us at 1.1 M atoms); the force pass and
                         // the list measured slower with it on MI355X
    cfg->bin_group = 1;
    cfg->nl_scheme = 3;  // 1.1 M atoms, thermalised: 0.35-0.37 ms per rebuild; per atom (0) 0.42 ms
}

int sgk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *sgk_last_error(const sgk_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int sgk_create(const SimParams *sp, const sgk_config *cfgIn, sgk_ctx **out) {
    if (!sp || !out) return fail(nullptr, SGK_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, SGK_E_HIP, "no HIP device");
    sgk_ctx *c = new sgk_ctx();
    {
        std::string msg;
        const int rc0 = sgk::prepare_ctx(sp, cfgIn, c, msg);
        if (rc0 != SGK_OK) {
            delete c;
            return fail(nullptr, rc0, msg);
        }
    }
    const sgk_config &cfg = c->cfg;
    (void)hipGetDevice(&c->device);
    int rc;
#define SGK_TRY(x)                   \
    do {                             \
        if ((rc = (x)) != SGK_OK) {  \
            g_create_error = c->err; \
            sgk_destroy(c);          \
            return rc;               \
        }                            \
    } while (0)
    {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
            delete c;
            return SGK_E_HIP;
        }
    }
    const DevParams &p = c->dp;
    const size_t R = c->mixed ? 8 : 4;
    const size_t nMol = p.nMol, nMe = std::max(p.nMolMe, 1), nC = std::max(p.nC, 1);
    SGK_TRY(dalloc(c, (char **)&c->pos, nMol * (c->mixed ? 32 : 16), true));
    SGK_TRY(dalloc(c, (char **)&c->posAlt, nMol * (c->mixed ? 32 : 16), true));
    SGK_TRY(dalloc(c, (char **)&c->vel, 3 * nMol * R, true));
    SGK_TRY(dalloc(c, (char **)&c->acc, 3 * nMol * R, true));
    SGK_TRY(dalloc(c, &c->rho, nMe, true));
    SGK_TRY(dalloc(c, &c->embF, nMe, true));
    SGK_TRY(dalloc(c, &c->embDF, nMe, true));
    SGK_TRY(dalloc(c, &c->fric, nMe, true));
    for (int b = 0; b < 2; ++b) {
        SGK_TRY(dalloc(c, &c->cellOf[b], nMol, false));
        SGK_TRY(dalloc(c, &c->cnt[b], (size_t)p.ncell, true));
        SGK_TRY(dalloc(c, &c->slots[b], (size_t)p.ncell * p.cellCap, false));
    }
    SGK_TRY(dalloc(c, &c->slotOf, nMol, false));
    SGK_TRY(dalloc(c, &c->cellPos, (size_t)p.ncell * p.cellCap, false));
    for (int b = 0; b < 2; ++b)  // -1: no cell yet (k_sortcells clears the outgoing buffer through it)
        SGK_TRY(hipMemsetAsync(c->cellOf[b], 0xff, 4 * nMol, c->stream) == hipSuccess
                    ? SGK_OK
                    : fail(c, SGK_E_HIP, "hipMemsetAsync of the cell index"));
    SGK_TRY(dalloc(c, &c->nlM, nMe * (size_t)p.nlCap, false));
    SGK_TRY(dalloc(c, &c->nnM, nMe, true));
    SGK_TRY(dalloc(c, &c->lj, nC * (size_t)p.ljCap, false));
    SGK_TRY(dalloc(c, &c->nlj, nC, true));
    SGK_TRY(dalloc(c, &c->cl, nC * (size_t)p.clCap, false));
    SGK_TRY(dalloc(c, &c->ncl, nC, true));
    SGK_TRY(dalloc(c, &c->bondC, nC, true));
    SGK_TRY(dalloc(c, &c->bondA, nC, true));
    if (p.nlScheme == 3 && nMol > 0) {
        const size_t nb = 2 * (size_t)p.ncell + 2;  // carbon and metal buckets of the list-build order
        SGK_TRY(dalloc(c, &c->nlOrder, nMol, false));
        SGK_TRY(dalloc(c, &c->ordCnt, nb, false));
        SGK_TRY(dalloc(c, &c->ordKey, nMol, false));
        SGK_TRY(dalloc(c, &c->ordSlot, nMol, false));
        SGK_TRY(dalloc(c, &c->ordBlock, (nb + sgk::kOrdChunk - 1) / sgk::kOrdChunk, false));
    }
    if (p.keepFull) {
        SGK_TRY(dalloc(c, &c->fullNN, nMol, true));
        SGK_TRY(dalloc(c, &c->fullNBL, nMol * (size_t)p.nlCap, false));
    }
    const size_t nParts = (size_t)c->part2Off;
    SGK_TRY(dalloc(c, (char **)&c->partials, (nParts + nParts / 64 + 2) * sgk::partial_bytes(c), true));
    if (cfg.rdf_complete && p.nMolMe > 0) {
        c->rdfGrid = sgk::rdf_grid(p, cfg.rdf_complete == 1 ? 2 : cfg.rdf_complete);
        if (c->rdfGrid.ncell == 0) {
            g_create_error = "rdf_complete needs rangeRdf below half of every box edge";
            sgk_destroy(c);
            return SGK_E_UNSUPPORTED;
        }
        const size_t nb = (size_t)c->rdfGrid.ncell + 1;
        SGK_TRY(dalloc(c, &c->rdfCnt, nb, false));
        SGK_TRY(dalloc(c, &c->rdfBlock, (nb + sgk::kOrdChunk - 1) / sgk::kOrdChunk, false));
        SGK_TRY(dalloc(c, &c->rdfKey, nMe, false));
        SGK_TRY(dalloc(c, &c->rdfSlot, nMe, false));
        SGK_TRY(dalloc(c, &c->rdfPos, nMe, false));
    }
    SGK_TRY(dalloc(c, &c->state, 1, true));
    SGK_TRY(dalloc(c, &c->records, (size_t)cfg.records_capacity, true));
    SGK_TRY(dalloc(c, &c->hist, (size_t)p.sizeHistRdf, true));
    if (hipHostMalloc((void **)&c->hostState, sizeof(DevState), 0) != hipSuccess) {
        g_create_error = "hipHostMalloc failed";
        sgk_destroy(c);
        return SGK_E_NOMEM;
    }
    std::memset(c->hostState, 0, sizeof(DevState));
    c->hostState->nebrNow = 1;
    c->hostState->step = sp->stepCount;
    c->hostState->nMolDeposited = sp->nMolDeposited;
    if (hipMemcpyAsync(c->state, c->hostState, sizeof(DevState), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        g_create_error = "state initialisation failed";
        sgk_destroy(c);
        return SGK_E_HIP;
    }
    c->hostStep = sp->stepCount;
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
#undef SGK_TRY
    *out = c;
    return SGK_OK;
}

int sgk_destroy(sgk_ctx *c) {
    if (!c) return SGK_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->adj) {
        sgk::adjoint_free(c->adj);
        delete c->adj;
    }
    for (auto &g : c->graphs)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    void *ptrs[] = {c->pos, c->posAlt, c->vel, c->acc, c->rho, c->embF, c->embDF, c->fric, c->cellOf[0], c->cellOf[1],
                    c->cnt[0], c->cnt[1], c->slots[0], c->slots[1], c->slotOf, c->cellPos, c->nlM, c->nnM, c->lj, c->nlj, c->cl,
                    c->ncl, c->bondC, c->bondA, c->fullNN, c->fullNBL, c->partials, c->state, c->records, c->hist,
                    c->nlOrder, c->ordCnt, c->ordKey, c->ordSlot, c->ordBlock, c->rdfCnt, c->rdfBlock, c->rdfKey,
                    c->rdfSlot, c->rdfPos};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (c->diffMem) (void)hipFree(c->diffMem);
    if (c->hostState) (void)hipHostFree(c->hostState);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SGK_OK;
}

// ---- state transfer --------------------------------------------------------------------------------
int sgk_upload(sgk_ctx *c, const sg_float4 *r, const sg_float3 *v, const sg_float3 *a) {
    if (!c) return SGK_E_ARG;
    const size_t n = (size_t)c->dp.nMol;
    if (r) {
        if (c->mixed) {
            std::vector<double> t(4 * n);
            for (size_t i = 0; i < n; ++i) {
                t[4 * i] = r[i].x;
                t[4 * i + 1] = r[i].y;
                t[4 * i + 2] = r[i].z;
                t[4 * i + 3] = r[i].w;
            }
            SGK_HIP(c, hipMemcpy(c->pos, t.data(), 32 * n, hipMemcpyHostToDevice));
        } else {
            SGK_HIP(c, hipMemcpy(c->pos, r, 16 * n, hipMemcpyHostToDevice));
        }
        c->haveList = false;
        c->haveCells = false;
        c->orderValid = false;
    }
    if (v) {
        if (c->mixed) {
            std::vector<double> t(3 * n);
            for (size_t i = 0; i < n; ++i) {
                t[i] = v[i].x;
                t[n + i] = v[i].y;
                t[2 * n + i] = v[i].z;
            }
            SGK_HIP(c, hipMemcpy(c->vel, t.data(), 8 * 3 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(3 * n);
            for (size_t i = 0; i < n; ++i) {
                t[i] = v[i].x;
                t[n + i] = v[i].y;
                t[2 * n + i] = v[i].z;
            }
            SGK_HIP(c, hipMemcpy(c->vel, t.data(), 4 * 3 * n, hipMemcpyHostToDevice));
        }
    }
    if (a) {
        if (c->mixed) {
            std::vector<double> t(3 * n);
            for (size_t i = 0; i < n; ++i) {
                t[i] = a[i].x;
                t[n + i] = a[i].y;
                t[2 * n + i] = a[i].z;
            }
            SGK_HIP(c, hipMemcpy(c->acc, t.data(), 8 * 3 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(3 * n);
            for (size_t i = 0; i < n; ++i) {
                t[i] = a[i].x;
                t[n + i] = a[i].y;
                t[2 * n + i] = a[i].z;
            }
            SGK_HIP(c, hipMemcpy(c->acc, t.data(), 4 * 3 * n, hipMemcpyHostToDevice));
        }
    }
    return SGK_OK;
}

int sgk_upload_f64(sgk_ctx *c, const double *r, const double *v, const double *a) {
    if (!c) return SGK_E_ARG;
    const size_t n = (size_t)c->dp.nMol;
    if (r) {
        if (c->mixed) {
            SGK_HIP(c, hipMemcpy(c->pos, r, 32 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(4 * n);
            for (size_t i = 0; i < 4 * n; ++i) t[i] = (float)r[i];
            SGK_HIP(c, hipMemcpy(c->pos, t.data(), 16 * n, hipMemcpyHostToDevice));
        }
        c->haveList = false;
        c->haveCells = false;
        c->orderValid = false;
    }
    if (v) {
        if (c->mixed) {
            std::vector<double> t(3 * n);
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) t[d * n + i] = v[3 * i + d];
            SGK_HIP(c, hipMemcpy(c->vel, t.data(), 8 * 3 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(3 * n);
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) t[d * n + i] = (float)v[3 * i + d];
            SGK_HIP(c, hipMemcpy(c->vel, t.data(), 4 * 3 * n, hipMemcpyHostToDevice));
        }
    }
    if (a) {
        if (c->mixed) {
            std::vector<double> t(3 * n);
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) t[d * n + i] = a[3 * i + d];
            SGK_HIP(c, hipMemcpy(c->acc, t.data(), 8 * 3 * n, hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(3 * n);
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) t[d * n + i] = (float)a[3 * i + d];
            SGK_HIP(c, hipMemcpy(c->acc, t.data(), 4 * 3 * n, hipMemcpyHostToDevice));
        }
    }
    return SGK_OK;
}

int sgk_download_f64(sgk_ctx *c, double *r, double *v, double *a) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->dp.nMol;
    if (r) {
        if (c->mixed) {
            SGK_HIP(c, hipMemcpy(r, c->pos, 32 * n, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(4 * n);
            SGK_HIP(c, hipMemcpy(t.data(), c->pos, 16 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < 4 * n; ++i) r[i] = t[i];
        }
    }
    if (v) {
        if (c->mixed) {
            std::vector<double> t(3 * n);
            SGK_HIP(c, hipMemcpy(t.data(), c->vel, 8 * 3 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) v[3 * i + d] = t[d * n + i];
        } else {
            std::vector<float> t(3 * n);
            SGK_HIP(c, hipMemcpy(t.data(), c->vel, 4 * 3 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) v[3 * i + d] = t[d * n + i];
        }
    }
    if (a) {
        if (c->mixed) {
            std::vector<double> t(3 * n);
            SGK_HIP(c, hipMemcpy(t.data(), c->acc, 8 * 3 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) a[3 * i + d] = t[d * n + i];
        } else {
            std::vector<float> t(3 * n);
            SGK_HIP(c, hipMemcpy(t.data(), c->acc, 4 * 3 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) a[3 * i + d] = t[d * n + i];
        }
    }
    return SGK_OK;
}

int sgk_download(sgk_ctx *c, sg_float4 *r, sg_float3 *v, sg_float3 *a) {
    if (!c) return SGK_E_ARG;
    const size_t n = (size_t)c->dp.nMol;
    std::vector<double> R(r ? 4 * n : 0), V(v ? 3 * n : 0), A(a ? 3 * n : 0);
    int rc = sgk_download_f64(c, r ? R.data() : nullptr, v ? V.data() : nullptr, a ? A.data() : nullptr);
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        if (r) r[i] = {(float)R[4 * i], (float)R[4 * i + 1], (float)R[4 * i + 2], (float)R[4 * i + 3]};
        if (v) v[i] = {(float)V[3 * i], (float)V[3 * i + 1], (float)V[3 * i + 2]};
        if (a) a[i] = {(float)A[3 * i], (float)A[3 * i + 1], (float)A[3 * i + 2]};
    }
    return SGK_OK;
}

int sgk_set_run_state(sgk_ctx *c, int stepCount, int nebrNow, double dispHi) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    SGK_HIP(c, hipMemcpy(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost));
    c->hostState->step = stepCount;
    c->hostState->nebrNow = nebrNow;
    c->hostState->dispHi = dispHi;
    SGK_HIP(c, hipMemcpy(c->state, c->hostState, sizeof(DevState), hipMemcpyHostToDevice));
    c->hostStep = stepCount;
    return SGK_OK;
}

int sgk_set_shear_state(sgk_ctx *c, double shear, double comX, double cmVelX) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    SGK_HIP(c, hipMemcpy(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost));
    c->hostState->shear = shear;
    c->hostState->comX = comX;
    c->hostState->cmVelX = cmVelX;
    SGK_HIP(c, hipMemcpy(c->state, c->hostState, sizeof(DevState), hipMemcpyHostToDevice));
    return SGK_OK;
}

// ---- stages ------------------------------------------------------------------------------------------
int sgk_bin(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    sgk::launch_bin(c, true);
    sgk::launch_sortcells(c, true);
    sgk::launch_commit_rebuild(c);
    SGK_HIP(c, hipGetLastError());
    c->haveCells = true;
    return check_flags(c);
}

int sgk_get_cells(sgk_ctx *c, int *cellOf, int *cellCount, int *cellSlots) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    SGK_HIP(c, hipMemcpy(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost));
    const int b = c->hostState->cellbuf;
    if (cellOf) SGK_HIP(c, hipMemcpy(cellOf, c->cellOf[b], 4 * (size_t)c->dp.nMol, hipMemcpyDeviceToHost));
    if (cellCount) SGK_HIP(c, hipMemcpy(cellCount, c->cnt[b], 4 * (size_t)c->dp.ncell, hipMemcpyDeviceToHost));
    if (cellSlots) {
        std::vector<int> cnt(c->dp.ncell);
        SGK_HIP(c, hipMemcpy(cnt.data(), c->cnt[b], 4 * (size_t)c->dp.ncell, hipMemcpyDeviceToHost));
        SGK_HIP(c, hipMemcpy(cellSlots, c->slots[b], 4 * (size_t)c->dp.ncell * c->dp.cellCap, hipMemcpyDeviceToHost));
        for (int q = 0; q < c->dp.ncell; ++q)  // slots past the count are undefined: fill -1
            for (int s = std::min(cnt[q], c->dp.cellCap); s < c->dp.cellCap; ++s) cellSlots[(size_t)q * c->dp.cellCap + s] = -1;
    }
    return SGK_OK;
}

int sgk_build_nebr(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    if (!c->haveCells) return fail(c, SGK_E_STATE, "no cells for the current positions: call sgk_bin first");
    sgk::launch_nl(c, true);
    SGK_HIP(c, hipGetLastError());
    int rc = check_flags(c);
    if (rc == SGK_OK) c->haveList = true;
    return rc;
}

int sgk_build_nl_order(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    sgk::build_nl_order(c);
    SGK_HIP(c, hipGetLastError());
    return SGK_OK;
}

int sgk_get_nebr(sgk_ctx *c, int *NN, int *NBL) {
    if (!c) return SGK_E_ARG;
    if (!c->dp.keepFull) return fail(c, SGK_E_STATE, "keep_full_nl was not set at creation");
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->dp.nMol;
    std::vector<int> nn(n);
    SGK_HIP(c, hipMemcpy(nn.data(), c->fullNN, 4 * n, hipMemcpyDeviceToHost));
    if (NN) std::memcpy(NN, nn.data(), 4 * n);
    if (NBL) {
        SGK_HIP(c, hipMemcpy(NBL, c->fullNBL, 4 * n * c->dp.nlCap, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i)
            for (int j = nn[i]; j < c->dp.nlCap; ++j) NBL[(size_t)j * n + i] = -1;
    }
    return SGK_OK;
}

int sgk_get_carbon_lists(sgk_ctx *c, int *nLj, int *ljIdx, int *ljPos, int *nClose, int *closeIdx, int *closePos) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    const size_t nC = (size_t)c->dp.nC;
    if (nC == 0) return SGK_OK;
    std::vector<int> nl(nC), nc(nC);
    std::vector<unsigned> lj(nC * c->dp.ljCap), cl(nC * c->dp.clCap);
    SGK_HIP(c, hipMemcpy(nl.data(), c->nlj, 4 * nC, hipMemcpyDeviceToHost));
    SGK_HIP(c, hipMemcpy(nc.data(), c->ncl, 4 * nC, hipMemcpyDeviceToHost));
    SGK_HIP(c, hipMemcpy(lj.data(), c->lj, 4 * lj.size(), hipMemcpyDeviceToHost));
    SGK_HIP(c, hipMemcpy(cl.data(), c->cl, 4 * cl.size(), hipMemcpyDeviceToHost));
    for (size_t ic = 0; ic < nC; ++ic) {
        if (nLj) nLj[ic] = nl[ic];
        if (nClose) nClose[ic] = nc[ic];
        for (int q = 0; q < c->dp.ljCap; ++q) {
            const unsigned e = q < nl[ic] ? lj[(size_t)q * nC + ic] : 0xffffffffu;
            if (ljIdx) ljIdx[ic * c->dp.ljCap + q] = q < nl[ic] ? (int)(e >> 8) : -1;
            if (ljPos) ljPos[ic * c->dp.ljCap + q] = q < nl[ic] ? (int)(e & 255u) : -1;
        }
        for (int q = 0; q < c->dp.clCap; ++q) {
            const unsigned e = q < nc[ic] ? cl[ic * (size_t)c->dp.clCap + q] : 0xffffffffu;
            if (closeIdx) closeIdx[ic * c->dp.clCap + q] = q < nc[ic] ? (int)(e >> 8) : -1;
            if (closePos) closePos[ic * c->dp.clCap + q] = q < nc[ic] ? (int)(e & 255u) : -1;
        }
    }
    return SGK_OK;
}

int sgk_compute_rho(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    if (!c->haveList) return fail(c, SGK_E_STATE, "no neighbour list: call sgk_bin and sgk_build_nebr first");
    sgk::launch_rho(c);
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_compute_forces(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    if (!c->haveList) return fail(c, SGK_E_STATE, "no neighbour list: call sgk_bin and sgk_build_nebr first");
    sgk::launch_rho(c);
    sgk::launch_force(c, false, false);
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_get_rho(sgk_ctx *c, double *rho) {
    if (!c || !rho) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> t(std::max(c->dp.nMolMe, 1));
    SGK_HIP(c, hipMemcpy(t.data(), c->rho, 4 * (size_t)c->dp.nMolMe, hipMemcpyDeviceToHost));
    for (int i = 0; i < c->dp.nMolMe; ++i) rho[i] = t[i];
    return SGK_OK;
}

int sgk_get_forces(sgk_ctx *c, double *acc, double *energy, double *fric) {
    if (!c) return SGK_E_ARG;
    const size_t n = (size_t)c->dp.nMol;
    if (acc) {
        int rc = sgk_download_f64(c, nullptr, nullptr, acc);
        if (rc) return rc;
    }
    if (energy) {
        std::vector<double> r(4 * n);
        int rc = sgk_download_f64(c, r.data(), nullptr, nullptr);
        if (rc) return rc;
        for (size_t i = 0; i < n; ++i) energy[i] = r[4 * i + 3];
    }
    if (fric && c->dp.nMolMe) {
        std::vector<float> t(c->dp.nMolMe);
        SGK_HIP(c, hipMemcpy(t.data(), c->fric, 4 * (size_t)c->dp.nMolMe, hipMemcpyDeviceToHost));
        for (int i = 0; i < c->dp.nMolMe; ++i) fric[i] = t[i];
    }
    return SGK_OK;
}

int sgk_leapfrog(sgk_ctx *c, int part) {
    if (!c) return SGK_E_ARG;
    if (part == 1) {
        sgk::launch_lf1(c, false);
        c->haveCells = false;
    } else
        sgk::launch_lf2(c);
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_thermostat(sgk_ctx *c, int stepCount, double vvSum) {
    if (!c) return SGK_E_ARG;
    sgk::launch_thermostat(c, stepCount, vvSum, 1);
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_eval_props(sgk_ctx *c, sgk_record *rec) {
    if (!c || !rec) return SGK_E_ARG;
    sgk::launch_props(c);
    sgk::launch_finalize(c, 1);
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipMemcpyAsync(rec, c->records, sizeof(sgk_record), hipMemcpyDeviceToHost, c->stream));
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_rdf_clear(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipMemsetAsync(c->hist, 0, 4 * (size_t)c->dp.sizeHistRdf, c->stream));
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_rdf_accumulate(sgk_ctx *c) {
    if (!c) return SGK_E_ARG;
    sgk::launch_rdf(c);
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_diffusion_enable(sgk_ctx *c, int maxTables) {
    if (!c) return SGK_E_ARG;
    const SimParams &s = c->sp;
    if (maxTables <= 0) return fail(c, SGK_E_ARG, "maxTables must be positive");
    if (s.nValDiffuse <= 0 || s.nBuffDiffuse <= 0 || s.stepDiffuse <= 0 || s.limitDiffuseAv <= 0)
        return fail(c, SGK_E_ARG, "diffusion sizes (nValDiffuse, nBuffDiffuse, stepDiffuse, limitDiffuseAv) must be positive");
    if (s.nBuffDiffuse > sgk::kMaxDiffuseBuffers) return fail(c, SGK_E_UNSUPPORTED, "more than 256 diffusion buffers");
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    if (c->diffMem) {
        SGK_HIP(c, hipFree(c->diffMem));
        c->diffMem = nullptr;
        c->diff = sgk::DiffusionArgs{};
    }
    const size_t nb = s.nBuffDiffuse, nv = s.nValDiffuse;
    const size_t ints = (nb + 2 + 1) / 2 * 2;  // counts and the two state words, padded to 8 bytes
    const size_t doubles = 6 * nb + nb * nv + nv + (size_t)maxTables * nv;
    const size_t bytes = 4 * ints + 8 * doubles;
    SGK_HIP(c, hipMalloc(&c->diffMem, bytes));
    c->bytes += (long long)bytes;
    sgk::DiffusionArgs &a = c->diff;
    a.count = (int *)c->diffMem;
    a.state = a.count + nb;
    double *d = (double *)((char *)c->diffMem + 4 * ints);
    a.org = d;
    a.tru = d + 3 * nb;
    a.rr = d + 6 * nb;
    a.av = a.rr + nb * nv;
    a.out = a.av + nv;
    a.nVal = (int)nv;
    a.nBuff = (int)nb;
    a.stepDiffuse = s.stepDiffuse;
    a.limitAv = s.limitDiffuseAv;
    a.cap = maxTables;
    a.deltaT = (double)s.deltaT;
    a.L[0] = (double)s.region.x;
    a.L[1] = (double)s.region.y;
    a.L[2] = (double)s.region.z;
    sgk::launch_diffuse_reset(c);
    sgk::drop_graphs(c);  // captured steps carry the sampling launches (or their absence)
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

int sgk_diffusion_get(sgk_ctx *c, double *tables, int maxTables, int *nTables) {
    if (!c) return SGK_E_ARG;
    if (!c->diff.count) return fail(c, SGK_E_STATE, "diffusion is not enabled: call sgk_diffusion_enable");
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    int st[2] = {0, 0};
    SGK_HIP(c, hipMemcpy(st, c->diff.state, sizeof(st), hipMemcpyDeviceToHost));
    if (nTables) *nTables = st[1];
    if (tables && maxTables > 0) {
        const int keep = std::min(std::min(st[1], maxTables), c->diff.cap);
        const size_t nv = (size_t)c->diff.nVal;
        for (int k = 0; k < keep; ++k) {
            const int idx = (st[1] - keep + k) % c->diff.cap;
            SGK_HIP(c, hipMemcpy(tables + (size_t)k * nv, c->diff.out + (size_t)idx * nv, 8 * nv, hipMemcpyDeviceToHost));
        }
    }
    return SGK_OK;
}

int sgk_rdf_get(sgk_ctx *c, unsigned int *hist) {
    if (!c || !hist) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    SGK_HIP(c, hipMemcpy(hist, c->hist, 4 * (size_t)c->dp.sizeHistRdf, hipMemcpyDeviceToHost));
    return SGK_OK;
}

}  // extern "C"

// ---- device-resident loop -----------------------------------------------------------------------------
// Per-step kernel-sequence flags known on the host: thermostat launch, shear launch, RDF launch,
// adhesion-energy record (adjoint windows).
std::string sgk::step_flag_string(sgk_ctx *c, int step0, int nsteps, int rdfEvery, bool observe) {
    const SimParams &s = c->sp;
    std::string f(nsteps, 'a');
    for (int k = 0; k < nsteps; ++k) {
        const int step = step0 + k + 1;
        int b = 0;
        const bool thermo = s.stepThermostat > 0 && step % s.stepThermostat == 0 &&
                            ((s.iRegime == 2 && step > s.stepEquil) || s.iRegime == 1);
        // with deltaF = 0 the shear stays 0 and ApplyShearK adds nothing: no launch needed
        const bool shear = s.iRegime == 2 && s.nMolMe > 0 && step > s.stepEquil + s.stepCool && s.deltaF != 0.f;
        const bool rdf = rdfEvery > 0 && step % rdfEvery == 0;
        // diffusion sample (SurfaceGrowth.cu:2315-2322), when the device buffers are enabled
        const bool msd = c->diff.count && s.iRegime == 2 && s.nMolMe > 0 && step > s.stepEquil &&
                         (step - s.stepEquil) % s.stepDiffuse == 0;
        b = (thermo ? 1 : 0) | (shear ? 2 : 0) | (rdf ? 4 : 0) | (observe ? 8 : 0) | (msd ? 16 : 0);
        f[k] = (char)('a' + b);
    }
    return f;
}

static void enqueue_steps(sgk_ctx *c, const std::string &flags) {
    sgk::launch_reset_records(c);
    // The force pass of a step also applies the next step's first half-kick when nothing touches the
    // velocities or accelerations in between (no thermostat, shear or deposition) and the step is
    // not the chunk's last (the host then reads full-step velocities): the next k_lf1 only drifts.
    // Such a force pass can also do that drift itself ("fused"): it writes the next positions into
    // posAlt, since the other atoms of the same pass still read the current ones, and the next step
    // starts with the binning alone (k_bin, which returns at once unless the lists are rebuilt).
    // The host swaps pos and posAlt after each fused step, so every launch here takes the current
    // buffer.  An even number of fused steps per chunk (the first pre-kicked step stays unfused
    // when the count is odd) returns the positions to the buffer the chunk started from: a
    // captured graph then starts and ends on the same buffer, and code outside the step loop sees
    // one position buffer.
    const size_t n = flags.size();
    std::vector<char> pre(n, 0);
    int nPre = 0;
    for (size_t k = 0; k < n; ++k) {
        const int b = flags[k] - 'a';
        pre[k] = c->dp.iRegime != 1 && !(b & 3) && k + 1 < n;
        nPre += pre[k];
    }
    bool unfuseFirst = (nPre & 1) != 0;
    bool kicked = false, drifted = false;
    for (size_t k = 0; k < n; ++k) {
        const int b = flags[k] - 'a';
        bool fuse = false;
        if (pre[k] && c->posAlt) {
            fuse = !unfuseFirst;
            unfuseFirst = false;
        }
        sgk::launch_insert(c);  // regime 1 only
        if (drifted)
            sgk::launch_bin(c, false);
        else
            sgk::launch_lf1(c, true, kicked);
        sgk::launch_sortcells(c, false);
        sgk::launch_nl(c, false);
        sgk::launch_rho(c);
        sgk::launch_force(c, true, true, pre[k], fuse ? c->posAlt : nullptr);
        kicked = pre[k];
        drifted = fuse;
        if (b & 1) sgk::launch_thermostat(c, 0, 0.0, 0);
        if (b & 2) sgk::launch_shear(c);
        // the samples below read this step's positions, still in pos
        if (b & 4) sgk::launch_rdf(c);
        if (b & 16) sgk::launch_diffuse(c);
        if (b & 8) sgk::launch_adh_obs(c, c->adj->d.part, c->adj->d.adh, 1);
        if (fuse) std::swap(c->pos, c->posAlt);
    }
}

int sgk::set_error(sgk_ctx *c, int code, const std::string &msg) { return fail(c, code, msg); }
int sgk::check_device_flags(sgk_ctx *c) { return check_flags(c); }

extern "C" {

int sgk_run(sgk_ctx *c, int nsteps, int rdfEvery, sgk_record *records) {
    return sgk::run_steps(c, nsteps, rdfEvery, records, nullptr);
}

}  // extern "C"

// The instantiated graph of a step pattern: from the cache, or captured now.  The cache holds at most
// 8 patterns per context; the least used is destroyed first (the stream has been synchronised
// since its last launch, by the run that launched it).
static int step_graph(sgk_ctx *c, const std::string &flags, hipGraphExec_t *out) {
    auto it = c->graphs.find(flags);
    if (it == c->graphs.end()) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        SGK_HIP(c, hipStreamIsCapturing(c->stream, &cs));
        if (cs != hipStreamCaptureStatusNone) return fail(c, SGK_E_STATE, "the engine stream is already being captured");
        if (c->graphs.size() >= 8) {
            auto victim = c->graphs.begin();
            for (auto g = c->graphs.begin(); g != c->graphs.end(); ++g)
                if (g->second.uses < victim->second.uses) victim = g;
            SGK_HIP(c, hipGraphExecDestroy(victim->second.exec));
            c->graphs.erase(victim);
        }
        hipGraph_t graph = nullptr;
        SGK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        g_capturing = true;
        enqueue_steps(c, flags);
        g_capturing = false;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        sgk::GraphEntry ge;
        e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
        it = c->graphs.emplace(flags, ge).first;
    }
    it->second.uses += 1;
    *out = it->second.exec;
    return SGK_OK;
}

extern "C" int sgk_prepare_run(sgk_ctx *c, int nsteps, int rdfEvery) {
    if (!c) return SGK_E_ARG;
    if (nsteps <= 0 || nsteps > c->cfg.records_capacity) return fail(c, SGK_E_ARG, "nsteps must be in 1..records_capacity");
    if (!c->cfg.use_graphs || c->profiling) return SGK_OK;
    hipGraphExec_t exec = nullptr;
    const int rc = step_graph(c, sgk::step_flag_string(c, c->hostStep, nsteps, rdfEvery, false), &exec);
    if (rc == SGK_OK) c->graphs[sgk::step_flag_string(c, c->hostStep, nsteps, rdfEvery, false)].uses -= 1;
    return rc;
}

int sgk::run_steps(sgk_ctx *c, int nsteps, int rdfEvery, sgk_record *records, double *adh) {
    if (!c) return SGK_E_ARG;
    if (nsteps <= 0) return SGK_OK;
    if (nsteps > c->cfg.records_capacity)
        return fail(c, SGK_E_ARG, "nsteps exceeds records_capacity; call sgk_run in chunks");
    if (adh && (!c->adj || c->adj->every < nsteps)) return fail(c, SGK_E_STATE, "no adhesion-energy buffer for the run");
    const std::string flags = sgk::step_flag_string(c, c->hostStep, nsteps, rdfEvery, adh != nullptr);
    if (!c->haveList) {  // positions uploaded since the last list, or none built: build it at the first step
        SGK_HIP(c, hipStreamSynchronize(c->stream));
        SGK_HIP(c, hipMemcpy(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost));
        if (!c->hostState->nebrNow) {
            c->hostState->nebrNow = 1;
            SGK_HIP(c, hipMemcpy(c->state, c->hostState, sizeof(DevState), hipMemcpyHostToDevice));
        }
    }
    SGK_HIP(c, hipEventRecord(c->ev0, c->stream));
    // the list-build order of nl_scheme 3 follows the cells of the current positions; it decides
    // only which lane builds which atom's lists, so it is refreshed now and then, outside the graph
    if (c->nlOrder && (!c->orderValid || c->orderAge >= sgk::kNlOrderSteps)) sgk::build_nl_order(c);
    if (c->cfg.use_graphs && !c->profiling) {
        hipGraphExec_t exec = nullptr;
        const int rc = step_graph(c, flags, &exec);
        if (rc != SGK_OK) return rc;
        SGK_HIP(c, hipGraphLaunch(exec, c->stream));
    } else {
        enqueue_steps(c, flags);
    }
    SGK_HIP(c, hipGetLastError());
    SGK_HIP(c, hipEventRecord(c->ev1, c->stream));
    if (records)
        SGK_HIP(c, hipMemcpyAsync(records, c->records, sizeof(sgk_record) * (size_t)nsteps, hipMemcpyDeviceToHost,
                                  c->stream));
    if (adh)
        SGK_HIP(c, hipMemcpyAsync(adh, c->adj->d.adh, sizeof(double) * (size_t)nsteps, hipMemcpyDeviceToHost, c->stream));
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->lastRunMs = ms;
    if (c->profiling) flush_profile(c);
    c->hostStep += nsteps;
    c->orderAge += nsteps;
    c->haveList = true;
    c->haveCells = false;  // cells are those of the last rebuild, not of the current positions
    return check_flags(c);
}

extern "C" {

double sgk_last_run_ms(sgk_ctx *c) { return c ? c->lastRunMs : 0.0; }

int sgk_set_profiling(sgk_ctx *c, int on) {
    if (!c) return SGK_E_ARG;
    c->profiling = on != 0;
    if (on) c->stats.clear();
    return SGK_OK;
}

int sgk_kernel_stats(sgk_ctx *c, sgk_kernel_stat *out, int max) {
    if (!c) return SGK_E_ARG;
    flush_profile(c);
    int n = 0;
    for (auto &kv : c->stats) {
        if (out && n < max) {
            std::memset(&out[n], 0, sizeof(sgk_kernel_stat));
            std::strncpy(out[n].name, kv.first.c_str(), sizeof(out[n].name) - 1);
            out[n].launches = kv.second.launches;
            out[n].ms = kv.second.ms;
        }
        ++n;
    }
    return n;
}

long long sgk_device_bytes(sgk_ctx *c) { return c ? c->bytes : 0; }

int sgk_struct_sizes(int *out, int max) {
    const int s[] = {(int)sizeof(sgk_config), (int)sizeof(sgk_record), (int)sizeof(sgk_kernel_stat),
                     (int)sizeof(sgk_adjoint_info)};
    int n = 0;
    for (; n < max && n < 4; ++n) out[n] = s[n];
    return n;
}

int sgk_set_deposited(sgk_ctx *c, int nMolDeposited) {
    if (!c) return SGK_E_ARG;
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    SGK_HIP(c, hipMemcpy(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost));
    c->hostState->nMolDeposited = nMolDeposited;
    SGK_HIP(c, hipMemcpy(c->state, c->hostState, sizeof(DevState), hipMemcpyHostToDevice));
    return SGK_OK;
}

static void set_shape(sgk_ctx *c, int G, int xcdMap) {
    c->metalLanes = G;
    c->dp.xcdMap = xcdMap;
    c->nbM = metal_blocks(c->dp.nMolMe, G, xcdMap);
    c->nbRho = c->nbM;
}

}  // extern "C"

void sgk::drop_graphs(sgk_ctx *c) {
    for (auto &g : c->graphs)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    c->graphs.clear();
}

extern "C" {

int sgk_autotune(sgk_ctx *c, int reps, sgk_tune_entry *entries, int maxEntries, int *nEntries) {
    if (!c) return SGK_E_ARG;
    if (!c->haveList) return fail(c, SGK_E_STATE, "no neighbour list: call sgk_bin and sgk_build_nebr (or sgk_run) first");
    if (reps <= 0) reps = 20;
    const bool prof = c->profiling;
    c->profiling = false;
    const int lanes[] = {1, 2, 4, 8, 16};
    const int base = c->dp.xcdMap & ~6;  // the density and force bits are tuned, the list bit is kept
    const int maps[] = {base | 2, base, base | 6, base | 4};
    int bestG = c->metalLanes, bestMap = c->dp.xcdMap, n = 0;
    double best = 1e300;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SGK_HIP(c, hipEventCreate(&e0));
    SGK_HIP(c, hipEventCreate(&e1));
    for (int G : lanes) {
        if (c->dp.nMolMe == 0 && G != lanes[0]) break;  // no metal pass to shape
        for (int m : maps) {
            set_shape(c, G, m);
            sgk::launch_rho(c);  // warm-up
            sgk::launch_force(c, false, false);
            SGK_HIP(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; ++r) {
                sgk::launch_rho(c);
                sgk::launch_force(c, false, false);
            }
            SGK_HIP(c, hipEventRecord(e1, c->stream));
            SGK_HIP(c, hipGetLastError());
            SGK_HIP(c, hipEventSynchronize(e1));
            float ms = 0.f;
            SGK_HIP(c, hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            if (entries && n < maxEntries) entries[n] = sgk_tune_entry{G, m, us};
            ++n;
            if (us < best) {
                best = us;
                bestG = G;
                bestMap = m;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_shape(c, bestG, bestMap);
    c->cfg.metal_lanes = bestG;
    c->cfg.xcd_map = bestMap;
    // leave acceleration, energy and friction as the chosen shape computes them
    sgk::launch_rho(c);
    sgk::launch_force(c, false, false);
    SGK_HIP(c, hipStreamSynchronize(c->stream));
    sgk::drop_graphs(c);  // captured launches carry the old grid and parameters
    c->profiling = prof;
    if (nEntries) *nEntries = n;
    return check_flags(c);
}

int sgk_metal_lanes(const sgk_ctx *c) { return c ? c->metalLanes : 0; }

}  // extern "C"
