// sgk_adjoint_api.hip - the adjoint C-ABI: a trajectory window run forward with checkpoints, the
// reverse sweep with recomputation, and the force-pass vector-Jacobian product on its own.
//
// Forward (sgk_adjoint_forward): a(r0) from a neighbour list built at r0, then the window's steps
// through the ordinary device loop (hipGraph chunks of `every` steps, the adhesion energy of each
// step recorded beside it).  Before each chunk a checkpoint keeps positions, velocities,
// accelerations, the neighbour structures and the run state; after the last, the final state
// including the cell buffers, so the context can be left exactly where the forward ended.
// Backward (sgk_adjoint_backward): for each segment, last to first, restore its checkpoint,
// recompute its steps eagerly (same kernels, so bitwise the stored trajectory - checked against
// the next checkpoint), keeping the positions after each step, the velocities before each
// thermostat and a copy of every neighbour list rebuilt on the way; then run the step transposes
// in reverse (sgk_adjoint.hip).  Memory: segments+1 checkpoints plus one segment of steps, instead
// of the whole window.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "sgk_adjoint.h"
#include "sgk_device.h"
#include "sgk_internal.h"

using sgk::Adjoint;
using sgk::AdjDev;
using sgk::Checkpoint;
using sgk::DevState;
using sgk::NlSnap;

namespace {

#define ADJ_HIP(c, expr)                                                                                   \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return sgk::set_error(c, SGK_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define ADJ_TRY(x)                 \
    do {                           \
        int rc_ = (x);             \
        if (rc_ != SGK_OK) return rc_; \
    } while (0)

template <class T>
int aalloc(sgk_ctx *c, Adjoint *a, T **p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (hipMalloc((void **)p, bytes) != hipSuccess) {
        *p = nullptr;
        return sgk::set_error(c, SGK_E_NOMEM, "adjoint: hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    a->allocs.push_back(*p);
    a->bytes += (long long)bytes;
    return SGK_OK;
}

struct Sizes {
    size_t N, nMe, nC, nlM, lj, cl, full;
};
Sizes sizes(const sgk_ctx *c) {
    const sgk::DevParams &p = c->dp;
    Sizes s;
    s.N = (size_t)p.nMol;
    s.nMe = (size_t)std::max(p.nMolMe, 1);
    s.nC = (size_t)std::max(p.nC, 1);
    s.nlM = s.nMe * p.nlCap;
    s.lj = s.nC * p.ljCap;
    s.cl = s.nC * p.clCap;
    s.full = p.keepFull ? s.N * p.nlCap : 0;
    return s;
}

int alloc_nl(sgk_ctx *c, Adjoint *a, NlSnap &n) {
    const Sizes z = sizes(c);
    ADJ_TRY(aalloc(c, a, &n.nlM, z.nlM));
    ADJ_TRY(aalloc(c, a, &n.nnM, z.nMe));
    ADJ_TRY(aalloc(c, a, &n.lj, z.lj));
    ADJ_TRY(aalloc(c, a, &n.nlj, z.nC));
    ADJ_TRY(aalloc(c, a, &n.cl, z.cl));
    ADJ_TRY(aalloc(c, a, &n.ncl, z.nC));
    if (z.full) {
        ADJ_TRY(aalloc(c, a, &n.fullNN, z.N));
        ADJ_TRY(aalloc(c, a, &n.fullNBL, z.full));
    }
    return SGK_OK;
}

// Copy the neighbour structures between the live context and a snapshot (toLive selects the way).
int copy_nl(sgk_ctx *c, const NlSnap &n, bool toLive) {
    const Sizes z = sizes(c);
    struct Item {
        void *live, *snap;
        size_t bytes;
    } items[] = {{c->nlM, n.nlM, 4 * z.nlM}, {c->nnM, n.nnM, 4 * z.nMe}, {c->lj, n.lj, 4 * z.lj},
                 {c->nlj, n.nlj, 4 * z.nC},  {c->cl, n.cl, 4 * z.cl},    {c->ncl, n.ncl, 4 * z.nC},
                 {c->fullNN, n.fullNN, 4 * (z.full ? z.N : 0)}, {c->fullNBL, n.fullNBL, 4 * z.full}};
    for (const Item &it : items) {
        if (!it.bytes) continue;
        ADJ_HIP(c, hipMemcpyAsync(toLive ? it.live : it.snap, toLive ? it.snap : it.live, it.bytes,
                                  hipMemcpyDeviceToDevice, c->stream));
    }
    return SGK_OK;
}

int alloc_vectors(sgk_ctx *c, Adjoint *a) {
    const Sizes z = sizes(c);
    AdjDev &d = a->d;
    for (double **v : {&d.Rb, &d.Vb, &d.Ab, &d.W, &d.Lam, &d.G}) ADJ_TRY(aalloc(c, a, v, 3 * z.N));
    for (double **v : {&d.rhoBar, &d.rho, &d.dF, &d.ddF}) ADJ_TRY(aalloc(c, a, v, z.nMe));
    ADJ_TRY(aalloc(c, a, &d.pbar, 4));
    ADJ_TRY(aalloc(c, a, &d.therm, 8));
    ADJ_TRY(aalloc(c, a, &d.part, 4 * ((z.N + 255) / 256) + 8));
    ADJ_TRY(aalloc(c, a, &d.mismatch, 1));
    ADJ_TRY(aalloc(c, a, &d.adh, (size_t)std::max(a->every, 1)));
    return SGK_OK;
}

// Adjoint storage of the context for a window shape (nsteps = 0: vector-Jacobian buffers only).
int ensure(sgk_ctx *c, int nsteps, int every) {
    Adjoint *a = c->adj;
    if (a && (nsteps == 0 ? !a->allocs.empty() : (a->nsteps == nsteps && a->every == every && !a->ck.empty())))
        return SGK_OK;
    if (a) {
        sgk::drop_graphs(c);  // graphs of adjoint windows record into the buffers freed here
        sgk::adjoint_free(a);
        delete a;
    }
    a = c->adj = new Adjoint();
    a->nsteps = nsteps;
    a->every = every;
    ADJ_HIP(c, hipEventCreate(&a->ev0));
    ADJ_HIP(c, hipEventCreate(&a->ev1));
    ADJ_TRY(alloc_vectors(c, a));
    if (nsteps == 0) return SGK_OK;
    const Sizes z = sizes(c);
    const sgk::DevParams &p = c->dp;
    a->segments = (nsteps + every - 1) / every;
    a->ck.resize(a->segments + 1);
    for (Checkpoint &k : a->ck) {
        ADJ_TRY(aalloc(c, a, (char **)&k.pos, 32 * z.N));
        ADJ_TRY(aalloc(c, a, (char **)&k.vel, 24 * z.N));
        ADJ_TRY(aalloc(c, a, (char **)&k.acc, 24 * z.N));
        ADJ_TRY(alloc_nl(c, a, k.nl));
    }
    Checkpoint &last = a->ck.back();  // the final state keeps the cell buffers too
    for (int b = 0; b < 2; ++b) {
        ADJ_TRY(aalloc(c, a, &last.cnt[b], (size_t)p.ncell));
        ADJ_TRY(aalloc(c, a, &last.slots[b], (size_t)p.ncell * p.cellCap));
        ADJ_TRY(aalloc(c, a, &last.cellOf[b], z.N));
    }
    ADJ_TRY(aalloc(c, a, &last.slotOf, z.N));
    ADJ_TRY(aalloc(c, a, &last.cellPos, (size_t)p.ncell * p.cellCap));
    ADJ_TRY(aalloc(c, a, (char **)&a->posStore, 32 * z.N * every));
    ADJ_TRY(aalloc(c, a, (char **)&a->velStore, 24 * z.N * every));
    return SGK_OK;
}

int cell_copy(sgk_ctx *c, const Checkpoint &k, bool toLive) {
    const sgk::DevParams &p = c->dp;
    const size_t N = (size_t)p.nMol, cells = (size_t)p.ncell, slots = cells * p.cellCap;
    struct Item {
        void *live, *snap;
        size_t bytes;
    } items[] = {{c->cnt[0], k.cnt[0], 4 * cells},      {c->cnt[1], k.cnt[1], 4 * cells},
                 {c->slots[0], k.slots[0], 4 * slots},  {c->slots[1], k.slots[1], 4 * slots},
                 {c->cellOf[0], k.cellOf[0], 4 * N},    {c->cellOf[1], k.cellOf[1], 4 * N},
                 {c->slotOf, k.slotOf, 4 * N},          {c->cellPos, k.cellPos, 16 * slots}};
    for (const Item &it : items)
        ADJ_HIP(c, hipMemcpyAsync(toLive ? it.live : it.snap, toLive ? it.snap : it.live, it.bytes,
                                  hipMemcpyDeviceToDevice, c->stream));
    return SGK_OK;
}

int save(sgk_ctx *c, Checkpoint &k) {
    const size_t N = (size_t)c->dp.nMol;
    ADJ_HIP(c, hipMemcpyAsync(k.pos, c->pos, 32 * N, hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(k.vel, c->vel, 24 * N, hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(k.acc, c->acc, 24 * N, hipMemcpyDeviceToDevice, c->stream));
    ADJ_TRY(copy_nl(c, k.nl, false));
    if (k.cellPos) ADJ_TRY(cell_copy(c, k, false));
    ADJ_HIP(c, hipMemcpyAsync(&k.st, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

// Restore a checkpoint into the live context.  A checkpoint without cell buffers keeps the live
// cell buffers and the index of the current one: the next rebuild bins from scratch into the
// buffer the protocol has cleared, so any consistent cell state serves.
int restore(sgk_ctx *c, const Checkpoint &k) {
    const size_t N = (size_t)c->dp.nMol;
    ADJ_HIP(c, hipMemcpyAsync(c->pos, k.pos, 32 * N, hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(c->vel, k.vel, 24 * N, hipMemcpyDeviceToDevice, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(c->acc, k.acc, 24 * N, hipMemcpyDeviceToDevice, c->stream));
    ADJ_TRY(copy_nl(c, k.nl, true));
    ADJ_HIP(c, hipMemcpyAsync(c->hostState, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    DevState s = k.st;
    if (k.cellPos) {
        ADJ_TRY(cell_copy(c, k, true));
    } else {
        s.cellbuf = c->hostState->cellbuf;
        s.rebuilds = c->hostState->rebuilds;
    }
    *c->hostState = s;
    ADJ_HIP(c, hipMemcpyAsync(c->state, c->hostState, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    c->hostStep = s.step;
    c->haveList = true;
    c->haveCells = false;
    return SGK_OK;
}

int check_window(sgk_ctx *c, int nsteps, int every) {
    const sgk::DevParams &p = c->dp;
    if (!c->mixed)
        return sgk::set_error(c, SGK_E_UNSUPPORTED, "adjoint: needs a mixed-precision context (FP64 state)");
    if (nsteps <= 0 || every <= 0) return sgk::set_error(c, SGK_E_ARG, "adjoint: nsteps and checkpoint_every must be > 0");
    if (every > c->cfg.records_capacity)
        return sgk::set_error(c, SGK_E_ARG, "adjoint: checkpoint_every exceeds records_capacity");
    if (p.iRegime == 1) return sgk::set_error(c, SGK_E_UNSUPPORTED, "adjoint: deposition (regime 1) is not differentiated");
    if (p.iRegime == 2 && p.nMolMe && p.deltaF != 0.f && c->hostStep + nsteps > p.stepEquil + p.stepCool)
        return sgk::set_error(c, SGK_E_UNSUPPORTED, "adjoint: the window reaches the shear phase");
    return SGK_OK;
}

// Host AoS <-> device SoA for 3-vectors.
int upload_vec(sgk_ctx *c, double *dev, const double *host) {
    const size_t N = (size_t)c->dp.nMol;
    std::vector<double> t(3 * N);
    for (size_t i = 0; i < N; ++i)
        for (int x = 0; x < 3; ++x) t[x * N + i] = host[3 * i + x];
    ADJ_HIP(c, hipMemcpy(dev, t.data(), 8 * 3 * N, hipMemcpyHostToDevice));
    return SGK_OK;
}
int download_vec(sgk_ctx *c, std::vector<double> &t, const double *dev) {
    const size_t N = (size_t)c->dp.nMol;
    t.resize(3 * N);
    ADJ_HIP(c, hipMemcpy(t.data(), dev, 8 * 3 * N, hipMemcpyDeviceToHost));
    return SGK_OK;
}

// Force-pass transpose at the live positions and lists, cotangent already in d.Lam.
void enqueue_vjp(sgk_ctx *c, const AdjDev &d, double wE) {
    sgk::launch_rho(c);  // the bond and near lists of the current positions
    sgk::launch_adj_rho(c, d);
    sgk::launch_adj_force(c, d, wE);
    sgk::launch_adj_rho2(c, d);
}

}  // namespace

void sgk::adjoint_free(Adjoint *a) {
    if (!a) return;
    for (void *q : a->allocs) (void)hipFree(q);
    a->allocs.clear();
    if (a->ev0) (void)hipEventDestroy(a->ev0);
    if (a->ev1) (void)hipEventDestroy(a->ev1);
    a->ev0 = a->ev1 = nullptr;
    a->ck.clear();
    a->slots.clear();
    a->bytes = 0;
    a->ready = false;
}

extern "C" {

int sgk_adjoint_forward(sgk_ctx *c, int nsteps, int every, double *obs, sgk_record *records) {
    if (!c) return SGK_E_ARG;
    ADJ_TRY(check_window(c, nsteps, every));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    ADJ_TRY(ensure(c, nsteps, every));
    Adjoint *a = c->adj;
    a->ready = false;
    a->step0 = c->hostStep;
    ADJ_HIP(c, hipEventRecord(a->ev0, c->stream));
    // a(r0) from a list built at r0
    sgk::launch_bin(c, true);
    sgk::launch_sortcells(c, true);
    sgk::launch_commit_rebuild(c);
    sgk::launch_nl(c, true);
    sgk::launch_rho(c);
    sgk::launch_force(c, false, false);
    ADJ_HIP(c, hipGetLastError());
    ADJ_TRY(sgk::check_device_flags(c));
    c->haveCells = true;
    c->haveList = true;
    ADJ_TRY(sgk_set_run_state(c, c->hostStep, 1, 0.0));  // the window opens with a rebuild
    std::vector<sgk_record> recs(nsteps);
    std::vector<double> adh(nsteps);
    for (int s = 0; s < a->segments; ++s) {
        ADJ_TRY(save(c, a->ck[s]));
        const int n = std::min(every, nsteps - s * every);
        ADJ_TRY(sgk::run_steps(c, n, 0, recs.data() + (size_t)s * every, adh.data() + (size_t)s * every));
    }
    ADJ_TRY(save(c, a->ck.back()));
    ADJ_HIP(c, hipEventRecord(a->ev1, c->stream));
    ADJ_HIP(c, hipEventSynchronize(a->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a->ev0, a->ev1);
    a->forwardMs = ms;
    const std::string flags = sgk::step_flag_string(c, a->step0, nsteps, 0, false);
    a->rebuilt.assign(nsteps, 0);
    a->thermo.assign(nsteps, '0');
    double e = 0.0;
    for (int k = 0; k < nsteps; ++k) {
        a->rebuilt[k] = recs[k].rebuilt;
        a->thermo[k] = ((flags[k] - 'a') & 1) ? '1' : '0';
        e += adh[k];
    }
    a->obs[0] = e / nsteps;
    a->obs[1] = c->dp.nMolMe ? recs.back().comSum[2] / c->dp.nMolMe - (double)c->dp.z_0 : 0.0;
    if (obs) {
        obs[0] = a->obs[0];
        obs[1] = a->obs[1];
    }
    if (records) std::memcpy(records, recs.data(), sizeof(sgk_record) * (size_t)nsteps);
    a->ready = true;
    return SGK_OK;
}

int sgk_adjoint_backward(sgk_ctx *c, const double *w, double *grad, double *dr0, double *dv0) {
    if (!c || !w || !grad) return SGK_E_ARG;
    Adjoint *a = c->adj;
    if (!a || !a->ready) return sgk::set_error(c, SGK_E_STATE, "adjoint: no forward window: call sgk_adjoint_forward first");
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    ADJ_HIP(c, hipEventRecord(a->ev0, c->stream));
    const AdjDev &d = a->d;
    const size_t N = (size_t)c->dp.nMol;
    const double wE = w[0] / a->nsteps;
    sgk::launch_adj_seed(c, d, w[1]);
    ADJ_HIP(c, hipMemsetAsync(d.mismatch, 0, sizeof(int), c->stream));
    for (int s = a->segments - 1; s >= 0; --s) {
        ADJ_TRY(restore(c, a->ck[s]));
        const int base = s * a->every, n = std::min(a->every, a->nsteps - base);
        // recompute the segment: the step loop's kernels, one step at a time
        sgk::launch_reset_records(c);
        std::vector<int> ver(n);
        int nver = 0;
        for (int k = 0; k < n; ++k) {
            sgk::launch_lf1(c, true);
            sgk::launch_sortcells(c, false);
            sgk::launch_nl(c, false);
            sgk::launch_rho(c);
            sgk::launch_force(c, true, true);
            ADJ_HIP(c, hipMemcpyAsync((char *)a->posStore + 32 * N * k, c->pos, 32 * N, hipMemcpyDeviceToDevice, c->stream));
            ADJ_HIP(c, hipMemcpyAsync((char *)a->velStore + 24 * N * k, c->vel, 24 * N, hipMemcpyDeviceToDevice, c->stream));
            if (a->thermo[base + k] == '1') sgk::launch_thermostat(c, 0, 0.0, 0);
            if (a->rebuilt[base + k]) {
                if (nver >= (int)a->slots.size()) {
                    a->slots.emplace_back();
                    ADJ_TRY(alloc_nl(c, a, a->slots.back()));
                }
                ADJ_TRY(copy_nl(c, a->slots[nver], false));
                ver[k] = ++nver;
            } else {
                ver[k] = k ? ver[k - 1] : 0;
            }
        }
        ADJ_HIP(c, hipGetLastError());
        const Checkpoint &next = a->ck[s + 1];
        sgk::launch_adj_cmp(c, d, next.pos, next.vel, next.acc);
        // the step transposes in reverse
        int cur = n ? ver[n - 1] : 0;
        for (int k = n - 1; k >= 0; --k) {
            if (ver[k] != cur) {
                ADJ_TRY(copy_nl(c, ver[k] == 0 ? a->ck[s].nl : a->slots[ver[k] - 1], true));
                cur = ver[k];
            }
            ADJ_HIP(c, hipMemcpyAsync(c->pos, (char *)a->posStore + 32 * N * k, 32 * N, hipMemcpyDeviceToDevice, c->stream));
            const double *vp = a->thermo[base + k] == '1' ? (const double *)((char *)a->velStore + 24 * N * k) : nullptr;
            if (vp) sgk::launch_adj_therm(c, d, vp, a->step0 + base + k + 1);
            sgk::launch_adj_pre(c, d, vp);
            enqueue_vjp(c, d, wE);
            sgk::launch_adj_post(c, d);
        }
        ADJ_HIP(c, hipGetLastError());
    }
    // the initial acceleration a_0 = a(r_0) with the list built at r_0
    ADJ_TRY(restore(c, a->ck[0]));
    sgk::launch_adj_lam0(c, d);
    enqueue_vjp(c, d, 0.0);
    ADJ_HIP(c, hipGetLastError());
    ADJ_HIP(c, hipMemcpyAsync(grad, d.pbar, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ADJ_HIP(c, hipMemcpyAsync(&a->mismatches, d.mismatch, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    if (dv0 || dr0) {
        std::vector<double> V, R, G;
        if (dv0) {
            ADJ_TRY(download_vec(c, V, d.Vb));
            for (size_t i = 0; i < N; ++i)
                for (int x = 0; x < 3; ++x) dv0[3 * i + x] = V[x * N + i];
        }
        if (dr0) {
            ADJ_TRY(download_vec(c, R, d.Rb));
            ADJ_TRY(download_vec(c, G, d.G));
            for (size_t i = 0; i < N; ++i)
                for (int x = 0; x < 3; ++x) dr0[3 * i + x] = R[x * N + i] + G[x * N + i];
        }
    }
    ADJ_TRY(restore(c, a->ck.back()));  // leave the context where the forward ended
    ADJ_HIP(c, hipEventRecord(a->ev1, c->stream));
    ADJ_HIP(c, hipEventSynchronize(a->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a->ev0, a->ev1);
    a->backwardMs = ms;
    return SGK_OK;
}

int sgk_adjoint_get_info(sgk_ctx *c, sgk_adjoint_info *out) {
    if (!c || !out) return SGK_E_ARG;
    std::memset(out, 0, sizeof(*out));
    const Adjoint *a = c->adj;
    if (!a) return SGK_OK;
    out->nsteps = a->nsteps;
    out->checkpoint_every = a->every;
    out->segments = a->segments;
    out->list_copies = (int)a->slots.size();
    out->mismatches = a->mismatches;
    out->bytes = a->bytes;
    out->forward_ms = a->forwardMs;
    out->backward_ms = a->backwardMs;
    out->obs[0] = a->obs[0];
    out->obs[1] = a->obs[1];
    return SGK_OK;
}

int sgk_force_vjp(sgk_ctx *c, const double *abar, double wAdh, double *rbar, double *pbar) {
    if (!c || !abar) return SGK_E_ARG;
    if (!c->mixed)
        return sgk::set_error(c, SGK_E_UNSUPPORTED, "force VJP: needs a mixed-precision context (FP64 state)");
    if (!c->haveList) return sgk::set_error(c, SGK_E_STATE, "no neighbour list: call sgk_bin and sgk_build_nebr first");
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    ADJ_TRY(ensure(c, 0, 0));
    const AdjDev &d = c->adj->d;
    ADJ_TRY(upload_vec(c, d.Lam, abar));
    const size_t N = (size_t)c->dp.nMol;
    ADJ_HIP(c, hipMemsetAsync(d.G, 0, 8 * 3 * N, c->stream));
    ADJ_HIP(c, hipMemsetAsync(d.rhoBar, 0, 8 * (size_t)std::max(c->dp.nMolMe, 1), c->stream));
    ADJ_HIP(c, hipMemsetAsync(d.pbar, 0, 4 * sizeof(double), c->stream));
    enqueue_vjp(c, d, wAdh);
    ADJ_HIP(c, hipGetLastError());
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    if (rbar) {
        std::vector<double> G;
        ADJ_TRY(download_vec(c, G, d.G));
        for (size_t i = 0; i < N; ++i)
            for (int x = 0; x < 3; ++x) rbar[3 * i + x] = G[x * N + i];
    }
    if (pbar) {
        ADJ_HIP(c, hipMemcpyAsync(pbar, d.pbar, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        ADJ_HIP(c, hipStreamSynchronize(c->stream));
    }
    return SGK_OK;
}

int sgk_adhesion_energy(sgk_ctx *c, double *e) {
    if (!c || !e) return SGK_E_ARG;
    if (!c->mixed) return sgk::set_error(c, SGK_E_UNSUPPORTED, "adhesion energy: needs a mixed-precision context");
    if (!c->haveList) return sgk::set_error(c, SGK_E_STATE, "no neighbour list: call sgk_bin and sgk_build_nebr first");
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    ADJ_TRY(ensure(c, 0, 0));
    sgk::launch_adh_obs(c, c->adj->d.part, c->adj->d.adh, 0);
    ADJ_HIP(c, hipGetLastError());
    ADJ_HIP(c, hipMemcpyAsync(e, c->adj->d.adh, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    return SGK_OK;
}

}  // extern "C"
