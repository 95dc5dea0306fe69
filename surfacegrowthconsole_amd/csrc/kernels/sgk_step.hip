// sgk_step.hip - integrator, thermostat, shear, reductions and the per-step bookkeeping.
//
// The reference copies nine reduction results to the host every step and decides on the host
// whether to rebuild the neighbour list (SurfaceGrowth.cu:2242-2395).  Here the force kernels leave
// one Partial per block and k_finalize (one workgroup) reduces them in a fixed order, evaluates the
// same host arithmetic on the device, updates dispHi / nebrNow / shear and appends an sgk_record,
// so consecutive steps run without any host round trip.
#include "sgk_batch.h"
#include "sgk_cells.h"
#include "sgk_device.h"
#include "sgk_internal.h"
#include "sgk_reduce.h"

namespace sgk {

// LeapfrogStepK part 1 (SurfaceGrowth.cu:397-434) fused with ApplyBoundaryCondK (436-453) and,
// when the device flag asks for a rebuild, with the cell binning of the new position
// (BinAtomsIntoCellsK, SurfaceGrowth.cu:547-609; see bin_atom).
// kicked: the previous force pass already applied this step's half-kick (integrate 2), so only the
// drift is left.  drifted: it also did the drift (into the buffer pos now points to), so only the
// binning is left.  rebuild: whether this step bins (< 0: the replica's own flag, st->nebrNow).
template <bool M>
__device__ __forceinline__ void lf1_body(const DevParams &p, int blk, typename Prec<M>::P4 *__restrict__ pos,
                                         typename Prec<M>::R *__restrict__ vel,
                                         const typename Prec<M>::R *__restrict__ acc, const CellBufs &bufs,
                                         DevState *st, int doBin, int kicked, int rebuild = -1,
                                         int drifted = 0) {
#pragma clang fp contract(off)
    using R = typename Prec<M>::R;
    const int n = blk * blockDim.x + threadIdx.x;
    if (n >= p.nMol) return;
    auto r = pos[n];
    if (drifted) {
        if (doBin && (rebuild < 0 ? st->nebrNow : rebuild))
            bin_atom(p, n, make_float3((float)r.x, (float)r.y, (float)r.z), bufs, st);
        return;
    }
    bool move = true;
    if (n >= p.nMolMe)
        move = r.x != (R)p.xLeft && r.x < (R)p.xRight && r.y != (R)p.yTop && r.y != (R)p.yBottom;
    R vx = R(0), vy = R(0), vz = R(0);
    if (move) {
        const R hdt = (R)(0.5f * p.deltaT);
        vx = vel[n];
        vy = vel[p.nMol + n];
        vz = vel[2 * p.nMol + n];
        if (!kicked) {
            vx = vx + hdt * acc[n];
            vy = vy + hdt * acc[p.nMol + n];
            vz = vz + hdt * acc[2 * p.nMol + n];
            vel[n] = vx;
            vel[p.nMol + n] = vy;
            vel[2 * p.nMol + n] = vz;
        }
    }
    drift_wrap<R>(p, r, move, vx, vy, vz);
    pos[n] = r;
    if (doBin && (rebuild < 0 ? st->nebrNow : rebuild))
        bin_atom(p, n, make_float3((float)r.x, (float)r.y, (float)r.z), bufs, st);
}

template <bool M>
__global__ __launch_bounds__(256) void k_lf1(DevParams p, typename Prec<M>::P4 *__restrict__ pos,
                                             typename Prec<M>::R *__restrict__ vel,
                                             const typename Prec<M>::R *__restrict__ acc,
                                             CellBufs bufs, DevState *st, int doBin, int kicked) {
    lf1_body<M>(p, blockIdx.x, pos, vel, acc, bufs, st, doBin, kicked);
}

// Batched: one workgroup of one replica per table entry.
// syncN > 0 (synchronised rebuilds over syncN replicas): the step rebuilds every replica's lists if
// any replica's rule asked for it.  Each workgroup ORs the flags the replicas' finalize left, and the
// first workgroup of each replica stores the OR for the step's later kernels; a workgroup that reads
// a flag while it is being stored sees the replica's own decision or the OR, so it gets the same OR.
// The flag is read and stored with relaxed atomics, so that overlap is defined behaviour.
template <bool M>
__global__ __launch_bounds__(256) void kb_lf1(BatchArgs a, int kicked, int syncN, int drifted) {
    using P4 = typename Prec<M>::P4;
    using R = typename Prec<M>::R;
    const int2 e = a.atomTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    DevState *st = a.state + e.x;
    int rebuild = -1;
    if (syncN > 0) {
        __shared__ int any;
        if (threadIdx.x < 64) {
            int f = 0;
            for (int r = threadIdx.x; r < syncN; r += 64)
                f |= __hip_atomic_load(&a.state[r].nebrNow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            f = __any(f) ? 1 : 0;
            if (threadIdx.x == 0) any = f;
        }
        __syncthreads();
        rebuild = any;
        if (e.y == 0 && threadIdx.x == 0)
            __hip_atomic_store(&st->nebrNow, rebuild, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    lf1_body<M>(B.p, e.y, (P4 *)a.pos + B.atomOff, (R *)a.vel + 3 * B.atomOff, (const R *)a.acc + 3 * B.atomOff,
                rep_cells(a, B), st, 1, kicked, rebuild, drifted);
}

// InsertAtomsK (SurfaceGrowth.cu:488-514) with the schedule of .cu:2148-2168, for the step about to
// run (st->step + 1): metal atoms of this step's group are placed at z = 0.49 Lz moving down at
// velMagDepos; atoms not yet due are parked above the box at rest (outside every cell).
template <bool M>
__global__ __launch_bounds__(256) void k_insert(DevParams p, typename Prec<M>::P4 *__restrict__ pos,
                                                typename Prec<M>::R *__restrict__ vel, const DevState *st) {
    using R = typename Prec<M>::R;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nMolMe) return;
    const int step = st->step + 1, dep = st->nMolDeposited;
    int toDeposit = 0;
    if (step > p.stepEquil) {
        if (dep >= p.nMolMe) return;
        if (p.stepDeposit > 0 && (step - p.stepEquil) % p.stepDeposit == 0) toDeposit = p.nMolToDeposit;
    }
    const int interval = dep + toDeposit;
    if (n >= dep && n < interval) {
        pos[n].z = (R)(0.49f * p.Lz);
        vel[2 * p.nMol + n] = (R)(-p.velMagDepos);
    }
    if (n >= interval) {
        pos[n].z = (R)(1.5f * p.Lz + 0.5f * (float)(n + 1) * p.Lz);
        vel[n] = R(0);
        vel[p.nMol + n] = R(0);
        vel[2 * p.nMol + n] = R(0);
    }
}

// LeapfrogStepK part 2 alone (the step loop fuses it into the force kernels).
template <bool M>
__global__ __launch_bounds__(256) void k_lf2(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                             typename Prec<M>::R *__restrict__ vel,
                                             const typename Prec<M>::R *__restrict__ acc) {
#pragma clang fp contract(off)
    using R = typename Prec<M>::R;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nMol) return;
    const auto r = pos[n];
    if (n >= p.nMolMe && !(r.x != (R)p.xLeft && r.x < (R)p.xRight && r.y != (R)p.yTop && r.y != (R)p.yBottom)) return;
    const R hdt = (R)(0.5f * p.deltaT);
    vel[n] = vel[n] + hdt * acc[n];
    vel[p.nMol + n] = vel[p.nMol + n] + hdt * acc[p.nMol + n];
    vel[2 * p.nMol + n] = vel[2 * p.nMol + n] + hdt * acc[2 * p.nMol + n];
}

// Per-atom contributions of the current state (the reductions of SurfaceGrowth.cu:1101-1891).
template <bool M>
__global__ __launch_bounds__(256) void k_props(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                               const typename Prec<M>::R *__restrict__ vel,
                                               const float *__restrict__ fric, Partial<typename Prec<M>::A> *part) {
    using A = typename Prec<M>::A;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    Partial<A> q;
    partial_identity(q);
    if (n < p.nMol) {
        const auto r = pos[n];
        const A vx = (A)vel[n], vy = (A)vel[p.nMol + n], vz = (A)vel[2 * p.nMol + n];
        const A v2 = vx * vx + vy * vy + vz * vz;
        q.u = (A)r.w;
        q.vmax = v2;
        if (n < p.nMolMe) {
            const A m = (A)p.massMe;
            q.fr = fric ? (A)fric[n] : A(0);
            q.com[0] = (A)r.x;
            q.com[1] = (A)r.y;
            q.com[2] = (A)r.z;
            q.cmv[0] = vx;
            q.cmv[1] = vy;
            q.cmv[2] = vz;
            q.vs[0] = m * vx;
            q.vs[1] = vy;
            q.vs[2] = vz;
            q.vv = m * v2;
            for (int d = 0; d < 3; ++d) q.bmin[d] = q.bmax[d] = q.com[d];
        } else {
            q.vs[0] = vx;
            q.vs[1] = vy;
            q.vs[2] = vz;
            q.vv = v2;
        }
    }
    const Partial<A> r = block_reduce_partial(q);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// The step reduction in two stages: k_reduce1 (one wave per chunk of kPartChunk force-pass
// partials, many waves) and k_finalize (one workgroup over the chunk totals, then the step's
// bookkeeping; mode 1 reduces k_props partials directly).  A single workgroup over all partials
// took 26 us at 1.1 M atoms.
constexpr int kFinalizeThreads = 256;
template <bool M>
__global__ __launch_bounds__(256) void k_reduce1(const Partial<typename Prec<M>::A> *__restrict__ part, int nParts,
                                                 Partial<typename Prec<M>::A> *__restrict__ part2) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nChunks = (nParts + kPartChunk - 1) / kPartChunk;
    if (j < nChunks) reduce_chunk(part + j * kPartChunk, min(kPartChunk, nParts - j * kPartChunk), part2 + j);
}

template <bool M>
__global__ __launch_bounds__(kFinalizeThreads) void k_finalize(DevParams p,
                                                               const Partial<typename Prec<M>::A> *__restrict__ part,
                                                               int nParts, DevState *st, sgk_record *records,
                                                               int recCap, int mode) {
    finalize_body<typename Prec<M>::A>(p, part, nParts, st, records, recCap, mode);
}

// Batched: both stages in one workgroup per replica - the waves reduce the chunks (the calls of
// k_reduce1, in the same order), then the workgroup finalizes - so a replica's reduction order is
// that of the replica run alone.  Up to kFinalizeLdsChunks chunk totals stay in LDS.
constexpr int kFinalizeLdsChunks = 16;
template <bool M>
__global__ __launch_bounds__(kFinalizeThreads) void kb_finalize(BatchArgs a) {
    using A = typename Prec<M>::A;
    __shared__ Partial<A> tot[kFinalizeLdsChunks];
    const BatchRep &B = a.reps[blockIdx.x];
    const Partial<A> *part = (const Partial<A> *)a.partials + B.partOff;
    Partial<A> *part2 = B.nChunks <= kFinalizeLdsChunks ? tot : (Partial<A> *)a.partials + B.part2Off;
    for (int j = threadIdx.x >> 6; j < B.nChunks; j += kFinalizeThreads / 64)
        reduce_chunk(part + j * kPartChunk, min(kPartChunk, B.nbForce - j * kPartChunk), part2 + j);
    __syncthreads();
    finalize_body<A>(B.p, part2, B.nChunks, a.state + blockIdx.x, a.records + (size_t)blockIdx.x * a.recCap,
                     a.recCap, 0);
}

// ApplyBerendsenThermostat (SurfaceGrowth.cu:455-485), kinetic energy of all atoms.
template <bool M>
__device__ __forceinline__ void thermostat_body(const DevParams &p, int blk, typename Prec<M>::R *__restrict__ vel,
                                                const DevState *st, int stepOverride, double vvOverride,
                                                int useOverride, int metalAll) {
    using R = typename Prec<M>::R;
    using A = typename Prec<M>::A;
    const int n = blk * blockDim.x + threadIdx.x;
    if (n >= p.nMol) return;
    const int step = useOverride ? stepOverride : st->step;
    const A vv = (A)(useOverride ? vvOverride : st->vvSum);
    const bool carbonOn = (p.iRegime == 2 && step > p.stepEquil) || p.iRegime == 1 || useOverride;
    const bool metalOn = p.iRegime == 2 && step > p.stepEquil && (metalAll || step < p.stepEquil + p.stepCool);
    if (!carbonOn) return;
    const A ke = vv * A(0.5f) / (A)p.nMol;
    A b;
    if (n >= p.nMolMe)
        b = sqrt(A(1.f) + (A)p.gamma * ((A)p.temperature * A(1.5f) * (A)p.kB / ke - A(1.f)));
    else if (metalOn)
        b = sqrt(A(1.f) + (A)p.gamma * ((A)p.temperature * (A)p.kB / ke - A(1.f)));
    else
        return;
    vel[n] = (R)(b * (A)vel[n]);
    vel[p.nMol + n] = (R)(b * (A)vel[p.nMol + n]);
    vel[2 * p.nMol + n] = (R)(b * (A)vel[2 * p.nMol + n]);
}

template <bool M>
__global__ __launch_bounds__(256) void k_thermostat(DevParams p, typename Prec<M>::R *__restrict__ vel,
                                                    const DevState *st, int stepOverride, double vvOverride,
                                                    int useOverride, int metalAll) {
    thermostat_body<M>(p, blockIdx.x, vel, st, stepOverride, vvOverride, useOverride, metalAll);
}

template <bool M>
__global__ __launch_bounds__(256) void kb_thermostat(BatchArgs a, int metalAll) {
    const int2 e = a.atomTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    thermostat_body<M>(B.p, e.y, (typename Prec<M>::R *)a.vel + 3 * B.atomOff, a.state + e.x, 0, 0.0, 0, metalAll);
}

// ApplyShearK (SurfaceGrowth.cu:517-534): metal atoms left of the centre of mass get the shear
// acceleration; the count goes into the step's record.
template <bool M>
__device__ __forceinline__ void shear_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                           typename Prec<M>::R *__restrict__ acc, const DevState *st,
                                           sgk_record *records, int recCap) {
    const int n = blk * blockDim.x + threadIdx.x;
    int hit = 0;
    if (n < p.nMolMe) {
        const float dr = (float)st->comX - (float)pos[n].x;
        if (dr > 0.f) {
            acc[n] += (typename Prec<M>::R)st->shear;
            hit = 1;
        }
    }
    const int cnt = __syncthreads_count(hit);
    const int slot = st->rec - 1;
    if (threadIdx.x == 0 && cnt && slot >= 0 && slot < recCap) atomicAdd(&records[slot].shearCount, (double)cnt);
}

template <bool M>
__global__ __launch_bounds__(256) void k_shear(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                               typename Prec<M>::R *__restrict__ acc, const DevState *st,
                                               sgk_record *records, int recCap) {
    shear_body<M>(p, blockIdx.x, pos, acc, st, records, recCap);
}

template <bool M>
__global__ __launch_bounds__(256) void kb_shear(BatchArgs a) {
    const int2 e = a.metalTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    if (B.p.deltaF == 0.f) return;  // uniform per workgroup; this replica's shear stays 0
    shear_body<M>(B.p, e.y, (const typename Prec<M>::P4 *)a.pos + B.atomOff,
                  (typename Prec<M>::R *)a.acc + 3 * B.atomOff, a.state + e.x, a.records + (size_t)e.x * a.recCap,
                  a.recCap);
}

// Centre-of-mass diffusion on the device (EvalDiffusion + AccumDiffusion, SurfaceGrowth.cu:2898-2952),
// one workgroup, in FP64: after the step's reduction, the metal centre of mass comSum / nMolMe of the
// step's record is one sample.  Each staggered buffer (one thread each) unwraps the sample next to
// its previous true position by Nint of the box (ComputeDefs.h:64-65) and stores |r_true - origin|^2
// at its count; a buffer that is full adds into the running average, in buffer order, and every
// limitAv completed buffers the average becomes a table D_j = av_j / (6 stepDiffuse dt limitAv j)
// in the ring `out` (tables counts them).  Counts start at -nb nVal / nBuff (InitDiffusion).
__device__ __forceinline__ double nint64(double x) { return x < 0. ? -(double)(int)(0.5 - x) : (double)(int)(0.5 + x); }

__global__ __launch_bounds__(256) void k_diffuse(DiffusionArgs a, const DevState *st, const sgk_record *records,
                                                 int nMolMe) {
#pragma clang fp contract(off)
    __shared__ double com[3];
    __shared__ int cnt[kMaxDiffuseBuffers];
    __shared__ int cav, ntab;
    const int t = threadIdx.x;
    if (t == 0) {
        const sgk_record &r = records[st->rec - 1];
        for (int k = 0; k < 3; ++k) com[k] = r.comSum[k] / (double)nMolMe;
        cav = a.state[0];
        ntab = a.state[1];
    }
    if (t < a.nBuff) cnt[t] = a.count[t];
    __syncthreads();
    if (t < a.nBuff) {
        double *org = a.org + 3 * t, *tr = a.tru + 3 * t;
        if (cnt[t] == 0)
            for (int k = 0; k < 3; ++k) org[k] = tr[k] = com[k];
        if (cnt[t] >= 0) {
            double acc = 0.;
            for (int k = 0; k < 3; ++k) {
                const double L = a.L[k];
                double dr = tr[k] - com[k];
                dr = dr / L;
                dr = nint64(dr);
                dr = dr * L;
                const double u = com[k] + dr;
                tr[k] = u;
                const double e = u - org[k];
                acc += e * e;
            }
            a.rr[(size_t)t * a.nVal + cnt[t]] = acc;
        }
        ++cnt[t];
    }
    __syncthreads();
    for (int nb = 0; nb < a.nBuff; ++nb) {
        if (cnt[nb] != a.nVal) continue;  // shared value: the branch is uniform
        for (int j = t; j < a.nVal; j += blockDim.x) a.av[j] = a.av[j] + a.rr[(size_t)nb * a.nVal + j];
        __syncthreads();
        if (t == 0) {
            cnt[nb] = 0;
            ++cav;
        }
        __syncthreads();
        if (cav == a.limitAv) {
            const double fac = 1. / ((double)(6 * a.stepDiffuse) * a.deltaT * (double)a.limitAv);
            double *out = a.out + (size_t)(ntab % a.cap) * a.nVal;
            for (int j = t; j < a.nVal; j += blockDim.x) {
                out[j] = j >= 1 ? a.av[j] * (fac / (double)j) : a.av[j];
                a.av[j] = 0.;
            }
            __syncthreads();
            if (t == 0) {
                ++ntab;
                cav = 0;
            }
            __syncthreads();
        }
    }
    if (t < a.nBuff) a.count[t] = cnt[t];
    if (t == 0) {
        a.state[0] = cav;
        a.state[1] = ntab;
    }
}

// Sets the buffers to their start (InitDiffusion + ZeroDiffusion, SurfaceGrowth.cu:2865-2881).
__global__ __launch_bounds__(256) void k_diffuse_reset(DiffusionArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.nBuff) a.count[t] = -t * a.nVal / a.nBuff;
    for (int j = t; j < a.nVal; j += gridDim.x * blockDim.x) a.av[j] = 0.;
    if (t == 0) {
        a.state[0] = 0;
        a.state[1] = 0;
    }
}

__global__ void k_reset_records(DevState *st) { st->rec = 0; }

__global__ void kb_reset_records(DevState *st, int nrep) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrep) st[r].rec = 0;
}

// ---- launchers ---------------------------------------------------------------------------------
static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

size_t partial_bytes_for(bool mixed) { return mixed ? sizeof(Partial<double>) : sizeof(Partial<float>); }
size_t partial_bytes(const sgk_ctx *c) { return partial_bytes_for(c->mixed); }

void launch_lf1(sgk_ctx *c, bool bin, bool kicked) {
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    if (c->mixed)
        hipLaunchKernelGGL(k_lf1<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (double4 *)c->pos, (double *)c->vel,
                           (const double *)c->acc, cell_bufs(c), c->state, (int)bin, (int)kicked);
    else
        hipLaunchKernelGGL(k_lf1<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (float4 *)c->pos, (float *)c->vel,
                           (const float *)c->acc, cell_bufs(c), c->state, (int)bin, (int)kicked);
    profile_end(c, "lf1");
}

void launch_insert(sgk_ctx *c) {
    if (c->dp.nMolMe == 0 || c->dp.iRegime != 1) return;
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMolMe, 256);
    if (c->mixed)
        hipLaunchKernelGGL(k_insert<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (double4 *)c->pos,
                           (double *)c->vel, c->state);
    else
        hipLaunchKernelGGL(k_insert<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (float4 *)c->pos,
                           (float *)c->vel, c->state);
    profile_end(c, "insert");
}

void launch_lf2(sgk_ctx *c) {
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    if (c->mixed)
        hipLaunchKernelGGL(k_lf2<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (const double4 *)c->pos,
                           (double *)c->vel, (const double *)c->acc);
    else
        hipLaunchKernelGGL(k_lf2<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (const float4 *)c->pos,
                           (float *)c->vel, (const float *)c->acc);
    profile_end(c, "lf2");
}

void launch_props(sgk_ctx *c) {
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    float *fric = (c->dp.iRegime != 0 && c->dp.nMolMe) ? c->fric : nullptr;
    if (c->mixed)
        hipLaunchKernelGGL(k_props<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (const double4 *)c->pos,
                           (const double *)c->vel, fric, (Partial<double> *)c->partials);
    else
        hipLaunchKernelGGL(k_props<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (const float4 *)c->pos,
                           (const float *)c->vel, fric, (Partial<float> *)c->partials);
    profile_end(c, "props");
}

template <bool M>
static void finalize_dispatch(sgk_ctx *c, int mode) {
    using A = typename Prec<M>::A;
    Partial<A> *part = (Partial<A> *)c->partials;
    if (mode == 1) {
        hipLaunchKernelGGL(k_finalize<M>, dim3(1), dim3(kFinalizeThreads), 0, c->stream, c->dp, part, c->nbProps,
                           c->state, c->records, 1, 1);
        return;
    }
    const int nParts = c->nbM + c->nbC, nChunks = (nParts + kPartChunk - 1) / kPartChunk;
    Partial<A> *part2 = part + c->part2Off;
    hipLaunchKernelGGL(k_reduce1<M>, dim3((nChunks + 3) / 4), dim3(256), 0, c->stream, part, nParts, part2);
    hipLaunchKernelGGL(k_finalize<M>, dim3(1), dim3(kFinalizeThreads), 0, c->stream, c->dp, part2, nChunks, c->state,
                       c->records, c->cfg.records_capacity, 0);
}

void launch_finalize(sgk_ctx *c, int mode) {
    profile_begin(c);
    if (c->mixed)
        finalize_dispatch<true>(c, mode);
    else
        finalize_dispatch<false>(c, mode);
    profile_end(c, "finalize");
}

void launch_thermostat(sgk_ctx *c, int stepOverride, double vvOverride, int useOverride) {
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    const int metalAll = c->cfg.metal_thermostat_all;
    if (c->mixed)
        hipLaunchKernelGGL(k_thermostat<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (double *)c->vel, c->state,
                           stepOverride, vvOverride, useOverride, metalAll);
    else
        hipLaunchKernelGGL(k_thermostat<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (float *)c->vel, c->state,
                           stepOverride, vvOverride, useOverride, metalAll);
    profile_end(c, "thermostat");
}

void launch_shear(sgk_ctx *c) {
    if (c->dp.nMolMe == 0) return;
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMolMe, 256);
    if (c->mixed)
        hipLaunchKernelGGL(k_shear<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (const double4 *)c->pos,
                           (double *)c->acc, c->state, c->records, c->cfg.records_capacity);
    else
        hipLaunchKernelGGL(k_shear<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (const float4 *)c->pos,
                           (float *)c->acc, c->state, c->records, c->cfg.records_capacity);
    profile_end(c, "shear");
}

void launch_reset_records(sgk_ctx *c) { hipLaunchKernelGGL(k_reset_records, dim3(1), dim3(1), 0, c->stream, c->state); }

void launch_diffuse(sgk_ctx *c) {
    if (!c->diff.count || c->dp.nMolMe == 0) return;
    profile_begin(c);
    hipLaunchKernelGGL(k_diffuse, dim3(1), dim3(256), 0, c->stream, c->diff, c->state, c->records, c->dp.nMolMe);
    profile_end(c, "diffuse");
}

void launch_diffuse_reset(sgk_ctx *c) {
    if (!c->diff.count) return;
    hipLaunchKernelGGL(k_diffuse_reset, dim3(blocks_for(std::max(c->diff.nVal, c->diff.nBuff), 256)), dim3(256), 0,
                       c->stream, c->diff);
}

// ---- batched launchers -----------------------------------------------------------------------
void launch_batch_lf1(sgk_batch *b, bool kicked, bool drifted) {
    const int syncN = b->cfg.sync_rebuilds && b->nrep > 1 ? b->nrep : 0;
    if (b->mixed)
        hipLaunchKernelGGL(kb_lf1<true>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), (int)kicked, syncN,
                           (int)drifted);
    else
        hipLaunchKernelGGL(kb_lf1<false>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), (int)kicked, syncN,
                           (int)drifted);
}

// One workgroup per replica closes the step (the rebuild flags of synchronised rebuilds are ORed at
// the start of the next step, in kb_lf1).
void launch_batch_finalize(sgk_batch *b) {
    if (b->mixed)
        hipLaunchKernelGGL(kb_finalize<true>, dim3(b->nrep), dim3(kFinalizeThreads), 0, b->stream, b->args());
    else
        hipLaunchKernelGGL(kb_finalize<false>, dim3(b->nrep), dim3(kFinalizeThreads), 0, b->stream, b->args());
}

void launch_batch_thermostat(sgk_batch *b) {
    const int metalAll = b->cfg.metal_thermostat_all;
    if (b->mixed)
        hipLaunchKernelGGL(kb_thermostat<true>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), metalAll);
    else
        hipLaunchKernelGGL(kb_thermostat<false>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), metalAll);
}

void launch_batch_shear(sgk_batch *b) {
    if (b->nMetalBlocks == 0) return;
    if (b->mixed)
        hipLaunchKernelGGL(kb_shear<true>, dim3(b->nMetalBlocks), dim3(256), 0, b->stream, b->args());
    else
        hipLaunchKernelGGL(kb_shear<false>, dim3(b->nMetalBlocks), dim3(256), 0, b->stream, b->args());
}

void launch_batch_reset_records(sgk_batch *b) {
    hipLaunchKernelGGL(kb_reset_records, dim3((b->nrep + 255) / 256), dim3(256), 0, b->stream, b->state, b->nrep);
}

}  // namespace sgk
