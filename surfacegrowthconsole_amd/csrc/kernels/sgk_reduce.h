// sgk_reduce.h - per-block partial reductions (the force and props kernels) and the per-step
// bookkeeping that k_finalize / kb_finalize run over them.
#pragma once

#include "sgk.h"
#include "sgk_device.h"

namespace sgk {

template <class T>
__device__ __forceinline__ void partial_identity(Partial<T> &q) {
    q.u = q.fr = q.vv = q.vmax = q.vir = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q.com[d] = q.cmv[d] = q.vs[d] = T(0);
        q.bmin[d] = T(3.0e38f);
        q.bmax[d] = T(-3.0e38f);
    }
}

template <class T>
__device__ __forceinline__ void partial_combine(Partial<T> &r, const Partial<T> &s) {
    r.u += s.u;
    r.vir += s.vir;
    r.fr += s.fr;
    r.vv += s.vv;
    r.vmax = s.vmax > r.vmax ? s.vmax : r.vmax;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        r.com[d] += s.com[d];
        r.cmv[d] += s.cmv[d];
        r.vs[d] += s.vs[d];
        r.bmin[d] = s.bmin[d] < r.bmin[d] ? s.bmin[d] : r.bmin[d];
        r.bmax[d] = s.bmax[d] > r.bmax[d] ? s.bmax[d] : r.bmax[d];
    }
}

// Wave reduction of a Partial (result in lane 63).  METAL = false reduces only the fields a
// carbon atom contributes (energy, virial, velocity sums, v^2 max); the others stay at their identity.
template <bool METAL, class T>
__device__ __forceinline__ void partial_wave(Partial<T> &q) {
    q.u = wave_reduce63<OpSum>(q.u);
    q.vir = wave_reduce63<OpSum>(q.vir);
    q.vv = wave_reduce63<OpSum>(q.vv);
    q.vmax = wave_reduce63<OpMax>(q.vmax);
#pragma unroll
    for (int d = 0; d < 3; ++d) q.vs[d] = wave_reduce63<OpSum>(q.vs[d]);
    if (METAL) {
        q.fr = wave_reduce63<OpSum>(q.fr);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            q.com[d] = wave_reduce63<OpSum>(q.com[d]);
            q.cmv[d] = wave_reduce63<OpSum>(q.cmv[d]);
            q.bmin[d] = wave_reduce63<OpMin>(q.bmin[d]);
            q.bmax[d] = wave_reduce63<OpMax>(q.bmax[d]);
        }
    }
}

// Reduce one Partial per thread of the block; thread 0 returns the total (fixed order).  Every
// thread of the block must call it.
template <bool METAL = true, class T>
__device__ Partial<T> block_reduce_partial(Partial<T> q) {
    __shared__ Partial<T> sh[16];
    partial_wave<METAL>(q);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 63) sh[w] = q;
    __syncthreads();
    Partial<T> r = sh[0];
    if (threadIdx.x == 0)
        for (int k = 1; k < nw; ++k) partial_combine(r, sh[k]);
    return r;
}

// First stage of the step reduction: one wave reduces a chunk of up to kPartChunk partials (lane k
// takes partial k; fixed order), the result goes to *out.  The chunk count is wave-uniform.
constexpr int kPartChunk = 64;
template <class T>
__device__ __forceinline__ void reduce_chunk(const Partial<T> *part, int n, Partial<T> *out) {
    const int lane = threadIdx.x & 63;
    Partial<T> q;
    partial_identity(q);
    if (lane < n) q = part[lane];
    partial_wave<true>(q);
    if (lane == 63) *out = q;
}

// Reduce nParts partials with the whole block (each thread a strided subset, then the block tree:
// the order is fixed, so the result is bitwise reproducible), then evaluate the reference's host
// arithmetic of SurfaceGrowth.cu:2307-2395 in thread 0.
// mode 0: one step - advance the step counter, commit a rebuild, update dispHi / nebrNow / shear,
//         append a record.  mode 1: evaluation only - write records[0], leave the state alone.
template <class A>
__device__ void finalize_body(const DevParams &p, const Partial<A> *part, int nParts, DevState *st,
                              sgk_record *records, int recCap, int mode) {
    Partial<A> q;
    partial_identity(q);
    for (int k = threadIdx.x; k < nParts; k += blockDim.x) partial_combine(q, part[k]);
    const Partial<A> P = block_reduce_partial(q);
    if (threadIdx.x != 0) return;
    const int nMe = p.nMolMe, nMol = p.nMol;
    sgk_record rec;
    rec.shear = 0;
    rec.shearCount = 0;
    // cmVel.val = cmVel.x / nMolMe (SurfaceGrowth.cu:2308)
    const A cmVelX = nMe ? P.cmv[0] / (A)nMe : A(0);
    // ComputeVvSumK subtracts cmVel.val from metal vx in regime 2 (SurfaceGrowth.cu:1230-1234):
    // sum m((vx-c)^2+vy^2+vz^2) = sum m|v|^2 - m c sum vx, since sum vx = nMe c.
    A vvSum = P.vv;
    if (p.iRegime == 2 && nMe) vvSum = vvSum - (A)p.massMe * cmVelX * P.cmv[0];
    // the metal bounding box includes the reference reduction's zero padding (SurfaceGrowth.cu:1671-1682)
    const bool pad = nMe != 1024 * 1024;
    for (int d = 0; d < 3; ++d) {
        rec.vSum[d] = (double)P.vs[d];
        rec.comSum[d] = nMe ? (double)P.com[d] : 0.0;
        rec.cmVelSum[d] = nMe ? (double)P.cmv[d] : 0.0;
        A mn = P.bmin[d], mx = P.bmax[d];
        if (pad) {
            mn = mn < A(0) ? mn : A(0);
            mx = mx > A(0) ? mx : A(0);
        }
        rec.bboxMin[d] = nMe ? (double)mn : 0.0;
        rec.bboxMax[d] = nMe ? (double)mx : 0.0;
    }
    const A ke = A(0.5f) * vvSum / (A)nMol, pe = P.u / (A)nMol;
    rec.kinEnergy = (double)ke;
    rec.potEnergy = (double)pe;
    rec.totEnergy = (double)(ke + pe);
    rec.vvSum = (double)vvSum;
    rec.vvMax = (double)P.vmax;
    rec.uSum = (double)P.u;
    rec.virial = (double)P.vir;
    rec.frict = (p.iRegime != 0 && nMe) ? (double)P.fr : 0.0;
    if (mode == 1) {
        rec.step = st->step;
        rec.rebuilt = 0;
        rec.nebrNow = st->nebrNow;
        rec.errors = st->errors;
        rec.dispHi = st->dispHi;
        rec.nMolDeposited = st->nMolDeposited;
        rec.pad = 0;
        records[0] = rec;
        return;
    }
    const int step = st->step + 1;
    st->step = step;
    rec.step = step;
    rec.rebuilt = st->nebrNow;
    A dispHi = (A)st->dispHi;
    if (st->nebrNow) {  // this step rebuilt: the new cell buffer becomes current (SurfaceGrowth.cu:2193-2195)
        const int old = st->cellbuf;
        st->cellbuf = old ^ 1;
        st->rebuilds += 1;
        dispHi = A(0);
    }
    const A dstep = (A)sqrt(P.vmax) * (A)p.deltaT;
    dispHi += dstep;  // SurfaceGrowth.cu:2390-2391
    // the reference's skin rule; a fixed cadence adds rebuilds but never removes one
    int nebrNow = dispHi > A(0.5f) * (A)p.rNebrShell;
    if (p.rebuildEvery > 0 && step % p.rebuildEvery == 0) nebrNow = 1;
    if ((double)(dispHi + dstep) > (double)p.dispBound && !nebrNow) st->errors |= SGK_FLAG_DISPLACEMENT;
    st->nebrNow = nebrNow;
    st->dispHi = (double)dispHi;
    st->vvSum = (double)vvSum;
    st->cmVelX = (double)cmVelX;
    st->comX = nMe ? (double)(P.com[0] / (A)nMe) : 0.0;
    // shear increment (SurfaceGrowth.cu:2398-2404)
    if (p.iRegime == 2 && nMe && step > p.stepEquil + p.stepCool) {
        if ((float)cmVelX < 0.005f) st->shear = (double)((A)st->shear + (A)p.deltaF);
    }
    // deposition schedule (SurfaceGrowth.cu:2148-2168): the step's group counts from now on
    if (p.iRegime == 1 && step > p.stepEquil && st->nMolDeposited < nMe && p.stepDeposit > 0 &&
        (step - p.stepEquil) % p.stepDeposit == 0)
        st->nMolDeposited += p.nMolToDeposit;
    rec.nMolDeposited = st->nMolDeposited;
    rec.pad = 0;
    rec.shear = st->shear;
    rec.nebrNow = nebrNow;
    rec.dispHi = st->dispHi;
    rec.errors = st->errors;
    if (st->rec < recCap) records[st->rec] = rec;
    st->rec += 1;
}

}  // namespace sgk
