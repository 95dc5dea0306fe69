// sgk_device.h - device-side types and math shared by the gfx950 kernels.
//
// Precision forms:
//   Plain (MIXED=false): positions/velocities float, sums float - the reference's FP32 throughout.
//   Mixed (MIXED=true):  positions/velocities double, sums double; displacement vectors, distances
//                        and small differences formed in FP64; pair arithmetic in FP32 or, with
//                        pair_fp64, FP64 (template parameter T of the force kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgk_params.h"

namespace sgk {

template <bool MIXED>
struct Prec;
template <>
struct Prec<false> {
    using R = float;   // storage / integration precision
    using P4 = float4; // x, y, z, E_pot
    using A = float;   // reduction accumulator
};
template <>
struct Prec<true> {
    using R = double;
    using P4 = double4;
    using A = double;
};

// Single-image periodic correction as in every reference force kernel (SurfaceGrowth.cu:752-759),
// evaluated in the storage precision and rounded to float afterwards.
template <class R>
__device__ __forceinline__ R wrap1(R d, R L) {
    if (d >= R(0.5f) * L)
        d -= L;
    else if (d < -R(0.5f) * L)
        d += L;
    return d;
}

template <class P4, class R>
__device__ __forceinline__ float3 min_image(const P4 &a, const P4 &b, R Lx, R Ly, R Lz) {
    return make_float3((float)wrap1<R>(a.x - b.x, Lx), (float)wrap1<R>(a.y - b.y, Ly),
                       (float)wrap1<R>(a.z - b.z, Lz));
}

// gfx950 dispatches workgroups to its 8 XCDs round robin: workgroup b runs on XCD b % 8, and each
// XCD has its own L2.  xcd_block maps the hardware index b of a range of n workgroups to a logical
// index so that every XCD works on one contiguous 1/8 of the range.  Atoms with nearby indices are
// spatial neighbours, so an XCD's gathers then stay inside the positions its own L2 holds.
// The map is a bijection of [0, n); a range that starts at a hardware index that is a multiple of
// 8 keeps the XCD of every workgroup.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int q = n / kXcds, r = n % kXcds, x = b % kXcds, k = b / kXcds;
    return x * q + min(x, r) + k;
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

// Pair-math precision T (float, or double in the mixed form's FP64-pair variant).
template <class T>
struct PM;
template <>
struct PM<float> {
    __device__ static __forceinline__ float rcp(float x) { return rcpf(x); }
    __device__ static __forceinline__ float sqrt(float x) { return sqrt_fast(x); }
    __device__ static __forceinline__ float exp(float x) { return expf(x); }
    __device__ static __forceinline__ float log(float x) { return logf(x); }
};
template <>
struct PM<double> {
    __device__ static __forceinline__ double rcp(double x) { return 1.0 / x; }
    __device__ static __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    __device__ static __forceinline__ double exp(double x) { return ::exp(x); }
    __device__ static __forceinline__ double log(double x) { return ::log(x); }
};

// x^19 and x^20 by multiplication (the reference calls powf(x, 19.f) and powf(x, 20.f)).
template <class T>
__device__ __forceinline__ void pow19_20(T x, T &p19, T &p20) {
    T x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    p19 = x16 * x2 * x;
    p20 = x16 * x4;
}

// Electron-density contribution f(r) (SurfaceGrowth.cu:61-69).
template <class T>
__device__ __forceinline__ T eam_f(const DevParams &p, T r) {
    T x = r * (T)p.rei;
    T l19, l20;
    pow19_20<T>(x - (T)p.lambda, l19, l20);
    return (T)p.fe * PM<T>::exp(-(T)p.beta * (x - T(1))) * PM<T>::rcp(T(1) + l20);
}

// phi(r), -phi'(r) and f'(r) in one pass (SurfaceGrowth.cu:46-58, 124-158), sharing the two
// exponentials and the powers.
template <class T>
__device__ __forceinline__ void eam_pair(const DevParams &p, T r, T &phi, T &mdphi, T &df) {
    const T rei = (T)p.rei, al = (T)p.alpha, be = (T)p.beta;
    T x = r * rei;
    T k19, k20, l19, l20;
    pow19_20<T>(x - (T)p.kappa, k19, k20);
    pow19_20<T>(x - (T)p.lambda, l19, l20);
    T ea = PM<T>::exp(-al * (x - T(1)));
    T eb = PM<T>::exp(-be * (x - T(1)));
    T iA = PM<T>::rcp(T(1) + k20), iB = PM<T>::rcp(T(1) + l20);
    T tA = (T)p.A * ea * iA, tB = (T)p.B * eb * iB;
    T gB = (be + T(20) * l19 * iB);
    phi = tA - tB;
    // -phi' = (A ea (alpha + 20 k19/dA)/dA - B eb (beta + 20 l19/dB)/dB) * rei
    mdphi = (tA * (al + T(20) * k19 * iA) - tB * gB) * rei;
    df = -(T)p.fe * rei * eb * gB * iB;
}

// Embedding F(rho) and F'(rho), three branches (SurfaceGrowth.cu:72-118, 160-210).
template <class T>
__device__ __forceinline__ T eam_embed(const DevParams &p, T rho, T &dF) {
    const T rhoe = (T)p.rhoe, rhon = T(0.85f) * rhoe, rhoo = T(1.15f) * rhoe, rhoei = (T)p.rhoei;
    if (rho < rhon) {
        T rhoni = rhoei * T(1.176471f);
        T t = rho * rhoni - T(1);
        dF = rhoni * ((T)p.Fn[1] + T(2) * (T)p.Fn[2] * t + T(3) * (T)p.Fn[3] * t * t);
        return (T)p.Fn[0] + (T)p.Fn[1] * t + (T)p.Fn[2] * t * t + (T)p.Fn[3] * t * t * t;
    }
    if (rho < rhoo) {
        T t = rho * rhoei - T(1);
        dF = rhoei * ((T)p.F[1] + T(2) * (T)p.F[2] * t + T(3) * (T)p.F[3] * t * t);
        return (T)p.F[0] + (T)p.F[1] * t + (T)p.F[2] * t * t + (T)p.F[3] * t * t * t;
    }
    // F = Fe (1 - ln x^eta) x^eta, F' = -Fe eta rhoei x^(eta-1) ln x^eta with x = rho/rhoe
    T x = rho * rhoei;
    T lx = PM<T>::log(x);
    T pe = PM<T>::exp((T)p.eta * lx);
    T lp = (T)p.eta * lx;
    dF = -(T)p.Fe * (T)p.eta * rhoei * (pe * PM<T>::rcp(x)) * lp;
    return (T)p.Fe * (T(1) - lp) * pe;
}

// Lennard-Jones force factor and energy (SurfaceGrowth.cu:880-884).
template <class T>
__device__ __forceinline__ T lj_pair(const DevParams &p, T rr, T &u) {
    T rri = (T)p.sigma2 * PM<T>::rcp(rr);
    T rri3 = rri * rri * rri;
    u = T(4) * (T)p.epsLJ * rri3 * (rri3 - T(1));
    return T(48) * (T)p.epsLJ * rri3 * (rri3 - T(0.5)) * rri * (T)p.inv_sigma2;
}

// Wave64 reductions through DPP-backed shuffles.
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
// Sum over aligned groups of G lanes (G a power of two <= 64).
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace sgk
