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

// Leapfrog drift and periodic wrap of one atom (LeapfrogStepK part 1 after its half-kick, and
// ApplyBoundaryCondK: SurfaceGrowth.cu:397-453): r + dt v if the atom moves, then each coordinate
// wrapped once into [-L/2, L/2).  k_lf1 and the force pass's fused drift both call it, so the two
// paths round alike.
template <class R, class P4>
__device__ __forceinline__ void drift_wrap(const DevParams &p, P4 &r, bool move, R vx, R vy, R vz) {
#pragma clang fp contract(off)
    if (move) {
        const R dt = (R)p.deltaT;
        r.x = r.x + dt * vx;
        r.y = r.y + dt * vy;
        r.z = r.z + dt * vz;
    }
    r.x = wrap1<R>(r.x, (R)p.Lx);
    r.y = wrap1<R>(r.y, (R)p.Ly);
    r.z = wrap1<R>(r.z, (R)p.Lz);
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
// 1/x in FP64: the hardware estimate refined by two Newton steps (within an ulp or two of the
// correctly rounded quotient, at half the instructions of an IEEE division).
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
// sqrt in FP64 from the hardware reciprocal square root: one Goldschmidt step (y ~ 1/sqrt(x),
// g ~ sqrt(x), h ~ y/2) and one Newton correction of g - within about an ulp, at half the
// instructions of the IEEE-rounded sequence.  Returns 0 for x <= 0.
__device__ __forceinline__ double sqrt64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return x > 0.0 ? __builtin_fma(d, h, g) : 0.0;
}
// atan(t) for |t| <= 1 in FP32: t P(t^2), P of degree 8 with P(0) = 1, the minimax fit of the
// relative error on [0, 1] made by fit_atan_unit.py (beside this file); max relative error 1.2e-7
// over every float in [2^-14, 1], FP32 evaluation included (tests/test_atan_unit.py).
__device__ __forceinline__ float atan_unit(float t) {
    const float t2 = t * t;
    float p = 0.0029206919f;
    p = p * t2 - 0.016367926f;
    p = p * t2 + 0.043211858f;
    p = p * t2 - 0.07552214f;
    p = p * t2 + 0.10666005f;
    p = p * t2 - 0.14211055f;
    p = p * t2 + 0.19993773f;
    p = p * t2 - 0.33333153f;
    p = p * t2 + 1.0f;
    return t * p;
}

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
    __device__ static __forceinline__ double rcp(double x) { return rcp64(x); }
    __device__ static __forceinline__ double sqrt(double x) { return sqrt64(x); }
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

// Lane exchange through DPP (data-parallel primitives of the VALU, no LDS round trip).  Lanes in
// rows that RM does not enable read `old`.  All lanes of the wave must be active.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_i32(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ float dpp(float old, float v) {
    return __int_as_float(dpp_i32<CTRL, RM>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp(double old, double v) {
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = dpp_i32<CTRL, RM>((int)o, (int)x);
    const int hi = dpp_i32<CTRL, RM>((int)(o >> 32), (int)(x >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// Full-row patterns (every lane has a source): no fallback value needed.
template <int CTRL>
__device__ __forceinline__ float dpp_all(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_all(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
enum : int {
    kQuadSwap1 = 0xB1,  // quad_perm [1,0,3,2]
    kQuadSwap2 = 0x4E,  // quad_perm [2,3,0,1]
    kRowMirror = 0x140,
    kRowHalfMirror = 0x141,
    kRowBcast15 = 0x142,
    kRowBcast31 = 0x143
};

struct OpSum {
    template <class T>
    __device__ static __forceinline__ T id() { return T(0); }
    template <class T>
    __device__ static __forceinline__ T f(T a, T b) { return a + b; }
};
struct OpMax {
    template <class T>
    __device__ static __forceinline__ T id() { return T(-3.0e38f); }
    template <class T>
    __device__ static __forceinline__ T f(T a, T b) { return b > a ? b : a; }
};
struct OpMin {
    template <class T>
    __device__ static __forceinline__ T id() { return T(3.0e38f); }
    template <class T>
    __device__ static __forceinline__ T f(T a, T b) { return b < a ? b : a; }
};

// Reduction over the wave; the result is valid in lane 63 (fixed order, so reproducible).
template <class Op, class T>
__device__ __forceinline__ T wave_reduce63(T v) {
    const T e = Op::template id<T>();
    v = Op::f(v, dpp_all<kQuadSwap1>(v));
    v = Op::f(v, dpp_all<kQuadSwap2>(v));
    v = Op::f(v, dpp_all<kRowHalfMirror>(v));
    v = Op::f(v, dpp_all<kRowMirror>(v));
    v = Op::f(v, dpp<kRowBcast15, 0xA>(e, v));
    v = Op::f(v, dpp<kRowBcast31, 0xC>(e, v));
    return v;
}

// Sum over aligned groups of G lanes (G a power of two <= 16); every lane of a group gets the
// same value (each step adds a commutative pair).
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16, "group of 1..16 lanes");
    if (G >= 2) v += dpp_all<kQuadSwap1>(v);
    if (G >= 4) v += dpp_all<kQuadSwap2>(v);
    if (G >= 8) v += dpp_all<kRowHalfMirror>(v);
    if (G >= 16) v += dpp_all<kRowMirror>(v);
    return v;
}

}  // namespace sgk
