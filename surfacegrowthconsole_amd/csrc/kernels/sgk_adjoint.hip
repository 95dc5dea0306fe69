// sgk_adjoint.hip - reverse-mode (adjoint) kernels of the MD step for gfx950.
//
// No reference counterpart: the reference integrates forward only.  For one step n of a window the
// reverse sweep needs the vector-Jacobian product of the force pass: for a cotangent Lam of the
// accelerations and a weight wE of the adhesion energy,
//   G    = d/dr   [ Lam . a(r) + wE E_adh(r) ]
//   pbar = d/ds_k [ same ],  s_k the scale factors of LJ epsilon, LJ sigma, embedding F, density f,
// in FP64 at the positions of step n with the neighbour structures the forward step used.  Every
// discrete decision (cutoff tests, bonds, neighbours of bonds, the LJ multiplicity of the carbon
// loop) is taken exactly as the force pass takes it (same FP64 tests on the same lists), so G and
// pbar are derivatives of the force law the forward step evaluated (SurfaceGrowth.cu:776-1091).
//
// Each atom owns its force expression a_i and scatters the derivatives of Lam_i . a_i with respect
// to every position the expression reads - its own, its list neighbours', and for carbon the
// neighbours of its bonded neighbours - with FP64 atomics (hardware global_atomic_add_f64).  The
// EAM density, read through F'(rho_i) and F'(rho_k), takes two passes: k_adj_force accumulates the
// density adjoint rhoBar, k_adj_rho2 turns it into position derivatives.  The atomics make G and
// pbar reproducible to rounding only (summation order); the forward recomputation is bitwise.
//
// Mixed-precision contexts only (positions double4, velocities/accelerations FP64 SoA).
#include "sgk_adjoint.h"
#include "sgk_device.h"
#include "sgk_internal.h"

namespace sgk {

namespace {

constexpr int kBs = 256;

// Displacement r_a - r_b with the single-image wrap and its square, FP64, rounded as the force
// pass rounds it (no contraction), so the cutoff tests decide identically.
__device__ __forceinline__ double dispd(const double4 &a, const double4 &b, const DevParams &p, double (&d)[3]) {
#pragma clang fp contract(off)
    d[0] = wrap1<double>(a.x - b.x, (double)p.Lx);
    d[1] = wrap1<double>(a.y - b.y, (double)p.Ly);
    d[2] = wrap1<double>(a.z - b.z, (double)p.Lz);
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

__device__ __forceinline__ double dot3(const double (&a)[3], const double (&b)[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Carbon atoms on the sheet perimeter stay fixed (SurfaceGrowth.cu:415-424), exact compares.
__device__ __forceinline__ bool moves(const DevParams &p, int n, const double4 &r) {
    if (n < p.nMolMe) return true;
    return r.x != (double)p.xLeft && r.x < (double)p.xRight && r.y != (double)p.yTop && r.y != (double)p.yBottom;
}

// Sum over a 256-thread block; the result is valid in thread 0.  Every thread must call it.
__device__ double block_sum(double v) {
    __shared__ double sh[kBs / 64];
    v = wave_reduce63<OpSum>(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBs / 64; ++w) r += sh[w];
    return r;
}

__device__ __forceinline__ void add_pbar(double *pbar, double p0, double p1, double p2, double p3) {
    const double s0 = block_sum(p0), s1 = block_sum(p1), s2 = block_sum(p2), s3 = block_sum(p3);
    if (threadIdx.x == 0) {
        if (s0 != 0.0) atomicAdd(&pbar[0], s0);
        if (s1 != 0.0) atomicAdd(&pbar[1], s1);
        if (s2 != 0.0) atomicAdd(&pbar[2], s2);
        if (s3 != 0.0) atomicAdd(&pbar[3], s3);
    }
}

__device__ __forceinline__ void scatter(double *G, int N, int k, double sgn, const double (&g)[3]) {
    atomicAdd(&G[k], sgn * g[0]);
    atomicAdd(&G[N + k], sgn * g[1]);
    atomicAdd(&G[2 * N + k], sgn * g[2]);
}

// ---- FP64 force-law derivatives ----------------------------------------------------------------
// Lennard-Jones (SurfaceGrowth.cu:880-884) with a = (sigma^2/rr)^3:
//   fc = 48 eps a (a - 1/2) / rr     (a_i += fc d)        U = 4 eps a (a - 1)
//   fcp = d fc / d rr = 24 eps (4a - 14a^2) / rr^2
//   sfc = sigma d fc / d sigma = 144 eps a (4a - 1) / rr,  sU = sigma dU/dsigma = fc rr,
//   and eps d/d eps is the identity.
struct LjD {
    double fc, fcp, U, sfc, sU;
};
__device__ __forceinline__ LjD lj_d(const DevParams &p, double rr) {
    const double sig = (double)p.sigmaLJ, eps = (double)p.epsLJ, s2 = sig * sig;
    const double rri = s2 / rr, a = rri * rri * rri;
    LjD o;
    o.fc = 48.0 * eps * a * (a - 0.5) * rri / s2;
    o.U = 4.0 * eps * a * (a - 1.0);
    o.fcp = 24.0 * eps * (4.0 * a - 14.0 * a * a) / (rr * rr);
    o.sfc = 144.0 * eps * a * (4.0 * a - 1.0) / rr;
    o.sU = o.fc * rr;
    return o;
}

__device__ __forceinline__ void pow18_20(double x, double &p18, double &p19, double &p20) {
    const double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    p18 = x16 * x2;
    p19 = p18 * x;
    p20 = x16 * x4;
}

// EAM pair functions (SurfaceGrowth.cu:46-69, 124-158).  Each term is g = C e^{-a(x-1)}/(1+q),
// q = (x-k)^20, x = r/re; with P = a + 20 (x-k)^19/(1+q):  g' = -rei g P and
// g'' = rei g (rei P^2 - P_r),  P_r = dP/dr = rei 20 (x-k)^18 (19 - q)/(1+q)^2.
struct EamD {
    double f, f1, f2, phi1, phi2;  // f, f', f'', phi', phi''
};
__device__ __forceinline__ EamD eam_d(const DevParams &p, double r) {
    const double rei = p.rei, al = p.alpha, be = p.beta, x = r * rei;
    double k18, k19, k20, l18, l19, l20;
    pow18_20(x - (double)p.kappa, k18, k19, k20);
    pow18_20(x - (double)p.lambda, l18, l19, l20);
    const double ia = 1.0 / (1.0 + k20), ib = 1.0 / (1.0 + l20);
    const double ga = (double)p.A * exp(-al * (x - 1.0)) * ia;
    const double eb = exp(-be * (x - 1.0));
    const double gb = (double)p.B * eb * ib, gf = (double)p.fe * eb * ib;
    const double Pa = al + 20.0 * k19 * ia, Pb = be + 20.0 * l19 * ib;
    const double Qa = rei * 20.0 * k18 * (19.0 - k20) * ia * ia, Qb = rei * 20.0 * l18 * (19.0 - l20) * ib * ib;
    const double ha = rei * (rei * Pa * Pa - Qa), hb = rei * (rei * Pb * Pb - Qb);
    EamD o;
    o.f = gf;
    o.f1 = -rei * gf * Pb;
    o.f2 = gf * hb;
    o.phi1 = -rei * (ga * Pa - gb * Pb);
    o.phi2 = ga * ha - gb * hb;
    return o;
}

// F'(rho) and F''(rho), three branches (SurfaceGrowth.cu:160-210).
__device__ __forceinline__ void embed_d(const DevParams &p, double rho, double &dF, double &ddF) {
    const double rhoe = p.rhoe, rhon = (double)0.85f * rhoe, rhoo = (double)1.15f * rhoe, rhoei = p.rhoei;
    if (rho < rhon) {
        const double rn = rhoei * (double)1.176471f, t = rho * rn - 1.0;
        dF = rn * ((double)p.Fn[1] + 2.0 * (double)p.Fn[2] * t + 3.0 * (double)p.Fn[3] * t * t);
        ddF = rn * rn * (2.0 * (double)p.Fn[2] + 6.0 * (double)p.Fn[3] * t);
        return;
    }
    if (rho < rhoo) {
        const double t = rho * rhoei - 1.0;
        dF = rhoei * ((double)p.F[1] + 2.0 * (double)p.F[2] * t + 3.0 * (double)p.F[3] * t * t);
        ddF = rhoei * rhoei * (2.0 * (double)p.F[2] + 6.0 * (double)p.F[3] * t);
        return;
    }
    // F = Fe (1 - eta ln x) x^eta, x = rho/rhoe: F' = -Fe eta^2 rhoei x^(eta-1) ln x,
    // F'' = -Fe eta^2 rhoei^2 x^(eta-2) ((eta-1) ln x + 1)
    const double x = rho * rhoei, eta = p.eta, lx = log(x), xe = exp(eta * lx);
    dF = -(double)p.Fe * eta * eta * rhoei * (xe / x) * lx;
    ddF = -(double)p.Fe * eta * eta * rhoei * rhoei * (xe / (x * x)) * ((eta - 1.0) * lx + 1.0);
}

// Angle factor T(c) = mu_theta (acos c - theta_0) / sqrt(1 - c^2) and T'(c) = (c T - mu_theta)/(1-c^2).
__device__ __forceinline__ void angle_t(const DevParams &p, double c, double &T, double &Tp) {
    const double s2 = 1.0 - c * c, mu = (double)p.mu_theta;
    T = mu * (acos(c) - (double)p.theta_0) / sqrt(s2);
    Tp = (c * T - mu) / s2;
}

// Neighbour-angle term of bond d = r_i - r_k (|d| = R) with e = r_m - r_k, rr2 = |e|^2
// (SurfaceGrowth.cu:953-968): c = (d.e)/(R rr2) (the reference divides by rr2), F = T(c) v with
// v = dc/dd = e/(R rr2) - c d/R^2.  For s = l.F: gd = ds/dd, ge = ds/de.
__device__ __forceinline__ void na_vjp(const DevParams &p, const double (&l)[3], const double (&d)[3], double R,
                                       const double (&e)[3], double rr2, double (&gd)[3], double (&ge)[3]) {
    const double iRr = 1.0 / (R * rr2), iR2 = 1.0 / (R * R);
    const double c = dot3(d, e) * iRr;
    double T, Tp;
    angle_t(p, c, T, Tp);
    const double P = dot3(l, d), E = dot3(l, e);
    const double Lv = E * iRr - c * P * iR2;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const double v = e[x] * iRr - c * d[x] * iR2;           // dc/dd
        const double ce = d[x] * iRr - 2.0 * c * e[x] / rr2;     // dc/de
        const double dLd = -E * d[x] * iRr * iR2 - v * P * iR2 - c * l[x] * iR2 + 2.0 * c * P * d[x] * iR2 * iR2;
        const double dLe = l[x] * iRr - 2.0 * E * e[x] * iRr / rr2 - P * iR2 * ce;
        gd[x] = Tp * Lv * v + T * dLd;
        ge[x] = Tp * Lv * ce + T * dLe;
    }
}

// Central angle of bonds da, db (SurfaceGrowth.cu:1016-1055): c = da.db/(Ra Rb),
// F = T(c) (dc/dda + dc/ddb).  For s = l.F: ga = ds/dda, gb = ds/ddb.
__device__ __forceinline__ void ang_vjp(const DevParams &p, const double (&l)[3], const double (&da)[3], double Ra,
                                        const double (&db)[3], double Rb, double (&ga)[3], double (&gb)[3]) {
    const double iab = 1.0 / (Ra * Rb), ia2 = 1.0 / (Ra * Ra), ib2 = 1.0 / (Rb * Rb);
    const double c = dot3(da, db) * iab;
    double T, Tp;
    angle_t(p, c, T, Tp);
    const double P = dot3(l, da), Q = dot3(l, db);
    const double K = P * ia2 + Q * ib2;
    const double Lu = (P + Q) * iab - c * K;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const double ca = db[x] * iab - c * da[x] * ia2;  // dc/dda
        const double cb = da[x] * iab - c * db[x] * ib2;  // dc/ddb
        const double dLa = l[x] * iab - (P + Q) * da[x] * iab * ia2 - ca * K - c * (l[x] * ia2 - 2.0 * P * da[x] * ia2 * ia2);
        const double dLb = l[x] * iab - (P + Q) * db[x] * iab * ib2 - cb * K - c * (l[x] * ib2 - 2.0 * Q * db[x] * ib2 * ib2);
        ga[x] = Tp * Lu * ca + T * dLa;
        gb[x] = Tp * Lu * cb + T * dLb;
    }
}

// ---- kernels ---------------------------------------------------------------------------------------
// Adhesion energy: block partial sums of the metal-side LJ pair energies.
__global__ __launch_bounds__(kBs) void k_adh_part(DevParams p, const double4 *__restrict__ pos,
                                                  const int *__restrict__ nlM, const int *__restrict__ nnM,
                                                  double *__restrict__ part) {
    const int i = blockIdx.x * kBs + threadIdx.x;
    double e = 0.0;
    if (i < p.nMolMe) {
        const double4 A = pos[i];
        const int *row = nlM + (size_t)i * p.nlCap;
        const int nn = nnM[i];
        for (int j = 0; j < nn; ++j) {
            const int k = row[j];
            if (k < p.nMolMe) continue;
            double d[3];
            const double rr = dispd(A, pos[k], p, d);
            if (rr < (double)p.rrCutLJ) e += lj_d(p, rr).U;
        }
    }
    const double s = block_sum(e);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Fixed-order total of the partials into out[st.rec - 1] (the record slot of the step just
// finished) or out[0].
__global__ void k_adh_fin(const double *__restrict__ part, int n, double *__restrict__ out, const DevState *st,
                          int recIndex) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += part[k];
    out[recIndex ? st->rec - 1 : 0] = s;
}

// FP64 density, F' and F'' of every metal atom (EamComputeRhoK, SurfaceGrowth.cu:716-773); G lanes
// per atom share its neighbour loop, as in the forward density pass.
template <int G>
__global__ __launch_bounds__(kBs) void k_adj_rho(DevParams p, const double4 *__restrict__ pos,
                                                 const int *__restrict__ nlM, const int *__restrict__ nnM,
                                                 double *__restrict__ rho, double *__restrict__ dF,
                                                 double *__restrict__ ddF) {
    const int gt = blockIdx.x * kBs + threadIdx.x, i = gt / G, l = gt % G;
    const bool active = i < p.nMolMe;
    double sum = 0.0;
    if (active) {
        const double4 A = pos[i];
        const int *row = nlM + (size_t)i * p.nlCap;
        const int nn = nnM[i];
        for (int j = l; j < nn; j += G) {
            const int k = row[j];
            if (k >= p.nMolMe) continue;
            double d[3];
            const double rr = dispd(A, pos[k], p, d);
            if (rr < (double)p.rrCutEam) sum += eam_d(p, sqrt(rr)).f;
        }
    }
    sum = group_sum<G>(sum);  // every lane of the wave takes part
    if (active && l == 0) {
        rho[i] = sum;
        embed_d(p, sum, dF[i], ddF[i]);
    }
}

struct AdjArgs {
    const double4 *pos;
    const int *nlM, *nnM, *nlj;
    const unsigned *lj;
    const uint4 *bondC, *bondA;
    AdjDev d;
};

__device__ void metal_pairs(const DevParams &p, int i, int gl, int stride, const AdjArgs &a, double wE,
                            double (&pb)[4], double (&gi)[3], double &rb);

// Metal atom i: EAM pairs and LJ with carbon (SurfaceGrowth.cu:856-896).  G lanes per atom: lane
// gl takes list entries gl, gl+G, ...; the atom's own derivative and density adjoint are summed over
// its lanes (every lane of the wave calls this).
template <int G>
__device__ void metal_vjp(const DevParams &p, int i, int gl, bool active, const AdjArgs &a, double wE,
                          double (&pb)[4]) {
    const int N = p.nMol;
    double gi[3] = {0.0, 0.0, 0.0}, rb = 0.0;
    if (active) metal_pairs(p, i, gl, G, a, wE, pb, gi, rb);
    for (int x = 0; x < 3; ++x) gi[x] = group_sum<G>(gi[x]);
    rb = group_sum<G>(rb);
    if (active && gl == 0) {
        scatter(a.d.G, N, i, 1.0, gi);
        if (rb != 0.0) atomicAdd(&a.d.rhoBar[i], rb);
    }
}

// The pair loop of metal_vjp over list entries gl, gl+stride, ...
__device__ void metal_pairs(const DevParams &p, int i, int gl, int stride, const AdjArgs &a, double wE,
                            double (&pb)[4], double (&gi)[3], double &rb) {
    const int N = p.nMol;
    const double im = 1.0 / (double)p.massMe;
    const double l[3] = {a.d.Lam[i] * im, a.d.Lam[N + i] * im, a.d.Lam[2 * N + i] * im};
    const bool hasL = l[0] != 0.0 || l[1] != 0.0 || l[2] != 0.0;
    if (!hasL && wE == 0.0) return;
    const double4 A = a.pos[i];
    const double Fi1 = a.d.dF[i], Fi2 = a.d.ddF[i];
    const int *row = a.nlM + (size_t)i * p.nlCap;
    const int nn = a.nnM[i];
    for (int j = gl; j < nn; j += stride) {
        const int k = row[j];
        double d[3];
        const double rr = dispd(A, a.pos[k], p, d);
        double g[3];
        if (k < p.nMolMe) {
            if (!hasL || !(rr < (double)p.rrCutEam)) continue;
            // a_i += c d / m, c = -(phi' + f' S)/r, S = F'(rho_i) + F'(rho_k)
            const double r = sqrt(rr);
            const EamD e = eam_d(p, r);
            const double S = Fi1 + a.d.dF[k];
            const double c = -(e.phi1 + e.f1 * S) / r;
            const double dc = (-(e.phi2 + e.f2 * S) - c) / r;
            const double P = dot3(l, d), q = dc * P / r;
            for (int x = 0; x < 3; ++x) g[x] = c * l[x] + q * d[x];
            const double t = -P * e.f1 / r;  // d s / d S
            rb += t * Fi2;
            atomicAdd(&a.d.rhoBar[k], t * a.d.ddF[k]);
            pb[2] += t * S;  // F' = s_F F0'
            pb[3] += t * S;  // f' = s_rho f0'
        } else {
            if (!(rr < (double)p.rrCutLJ)) continue;
            const LjD L = lj_d(p, rr);
            const double P = dot3(l, d);
            for (int x = 0; x < 3; ++x) g[x] = L.fc * l[x] + 2.0 * L.fcp * P * d[x] - wE * L.fc * d[x];
            pb[0] += P * L.fc + wE * L.U;
            pb[1] += P * L.sfc + wE * L.sU;
        }
        for (int x = 0; x < 3; ++x) gi[x] += g[x];
        scatter(a.d.G, N, k, -1.0, g);
    }
}

// Carbon atom i: bonds, plane, angle terms and LJ with metal (SurfaceGrowth.cu:899-1062), the
// structure of the force pass's carbon_body: bonds and near neighbours from bondC/bondA, the LJ
// multiplicity of the reused loop counter.
__device__ void carbon_vjp(const DevParams &p, int ic, const AdjArgs &a, double (&pb)[4]) {
    const int N = p.nMol, i = p.nMolMe + ic;
    const double l[3] = {a.d.Lam[i], a.d.Lam[N + i], a.d.Lam[2 * N + i]};
    if (l[0] == 0.0 && l[1] == 0.0 && l[2] == 0.0) return;  // frozen, or no cotangent
    const double4 A = a.pos[i];
    const uint4 nbi = a.bondC[ic];
    const int counti = (int)nbi.w;
    const unsigned be[3] = {nbi.x, nbi.y, nbi.z};
    const int j0 = counti == 3 ? (int)(nbi.z & 255u) : -1;
    int kb[3];
    double4 Bk[3];
    double d[3][3], R[3], gk[3][3];
    double gi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        kb[b] = b < counti ? (int)(be[b] >> 8) : i;
        Bk[b] = a.pos[kb[b]];
        R[b] = 1.0;
        for (int x = 0; x < 3; ++x) gk[b][x] = d[b][x] = 0.0;
        if (b < counti) R[b] = sqrt(dispd(A, Bk[b], p, d[b]));
    }
    const double mup = (double)p.mu_p;
    // neighbour-angle terms: k's first two near neighbours m != i (SurfaceGrowth.cu:940-978)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        if (b >= counti) break;
        const uint4 nbk = a.bondA[kb[b] - p.nMolMe];
        const unsigned me[3] = {nbk.x, nbk.y, nbk.z};
        const int na = (int)nbk.w;
        int countk = 0;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            if (t >= na || countk >= 2 || (int)(me[t] >> 8) == i) continue;
            const int m = (int)(me[t] >> 8);
            double e[3];
            const double rr2 = dispd(Bk[b], a.pos[m], p, e);
            for (int x = 0; x < 3; ++x) e[x] = -e[x];  // r_m - r_k
            double gd[3], ge[3];
            na_vjp(p, l, d[b], R[b], e, rr2, gd, ge);
            for (int x = 0; x < 3; ++x) {
                gi[x] += gd[x];
                gk[b][x] -= gd[x] + ge[x];
            }
            ge[2] += -mup * l[2] / 9.0;  // F_z += -mu_p (z_m - z_0)/9
            scatter(a.d.G, N, m, 1.0, ge);
            ++countk;
        }
    }
    if (counti == 3) {  // SurfaceGrowth.cu:985-1061
        const double mur = (double)p.mu_r, c23 = (double)0.666666667f;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            // F = -mu_r (R-1)/R d
            const double g = -mur * (1.0 - 1.0 / R[b]), gp = -mur / (R[b] * R[b]);
            const double P = dot3(l, d[b]);
            for (int x = 0; x < 3; ++x) {
                const double v = g * l[x] + gp * P * d[b][x] / R[b];
                gi[x] += v;
                gk[b][x] -= v;
            }
            // F_z += -(2/3) mu_p (2 (z_i - z_0) - sum_b (z_kb - z_0))
            gk[b][2] += c23 * mup * l[2];
        }
        gi[2] += -2.0 * c23 * mup * l[2];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int ia = q, ib = q == 2 ? 0 : q + 1;
            double ga[3], gb[3];
            ang_vjp(p, l, d[ia], R[ia], d[ib], R[ib], ga, gb);
            for (int x = 0; x < 3; ++x) {
                gi[x] += ga[x] + gb[x];
                gk[ia][x] -= ga[x];
                gk[ib][x] -= gb[x];
            }
        }
    }
    // LJ with metal: entries at list positions <= j0 count once, 4..j0 twice, and with j0 = 2
    // position 3 not at all (SurfaceGrowth.cu:989)
    const int nl = a.nlj[ic];
    for (int qq = 0; qq < nl; ++qq) {
        const unsigned e = a.lj[(size_t)qq * p.nC + ic];
        const int k = (int)(e >> 8), pk = (int)(e & 255u);
        const int mult = j0 < 0 ? 1 : (int)(pk <= j0) + (int)(pk >= 4);
        if (mult == 0) continue;
        double dk[3];
        const double rr = dispd(A, a.pos[k], p, dk);
        if (!(rr < (double)p.rrCutLJ)) continue;
        const LjD L = lj_d(p, rr);
        const double P = dot3(l, dk);
        double g[3];
        for (int x = 0; x < 3; ++x) g[x] = mult * (L.fc * l[x] + 2.0 * L.fcp * P * dk[x]);
        for (int x = 0; x < 3; ++x) gi[x] += g[x];
        scatter(a.d.G, N, k, -1.0, g);
        pb[0] += mult * P * L.fc;
        pb[1] += mult * P * L.sfc;
    }
    scatter(a.d.G, N, i, 1.0, gi);
#pragma unroll
    for (int b = 0; b < 3; ++b)
        if (b < counti) scatter(a.d.G, N, kb[b], 1.0, gk[b]);
}

// blocks [0, nbM): metal atoms, G lanes each; [nbM, nbM + nbC): carbon atoms, one thread each.
template <int G>
__global__ __launch_bounds__(kBs) void k_adj_force(DevParams p, AdjArgs a, int nbM, double wE) {
    double pb[4] = {0.0, 0.0, 0.0, 0.0};
    if ((int)blockIdx.x < nbM) {
        const int gt = blockIdx.x * kBs + threadIdx.x, i = gt / G;
        metal_vjp<G>(p, i, gt % G, i < p.nMolMe, a, wE, pb);
    } else {
        const int ic = (blockIdx.x - nbM) * kBs + threadIdx.x;
        if (ic < p.nC) carbon_vjp(p, ic, a, pb);
    }
    add_pbar(a.d.pbar, pb[0], pb[1], pb[2], pb[3]);
}

// Density adjoint to positions: rho_i = sum_k f(r_ik), so d/dr_i += rhoBar_i f'(r) d/r and
// d/dr_k -= the same; and rho = s_rho rho0 gives pbar_density += rhoBar_i rho_i.
template <int G>
__global__ __launch_bounds__(kBs) void k_adj_rho2(DevParams p, AdjArgs a) {
    const int gt = blockIdx.x * kBs + threadIdx.x, i = gt / G, l = gt % G;
    const bool active = i < p.nMolMe;
    const int N = p.nMol;
    double p3 = 0.0, gi[3] = {0.0, 0.0, 0.0};
    if (active) {
        const double rb = a.d.rhoBar[i];
        if (l == 0) p3 = rb * a.d.rho[i];
        if (rb != 0.0) {
            const double4 A = a.pos[i];
            const int *row = a.nlM + (size_t)i * p.nlCap;
            const int nn = a.nnM[i];
            for (int j = l; j < nn; j += G) {
                const int k = row[j];
                if (k >= p.nMolMe) continue;
                double d[3];
                const double rr = dispd(A, a.pos[k], p, d);
                if (!(rr < (double)p.rrCutEam)) continue;
                const double r = sqrt(rr);
                const double s = rb * eam_d(p, r).f1 / r;
                double g[3] = {s * d[0], s * d[1], s * d[2]};
                for (int x = 0; x < 3; ++x) gi[x] += g[x];
                scatter(a.d.G, N, k, -1.0, g);
            }
        }
    }
    for (int x = 0; x < 3; ++x) gi[x] = group_sum<G>(gi[x]);
    if (active && l == 0) scatter(a.d.G, N, i, 1.0, gi);
    add_pbar(a.d.pbar, 0.0, 0.0, 0.0, p3);
}

// Thermostat transpose, part 1: block sums of sum m|v|^2, sum_metal vx, sum_carbon Vb.v and
// sum_metal Vb.v over the velocities v before the thermostat.
__global__ __launch_bounds__(kBs) void k_adj_therm_part(DevParams p, const double *__restrict__ vpre,
                                                        const double *__restrict__ Vb, double *__restrict__ part) {
    const int N = p.nMol, n = blockIdx.x * kBs + threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (n < N) {
        const double vx = vpre[n], vy = vpre[N + n], vz = vpre[2 * N + n];
        const double v2 = vx * vx + vy * vy + vz * vz;
        const double dt = Vb[n] * vx + Vb[N + n] * vy + Vb[2 * N + n] * vz;
        if (n < p.nMolMe) {
            s[0] = (double)p.massMe * v2;
            s[1] = vx;
            s[3] = dt;
        } else {
            s[0] = v2;
            s[2] = dt;
        }
    }
    for (int q = 0; q < 4; ++q) {
        const double t = block_sum(s[q]);
        if (threadIdx.x == 0) part[blockIdx.x * 4 + q] = t;
    }
}

// Part 2 (one thread): the Berendsen factors of ApplyBerendsenThermostat (SurfaceGrowth.cu:455-485)
// from vv = sum m|v - vcm_x|^2 (regime 2 removes the metal centre-of-mass x velocity), and
// S = sum_n (Vb_n . v_n) d beta_n / d vv.  therm = {beta_C, beta_M, S, vcm_x}.
__global__ void k_adj_therm_fin(DevParams p, const double *__restrict__ part, int nParts, double *__restrict__ therm,
                                int step, int metalAll) {
    if (threadIdx.x != 0) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < nParts; ++k)
        for (int q = 0; q < 4; ++q) s[q] += part[k * 4 + q];
    const int nMe = p.nMolMe;
    const double m = (double)p.massMe;
    const double c = nMe ? s[1] / nMe : 0.0;
    double vv = s[0];
    if (p.iRegime == 2 && nMe) vv = vv - m * c * s[1];
    const double ke = vv * (double)0.5f / (double)p.nMol;
    const double g = (double)p.gamma;
    const double c0C = (double)p.temperature * (double)1.5f * (double)p.kB, c0M = (double)p.temperature * (double)p.kB;
    const double bC = sqrt(1.0 + g * (c0C / ke - 1.0)), bM = sqrt(1.0 + g * (c0M / ke - 1.0));
    const double dke = 0.5 / (double)p.nMol;  // d ke / d vv
    const double dbC = -g * c0C / (2.0 * bC * ke * ke) * dke, dbM = -g * c0M / (2.0 * bM * ke * ke) * dke;
    const bool metalOn = p.iRegime == 2 && step > p.stepEquil && (metalAll || step < p.stepEquil + p.stepCool);
    therm[0] = bC;
    therm[1] = metalOn ? bM : 1.0;
    therm[2] = s[2] * dbC + (metalOn ? s[3] * dbM : 0.0);
    therm[3] = p.iRegime == 2 ? c : 0.0;
}

// Before the force transpose of step n: W = adjoint of the velocities before the thermostat
// (beta Vb + S dvv/dv on thermostat steps, Vb otherwise), Lam = mask (Ab + hdt W); clears G, rhoBar.
__global__ __launch_bounds__(kBs) void k_adj_pre(DevParams p, const double4 *__restrict__ pos, AdjDev d,
                                                 const double *__restrict__ vpre) {
    const int N = p.nMol, n = blockIdx.x * kBs + threadIdx.x;
    if (n >= N) return;
    const bool f = moves(p, n, pos[n]);
    const double hdt = (double)(0.5f * p.deltaT);
    double w[3] = {d.Vb[n], d.Vb[N + n], d.Vb[2 * N + n]};
    if (vpre) {
        const double beta = n >= p.nMolMe ? d.therm[0] : d.therm[1], S = d.therm[2];
        const double v[3] = {vpre[n], vpre[N + n], vpre[2 * N + n]};
        const double m = n < p.nMolMe ? (double)p.massMe : 1.0;
        const double vx = n < p.nMolMe ? v[0] - d.therm[3] : v[0];
        const double gv[3] = {2.0 * m * vx, 2.0 * m * v[1], 2.0 * m * v[2]};
        for (int x = 0; x < 3; ++x) w[x] = beta * w[x] + S * gv[x];
    }
    for (int x = 0; x < 3; ++x) {
        d.W[x * N + n] = w[x];
        d.Lam[x * N + n] = f ? d.Ab[x * N + n] + hdt * w[x] : 0.0;
        d.G[x * N + n] = 0.0;
    }
    if (n < p.nMolMe) d.rhoBar[n] = 0.0;
}

// After it: Rb += G; Vb = W + mask dt Rb; Ab = mask hdt Vb  (the transpose of leapfrog part 1).
__global__ __launch_bounds__(kBs) void k_adj_post(DevParams p, const double4 *__restrict__ pos, AdjDev d) {
    const int N = p.nMol, n = blockIdx.x * kBs + threadIdx.x;
    if (n >= N) return;
    const bool f = moves(p, n, pos[n]);
    const double hdt = (double)(0.5f * p.deltaT), dt = (double)p.deltaT;
    for (int x = 0; x < 3; ++x) {
        const double rb = d.Rb[x * N + n] + d.G[x * N + n];
        const double vb = d.W[x * N + n] + (f ? dt * rb : 0.0);
        d.Rb[x * N + n] = rb;
        d.Vb[x * N + n] = vb;
        d.Ab[x * N + n] = f ? hdt * vb : 0.0;
    }
}

// The initial acceleration a_0 = a(r_0): cotangent Ab, clears G, rhoBar.
__global__ __launch_bounds__(kBs) void k_adj_lam0(DevParams p, AdjDev d) {
    const int N = p.nMol, n = blockIdx.x * kBs + threadIdx.x;
    if (n >= N) return;
    for (int x = 0; x < 3; ++x) {
        d.Lam[x * N + n] = d.Ab[x * N + n];
        d.G[x * N + n] = 0.0;
    }
    if (n < p.nMolMe) d.rhoBar[n] = 0.0;
}

// Terminal adjoint: d (wH * height) / d r_W = wH / nMolMe on the metal z; zero elsewhere.
__global__ __launch_bounds__(kBs) void k_adj_seed(DevParams p, AdjDev d, double wH) {
    const int N = p.nMol, n = blockIdx.x * kBs + threadIdx.x;
    if (n == 0)
        for (int k = 0; k < 4; ++k) d.pbar[k] = 0.0;
    if (n >= N) return;
    for (int x = 0; x < 3; ++x) {
        d.Rb[x * N + n] = (x == 2 && n < p.nMolMe) ? wH / (double)p.nMolMe : 0.0;
        d.Vb[x * N + n] = 0.0;
        d.Ab[x * N + n] = 0.0;
    }
}

// Bitwise comparison of the recomputed state with a stored one.
__global__ __launch_bounds__(kBs) void k_adj_cmp(DevParams p, const double4 *__restrict__ pos,
                                                 const double *__restrict__ vel, const double *__restrict__ acc,
                                                 const double4 *__restrict__ pos2, const double *__restrict__ vel2,
                                                 const double *__restrict__ acc2, int *__restrict__ mismatch) {
    const int N = p.nMol, n = blockIdx.x * kBs + threadIdx.x;
    if (n >= N) return;
    const double4 a = pos[n], b = pos2[n];
    bool same = __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
                __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
                __double_as_longlong(a.z) == __double_as_longlong(b.z) &&
                __double_as_longlong(a.w) == __double_as_longlong(b.w);
    for (int x = 0; x < 3; ++x) {
        same = same && __double_as_longlong(vel[x * N + n]) == __double_as_longlong(vel2[x * N + n]);
        same = same && __double_as_longlong(acc[x * N + n]) == __double_as_longlong(acc2[x * N + n]);
    }
    if (!same) atomicAdd(mismatch, 1);
}

inline unsigned blocks(long long n) { return (unsigned)((n + kBs - 1) / kBs); }

AdjArgs adj_args(sgk_ctx *c, const AdjDev &d) {
    AdjArgs a;
    a.pos = (const double4 *)c->pos;
    a.nlM = c->nlM;
    a.nnM = c->nnM;
    a.nlj = c->nlj;
    a.lj = c->lj;
    a.bondC = c->bondC;
    a.bondA = c->bondA;
    a.d = d;
    return a;
}

}  // namespace

// ---- launchers ---------------------------------------------------------------------------------------
void launch_adh_obs(sgk_ctx *c, double *part, double *out, int recIndex) {
    const unsigned g = blocks(std::max(c->dp.nMolMe, 1));
    hipLaunchKernelGGL(k_adh_part, dim3(g), dim3(kBs), 0, c->stream, c->dp, (const double4 *)c->pos, c->nlM, c->nnM,
                       part);
    hipLaunchKernelGGL(k_adh_fin, dim3(1), dim3(64), 0, c->stream, part, (int)g, out, c->state, recIndex);
}

// Lanes per metal atom of the metal kernels: the context's forward choice (sgk_config.metal_lanes).
#define SGK_ADJ_G(call)            \
    switch (c->metalLanes) {       \
        case 1: call(1); break;    \
        case 2: call(2); break;    \
        case 4: call(4); break;    \
        case 16: call(16); break;  \
        default: call(8); break;   \
    }

void launch_adj_rho(sgk_ctx *c, const AdjDev &d) {
    if (c->dp.nMolMe == 0) return;
    const unsigned g = blocks((long long)c->dp.nMolMe * c->metalLanes);
#define K(G)                                                                                                 \
    hipLaunchKernelGGL(k_adj_rho<G>, dim3(g), dim3(kBs), 0, c->stream, c->dp, (const double4 *)c->pos, c->nlM, \
                       c->nnM, d.rho, d.dF, d.ddF)
    SGK_ADJ_G(K)
#undef K
}

void launch_adj_force(sgk_ctx *c, const AdjDev &d, double wE) {
    const int nbM = (int)blocks((long long)c->dp.nMolMe * c->metalLanes), nbC = (int)blocks(c->dp.nC);
    if (nbM + nbC == 0) return;
#define K(G) \
    hipLaunchKernelGGL(k_adj_force<G>, dim3(nbM + nbC), dim3(kBs), 0, c->stream, c->dp, adj_args(c, d), nbM, wE)
    SGK_ADJ_G(K)
#undef K
}

void launch_adj_rho2(sgk_ctx *c, const AdjDev &d) {
    if (c->dp.nMolMe == 0) return;
    const unsigned g = blocks((long long)c->dp.nMolMe * c->metalLanes);
#define K(G) hipLaunchKernelGGL(k_adj_rho2<G>, dim3(g), dim3(kBs), 0, c->stream, c->dp, adj_args(c, d))
    SGK_ADJ_G(K)
#undef K
}

void launch_adj_therm(sgk_ctx *c, const AdjDev &d, const double *velPre, int step) {
    const unsigned g = blocks(c->dp.nMol);
    hipLaunchKernelGGL(k_adj_therm_part, dim3(g), dim3(kBs), 0, c->stream, c->dp, velPre, d.Vb, d.part);
    hipLaunchKernelGGL(k_adj_therm_fin, dim3(1), dim3(64), 0, c->stream, c->dp, d.part, (int)g, d.therm, step,
                       c->cfg.metal_thermostat_all);
}

void launch_adj_pre(sgk_ctx *c, const AdjDev &d, const double *velPre) {
    hipLaunchKernelGGL(k_adj_pre, dim3(blocks(c->dp.nMol)), dim3(kBs), 0, c->stream, c->dp, (const double4 *)c->pos, d,
                       velPre);
}

void launch_adj_post(sgk_ctx *c, const AdjDev &d) {
    hipLaunchKernelGGL(k_adj_post, dim3(blocks(c->dp.nMol)), dim3(kBs), 0, c->stream, c->dp, (const double4 *)c->pos,
                       d);
}

void launch_adj_lam0(sgk_ctx *c, const AdjDev &d) {
    hipLaunchKernelGGL(k_adj_lam0, dim3(blocks(c->dp.nMol)), dim3(kBs), 0, c->stream, c->dp, d);
}

void launch_adj_seed(sgk_ctx *c, const AdjDev &d, double wHeight) {
    hipLaunchKernelGGL(k_adj_seed, dim3(blocks(c->dp.nMol)), dim3(kBs), 0, c->stream, c->dp, d, wHeight);
}

void launch_adj_cmp(sgk_ctx *c, const AdjDev &d, const void *pos, const void *vel, const void *acc) {
    hipLaunchKernelGGL(k_adj_cmp, dim3(blocks(c->dp.nMol)), dim3(kBs), 0, c->stream, c->dp, (const double4 *)c->pos,
                       (const double *)c->vel, (const double *)c->acc, (const double4 *)pos, (const double *)vel,
                       (const double *)acc, d.mismatch);
}

}  // namespace sgk
