// sgk_force.hip - EAM density and the force pass for gfx950.
//
// The reference evaluates everything in one kernel with one thread per atom and one pass over the
// neighbour list (ComputeForcesK, SurfaceGrowth.cu:776-1091), after EamComputeRhoK
// (SurfaceGrowth.cu:716-773).  Here:
//   k_rho    blocks [0, nbRho): G lanes per metal atom split its neighbour list; rho, F(rho),
//            F'(rho) stored, so the force pass reads F'(rho_k) instead of recomputing it per pair.
//            blocks [nbRho, nbRho + nbC): one thread per carbon scans its close list once and stores
//            the first three carbons within rCutC (its bonds) and the first three atoms of any kind
//            within rCutC (what a bonded neighbour's angle terms need), so the force pass reads
//            four near lists instead of scanning four close lists.
//   k_force  one launch: blocks [0, nbM) are the metal pass (G lanes per atom: EAM pairs + LJ with
//            carbon), blocks [nbM, nbM + nbC) the carbon pass (one thread per atom: bond, plane and
//            angle terms from the close list + LJ with metal).  The epilogue writes acceleration,
//            energy, friction, applies leapfrog part 2 and leaves one Partial per block; k_finalize
//            (sgk_step.hip), the next launch, reduces them and runs the step bookkeeping.  (A
//            last-block reduction inside k_force needs a device-scope release fence per block, which
//            on gfx950 writes back and invalidates the XCD's L2 - it cost more than the pass itself.)
// Precision: displacement vectors, cutoff tests and, for the stiff graphene terms, the bond-length
// and angle deviations are formed in the storage precision R (FP64 in the mixed form).  EAM and LJ
// pair arithmetic runs in T (FP32, or FP64 with pair_fp64); graphene force-law arithmetic is FP32.
// Gathers are issued in chunks of predicated loads so several are in flight per thread.
#include "sgk_batch.h"
#include "sgk_device.h"
#include "sgk_internal.h"
#include "sgk_reduce.h"

namespace sgk {

// Displacement and its square in storage precision; cutoff decisions are made on this value.
// The single-image correction changes nothing unless some component reaches half the box, so it
// runs only then: three compares per pair on the common path, and a branch that a wave skips
// unless one of its pairs crosses a periodic face.  The result is the same in every case.
template <class R, class P4>
__device__ __forceinline__ R disp(const P4 &a, const P4 &b, const DevParams &p, R &dx, R &dy, R &dz) {
#pragma clang fp contract(off)
    dx = a.x - b.x;
    dy = a.y - b.y;
    dz = a.z - b.z;
    if (fabs(dx) >= R(0.5f) * (R)p.Lx || fabs(dy) >= R(0.5f) * (R)p.Ly || fabs(dz) >= R(0.5f) * (R)p.Lz) {
        dx = wrap1<R>(dx, (R)p.Lx);
        dy = wrap1<R>(dy, (R)p.Ly);
        dz = wrap1<R>(dz, (R)p.Lz);
    }
    return dx * dx + dy * dy + dz * dz;
}

constexpr int kChunk = 8;  // neighbours gathered per round

// N consecutive list entries (N a multiple of 4, 16-byte loads; rows are 32-byte aligned); entries
// at or past `valid` are replaced by `pad` so they index a real atom.
template <int N>
__device__ __forceinline__ void load_chunk(const unsigned *row, int valid, unsigned pad, unsigned (&ent)[N]) {
#pragma unroll
    for (int h = 0; h < N / 4; ++h) {
        const uint4 v = *(const uint4 *)(row + 4 * h);
        ent[4 * h] = v.x;
        ent[4 * h + 1] = v.y;
        ent[4 * h + 2] = v.z;
        ent[4 * h + 3] = v.w;
    }
#pragma unroll
    for (int t = 0; t < N; ++t)
        if (t >= valid) ent[t] = pad;
}

// EamComputeRhoK (SurfaceGrowth.cu:716-773) with G lanes per atom; pair math in T.
template <bool M, class T, int G>
__device__ __forceinline__ void rho_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                         const int *__restrict__ nlM, const int *__restrict__ nnM,
                                         float *__restrict__ rho, float *__restrict__ embF,
                                         float *__restrict__ embDF) {
    using R = typename Prec<M>::R;
    using P4 = typename Prec<M>::P4;
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / G, l = gt % G;
    const bool active = i < p.nMolMe;
    T sum = T(0);
    if (active) {
        const P4 A = pos[i];
        const int nn = nnM[i];
        const int *row = nlM + (size_t)i * p.nlCap;
        // gathers per round: 2 with FP64 pair math (88 VGPRs, 5 waves per SIMD), 4 with FP32; lane l
        // sums entries l, l + G, l + 2G, ... in order either way
        constexpr int MC = sizeof(T) == 8 ? 2 : 4;
        int kn[MC];  // the next round's list entries, loaded a round ahead of their position gathers
#pragma unroll
        for (int q = 0; q < MC; ++q) {
            const int j = l + q * G;
            kn[q] = j < nn ? row[j] : i;
        }
        for (int base = l; base < nn; base += G * MC) {
            int k[MC];
            P4 B[MC];
#pragma unroll
            for (int q = 0; q < MC; ++q) k[q] = kn[q];
#pragma unroll
            for (int q = 0; q < MC; ++q) B[q] = pos[k[q]];
#pragma unroll
            for (int q = 0; q < MC; ++q) {
                const int j = base + G * MC + q * G;
                kn[q] = j < nn ? row[j] : i;
            }
#pragma unroll
            for (int q = 0; q < MC; ++q) {
                if (k[q] >= p.nMolMe || base + q * G >= nn) continue;
                R dx, dy, dz;
                const R rr = disp<R>(A, B[q], p, dx, dy, dz);
                if (rr < (R)p.rrCutEam) sum += eam_f<T>(p, PM<T>::sqrt((T)rr));
            }
        }
    }
    sum = group_sum<G>(sum);
    if (active && l == 0) {
        T dF;
        rho[i] = (float)sum;
        embF[i] = (float)eam_embed<T>(p, sum, dF);
        embDF[i] = (float)dF;
    }
}

// Near lists of one carbon from its close list at the current positions (the bond search of
// SurfaceGrowth.cu:917-931 and the neighbour search of 940-951, cutoff tests in storage precision):
//   bondC = {first three carbons with r^2 < rrCutC, as k<<8 | list position; count}
//   bondA = {first three atoms of any kind with r^2 < rrCutC; count}
// The close list holds every neighbour-list entry within rCutC (the displacement bound keeps it
// exact), in list order, so these are the entries the reference's loops find.
template <bool M>
__device__ __forceinline__ void bonds_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                           const unsigned *__restrict__ cl, const int *__restrict__ ncl,
                                           uint4 *__restrict__ bondC, uint4 *__restrict__ bondA) {
    using R = typename Prec<M>::R;
    using P4 = typename Prec<M>::P4;
    const int ic = blk * blockDim.x + threadIdx.x;
    if (ic >= p.nC) return;
    const int i = p.nMolMe + ic;
    const P4 Ai = pos[i];
    unsigned cb[3] = {0u, 0u, 0u}, ab[3] = {0u, 0u, 0u};
    int nc = 0, na = 0;
    const int nclose = ncl[ic];
    // 4 entries a round with FP64 positions, 8 with FP32 (register pressure of the density launch)
    constexpr int CB = sizeof(R) == 8 ? 4 : kChunk;
    for (int base = 0; base < nclose && (nc < 3 || na < 3); base += CB) {
        unsigned ent[CB];
        P4 B[CB];
        load_chunk(cl + (size_t)ic * p.clCap + base, nclose - base, (unsigned)i << 8, ent);
#pragma unroll
        for (int t = 0; t < CB; ++t) B[t] = pos[ent[t] >> 8];
#pragma unroll
        for (int t = 0; t < CB; ++t) {
            if (base + t >= nclose) continue;
            R dx, dy, dz;
            const R rr = disp<R>(Ai, B[t], p, dx, dy, dz);
            if (!(rr < (R)p.rrCutC)) continue;
            if (na < 3) ab[na++] = ent[t];
            if (nc < 3 && (int)(ent[t] >> 8) >= p.nMolMe) cb[nc++] = ent[t];
        }
    }
    bondC[ic] = make_uint4(cb[0], cb[1], cb[2], (unsigned)nc);
    bondA[ic] = make_uint4(ab[0], ab[1], ab[2], (unsigned)na);
}

// One workgroup of the density launch: blocks [0, nbRho) metal density, the rest carbon near lists.
template <bool M, class T, int G>
__device__ __forceinline__ void rho_block(const DevParams &p, int b, int nbRho, int nb, bool xm,
                                          const typename Prec<M>::P4 *__restrict__ pos, const int *__restrict__ nlM,
                                          const int *__restrict__ nnM, float *__restrict__ rho,
                                          float *__restrict__ embF, float *__restrict__ embDF,
                                          const unsigned *__restrict__ cl, const int *__restrict__ ncl,
                                          uint4 *__restrict__ bondC, uint4 *__restrict__ bondA) {
    if (b < nbRho)
        rho_body<M, T, G>(p, xm ? xcd_block(b, nbRho) : b, pos, nlM, nnM, rho, embF, embDF);
    else
        bonds_body<M>(p, xm ? xcd_block(b - nbRho, nb - nbRho) : b - nbRho, pos, cl, ncl, bondC, bondA);
}

template <bool M, class T, int G>
__global__ __launch_bounds__(256) void k_rho(DevParams p, int nbRho, const typename Prec<M>::P4 *__restrict__ pos,
                                             const int *__restrict__ nlM, const int *__restrict__ nnM,
                                             float *__restrict__ rho, float *__restrict__ embF,
                                             float *__restrict__ embDF, const unsigned *__restrict__ cl,
                                             const int *__restrict__ ncl, uint4 *__restrict__ bondC,
                                             uint4 *__restrict__ bondA) {
    // nbRho is a multiple of 8: both ranges keep the XCD of each block
    rho_block<M, T, G>(p, blockIdx.x, nbRho, gridDim.x, p.xcdMap & 2, pos, nlM, nnM, rho, embF, embDF, cl, ncl,
                       bondC, bondA);
}

template <bool M, class T, int G>
__global__ __launch_bounds__(256) void kb_rho(BatchArgs a) {
    const int2 e = a.forceTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const DevParams &p = B.p;
    rho_block<M, T, G>(p, e.y, B.nbM, B.nbForce, false, (const typename Prec<M>::P4 *)a.pos + B.atomOff,
                       a.nlM + B.metalOff * p.nlCap, a.nnM + B.metalOff, a.rho + B.metalOff, a.embF + B.metalOff,
                       a.embDF + B.metalOff, a.cl + B.carbonOff * p.clCap, a.ncl + B.carbonOff,
                       a.bondC + B.carbonOff, a.bondA + B.carbonOff);
}

// ---- metal pass -----------------------------------------------------------------------------
// ComputeForcesK metal branch (SurfaceGrowth.cu:856-896, 1069-1082), G lanes per atom; pair math
// and per-atom sums in T.  Returns the atom's contribution to the block reductions.
template <bool M, class T, int G>
__device__ __forceinline__ Partial<typename Prec<M>::A> metal_body(
    const DevParams &p, int blk, typename Prec<M>::P4 *__restrict__ pos, typename Prec<M>::P4 *__restrict__ posOut,
    typename Prec<M>::R *__restrict__ vel, typename Prec<M>::R *__restrict__ acc, const int *__restrict__ nlM,
    const int *__restrict__ nnM, const float *__restrict__ embF, const float *__restrict__ embDF,
    float *__restrict__ fric, int integrate) {
    using R = typename Prec<M>::R;
    using A = typename Prec<M>::A;
    using P4 = typename Prec<M>::P4;
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / G, l = gt % G;
    const bool active = i < p.nMolMe;
    T fx = T(0), fy = T(0), fz = T(0), frx = T(0), u = T(0), w = T(0);
    P4 Ai;
    if (active) {
        Ai = pos[i];
        const T dFi = (T)embDF[i];
        const int nn = nnM[i];
        const int *row = nlM + (size_t)i * p.nlCap;
        // Gathers in flight per lane: 4 with FP32 pair math; 2 with FP64, whose pair temporaries
        // would otherwise push the kernel past 128 VGPRs (3 waves per SIMD instead of 4).  Lane l
        // takes entries l, l + G, l + 2G, ... in order either way, so the sums do not change.
        constexpr int MC = sizeof(T) == 8 ? 2 : 4;
        for (int base = l; base < nn; base += G * MC) {
            int k[MC];
            P4 B[MC];
            float dFk[MC];
#pragma unroll
            for (int q = 0; q < MC; ++q) {
                const int j = base + q * G;
                k[q] = j < nn ? row[j] : i;
            }
#pragma unroll
            for (int q = 0; q < MC; ++q) {
                B[q] = pos[k[q]];
                dFk[q] = k[q] < p.nMolMe ? embDF[k[q]] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < MC; ++q) {
                if (base + q * G >= nn) continue;
                R dxr, dyr, dzr;
                const R rrr = disp<R>(Ai, B[q], p, dxr, dyr, dzr);
                const T dx = (T)dxr, dy = (T)dyr, dz = (T)dzr, rr = (T)rrr;
                if (k[q] < p.nMolMe) {
                    if (rrr < (R)p.rrCutEam) {
                        const T r = PM<T>::sqrt(rr);
                        T phi, mdphi, df;
                        eam_pair<T>(p, r, phi, mdphi, df);
                        const T fc = (mdphi - df * (dFi + (T)dFk[q])) * PM<T>::rcp(r);
                        fx += fc * dx;
                        fy += fc * dy;
                        fz += fc * dz;
                        w += fc * rr;
                        u += phi;
                    }
                } else if (rrr < (R)p.rrCutLJ) {
                    T ul;
                    const T fc = lj_pair<T>(p, rr, ul);
                    fx += fc * dx;
                    fy += fc * dy;
                    fz += fc * dz;
                    frx += fc * dx;
                    w += fc * rr;
                    u += ul;
                }
            }
        }
    }
    fx = group_sum<G>(fx);
    fy = group_sum<G>(fy);
    fz = group_sum<G>(fz);
    frx = group_sum<G>(frx);
    u = group_sum<G>(u);
    w = group_sum<G>(w);
    Partial<A> q;
    partial_identity(q);
    if (active && l == 0) {
        const R ax = (R)(fx / (T)p.massMe), ay = (R)(fy / (T)p.massMe), az = (R)(fz / (T)p.massMe);
        if (integrate != 2) {  // after the next step's half-kick here, nothing reads it before it is rewritten
            acc[i] = ax;
            acc[p.nMol + i] = ay;
            acc[2 * p.nMol + i] = az;
        }
        const T e = T(0.5f) * u + (T)embF[i];
        fric[i] = (float)frx;
        R vx = vel[i], vy = vel[p.nMol + i], vz = vel[2 * p.nMol + i];
        R nx = vx, ny = vy, nz = vz;
        if (integrate) {
#pragma clang fp contract(off)
            const R hdt = (R)(0.5f * p.deltaT);
            vx = vx + hdt * ax;
            vy = vy + hdt * ay;
            vz = vz + hdt * az;
            // integrate 2: also the next step's first half-kick, the same operations k_lf1 would do
            nx = integrate == 2 ? vx + hdt * ax : vx;
            ny = integrate == 2 ? vy + hdt * ay : vy;
            nz = integrate == 2 ? vz + hdt * az : vz;
            vel[i] = nx;
            vel[p.nMol + i] = ny;
            vel[2 * p.nMol + i] = nz;
        }
        if (posOut) {  // integrate 2 with the next step's drift: the position it starts from
            P4 rn = Ai;
            rn.w = (R)e;
            drift_wrap<R>(p, rn, true, nx, ny, nz);
            posOut[i] = rn;
        } else {
            pos[i].w = (R)e;
        }
        const A m = (A)p.massMe;
        q.u = (A)e;
        q.vir = (A)(T(0.5f) * w);
        q.fr = (A)frx;
        q.com[0] = (A)Ai.x;
        q.com[1] = (A)Ai.y;
        q.com[2] = (A)Ai.z;
        q.cmv[0] = (A)vx;
        q.cmv[1] = (A)vy;
        q.cmv[2] = (A)vz;
        q.vs[0] = m * (A)vx;
        q.vs[1] = (A)vy;
        q.vs[2] = (A)vz;
        const A v2 = (A)vx * (A)vx + (A)vy * (A)vy + (A)vz * (A)vz;
        q.vv = m * v2;
        q.vmax = v2;
        for (int d = 0; d < 3; ++d) q.bmin[d] = q.bmax[d] = q.com[d];
    }
    return q;
}

// ---- carbon pass ----------------------------------------------------------------------------
// Angle deviation acos(c) - theta_0.  Plain form: as the reference, acosf.  Mixed form: the arc of
// u = sin(theta - 2pi/3) over v = cos(theta - 2pi/3) in FP32 - the deviation is an FP32 quantity
// of the force law - plus the constant (2pi/3 - theta_0) because theta_0 is the float 2pi/3 of the
// reference.  u = -s/2 - h c (h = sin(2pi/3), s = sin(theta)) cancels near the equilibrium angle;
// with c = -1/2 + d and h - s = (h^2 - s^2)/(h + s) = d (d - 1)/(h + s) it becomes
// u = d ((d - 1)/(2 (h + s)) - h), which does not.  Only d = c + 1/2 is formed in FP64; the rest,
// s included, is FP32 at full relative precision.  Within 45 degrees of 2pi/3 (every bonded angle of
// a sheet that holds together) the arc is atan of u/v by polynomial, elsewhere atan2f; both carry a
// relative error of a few ulp.  s is returned for the 1/sin(theta) factor.
template <class R>
struct AngleDev;
template <>
struct AngleDev<float> {
    __device__ static __forceinline__ float eval(float c, const DevParams &p, float &s) {
        s = sqrtf(1.f - c * c);
        return acosf(c) - p.theta_0;
    }
};
template <>
struct AngleDev<double> {
    __device__ static __forceinline__ float eval(double c, const DevParams &p, float &s) {
        const float h = 0.866025403784f;           // sin(2pi/3)
        const float d = (float)(c + 0.5);          // c - cos(2pi/3), from the FP64 cosine
        const float fc = (float)c;
        s = sqrtf(fmaxf(1.f - fc * fc, 0.f));
        const float fu = d * (0.5f * (d - 1.f) * rcpf(h + s) - h);  // sin(theta - 2pi/3)
        const float fv = -0.5f * fc + h * s;                         // cos(theta - 2pi/3)
        const float a = (fv > 0.f && fabsf(fu) <= fv) ? atan_unit(fu * rcpf(fv)) : atan2f(fu, fv);
        return a + (float)(2.0943951023931954923 - (double)p.theta_0);
    }
};

template <class R>
struct V3 {
    R x, y, z;
};

// Reciprocal and square root in the displacement precision R.
__device__ __forceinline__ float rcpR(float x) { return rcpf(x); }
__device__ __forceinline__ double rcpR(double x) { return rcp64(x); }
__device__ __forceinline__ float sqrtR(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrtR(double x) { return sqrt64(x); }

// Graphene-law accumulation in the storage precision: on a hot sheet a carbon's terms reach ~1e3 and
// cancel to a force of ~1e2, so the mixed form sums them in FP64 (the terms themselves stay FP32);
// the plain form sums in FP32, as the reference.
__device__ __forceinline__ void gacc(float &a, float t) { a += t; }
__device__ __forceinline__ void gacc(double &a, float t) { a += (double)t; }

// Neighbour-angle terms of bond (i,k) with k's bonded neighbour m (SurfaceGrowth.cu:953-978).
// rik: r_i - r_k, aik = |rik|; e: r_m - r_k, rr2 = |e|^2 (the reference divides by rr2 where |e|
// is meant; kept).
template <class R>
__device__ __forceinline__ void neighbour_angle(const DevParams &p, const V3<R> &rik, R aik, const V3<R> &e, R rr2,
                                                float mz_rel, R &fx, R &fy, R &fz) {
    const R c = (rik.x * e.x + rik.y * e.y + rik.z * e.z) * rcpR(aik * rr2);
    float s;
    const float dth = AngleDev<R>::eval(c, p, s);
    const float tmp = p.mu_theta * dth * rcpf(s);
    const float fa = (float)aik, frr = (float)rr2;
    float fc = tmp * rcpf(fa * frr);
    gacc(fx, fc * (float)e.x);
    gacc(fy, fc * (float)e.y);
    gacc(fz, fc * (float)e.z);
    fc = -tmp * (float)c * rcpf(fa * fa);
    gacc(fx, fc * (float)rik.x);
    gacc(fy, fc * (float)rik.y);
    gacc(fz, fc * (float)rik.z);
    gacc(fz, -p.mu_p * mz_rel * (1.f / 9.f));
}

// ComputeForcesK carbon branch (SurfaceGrowth.cu:899-1062, 1083-1088), one thread per carbon.
template <bool M, class T>
__device__ __forceinline__ Partial<typename Prec<M>::A> carbon_body(
    const DevParams &p, int blk, typename Prec<M>::P4 *__restrict__ pos, typename Prec<M>::P4 *__restrict__ posOut,
    typename Prec<M>::R *__restrict__ vel, typename Prec<M>::R *__restrict__ acc, const unsigned *__restrict__ lj,
    const int *__restrict__ nlj, const uint4 *__restrict__ bondC, const uint4 *__restrict__ bondA, int integrate) {
    using R = typename Prec<M>::R;
    using A = typename Prec<M>::A;
    using P4 = typename Prec<M>::P4;
    const int ic = blk * blockDim.x + threadIdx.x;
    Partial<A> q;
    partial_identity(q);
    if (ic >= p.nC) return q;
    const int i = p.nMolMe + ic;
    const P4 Ai = pos[i];
    const R z0 = (R)p.z_0;
    R fx = R(0), fy = R(0), fz = R(0), u = R(0);
    V3<R> rnn[3];
    R absR[3];
    float dZ[3];
    int kb[3];
    P4 Bk[3];
    // Bonds: the first three carbon neighbours, in list order, closer than rCutC (.cu:921).
    const uint4 nbi = bondC[ic];
    const int counti = (int)nbi.w;
    const unsigned be[3] = {nbi.x, nbi.y, nbi.z};
    const int j0 = counti == 3 ? (int)(nbi.z & 255u) : -1;
    uint4 nbk[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        kb[b] = b < counti ? (int)(be[b] >> 8) : i;
        Bk[b] = pos[kb[b]];
        nbk[b] = bondA[b < counti ? kb[b] - p.nMolMe : ic];
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        if (b >= counti) continue;
        const R rrr = disp<R>(Ai, Bk[b], p, rnn[b].x, rnn[b].y, rnn[b].z);
        absR[b] = sqrtR(rrr);
        dZ[b] = (float)(Bk[b].z - z0);
    }
    // For each bond k: k's own near neighbours m != i, the first two in k's list order (.cu:940-978).
#pragma unroll
    for (int b = 0; b < 3; ++b) {  // unrolled: every array index is a constant, nothing goes to scratch
        if (b >= counti) break;
        // The two partners are picked before any arithmetic: where i sits in k's list differs from
        // lane to lane, and a loop over all three entries would issue three evaluations per bond.
        const unsigned me[3] = {nbk[b].x, nbk[b].y, nbk[b].z};
        const int na = (int)nbk[b].w;
        int ms[2] = {i, i}, nsel = 0;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int m = (int)(me[t] >> 8);
            if (t < na && m != i) {
                if (nsel == 0)
                    ms[0] = m;
                else if (nsel == 1)
                    ms[1] = m;
                ++nsel;
            }
        }
        P4 Bm[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) Bm[t] = pos[ms[t]];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (t >= nsel) break;
            V3<R> e;
            const R rr2 = disp<R>(Bk[b], Bm[t], p, e.x, e.y, e.z);
            const V3<R> em = {-e.x, -e.y, -e.z};
            neighbour_angle<R>(p, rnn[b], absR[b], em, rr2, (float)(Bm[t].z - z0), fx, fy, fz);
        }
    }
    if (counti == 3) {  // SurfaceGrowth.cu:985-1061
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float dr = (float)(absR[b] - R(1));
            const float fc = -p.mu_r * dr * rcpf((float)absR[b]);
            gacc(fx, fc * (float)rnn[b].x);
            gacc(fy, fc * (float)rnn[b].y);
            gacc(fz, fc * (float)rnn[b].z);
            gacc(u, p.mu_r * dr * dr);
        }
        const float zi = (float)(Ai.z - z0);
        float dz = 2.f * zi - (dZ[0] + dZ[1] + dZ[2]);
        gacc(fz, -0.666666667f * p.mu_p * dz);
        dz = zi - 0.33333333f * (dZ[0] + dZ[1] + dZ[2]);
        gacc(u, p.mu_p * dz * dz);
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) {
            const int a = qq, b = qq == 2 ? 0 : qq + 1;
            const R c = (rnn[a].x * rnn[b].x + rnn[a].y * rnn[b].y + rnn[a].z * rnn[b].z) * rcpR(absR[a] * absR[b]);
            float s;
            const float dth = AngleDev<R>::eval(c, p, s);
            const float tmp = p.mu_theta * dth * rcpf(s);
            const float ra = (float)absR[a], rb = (float)absR[b], fcos = (float)c;
            const float ira = rcpf(ra), irb = rcpf(rb);
            const float inv = ira * irb;
            const float fcb = tmp * (1.f - ra * fcos * irb) * inv;
            const float fca = tmp * (1.f - rb * fcos * ira) * inv;
            gacc(fx, fcb * (float)rnn[b].x + fca * (float)rnn[a].x);
            gacc(fy, fcb * (float)rnn[b].y + fca * (float)rnn[a].y);
            gacc(fz, fcb * (float)rnn[b].z + fca * (float)rnn[a].z);
            gacc(u, p.mu_theta * dth * dth);
        }
    }
    // LJ with metal.  The reference reuses its neighbour-loop counter for the bond loop
    // (SurfaceGrowth.cu:989), so after the third bond (list position j0) the loop resumes at
    // position 4: entries at positions <= j0 count once, positions 4..j0 twice, and with j0 = 2
    // position 3 not at all.
    T gx = (T)fx, gy = (T)fy, gz = (T)fz, gu = (T)u, gw = T(0);
    const int nl = nlj[ic];
    for (int qq = 0; qq < nl; ++qq) {
        const unsigned e = lj[(size_t)qq * p.nC + ic];
        const int k = (int)(e >> 8), pk = (int)(e & 255u);
        const int mult = j0 < 0 ? 1 : (int)(pk <= j0) + (int)(pk >= 4);
        if (mult == 0) continue;
        R dxr, dyr, dzr;
        const R rrr = disp<R>(Ai, pos[k], p, dxr, dyr, dzr);
        if (!(rrr < (R)p.rrCutLJ)) continue;
        T ul;
        const T fc = lj_pair<T>(p, (T)rrr, ul);
        for (int t = 0; t < mult; ++t) {
            gx += fc * (T)dxr;
            gy += fc * (T)dyr;
            gz += fc * (T)dzr;
            gu += ul;
            gw += fc * (T)rrr;
        }
    }
    const R ax = (R)gx, ay = (R)gy, az = (R)gz;  // carbon mass is the unit
    if (integrate != 2) {  // after the next step's half-kick here, nothing reads it before it is rewritten
        acc[i] = ax;
        acc[p.nMol + i] = ay;
        acc[2 * p.nMol + i] = az;
    }
    const float e = (float)(T(0.5f) * gu);
    R vx = vel[i], vy = vel[p.nMol + i], vz = vel[2 * p.nMol + i];
    R nx = vx, ny = vy, nz = vz;
    const bool move = Ai.x != (R)p.xLeft && Ai.x < (R)p.xRight && Ai.y != (R)p.yTop && Ai.y != (R)p.yBottom;
    if (integrate && move) {
#pragma clang fp contract(off)
        const R hdt = (R)(0.5f * p.deltaT);
        vx = vx + hdt * ax;
        vy = vy + hdt * ay;
        vz = vz + hdt * az;
        nx = integrate == 2 ? vx + hdt * ax : vx;  // 2: with the next step's first half-kick
        ny = integrate == 2 ? vy + hdt * ay : vy;
        nz = integrate == 2 ? vz + hdt * az : vz;
        vel[i] = nx;
        vel[p.nMol + i] = ny;
        vel[2 * p.nMol + i] = nz;
    }
    if (posOut) {  // integrate 2 with the next step's drift (k_lf1's move test: Ai is its position)
        P4 rn = Ai;
        rn.w = (R)e;
        drift_wrap<R>(p, rn, move, nx, ny, nz);
        posOut[i] = rn;
    } else {
        pos[i].w = (R)e;
    }
    q.u = (A)e;
    q.vir = (A)(T(0.5f) * gw);
    q.vs[0] = (A)vx;
    q.vs[1] = (A)vy;
    q.vs[2] = (A)vz;
    const A v2 = (A)vx * (A)vx + (A)vy * (A)vy + (A)vz * (A)vz;
    q.vv = v2;
    q.vmax = v2;
    return q;
}

// One workgroup of the force launch (blocks [0, nbM) metal, [nbM, nb) carbon): forces, energies,
// leapfrog part 2 and the block's Partial in part[b].
template <bool M, class T, int G>
__device__ __forceinline__ void force_block(const DevParams &p, int b, int nbM, int nb, bool xm,
                                            typename Prec<M>::P4 *__restrict__ pos,
                                            typename Prec<M>::P4 *__restrict__ posOut,
                                            typename Prec<M>::R *__restrict__ vel, typename Prec<M>::R *__restrict__ acc,
                                            const int *__restrict__ nlM, const int *__restrict__ nnM,
                                            const float *__restrict__ embF, const float *__restrict__ embDF,
                                            float *__restrict__ fric, const unsigned *__restrict__ lj,
                                            const int *__restrict__ nlj, const uint4 *__restrict__ bondC,
                                            const uint4 *__restrict__ bondA, Partial<typename Prec<M>::A> *part,
                                            int integrate) {
    using A = typename Prec<M>::A;
    // Each branch reduces and stores on its own: a Partial merged across the two branches keeps
    // both bodies' registers live at once (154 VGPRs in the FP64 form against 120 apart).
    if (b < nbM) {
        const Partial<A> q = metal_body<M, T, G>(p, xm ? xcd_block(b, nbM) : b, pos, posOut, vel, acc, nlM, nnM, embF,
                                                 embDF, fric, integrate);
        const Partial<A> r = block_reduce_partial<true>(q);
        if (threadIdx.x == 0) part[b] = r;
    } else {
        const Partial<A> q = carbon_body<M, T>(p, xm ? xcd_block(b - nbM, nb - nbM) : b - nbM, pos, posOut, vel, acc,
                                               lj, nlj, bondC, bondA, integrate);
        const Partial<A> r = block_reduce_partial<false>(q);
        if (threadIdx.x == 0) part[b] = r;
    }
}

template <bool M, class T, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_force(DevParams p, int nbM, typename Prec<M>::P4 *__restrict__ pos,
                                               typename Prec<M>::P4 *__restrict__ posOut,
                                               typename Prec<M>::R *__restrict__ vel, typename Prec<M>::R *__restrict__ acc,
                                               const int *__restrict__ nlM, const int *__restrict__ nnM,
                                               const float *__restrict__ embF, const float *__restrict__ embDF,
                                               float *__restrict__ fric, const unsigned *__restrict__ lj,
                                               const int *__restrict__ nlj, const uint4 *__restrict__ bondC,
                                               const uint4 *__restrict__ bondA, Partial<typename Prec<M>::A> *part,
                                               int integrate) {
    // nbM is a multiple of 8: both ranges keep the XCD of each block
    force_block<M, T, G>(p, blockIdx.x, nbM, gridDim.x, p.xcdMap & 4, pos, posOut, vel, acc, nlM, nnM, embF, embDF,
                         fric, lj, nlj, bondC, bondA, part, integrate);
}

// Batched: each replica's workgroups write into its own range of partials (kb_finalize reduces).
template <bool M, class T, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kb_force(BatchArgs a, int integrate) {
    using R = typename Prec<M>::R;
    using A = typename Prec<M>::A;
    const int2 e = a.forceTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const DevParams &p = B.p;
    force_block<M, T, G>(p, e.y, B.nbM, B.nbForce, false, (typename Prec<M>::P4 *)a.pos + B.atomOff,
                         a.posOut ? (typename Prec<M>::P4 *)a.posOut + B.atomOff : nullptr,
                         (R *)a.vel + 3 * B.atomOff, (R *)a.acc + 3 * B.atomOff, a.nlM + B.metalOff * p.nlCap,
                         a.nnM + B.metalOff, a.embF + B.metalOff, a.embDF + B.metalOff, a.fric + B.metalOff,
                         a.lj + B.carbonOff * p.ljCap, a.nlj + B.carbonOff, a.bondC + B.carbonOff,
                         a.bondA + B.carbonOff, (Partial<A> *)a.partials + B.partOff, integrate);
}

// ---- launchers ---------------------------------------------------------------------------------

template <bool M, class T>
static void rho_dispatch(sgk_ctx *c) {
    using P4 = typename Prec<M>::P4;
    const int G = c->metalLanes;
    const unsigned g = (unsigned)(c->nbRho + c->nbC);
    const P4 *pos = (const P4 *)c->pos;
#define SGK_RHO(GG)                                                                                              \
    hipLaunchKernelGGL((k_rho<M, T, GG>), dim3(g), dim3(256), 0, c->stream, c->dp, c->nbRho, pos, c->nlM, c->nnM, \
                       c->rho, c->embF, c->embDF, c->cl, c->ncl, c->bondC, c->bondA)
    switch (G) {
        case 1: SGK_RHO(1); break;
        case 2: SGK_RHO(2); break;
        case 4: SGK_RHO(4); break;
        case 16: SGK_RHO(16); break;
        default: SGK_RHO(8); break;
    }
#undef SGK_RHO
}

template <bool M, class T>
static void force_dispatch(sgk_ctx *c, int integrate, void *posOut) {
    using P4 = typename Prec<M>::P4;
    using R = typename Prec<M>::R;
    using A = typename Prec<M>::A;
    const int G = c->metalLanes;
    const unsigned g = (unsigned)(c->nbM + c->nbC);
    P4 *pos = (P4 *)c->pos;
    R *vel = (R *)c->vel;
    Partial<A> *part = (Partial<A> *)c->partials;
#define SGK_F(GG)                                                                                                \
    hipLaunchKernelGGL((k_force<M, T, GG>), dim3(g), dim3(256), 0, c->stream, c->dp, c->nbM, pos, (P4 *)posOut,  \
                       vel, (R *)c->acc,                                                                        \
                       c->nlM, c->nnM, c->embF, c->embDF, c->fric, c->lj, c->nlj, c->bondC, c->bondA, part,     \
                       integrate)
    switch (G) {
        case 1: SGK_F(1); break;
        case 2: SGK_F(2); break;
        case 4: SGK_F(4); break;
        case 16: SGK_F(16); break;
        default: SGK_F(8); break;
    }
#undef SGK_F
}

void launch_rho(sgk_ctx *c) {
    if (c->nbRho + c->nbC == 0) return;
    profile_begin(c);
    if (!c->mixed)
        rho_dispatch<false, float>(c);
    else if (c->cfg.pair_fp64)
        rho_dispatch<true, double>(c);
    else
        rho_dispatch<true, float>(c);
    profile_end(c, "rho");
}

int batch_lane_index(int G) { return G == 1 ? 0 : G == 2 ? 1 : G == 4 ? 2 : G == 8 ? 3 : 4; }

// One launch per lane count present: the replicas with G lanes per metal atom, through their own
// block table (the kernel body is templated on G).
template <bool M, class T>
static void batch_dispatch(sgk_batch *b, bool force, int integrate, void *posOut) {
    static const int kinds[sgk_batch::kLaneKinds] = {1, 2, 4, 8, 16};
    for (int q = 0; q < sgk_batch::kLaneKinds; ++q) {
        if (b->grpN[q] == 0) continue;
        BatchArgs a = b->args();
        a.posOut = posOut;
        a.forceTab = b->forceTab + b->grpOff[q];
        const dim3 g(b->grpN[q]), t(256);
#define SGK_B(GG)                                                                            \
    do {                                                                                     \
        if (force)                                                                           \
            hipLaunchKernelGGL((kb_force<M, T, GG>), g, t, 0, b->stream, a, integrate);     \
        else                                                                                 \
            hipLaunchKernelGGL((kb_rho<M, T, GG>), g, t, 0, b->stream, a);                  \
    } while (0)
        switch (kinds[q]) {
            case 1: SGK_B(1); break;
            case 2: SGK_B(2); break;
            case 4: SGK_B(4); break;
            case 16: SGK_B(16); break;
            default: SGK_B(8); break;
        }
#undef SGK_B
    }
}

static void batch_launch(sgk_batch *b, bool force, int integrate, void *posOut = nullptr) {
    if (b->nForceBlocks == 0) return;
    if (integrate != 2) posOut = nullptr;  // the drift follows the next step's half-kick only
    if (!b->mixed)
        batch_dispatch<false, float>(b, force, integrate, posOut);
    else if (b->cfg.pair_fp64)
        batch_dispatch<true, double>(b, force, integrate, posOut);
    else
        batch_dispatch<true, float>(b, force, integrate, posOut);
}

void launch_batch_rho(sgk_batch *b) { batch_launch(b, false, 0); }
void launch_batch_force(sgk_batch *b, bool prekick, void *posOut) { batch_launch(b, true, prekick ? 2 : 1, posOut); }
void launch_batch_force_eval(sgk_batch *b) { batch_launch(b, true, 0); }

void launch_force(sgk_ctx *c, bool integrate, bool finalize, bool prekick, void *posOut) {
    if (c->nbM + c->nbC == 0) return;
    profile_begin(c);
    const int mode = integrate ? (prekick ? 2 : 1) : 0;
    if (mode != 2) posOut = nullptr;  // the drift follows the next step's half-kick only
    if (!c->mixed)
        force_dispatch<false, float>(c, mode, posOut);
    else if (c->cfg.pair_fp64)
        force_dispatch<true, double>(c, mode, posOut);
    else
        force_dispatch<true, float>(c, mode, posOut);
    profile_end(c, "force");
    if (finalize) launch_finalize(c, 0);
}

}  // namespace sgk
