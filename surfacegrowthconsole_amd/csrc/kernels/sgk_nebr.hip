// sgk_nebr.hip - cell binning, neighbour lists and the RDF histogram for gfx950.
//
// Rebuild protocol (two cell buffers, cur = state.cellbuf, nxt = cur ^ 1):
//   binning     cell index per atom with the reference formula, slot by atomicAdd into nxt
//               (bin_atom; in the step loop it runs inside the leapfrog kernel).
//   k_sortcells per atom: the atom that took slot 0 of its cell has its wave sort that cell's
//               slots ascending (bitonic, in registers) and write the cell-ordered float
//               positions with each atom's index in w (padded with a far-away point to a multiple
//               of 8), so a list walk reads one 16-byte entry per candidate; every atom clears its
//               cell's count in cur for the rebuild after this one.  No global counter is involved.
//   k_nlc       the lists in the reference order from the sorted cells: lanes per metal atom, one
//               lane per carbon in cell order (nl_scheme 3, the default; schemes 0-2: k_nl, k_nlt).
//   commit      cellbuf = nxt (the step reduction, or k_commit).
// The API path (sgk_bin) bins with k_bin.  In the device-resident loop the kernels return at once
// unless state.nebrNow is set, so a captured step needs no host decision.
//
// Cell index, neighbour-list membership and order, and RDF bins are integer results defined by the
// reference's FP32 arithmetic; both precision forms evaluate them in FP32 on the float-rounded
// positions.  (The graphene sheet lies exactly on a cell boundary in z by construction, so the
// cell of every carbon atom - and the list order the carbon force loop depends on - is decided by
// that rounding.)
#include "sgk_batch.h"
#include "sgk_cells.h"
#include "sgk_device.h"
#include "sgk_internal.h"

namespace sgk {

// Offset table of ComputeDefs.h:221-234: offset o -> (s[o/9], s[(o/3)%3], s[o%3]), s = {0,1,-1}.
__device__ __forceinline__ int off_comp(int v) { return v == 0 ? 0 : (v == 1 ? 1 : -1); }

// Neighbour cell of (ccx,ccy,ccz) for offset o with the cell-wrap shift (SurfaceGrowth.cu:641-674).
template <class R>
__device__ __forceinline__ int stencil_cell(const DevParams &p, int ccx, int ccy, int ccz, int o, R &shx, R &shy,
                                            R &shz) {
    int mx = ccx + off_comp(o / 9), my = ccy + off_comp((o / 3) % 3), mz = ccz + off_comp(o % 3);
    shx = shy = shz = R(0);
    if (mx >= p.cx) {
        mx = 0;
        shx = (R)p.Lx;
    } else if (mx < 0) {
        mx = p.cx - 1;
        shx = -(R)p.Lx;
    }
    if (my >= p.cy) {
        my = 0;
        shy = (R)p.Ly;
    } else if (my < 0) {
        my = p.cy - 1;
        shy = -(R)p.Ly;
    }
    if (mz >= p.cz) {
        mz = 0;
        shz = (R)p.Lz;
    } else if (mz < 0) {
        mz = p.cz - 1;
        shz = -(R)p.Lz;
    }
    return (mz * p.cy + my) * p.cx + mx;
}

// Float copy of a position (exact in the plain form, rounded in the mixed form).
template <class P4>
__device__ __forceinline__ float3 f3(const P4 &q) {
    return make_float3((float)q.x, (float)q.y, (float)q.z);
}

// Distance test of the neighbour-list build and RDF: ((A - B) - shift)^2 in FP32 with every
// operation rounded separately, the canonical evaluation the oracle uses.  SHIFT = false leaves out
// the subtraction of shifts known to be zero (x - 0 is x exactly).
template <bool SHIFT = true>
__device__ __forceinline__ float shifted_rr(const float3 &a, const float3 &b, float shx, float shy, float shz) {
#pragma clang fp contract(off)
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    if (SHIFT) {
        dx = dx - shx;
        dy = dy - shy;
        dz = dz - shz;
    }
    return dx * dx + dy * dy + dz * dz;
}

// BinAtomsIntoCellsK (SurfaceGrowth.cu:547-609) as an O(N) pass (API path).
template <bool M>
__global__ void k_bin(DevParams p, const typename Prec<M>::P4 *__restrict__ pos, CellBufs bufs, DevState *st,
                      int force) {
    if (!force && !st->nebrNow) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nMol) return;
    bin_atom(p, i, f3(pos[i]), bufs, st);
}

// Coordinate of the padding entries of the cell-ordered positions: its squared distance to any atom
// (~3e36) is finite in FP32 and beyond every cutoff.
constexpr float kFarAway = 1.0e18f;

// Sort the cells owned by the lanes in `own` (cell index cn, count ocnt in the owner lane), 64 / W
// cells per pass: lanes [g W, g W + W) sort the g-th cell of the pass with a bitonic network in
// registers, then store the sorted slots and the cell-ordered float positions.  Wave-collective.
template <int W, class P4>
__device__ __forceinline__ void sort_cells(const DevParams &p, const CellBufs &bufs, int b, const P4 *__restrict__ pos,
                                           int cn, int ocnt, unsigned long long own) {
    constexpr int kPer = 64 / W;
    const int lane = threadIdx.x & 63, g = lane / W, idx = lane % W;
    while (own) {
        int src = -1;  // owner lane of this lane's cell in the pass
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int l = own ? __ffsll((long long)own) - 1 : -1;
            if (own) own &= own - 1ull;
            if (g == q) src = l;
        }
        // every lane takes part in both exchanges (a lane that skipped one would supply no data)
        const int c = __shfl(cn, src < 0 ? 0 : src, 64);
        const int cntSrc = __shfl(ocnt, src < 0 ? 0 : src, 64);
        const int cnt = src < 0 ? 0 : cntSrc;
        int *s = bufs.slotsB(b) + (size_t)(src < 0 ? 0 : c) * p.cellCap;
        int v = idx < cnt ? s[idx] : 0x7fffffff;
#pragma unroll
        for (int k = 2; k <= W; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int w = __shfl_xor(v, j, 64);
                const bool lower = (idx & j) == 0, up = (idx & k) == 0;
                v = (lower == up) ? min(v, w) : max(v, w);
            }
        }
        if (idx < cnt) {
            s[idx] = v;
            const float3 q = f3(pos[v]);
            bufs.cellPos[(size_t)c * p.cellCap + idx] = make_float4(q.x, q.y, q.z, __int_as_float(v));
        } else if (idx < ((cnt + 7) & ~7)) {
            // padding up to the next multiple of 8: a position no list radius reaches, so a walk in
            // chunks of 8 needs no per-candidate bound test
            bufs.cellPos[(size_t)c * p.cellCap + idx] =
                make_float4(kFarAway, kFarAway, kFarAway, __int_as_float(0x7fffffff));
        }
    }
}

// Sort each occupied cell of the new buffer (<= 64 atoms, one per lane) ascending and store its
// atoms' float positions in slot order; clear the counts of the outgoing buffer.  One thread per
// atom; the owner of a cell is the atom holding slot 0, and the owner's wave sorts the cell.
template <bool M>
__device__ __forceinline__ void sortcells_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                               const CellBufs &bufs, DevState *st, int force) {
    if (!force && !st->nebrNow) return;
    const int b = st->cellbuf ^ 1, old = st->cellbuf;
    const int n = blk * blockDim.x + threadIdx.x;
    const bool in = n < p.nMol;  // every lane stays for the wave-wide sort
    int cn = -1;
    if (in) {
        const int co = bufs.cellOfB(old)[n];
        if (co >= 0) bufs.cntB(old)[co] = 0;
        cn = bufs.cellOfB(b)[n];
    }
    const bool owner = in && cn >= 0 && bufs.slotOf[n] == 0;
    const int ocnt = owner ? min(bufs.cntB(b)[cn], p.cellCap) : 0;
    // Cells of up to 16 atoms (most of them) are sorted four at a time, 16 lanes each; up to 32, two
    // at a time; larger ones with the whole wave.
    sort_cells<16>(p, bufs, b, pos, cn, ocnt, __ballot(owner && ocnt <= 16));
    sort_cells<32>(p, bufs, b, pos, cn, ocnt, __ballot(owner && ocnt > 16 && ocnt <= 32));
    sort_cells<64>(p, bufs, b, pos, cn, ocnt, __ballot(owner && ocnt > 32));
}

template <bool M>
__global__ __launch_bounds__(256) void k_sortcells(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                                   CellBufs bufs, DevState *st, int force) {
    sortcells_body<M>(p, blockIdx.x, pos, bufs, st, force);
}

template <bool M>
__global__ __launch_bounds__(256) void kb_sortcells(BatchArgs a) {
    const int2 e = a.atomTab[blockIdx.x];
    DevState *st = a.state + e.x;
    if (!st->nebrNow) return;  // tested before the replica's table entry is read: most steps rebuild nothing
    const BatchRep &B = a.reps[e.x];
    sortcells_body<M>(B.p, e.y, (const typename Prec<M>::P4 *)a.pos + B.atomOff, rep_cells(a, B), st, 0);
}

// BuildNebrListK (SurfaceGrowth.cu:612-711), one thread per atom over cells whose slots are sorted
// ascending (k_sortcells), neighbours appended in the reference order: 27 offset cells in table
// order, atoms of a cell ascending.  Slots are read 8 at a time with vector loads and their
// positions gathered together; the next cell's count is prefetched.  Outputs:
//   metal atom:  nlM[i*nlCap + j] = k                     (all neighbours)
//   carbon atom: lj[q*nC + ic] = k<<8 | j   for metal k   (LJ partners with list position j)
//                cl[ic*clCap + q] = k<<8 | j if r^2 < closeR2 (bond and angle candidates)
//   keepFull:    fullNBL[j*nMol + i] = k, fullNN[i]      (reference layout)
// Append neighbour k (at squared distance rr) to atom i's lists; nn, nljv, nclv are its running
// counts (list position, metal partners, close partners).
__device__ __forceinline__ void nl_emit(const DevParams &p, int i, bool metal, int ic, int k, float rr, int &nn,
                                        int &nljv, int &nclv, int &err, int *__restrict__ nlM,
                                        unsigned *__restrict__ lj, unsigned *__restrict__ cl,
                                        int *__restrict__ fullNBL) {
    if (nn < p.nlCap) {
        if (p.keepFull) fullNBL[(size_t)nn * p.nMol + i] = k;
        if (metal) {
            nlM[(size_t)i * p.nlCap + nn] = k;
        } else {
            const unsigned ent = ((unsigned)k << 8) | (unsigned)nn;
            if (k < p.nMolMe) {
                if (nljv < p.ljCap)
                    lj[(size_t)nljv * p.nC + ic] = ent;
                else
                    err |= SGK_FLAG_LJ_OVERFLOW;
                ++nljv;
            }
            if (rr < p.closeR2) {
                if (nclv < p.clCap)
                    cl[(size_t)ic * p.clCap + nclv] = ent;
                else
                    err |= SGK_FLAG_CLOSE_OVERFLOW;
                ++nclv;
            }
        }
    } else {
        err |= SGK_FLAG_NL_OVERFLOW;
    }
    ++nn;
}

// Counts of atom i's lists and its error flags.
__device__ __forceinline__ void nl_finish(const DevParams &p, int i, bool metal, int ic, int nn, int nljv, int nclv,
                                          int err, int *__restrict__ nnM, int *__restrict__ nlj,
                                          int *__restrict__ ncl, int *__restrict__ fullNN, DevState *st) {
    nn = min(nn, p.nlCap);
    if (p.keepFull) fullNN[i] = nn;
    if (metal) {
        nnM[i] = nn;
    } else {
        nlj[ic] = min(nljv, p.ljCap);
        ncl[ic] = min(nclv, p.clCap);
    }
    if (err) atomicOr(&st->errors, err);
}

template <bool M>
__device__ __forceinline__ void nl_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                        const CellBufs &bufs, DevState *st, int force, int fromNext,
                                        int *__restrict__ nlM, int *__restrict__ nnM, unsigned *__restrict__ lj,
                                        int *__restrict__ nlj, unsigned *__restrict__ cl, int *__restrict__ ncl,
                                        int *__restrict__ fullNN, int *__restrict__ fullNBL) {
    if (!force && !st->nebrNow) return;
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= p.nMol) return;
    const int b = fromNext ? (st->cellbuf ^ 1) : st->cellbuf;
    const int *cnt = bufs.cntB(b);
    const float4 *cpos = bufs.cellPos;
    const bool metal = i < p.nMolMe;
    const int ic = i - p.nMolMe;
    const int c = bufs.cellOfB(b)[i];
    int nn = 0, nljv = 0, nclv = 0, err = 0;
    if (c >= 0) {
        const int ccz = c / (p.cx * p.cy);
        const int ccy = (c - ccz * p.cx * p.cy) / p.cx;
        const int ccx = c - (ccz * p.cy + ccy) * p.cx;
        const float3 A = f3(pos[i]);
        const float rrN = p.rrNebr;
        float shx, shy, shz;
        int C = stencil_cell<float>(p, ccx, ccy, ccz, 0, shx, shy, shz);
        int ncur = cnt[C];
        for (int o = 0; o < 27; ++o) {
            float nshx = 0.f, nshy = 0.f, nshz = 0.f;
            int Cn = C, nnext = 0;
            if (o < 26) {  // prefetch the next cell's count
                Cn = stencil_cell<float>(p, ccx, ccy, ccz, o + 1, nshx, nshy, nshz);
                nnext = cnt[Cn];
            }
            const int n = min(ncur, p.cellCap);
            const float4 *cp = cpos + (size_t)C * p.cellCap;
            for (int base = 0; base < n; base += 8) {
                float4 P[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) P[t] = cp[base + t];
                int k[8];  // the slot's atom index rides in w
#pragma unroll
                for (int t = 0; t < 8; ++t) k[t] = __float_as_int(P[t].w);
                float3 B[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) B[t] = make_float3(P[t].x, P[t].y, P[t].z);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    if (base + t >= n) continue;
                    const float rr = shifted_rr(A, B[t], shx, shy, shz);
                    if (!(rr < rrN && k[t] != i)) continue;
                    nl_emit(p, i, metal, ic, k[t], rr, nn, nljv, nclv, err, nlM, lj, cl, fullNBL);
                }
            }
            C = Cn;
            ncur = nnext;
            shx = nshx;
            shy = nshy;
            shz = nshz;
        }
    }
    nl_finish(p, i, metal, ic, nn, nljv, nclv, err, nnM, nlj, ncl, fullNN, st);
}

// BuildNebrListK (SurfaceGrowth.cu:612-711) with the candidates staged in LDS.  The 256 atoms of a
// workgroup are consecutive indices - spatial neighbours, since the generators emit atoms along
// lattice rows - so the cells they can reach form a small box: their cells' bounding box plus one
// cell on every side.  The workgroup loads every atom of that box once (counts, a block scan, then
// positions and indices) and each thread walks its atom's 27 cells in LDS, in the reference order,
// with the same FP32 distance test and the same cell-wrap shift as nl_body.  A workgroup whose box
// exceeds the LDS budget (a metal cluster's rows, a row crossing the periodic boundary) falls back
// to nl_body; the decision is uniform over the workgroup.
constexpr int kNlMaxHalo = 1024;  // cells of the box
constexpr int kNlMaxEnt = 2048;   // atoms in those cells

// Exclusive prefix sum of a[0..n) (n <= 4 * blockDim.x, blockDim.x = 256) in place; a[n] = total.
__device__ __forceinline__ void block_scan4(int *a, int n) {
    __shared__ int wsum[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int v[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int h = 4 * t + q;
        v[q] = h < n ? a[h] : 0;
        sum += v[q];
    }
    int inc = sum;  // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(inc, d, 64);
        if (lane >= d) inc += u;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int base = inc - sum;
    for (int k = 0; k < w; ++k) base += wsum[k];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int h = 4 * t + q;
        if (h < n) a[h] = base;
        base += v[q];
    }
    if (t == blockDim.x - 1) a[n] = base;
    __syncthreads();
}

template <bool M>
__device__ __forceinline__ void nlt_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                         const CellBufs &bufs, DevState *st, int force, int fromNext,
                                         int *__restrict__ nlM, int *__restrict__ nnM, unsigned *__restrict__ lj,
                                         int *__restrict__ nlj, unsigned *__restrict__ cl, int *__restrict__ ncl,
                                         int *__restrict__ fullNN, int *__restrict__ fullNBL) {
    if (!force && !st->nebrNow) return;
    __shared__ int sStart[kNlMaxHalo + 1];
    __shared__ float4 sEnt[kNlMaxEnt];
    __shared__ int sLo[3], sHi[3], sFall;
    const int i = blk * blockDim.x + threadIdx.x;
    const bool in = i < p.nMol;
    const int b = fromNext ? (st->cellbuf ^ 1) : st->cellbuf;
    const int c = in ? bufs.cellOfB(b)[i] : -1;
    int ccx = 0, ccy = 0, ccz = 0;
    if (c >= 0) {
        ccz = c / (p.cx * p.cy);
        ccy = (c - ccz * p.cx * p.cy) / p.cx;
        ccx = c - (ccz * p.cy + ccy) * p.cx;
    }
    if (threadIdx.x == 0) {
        sLo[0] = sLo[1] = sLo[2] = 0x7fffffff;
        sHi[0] = sHi[1] = sHi[2] = -1;
    }
    __syncthreads();
    if (c >= 0) {
        atomicMin(&sLo[0], ccx);
        atomicMin(&sLo[1], ccy);
        atomicMin(&sLo[2], ccz);
        atomicMax(&sHi[0], ccx);
        atomicMax(&sHi[1], ccy);
        atomicMax(&sHi[2], ccz);
    }
    __syncthreads();
    if (sHi[0] < 0) {  // no binned atom in the workgroup: empty lists
        if (in) nl_finish(p, i, i < p.nMolMe, i - p.nMolMe, 0, 0, 0, 0, nnM, nlj, ncl, fullNN, st);
        return;
    }
    const int x0 = sLo[0] - 1, y0 = sLo[1] - 1, z0 = sLo[2] - 1;  // box origin (grid coordinates)
    const int BX = sHi[0] - sLo[0] + 3, BY = sHi[1] - sLo[1] + 3, BZ = sHi[2] - sLo[2] + 3;
    const long long Hl = (long long)BX * BY * BZ;
    const int *cnt = bufs.cntB(b);
    // box cell h -> grid cell, with the reference's single wrap per axis (stencil_cell)
    auto grid_cell = [&](int h) {
        const int hz = h / (BX * BY), hy = (h / BX) % BY, hx = h % BX;
        int gx = x0 + hx, gy = y0 + hy, gz = z0 + hz;
        gx = gx >= p.cx ? 0 : (gx < 0 ? p.cx - 1 : gx);
        gy = gy >= p.cy ? 0 : (gy < 0 ? p.cy - 1 : gy);
        gz = gz >= p.cz ? 0 : (gz < 0 ? p.cz - 1 : gz);
        return (gz * p.cy + gy) * p.cx + gx;
    };
    const int H = Hl <= kNlMaxHalo ? (int)Hl : 0;
    if (threadIdx.x == 0) sFall = Hl > kNlMaxHalo;
    for (int h = threadIdx.x; h < H; h += blockDim.x) sStart[h] = min(cnt[grid_cell(h)], p.cellCap);
    __syncthreads();
    if (!sFall) {
        block_scan4(sStart, H);
        if (threadIdx.x == 0) sFall = sStart[H] > kNlMaxEnt;
        __syncthreads();
    }
    if (sFall) {  // uniform: the whole workgroup walks the cells in global memory
        nl_body<M>(p, blk, pos, bufs, st, force, fromNext, nlM, nnM, lj, nlj, cl, ncl, fullNN, fullNBL);
        return;
    }
    const int E = sStart[H];
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        int lo = 0, hi = H;  // last box cell whose start is <= e
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sStart[mid] <= e) lo = mid; else hi = mid;
        }
        const size_t g = (size_t)grid_cell(lo) * p.cellCap + (e - sStart[lo]);
        sEnt[e] = bufs.cellPos[g];  // w holds the slot's atom index
    }
    __syncthreads();
    if (!in) return;
    const bool metal = i < p.nMolMe;
    const int ic = i - p.nMolMe;
    int nn = 0, nljv = 0, nclv = 0, err = 0;
    if (c >= 0) {
        const float3 A = f3(pos[i]);
        const float rrN = p.rrNebr;
        for (int o = 0; o < 27; ++o) {
            float shx, shy, shz;
            (void)stencil_cell<float>(p, ccx, ccy, ccz, o, shx, shy, shz);
            const int h = (ccx + off_comp(o / 9) - x0) + BX * ((ccy + off_comp((o / 3) % 3) - y0) +
                                                               BY * (ccz + off_comp(o % 3) - z0));
            const int e1 = sStart[h + 1];
            for (int e = sStart[h]; e < e1; ++e) {
                const float4 q = sEnt[e];
                const int k = __float_as_int(q.w);
                const float rr = shifted_rr(A, make_float3(q.x, q.y, q.z), shx, shy, shz);
                if (!(rr < rrN && k != i)) continue;
                nl_emit(p, i, metal, ic, k, rr, nn, nljv, nclv, err, nlM, lj, cl, fullNBL);
            }
        }
    }
    nl_finish(p, i, metal, ic, nn, nljv, nclv, err, nnM, nlj, ncl, fullNN, st);
}

// BuildNebrListK (SurfaceGrowth.cu:612-711), one wave per occupied cell (the wave of the cell's
// slot-0 atom).  The candidates of the 27 neighbour cells are laid over the lanes in the reference
// order (offset table order, then ascending slot), 64 at a time; for each atom of the cell in turn
// a ballot of the lanes that pass the distance test gives the new entries, and a prefix popcount
// their positions - so order and positions are those of the per-atom walk.  The candidates of a
// chunk are loaded once for all atoms of the cell, in one round trip.  Output as nl_body.
template <bool M>
__device__ __forceinline__ void nlw_body(const DevParams &p, int blk, const CellBufs &bufs, DevState *st,
                                         int force, int fromNext, int *__restrict__ nlM, int *__restrict__ nnM,
                                         unsigned *__restrict__ lj, int *__restrict__ nlj, unsigned *__restrict__ cl,
                                         int *__restrict__ ncl, int *__restrict__ fullNN, int *__restrict__ fullNBL) {
    if (!force && !st->nebrNow) return;
    const int b = fromNext ? (st->cellbuf ^ 1) : st->cellbuf;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int n = blk * blockDim.x + threadIdx.x;
    const bool in = n < p.nMol;  // every lane stays: the work below is wave-wide
    const int cap = p.cellCap;
    const int *cnt = bufs.cntB(b);
    const int *slots = bufs.slotsB(b);
    const float4 *cpos = bufs.cellPos;
    int cn = -1;
    if (in) {
        cn = bufs.cellOfB(b)[n];
        if (cn < 0) {  // not binned: empty lists
            if (n < p.nMolMe)
                nnM[n] = 0;
            else {
                nlj[n - p.nMolMe] = 0;
                ncl[n - p.nMolMe] = 0;
            }
            if (p.keepFull) fullNN[n] = 0;
        }
    }
    unsigned long long todo = __ballot(in && cn >= 0 && bufs.slotOf[n] == 0);
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        todo &= todo - 1ull;
        const int c = __shfl(cn, l, 64);
        const int nc = min(cnt[c], cap);
        // the cell's own atoms: lane t holds atom t (ascending) and its position
        int myI = -1;
        float4 myA = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane < nc) {
            myI = slots[(size_t)c * cap + lane];
            myA = cpos[(size_t)c * cap + lane];
        }
        const int ccz = c / (p.cx * p.cy);
        const int ccy = (c - ccz * p.cx * p.cy) / p.cx;
        const int ccx = c - (ccz * p.cy + ccy) * p.cx;
        // lane o < 27: neighbour cell o, its count and the exclusive prefix of counts
        int co = 0;
        if (lane < 27) {
            float shx, shy, shz;
            const int Co = stencil_cell<float>(p, ccx, ccy, ccz, lane, shx, shy, shz);
            co = min(cnt[Co], cap);
        }
        int off = co;  // inclusive scan over the 27 lanes
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const int v = __shfl_up(off, d, 64);
            if (lane >= d) off += v;
        }
        const int total = __shfl(off, 26, 64);
        off -= co;  // exclusive
        // per-atom counters live in lane t
        int nn = 0, nljv = 0, nclv = 0, err = 0;
        for (int base = 0; base < total; base += 64) {
            const int j = base + lane;
            const bool valid = j < total;
            // the neighbour cell of candidate j: last o with off_o <= j (uniform loop over 27)
            int o = 0;
#pragma unroll
            for (int q = 1; q < 27; ++q)
                if (__shfl(off, q, 64) <= j) o = q;
            float shx = 0.f, shy = 0.f, shz = 0.f;
            const int Co = stencil_cell<float>(p, ccx, ccy, ccz, o, shx, shy, shz);
            const int slot = j - __shfl(off, o, 64);
            int k = -1;
            float4 B = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid) {
                k = slots[(size_t)Co * cap + slot];
                B = cpos[(size_t)Co * cap + slot];
            }
            const float3 Bf = make_float3(B.x, B.y, B.z);
            for (int t = 0; t < nc; ++t) {
                const int it = __shfl(myI, t, 64);
                const float3 At = make_float3(__shfl(myA.x, t, 64), __shfl(myA.y, t, 64), __shfl(myA.z, t, 64));
                const float rr = shifted_rr(At, Bf, shx, shy, shz);
                const bool hit = valid && rr < p.rrNebr && k != it;
                const unsigned long long m = __ballot(hit);
                if (!m) continue;
                const int nnT = __shfl(nn, t, 64);
                const int posn = nnT + __popcll(m & below);
                if (it < p.nMolMe) {
                    if (hit) {
                        if (posn < p.nlCap) {
                            nlM[(size_t)it * p.nlCap + posn] = k;
                            if (p.keepFull) fullNBL[(size_t)posn * p.nMol + it] = k;
                        } else {
                            err |= SGK_FLAG_NL_OVERFLOW;
                        }
                    }
                } else {
                    const int ict = it - p.nMolMe;
                    const unsigned long long mlj = __ballot(hit && k < p.nMolMe);
                    const unsigned long long mcl = __ballot(hit && rr < p.closeR2);
                    const int ljT = __shfl(nljv, t, 64), clT = __shfl(nclv, t, 64);
                    if (hit) {
                        if (posn < p.nlCap) {
                            if (p.keepFull) fullNBL[(size_t)posn * p.nMol + it] = k;
                            const unsigned ent = ((unsigned)k << 8) | (unsigned)posn;
                            if (k < p.nMolMe) {
                                const int q = ljT + __popcll(mlj & below);
                                if (q < p.ljCap)
                                    lj[(size_t)q * p.nC + ict] = ent;
                                else
                                    err |= SGK_FLAG_LJ_OVERFLOW;
                            }
                            if (rr < p.closeR2) {
                                const int q = clT + __popcll(mcl & below);
                                if (q < p.clCap)
                                    cl[(size_t)ict * p.clCap + q] = ent;
                                else
                                    err |= SGK_FLAG_CLOSE_OVERFLOW;
                            }
                        } else {
                            err |= SGK_FLAG_NL_OVERFLOW;
                        }
                    }
                    if (lane == t) {
                        nljv += __popcll(mlj);
                        nclv += __popcll(mcl);
                    }
                }
                if (lane == t) nn += __popcll(m);
            }
        }
        // lane t writes atom t's counts (a cut list past the capacity keeps its first entries,
        // as the reference's list of iNebrMax entries does, and raises the overflow flag)
        if (lane < nc) {
            const int nnc = min(nn, p.nlCap);
            if (p.keepFull) fullNN[myI] = nnc;
            if (myI < p.nMolMe) {
                nnM[myI] = nnc;
            } else {
                nlj[myI - p.nMolMe] = min(nljv, p.ljCap);
                ncl[myI - p.nMolMe] = min(nclv, p.clCap);
            }
        }
        if (err) atomicOr(&st->errors, err);
    }
}

template <bool M>
__global__ __launch_bounds__(256) void k_nl(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                            CellBufs bufs, DevState *st, int force,
                                            int fromNext, int *__restrict__ nlM, int *__restrict__ nnM,
                                            unsigned *__restrict__ lj, int *__restrict__ nlj,
                                            unsigned *__restrict__ cl, int *__restrict__ ncl,
                                            int *__restrict__ fullNN, int *__restrict__ fullNBL) {
    const int blk = (p.xcdMap & 1) ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    if (p.nlScheme == 1)
        nlw_body<M>(p, blk, bufs, st, force, fromNext, nlM, nnM, lj, nlj, cl, ncl, fullNN, fullNBL);
    else
        nl_body<M>(p, blk, pos, bufs, st, force, fromNext, nlM, nnM, lj, nlj, cl, ncl, fullNN, fullNBL);
}

// The LDS-staged scheme in a kernel of its own (its 36 KB of LDS would limit the others' occupancy).
template <bool M>
__global__ __launch_bounds__(256) void k_nlt(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                             CellBufs bufs, DevState *st, int force,
                                             int fromNext, int *__restrict__ nlM, int *__restrict__ nnM,
                                             unsigned *__restrict__ lj, int *__restrict__ nlj,
                                             unsigned *__restrict__ cl, int *__restrict__ ncl,
                                             int *__restrict__ fullNN, int *__restrict__ fullNBL) {
    const int blk = (p.xcdMap & 1) ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    nlt_body<M>(p, blk, pos, bufs, st, force, fromNext, nlM, nnM, lj, nlj, cl, ncl, fullNN, fullNBL);
}

// ---- cell-ordered scheme (nl_scheme 3) ---------------------------------------------------------
// BuildNebrListK (SurfaceGrowth.cu:612-711) in two parts of one launch; output as nl_body, bit for
// bit.  The per-atom walk above is bound at once by the L1 (a wave's 64 atoms lie in ~20 different
// cells, so each gather touches ~20 cache lines), by VALU issue and by the scalar unit, and its metal
// atoms - ~480 candidates each, one lane per atom - finish last.  Here:
//   metal part  (blocks [0, nbM)): kNlMetalLanes lanes per metal atom.  The lanes of an atom take
//               two candidates each of every chunk of 8 in the walk order; a ballot over the group
//               gives each hit its list position.
//   carbon part (blocks [nbM, nbM + nbC)): one lane per carbon.
//   Both visit their atoms in the order of their cells (nlOrder, built by build_nl_order; index
//   order without one), so the lanes of a wave hold a few adjacent cells that share their neighbour
//   cells and their candidate loads hit the same cache lines.  Both walk the 27 cells as 9 columns:
//   offsets 3q, 3q+1 and 3q+2 of the table differ only in z, so a column's three cells are taken
//   back to back, the carbon walk in chunks of 4 and the metal one in chunks of 8 that may straddle
//   cells.  A graphene column split over two z layers is ~12 candidates: 3 chunks, where cell by
//   cell it took two cells padded to 8.  Each candidate is one 16-byte load (position and atom
//   index); there is no self test outside the atom's own column unless a short grid axis wraps a
//   neighbour offset onto it.

// Lists of metal atom i, G lanes per atom (l = lane within the group); called by whole groups.  Per
// group step the lanes test candidates base + l and base + G + l of the column (base a multiple of 8).
template <bool M, int G, bool KEEP>
__device__ __forceinline__ void nlm_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                         const CellBufs &bufs, int b, DevState *st, const int *__restrict__ order,
                                         int *__restrict__ nlM, int *__restrict__ nnM, int *__restrict__ fullNN,
                                         int *__restrict__ fullNBL) {
    static_assert(G == 4 || G == 8 || G == 16, "two candidates per lane, 2G a step");
    const int gt = blk * blockDim.x + threadIdx.x;
    const int g = gt / G, l = gt % G;
    const bool active = g < p.nMolMe;
    const int i = active && order ? order[p.nC + g] : g;  // metal atoms by cell (index order without one)
    const int c = active ? bufs.cellOfB(b)[i] : -1;
    const int gshift = (threadIdx.x & 63) & ~(G - 1);
    const unsigned gmask = (1u << G) - 1u, below = (1u << l) - 1u;
    int nn = 0;
    if (c >= 0) {
        const int ccz = c / (p.cx * p.cy);
        const int ccy = (c - ccz * p.cx * p.cy) / p.cx;
        const int ccx = c - (ccz * p.cy + ccy) * p.cx;
        const float3 A = f3(pos[i]);
        const float rrN = p.rrNebr;
        const int *cnt = bufs.cntB(b);
        const bool wrapSelf = p.cx < 3 || p.cy < 3 || p.cz < 3;
        // one column (offsets 3q, 3q+1, 3q+2: same x and y shift) at a time, its cells back to back
        const int cap = p.cellCap;
        const float4 *cpos = bufs.cellPos;
        for (int q = 0; q < 9; ++q) {
            float shx, shy, sz0, sz1, sz2, tx, ty;
            const int C0 = stencil_cell<float>(p, ccx, ccy, ccz, 3 * q, shx, shy, sz0);
            const int C1 = stencil_cell<float>(p, ccx, ccy, ccz, 3 * q + 1, tx, ty, sz1);
            const int C2 = stencil_cell<float>(p, ccx, ccy, ccz, 3 * q + 2, tx, ty, sz2);
            const int n0 = min(cnt[C0], cap), n1 = min(cnt[C1], cap), n2 = min(cnt[C2], cap);
            const int n01 = n0 + n1, tot = n01 + n2;
            const bool self = q == 0 || wrapSelf;
            for (int base = 0; base < tot; base += 2 * G) {
                const int t0 = base + l, t1 = base + G + l;
                const int v0 = min(t0, tot - 1), v1 = min(t1, tot - 1);
                const int c0 = v0 < n0 ? C0 : (v0 < n01 ? C1 : C2), c1 = v1 < n0 ? C0 : (v1 < n01 ? C1 : C2);
                const int s0 = v0 < n0 ? v0 : (v0 < n01 ? v0 - n0 : v0 - n01);
                const int s1 = v1 < n0 ? v1 : (v1 < n01 ? v1 - n0 : v1 - n01);
                const float z0 = v0 < n0 ? sz0 : (v0 < n01 ? sz1 : sz2), z1 = v1 < n0 ? sz0 : (v1 < n01 ? sz1 : sz2);
                const float4 q0 = cpos[(size_t)c0 * cap + s0], q1 = cpos[(size_t)c1 * cap + s1];
                const int k0 = __float_as_int(q0.w), k1 = __float_as_int(q1.w);
                const float rr0 = shifted_rr(A, make_float3(q0.x, q0.y, q0.z), shx, shy, z0);
                const float rr1 = shifted_rr(A, make_float3(q1.x, q1.y, q1.z), shx, shy, z1);
                const bool h0 = t0 < tot && rr0 < rrN && (!self || k0 != i);
                const bool h1 = t1 < tot && rr1 < rrN && (!self || k1 != i);
                const unsigned g0 = (unsigned)(__ballot(h0) >> gshift) & gmask;
                const unsigned g1 = (unsigned)(__ballot(h1) >> gshift) & gmask;
                const int at0 = nn + __popc(g0 & below), at1 = nn + __popc(g0) + __popc(g1 & below);
                if (h0 && at0 < p.nlCap) {
                    nlM[(size_t)i * p.nlCap + at0] = k0;
                    if (KEEP) fullNBL[(size_t)at0 * p.nMol + i] = k0;
                }
                if (h1 && at1 < p.nlCap) {
                    nlM[(size_t)i * p.nlCap + at1] = k1;
                    if (KEEP) fullNBL[(size_t)at1 * p.nMol + i] = k1;
                }
                nn += __popc(g0) + __popc(g1);
            }
        }
    }
    if (active && l == 0)
        nl_finish(p, i, true, i - p.nMolMe, nn, 0, 0, nn > p.nlCap ? SGK_FLAG_NL_OVERFLOW : 0, nnM, nullptr, nullptr,
                  fullNN, st);
}

// One chunk of 4 candidates of a carbon atom's column walk (SELF: the column may hold the atom
// itself): candidates already loaded, each with its own z shift (the three cells of a column share
// the x and y shift); t >= valid are not candidates.
template <bool SELF, bool KEEP, bool SHIFT>
__device__ __forceinline__ void nlc_chunk(const DevParams &p, int i, int ic, const float3 &A, float shx, float shy,
                                          const float (&shz)[4], const float4 (&P)[4], int valid, int &nn,
                                          int &nljv, int &nclv, int &err, unsigned *__restrict__ lj,
                                          unsigned *__restrict__ cl, int *__restrict__ fullNBL) {
    int k[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) k[t] = __float_as_int(P[t].w);
    const float rrN = p.rrNebr, rrC = p.closeR2;
    const int nMe = p.nMolMe;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float rr = shifted_rr<SHIFT>(A, make_float3(P[t].x, P[t].y, P[t].z), shx, shy, shz[t]);
        bool hit = rr < rrN && t < valid;
        if (SELF) hit = hit && k[t] != i;
        if (hit && (k[t] < nMe || rr < rrC)) {
            if (nn < p.nlCap) {
                const unsigned ent = ((unsigned)k[t] << 8) | (unsigned)nn;
                if (k[t] < nMe) {
                    if (nljv < p.ljCap)
                        lj[(size_t)nljv * p.nC + ic] = ent;
                    else
                        err |= SGK_FLAG_LJ_OVERFLOW;
                    ++nljv;
                }
                if (rr < rrC) {
                    if (nclv < p.clCap)
                        cl[(size_t)ic * p.clCap + nclv] = ent;
                    else
                        err |= SGK_FLAG_CLOSE_OVERFLOW;
                    ++nclv;
                }
            }
        }
        if (KEEP && hit && nn < p.nlCap) fullNBL[(size_t)nn * p.nMol + i] = k[t];
        nn += hit ? 1 : 0;
    }
}

// Lists of carbon atom i in cell c (one lane).
template <bool KEEP>
__device__ __forceinline__ void nlc_walk(const DevParams &p, int i, int c, const float3 A, const int *__restrict__ cnt,
                                         const float4 *__restrict__ cpos,
                                         DevState *st, unsigned *__restrict__ lj, int *__restrict__ nlj,
                                         unsigned *__restrict__ cl, int *__restrict__ ncl,
                                         int *__restrict__ fullNN, int *__restrict__ fullNBL) {
    const int ic = i - p.nMolMe;
    const int ccz = c / (p.cx * p.cy);
    const int ccy = (c - ccz * p.cx * p.cy) / p.cx;
    const int ccx = c - (ccz * p.cy + ccy) * p.cx;
    const bool wrapSelf = p.cx < 3 || p.cy < 3 || p.cz < 3;
    int nn = 0, nljv = 0, nclv = 0, err = 0;
    // The offsets 3q, 3q+1, 3q+2 of the table share their x and y components (the z component is
    // o % 3), so the walk takes them as one column: its three cells back to back, in chunks of 4
    // that may straddle cells.  A sheet split over two z layers then walks ~12 slots per column
    // instead of two padded cells of 8.
    const int cap = p.cellCap;
    for (int q = 0; q < 9; ++q) {
        float shx, shy, sz0, sz1, sz2, tx, ty;
        const int C0 = stencil_cell<float>(p, ccx, ccy, ccz, 3 * q, shx, shy, sz0);
        const int C1 = stencil_cell<float>(p, ccx, ccy, ccz, 3 * q + 1, tx, ty, sz1);
        const int C2 = stencil_cell<float>(p, ccx, ccy, ccz, 3 * q + 2, tx, ty, sz2);
        const int n0 = min(cnt[C0], cap), n1 = min(cnt[C1], cap), n2 = min(cnt[C2], cap);
        const int n01 = n0 + n1, tot = n01 + n2;
        // candidate v of the column is entry v + o_k of the cell-ordered positions (k its cell);
        // indices fit 32 bits (ncell * cellCap < 2^31 is checked at set-up)
        const int o0 = C0 * cap, o1 = C1 * cap - n0, o2 = C2 * cap - n01;
        const bool self = q == 0 || wrapSelf;
        // no lane of the wave sees this column through a periodic face: the shifts are all zero
        const bool plain = __all(shx == 0.f && shy == 0.f && sz0 == 0.f && sz1 == 0.f && sz2 == 0.f);
#define SGK_NLC_COLUMN(SELF, SHIFT)                                                                          \
    for (int base = 0; base < tot; base += 4) {                                                             \
        float4 P[4];                                                                                        \
        float shz[4];                                                                                       \
        _Pragma("unroll") for (int t = 0; t < 4; ++t) {                                                     \
            const int v = min(base + t, tot - 1);                                                           \
            const bool a0 = v < n0, a1 = v < n01;                                                           \
            const int o = a0 ? o0 : (a1 ? o1 : o2);                                                         \
            shz[t] = a0 ? sz0 : (a1 ? sz1 : sz2);                                                           \
            P[t] = cpos[(unsigned)(o + v)];                                                                 \
        }                                                                                                   \
        nlc_chunk<SELF, KEEP, SHIFT>(p, i, ic, A, shx, shy, shz, P, tot - base, nn, nljv, nclv, err, lj, cl, \
                                     fullNBL);                                                              \
    }
        if (self) {
            if (plain) {
                SGK_NLC_COLUMN(true, false);
            } else {
                SGK_NLC_COLUMN(true, true);
            }
        } else {
            if (plain) {
                SGK_NLC_COLUMN(false, false);
            } else {
                SGK_NLC_COLUMN(false, true);
            }
        }
#undef SGK_NLC_COLUMN
    }
    if (nn > p.nlCap) err |= SGK_FLAG_NL_OVERFLOW;
    nl_finish(p, i, false, ic, nn, nljv, nclv, err, nullptr, nlj, ncl, fullNN, st);
}

// Carbon part: lane t builds the lists of carbon order[t] (index order without an order).
template <bool M, bool KEEP>
__device__ __forceinline__ void nlc_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                         const CellBufs &bufs, int b, DevState *st, const int *__restrict__ order,
                                         unsigned *__restrict__ lj, int *__restrict__ nlj, unsigned *__restrict__ cl,
                                         int *__restrict__ ncl, int *__restrict__ fullNN,
                                         int *__restrict__ fullNBL) {
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= p.nC) return;
    const int i = order ? order[t] : p.nMolMe + t;
    const int c = bufs.cellOfB(b)[i];
    if (c < 0) {  // not binned: empty lists
        nl_finish(p, i, false, i - p.nMolMe, 0, 0, 0, 0, nullptr, nlj, ncl, fullNN, st);
        return;
    }
    nlc_walk<KEEP>(p, i, c, f3(pos[i]), bufs.cntB(b), bufs.cellPos, st, lj, nlj, cl, ncl, fullNN,
                   fullNBL);
}

template <bool M, bool KEEP>
__global__ __launch_bounds__(256) void k_nlc(DevParams p, int nbM, const typename Prec<M>::P4 *__restrict__ pos,
                                             CellBufs bufs, DevState *st, int force, int fromNext,
                                             const int *__restrict__ order, int *__restrict__ nlM,
                                             int *__restrict__ nnM, unsigned *__restrict__ lj,
                                             int *__restrict__ nlj, unsigned *__restrict__ cl,
                                             int *__restrict__ ncl, int *__restrict__ fullNN,
                                             int *__restrict__ fullNBL) {
    if (!force && !st->nebrNow) return;
    const int b = fromNext ? (st->cellbuf ^ 1) : st->cellbuf;
    // xcd_map bit 0: each XCD walks a contiguous 1/8 of each range, so the cells a workgroup shares
    // with the next ones in the list-build order stay in one L2 (the lists do not depend on it)
    const bool xm = p.xcdMap & 1;
    const int bb = blockIdx.x;
    if (bb < nbM)
        nlm_body<M, kNlMetalLanes, KEEP>(p, xm ? xcd_block(bb, nbM) : bb, pos, bufs, b, st, order, nlM, nnM, fullNN,
                                         fullNBL);
    else
        nlc_body<M, KEEP>(p, xm ? xcd_block(bb - nbM, gridDim.x - nbM) : bb - nbM, pos, bufs, b, st, order, lj, nlj,
                          cl, ncl, fullNN, fullNBL);
}

// ---- list-build order (nl_scheme 3) ----------------------------------------------------------------
// A counting sort of the atoms by the cell of their current position (the binning formula, cells
// numbered column by column so that an atom drifting to the next z layer - a sheet lying on a cell
// boundary splits across two - stays next to its column; one bucket for positions outside the grid):
// carbons in buckets [0, ncell], metal atoms in [ncell + 1, 2 ncell + 1], so that nlOrder[0, nC) are
// the carbons and nlOrder[nC, nMol) the metal atoms.  It is built outside the step loop and ages as
// atoms move; it decides which lanes build which atom's lists, never the lists.
// kOrdChunk (sgk_internal.h): buckets per scan block, 4 per thread, read as one 16-byte load

// Bucket of atom n in the list-build order: its column-major cell, carbons first.
__device__ __forceinline__ int ord_bucket(const DevParams &p, const float3 q, int n) {
#pragma clang fp contract(off)
    const float sx = q.x + 0.5f * p.Lx, sy = q.y + 0.5f * p.Ly, sz = q.z + 0.5f * p.Lz;
    const int cx = (int)(sx * p.iwx), cy = (int)(sy * p.iwy), cz = (int)(sz * p.iwz);
    const int c = (cy * p.cx + cx) * p.cz + cz;  // column-major: a column's z cells are adjacent
    int k = (c < 0 || c >= p.ncell) ? p.ncell : c;
    if (n < p.nMolMe) k += p.ncell + 1;  // metal buckets after the carbon ones
    return k;
}

template <bool M>
__global__ __launch_bounds__(256) void k_ord_count(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                                   int *__restrict__ cnt, int *__restrict__ key,
                                                   int *__restrict__ slot) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nMol) return;
    const int k = ord_bucket(p, f3(pos[n]), n);
    // lanes with the same bucket share one atomicAdd, as in binning
    const int lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(1), grp = 1ull << lane;
    int leader = lane;
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        const int kl = __shfl(k, l, 64);
        const unsigned long long m = __ballot(k == kl);
        if (k == kl) {
            grp = m;
            leader = l;
        }
        todo &= ~m;
    }
    int base = 0;
    if (lane == leader) base = atomicAdd(&cnt[k], __popcll(grp));
    base = __shfl(base, leader, 64);
    key[n] = k;
    slot[n] = base + __popcll(grp & below);
}

// Total of each block of kOrdChunk buckets [0, nb).
// The 4 counts of thread t of scan block b: [b kOrdChunk + 4t, +4), zero past nb.  Bucket arrays are
// allocated 256-byte aligned, so the full groups are single 16-byte loads.
__device__ __forceinline__ int4 ord_load4(const int *__restrict__ cnt, int nb, long long base) {
    if (base + 3 < nb) return *(const int4 *)(cnt + base);
    int4 v = make_int4(0, 0, 0, 0);
    if (base < nb) v.x = cnt[base];
    if (base + 1 < nb) v.y = cnt[base + 1];
    if (base + 2 < nb) v.z = cnt[base + 2];
    return v;
}

__global__ __launch_bounds__(256) void k_ord_blocksum(int nb, const int *__restrict__ cnt, int *__restrict__ blockSum) {
    __shared__ int sh[256 + 1];
    const long long base = (long long)blockIdx.x * kOrdChunk + threadIdx.x * 4;
    const int4 v = ord_load4(cnt, nb, base);
    sh[threadIdx.x] = v.x + v.y + v.z + v.w;
    __syncthreads();
    block_scan4(sh, blockDim.x);
    if (threadIdx.x == 0) blockSum[blockIdx.x] = sh[blockDim.x];
}

// Counts -> exclusive starts, in place: the block totals before this block, then the block's own scan.
__global__ __launch_bounds__(256) void k_ord_scan(int nb, int *__restrict__ cnt, const int *__restrict__ blockSum) {
    __shared__ int sh[256 + 1];
    int pre = 0;
    for (int j = threadIdx.x; j < (int)blockIdx.x; j += blockDim.x) pre += blockSum[j];
    sh[threadIdx.x] = pre;
    __syncthreads();
    block_scan4(sh, blockDim.x);
    pre = sh[blockDim.x];
    __syncthreads();
    const long long base = (long long)blockIdx.x * kOrdChunk + threadIdx.x * 4;
    const int4 v = ord_load4(cnt, nb, base);
    sh[threadIdx.x] = v.x + v.y + v.z + v.w;
    __syncthreads();
    block_scan4(sh, blockDim.x);
    const int r0 = pre + sh[threadIdx.x], r1 = r0 + v.x, r2 = r1 + v.y, r3 = r2 + v.z;
    if (base + 3 < nb) {
        *(int4 *)(cnt + base) = make_int4(r0, r1, r2, r3);
    } else {
        if (base < nb) cnt[base] = r0;
        if (base + 1 < nb) cnt[base + 1] = r1;
        if (base + 2 < nb) cnt[base + 2] = r2;
    }
}

__global__ __launch_bounds__(256) void k_ord_scatter(DevParams p, const int *__restrict__ start,
                                                     const int *__restrict__ key, const int *__restrict__ slot,
                                                     int *__restrict__ order) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nMol) return;
    order[start[key[n]] + slot[n]] = n;
}

// Batched: the step path (lists built from the new cell buffer before the commit).
template <bool M>
__global__ __launch_bounds__(256) void kb_nl(BatchArgs a) {
    const int2 e = a.atomTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const DevParams &p = B.p;
    if (p.nlScheme == 1)
        nlw_body<M>(p, e.y, rep_cells(a, B), a.state + e.x, 0, 1, a.nlM + B.metalOff * p.nlCap, a.nnM + B.metalOff,
                    a.lj + B.carbonOff * p.ljCap, a.nlj + B.carbonOff, a.cl + B.carbonOff * p.clCap,
                    a.ncl + B.carbonOff, nullptr, nullptr);
    else
        nl_body<M>(p, e.y, (const typename Prec<M>::P4 *)a.pos + B.atomOff, rep_cells(a, B), a.state + e.x, 0, 1,
                   a.nlM + B.metalOff * p.nlCap, a.nnM + B.metalOff, a.lj + B.carbonOff * p.ljCap,
                   a.nlj + B.carbonOff, a.cl + B.carbonOff * p.clCap, a.ncl + B.carbonOff, nullptr, nullptr);
}

template <bool M>
__global__ __launch_bounds__(256) void kb_nlt(BatchArgs a) {
    const int2 e = a.atomTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const DevParams &p = B.p;
    nlt_body<M>(p, e.y, (const typename Prec<M>::P4 *)a.pos + B.atomOff, rep_cells(a, B), a.state + e.x, 0, 1,
                a.nlM + B.metalOff * p.nlCap, a.nnM + B.metalOff, a.lj + B.carbonOff * p.ljCap, a.nlj + B.carbonOff,
                a.cl + B.carbonOff * p.clCap, a.ncl + B.carbonOff, nullptr, nullptr);
}

// Batched list-build order: one counting sort over every replica, replica r's buckets at
// [ordOff, ordOff + 2 ncell + 2), so its atoms land at [atomOff, atomOff + nMol) of the order, by
// local index.  The slot within a bucket comes from a plain atomicAdd: it decides only which lane
// builds which atom's lists.
template <bool M>
__global__ __launch_bounds__(256) void kb_ord_count(BatchArgs a, int *__restrict__ cnt, int *__restrict__ key,
                                                    int *__restrict__ slot) {
    const int2 e = a.atomTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const int n = e.y * blockDim.x + threadIdx.x;
    if (n >= B.p.nMol) return;
    const int k = B.ordOff + ord_bucket(B.p, f3(((const typename Prec<M>::P4 *)a.pos)[B.atomOff + n]), n);
    key[B.atomOff + n] = k;
    slot[B.atomOff + n] = atomicAdd(&cnt[k], 1);
}

__global__ __launch_bounds__(256) void kb_ord_scatter(BatchArgs a, const int *__restrict__ start,
                                                      const int *__restrict__ key, const int *__restrict__ slot,
                                                      int *__restrict__ order) {
    const int2 e = a.atomTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const int n = e.y * blockDim.x + threadIdx.x;
    if (n >= B.p.nMol) return;
    order[start[key[B.atomOff + n]] + slot[B.atomOff + n]] = n;
}

// Batched cell-ordered list build (the step path: from the new cell buffer before the commit, only
// replicas that rebuild this step).
template <bool M>
__global__ __launch_bounds__(256) void kb_nlc(BatchArgs a) {
    const int2 e = a.nlTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    const DevParams &p = B.p;
    DevState *st = a.state + e.x;
    if (!st->nebrNow) return;
    const int b = st->cellbuf ^ 1;
    const typename Prec<M>::P4 *pos = (const typename Prec<M>::P4 *)a.pos + B.atomOff;
    const int *order = a.order ? a.order + B.atomOff : nullptr;
    if (e.y < B.nbNlM)
        nlm_body<M, kNlMetalLanes, false>(p, e.y, pos, rep_cells(a, B), b, st, order, a.nlM + B.metalOff * p.nlCap,
                                          a.nnM + B.metalOff, nullptr, nullptr);
    else
        nlc_body<M, false>(p, e.y - B.nbNlM, pos, rep_cells(a, B), b, st, order, a.lj + B.carbonOff * p.ljCap,
                           a.nlj + B.carbonOff, a.cl + B.carbonOff * p.clCap, a.ncl + B.carbonOff, nullptr, nullptr);
}

// Make the freshly built buffer current (API path; the step path does this in its reduction).
__global__ void k_commit(DevState *st) {
    st->cellbuf ^= 1;
    st->rebuilds += 1;
}

// EvalRdfK (SurfaceGrowth.cu:1895-1999): metal-metal pairs of the 27-cell stencil of the current
// cells, r^2 < rangeRdf^2, bin = int(sqrt(r^2) * intervalRdf).  One thread per metal atom.
// Histogram in LDS, merged with one atomic per bin and block; counts are order independent.
template <bool M>
__device__ __forceinline__ void rdf_body(const DevParams &p, int blk, const typename Prec<M>::P4 *__restrict__ pos,
                                         const CellBufs &bufs, const DevState *st, unsigned *__restrict__ hist,
                                         unsigned *sh, int *errors) {
    for (int b = threadIdx.x; b < p.sizeHistRdf; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    const int i = blk * blockDim.x + threadIdx.x;
    const int c = i < p.nMolMe ? bufs.cellOfB(st->cellbuf)[i] : -1;
    int edge = 0;
    if (c >= 0) {
        const int cb = st->cellbuf;
        const int *cnt = bufs.cntB(cb);
        const int *slots = bufs.slotsB(cb);
        const int ccz = c / (p.cx * p.cy);
        const int ccy = (c - ccz * p.cx * p.cy) / p.cx;
        const int ccx = c - (ccz * p.cy + ccy) * p.cx;
        const float3 A = f3(pos[i]);
        const float range2 = p.rangeRdf2;
        for (int o = 0; o < 27; ++o) {
            float shx, shy, shz;
            const int C = stencil_cell<float>(p, ccx, ccy, ccz, o, shx, shy, shz);
            const int n = min(cnt[C], p.cellCap);
            const int *s = slots + (size_t)C * p.cellCap;
            for (int t = 0; t < n; ++t) {
                const int k = s[t];
                if (k >= p.nMolMe) continue;
                const float rr = shifted_rr(A, f3(pos[k]), shx, shy, shz);
                if (rr < range2 && k != i) {
                    const int bin = (int)(sqrtf(rr) * p.intervalRdf);  // correctly rounded sqrt
                    if (bin >= p.sizeHistRdf)
                        edge = 1;
                    else
                        atomicAdd(&sh[bin], 1u);
                }
            }
        }
    }
    if (edge) atomicOr(errors, SGK_FLAG_RDF_EDGE);
    __syncthreads();
    for (int b = threadIdx.x; b < p.sizeHistRdf; b += blockDim.x)
        if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

template <bool M>
__global__ __launch_bounds__(256) void k_rdf(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                             CellBufs bufs, DevState *st, unsigned *__restrict__ hist) {
    extern __shared__ unsigned sh[];
    rdf_body<M>(p, blockIdx.x, pos, bufs, st, hist, sh, &st->errors);
}

template <bool M>
__global__ __launch_bounds__(256) void kb_rdf(BatchArgs a) {
    extern __shared__ unsigned sh[];
    const int2 e = a.metalTab[blockIdx.x];
    const BatchRep &B = a.reps[e.x];
    rdf_body<M>(B.p, e.y, (const typename Prec<M>::P4 *)a.pos + B.atomOff, rep_cells(a, B), a.state + e.x,
                a.hist + (size_t)e.x * B.p.sizeHistRdf, sh, &a.state[e.x].errors);
}

// ---- complete-range RDF (no counterpart in the reference) ------------------------------------
// The reference stencil above finds every pair only up to one cell width.  This variant counts every
// ordered metal-metal pair within rangeRdf at the current positions: metal atoms are sorted into a
// grid of cells at least rangeRdf/S wide (count, scan, scatter of float positions with the atom index
// in w), then each atom's lanes walk the (2S+1)^3 cells around it, skipping those whose box lies
// beyond rangeRdf from the atom's cell (none for S = 2; S = 2 is the default and measured fastest on
// a 100,000-atom cluster: finer cells cull more of the cube but cost more per cell).  A pair is counted once, from the atom in the lower cell
// (or, in one cell, the lower index), and adds 2 to its bin - the reference's histogram counts both
// orders.  Displacements use the single-image wrap, so rangeRdf must stay below half the box; bins
// use the reference's formula.
__device__ __forceinline__ int rdf_cell_axis(float x, float L, float iw, int n) {
    const int c = (int)((x + 0.5f * L) * iw);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

template <bool M>
__global__ __launch_bounds__(256) void k_rdfc_count(DevParams p, RdfGrid g, const typename Prec<M>::P4 *__restrict__ pos,
                                                    int *__restrict__ cnt, int *__restrict__ key, int *__restrict__ slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nMolMe) return;
    const float3 q = f3(pos[i]);
    const int c = (rdf_cell_axis(q.z, p.Lz, g.iwz, g.nz) * g.ny + rdf_cell_axis(q.y, p.Ly, g.iwy, g.ny)) * g.nx +
                  rdf_cell_axis(q.x, p.Lx, g.iwx, g.nx);
    key[i] = c;
    slot[i] = atomicAdd(&cnt[c], 1);
}

template <bool M>
__global__ __launch_bounds__(256) void k_rdfc_scatter(DevParams p, const typename Prec<M>::P4 *__restrict__ pos,
                                                      const int *__restrict__ start, const int *__restrict__ key,
                                                      const int *__restrict__ slot, float4 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nMolMe) return;
    const float3 q = f3(pos[i]);
    out[start[key[i]] + slot[i]] = make_float4(q.x, q.y, q.z, __int_as_float(i));
}

__device__ __forceinline__ float wrap_once(float d, float L) {
    return d >= 0.5f * L ? d - L : (d < -0.5f * L ? d + L : d);
}

constexpr int kRdfLanes = 8;

template <bool M>
__global__ __launch_bounds__(256) void k_rdfc_pairs(DevParams p, RdfGrid g, const typename Prec<M>::P4 *__restrict__ pos,
                                                    const int *__restrict__ start, const int *__restrict__ key,
                                                    const float4 *__restrict__ sorted, unsigned *__restrict__ hist,
                                                    int *errors) {
#pragma clang fp contract(off)
    extern __shared__ unsigned sh[];
    for (int b = threadIdx.x; b < p.sizeHistRdf; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = gt / kRdfLanes, l = gt % kRdfLanes;
    int edge = 0;
    if (q < p.nMolMe) {
        // atoms in cell order: the atoms of a wave share their stencil
        const float4 me = sorted[q];
        const int i = __float_as_int(me.w);
        const float3 A = make_float3(me.x, me.y, me.z);
        const int ci = key[i];
        const int cz = ci / (g.nx * g.ny), cy = (ci / g.nx) % g.ny, cx = ci % g.nx;
        // stencil extents per axis: 2s+1 cells around the atom's, or the whole axis once
        const int tx = 2 * g.sx + 1 < g.nx ? 2 * g.sx + 1 : g.nx;
        const int ty = 2 * g.sy + 1 < g.ny ? 2 * g.sy + 1 : g.ny;
        const int tz = 2 * g.sz + 1 < g.nz ? 2 * g.sz + 1 : g.nz;
        const int x0 = tx < g.nx ? cx - g.sx : 0, y0 = ty < g.ny ? cy - g.sy : 0, z0 = tz < g.nz ? cz - g.sz : 0;
        const int ncells = tx * ty * tz;
        const float range2 = p.rangeRdf2;
        for (int t = l; t < ncells; t += kRdfLanes) {
            const int ox = t % tx, oy = (t / tx) % ty, oz = t / (tx * ty);
            // closest approach of the two cell boxes along each axis (none where the stencil spans
            // the whole axis): beyond rangeRdf, the cell holds no pair in range
            const float gx = tx < g.nx ? (float)max(abs(ox - g.sx) - 1, 0) * g.wx : 0.f;
            const float gy = ty < g.ny ? (float)max(abs(oy - g.sy) - 1, 0) * g.wy : 0.f;
            const float gz = tz < g.nz ? (float)max(abs(oz - g.sz) - 1, 0) * g.wz : 0.f;
            if (gx * gx + gy * gy + gz * gz >= g.cull2) continue;
            int x = x0 + ox, y = y0 + oy, z = z0 + oz;
            x += x < 0 ? g.nx : (x >= g.nx ? -g.nx : 0);
            y += y < 0 ? g.ny : (y >= g.ny ? -g.ny : 0);
            z += z < 0 ? g.nz : (z >= g.nz ? -g.nz : 0);
            const int c = (z * g.ny + y) * g.nx + x;
            if (c < ci) continue;  // this cell's pairs with i are counted from the other side
            const int b0 = start[c], b1 = start[c + 1];
            for (int j = b0; j < b1; ++j) {
                const float4 e = sorted[j];
                if (c == ci && __float_as_int(e.w) <= i) continue;
                const float dx = wrap_once(A.x - e.x, p.Lx), dy = wrap_once(A.y - e.y, p.Ly),
                            dz = wrap_once(A.z - e.z, p.Lz);
                const float rr = dx * dx + dy * dy + dz * dz;
                if (rr < range2) {
                    const int bin = (int)(sqrtf(rr) * p.intervalRdf);
                    if (bin >= p.sizeHistRdf)
                        edge = 1;
                    else
                        atomicAdd(&sh[bin], 2u);
                }
            }
        }
    }
    if (edge) atomicOr(errors, SGK_FLAG_RDF_EDGE);
    __syncthreads();
    for (int b = threadIdx.x; b < p.sizeHistRdf; b += blockDim.x)
        if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

__global__ __launch_bounds__(256) void k_fill_i32(int *__restrict__ a, long long n, int v) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        a[t] = v;
}

// ---- launchers ---------------------------------------------------------------------------------
static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// The RDF grid of a system: cells at least rangeRdf/S wide per axis (a pair within range then lies
// within S cells); 0 cells when rangeRdf reaches half the box (the single-image wrap would miss
// pairs).
RdfGrid rdf_grid(const DevParams &p, int S) {
    RdfGrid g{};
    const float range = sqrtf(p.rangeRdf2);
    if (!(range > 0.f) || range >= 0.5f * p.Lx || range >= 0.5f * p.Ly || range >= 0.5f * p.Lz) return g;
    S = std::min(std::max(S, 2), 6);
    // 1e-4 wider than range/S: the float rounding of a cell index (~1e-7 of the box) can then never
    // put a pair within range S + 1 cells apart
    const float w = range / (float)S * 1.0001f;
    g.nx = std::max(1, (int)(p.Lx / w));
    g.ny = std::max(1, (int)(p.Ly / w));
    g.nz = std::max(1, (int)(p.Lz / w));
    // the cell index uses (x + L/2) * n / L, so each cell is at least L/n >= w wide
    g.iwx = (float)g.nx / p.Lx;
    g.iwy = (float)g.ny / p.Ly;
    g.iwz = (float)g.nz / p.Lz;
    // culling uses the nominal width range/S (<= L/n) and a range 0.1 % long, far above the ~1e-7
    // of the box by which a rounded cell index can misplace an atom
    g.wx = g.wy = g.wz = range / (float)S;
    g.cull2 = range * range * 1.002f;
    g.sx = g.sy = g.sz = S;
    g.ncell = g.nx * g.ny * g.nz;
    return g;
}

void launch_bin(sgk_ctx *c, bool force) {
    if (c->dp.nMol == 0) return;
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    if (c->mixed)
        hipLaunchKernelGGL(k_bin<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (const double4 *)c->pos, cell_bufs(c),
                           c->state, (int)force);
    else
        hipLaunchKernelGGL(k_bin<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (const float4 *)c->pos, cell_bufs(c),
                           c->state, (int)force);
    profile_end(c, "bin");
}

void launch_sortcells(sgk_ctx *c, bool force) {
    if (c->dp.nMol == 0) return;
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    if (c->mixed)
        hipLaunchKernelGGL(k_sortcells<true>, dim3(g), dim3(256), 0, c->stream, c->dp, (const double4 *)c->pos,
                           cell_bufs(c), c->state, (int)force);
    else
        hipLaunchKernelGGL(k_sortcells<false>, dim3(g), dim3(256), 0, c->stream, c->dp, (const float4 *)c->pos,
                           cell_bufs(c), c->state, (int)force);
    profile_end(c, "sortcells");
}

void launch_nl(sgk_ctx *c, bool force) {
    if (c->dp.nMol == 0) return;
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMol, 256);
    const int fromNext = force ? 0 : 1;  // step path: build from the new buffer before the commit
    if (c->dp.nlScheme == 3) {
        const int nbM = (int)blocks_for((long long)c->dp.nMolMe * kNlMetalLanes, 256);
        const unsigned gc = (unsigned)nbM + blocks_for(c->dp.nC, 256);
        const int *order = c->orderValid ? c->nlOrder : nullptr;
#define SGK_NLC(MM, KK, P4)                                                                                      \
    hipLaunchKernelGGL((k_nlc<MM, KK>), dim3(gc), dim3(256), 0, c->stream, c->dp, nbM, (const P4 *)c->pos,           \
                       cell_bufs(c), c->state, (int)force, fromNext, order, c->nlM, c->nnM, c->lj, c->nlj, c->cl,  \
                       c->ncl, c->fullNN, c->fullNBL)
        const bool keep = c->dp.keepFull != 0;
        if (c->mixed) {
            if (keep) SGK_NLC(true, true, double4); else SGK_NLC(true, false, double4);
        } else {
            if (keep) SGK_NLC(false, true, float4); else SGK_NLC(false, false, float4);
        }
#undef SGK_NLC
        profile_end(c, "nl");
        return;
    }
    const bool tiled = c->dp.nlScheme == 2;
#define SGK_NL(K, MM, P4)                                                                                         \
    hipLaunchKernelGGL(K<MM>, dim3(g), dim3(256), 0, c->stream, c->dp, (const P4 *)c->pos, cell_bufs(c), c->state, \
                       (int)force, fromNext, c->nlM, c->nnM, c->lj, c->nlj, c->cl, c->ncl, c->fullNN, c->fullNBL)
    if (c->mixed) {
        if (tiled) SGK_NL(k_nlt, true, double4); else SGK_NL(k_nl, true, double4);
    } else {
        if (tiled) SGK_NL(k_nlt, false, float4); else SGK_NL(k_nl, false, float4);
    }
#undef SGK_NL
    profile_end(c, "nl");
}

void build_nl_order(sgk_ctx *c) {
    const DevParams &p = c->dp;
    if (!c->nlOrder || p.nMol == 0) return;
    const int nb = 2 * p.ncell + 2;
    const unsigned nblk = blocks_for(nb, kOrdChunk);
    profile_begin(c);
    (void)hipMemsetAsync(c->ordCnt, 0, sizeof(int) * (size_t)nb, c->stream);
    if (c->mixed)
        hipLaunchKernelGGL(k_ord_count<true>, dim3(blocks_for(p.nMol, 256)), dim3(256), 0, c->stream, p,
                           (const double4 *)c->pos, c->ordCnt, c->ordKey, c->ordSlot);
    else
        hipLaunchKernelGGL(k_ord_count<false>, dim3(blocks_for(p.nMol, 256)), dim3(256), 0, c->stream, p,
                           (const float4 *)c->pos, c->ordCnt, c->ordKey, c->ordSlot);
    hipLaunchKernelGGL(k_ord_blocksum, dim3(nblk), dim3(256), 0, c->stream, nb, c->ordCnt, c->ordBlock);
    hipLaunchKernelGGL(k_ord_scan, dim3(nblk), dim3(256), 0, c->stream, nb, c->ordCnt, c->ordBlock);
    hipLaunchKernelGGL(k_ord_scatter, dim3(blocks_for(p.nMol, 256)), dim3(256), 0, c->stream, p, c->ordCnt, c->ordKey,
                       c->ordSlot, c->nlOrder);
    profile_end(c, "nlorder");
    c->orderValid = true;
    c->orderAge = 0;
}

void launch_batch_sortcells(sgk_batch *b) {
    if (b->mixed)
        hipLaunchKernelGGL(kb_sortcells<true>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args());
    else
        hipLaunchKernelGGL(kb_sortcells<false>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args());
}

void build_batch_nl_order(sgk_batch *b) {
    if (!b->order || b->nAtoms == 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min(blocks_for(b->nOrdBuckets, 256), 4096u)), dim3(256), 0, b->stream,
                       b->ordCnt, b->nOrdBuckets, 0);
    if (b->mixed)
        hipLaunchKernelGGL(kb_ord_count<true>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), b->ordCnt,
                           b->ordKey, b->ordSlot);
    else
        hipLaunchKernelGGL(kb_ord_count<false>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), b->ordCnt,
                           b->ordKey, b->ordSlot);
    const int nb = (int)b->nOrdBuckets;
    const unsigned nblk = blocks_for(nb, kOrdChunk);
    hipLaunchKernelGGL(k_ord_blocksum, dim3(nblk), dim3(256), 0, b->stream, nb, b->ordCnt, b->ordBlock);
    hipLaunchKernelGGL(k_ord_scan, dim3(nblk), dim3(256), 0, b->stream, nb, b->ordCnt, b->ordBlock);
    hipLaunchKernelGGL(kb_ord_scatter, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args(), b->ordCnt, b->ordKey,
                       b->ordSlot, b->order);
    b->orderValid = true;
    b->orderAge = 0;
}

void launch_batch_nl(sgk_batch *b) {
    if (b->cfg.nl_scheme == 3 && b->nlTab) {
        if (b->mixed)
            hipLaunchKernelGGL(kb_nlc<true>, dim3(b->nNlBlocks), dim3(256), 0, b->stream, b->args());
        else
            hipLaunchKernelGGL(kb_nlc<false>, dim3(b->nNlBlocks), dim3(256), 0, b->stream, b->args());
        return;
    }
    const bool tiled = b->cfg.nl_scheme == 2;
    if (b->mixed) {
        if (tiled)
            hipLaunchKernelGGL(kb_nlt<true>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args());
        else
            hipLaunchKernelGGL(kb_nl<true>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args());
    } else {
        if (tiled)
            hipLaunchKernelGGL(kb_nlt<false>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args());
        else
            hipLaunchKernelGGL(kb_nl<false>, dim3(b->nAtomBlocks), dim3(256), 0, b->stream, b->args());
    }
}

void launch_batch_rdf(sgk_batch *b) {
    if (b->nMetalBlocks == 0) return;
    const size_t shm = sizeof(unsigned) * (size_t)b->sp[0].sizeHistRdf;
    if (b->mixed)
        hipLaunchKernelGGL(kb_rdf<true>, dim3(b->nMetalBlocks), dim3(256), shm, b->stream, b->args());
    else
        hipLaunchKernelGGL(kb_rdf<false>, dim3(b->nMetalBlocks), dim3(256), shm, b->stream, b->args());
}

void launch_commit_rebuild(sgk_ctx *c) {
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(1), 0, c->stream, c->state);
}

void launch_rdf(sgk_ctx *c) {
    if (c->dp.nMolMe == 0) return;
    profile_begin(c);
    const unsigned g = blocks_for(c->dp.nMolMe, 256);
    const size_t shm = sizeof(unsigned) * (size_t)c->dp.sizeHistRdf;
    if (c->rdfPos) {  // complete-range variant
        const DevParams &p = c->dp;
        const RdfGrid &G = c->rdfGrid;
        const int nb = G.ncell + 1;  // the last bucket stays empty: its start is the total
        hipLaunchKernelGGL(k_fill_i32, dim3(std::min(blocks_for(nb, 256), 4096u)), dim3(256), 0, c->stream, c->rdfCnt,
                           (long long)nb, 0);
        if (c->mixed)
            hipLaunchKernelGGL(k_rdfc_count<true>, dim3(g), dim3(256), 0, c->stream, p, G, (const double4 *)c->pos,
                               c->rdfCnt, c->rdfKey, c->rdfSlot);
        else
            hipLaunchKernelGGL(k_rdfc_count<false>, dim3(g), dim3(256), 0, c->stream, p, G, (const float4 *)c->pos,
                               c->rdfCnt, c->rdfKey, c->rdfSlot);
        const unsigned nblk = blocks_for(nb, kOrdChunk);
        hipLaunchKernelGGL(k_ord_blocksum, dim3(nblk), dim3(256), 0, c->stream, nb, c->rdfCnt, c->rdfBlock);
        hipLaunchKernelGGL(k_ord_scan, dim3(nblk), dim3(256), 0, c->stream, nb, c->rdfCnt, c->rdfBlock);
        const unsigned gp = blocks_for((long long)p.nMolMe * kRdfLanes, 256);
        if (c->mixed) {
            hipLaunchKernelGGL(k_rdfc_scatter<true>, dim3(g), dim3(256), 0, c->stream, p, (const double4 *)c->pos,
                               c->rdfCnt, c->rdfKey, c->rdfSlot, c->rdfPos);
            hipLaunchKernelGGL(k_rdfc_pairs<true>, dim3(gp), dim3(256), shm, c->stream, p, G, (const double4 *)c->pos,
                               c->rdfCnt, c->rdfKey, c->rdfPos, c->hist, &c->state->errors);
        } else {
            hipLaunchKernelGGL(k_rdfc_scatter<false>, dim3(g), dim3(256), 0, c->stream, p, (const float4 *)c->pos,
                               c->rdfCnt, c->rdfKey, c->rdfSlot, c->rdfPos);
            hipLaunchKernelGGL(k_rdfc_pairs<false>, dim3(gp), dim3(256), shm, c->stream, p, G, (const float4 *)c->pos,
                               c->rdfCnt, c->rdfKey, c->rdfPos, c->hist, &c->state->errors);
        }
        profile_end(c, "rdf");
        return;
    }
    if (c->mixed)
        hipLaunchKernelGGL(k_rdf<true>, dim3(g), dim3(256), shm, c->stream, c->dp, (const double4 *)c->pos, cell_bufs(c),
                           c->state, c->hist);
    else
        hipLaunchKernelGGL(k_rdf<false>, dim3(g), dim3(256), shm, c->stream, c->dp, (const float4 *)c->pos, cell_bufs(c),
                           c->state, c->hist);
    profile_end(c, "rdf");
}

}  // namespace sgk
