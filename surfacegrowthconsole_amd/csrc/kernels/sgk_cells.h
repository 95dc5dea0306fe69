// sgk_cells.h - cell buffers and the per-atom binning step shared by the binning kernels.
#pragma once

#include "sgk.h"
#include "sgk_device.h"
#include "sgk_internal.h"

namespace sgk {

// Two cell buffers; state.cellbuf names the current one (see the rebuild protocol in sgk_nebr.hip).
struct CellBufs {
    int *cnt[2];
    int *slots[2];
    int *nonempty[2];
};

inline CellBufs cell_bufs(sgk_ctx *c) {
    CellBufs b;
    for (int k = 0; k < 2; ++k) {
        b.cnt[k] = c->cnt[k];
        b.slots[k] = c->slots[k];
        b.nonempty[k] = c->nonempty[k];
    }
    return b;
}

// Cell of atom n (position q, float) with the reference's FP32 formula
// c = VLinear(int((r + L/2) * invWidth)) (SurfaceGrowth.cu:594-600), truncation toward zero, and
// its slot in the new buffer (cellbuf ^ 1) by atomicAdd; the first atom of a cell appends the cell
// to the buffer's non-empty list.  Slot order is arbitrary; readers sort.
__device__ __forceinline__ void bin_atom(const DevParams &p, int n, float3 q, int *__restrict__ cellOf,
                                         const CellBufs &bufs, DevState *st) {
#pragma clang fp contract(off)
    const float sx = q.x + 0.5f * p.Lx;
    const float sy = q.y + 0.5f * p.Ly;
    const float sz = q.z + 0.5f * p.Lz;
    const int cx = (int)(sx * p.iwx), cy = (int)(sy * p.iwy), cz = (int)(sz * p.iwz);
    const int c = (cz * p.cy + cy) * p.cx + cx;
    if (c < 0 || c >= p.ncell) {  // no reference block has this index: the atom is not binned
        cellOf[n] = -1;
        atomicOr(&st->errors, SGK_FLAG_UNBINNED);
        return;
    }
    cellOf[n] = c;
    const int b = st->cellbuf ^ 1;
    const int slot = atomicAdd(&bufs.cnt[b][c], 1);
    // Append newly occupied cells with one atomic per wave (a single counter hit once per cell
    // serialises at the L2).
    const unsigned long long first = __ballot(slot == 0);
    if (first) {
        const int lane = __lane_id();
        const int leader = __ffsll((long long)first) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&st->nonempty[b], __popcll(first));
        base = __shfl(base, leader, 64);
        if (slot == 0) bufs.nonempty[b][base + __popcll(first & ((1ull << lane) - 1ull))] = c;
    }
    if (slot < p.cellCap)
        bufs.slots[b][(size_t)c * p.cellCap + slot] = n;
    else
        atomicOr(&st->errors, SGK_FLAG_CELL_OVERFLOW);
}

}  // namespace sgk
