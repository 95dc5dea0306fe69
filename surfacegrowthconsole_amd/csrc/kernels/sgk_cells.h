// sgk_cells.h - cell buffers and the per-atom binning step shared by the binning kernels.
#pragma once

#include "sgk.h"
#include "sgk_device.h"
#include "sgk_internal.h"

namespace sgk {

// The cell buffers of a single-system context.
inline CellBufs cell_bufs(sgk_ctx *c) {
    CellBufs b;
    for (int k = 0; k < 2; ++k) {
        b.cnt[k] = c->cnt[k];
        b.slots[k] = c->slots[k];
        b.cellOf[k] = c->cellOf[k];
    }
    b.slotOf = c->slotOf;
    b.cellPos = c->cellPos;
    return b;
}

// Cell of atom n (position q, float) with the reference's FP32 formula
// c = VLinear(int((r + L/2) * invWidth)) (SurfaceGrowth.cu:594-600), truncation toward zero, and
// its slot in the new buffer (cellbuf ^ 1) by atomicAdd.  Slot order is arbitrary; k_sortcells
// sorts.  Called by every active lane of a wave together (wave-collective).
__device__ __forceinline__ void bin_atom(const DevParams &p, int n, float3 q, const CellBufs &bufs, DevState *st) {
#pragma clang fp contract(off)
    const float sx = q.x + 0.5f * p.Lx;
    const float sy = q.y + 0.5f * p.Ly;
    const float sz = q.z + 0.5f * p.Lz;
    const int cx = (int)(sx * p.iwx), cy = (int)(sy * p.iwy), cz = (int)(sz * p.iwz);
    const int c = (cz * p.cy + cy) * p.cx + cx;
    const int b = st->cellbuf ^ 1;
    if (c < 0 || c >= p.ncell) {  // no reference block has this index: the atom is not binned
        bufs.cellOfB(b)[n] = -1;
        atomicOr(&st->errors, SGK_FLAG_UNBINNED);
        return;
    }
    bufs.cellOfB(b)[n] = c;
    // Lanes of a wave that fall in the same cell (consecutive atoms are spatial neighbours) take
    // their slots with one atomic: group the active lanes by cell, one pass per distinct cell,
    // then one atomicAdd per group leader.  Device-scope atomics on one address serialise, so
    // their number, not bandwidth, bounds binning.
    const int lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = p.binGroup ? __ballot(1) : 0ull, grp = 1ull << lane;
    int leader = lane;
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        const int cl = __shfl(c, l, 64);
        const unsigned long long m = __ballot(c == cl);
        if (c == cl) {
            grp = m;
            leader = l;
        }
        todo &= ~m;
    }
    int base = 0;
    if (lane == leader) base = atomicAdd(&bufs.cntB(b)[c], __popcll(grp));
    base = __shfl(base, leader, 64);
    const int slot = base + __popcll(grp & below);
    bufs.slotOf[n] = slot;
    if (slot < p.cellCap)
        bufs.slotsB(b)[(size_t)c * p.cellCap + slot] = n;
    else
        atomicOr(&st->errors, SGK_FLAG_CELL_OVERFLOW);
}

}  // namespace sgk
