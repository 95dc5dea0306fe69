// sgk_batch.h - tables of the batched engine: many replicas of uneven size advance together, one
// launch per stage.  Each replica's atoms, cells and lists occupy their own ranges of shared
// arrays; a block table maps every workgroup of a batched launch to (replica, block within the
// replica), so the per-atom code of the single-system kernels runs unchanged on the replica's
// ranges, and a replica's results are those of the replica run alone.
#pragma once

#include <hip/hip_runtime.h>

#include "sgk.h"
#include "sgk_params.h"

namespace sgk {

struct BatchRep {
    DevParams p;
    long long atomOff;   // first atom (pos, vel, acc, cellOf, slotOf; vel/acc blocks of 3*nMol at 3*atomOff)
    long long metalOff;  // first metal atom (rho, embF, embDF, fric, nnM; nlM rows)
    long long carbonOff; // first carbon atom (lj, nlj, cl, ncl, bondC, bondA)
    long long cellOff;   // first cell (cnt, slots, cellPos)
    int partOff;         // first force-pass partial
    int nbForce;         // force-pass workgroups: nbM metal then carbon
    int nbM;             // metal workgroups of the density and force passes
    int part2Off;        // first chunk total of the two-stage step reduction
    int nChunks;         // chunks of kPartChunk force-pass partials
    int nbNlM;           // workgroups of the metal part of the cell-ordered list build (4 lanes per atom)
    int ordOff;          // first bucket of the replica's list-build order (2 ncell + 2 buckets)
    int pad;
};

struct BatchArgs {
    const BatchRep *reps;
    const int2 *atomTab;   // (replica, block) per workgroup over each replica's atoms
    const int2 *forceTab;  // (replica, block) over each replica's nbM + nbC force/density blocks
    const int2 *metalTab;  // (replica, block) over each replica's metal atoms
    const int2 *nlTab;     // (replica, block) over each replica's cell-ordered list-build blocks
    const int *order;      // list-build order: replica r's atoms (local indices) at [atomOff, atomOff + nMol)
    void *pos, *vel, *acc;
    void *posOut;          // the force pass's fused drift writes here (null: no drift)
    float *rho, *embF, *embDF, *fric;
    int *cellOf[2], *cnt[2], *slots[2], *slotOf;
    float4 *cellPos;
    int *nlM, *nnM, *nlj, *ncl;
    unsigned *lj, *cl;
    uint4 *bondC, *bondA;
    void *partials;
    DevState *state;
    sgk_record *records;
    int recCap;
    unsigned *hist;
};

// The cell buffers of one replica.
__device__ __forceinline__ CellBufs rep_cells(const BatchArgs &a, const BatchRep &R) {
    CellBufs b;
    const long long cap = R.p.cellCap;
    for (int k = 0; k < 2; ++k) {
        b.cnt[k] = a.cnt[k] + R.cellOff;
        b.slots[k] = a.slots[k] + R.cellOff * cap;
        b.cellOf[k] = a.cellOf[k] + R.atomOff;
    }
    b.slotOf = a.slotOf + R.atomOff;
    b.cellPos = a.cellPos + R.cellOff * cap;
    return b;
}

}  // namespace sgk
