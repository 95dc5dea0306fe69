// sgk_noalloc.hip - the force evaluation of sgk_create + sgk_bin + sgk_build_nebr +
// sgk_compute_forces in one call that allocates nothing (sgk_forces_noalloc).
//
// Positions, velocities, accelerations and friction are the caller's device buffers.  Everything
// else the kernels need - both cell buffers, the cell-ordered positions, the metal and carbon
// lists, the near lists, rho/F/F', the block partials and the run state - is carved from one
// file-scope device array, `scratch`, of SGK_NOALLOC_SCRATCH_BYTES.  Each GPU that loads the library
// holds its own copy of the array until the process ends; its device address is looked up once per
// device with hipGetSymbolAddress and kept per device, and only that address is handed to kernels.
// The call enqueues kernels on the caller's stream and returns: no allocation, no memset (the
// buffers that must start cleared are cleared by a fill kernel), no copy and no synchronisation.
// Calls that overlap on one GPU share the array and race; the library does not serialise them.
#include <mutex>
#include <string>

#include "sgk_internal.h"

// The scratch array.  In host code its name stands for a host-side shadow, not a device address:
// kernels get the address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[SGK_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device (null: not looked up)
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// Byte offsets of the carved buffers: each starts on a 256-byte boundary.
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        const size_t at = off;
        off += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return at;
    }
};

struct Layout {
    size_t cellOf[2], cnt[2], slots[2], slotOf, cellPos, nlM, nnM, lj, nlj, cl, ncl, bondC, bondA, rho, embF, embDF,
        fric, zeroVel, partials, state, total;
    size_t zeroVelBytes;
};

Layout plan(const sgk_ctx &c, bool ownFric, bool ownVel) {
    const sgk::DevParams &p = c.dp;
    const size_t nMol = p.nMol, nMe = std::max(p.nMolMe, 1), nC = std::max(p.nC, 1), ncell = p.ncell;
    const size_t R = c.mixed ? 8 : 4;
    Layout L{};
    Carve k;
    for (int b = 0; b < 2; ++b) {
        L.cellOf[b] = k.take(4 * nMol);
        L.cnt[b] = k.take(4 * ncell);
        L.slots[b] = k.take(4 * ncell * p.cellCap);
    }
    L.slotOf = k.take(4 * nMol);
    L.cellPos = k.take(16 * ncell * p.cellCap);
    L.nlM = k.take(4 * nMe * p.nlCap);
    L.nnM = k.take(4 * nMe);
    L.lj = k.take(4 * nC * p.ljCap);
    L.nlj = k.take(4 * nC);
    L.cl = k.take(4 * nC * p.clCap);
    L.ncl = k.take(4 * nC);
    L.bondC = k.take(16 * nC);
    L.bondA = k.take(16 * nC);
    L.rho = k.take(4 * nMe);
    L.embF = k.take(4 * nMe);
    L.embDF = k.take(4 * nMe);
    L.fric = ownFric ? k.take(4 * nMe) : 0;
    L.zeroVelBytes = ownVel ? 3 * nMol * R : 0;
    L.zeroVel = ownVel ? k.take(L.zeroVelBytes) : 0;
    const size_t nParts = (size_t)c.part2Off;
    L.partials = k.take((nParts + nParts / 64 + 2) * sgk::partial_bytes_for(c.mixed));
    L.state = k.take(sizeof(sgk::DevState));
    L.total = k.off;
    return L;
}

// Clears what a fresh context starts with: both cell indices at -1, both cell counts at 0, the
// stand-in velocities at 0, and the run state (rebuild requested, step and deposition counters).
struct InitArgs {
    int *cellOf0, *cellOf1, *cnt0, *cnt1;
    unsigned *zero;
    long long nMol, ncell, nZero;
    sgk::DevState *state;
    int step, nMolDeposited;
};

__global__ __launch_bounds__(256) void k_noalloc_init(InitArgs a) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long long t = t0; t < a.nMol; t += stride) {
        a.cellOf0[t] = -1;
        a.cellOf1[t] = -1;
    }
    for (long long t = t0; t < a.ncell; t += stride) {
        a.cnt0[t] = 0;
        a.cnt1[t] = 0;
    }
    for (long long t = t0; t < a.nZero; t += stride) a.zero[t] = 0u;
    if (t0 == 0) {
        sgk::DevState s;
        s.step = a.step;
        s.nebrNow = 1;
        s.errors = 0;
        s.cellbuf = 0;
        s.rec = 0;
        s.rebuilds = 0;
        s.dispHi = 0.0;
        s.shear = 0.0;
        s.vvSum = 0.0;
        s.comX = 0.0;
        s.cmVelX = 0.0;
        s.shearCount = 0u;
        s.nMolDeposited = a.nMolDeposited;
        *a.state = s;
    }
}

__global__ void k_noalloc_flags(const sgk::DevState *state, int *flags) { *flags = state->errors; }

}  // namespace

extern "C" {

const char *sgk_noalloc_last_error(void) { return g_err.c_str(); }

long long sgk_noalloc_scratch_bytes(const SimParams *sp, const sgk_config *cfg) {
    if (!sp) return -1;
    sgk_ctx c;
    std::string msg;
    if (sgk::prepare_ctx(sp, cfg, &c, msg) != SGK_OK) return -1;
    return (long long)plan(c, true, true).total;
}

int sgk_forces_noalloc(const SimParams *sp, const sgk_config *cfg, void *pos, const void *vel, void *acc,
                       float *fric, int *flags, void *stream) {
    if (!sp || !pos || !acc) return fail(SGK_E_ARG, "null argument (parameters, positions and accelerations are required)");
    sgk_ctx c;
    {
        std::string msg;
        const int rc = sgk::prepare_ctx(sp, cfg, &c, msg);
        if (rc != SGK_OK) return fail(rc, msg);
    }
    if (c.dp.keepFull) return fail(SGK_E_UNSUPPORTED, "keep_full_nl needs the reference-layout list: not available without allocation");
    const Layout L = plan(c, fric == nullptr, vel == nullptr);
    if (L.total > (size_t)SGK_NOALLOC_SCRATCH_BYTES)
        return fail(SGK_E_UNSUPPORTED, "system needs " + std::to_string(L.total) + " bytes of scratch; the limit without allocation is " +
                                           std::to_string((long long)SGK_NOALLOC_SCRATCH_BYTES) + " bytes (use sgk_create)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return fail(SGK_E_HIP, "no usable current HIP device");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_scratch[dev]) {
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr)
                return fail(SGK_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            g_scratch[dev] = ptr;
        }
        base = g_scratch[dev];
    }
    char *b = (char *)base;
    c.device = dev;
    c.stream = (hipStream_t)stream;
    c.pos = pos;
    c.vel = vel ? const_cast<void *>(vel) : (void *)(b + L.zeroVel);
    c.acc = acc;
    c.fric = fric ? fric : (float *)(b + L.fric);
    for (int k = 0; k < 2; ++k) {
        c.cellOf[k] = (int *)(b + L.cellOf[k]);
        c.cnt[k] = (int *)(b + L.cnt[k]);
        c.slots[k] = (int *)(b + L.slots[k]);
    }
    c.slotOf = (int *)(b + L.slotOf);
    c.cellPos = (float4 *)(b + L.cellPos);
    c.nlM = (int *)(b + L.nlM);
    c.nnM = (int *)(b + L.nnM);
    c.lj = (unsigned *)(b + L.lj);
    c.nlj = (int *)(b + L.nlj);
    c.cl = (unsigned *)(b + L.cl);
    c.ncl = (int *)(b + L.ncl);
    c.bondC = (uint4 *)(b + L.bondC);
    c.bondA = (uint4 *)(b + L.bondA);
    c.rho = (float *)(b + L.rho);
    c.embF = (float *)(b + L.embF);
    c.embDF = (float *)(b + L.embDF);
    c.partials = b + L.partials;
    c.state = (sgk::DevState *)(b + L.state);
    c.orderValid = false;  // lists in cell order of index (same lists, no order buffers)

    InitArgs ia;
    ia.cellOf0 = c.cellOf[0];
    ia.cellOf1 = c.cellOf[1];
    ia.cnt0 = c.cnt[0];
    ia.cnt1 = c.cnt[1];
    ia.zero = (unsigned *)(b + L.zeroVel);
    ia.nMol = c.dp.nMol;
    ia.ncell = c.dp.ncell;
    ia.nZero = (long long)(L.zeroVelBytes / 4);
    ia.state = c.state;
    ia.step = sp->stepCount;
    ia.nMolDeposited = sp->nMolDeposited;
    const long long work = std::max<long long>(std::max(ia.nMol, ia.ncell), ia.nZero);
    const unsigned g = (unsigned)std::min<long long>((work + 255) / 256, 4096);
    hipLaunchKernelGGL(k_noalloc_init, dim3(std::max(g, 1u)), dim3(256), 0, c.stream, ia);
    // the stage sequence of sgk_bin, sgk_build_nebr and sgk_compute_forces
    sgk::launch_bin(&c, true);
    sgk::launch_sortcells(&c, true);
    sgk::launch_commit_rebuild(&c);
    sgk::launch_nl(&c, true);
    sgk::launch_rho(&c);
    sgk::launch_force(&c, false, false);
    if (flags) hipLaunchKernelGGL(k_noalloc_flags, dim3(1), dim3(1), 0, c.stream, c.state, flags);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SGK_E_HIP, std::string("launch: ") + hipGetErrorString(e));
    return SGK_OK;
}

}  // extern "C"
