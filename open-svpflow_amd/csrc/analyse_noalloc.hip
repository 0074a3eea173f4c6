// svpf_analyse_pair_noalloc: svpf_analyse_pair for callers that cannot allocate device
// memory (serving loops and harnesses that forbid hipMalloc).  Its workspace, sort
// storage and descriptor blob are carved from one file-scope device array instead of
// the context's grow-only buffers.
//
// The array's name is a host-side stand-in in host code: its device address on each GPU
// comes from hipGetSymbolAddress, resolved once per device by svpf_create (with that
// device current) and kept in `g_scratch`.  Only that address is passed on.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind; the
// kernels initialise everything a call reads before reading it (as they do for the
// context's reused workspace).  Calls that overlap on one GPU -- from two threads or two
// contexts -- race on the array and must be serialised by the caller.
#include <mutex>
#include <string>

#include "svpf_internal.h"

namespace {

// 128 MiB: one pair of 2560x1440 frames with 8x8 blocks and overlap 4 (the densest
// configuration of the serving workloads) needs about 112 MiB.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices] = {};

}  // namespace

namespace svpf {

int noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (device < 0 || device >= kMaxDevices) return fail(SVPF_ERR_ARG, "svpf: device index out of range");
    if (g_scratch[device]) return SVPF_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SVPF_ERR_HIP, std::string("svpf: scratch address lookup failed: ") + hipGetErrorString(e));
    g_scratch[device] = (uint8_t*)p;
    return SVPF_OK;
}

}  // namespace svpf

extern "C" int svpf_analyse_pair_noalloc(svpf_ctx* ctx, const svpf_analyse_params* p, const svpf_frame* d_cur,
                                         const svpf_frame* d_ref, uint8_t* d_payload) {
    using namespace svpf;
    if (!ctx || !p || !d_cur || !d_ref || !d_payload) return fail(SVPF_ERR_ARG, "svpf_analyse_pair_noalloc: null argument");
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(SVPF_ERR_ARG, "svpf_analyse_pair_noalloc: bad device");
    uint8_t* base;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        base = g_scratch[ctx->device];
    }
    if (!base) return fail(SVPF_ERR_STATE, "svpf_analyse_pair_noalloc: scratch memory not resolved on this device");
    SVPF_HIP(hipSetDevice(ctx->device));
    const ExtWorkspace ws{base, kScratchBytes};
    const svpf_analyse_params* ps[1] = {p};
    return analyse_run(ctx, 1, ps, d_cur, d_ref, &d_payload, &ws);
}
