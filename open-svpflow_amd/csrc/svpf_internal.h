// Internal declarations shared by the libsvpf translation units (host side and
// kernel launchers).  Not part of the public C-ABI (include/svpf.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/svpf.h"

struct svpf_ctx {
    int device;
    hipStream_t stream;
    hipEvent_t t0, t1;
    // Descriptor arrays for launches: pinned staging buffer + device copy, grown on
    // demand.  `desc_done` marks when the last upload finished reading the staging buffer.
    void* h_desc = nullptr;
    void* d_desc = nullptr;
    size_t desc_cap = 0;
    hipEvent_t desc_done = nullptr;
    // Analyse workspace (device), grown on demand and reused across calls.
    void* d_work = nullptr;
    size_t work_cap = 0;
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_cap = 0;
    uint32_t* h_count = nullptr;  // pinned words for reading back sweep queue lengths
    size_t h_count_cap = 0;
    // Analyse: parallel sweeps per level before the sequential fallback (0 = the default).
    int max_sweeps = 0;
    // Analyse: how many visit-order successors a later sweep evaluates at once after a
    // block changes (0: none, every change waits for the next sweep).
    int chase = 0;
    double slot_scale = 1.0;   // factor on the finest level's sweep-0 slot count
    int fine_run = 32;         // finest-level positions per sweep-0 item
    int coarse_slots = 256;    // sweep-0 slots of the coarser levels
    int coarse_run = 16;       // coarser-level positions per sweep-0 item
    int sweep_grid = 2048;     // workgroups of a later sweep
    // Analyse: later sweeps replayed from captured hipGraphs (one per level launch
    // configuration, least recently used evicted beyond kSweepGraphs).
    bool sweep_graphs = true;
    int fast_list_min = 16384;  // later sweeps with a list this long run the tile-plan kernel (-1: never)
    struct SweepGraph {
        std::string key;  // the launch arguments the graph was captured with
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    std::vector<SweepGraph> graphs;
    unsigned long long graph_clock = 0;
    // Analyse sweep statistics per level since profiling was last enabled (device, 16 x 4,
    // added by a kernel so that profiling adds no host round trip):
    // {level runs, sweeps launched, evaluations skipped, evaluations run}
    unsigned long long* d_sweep_stats = nullptr;
    // Optional kernel timing: HIP event pairs recorded around launches on `stream`.
    struct ProfMark {
        int kernel;
        hipEvent_t a, b;
    };
    bool prof = false;
    std::vector<ProfMark> prof_marks;
    std::vector<hipEvent_t> prof_pool;
};

namespace svpf {

// One YUV420P8 frame's three planes (device or host pointers, byte strides).
struct Planes {
    uint8_t* p[3];
    int stride[3];
};

// Super-frame geometry for one clip (super_opts.rs:70-95 formulas).
struct SuperGeom {
    int width, height;     // luma size of the source
    int pel, levels;
    int gpu, scale_down, scale_up, full;
    int rows[3];           // plane heights of the super frame (luma super_h, chroma super_h/2)
    int level_row[3][16];  // first row of each level inside each plane
    int level_w[3][16];    // level width per plane
    int level_h[3][16];    // level height per plane
};

int reduce_dim(int dim, int n);
int auto_levels(int width, int height);
int super_plane_height(int base_h, int pel, int levels, bool full);
SuperGeom make_super_geom(const svpf_super_params& p);

// Error reporting for the C-ABI: thread-local last error string.
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// Kernel timing (svpf_profile_*): kernels that are timed one launch at a time.
enum ProfKernel { kProfSuper = 0, kProfSearch, kProfAnalyseAux, kProfPrep, kProfRender, kProfCount };
// Records the start event of one launch of `kernel` when profiling is on; returns a
// handle for prof_end (or -1).
int prof_begin(svpf_ctx* ctx, int kernel);
void prof_end(svpf_ctx* ctx, int handle);

#define SVPF_HIP(call)                                                                 \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::svpf::fail(SVPF_ERR_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace svpf

// Stream-ordered upload of a small descriptor table (capi_core.cpp).
int svpf_upload_desc(svpf_ctx* ctx, const void* data, size_t bytes, void** d_out);

namespace svpf {

// Analyse over caller-provided device memory instead of the context's grow-only
// buffers: the descriptor blob, the workspace and the sort storage are carved from
// [base, base + cap); nothing is allocated.
struct ExtWorkspace {
    uint8_t* base;
    size_t cap;
};
int analyse_run(svpf_ctx* ctx, int n, const svpf_analyse_params* const* params, const svpf_frame* d_cur,
                const svpf_frame* d_ref, uint8_t* const* d_payload, const ExtWorkspace* ext = nullptr);
// Resolves the allocation-free entry point's scratch array on `device` (which must be
// current); once per device (analyse_noalloc.hip).
int noalloc_scratch_init(int device);

// Kernel launchers (super.hip)
hipError_t launch_super_level0(const SuperGeom& g, const Planes* src, const Planes* dst, int nframes,
                               int max_stride, hipStream_t s);
hipError_t launch_super_reduce(const SuperGeom& g, const Planes* dst, int nframes, int level, hipStream_t s);

}  // namespace svpf
