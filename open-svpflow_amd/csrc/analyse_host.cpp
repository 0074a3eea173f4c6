// Host side of the Analyse search: option parsing (analyse_opts.rs) and the batched
// level pipeline that drives the kernels of analyse.hip.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "analyse_dev.h"
#include "options.h"
#include "svpf_internal.h"

namespace svpf {

hipError_t launch_search_level(const Chain* d_chains, int nchains, hipStream_t s);
hipError_t launch_sweep_init(const Chain* d_chains, int nchains, int max_n, const SweepState& st, hipStream_t s);
hipError_t launch_sweep0(const Chain* d_chains, int nchains, int max_n, const SweepState& st, hipStream_t s,
                         int level, bool fast, bool l0, bool lean, double slot_scale, int coarse_slots, int fine_run,
                         int coarse_run);
hipError_t launch_sweep(const Chain* d_chains, int grid, const SweepState& st, int sweep, hipStream_t s, bool l0);
hipError_t launch_sweep_fast(const Chain* d_chains, int grid, const SweepState& st, int sweep, hipStream_t s, bool l0);
hipError_t launch_sweep_stats(const SweepState& st, int sweeps, long long total_blocks, unsigned long long* out4,
                              hipStream_t s);
hipError_t launch_deferred(const Chain* d_chains, int grid, const SweepState& st, int sweep, hipStream_t s, bool l0);
hipError_t launch_sweep_scan(const Chain* d_chains, int nchains, const SweepState& st, int sweep, hipStream_t s);
hipError_t launch_validate(const Chain* d_chains, int nchains, int max_n, const SweepState& st, hipStream_t s);
hipError_t launch_interp(const void* d_tr, int ntr, int max_n, hipStream_t s);
hipError_t launch_global(const void* d_jobs, int njobs, hipStream_t s);
hipError_t launch_sort_keys(const void* d_jobs, int njobs, int max_n, hipStream_t s);
hipError_t segmented_sort(void* d_temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, int n, int nseg, const int* d_begin,
                          const int* d_end, hipStream_t s);
hipError_t launch_child_order(const void* d_jobs, int njobs, hipStream_t s);
hipError_t launch_pack(const void* d_jobs, int njobs, int max_count, int denom, hipStream_t s);
hipError_t launch_refine(const void* d_jobs, int njobs, int max_n, hipStream_t s);

namespace {

int32_t sat_mul(int32_t a, int32_t b) {
    int64_t r = (int64_t)a * b;
    return (int32_t)std::min<int64_t>(std::max<int64_t>(r, INT32_MIN), INT32_MAX);
}
int32_t f64_to_i32(double v) {  // Rust `as i32`
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}
// f64::powi as compiled (repeated squaring, compiler-rt __powidf2).
double powi_rt(double a, int b) {
    bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

int auto_search_levels(int w, int h, int bw, int bh, int ox, int oy) {  // analyse_opts.rs:271-295
    int sx = std::max(bw - ox, 1), sy = std::max(bh - oy, 1);
    int gw = std::max((w - ox) / sx, 0), gh = std::max((h - oy) / sy, 0);
    if (gw <= 0 || gh <= 0) return 1;
    int ex = ox + sx * gw, ey = oy + sy * gh, lv = 0;
    while (lv <= 12) {
        if (((ex >> lv) - ox) / sx <= 0 || ((ey >> lv) - oy) / sy <= 0) break;
        lv++;
    }
    return std::max(std::min(lv, 13), 1);
}

void output_block(const svpf_analyse_params& p, int* bw, int* bh, int* ox, int* oy) {  // :221-232
    if (p.has_refine) {
        *bw = std::max(p.block_w / 2, 4); *bh = std::max(p.block_h / 2, 4);
        *ox = std::max(p.overlap_x / 2, 0); *oy = std::max(p.overlap_y / 2, 0);
    } else {
        *bw = p.block_w; *bh = p.block_h; *ox = p.overlap_x; *oy = p.overlap_y;
    }
}

void out_grid(const svpf_analyse_params& p, int bw, int bh, int ox, int oy, int* gw, int* gh) {  // :234-240
    int sx = std::max(bw - ox, 1), sy = std::max(bh - oy, 1);
    *gw = std::max((p.width - ox) / sx, 1);
    *gh = std::max((p.height - oy) / sy, 1);
}

void level_grid(const svpf_analyse_params& p, int level, int* gw, int* gh) {  // analyse_search.rs:1453-1463
    int sx = std::max(p.block_w - p.overlap_x, 1), sy = std::max(p.block_h - p.overlap_y, 1);
    int gw0 = std::max((p.width - p.overlap_x) / sx, 0), gh0 = std::max((p.height - p.overlap_y) / sy, 0);
    int ex = p.overlap_x + sx * gw0, ey = p.overlap_y + sy * gh0;
    *gw = std::max(((ex >> level) - p.overlap_x) / sx, 1) + (level > 0);
    *gh = std::max(((ey >> level) - p.overlap_y) / sy, 1) + (level > 0);
}

void header_words(const svpf_analyse_params& p, int32_t h[16]) {  // pack_vector_frame header, :2757-2785
    int bw, bh, ox, oy, gw, gh;
    output_block(p, &bw, &bh, &ox, &oy);
    out_grid(p, bw, bh, ox, oy, &gw, &gh);
    int flags = ((p.vectors & 1) ? 1 : 0) | ((p.vectors & 2) ? 2 : 0);
    int32_t v[16] = {16, 0xA0, flags, bw, bh, p.pel, p.has_refine ? 1 : p.levels, 0, p.width, p.height, ox, oy,
                     gw, gh, p.selector, p.delta};
    std::memcpy(h, v, sizeof v);
}

// Static visit orders (analyse_search.rs:302-336): meander of the interior rows, then
// the first/last rows, then the first/last columns; coarse levels add the edge strip.
void initial_order(int gw, int gh, std::vector<int32_t>& o) {
    for (int y = 1; y < gh - 1; y++) {
        if ((y & 1) == 0) for (int x = 1; x < gw - 1; x++) o.push_back(x | (y << 16));
        else for (int x = gw - 2; x >= 1; x--) o.push_back(x | (y << 16));
    }
    for (int x = 0; x < gw; x++) {
        o.push_back(x);
        if (gh > 1) o.push_back(x | ((gh - 1) << 16));
    }
    for (int y = 1; y < gh - 1; y++) {
        o.push_back(y << 16);
        if (gw > 1) o.push_back((gw - 1) | (y << 16));
    }
}

std::vector<int32_t> level_order(int gw, int gh, int level) {
    std::vector<int32_t> o;
    int e = level > 0 ? 1 : 0;
    initial_order(gw - e, gh - e, o);
    if (e) {
        for (int x = 0; x <= gw - 1; x++) o.push_back(x | ((gh - 1) << 16));
        for (int y = 0; y < gh - 1; y++) o.push_back((gw - 1) | (y << 16));
    }
    return o;
}

// Byte-aligned bump allocator over one blob (host staging or device workspace).
struct Carve {
    size_t off = 0;
    size_t take(size_t n, size_t align = 256) {
        off = (off + align - 1) / align * align;
        size_t r = off;
        off += n;
        return r;
    }
};

struct ClipPlan {
    const svpf_analyse_params* p;
    int levels;
    int gw[16], gh[16];
    LevelGeom geom[16];
    size_t geom_off;       // in the descriptor blob
    size_t order_top_off;  // static order of the top level (blob)
    size_t order0_off;     // static meander of level 0 (blob)
    size_t header_off;     // 16 header words (blob)
};

constexpr int kMaxSweeps = 4096;
constexpr int kNotConverged = 1;


constexpr int kFastListMin = 16384;  // default of svpf_analyse_tuning::fast_list_min
constexpr int kGraphSweeps = 6;  // sweeps per captured graph: lists by s mod 2, queue lengths by s mod 3
constexpr size_t kSweepGraphs = 16;

// The executable graph of kGraphSweeps later sweeps (each followed by its count rebuild
// when `scan`) for these launch arguments, captured on first use.  Sweep s reads its
// list by s mod 2 and its queue length by s mod 3, so the graph captured for sweeps
// 1 .. 6 replays sweeps 6k + 1 .. 6k + 6 exactly.
int sweep_graph(svpf_ctx* ctx, const Chain* dch, int nch, int grid, const SweepState& st, bool scan, bool l0,
                hipGraphExec_t* out) {
    struct Key {
        const Chain* dch;
        SweepState st;
        int nch, grid, scan, l0;
    } k;
    std::memset(&k, 0, sizeof k);  // padding bytes take part in the comparison
    k.dch = dch;
    k.st = st;
    k.nch = nch;
    k.grid = grid;
    k.scan = scan;
    k.l0 = l0;
    const std::string key(reinterpret_cast<const char*>(&k), sizeof k);
    for (auto& g : ctx->graphs)
        if (g.key == key) {
            g.used = ++ctx->graph_clock;
            *out = g.exec;
            return SVPF_OK;
        }
    if (ctx->graphs.size() >= kSweepGraphs) {  // evict the least recently used
        size_t v = 0;
        for (size_t i = 1; i < ctx->graphs.size(); i++)
            if (ctx->graphs[i].used < ctx->graphs[v].used) v = i;
        SVPF_HIP(hipStreamSynchronize(ctx->stream));  // it may still be running
        (void)hipGraphExecDestroy(ctx->graphs[v].exec);
        (void)hipGraphDestroy(ctx->graphs[v].graph);
        ctx->graphs.erase(ctx->graphs.begin() + (long)v);
    }
    svpf_ctx::SweepGraph g;
    g.key = key;
    SVPF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int j = 1; j <= kGraphSweeps && e == hipSuccess; j++) {
        e = launch_sweep(dch, grid, st, j, ctx->stream, l0);
        if (e == hipSuccess && scan) e = launch_sweep_scan(dch, nch, st, j, ctx->stream);
    }
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g.graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g.graph) (void)hipGraphDestroy(g.graph);
        return fail(SVPF_ERR_HIP, std::string("svpf_analyse: sweep graph capture: ") + hipGetErrorString(e));
    }
    g.used = ++ctx->graph_clock;
    ctx->graphs.push_back(g);
    *out = g.exec;
    return SVPF_OK;
}

// Parallel sweeps of one level until no block changes (each sweep evaluates the blocks
// whose earlier-visited neighbours changed in the previous sweep).  The fixed point is
// unique -- every block depends only on blocks visited before it -- and equals the
// sequential result.  The later sweeps go in chunks: replays of a captured graph of
// kGraphSweeps sweeps (or, with the chase, without graphs or in the allocation-free
// entry point, launches one by one); the host reads the next queue length after each
// chunk, and sweeps launched past convergence find an empty queue and return at once.
// fast: every chain's searched block is a 4x4 or 8x8 square, and sweep 0 (every block
// of the level) runs the tile-plan kernel, then the general one over the blocks it
// passed on.  l0: the finest level's instantiation; lean: every chain's pnew >= -256,
// so the tile-plan kernel runs its steps as straight-line code.
int run_sweeps(svpf_ctx* ctx, const Chain* dch, int nch, int max_n, size_t total_blocks, const SweepState& st,
               bool scan, int level, bool fast, bool l0, bool lean, bool no_alloc) {
    const size_t words = 4;
    uint32_t local_count = 0;  // no_alloc: read back through pageable memory (a synchronous copy)
    uint32_t* h_count = no_alloc ? &local_count : nullptr;
    if (!no_alloc && ctx->h_count_cap < words) {
        if (ctx->h_count) (void)hipHostFree(ctx->h_count);
        ctx->h_count = nullptr;
        ctx->h_count_cap = 0;
        SVPF_HIP(hipHostMalloc((void**)&ctx->h_count, 4 * words, hipHostMallocDefault));
        ctx->h_count_cap = words;
    }
    if (!no_alloc) h_count = ctx->h_count;
    const int grid = (int)std::min<size_t>(std::max<size_t>(total_blocks, 1), (size_t)ctx->sweep_grid);
    SVPF_HIP(launch_sweep0(dch, nch, max_n, st, ctx->stream, level, fast, l0, lean, ctx->slot_scale, ctx->coarse_slots,
                           ctx->fine_run, ctx->coarse_run));
    if (fast) SVPF_HIP(launch_deferred(dch, grid, st, 0, ctx->stream, l0));
    if (scan) SVPF_HIP(launch_sweep_scan(dch, nch, st, 0, ctx->stream));
    SVPF_HIP(launch_validate(dch, nch, max_n, st, ctx->stream));
    // the chase claims blocks by absolute sweep number, so its sweeps are not replayable
    hipGraphExec_t exec = nullptr;
    if (ctx->sweep_graphs && !no_alloc && st.chase == 0) {
        const int rc = sweep_graph(ctx, dch, nch, grid, st, scan, l0, &exec);
        if (rc) return rc;
    }
    int s = 1;
    auto converged = [&]() -> int {
        if (ctx->prof)  // statistics for the benchmark, added on the device
            SVPF_HIP(launch_sweep_stats(st, s, (long long)total_blocks, ctx->d_sweep_stats + 4 * std::min(level, 15),
                                        ctx->stream));
        return SVPF_OK;
    };
    // Levels whose sweep 0 ran the straight-line tile-plan kernel: while a list holds at
    // least fast_list_min blocks, the later sweep runs that kernel too (the general
    // instantiation takes about three times as long per block at the coarse levels, a
    // third longer at the finest), and deferred_kernel completes the blocks it passed on;
    // then single launches of the general kernel up to the next graph boundary.
    if (fast && (lean || l0) && st.chase == 0 && ctx->fast_list_min > 0) {
        for (;;) {
            SVPF_HIP(hipMemcpyAsync(h_count, st.counts + s % 3, 4, hipMemcpyDeviceToHost, ctx->stream));
            SVPF_HIP(hipStreamSynchronize(ctx->stream));
            if (*h_count == 0) return converged();
            if (*h_count < (uint32_t)ctx->fast_list_min) break;
            if (s + 1 > st.max_sweeps) return kNotConverged;
            SVPF_HIP(hipMemsetAsync(st.dcount + kDcDeferred, 0, 4, ctx->stream));
            SVPF_HIP(launch_sweep_fast(dch, grid, st, s, ctx->stream, l0));
            SVPF_HIP(launch_deferred(dch, grid, st, s, ctx->stream, l0));
            if (scan) SVPF_HIP(launch_sweep_scan(dch, nch, st, s, ctx->stream));
            s++;
        }
        while (exec && (s - 1) % kGraphSweeps != 0) {
            if (s + 1 > st.max_sweeps) return kNotConverged;
            SVPF_HIP(launch_sweep(dch, grid, st, s, ctx->stream, l0));
            if (scan) SVPF_HIP(launch_sweep_scan(dch, nch, st, s, ctx->stream));
            s++;
        }
    }
    int chunk = exec ? kGraphSweeps : 8;
    for (;;) {
        if (s + chunk > st.max_sweeps) return kNotConverged;
        if (exec) {
            for (int k = 0; k < chunk; k += kGraphSweeps) SVPF_HIP(hipGraphLaunch(exec, ctx->stream));
        } else {
            for (int k = 0; k < chunk; k++) {
                SVPF_HIP(launch_sweep(dch, grid, st, s + k, ctx->stream, l0));
                if (scan) SVPF_HIP(launch_sweep_scan(dch, nch, st, s + k, ctx->stream));
            }
        }
        s += chunk;
        SVPF_HIP(hipMemcpyAsync(h_count, st.counts + s % 3, 4, hipMemcpyDeviceToHost, ctx->stream));
        SVPF_HIP(hipStreamSynchronize(ctx->stream));
        if (*h_count == 0) return converged();
        chunk = std::min(chunk * 2, exec ? 10 * kGraphSweeps : 64);
    }
}

}  // namespace

int analyse_run(svpf_ctx* ctx, int n, const svpf_analyse_params* const* params, const svpf_frame* d_cur,
                const svpf_frame* d_ref, uint8_t* const* d_payload, const ExtWorkspace* ext) {
    if (n <= 0) return SVPF_OK;
    // ---- per-clip plans (clips are identified by their params pointer)
    std::vector<ClipPlan> clips;
    std::vector<int> clip_of(n);
    for (int i = 0; i < n; i++) {
        const svpf_analyse_params* p = params[i];
        if (!p) return fail(SVPF_ERR_ARG, "svpf_analyse: null params");
        int c = -1;
        for (size_t k = 0; k < clips.size(); k++)
            if (clips[k].p == p) c = (int)k;
        if (c < 0) {
            if (p->levels < 1 || p->levels > 15 || p->levels > p->super_levels)
                return fail(SVPF_ERR_ARG, "svpf_analyse: unsupported level count");
            if (p->vectors < 1 || p->vectors > 3) return fail(SVPF_ERR_ARG, "svpf_analyse: bad vectors");
            ClipPlan cp{};
            cp.p = p;
            cp.levels = p->levels;
            for (int l = 0; l < cp.levels; l++) {
                level_grid(*p, l, &cp.gw[l], &cp.gh[l]);
                if (cp.gw[l] < 2 || cp.gh[l] < 2) return fail(SVPF_ERR_ARG, "svpf_analyse: frame too small for the block grid");
                if (cp.gw[l] >= 65536 || cp.gh[l] >= 32768) return fail(SVPF_ERR_ARG, "svpf_analyse: grid too large");
                LevelGeom& g = cp.geom[l];
                g.level = l;
                g.gw = cp.gw[l];
                g.gh = cp.gh[l];
                g.lw = reduce_dim(p->width, l);
                g.lh = reduce_dim(p->height, l);
                int y = 0, cy = 0;
                for (int lv = 0; lv < l; lv++) {  // level_y_offset / chroma_level_offset (:49-58, :2704-2713)
                    int h = reduce_dim(p->height, lv);
                    int sub = lv == 0 ? p->pel * p->pel : 1;
                    y += h * sub;
                    cy += std::max(h / 2, 1) * sub;
                }
                g.y_off = y;
                g.c_off = cy;
                g.clw = g.lw / 2;
                g.clh = g.lh / 2;
                g.pel = l == 0 ? std::max(p->pel, 1) : 1;
                g.bw = p->block_w;
                g.bh = p->block_h;
                g.step_x = std::max(p->block_w - p->overlap_x, 1);
                g.step_y = std::max(p->block_h - p->overlap_y, 1);
                g.smallest = l == p->levels - 1;
                g.satd = l == 0 ? p->fine_satd : p->coarse_satd;
                g.search_type = l == 0 ? p->search_type : p->coarse_search_type;
                {  // search_radius (:1465-1476)
                    int d = l == 0 ? p->distance : p->coarse_distance;
                    g.radius = d == 0 ? ((l == 0 && p->levels > 1) ? 0 : 10) : std::abs(d);
                }
                g.lambda_base = g.smallest ? 0
                                           : f64_to_i32((double)p->lambda / powi_rt(p->plevel, p->levels - l - 1) /
                                                        (double)std::max(g.pel * g.pel, 1));  // exact_lambda
                g.lsad = g.smallest ? p->lsad / 3 : p->lsad;
                g.pnew = p->pnew;
                g.pglobal = p->pglobal;
                g.pzero = p->pzero;
                g.pnbour = p->pnbour;
                g.pprev = p->prev;
                g.early = l == 0 && p->distance == 0;
                g.bad_sad = p->coarse_bad_sad;
                g.bad_range = p->coarse_bad_range;
                g.adaptive = p->coarse_distance < 0;
            }
            clips.push_back(cp);
            c = (int)clips.size() - 1;
        }
        clip_of[i] = c;
    }
    int max_levels = 0;
    for (auto& c : clips) max_levels = std::max(max_levels, c.levels);
    const int ndir_of = 2;  // chains per pair: [fwd, bwd] (bwd unused when vectors == 2, etc.)

    // ---- descriptor blob (host) layout
    Carve blob;
    for (auto& c : clips) {
        c.geom_off = blob.take(sizeof(LevelGeom) * 16);
        c.header_off = blob.take(64);
        c.order_top_off = blob.take(sizeof(int32_t) * (size_t)c.gw[c.levels - 1] * c.gh[c.levels - 1]);
        c.order0_off = blob.take(sizeof(int32_t) * (size_t)c.gw[0] * c.gh[0]);
    }
    const int nchains = n * ndir_of;
    // per level step: chains, transitions, global jobs, sort jobs, order jobs, segments
    size_t chain_off[16], tr_off[16], gj_off[16], sj_off[16], oj_off[16], seg_off[16];
    for (int l = 0; l < max_levels; l++) {
        chain_off[l] = blob.take(sizeof(Chain) * nchains);
        tr_off[l] = blob.take(sizeof(Transition) * nchains);
        gj_off[l] = blob.take(sizeof(GlobalJob) * nchains);
        sj_off[l] = blob.take(sizeof(SortJob) * nchains);
        oj_off[l] = blob.take(sizeof(OrderJob) * nchains);
        seg_off[l] = blob.take(sizeof(int) * 2 * (nchains + 1));
    }
    size_t pack_off = blob.take(sizeof(PackJob) * nchains);
    size_t refine_off = blob.take(sizeof(RefineJob) * nchains);
    size_t init_off = blob.take(sizeof(Transition) * nchains);
    std::vector<uint8_t> host(blob.off, 0);

    // ---- device workspace layout (per chain per level buffers)
    Carve work;
    struct ChainBufs {
        size_t field[16], res[16], opos[16], proj[16], key[16], order[16], glob[16];
        size_t refined;  // the refine pass's field (output grid), clips with refine[] only
        uint32_t base[16];  // first index in the level's sweep flag arrays
    };
    std::vector<ChainBufs> cb(nchains);
    size_t sort_total[16] = {};
    size_t sort_keys_in[16], sort_keys_out[16], sort_vals_in[16], sort_vals_out[16];
    for (int ci = 0; ci < nchains; ci++) {
        const ClipPlan& c = clips[clip_of[ci / 2]];
        for (int l = 0; l < c.levels; l++) {
            size_t cnt = (size_t)c.gw[l] * c.gh[l];
            cb[ci].field[l] = work.take(sizeof(DVec) * cnt);
            cb[ci].res[l] = work.take(sizeof(DVec) * cnt);
            cb[ci].opos[l] = work.take(4 * cnt);
            cb[ci].proj[l] = work.take(sizeof(DVec) * cnt);
            cb[ci].key[l] = work.take(8 * cnt);
            cb[ci].order[l] = work.take(4 * cnt);
            cb[ci].glob[l] = work.take(8);
        }
        if (c.p->has_refine) {
            int bw, bh, ox, oy, gw, gh;
            output_block(*c.p, &bw, &bh, &ox, &oy);
            out_grid(*c.p, bw, bh, ox, oy, &gw, &gh);
            cb[ci].refined = work.take(sizeof(DVec) * (size_t)gw * gh);
        }
    }
    // sweep state per level: flags x2 (one entry per block of every chain), worklists x2, counters
    size_t level_blocks[16] = {}, flags_off[16][2], wl_off[16][2], counts_off[16], found_off[16], badf_off[16],
        cpre_off[16], minpos_off[16], memo_off[16], fseq_off[16], bseq_off[16], cseq_off[16], seed_off[16], dl_off[16],
        stamp_off[16];
    int level_max_n[16] = {};
    for (int l = 0; l < max_levels; l++) {
        size_t tot = 0;
        for (int ci = 0; ci < nchains; ci++) {
            const ClipPlan& c = clips[clip_of[ci / 2]];
            if (l >= c.levels) continue;
            cb[ci].base[l] = (uint32_t)tot;
            tot += (size_t)c.gw[l] * c.gh[l];
            level_max_n[l] = std::max(level_max_n[l], c.gw[l] * c.gh[l]);
        }
        if (tot >= 0xFFFFFFFFull) return fail(SVPF_ERR_ARG, "svpf_analyse: batch too large");
        level_blocks[l] = tot;
        for (int q = 0; q < 2; q++) {
            flags_off[l][q] = work.take(4 * tot + 4);
            wl_off[l][q] = work.take(8 * tot + 8);
        }
        counts_off[l] = work.take(8 * (size_t)kCntWords);  // counts, then dcount
        dl_off[l] = work.take(8 * tot + 8);
        found_off[l] = work.take(4 * tot + 4);
        badf_off[l] = work.take(4 * tot + 4);
        cpre_off[l] = work.take(4 * tot + 4);
        minpos_off[l] = work.take(4 * (size_t)nchains + 4);
        memo_off[l] = work.take(sizeof(Memo) * (tot + 1));
        fseq_off[l] = work.take(4 * tot + 4);
        bseq_off[l] = work.take(4 * tot + 4);
        cseq_off[l] = work.take(4 * tot + 4);
        seed_off[l] = work.take(sizeof(SeedRec) * (tot + 1));
        stamp_off[l] = work.take(4 * tot + 4);
    }
    for (int l = 0; l < max_levels; l++) {
        size_t tot = 0;
        for (int ci = 0; ci < nchains; ci++) {
            const ClipPlan& c = clips[clip_of[ci / 2]];
            if (l < c.levels) tot += (size_t)(c.gw[l] - 1) * (c.gh[l] - 1);
        }
        sort_total[l] = tot;
        sort_keys_in[l] = work.take(4 * tot + 4);
        sort_keys_out[l] = work.take(4 * tot + 4);
        sort_vals_in[l] = work.take(4 * tot + 4);
        sort_vals_out[l] = work.take(4 * tot + 4);
    }
    // sort temp storage (max over levels)
    size_t sort_tmp = 0;
    for (int l = 1; l < max_levels; l++) {
        if (!sort_total[l]) continue;
        size_t bytes = 0;
        SVPF_HIP(segmented_sort(nullptr, &bytes, nullptr, nullptr, nullptr, nullptr, (int)sort_total[l], nchains,
                                nullptr, nullptr, ctx->stream));
        sort_tmp = std::max(sort_tmp, bytes);
    }
    uint8_t* W = nullptr;
    void* sort_ptr = nullptr;
    size_t sort_cap = 0;
    void* d_blob = nullptr;
    if (ext) {
        // one external buffer: descriptor blob, then the workspace, then the sort storage
        const size_t work_at = (blob.off + 255) / 256 * 256;
        const size_t sort_at = (work_at + work.off + 255) / 256 * 256;
        const size_t need = sort_at + sort_tmp;
        if (need > ext->cap)
            return fail(SVPF_ERR_ARG, "svpf_analyse_pair_noalloc: this pair needs " + std::to_string(need) +
                                          " bytes of scratch memory, more than its maximum of " +
                                          std::to_string(ext->cap));
        d_blob = ext->base;
        W = ext->base + work_at;
        sort_ptr = ext->base + sort_at;
        sort_cap = sort_tmp;
    } else {
        if (work.off > ctx->work_cap) {
            SVPF_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->d_work) (void)hipFree(ctx->d_work);
            ctx->d_work = nullptr;
            ctx->work_cap = 0;
            size_t cap = work.off + work.off / 4;
            hipError_t e = hipMalloc(&ctx->d_work, cap);
            if (e != hipSuccess) return fail(SVPF_ERR_NOMEM, "svpf_analyse: workspace allocation failed");
            ctx->work_cap = cap;
        }
        if (sort_tmp > ctx->sort_tmp_cap) {
            SVPF_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->d_sort_tmp) (void)hipFree(ctx->d_sort_tmp);
            ctx->d_sort_tmp = nullptr;
            ctx->sort_tmp_cap = 0;
            hipError_t e = hipMalloc(&ctx->d_sort_tmp, sort_tmp);
            if (e != hipSuccess) return fail(SVPF_ERR_NOMEM, "svpf_analyse: sort workspace allocation failed");
            ctx->sort_tmp_cap = sort_tmp;
        }
        W = (uint8_t*)ctx->d_work;
        sort_ptr = ctx->d_sort_tmp;
        sort_cap = ctx->sort_tmp_cap;
    }

    // ---- fill static parts of the blob
    for (auto& c : clips) {
        std::memcpy(&host[c.geom_off], c.geom, sizeof(LevelGeom) * c.levels);
        int32_t h[16];
        header_words(*c.p, h);
        std::memcpy(&host[c.header_off], h, 64);
        auto top = level_order(c.gw[c.levels - 1], c.gh[c.levels - 1], c.levels - 1);
        std::memcpy(&host[c.order_top_off], top.data(), top.size() * 4);
        std::vector<int32_t> o0;
        initial_order(c.gw[0], c.gh[0], o0);
        std::memcpy(&host[c.order0_off], o0.data(), o0.size() * 4);
    }
    // Device addresses of blob entries are only known after the upload; pointers into
    // the blob are written as offsets first and patched below.
    auto blob_ptr = [&](size_t off) -> uint8_t* { return (uint8_t*)d_blob + off; };
    // The upload needs the final content, which includes device pointers into the blob:
    // reserve the device buffer first by uploading the current (partial) blob.
    int rc = SVPF_OK;
    if (!ext) {
        rc = svpf_upload_desc(ctx, host.data(), host.size(), &d_blob);
        if (rc) return rc;
    }

    // ---- per-step tables
    auto chain_alive = [&](int ci) {
        const svpf_analyse_params& p = *clips[clip_of[ci / 2]].p;
        return (ci & 1) ? (p.vectors & 1) != 0 : (p.vectors & 2) != 0;  // even = fwd, odd = bwd
    };
    bool too_large = false;
    auto to_splanes = [&](const svpf_frame& f, const svpf_analyse_params& p) {
        SPlanes s;
        int yh = super_plane_height(p.height, p.pel, p.super_levels, true), uvh = yh / 2;
        s.p[0] = f.y; s.p[1] = f.u; s.p[2] = f.v;
        s.stride[0] = f.stride_y; s.stride[1] = f.stride_u; s.stride[2] = f.stride_v;
        s.len[0] = (int64_t)f.stride_y * yh; s.len[1] = (int64_t)f.stride_u * uvh; s.len[2] = (int64_t)f.stride_v * uvh;
        for (int k = 0; k < 3; k++)
            if (s.len[k] >= INT32_MAX) too_large = true;  // the kernels index planes with 32-bit offsets
        return s;
    };
    int nch_step[16] = {}, ntr_step[16] = {}, nsort_step[16] = {}, max_tr_n[16] = {}, max_sort_n[16] = {};
    bool seq_step[16] = {}, bad_step[16] = {};  // sequential fallback / bad-block count rebuilds
    bool plan_step[16];                         // every searched block a 4x4 or 8x8 square
    bool l0_step[16];                           // the finest level of multi-level searches, pnew >= -256
    bool lean_step[16];                         // every chain's pnew >= -256 (no serial replay)
    for (bool& f : plan_step) f = true;
    for (int l = 0; l < 16; l++) l0_step[l] = l == 0;
    for (bool& f : lean_step) f = true;
    Transition* inits = (Transition*)&host[init_off];
    int ninit = 0, max_init_n = 0;
    for (int step = max_levels - 1; step >= 0; step--) {
        const int l = step;  // aligned at level 0: every clip runs level l at step l
        Chain* chains = (Chain*)&host[chain_off[l]];
        Transition* trs = (Transition*)&host[tr_off[l]];
        GlobalJob* gjs = (GlobalJob*)&host[gj_off[l]];
        SortJob* sjs = (SortJob*)&host[sj_off[l]];
        OrderJob* ojs = (OrderJob*)&host[oj_off[l]];
        int* seg = (int*)&host[seg_off[l]];
        int nc = 0, nt = 0, ns = 0;
        int seg_cursor = 0;
        for (int ci = 0; ci < nchains; ci++) {
            if (!chain_alive(ci)) continue;
            const int pi = ci / 2;
            const ClipPlan& c = clips[clip_of[pi]];
            const svpf_analyse_params& p = *c.p;
            if (l >= c.levels) continue;
            const bool top = l == c.levels - 1;
            const bool bidir = p.vectors == 3;
            const int partner = bidir ? (ci ^ 1) : ci;
            Chain ch{};
            // fwd searches cur against ref; bwd searches ref against cur (analyse_search.rs:160-191)
            ch.src = to_splanes((ci & 1) ? d_ref[pi] : d_cur[pi], p);
            ch.ref = to_splanes((ci & 1) ? d_cur[pi] : d_ref[pi], p);
            ch.in = (const DVec*)(W + cb[ci].field[l]);
            ch.res = (DVec*)(W + cb[ci].res[l]);
            ch.opos = (int32_t*)(W + cb[ci].opos[l]);
            ch.base = cb[ci].base[l];
            ch.prev = top ? nullptr : (const DVec*)(W + cb[partner].proj[l]);
            if (top) ch.order = (const int32_t*)blob_ptr(c.order_top_off);
            else if (l == 0) ch.order = (const int32_t*)blob_ptr(c.order0_off);
            else ch.order = (const int32_t*)(W + cb[ci].order[l]);
            ch.geom = c.geom[l];
            ch.global = top ? nullptr : (const int32_t*)(W + cb[ci].glob[l]);
            chains[nc++] = ch;
            if (top) {
                // clear the top level's field (the workspace is reused across calls)
                Transition t{};
                t.coarse = nullptr;
                t.next = (DVec*)(W + cb[ci].field[l]);
                t.proj_key = (unsigned long long*)(W + cb[ci].key[l]);
                t.proj = (DVec*)(W + cb[ci].proj[l]);
                t.gw = c.gw[l];
                t.gh = c.gh[l];
                inits[ninit++] = t;
                max_init_n = std::max(max_init_n, c.gw[l] * c.gh[l]);
            }
            if (!top) {
                // transition from level l+1 into level l
                Transition t{};
                t.coarse = (const DVec*)(W + cb[ci].res[l + 1]);
                t.next = (DVec*)(W + cb[ci].field[l]);
                t.proj_key = (unsigned long long*)(W + cb[ci].key[l]);
                t.proj = (DVec*)(W + cb[ci].proj[l]);
                t.cgw = c.gw[l + 1]; t.cgh = c.gh[l + 1]; t.gw = c.gw[l]; t.gh = c.gh[l];
                t.bw = p.block_w; t.bh = p.block_h; t.ox = p.overlap_x; t.oy = p.overlap_y;
                t.pel = l == 0 ? std::max(p.pel, 1) : 1;
                trs[nt] = t;
                GlobalJob gj{};
                gj.field = t.coarse;
                gj.n = c.gw[l + 1] * c.gh[l + 1];
                gj.lw = reduce_dim(p.width, l + 1);
                gj.lh = reduce_dim(p.height, l + 1);
                gj.out = (int*)(W + cb[ci].glob[l]);
                gjs[nt] = gj;
                max_tr_n[l] = std::max(max_tr_n[l], c.gw[l] * c.gh[l]);
                nt++;
                if (l > 0) {  // the coarse field's score order defines level l's visit order
                    SortJob sj{};
                    sj.field = t.coarse;
                    sj.keys = (uint32_t*)(W + sort_keys_in[l + 1]);
                    sj.vals = (uint32_t*)(W + sort_vals_in[l + 1]);
                    sj.gw = c.gw[l + 1];
                    sj.gh = c.gh[l + 1];
                    uint32_t v = (uint32_t)std::max(p.block_w * p.block_h / 8, 1);
                    sj.shift = 31 - __builtin_clz(v) + 1;
                    sj.seg_begin = seg_cursor;
                    int cnt = (sj.gw - 1) * (sj.gh - 1);
                    seg[ns] = seg_cursor;
                    seg[nchains + 1 + ns] = seg_cursor + cnt;
                    sjs[ns] = sj;
                    OrderJob oj{};
                    oj.parents = (const uint32_t*)(W + sort_vals_out[l + 1]) + seg_cursor;
                    oj.order = (int32_t*)(W + cb[ci].order[l]);
                    oj.cgw = c.gw[l + 1]; oj.cgh = c.gh[l + 1]; oj.gw = c.gw[l]; oj.gh = c.gh[l];
                    ojs[ns] = oj;
                    max_sort_n[l] = std::max(max_sort_n[l], cnt);
                    seg_cursor += cnt;
                    ns++;
                }
            }
        }
        nch_step[l] = nc;
        ntr_step[l] = nt;
        nsort_step[l] = ns;
        for (auto& c : clips) {
            if (l > 0 && l < c.levels && c.geom[l].bad_sad > 0) bad_step[l] = true;
            if (l < c.levels) {
                const LevelGeom& g = c.geom[l];
                const int sw = g.smallest ? g.bw >> 1 : g.bw, sh = g.smallest ? g.bh >> 1 : g.bh;
                if (sw != sh || (sw != 4 && sw != 8)) plan_step[l] = false;
                if (g.level != 0 || g.smallest || g.pnew < -256) l0_step[l] = false;
                if (g.pnew < -256) lean_step[l] = false;
            }
        }
    }
    if (too_large) return fail(SVPF_ERR_ARG, "svpf_analyse: super-frame plane of 2 GiB or more");
    PackJob* pj = (PackJob*)&host[pack_off];
    int npack = 0, max_count = 0;
    for (int ci = 0; ci < nchains; ci++) {
        if (!chain_alive(ci)) continue;
        const int pi = ci / 2;
        const ClipPlan& c = clips[clip_of[pi]];
        const svpf_analyse_params& p = *c.p;
        int bw, bh, ox, oy, gw, gh;
        output_block(p, &bw, &bh, &ox, &oy);
        out_grid(p, bw, bh, ox, oy, &gw, &gh);
        const int count = gw * gh;
        if (!p.has_refine && count != c.gw[0] * c.gh[0]) return fail(SVPF_ERR_ARG, "svpf_analyse: output grid mismatch");
        // region order: bwd ("previous") first, then fwd ("current")
        const bool is_bwd = (ci & 1) != 0;
        const size_t region = 4 + (size_t)count * 8;
        const size_t off = 0x40 + ((!is_bwd && (p.vectors & 1)) ? region : 0);
        PackJob j{};
        j.field = (const DVec*)(W + (p.has_refine ? cb[ci].refined : cb[ci].res[0]));
        j.out = d_payload[pi] + off;
        j.payload = d_payload[pi];
        j.header = (const int32_t*)blob_ptr(c.header_off);
        j.count = count;
        pj[npack++] = j;
        max_count = std::max(max_count, count);
        if (((uintptr_t)d_payload[pi]) & 3) return fail(SVPF_ERR_ARG, "svpf_analyse: payload must be 4-byte aligned");
    }
    // refine pass jobs (exact_recalculate) of the clips that ask for it
    RefineJob* rj = (RefineJob*)&host[refine_off];
    int nrefine = 0, max_refine = 0;
    for (int ci = 0; ci < nchains; ci++) {
        if (!chain_alive(ci)) continue;
        const int pi = ci / 2;
        const ClipPlan& c = clips[clip_of[pi]];
        const svpf_analyse_params& p = *c.p;
        if (!p.has_refine) continue;
        int bw, bh, ox, oy, gw, gh;
        output_block(p, &bw, &bh, &ox, &oy);
        out_grid(p, bw, bh, ox, oy, &gw, &gh);
        RefineJob j{};
        j.ch.src = to_splanes((ci & 1) ? d_ref[pi] : d_cur[pi], p);
        j.ch.ref = to_splanes((ci & 1) ? d_cur[pi] : d_ref[pi], p);
        j.ch.in = (const DVec*)(W + cb[ci].res[0]);
        j.ch.res = (DVec*)(W + cb[ci].refined);
        LevelGeom g = c.geom[0];
        g.gw = gw;
        g.gh = gh;
        g.bw = bw;
        g.bh = bh;
        g.step_x = std::max(bw - ox, 1);
        g.step_y = std::max(bh - oy, 1);
        g.smallest = 0;
        g.satd = p.refine_satd;
        g.search_type = p.refine_search_type;
        g.radius = std::abs(p.refine_distance);
        g.early = 0;
        g.bad_sad = 0;
        g.bad_range = 0;
        g.adaptive = 0;
        j.ch.geom = g;
        j.old_gw = c.gw[0];
        j.old_gh = c.gh[0];
        j.old_bw = p.block_w;
        j.old_bh = p.block_h;
        j.old_sx = std::max(p.block_w - p.overlap_x, 1);
        j.old_sy = std::max(p.block_h - p.overlap_y, 1);
        j.thsad = p.refine_thsad;
        j.lambda = p.lambda;
        j.pnew = p.pnew;
        j.type = p.refine_search_type;
        j.distance = p.refine_distance;
        rj[nrefine++] = j;
        max_refine = std::max(max_refine, gw * gh);
    }
    // rescale_scores denominator per clip: floor(sqrt(W^2 + H^2)) * output block width
    int denom = -1;
    for (auto& c : clips) {
        int bw, bh, ox, oy;
        output_block(*c.p, &bw, &bh, &ox, &oy);
        double f = std::sqrt(powi_rt((double)c.p->width, 2) + powi_rt((double)c.p->height, 2));
        int diag = std::max(f64_to_i32(f), 0);
        int d = std::max(sat_mul(diag, std::max(bw, 1)), 1);
        if (denom >= 0 && d != denom) denom = -2;
        else if (denom == -1) denom = d;
    }
    if (denom < 0) {
        // clips with different denominators: pack per clip below (one launch per clip)
    }
    // final upload with every device pointer in place
    if (ext) {
        // a copy from pageable memory, ordered after the previous call's kernels on the
        // stream; the wait lets `host` go out of scope whatever the runtime's staging
        SVPF_HIP(hipMemcpyAsync(d_blob, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
        SVPF_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        void* first_blob = d_blob;
        rc = svpf_upload_desc(ctx, host.data(), host.size(), &d_blob);
        if (rc) return rc;
        if (d_blob != first_blob) return fail(SVPF_ERR_STATE, "svpf_analyse: descriptor buffer moved");
    }

    // ---- launches
    if (ninit) SVPF_HIP(launch_interp(blob_ptr(init_off), ninit, max_init_n, ctx->stream));
    for (int l = max_levels - 1; l >= 0; l--) {
        if (ntr_step[l]) {
            SVPF_HIP(launch_interp(blob_ptr(tr_off[l]), ntr_step[l], max_tr_n[l], ctx->stream));
            SVPF_HIP(launch_global(blob_ptr(gj_off[l]), ntr_step[l], ctx->stream));
        }
        if (nsort_step[l]) {
            SVPF_HIP(launch_sort_keys(blob_ptr(sj_off[l]), nsort_step[l], max_sort_n[l], ctx->stream));
            size_t bytes = sort_cap;
            const int* segp = (const int*)blob_ptr(seg_off[l]);
            const int total = (int)sort_total[l + 1];
            SVPF_HIP(segmented_sort(sort_ptr, &bytes, (const uint32_t*)(W + sort_keys_in[l + 1]),
                                    (uint32_t*)(W + sort_keys_out[l + 1]), (const uint32_t*)(W + sort_vals_in[l + 1]),
                                    (uint32_t*)(W + sort_vals_out[l + 1]), total, nsort_step[l], segp,
                                    segp + nchains + 1, ctx->stream));
            SVPF_HIP(launch_child_order(blob_ptr(oj_off[l]), nsort_step[l], ctx->stream));
        }
        if (nch_step[l]) {
            const int ph = prof_begin(ctx, kProfSearch);
            const Chain* dch = (const Chain*)blob_ptr(chain_off[l]);
            SweepState st{};
            for (int q = 0; q < 2; q++) {
                st.wl[q] = (unsigned long long*)(W + wl_off[l][q]);
                st.flags[q] = (uint32_t*)(W + flags_off[l][q]);
            }
            st.counts = (uint32_t*)(W + counts_off[l]);
            st.dcount = st.counts + kCntWords;
            st.dl = (unsigned long long*)(W + dl_off[l]);
            st.found = (uint32_t*)(W + found_off[l]);
            st.badf = (uint32_t*)(W + badf_off[l]);
            st.cpre = (int32_t*)(W + cpre_off[l]);
            st.minpos = (uint32_t*)(W + minpos_off[l]);
            st.memo = (Memo*)(W + memo_off[l]);
            st.fseq = (uint32_t*)(W + fseq_off[l]);
            st.bseq = (uint32_t*)(W + bseq_off[l]);
            st.cseq = (int32_t*)(W + cseq_off[l]);
            st.seeds = (SeedRec*)(W + seed_off[l]);
            st.stamp = (uint32_t*)(W + stamp_off[l]);
            st.chase = ctx->chase;
            st.max_sweeps = ctx->max_sweeps > 0 ? ctx->max_sweeps : kMaxSweeps;
            SVPF_HIP(launch_sweep_init(dch, nch_step[l], level_max_n[l], st, ctx->stream));
            int rc2 = seq_step[l] ? SVPF_OK
                                  : run_sweeps(ctx, dch, nch_step[l], level_max_n[l], level_blocks[l], st, bad_step[l],
                                               l, plan_step[l], l0_step[l], lean_step[l], ext != nullptr);
            if (rc2 == kNotConverged || seq_step[l])
                SVPF_HIP(launch_search_level(dch, nch_step[l], ctx->stream));
            else if (rc2)
                return rc2;
            prof_end(ctx, ph);
        }
    }
    if (nrefine) SVPF_HIP(launch_refine(blob_ptr(refine_off), nrefine, max_refine, ctx->stream));
    if (denom > 0) {
        SVPF_HIP(launch_pack(blob_ptr(pack_off), npack, max_count, denom, ctx->stream));
    } else {
        // per-clip denominators
        int k = 0;
        for (int ci = 0; ci < nchains; ci++) {
            if (!chain_alive(ci)) continue;
            const svpf_analyse_params& p = *clips[clip_of[ci / 2]].p;
            int bw, bh, ox, oy;
            output_block(p, &bw, &bh, &ox, &oy);
            double f = std::sqrt(powi_rt((double)p.width, 2) + powi_rt((double)p.height, 2));
            int d = std::max(sat_mul(std::max(f64_to_i32(f), 0), std::max(bw, 1)), 1);
            SVPF_HIP(launch_pack(blob_ptr(pack_off + sizeof(PackJob) * k), 1, pj[k].count, d, ctx->stream));
            k++;
        }
    }
    return SVPF_OK;
}

}  // namespace svpf

using namespace svpf;

extern "C" {

int svpf_analyse_params_from_opt(const char* opt, size_t opt_len, int64_t sdata, int32_t width, int32_t height,
                                 svpf_analyse_params* out, int32_t* vdata16, int32_t* payload_len, char* err,
                                 size_t err_len) {
    auto reject = [&](const char* msg) {
        if (err && err_len) {
            std::strncpy(err, msg, err_len - 1);
            err[err_len - 1] = 0;
        }
        return fail(SVPF_ERR_OPT, msg);
    };
    if (!out) return fail(SVPF_ERR_ARG, "svpf_analyse_params_from_opt: out is null");
    Opt o;
    bool have = opt && opt_len && Opt::parse(opt, opt_len, &o, nullptr);
    const Opt* op = have ? &o : nullptr;
    svpf_analyse_params p{};
    // unpack_sdata (analyse_opts.rs:243-257)
    uint64_t raw;
    std::memcpy(&raw, &sdata, 8);
    p.height = (int)(raw & 0xFFFF);
    p.width = (int)((raw >> 16) & 0xFFFF);
    p.pel = (int)((raw >> 32) & 0xFF);
    p.super_levels = (int)((raw >> 40) & 0xFF);
    p.selector = (int)((raw >> 48) & 0xFF);
    if (p.width <= 0 || p.height <= 0 || p.pel < 1 || p.pel > 4 || p.super_levels <= 0)
        return reject("SVAnalyse: invalid super clip params");
    int64_t v;
    double f;
    bool b;
    auto geti = [&](std::initializer_list<const char*> a, int64_t dflt) { return (op && op->get_int(a, &v)) ? v : dflt; };
    auto geti2 = [&](std::initializer_list<const char*> a, std::initializer_list<const char*> b2, int64_t dflt) {
        if (op && op->get_int(a, &v)) return v;
        if (op && op->get_int(b2, &v)) return v;
        return dflt;
    };
    auto getf2 = [&](std::initializer_list<const char*> a, std::initializer_list<const char*> b2, double dflt) {
        if (op && op->get_float(a, &f)) return f;
        if (op && op->get_float(b2, &f)) return f;
        return dflt;
    };
    auto getb2 = [&](std::initializer_list<const char*> a, std::initializer_list<const char*> b2, bool dflt) {
        if (op && op->get_bool(a, &b)) return b;
        if (op && op->get_bool(b2, &b)) return b;
        return dflt;
    };
    p.vectors = (int32_t)geti({"vectors"}, 3);
    if (p.vectors < 1 || p.vectors > 3) return reject("SVAnalyse: vectors must be 1, 2 or 3");
    p.block_w = (int32_t)geti({"block", "w"}, 16);
    p.block_h = (int32_t)geti({"block", "h"}, p.block_w);
    if (p.block_w != 4 && p.block_w != 8 && p.block_w != 16 && p.block_w != 32)
        return reject("SVAnalyse: block.w must be 4/8/16/32");
    p.overlap_mode = (int32_t)geti({"block", "overlap"}, 2);
    switch (p.overlap_mode) {  // overlap_from_mode (:259-269)
        case 0: p.overlap_x = 0; p.overlap_y = 0; break;
        case 1: p.overlap_x = p.block_w / 8; p.overlap_y = p.block_h / 8; break;
        case 2: p.overlap_x = p.block_w / 4; p.overlap_y = p.block_h / 4; break;
        case 3: p.overlap_x = p.block_w / 2; p.overlap_y = p.block_h / 2; break;
        default: return reject("SVAnalyse: overlap must be 0, 1 (=1/8*block), 2 (=1/4*block) or 3 (=1/2*block)");
    }
    p.delta = (int32_t)geti2({"special", "delta"}, {"delta"}, 1);
    double lambda = getf2({"main", "penalty", "lambda"}, {"penalty", "lambda"}, 10.0);
    int32_t lsad = (int32_t)geti2({"main", "penalty", "lsad"}, {"penalty", "lsad"}, 8000);
    p.plevel = std::fmax(getf2({"main", "penalty", "plevel"}, {"penalty", "plevel"}, 1.5), 0.01);
    p.pnew = (int32_t)geti2({"main", "penalty", "pnew"}, {"penalty", "pnew"}, 50);
    p.pglobal = (int32_t)geti2({"main", "penalty", "pglobal"}, {"penalty", "pglobal"}, 50);
    p.pzero = (int32_t)geti2({"main", "penalty", "pzero"}, {"penalty", "pzero"}, 100);
    p.pnbour = (int32_t)geti2({"main", "penalty", "pnbour"}, {"penalty", "pnbour"}, 50);
    p.prev = (int32_t)geti2({"main", "penalty", "prev"}, {"penalty", "prev"}, 0);
    p.search_type = (int32_t)geti2({"main", "search", "type"}, {"search", "type"}, 4);
    p.distance = (int32_t)geti2({"main", "search", "distance"}, {"search", "distance"}, 2);
    p.coarse_search_type = (int32_t)geti2({"main", "search", "coarse", "type"}, {"search", "coarse", "type"}, p.search_type);
    p.coarse_distance = (int32_t)geti2({"main", "search", "coarse", "distance"}, {"search", "coarse", "distance"}, 0);
    p.coarse_trymany = getb2({"main", "search", "coarse", "trymany"}, {"search", "coarse", "trymany"}, false);
    int32_t bad_sad = (int32_t)geti2({"main", "search", "coarse", "bad", "sad"}, {"search", "coarse", "bad", "sad"}, 1000);
    p.coarse_bad_range = (int32_t)geti({"main", "search", "coarse", "bad", "range"}, -24);
    p.fine_satd = getb2({"main", "search", "satd"}, {"search", "satd"}, false);
    p.coarse_satd = (op && op->get_bool({"main", "search", "coarse", "satd"}, &b)) ? b : true;
    p.sort = getb2({"main", "search", "sort"}, {"sort"}, true);
    const Opt* refine = op ? op->get_array_first({"refine"}) : nullptr;
    p.has_refine = 0;
    if (refine && refine->get_int({"thsad"}, &v)) {
        p.has_refine = 1;
        uint32_t m = (uint32_t)std::max((p.block_w / 2) * (p.block_h / 2), 1);
        p.refine_thsad = (int32_t)((uint32_t)(int32_t)v * m) / 32;  // wrapping i32 multiply, then / 32
    }
    p.refine_search_type = (refine && refine->get_int({"search", "type"}, &v)) ? (int32_t)v : p.search_type;
    p.refine_distance = (refine && refine->get_int({"search", "distance"}, &v)) ? (int32_t)v : p.distance;
    p.refine_satd = (refine && refine->get_bool({"search", "satd"}, &b)) ? b : p.fine_satd;
    int levels = std::max(std::min(auto_search_levels(p.width, p.height, p.block_w, p.block_h, p.overlap_x,
                                                      p.overlap_y), p.super_levels), 1);
    if (op && op->get_int({"main", "levels"}, &v) && v > 0) levels = std::max(std::min((int)v, p.super_levels), 1);
    p.levels = levels;
    double area = (double)std::max(p.block_w * p.block_h, 1);
    p.lambda = f64_to_i32(2.0 * area * 1000.0 * 0.015625 * lambda);
    int area_scale = std::max(p.block_w * p.block_h, 1) / 32;
    p.lsad = sat_mul(lsad, area_scale);
    p.coarse_bad_sad = sat_mul(bad_sad, area_scale);
    if (width > 0) p.width = width;    // get_frame glue: the source clip's size wins
    if (height > 0) p.height = height;
    *out = p;
    int bw, bh, ox, oy, gw, gh;
    output_block(p, &bw, &bh, &ox, &oy);
    out_grid(p, bw, bh, ox, oy, &gw, &gh);
    if (vdata16) {
        int32_t h[16] = {0xA0, p.vectors == 1 ? 1 : (p.vectors == 2 ? 2 : 3), bw, bh, p.pel,
                         p.has_refine ? 1 : p.levels, 0, p.width, p.height, ox, oy, gw, gh, p.selector, p.delta, 0};
        std::memcpy(vdata16, h, sizeof h);
    }
    if (payload_len) *payload_len = 0x40 + ((p.vectors == 3) ? 2 : 1) * (4 + gw * gh * 8);
    return SVPF_OK;
}

int svpf_analyse_multi(svpf_ctx* ctx, int32_t n, const svpf_analyse_params* const* params, const svpf_frame* d_cur,
                       const svpf_frame* d_ref, uint8_t* const* d_payload) {
    if (!ctx || n < 0 || (n > 0 && (!params || !d_cur || !d_ref || !d_payload)))
        return fail(SVPF_ERR_ARG, "svpf_analyse: null argument");
    SVPF_HIP(hipSetDevice(ctx->device));
    return analyse_run(ctx, n, params, d_cur, d_ref, d_payload);
}

int svpf_analyse_batch(svpf_ctx* ctx, const svpf_analyse_params* p, int32_t n, const svpf_frame* d_cur,
                       const svpf_frame* d_ref, uint8_t* const* d_payload) {
    std::vector<const svpf_analyse_params*> ps((size_t)std::max(n, 0), p);
    return svpf_analyse_multi(ctx, n, ps.data(), d_cur, d_ref, d_payload);
}

int svpf_analyse_pair(svpf_ctx* ctx, const svpf_analyse_params* p, const svpf_frame* d_cur, const svpf_frame* d_ref,
                      uint8_t* d_payload) {
    return svpf_analyse_batch(ctx, p, 1, d_cur, d_ref, &d_payload);
}

int svpf_analyse_set_tuning(svpf_ctx* ctx, const svpf_analyse_tuning* t) {
    if (!ctx || !t) return fail(SVPF_ERR_ARG, "svpf_analyse_set_tuning: null argument");
    if (t->fine_run != 0 && t->fine_run != 1 && t->fine_run != 2 && t->fine_run != 4 && t->fine_run != 8 &&
        t->fine_run != 16 && t->fine_run != 32)
        return fail(SVPF_ERR_ARG, "svpf_analyse_set_tuning: fine_run must be 1, 2, 4, 8, 16 or 32");
    if (t->coarse_run != 0 && t->coarse_run != 4 && t->coarse_run != 8 && t->coarse_run != 16 && t->coarse_run != 32)
        return fail(SVPF_ERR_ARG, "svpf_analyse_set_tuning: coarse_run must be 4, 8, 16 or 32");
    if (t->chase > 65536 || t->slot_scale < 0.0 || t->slot_scale > 16.0 || t->coarse_slots < 0 ||
        t->coarse_slots > (1 << 20) || t->sweep_grid < 0 || t->sweep_grid > (1 << 20) || t->graphs < -1 ||
        t->graphs > 1 || t->fast_list_min < -1)
        return fail(SVPF_ERR_ARG, "svpf_analyse_set_tuning: setting out of range");
    ctx->chase = std::max(t->chase, 0);
    ctx->slot_scale = t->slot_scale == 0.0 ? 1.0 : t->slot_scale;
    ctx->fine_run = t->fine_run == 0 ? 32 : t->fine_run;
    ctx->coarse_slots = t->coarse_slots == 0 ? 256 : t->coarse_slots;
    ctx->coarse_run = t->coarse_run == 0 ? 16 : t->coarse_run;
    ctx->sweep_grid = t->sweep_grid == 0 ? 2048 : t->sweep_grid;
    ctx->sweep_graphs = t->graphs >= 0;
    ctx->fast_list_min = t->fast_list_min == 0 ? kFastListMin : t->fast_list_min;
    return SVPF_OK;
}

int svpf_analyse_set_max_sweeps(svpf_ctx* ctx, int32_t max_sweeps) {
    if (!ctx || max_sweeps < 0 || max_sweeps > kMaxSweeps)
        return fail(SVPF_ERR_ARG, "svpf_analyse_set_max_sweeps: bad argument");
    ctx->max_sweeps = max_sweeps;
    return SVPF_OK;
}

}  // extern "C"
