// C-ABI runtime and Super entry points of libsvpf (include/svpf.h).
#include <cstring>
#include <mutex>

#include "options.h"
#include "svpf_internal.h"


namespace svpf {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int prof_begin(svpf_ctx* ctx, int kernel) {
    if (!ctx->prof) return -1;
    hipEvent_t ev[2];
    for (auto& e : ev) {
        if (!ctx->prof_pool.empty()) {
            e = ctx->prof_pool.back();
            ctx->prof_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            return -1;
        }
    }
    if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) {
        ctx->prof_pool.push_back(ev[0]);
        ctx->prof_pool.push_back(ev[1]);
        return -1;
    }
    ctx->prof_marks.push_back(svpf_ctx::ProfMark{kernel, ev[0], ev[1]});
    return (int)ctx->prof_marks.size() - 1;
}

void prof_end(svpf_ctx* ctx, int handle) {
    if (handle < 0 || handle >= (int)ctx->prof_marks.size()) return;
    (void)hipEventRecord(ctx->prof_marks[(size_t)handle].b, ctx->stream);
}

static void prof_clear(svpf_ctx* ctx) {
    for (auto& m : ctx->prof_marks) {
        ctx->prof_pool.push_back(m.a);
        ctx->prof_pool.push_back(m.b);
    }
    ctx->prof_marks.clear();
}

int reduce_dim(int dim, int n) {  // super_opts.rs:70-75
    for (int i = 0; i < n; i++) dim = 2 * (dim / 4);
    return dim;
}

int auto_levels(int width, int height) {  // super_opts.rs:77-87
    int lv = 0;
    while (reduce_dim(height, lv) >= 4 && reduce_dim(width, lv) >= 4) lv++;
    return lv;
}

int super_plane_height(int base_h, int pel, int levels, bool full) {  // super_opts.rs:89-95
    int rows = full ? pel * pel * base_h : 0;
    for (int lv = 1; lv < levels; lv++) rows += reduce_dim(base_h, lv);
    return rows;
}

SuperGeom make_super_geom(const svpf_super_params& p) {
    SuperGeom g{};
    g.width = p.width;
    g.height = p.height;
    g.pel = p.pel;
    g.levels = p.levels;
    g.gpu = p.gpu;
    g.scale_down = p.scale_down;
    g.scale_up = p.scale_up;
    g.full = p.full;
    int sh = super_plane_height(p.height, p.pel, p.levels, p.full != 0);
    g.rows[0] = sh;
    g.rows[1] = g.rows[2] = sh / 2;
    for (int plane = 0; plane < 3; plane++) {
        int row = 0;
        for (int lv = 0; lv < p.levels && lv < 16; lv++) {
            int w = reduce_dim(p.width, lv), h = reduce_dim(p.height, lv);
            if (plane) { w /= 2; h /= 2; }
            g.level_row[plane][lv] = row;
            g.level_w[plane][lv] = w;
            g.level_h[plane][lv] = h;
            row += h * ((lv == 0 && p.full) ? p.pel * p.pel : 1);
        }
    }
    return g;
}

}  // namespace svpf

// Stream-ordered upload of a small descriptor table; returns its device address.
int svpf_upload_desc(svpf_ctx* ctx, const void* data, size_t bytes, void** d_out) {
    using namespace svpf;
    SVPF_HIP(hipEventSynchronize(ctx->desc_done));  // staging buffer free again
    if (bytes > ctx->desc_cap) {
        SVPF_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->d_desc) (void)hipFree(ctx->d_desc);
        if (ctx->h_desc) (void)hipHostFree(ctx->h_desc);
        ctx->d_desc = ctx->h_desc = nullptr;
        ctx->desc_cap = 0;
        size_t cap = bytes < 65536 ? 65536 : bytes;
        SVPF_HIP(hipMalloc(&ctx->d_desc, cap));
        SVPF_HIP(hipHostMalloc(&ctx->h_desc, cap, hipHostMallocDefault));
        ctx->desc_cap = cap;
    }
    std::memcpy(ctx->h_desc, data, bytes);
    SVPF_HIP(hipMemcpyAsync(ctx->d_desc, ctx->h_desc, bytes, hipMemcpyHostToDevice, ctx->stream));
    SVPF_HIP(hipEventRecord(ctx->desc_done, ctx->stream));
    *d_out = ctx->d_desc;
    return SVPF_OK;
}

namespace svpf {

static Planes to_planes(const svpf_frame& f) {
    Planes p;
    p.p[0] = f.y; p.p[1] = f.u; p.p[2] = f.v;
    p.stride[0] = f.stride_y; p.stride[1] = f.stride_u; p.stride[2] = f.stride_v;
    return p;
}

}  // namespace svpf

using namespace svpf;

extern "C" {

const char* svpf_last_error(void) { return g_error.c_str(); }

int svpf_version(void) { return 1174405393; }  // svpflow1/src/abi.rs:3-8, svpflow2/src/strings.rs:1

int svpf_create(int device, svpf_ctx** out) {
    if (!out) return fail(SVPF_ERR_ARG, "svpf_create: out is null");
    *out = nullptr;
    int count = 0;
    SVPF_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(SVPF_ERR_ARG, "svpf_create: no such device");
    SVPF_HIP(hipSetDevice(device));
    auto* c = new svpf_ctx{};
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->t0);
    if (e == hipSuccess) e = hipEventCreate(&c->t1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->desc_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(c->desc_done, c->stream);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_sweep_stats, 16 * 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_sweep_stats, 0, 16 * 4 * sizeof(unsigned long long), c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(SVPF_ERR_HIP, std::string("svpf_create: ") + hipGetErrorString(e));
    }
    // device address of svpf_analyse_pair_noalloc's scratch array on this device
    const int rc = noalloc_scratch_init(device);
    if (rc != SVPF_OK) {
        svpf_destroy(c);
        return rc;
    }
    *out = c;
    return SVPF_OK;
}

void svpf_destroy(svpf_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_desc) (void)hipFree(ctx->d_desc);
    if (ctx->h_desc) (void)hipHostFree(ctx->h_desc);
    prof_clear(ctx);
    for (auto e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (ctx->h_count) (void)hipHostFree(ctx->h_count);
    for (auto& g : ctx->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    if (ctx->d_sweep_stats) (void)hipFree(ctx->d_sweep_stats);
    (void)hipEventDestroy(ctx->desc_done);
    (void)hipEventDestroy(ctx->t0);
    (void)hipEventDestroy(ctx->t1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int svpf_synchronize(svpf_ctx* ctx) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    return SVPF_OK;
}

void* svpf_stream(svpf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int svpf_device_alloc(svpf_ctx* ctx, size_t bytes, void** d_out) {
    if (!ctx || !d_out) return fail(SVPF_ERR_ARG, "svpf_device_alloc: null argument");
    SVPF_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(d_out, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(SVPF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return SVPF_OK;
}

int svpf_device_free(svpf_ctx* ctx, void* d_ptr) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    if (!d_ptr) return SVPF_OK;
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    SVPF_HIP(hipFree(d_ptr));
    return SVPF_OK;
}

int svpf_copy_h2d(svpf_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    SVPF_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    return SVPF_OK;
}

int svpf_copy_d2h(svpf_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    SVPF_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    return SVPF_OK;
}

int svpf_copy_d2d(svpf_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    SVPF_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SVPF_OK;
}

int svpf_memset_d(svpf_ctx* ctx, void* d_dst, int value, size_t bytes) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    SVPF_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    return SVPF_OK;
}

int svpf_timer_start(svpf_ctx* ctx) {
    if (!ctx) return fail(SVPF_ERR_ARG, "null context");
    SVPF_HIP(hipEventRecord(ctx->t0, ctx->stream));
    return SVPF_OK;
}

int svpf_timer_stop(svpf_ctx* ctx, float* ms) {
    if (!ctx || !ms) return fail(SVPF_ERR_ARG, "null argument");
    SVPF_HIP(hipEventRecord(ctx->t1, ctx->stream));
    SVPF_HIP(hipEventSynchronize(ctx->t1));
    SVPF_HIP(hipEventElapsedTime(ms, ctx->t0, ctx->t1));
    return SVPF_OK;
}

// ------------------------------------------------------------------ Super

int svpf_super_params_from_opt(const char* opt, size_t opt_len, int32_t width, int32_t height,
                               svpf_super_params* out, int64_t* sdata, char* err, size_t err_len) {
    auto reject = [&](const char* msg) {
        if (err && err_len) {
            std::strncpy(err, msg, err_len - 1);
            err[err_len - 1] = 0;
        }
        return fail(SVPF_ERR_OPT, msg);
    };
    if (!out) return fail(SVPF_ERR_ARG, "svpf_super_params_from_opt: out is null");
    Opt o;
    bool have = false;
    if (opt && opt_len) have = Opt::parse(opt, opt_len, &o, nullptr);  // unparsable -> defaults
    const Opt* op = have ? &o : nullptr;
    int64_t v;
    bool b;
    svpf_super_params p{};
    p.pel = (op && op->get_int({"pel"}, &v)) ? (int32_t)v : 2;
    if (p.pel != 1 && p.pel != 2 && p.pel != 4) return reject("SVSuper: pel has to be 1 or 2 or 4");
    p.gpu = (op && op->get_int({"gpu"}, &v)) ? (int32_t)v : 1;
    if (p.gpu < 0 || p.gpu > 2) return reject("SVSuper: gpu has to be 0 or 1 or 2");
    p.scale_down = (op && op->get_int({"scale", "down"}, &v)) ? (int32_t)v : 4;
    p.scale_up = (op && op->get_int({"scale", "up"}, &v)) ? (int32_t)v : 2;
    if (p.gpu != 0) p.scale_up = 0;
    p.full = (op && op->get_bool({"full"}, &b)) ? b : 1;
    if (p.pel >= 2 && !p.full) return reject("SVSuper: can't ignore finest level with high pel value");
    p.rc = (op && op->get_bool({"rc"}, &b)) ? b : 0;
    p.levels = auto_levels(width, height);
    p.width = width;
    p.height = height;
    *out = p;
    if (sdata) {
        uint64_t flags = (p.rc ? 1u : 0u) | (p.full ? 0u : 2u);
        uint64_t raw = (uint64_t)(uint16_t)p.height | ((uint64_t)(uint16_t)p.width << 16) |
                       ((uint64_t)(uint8_t)p.pel << 32) | ((uint64_t)(uint8_t)p.levels << 40) |
                       ((uint64_t)(uint8_t)p.gpu << 48) | (flags << 56);
        std::memcpy(sdata, &raw, 8);
    }
    return SVPF_OK;
}

int32_t svpf_super_height(const svpf_super_params* p) {
    return p ? super_plane_height(p->height, p->pel, p->levels, p->full != 0) : 0;
}

int svpf_super_build_batch(svpf_ctx* ctx, const svpf_super_params* p, int32_t n, const svpf_frame* d_src,
                           const svpf_frame* d_dst) {
    if (!ctx || !p || !d_src || !d_dst || n < 0) return fail(SVPF_ERR_ARG, "svpf_super_build: null argument");
    if (n == 0 || p->levels <= 0) return SVPF_OK;
    if (p->levels > 16 || p->width < 4 || p->height < 4 || (p->width & 1) || (p->height & 1))
        return fail(SVPF_ERR_ARG, "svpf_super_build: unsupported frame size");
    SuperGeom g = make_super_geom(*p);
    int max_stride = 0;
    std::vector<Planes> desc(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        desc[i] = to_planes(d_src[i]);
        desc[n + i] = to_planes(d_dst[i]);
        for (int k = 0; k < 3; k++) {
            int need_src = k ? p->width / 2 : p->width;
            if (!desc[i].p[k] || !desc[n + i].p[k]) return fail(SVPF_ERR_ARG, "svpf_super_build: null plane");
            if (desc[i].stride[k] < need_src || desc[n + i].stride[k] < need_src)
                return fail(SVPF_ERR_ARG, "svpf_super_build: stride smaller than the plane width");
            if (desc[n + i].stride[k] > max_stride) max_stride = desc[n + i].stride[k];
        }
    }
    SVPF_HIP(hipSetDevice(ctx->device));
    void* d_desc = nullptr;
    int rc = svpf_upload_desc(ctx, desc.data(), desc.size() * sizeof(Planes), &d_desc);
    if (rc) return rc;
    const Planes* ds = (const Planes*)d_desc;
    const Planes* dd = ds + n;
    const int ph = prof_begin(ctx, kProfSuper);
    SVPF_HIP(launch_super_level0(g, ds, dd, n, max_stride, ctx->stream));
    for (int lv = 0; lv + 1 < p->levels; lv++) SVPF_HIP(launch_super_reduce(g, dd, n, lv, ctx->stream));
    prof_end(ctx, ph);
    return SVPF_OK;
}

int svpf_super_build(svpf_ctx* ctx, const svpf_super_params* p, const svpf_frame* d_src, const svpf_frame* d_dst) {
    return svpf_super_build_batch(ctx, p, 1, d_src, d_dst);
}

int svpf_profile_enable(svpf_ctx* ctx, int on) {
    if (!ctx) return fail(SVPF_ERR_ARG, "svpf_profile_enable: null context");
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    prof_clear(ctx);
    ctx->prof = on != 0;
    SVPF_HIP(hipMemsetAsync(ctx->d_sweep_stats, 0, 16 * 4 * sizeof(unsigned long long), ctx->stream));
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    return SVPF_OK;
}

int svpf_analyse_stats(svpf_ctx* ctx, int32_t level, int64_t* out4) {
    if (!ctx || !out4 || level < 0 || level >= 16) return fail(SVPF_ERR_ARG, "svpf_analyse_stats: bad argument");
    unsigned long long t[4];
    SVPF_HIP(hipMemcpyAsync(t, ctx->d_sweep_stats + 4 * level, sizeof t, hipMemcpyDeviceToHost, ctx->stream));
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; k++) out4[k] = (int64_t)t[k];
    return SVPF_OK;
}

int svpf_profile_read(svpf_ctx* ctx, const char* kernel, double* total_ms, int32_t* launches) {
    static const char* const kNames[kProfCount] = {"super", "search", "analyse_aux", "smooth_prep", "render"};
    if (!ctx || !kernel || !total_ms || !launches) return fail(SVPF_ERR_ARG, "svpf_profile_read: null argument");
    int id = -1;
    for (int k = 0; k < kProfCount; k++)
        if (std::strcmp(kernel, kNames[k]) == 0) id = k;
    if (id < 0) return fail(SVPF_ERR_ARG, std::string("svpf_profile_read: unknown kernel ") + kernel);
    SVPF_HIP(hipStreamSynchronize(ctx->stream));
    double sum = 0;
    int n = 0;
    for (const auto& m : ctx->prof_marks) {
        if (m.kernel != id) continue;
        float ms = 0;
        SVPF_HIP(hipEventElapsedTime(&ms, m.a, m.b));
        sum += ms;
        n++;
    }
    *total_ms = sum;
    *launches = n;
    return SVPF_OK;
}

}  // extern "C"
