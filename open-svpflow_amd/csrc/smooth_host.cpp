// Host side of the SmoothFps render: option parsing (crates/svpflow2/src/options.rs),
// frame timing (crates/svpflow-core/src/frame_math.rs), the per-output-frame decisions
// that do not depend on the vectors, and the render session that drives smooth.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "options.h"
#include "smooth_dev.h"
#include "svpf_internal.h"

namespace svpf {

hipError_t launch_smooth_prep(const PrepJob* d_jobs, int njobs, const SmoothGeom* d_geom, int count, int mc,
                              hipStream_t s);
hipError_t launch_smooth_masks(const FrameJob* d_jobs, FramePlan* d_plans, int njobs, const SmoothGeom* d_geom, int count,
                               int mw, int mh, bool coverage, hipStream_t s);
hipError_t launch_smooth_render(const FrameJob* d_jobs, const FramePlan* d_plans, int njobs, const SmoothGeom* d_geom,
                                int width, int height, int mw, hipStream_t s);
hipError_t launch_smooth_render_float(const FrameJob* d_jobs, const FramePlan* d_plans, int njobs,
                                      const SmoothGeom* d_geom, int width, int height, hipStream_t s);
hipError_t launch_smooth_light(const FrameJob* d_jobs, const FramePlan* d_plans, int njobs, const SmoothGeom* d_geom,
                               const SmoothGeom& g, hipStream_t s);

namespace {

int32_t trunc_i32(double v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)std::trunc(v);
}
int32_t sat32(int64_t v) { return (int32_t)std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX); }
int64_t smul(int64_t a, int64_t b) {
    __int128 r = (__int128)a * b;
    return r > INT64_MAX ? INT64_MAX : (r < INT64_MIN ? INT64_MIN : (int64_t)r);
}

// frame_math.rs Timing
struct Timing {
    int64_t fps_num = 0, fps_den = 1, frame_num = 0, frame_den = 1;
    static void norm(int64_t n, int64_t d, int64_t* on, int64_t* od) {  // :142-151
        if (n == 0 || d == 0) { *on = 0; *od = 1; return; }
        uint64_t a = (uint64_t)(n < 0 ? -n : n), b = (uint64_t)(d < 0 ? -d : d);
        while (b) { uint64_t t = a % b; a = b; b = t; }
        *on = n / (int64_t)a;
        *od = d / (int64_t)a;
    }
    static Timing make(int64_t sn, int64_t sd, bool abs, int64_t rn, int64_t rd) {  // :10-37
        Timing t;
        if (abs) norm(rn, rd, &t.fps_num, &t.fps_den);
        else norm(smul(sn, rn), smul(sd, rd), &t.fps_num, &t.fps_den);
        norm(smul(sd, t.fps_num), smul(sn, t.fps_den), &t.frame_num, &t.frame_den);
        return t;
    }
    int source_frame(int f, bool nearest) const {  // :39-49
        if (frame_num == 0) return f;
        int64_t s = smul(f, frame_den);
        return sat32((nearest ? s + frame_num / 2 : s) / frame_num);
    }
    int phase_256(int f, int sf) const {  // :51-67
        if (frame_num <= 0) return 0;
        __int128 rem = (__int128)f * frame_den - (__int128)sf * frame_num;
        if (rem <= 0) return 0;
        __int128 v = (rem * 512 + frame_num) / ((__int128)frame_num * 2);
        return (int)std::min<__int128>(v, INT32_MAX);
    }
    double raw_phase_256(int f, int sf) const { return ((double)f * (double)frame_den / (double)frame_num - (double)sf) * 256.0; }
    int scene_phase_256(double raw, int64_t mode) const {  // :88-126
        const double step = (double)frame_den * 256.0 / (double)frame_num;
        if (!std::isfinite(raw) || !std::isfinite(step) || step <= 0.0) return 0;
        const double lf = std::floor(raw / step), lr = raw - lf * step;
        int left;
        if (mode > 1) left = trunc_i32(lr) > 0 ? trunc_i32(lf) : std::max(trunc_i32(lf) - 1, 0);
        else left = trunc_i32(lf) + (trunc_i32(lr) >= trunc_i32(std::fabs(lr - step)) ? 1 : 0);
        const double rf = std::floor((256.0 - raw - 0.001) / step), rr = 256.0 - (rf * step + raw);
        const int right = trunc_i32(rf) + (mode > 1 ? (std::fabs(rr - step) < 0.1 ? 1 : 0)
                                                    : (trunc_i32(rr) > trunc_i32(std::fabs(rr - step)) ? 1 : 0));
        const int total = left + right;
        if (total == 0) return 0;
        const int phase = left * 256 / total;
        if ((mode & ~2) == 0) return phase;
        return left > right ? 256 - (right * 256 / total) / 2 : phase / 2;
    }
    int frames(int n) const {  // scale_frame_count, :128-135
        if (n <= 0 || frame_den == 0) return n;
        return sat32(smul(n - 1, frame_num) / frame_den + 1);
    }
};

int cpu_map(int v) {  // renderer.rs:3-6, 3186-3193
    static const int M[32] = {0, 1, 2, 2, -1, 3, 3, 3, -1, -1, -1, 4, -1, 4, -1, 4,
                              -1, -1, -1, -1, -1, -1, -1, -1, 5, -1, -1, -1, 5, -1, -1, 5};
    return (v >= 1 && v <= 32) ? M[v - 1] : -1;
}

void weights(int total, int shift, int32_t out[32][2]) {  // interpolation_weights, renderer.rs:2740-2770
    for (int i = 0; i < 32; i++) out[i][0] = out[i][1] = 0;
    if (total <= 0) return;
    const int scale = 1 << std::max(shift, 0);
    for (int p = 0; p < std::min(total, 32); p++) {
        if (shift < 0) { out[p][0] = total - p; out[p][1] = p; continue; }
        auto coeff = [&](int v) { return trunc_i32(std::ceil((double)(v * scale) / (double)total - 0.5)); };
        out[p][0] = coeff(total - p);
        out[p][1] = coeff(p);
    }
}

struct Opts {  // the options.rs fields the render reads
    bool rate_abs = false, has_num = false, has_den = false;
    int64_t num = 0, den = 0, algo = 21, scene_mode = 3, cover = 100, area = 0;
    bool blend = false, force13 = true, block = false;
    int adaptive[3] = {-1, -1, 1};
    double luma = 1.5, area_sharp = 1.0, area_scale = 1.0;
    int64_t blocks = 20, zero = 200, m1 = 1600, m2 = 2800, scene = 4000;
    Timing timing(int64_t sn, int64_t sd) const {  // options.rs:246-254
        int64_t n = has_num ? num : (rate_abs ? 60 : 2), d = has_den ? den : 1;
        return Timing::make(sn, sd, rate_abs, n == 0 ? 1 : n, d == 0 ? 1 : d);
    }
};

const char* parse_opts(const Opt* op, int64_t fps_num, int64_t fps_den, int vbw, int vbh, Opts* o) {
    int64_t v;
    double f;
    bool b;
    auto gi = [&](std::initializer_list<const char*> k, int64_t d) { return (op && op->get_int(k, &v)) ? v : d; };
    auto gb = [&](std::initializer_list<const char*> k, bool d) { return (op && op->get_bool(k, &b)) ? b : d; };
    auto gf = [&](std::initializer_list<const char*> k, double d) { return (op && op->get_float(k, &f)) ? f : d; };
    o->rate_abs = gb({"rate", "abs"}, false);
    o->has_num = op && op->get_int({"rate", "num"}, &o->num);
    o->has_den = op && op->get_int({"rate", "den"}, &o->den);
    o->scene_mode = gi({"scene", "mode"}, 3);
    o->blend = gb({"scene", "blend"}, false);
    if (o->scene_mode == 3) {
        const int64_t a = gi({"scene", "adaptive"}, 210);
        for (int k = 0, d = 1; k < 3; k++, d *= 10) o->adaptive[k] = (int)std::min<int64_t>((a / d) % 10, 3) - 1;
    }
    o->force13 = gb({"scene", "force13"}, true);
    o->luma = std::min(std::max(gf({"scene", "luma"}, 1.5), 0.0), 5.0);
    const int64_t bl = gi({"scene", "limits", "blocks"}, 20);
    o->blocks = (bl >= 0 && bl < 100) ? bl : 100;
    o->zero = gi({"scene", "limits", "zero"}, 200);
    o->m1 = gi({"scene", "limits", "m1"}, 1600);
    o->m2 = gi({"scene", "limits", "m2"}, 2800);
    o->scene = gi({"scene", "limits", "scene"}, 4000);
    o->area = gi({"mask", "area"}, 0);
    o->cover = std::max<int64_t>(gi({"mask", "cover"}, 100), 0);
    o->area_sharp = gf({"mask", "area_sharp"}, 1.0);
    o->algo = gi({"algo"}, 21);
    o->block = gb({"block"}, false);
    const int64_t a = o->algo;
    if (!(a == 1 || a == 2 || a == 11 || a == 13 || a == 21 || a == 22 || a == 23 || (a >= 90 && a <= 100)))
        return "SVSmoothFps: incorrect 'algo' value";
    if (o->scene_mode < 0 || o->scene_mode > 3) return "SVSmoothFps: 'scene.mode' must be in [0;3]";
    const Timing t = o->timing(fps_num, fps_den);
    auto below = [&](int64_t lim) { return (__int128)t.frame_num < (__int128)lim * t.frame_den; };
    if ((o->scene_mode == 2 && below(2)) || (o->scene_mode == 1 && below(1)))
        return "SVSmoothFps: 'scene.mode' value isn't supported for selected frame rate";
    if (o->scene_mode == 3 && below(1)) {  // normalize_scene_mode
        o->scene_mode = 0;
        o->adaptive[0] = -1; o->adaptive[1] = -1; o->adaptive[2] = 1;
    }
    auto sc = [&](int64_t x) { return (int64_t)trunc_i32((double)x * (double)vbw * (double)vbh / 32.0); };
    o->zero = sc(o->zero); o->m1 = sc(o->m1); o->m2 = sc(o->m2); o->scene = sc(o->scene);
    if (o->area > 0) o->area_scale = (double)sat32(o->area) / 15000.0;
    if (o->cover == 0 && o->algo >= 21 && o->algo <= 31) o->algo = o->algo == 22 ? 13 : 11;
    return nullptr;
}

// Options::padding (options.rs:256-277, pad_x/pad_y :788-796): the border the light.*
// options add around the output, per side.
void light_padding(const Opt* op, int width, int height, int* pad_x, int* pad_y) {
    *pad_x = *pad_y = 0;
    if (width <= 0 || height <= 0) return;
    double f;
    auto gf = [&](std::initializer_list<const char*> k, double d) { return (op && op->get_float(k, &f)) ? f : d; };
    const double w = (double)width, h = (double)height, src_aspect = w / h;
    const double sar = gf({"light", "sar"}, 1.0);
    double aspect = gf({"light", "aspect"}, 0.0) / sar;
    if (aspect < 0.01) aspect = src_aspect;  // (a NaN aspect stays NaN, as in the reference)
    const double zoom = std::max(gf({"light", "zoom"}, 0.0), 0.0);
    int tw = width, th = height;
    if (std::fabs(aspect - src_aspect) <= 0.01) {
    } else if (src_aspect <= aspect) {
        tw = trunc_i32(aspect * h);
    } else {
        th = trunc_i32(w / aspect);
    }
    const int x = (int)((int64_t)trunc_i32((double)tw * zoom / 100.0) + tw - width);
    const int y = (int)((int64_t)trunc_i32((double)th * zoom / 100.0) + th - height);
    const int hx = x / 2, hy = y / 2;
    *pad_x = (hx + 3 * (hx + 1 < 0 ? 1 : 0) + 1) & ~3;
    *pad_y = (hy + (hy + 1 < 0 ? 1 : 0) + 1) & ~1;
}

// LightParams (options.rs:279-286): border 12, lights 16 (at least 1), length 100, cell 1.0
// (at least 0.1).
struct LightSet {
    int border, lights, length;
    double cell;
};
LightSet light_params(const Opt* op) {
    int64_t v;
    double f;
    auto gi = [&](std::initializer_list<const char*> k, int64_t d) { return (op && op->get_int(k, &v)) ? v : d; };
    auto sat = [](int64_t x) { return (int)std::min<int64_t>(std::max<int64_t>(x, INT32_MIN), INT32_MAX); };
    LightSet l;
    l.border = sat(gi({"light", "border"}, 12));
    l.lights = std::max(sat(gi({"light", "lights"}, 16)), 1);
    l.length = sat(gi({"light", "length"}, 100));
    l.cell = std::max((op && op->get_float({"light", "cell"}, &f)) ? f : 1.0, 0.1);
    return l;
}

// Sizes this build draws the light border for (each a deviation only for settings far
// outside the reference's use): frames of 4 to 8192 pixels a side, border <= 255 samples
// and no deeper than the frame, at most 4096 control points a side, cell <= 1000.
const char* light_limits(const Opt* op, int width, int height, int pad_x, int pad_y) {
    const LightSet l = light_params(op);
    if (pad_x < 0 || pad_y < 0) return "SVSmoothFps: light padding would shrink the frame";
    if (width < 4 || height < 4 || width > kLightMaxDim || height > kLightMaxDim)
        return "SVSmoothFps: light border needs a frame of 4 to 8192 pixels a side";
    if (l.border > 255) return "SVSmoothFps: light.border above 255 is not supported";
    if (!(l.cell <= 1000.0)) return "SVSmoothFps: light.cell above 1000 is not supported";
    for (int p = 0; p < 2; p++) {
        const int w = p ? width / 2 : width, h = p ? height / 2 : height;
        const int b = std::max(p ? l.border / 2 : l.border, 1);
        if ((pad_y > 0 && b > h) || (pad_x > 0 && b > w)) return "SVSmoothFps: light.border deeper than the frame";
        if (l.lights > kLightMaxLights || (pad_x > 0 && (int64_t)l.lights * h / w > kLightMaxLights))
            return "SVSmoothFps: light.lights above 4096 a side is not supported";
    }
    return nullptr;
}

// corner_geometry (light.rs:350-373) of a w x h corner as (mix, weight) pairs, built here
// with the host's IEEE double arithmetic and sqrt, as the reference builds it.
std::vector<double> corner_table(int w, int h, int length) {
    std::vector<double> t;
    t.reserve(2 * (size_t)w * h);
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            const double x = (double)(j + 1) / (double)w;
            const double y = (double)(i + 1) / (double)h;
            const double distance = std::sqrt(x * x + y * y) * 100.0;
            t.push_back(x > y ? y / x * 0.5 : (y >= 0.0001 ? 1.0 - x / y * 0.5 : 1.0));
            t.push_back(length > 0 ? std::max(1.0 - distance / (double)length, 0.0) : 0.0);
        }
    return t;
}

// scale_magnitude (renderer.rs:3245-3259) for every 24-bit magnitude, evaluated with this
// host's libm pow in the reference's operation order.  The reference computes the mask on
// the CPU with the same expression, so the table is that computation, not an
// approximation of it; 2^24 entries cover every score a vector payload can carry.
std::vector<uint8_t> magnitude_table(int area, double scale, double expo) {
    const size_t n = size_t(1) << 24;
    std::vector<uint8_t> t(n);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < hw; w++)
        th.emplace_back([&, w] {
            for (size_t m = n * w / hw; m < n * (w + 1) / hw; m++) {
                const double v = std::pow(4.0 * (double)(uint32_t)m * scale / (double)area, expo) * 255.0;
                t[m] = (!std::isfinite(v) || v <= 0.0) ? 0 : (v >= 255.0 ? 255 : (uint8_t)v);
            }
        });
    for (auto& x : th) x.join();
    return t;
}

}  // namespace
}  // namespace svpf

using namespace svpf;

struct svpf_smooth {
    svpf_ctx* ctx;
    Opts opt;
    Timing timing;
    int width, height, nframes, out_frames;
    int64_t sdata;
    int32_t vdata[16];
    SmoothGeom geom;
    SmoothGeom* d_geom = nullptr;
    uint8_t* d_mag_lut = nullptr;
    double* d_corner = nullptr;  // light border corner tables (luma, chroma)
    int count, mw, mh, mc;
    std::vector<svpf_frame> src, sup;
    std::vector<uint8_t*> vec;
    // per source frame prep buffers (one allocation each, carved)
    std::vector<uint8_t*> prep;
    std::vector<char> prepped;
    size_t prep_bytes = 0;
    // per output frame scratch for a batch
    uint8_t* scratch = nullptr;
    size_t scratch_frames = 0, frame_bytes = 0;
};

namespace {

struct PrepLayout {
    size_t vec[2], mv[4], mag[2], stats, total;
};
PrepLayout prep_layout(const svpf_smooth* s) {
    PrepLayout L{};
    size_t off = 0;
    auto take = [&](size_t n) { size_t r = off; off += (n + 255) / 256 * 256; return r; };
    for (int k = 0; k < 2; k++) L.vec[k] = take(4 * (size_t)s->count);
    for (int k = 0; k < 4; k++) L.mv[k] = take(2 * (size_t)s->mc);
    for (int k = 0; k < 2; k++) L.mag[k] = take((size_t)s->mc);
    L.stats = take(64);
    L.total = off;
    return L;
}
struct FrameLayout {
    size_t acc[2], cov[2], maxmag, prof, draws, total;
};
FrameLayout frame_layout(const svpf_smooth* s) {
    FrameLayout L{};
    size_t off = 0;
    auto take = [&](size_t n) { size_t r = off; off += (n + 255) / 256 * 256; return r; };
    for (int k = 0; k < 2; k++) L.acc[k] = take(4 * (size_t)(s->mw + 2) * (s->mh + 2));
    for (int k = 0; k < 2; k++) L.cov[k] = take((size_t)s->mc);
    L.maxmag = take((size_t)s->mc);
    L.prof = take(4 * (size_t)s->geom.light_prof_len);
    L.draws = take(4 * (size_t)s->geom.light_draws);
    L.total = off;
    return L;
}

int ensure_prep(svpf_smooth* s, const std::vector<int>& frames) {
    std::vector<PrepJob> jobs;
    const PrepLayout L = prep_layout(s);
    for (int f : frames) {
        if (s->prepped[(size_t)f]) continue;
        if (!s->vec[(size_t)f]) return fail(SVPF_ERR_STATE, "svpf_smooth: vector payload of a source frame is not set");
        if (!s->prep[(size_t)f]) {
            hipError_t e = hipMalloc(&s->prep[(size_t)f], L.total);
            if (e != hipSuccess) return fail(SVPF_ERR_NOMEM, "svpf_smooth: prep allocation failed");
        }
        uint8_t* b = s->prep[(size_t)f];
        PrepJob j{};
        j.payload = s->vec[(size_t)f];
        for (int k = 0; k < 2; k++) j.vec[k] = (int16_t*)(b + L.vec[k]);
        for (int k = 0; k < 4; k++) j.mv[k] = (uint16_t*)(b + L.mv[k]);
        for (int k = 0; k < 2; k++) j.mag[k] = b + L.mag[k];
        j.stats = (int32_t*)(b + L.stats);
        jobs.push_back(j);
        s->prepped[(size_t)f] = 1;
    }
    if (jobs.empty()) return SVPF_OK;
    void* d = nullptr;
    int rc = svpf_upload_desc(s->ctx, jobs.data(), jobs.size() * sizeof(PrepJob), &d);
    if (rc) return rc;
    const int ph = prof_begin(s->ctx, kProfPrep);
    SVPF_HIP(launch_smooth_prep((const PrepJob*)d, (int)jobs.size(), s->d_geom, s->count, s->mc, s->ctx->stream));
    prof_end(s->ctx, ph);
    return SVPF_OK;
}

}  // namespace

extern "C" {

int svpf_smooth_create(svpf_ctx* ctx, const char* opt, size_t opt_len, int32_t width, int32_t height, int64_t fps_num,
                       int64_t fps_den, int32_t nframes, int64_t sdata, const int32_t* vdata16, svpf_smooth** out,
                       int32_t* out_frames, char* err, size_t err_len) {
    auto reject = [&](const char* msg) {
        if (err && err_len) {
            std::strncpy(err, msg, err_len - 1);
            err[err_len - 1] = 0;
        }
        return fail(SVPF_ERR_OPT, msg);
    };
    // ctx == NULL (with out == NULL): validate only (a plugin's create step)
    if ((!ctx) != (!out) || !vdata16 || width < 8 || height < 8 || (width & 1) || (height & 1) || nframes < 1)
        return fail(SVPF_ERR_ARG, "svpf_smooth_create: bad argument");
    // create-time checks in the order of filter.rs:60-130
    if (fps_num <= 0 || fps_den <= 0) return reject("SVSmoothFps: unable to determine source frame rate");
    if (vdata16[0] != 0xA0) return reject("SVSmoothFps: invalid vectors stream [2]");  // validate_vectors, :578-633
    if (vdata16[7] != width || vdata16[8] != height)
        return reject("SVSmoothFps: source and vectors frame sizes are different");
    Opt o;
    bool have = opt && opt_len && Opt::parse(opt, opt_len, &o, nullptr);
    auto* s = new svpf_smooth{};
    s->ctx = ctx;
    std::memcpy(s->vdata, vdata16, 64);
    const int vbw = vdata16[2], vbh = vdata16[3];
    if (const char* e = parse_opts(have ? &o : nullptr, fps_num, fps_den, vbw, vbh, &s->opt)) {
        delete s;
        return reject(e);
    }
    // validate_super (:646-669, metadata.rs:116-137)
    uint64_t raw;
    std::memcpy(&raw, &sdata, 8);
    const int s_marker = (int)((raw >> 32) & 0xFF), s_w = (int)((raw >> 16) & 0xFFFF), s_h = (int)(raw & 0xFFFF);
    const int s_sel = (int)((raw >> 48) & 0xFF), s_lim = (int)((raw >> 56) & 0xFF);
    if (!(s_marker == vdata16[4] && s_w == vdata16[7] && s_h == vdata16[8] &&
          (s_lim >= 4 || s_sel >= 3 || s_sel == vdata16[13]))) {
        delete s;
        return reject("SVSmoothFps: invalid super clip params");
    }
    if (vdata16[14] != 1) {
        delete s;
        return reject("SVSmoothFps: vectors with delta>1 are not allowed.");
    }
    int64_t cubic = 0;
    if (have && o.get_int({"cubic"}, &cubic) && cubic > 0 && vdata16[13] == 0) {  // validate_cubic, :681-694
        delete s;
        return reject("SVSmoothFps: 'cubic' mode isn't available with CPU rendering");
    }
    int pad_x, pad_y;
    light_padding(have ? &o : nullptr, width, height, &pad_x, &pad_y);
    if (pad_x != 0 || pad_y != 0) {
        // the output is widened by the padding and the light border drawn there (core.rs
        // :603-613, 1866-1897, light.rs); the limits keep every profile inside the frame
        // and its sums in 32 bits
        if (const char* e = light_limits(have ? &o : nullptr, width, height, pad_x, pad_y)) {
            delete s;
            return reject(e);
        }
    }
    s->timing = s->opt.timing(fps_num, fps_den);
    s->width = width;
    s->height = height;
    s->nframes = nframes;
    s->out_frames = s->timing.frames(nframes);
    s->sdata = sdata;
    if (!ctx) {
        if (out_frames) *out_frames = s->out_frames;
        delete s;
        return SVPF_OK;
    }
    SmoothGeom& g = s->geom;
    std::memset(&g, 0, sizeof g);
    g.width = width;
    g.height = height;
    g.vbw = vbw;
    g.vbh = vbh;
    g.vox = vdata16[9];
    g.voy = vdata16[10];
    g.gw = vdata16[11];
    g.gh = vdata16[12];
    g.marker = vdata16[4];
    g.flags = vdata16[1];
    g.count = std::max(g.gw, 0) * std::max(g.gh, 0);
    // bytes of the Analyse payload: header + one region per direction (pack layout of
    // svpflow1 Analyse); a region the flags name but the payload lacks fails decoding.
    g.payload_len = 0x40 + (((g.flags & 1) ? 1 : 0) + ((g.flags & 2) ? 1 : 0)) * (4 + 8 * g.count);
    auto ext = [](int block, int overlap, int grid, int source) {  // extended_grid, metadata.rs:238-241
        return grid + (((block - overlap) * grid + overlap) < source ? 1 : 0);
    };
    g.mw = ext(vbw, g.vox, g.gw, vdata16[7]);
    g.mh = ext(vbh, g.voy, g.gh, vdata16[8]);
    if (g.gw <= 0 || g.gh <= 0 || vbw <= 0 || vbh <= 0 || vbw * vbh > 4096) {
        delete s;
        return reject("SVSmoothFps: invalid vectors stream [1]");
    }
    g.cbw = std::max(vbw - g.vox, 1);
    g.cbh = std::max(vbh - g.voy, 1);
    g.cox = vbw / 2;
    g.coy = vbh / 2;
    // render_mode / request_super (filter.rs:300-313, 680-688, 728-751)
    const double ratio = (double)s->timing.frame_num / (double)s->timing.frame_den;
    const int render_mode = std::fabs(ratio - 1.0) < 0.001 ? 0 : (vdata16[13] < 1 ? 1 : 2);
    const bool request_super = (render_mode == 1 || render_mode == 2) && g.marker != 1;
    g.use_super = request_super && (s_marker == 1 || s_marker == 2 || s_marker == 4);
    g.super_shift = s_marker == 4 ? 2 : (s_marker == 2 ? 1 : 0);
    g.source_step = g.use_super ? std::max(s_marker, 1) : 1;
    g.super_rows = super_plane_height(height, std::max(s_marker, 1), (int)((raw >> 40) & 0xFF), !(s_lim & 2));
    if (render_mode == 1 && ((g.vox & 1) || (g.voy & 1))) {  // validate_overlap, filter.rs:696-709
        delete s;
        return reject("SVSmoothFps: overlap must be even with CPU rendering");
    }
    g.xshift[0] = cpu_map(g.cbw);
    g.yshift[0] = cpu_map(g.cbh);
    g.xshift[1] = cpu_map(g.cbw / 2);
    g.yshift[1] = cpu_map(g.cbh / 2);
    weights(g.cbw, g.xshift[0], g.xw[0]);
    weights(g.cbh, g.yshift[0], g.yw[0]);
    weights(g.cbw / 2, g.xshift[1], g.xw[1]);
    weights(g.cbh / 2, g.yshift[1], g.yw[1]);
    g.interp = !s->opt.block;
    for (int c = 0; c < 2; c++) {  // render_tiles (renderer.rs:1925-2010) per plane kind, as render_kernel reads them
        const int pw = c ? width / 2 : width, ph = c ? height / 2 : height;
        const int bw = c ? g.cbw / 2 : g.cbw, bh = c ? g.cbh / 2 : g.cbh;
        const int ox = g.interp ? (c ? g.cox / 2 : g.cox) : 0, oy = g.interp ? (c ? g.coy / 2 : g.coy) : 0;
        g.tile_last_x[c] = (bw > 0 && ox < pw) ? std::min(g.mw - 1, std::max((pw - ox + bw - 1) / bw - 1, 0)) : -1;
        g.tile_last_y[c] = (bh > 0 && oy < ph) ? std::min(g.mh - 1, std::max((ph - oy + bh - 1) / bh - 1, 0)) : -1;
        g.rcp_bh[c] = bh > 1 ? (uint32_t)((((uint64_t)1 << 32) + (uint64_t)bh - 1) / (uint64_t)bh) : 0u;
    }
    g.algo_req = (int)s->opt.algo;
    g.force13 = s->opt.force13;
    g.mask_area = s->opt.area > 0;
    g.mask_cover = (int)std::min<int64_t>(s->opt.cover, INT32_MAX);
    g.area_scale = s->opt.area_scale;
    g.area_sharp = s->opt.area_sharp;
    g.lim_zero = sat32(s->opt.zero);
    g.lim_m1 = sat32(s->opt.m1);
    g.lim_m2 = sat32(s->opt.m2);
    g.lim_scene = sat32(s->opt.scene);
    g.lim_blocks = sat32(s->opt.blocks);
    g.nframes = nframes;
    const double denom = g.marker == 3 ? 510.0 : 255.0;
    for (int i = 0; i <= 510; i++) {  // luma_byte (metadata.rs:523-526) for every possible (a + b)
        const int sc = trunc_i32(std::pow((double)i / denom, s->opt.luma) * 255.0);
        g.luma_lut[i] = (uint8_t)(std::max(sc, 20) & 0xFF);
    }
    for (int t = 0; t < 256; t++) {  // linear_luma (gpu.rs): the float renderer's luma image
        const float v = (float)t / 255.0f;
        g.lin_lut[t] = v < 0.081f ? v / 4.5f : std::pow((v + 0.099f) / 1.099f, 1.0f / 0.45f);
    }
    const int area = std::max(vbw * vbh, 1);
    const double strength = (double)g.mask_cover / 100.0;
    for (int r = 0; r <= area && r <= 4096; r++) {  // coverage_byte LUT (renderer.rs:3112-3117, 3262-3272)
        const double vv = (double)r * strength * 256.0 / (double)area;
        int b = 0;
        if (std::isfinite(vv)) {
            const int iv = trunc_i32(vv);
            b = iv >= 255 ? 255 : (iv & 0xFF);
        }
        g.cover_lut[r] = (uint8_t)b;
    }
    // light border geometry (light.rs:54-235): per plane two profiles per padded side,
    // one random draw per strip row and per corner row after 15 skipped (even frames) and
    // the six noise tables
    g.pad_x = pad_x;
    g.pad_y = pad_y;
    std::vector<double> corners[2];
    if (pad_x || pad_y) {
        const LightSet l = light_params(have ? &o : nullptr);
        g.light_border = l.border;
        g.light_lights = l.lights;
        g.light_length = l.length;
        g.light_cell = l.cell;
        g.light_draws = 15 + 6 * 128;
        for (int p = 0; p < 3; p++) {
            const int w = p ? width / 2 : width, h = p ? height / 2 : height;
            const int px = p ? pad_x / 2 : pad_x, py = p ? pad_y / 2 : pad_y;
            if (py > 0) { g.light_slots += 2; g.light_prof_len += 2 * w; g.light_draws += 2 * py; }
            if (px > 0) { g.light_slots += 2; g.light_prof_len += 2 * h; g.light_draws += 2 * px; }
            if (px > 0 && py > 0) g.light_draws += 4 * py;
            g.light_pixels += (w + 2 * px) * (h + 2 * py) - w * h;
        }
        if (pad_x > 0 && pad_y > 0) {
            corners[0] = corner_table(pad_x, pad_y, l.length);
            corners[1] = corner_table(pad_x / 2, pad_y / 2, l.length);
        }
    }
    s->count = g.count;
    s->mw = g.mw;
    s->mh = g.mh;
    s->mc = g.mw * g.mh;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess && !corners[0].empty()) {
        const size_t n0 = corners[0].size() * sizeof(double), n1 = corners[1].size() * sizeof(double);
        e = hipMalloc(&s->d_corner, n0 + n1);
        if (e == hipSuccess) e = hipMemcpy(s->d_corner, corners[0].data(), n0, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(s->d_corner + corners[0].size(), corners[1].data(), n1, hipMemcpyHostToDevice);
        g.corner_geo[0] = s->d_corner;
        g.corner_geo[1] = s->d_corner + corners[0].size();
    }
    if (e == hipSuccess && g.mask_area && g.area_sharp != 1.0) {
        const std::vector<uint8_t> t = magnitude_table(area, g.area_scale, g.area_sharp);
        e = hipMalloc(&s->d_mag_lut, t.size());
        if (e == hipSuccess) e = hipMemcpy(s->d_mag_lut, t.data(), t.size(), hipMemcpyHostToDevice);
        g.mag_lut = s->d_mag_lut;
    }
    if (e == hipSuccess) e = hipMalloc(&s->d_geom, sizeof(SmoothGeom));
    if (e == hipSuccess) e = hipMemcpy(s->d_geom, &g, sizeof g, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (s->d_mag_lut) (void)hipFree(s->d_mag_lut);
        if (s->d_corner) (void)hipFree(s->d_corner);
        if (s->d_geom) (void)hipFree(s->d_geom);
        delete s;
        return fail(SVPF_ERR_HIP, std::string("svpf_smooth_create: ") + hipGetErrorString(e));
    }
    s->src.assign((size_t)nframes, svpf_frame{});
    s->sup.assign((size_t)nframes, svpf_frame{});
    s->vec.assign((size_t)nframes, nullptr);
    s->prep.assign((size_t)nframes, nullptr);
    s->prepped.assign((size_t)nframes, 0);
    *out = s;
    if (out_frames) *out_frames = s->out_frames;
    return SVPF_OK;
}

void svpf_smooth_destroy(svpf_smooth* s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto* p : s->prep)
        if (p) (void)hipFree(p);
    if (s->scratch) (void)hipFree(s->scratch);
    if (s->d_geom) (void)hipFree(s->d_geom);
    if (s->d_mag_lut) (void)hipFree(s->d_mag_lut);
    if (s->d_corner) (void)hipFree(s->d_corner);
    delete s;
}

int svpf_smooth_set_sources(svpf_smooth* s, int32_t first, int32_t count, const svpf_frame* d_src,
                            const svpf_frame* d_super, uint8_t* const* d_vectors) {
    if (!s || first < 0 || count < 0 || first + count > s->nframes || !d_src || !d_vectors)
        return fail(SVPF_ERR_ARG, "svpf_smooth_set_sources: bad argument");
    for (int i = 0; i < count; i++) {
        const size_t f = (size_t)(first + i);
        s->src[f] = d_src[i];
        if (d_super) s->sup[f] = d_super[i];
        if (s->vec[f] != d_vectors[i]) s->prepped[f] = 0;
        s->vec[f] = d_vectors[i];
        s->prepped[f] = 0;  // contents may have changed
    }
    return SVPF_OK;
}

int svpf_smooth_render(svpf_smooth* s, int32_t k0, int32_t count, const svpf_frame* d_out, int32_t* h_decision) {
    if (!s || k0 < 0 || count < 0 || k0 + count > s->out_frames || (count && !d_out))
        return fail(SVPF_ERR_ARG, "svpf_smooth_render: bad argument");
    if (count == 0) return SVPF_OK;
    const SmoothGeom& g = s->geom;
    const Opts& o = s->opt;
    const Timing& t = s->timing;
    const int64_t sm = o.scene_mode;
    const double ratio = (double)t.frame_num / (double)t.frame_den;
    const int N = s->nframes;
    // host part of the decisions (core.rs:2054-2077, 765-800, 2132-2147)
    std::vector<FramePlan> plans((size_t)count);
    std::vector<int> need;
    bool coverage = false;
    for (int i = 0; i < count; i++) {
        const int k = k0 + i;
        FramePlan& p = plans[(size_t)i];
        std::memset(&p, 0, sizeof p);
        const bool nearest = k > 0 && !o.blend && sm != 0 && sm != 2;
        const int n = t.source_frame(k, nearest);
        int phase = 0;
        if (t.frame_num > 0) {
            if (sm == 1 || sm == 2) phase = t.scene_phase_256(t.raw_phase_256(k, n), sm == 1 ? 0 : 2);
            else phase = t.phase_256(k, n);
        }
        p.k = k;
        p.n = std::min(std::max(n, 0), N - 1);
        p.n1 = std::min(std::max(n + 1, 0), N - 1);
        p.raw_phase = phase;
        const int a = (int)o.algo;
        const bool algo_ok = a == 1 || a == 2 || a == 11 || a == 13 || a == 21 || a == 22 || a == 23;
        p.early = (k == 0 || phase == 0 || phase == 256 || (!o.blend && sm == 3 && ratio < 2.0) || !algo_ok) ? 1 : 0;
        const int cl = std::min(std::max(phase, 0), 256);
        for (int c = 0; c < 3; c++) {
            int ph = cl;
            if (sm == 3) {
                if (ratio < 2.0) ph = 0;
                else if (o.adaptive[c] >= 0) ph = t.scene_phase_256(t.raw_phase_256(k, n), o.adaptive[c]);
            }
            p.phase_for[c] = ph;
        }
        if (!p.early) {
            need.push_back(p.n);
            if (a == 23) {  // vector_neighbors reads n-1 and n+1 (core.rs:2456-2506)
                need.push_back(std::min(std::max(p.n - 1, 0), N - 1));
                need.push_back(p.n1);
            }
            if (a >= 21) coverage = true;
        }
    }
    int rc = ensure_prep(s, need);
    if (rc) return rc;
    const FrameLayout FL = frame_layout(s);
    if ((size_t)count > s->scratch_frames) {
        SVPF_HIP(hipStreamSynchronize(s->ctx->stream));
        if (s->scratch) (void)hipFree(s->scratch);
        s->scratch = nullptr;
        s->scratch_frames = 0;
        hipError_t e = hipMalloc(&s->scratch, FL.total * (size_t)count);
        if (e != hipSuccess) return fail(SVPF_ERR_NOMEM, "svpf_smooth_render: scratch allocation failed");
        s->scratch_frames = (size_t)count;
    }
    const PrepLayout PL = prep_layout(s);
    // Frames that are copies whatever the vectors say never read the prep buffers the
    // plan kernel would otherwise look at; they point at any prepped frame.
    uint8_t* any = nullptr;
    for (size_t f = 0; f < s->prep.size() && !any; f++)
        if (s->prepped[f]) any = s->prep[f];
    if (!any) {
        std::vector<int> one{plans[0].n};
        if (!s->vec[(size_t)plans[0].n]) return fail(SVPF_ERR_STATE, "svpf_smooth_render: vector payload not set");
        rc = ensure_prep(s, one);
        if (rc) return rc;
        any = s->prep[(size_t)plans[0].n];
    }
    auto prep_of = [&](int f) -> uint8_t* { return s->prepped[(size_t)f] ? s->prep[(size_t)f] : any; };
    std::vector<uint8_t> blob((size_t)count * (sizeof(FrameJob) + sizeof(FramePlan)));
    FrameJob* jobs = (FrameJob*)blob.data();
    std::memcpy(blob.data() + (size_t)count * sizeof(FrameJob), plans.data(), (size_t)count * sizeof(FramePlan));
    for (int i = 0; i < count; i++) {
        const FramePlan& p = plans[(size_t)i];
        uint8_t* fs = s->scratch + FL.total * (size_t)i;
        FrameJob& j = jobs[i];
        std::memset(&j, 0, sizeof j);
        const int n = p.n, nn = p.n1, np = std::min(std::max(p.n - 1, 0), N - 1);
        uint8_t* pn = prep_of(n);
        uint8_t* pp = prep_of(np);
        uint8_t* px = prep_of(nn);
        j.stats_n = (const int32_t*)(pn + PL.stats);
        j.stats_prev = (const int32_t*)(pp + PL.stats);
        j.stats_next = (const int32_t*)(px + PL.stats);
        for (int k = 0; k < 4; k++) j.mv_n[k] = (const uint16_t*)(pn + PL.mv[k]);
        j.mv2[0] = (const uint16_t*)(px + PL.mv[0]);
        j.mv2[1] = (const uint16_t*)(px + PL.mv[1]);
        j.mv3[0] = (const uint16_t*)(pp + PL.mv[2]);
        j.mv3[1] = (const uint16_t*)(pp + PL.mv[3]);
        j.mag[0] = pn + PL.mag[0];
        j.mag[1] = pn + PL.mag[1];
        for (int k = 0; k < 2; k++) {
            j.cov_acc[k] = (int32_t*)(fs + FL.acc[k]);
            j.cov[k] = fs + FL.cov[k];
        }
        j.maxmag = fs + FL.maxmag;
        j.vec_a = (const int16_t*)(pn + PL.vec[0]);
        j.vec_b = (const int16_t*)(pn + PL.vec[1]);
        const svpf_frame& f0 = s->src[(size_t)n];
        const svpf_frame& f1 = s->src[(size_t)nn];
        if (!f0.y || !f1.y) return fail(SVPF_ERR_STATE, "svpf_smooth_render: source frame not set");
        if (f1.stride_y != f0.stride_y || f1.stride_u != f0.stride_u || f1.stride_v != f0.stride_v)
            return fail(SVPF_ERR_ARG, "svpf_smooth_render: source frames must share strides");
        j.src0[0] = f0.y; j.src0[1] = f0.u; j.src0[2] = f0.v;
        j.src1[0] = f1.y; j.src1[1] = f1.u; j.src1[2] = f1.v;
        j.sstride[0] = f0.stride_y; j.sstride[1] = f0.stride_u; j.sstride[2] = f0.stride_v;
        if (g.use_super) {
            const svpf_frame& u0 = s->sup[(size_t)n];
            const svpf_frame& u1 = s->sup[(size_t)nn];
            if (!u0.y || !u1.y) return fail(SVPF_ERR_STATE, "svpf_smooth_render: super frame not set");
            if (u1.stride_y != u0.stride_y || u1.stride_u != u0.stride_u || u1.stride_v != u0.stride_v)
                return fail(SVPF_ERR_ARG, "svpf_smooth_render: super frames must share strides");
            j.sup0[0] = u0.y; j.sup0[1] = u0.u; j.sup0[2] = u0.v;
            j.sup1[0] = u1.y; j.sup1[1] = u1.u; j.sup1[2] = u1.v;
            j.supstride[0] = u0.stride_y; j.supstride[1] = u0.stride_u; j.supstride[2] = u0.stride_v;
        }
        const svpf_frame& d = d_out[i];
        if (!d.y || !d.u || !d.v) return fail(SVPF_ERR_ARG, "svpf_smooth_render: output frame not set");
        const int pad_x = g.pad_x, pad_y = g.pad_y;
        if (d.stride_y < s->width + 2 * pad_x || d.stride_u < s->width / 2 + pad_x ||
            d.stride_v < s->width / 2 + pad_x)
            return fail(SVPF_ERR_ARG, "svpf_smooth_render: output stride below the output width");
        // the frame goes at (pad_x, pad_y) of a padded output (chroma at half), the light
        // border around it
        j.obase[0] = d.y; j.obase[1] = d.u; j.obase[2] = d.v;
        j.dstride[0] = d.stride_y; j.dstride[1] = d.stride_u; j.dstride[2] = d.stride_v;
        j.dst[0] = d.y + (size_t)pad_y * d.stride_y + pad_x;
        j.dst[1] = d.u + (size_t)(pad_y / 2) * d.stride_u + pad_x / 2;
        j.dst[2] = d.v + (size_t)(pad_y / 2) * d.stride_v + pad_x / 2;
        j.prof = (float*)(fs + FL.prof);
        j.draws = (int32_t*)(fs + FL.draws);
    }
    void* d_up = nullptr;
    rc = svpf_upload_desc(s->ctx, blob.data(), blob.size(), &d_up);
    if (rc) return rc;
    FramePlan* d_plans = (FramePlan*)((uint8_t*)d_up + (size_t)count * sizeof(FrameJob));
    int ph = prof_begin(s->ctx, kProfPrep);
    SVPF_HIP(launch_smooth_masks((const FrameJob*)d_up, d_plans, count, s->d_geom, s->count, s->mw, s->mh, coverage,
                                 s->ctx->stream));
    prof_end(s->ctx, ph);
    ph = prof_begin(s->ctx, kProfRender);
    if (s->geom.float_render)
        SVPF_HIP(launch_smooth_render_float((const FrameJob*)d_up, d_plans, count, s->d_geom, s->width, s->height,
                                            s->ctx->stream));
    else
        SVPF_HIP(launch_smooth_render((const FrameJob*)d_up, d_plans, count, s->d_geom, s->width, s->height, s->mw,
                                      s->ctx->stream));
    SVPF_HIP(launch_smooth_light((const FrameJob*)d_up, d_plans, count, s->d_geom, g, s->ctx->stream));
    prof_end(s->ctx, ph);
    if (h_decision) {  // (decision, algo) per frame; synchronizes
        SVPF_HIP(hipMemcpyAsync(plans.data(), d_plans, (size_t)count * sizeof(FramePlan), hipMemcpyDeviceToHost,
                                s->ctx->stream));
        SVPF_HIP(hipStreamSynchronize(s->ctx->stream));
        for (int i = 0; i < count; i++) {
            h_decision[2 * i] = plans[(size_t)i].decision;
            h_decision[2 * i + 1] = plans[(size_t)i].algo;
        }
    }
    return SVPF_OK;
}

int svpf_smooth_set_float_render(svpf_smooth* s, int32_t on) {
    if (!s) return fail(SVPF_ERR_ARG, "svpf_smooth_set_float_render: null session");
    // the reference renders on its GPU only with interpolated vectors and without scene
    // blending (core.rs:1429); other sessions keep the integer renderer
    const int32_t v = (on && s->geom.interp && !s->opt.blend) ? 1 : 0;
    if (v == s->geom.float_render) return SVPF_OK;
    SVPF_HIP(hipStreamSynchronize(s->ctx->stream));  // queued renders read the geometry
    s->geom.float_render = v;
    SVPF_HIP(hipMemcpy(&s->d_geom->float_render, &v, sizeof v, hipMemcpyHostToDevice));
    return SVPF_OK;
}

int32_t svpf_smooth_output_frames(const svpf_smooth* s) { return s ? s->out_frames : 0; }

int svpf_smooth_output_size(const svpf_smooth* s, int32_t* width, int32_t* height) {
    if (!s || !width || !height) return fail(SVPF_ERR_ARG, "svpf_smooth_output_size: bad argument");
    *width = s->width + 2 * s->geom.pad_x;  // output_info (core.rs:603-613)
    *height = s->height + 2 * s->geom.pad_y;
    return SVPF_OK;
}

int svpf_smooth_output_rate(const svpf_smooth* s, int64_t* fps_num, int64_t* fps_den) {
    if (!s || !fps_num || !fps_den) return fail(SVPF_ERR_ARG, "svpf_smooth_output_rate: bad argument");
    *fps_num = s->timing.fps_num;  // output_info (core.rs:603-613)
    *fps_den = s->timing.fps_den;
    return SVPF_OK;
}

int svpf_smooth_frame_info(const svpf_smooth* s, int32_t k, svpf_smooth_frame* out) {
    if (!s || !out || k < 0 || k >= s->out_frames) return fail(SVPF_ERR_ARG, "svpf_smooth_frame_info: bad argument");
    const Opts& o = s->opt;
    const Timing& t = s->timing;
    const int N = s->nframes;
    // FilterState::source_frame (core.rs:2054-2077): nearest except scene modes 0 and 2
    const bool nearest = k > 0 && !o.blend && o.scene_mode != 0 && o.scene_mode != 2;
    const int n = t.source_frame(k, nearest);
    out->source = std::min(std::max(n, 0), N - 1);
    // what svpf_smooth_render reads for frame k: sources and Super frames n, n + 1 and
    // the vectors of n - 1 .. n + 1 (algorithm 23's neighbours)
    out->first = std::max(out->source - 1, 0);
    out->last = std::min(out->source + 1, N - 1);
    out->raw_phase = t.frame_num > 0 ? t.raw_phase_256(k, n) : 0.0;  // core.rs:539-542
    out->ratio = t.frame_den != 0 ? (double)t.frame_num / (double)t.frame_den : 0.0;
    return SVPF_OK;
}

int svpf_smooth_release(svpf_smooth* s, int32_t first, int32_t count) {
    if (!s || first < 0 || count < 0 || first + count > s->nframes)
        return fail(SVPF_ERR_ARG, "svpf_smooth_release: bad argument");
    bool any = false;
    for (int i = 0; i < count; i++) any = any || s->prep[(size_t)(first + i)] != nullptr;
    if (any) SVPF_HIP(hipStreamSynchronize(s->ctx->stream));  // queued work may still read them
    for (int i = 0; i < count; i++) {
        const size_t f = (size_t)(first + i);
        if (s->prep[f]) (void)hipFree(s->prep[f]);
        s->prep[f] = nullptr;
        s->prepped[f] = 0;
        s->vec[f] = nullptr;
        s->src[f] = svpf_frame{};
        s->sup[f] = svpf_frame{};
    }
    return SVPF_OK;
}

}  // extern "C"
