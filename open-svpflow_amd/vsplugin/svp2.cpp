// svp2 plugin shell (VapourSynth API 3): svpflow2's SmoothFps over libsvpf.  It
// registers what the reference registers for SmoothFps (crates/svpflow2/src/abi.rs:
// 17-49, strings.rs:3-17), checks the arguments in the order of create_filter
// (filter.rs:44-139, the option checks inside svpf_smooth_create), requests the
// frames an output frame reads (core.rs:150-201) and writes the output's duration and
// PTS (frame.rs:111-138).  The render runs on the GPU: the filter holds one context
// and one render session, and keeps the last few source frames (planes, Super frame
// and vector payload) on the device.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"

using namespace vs3;
using svpvs::DevFrame;

namespace {

constexpr int kSlots = 6;  // source frames kept on the device (an output frame reads at most 3)

struct Timing {
    int64_t num = 0, den = 1, pts = 0;
};

// _DurationNum / _DurationDen / _PTS of a frame (read_timing, frame.rs:275-300).
Timing read_timing(const Table& t, const FrameRef* f) {
    Timing r;
    const auto props_of = t.get_frame_props_ro();
    const auto get_int = t.prop_get_int();
    const Map* m = (props_of && get_int && f) ? props_of(f) : nullptr;
    if (!m) return r;
    int e1 = 0, e2 = 0, e3 = 0;
    const int64_t num = get_int(m, "_DurationNum", 0, &e1), den = get_int(m, "_DurationDen", 0, &e2);
    if (!e1 && !e2 && num >= 0 && den >= 0) {
        r.num = num;
        r.den = den;
    }
    const int64_t pts = get_int(m, "_PTS", 0, &e3);
    if (!e3) r.pts = pts;
    return r;
}

int64_t trunc_i64(double v) {
    if (std::isnan(v)) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

double seconds(int64_t num, int64_t den) { return den == 0 ? 0.0 : (double)num / (double)den; }

// copy_interpolated_timing (frame.rs:111-138): the output's share of the source
// durations at its phase, and its PTS moved by the phase.
void write_timing(const Table& t, const FrameRef* src, const FrameRef* next, FrameRef* out, double raw_phase,
                  double ratio) {
    Timing tm = read_timing(t, src);
    if (ratio > 0.0 && std::isfinite(ratio) && std::isfinite(raw_phase)) {
        const double cur = seconds(tm.num, tm.den);
        const Timing nt = read_timing(t, next);
        const double nxt = seconds(nt.num, nt.den);
        const double end = 256.0 / ratio + raw_phase;
        const double dur = end < 256.0 ? cur / ratio : ((256.0 - raw_phase) * cur + (end - 256.0) * nxt) / 256.0;
        tm.num = trunc_i64(dur * 10000000.0);
        tm.den = 10000000;
        if (tm.pts != 0) tm.pts = trunc_i64((double)tm.pts + cur * 1000000.0 * raw_phase / 256.0);
    }
    const auto props_of = t.get_frame_props_rw();
    const auto set_int = t.prop_set_int();
    Map* m = (props_of && set_int) ? props_of(out) : nullptr;
    if (!m) return;
    if (tm.num >= 0 && tm.den >= 0) {
        set_int(m, "_DurationNum", tm.num, kPaReplace);
        set_int(m, "_DurationDen", tm.den, kPaReplace);
    }
    set_int(m, "_PTS", tm.pts, kPaReplace);
}

struct Slot {
    int frame = -1;
    uint64_t used = 0;
    DevFrame src, sup;
    uint8_t* payload = nullptr;
};

struct SmoothFilter {
    Node* clip = nullptr;
    Node* super_clip = nullptr;
    Node* vectors = nullptr;
    Node* src = nullptr;
    VideoInfo vi{};      // output
    int src_frames = 0;  // source clip length
    int super_rows = 0;
    int32_t payload_len = 0;
    std::mutex mu;
    svpf_ctx* ctx = nullptr;
    svpf_smooth* session = nullptr;
    Slot slot[kSlots];
    uint64_t tick = 0;
    DevFrame out;

    // Source frame `idx` (its planes, Super frame and payload) bound to the session.
    bool bind(const Table& t, int idx, const FrameRef* s, const FrameRef* p, const FrameRef* v, int lo, int hi) {
        for (Slot& sl : slot)
            if (sl.frame == idx) {
                sl.used = ++tick;
                return true;
            }
        Slot* victim = nullptr;  // least recently used slot outside [lo, hi]
        for (Slot& sl : slot)
            if (sl.frame < lo || sl.frame > hi)
                if (!victim || sl.used < victim->used) victim = &sl;
        if (!victim) return false;
        if (victim->frame >= 0) svpf_smooth_release(session, victim->frame, 1);
        victim->frame = -1;
        if (!victim->payload) {
            void* d = nullptr;
            if (svpf_device_alloc(ctx, ((size_t)payload_len + 3) & ~(size_t)3, &d) != SVPF_OK) return false;
            victim->payload = (uint8_t*)d;
        }
        const uint8_t* row = t.get_read_ptr()(v, 0);
        if (!row || !svpvs::upload(ctx, victim->src, t, s, vi.height) ||
            !svpvs::upload(ctx, victim->sup, t, p, super_rows) ||
            svpf_copy_h2d(ctx, victim->payload, row, (size_t)payload_len) != SVPF_OK)
            return false;
        const svpf_frame fs = victim->src.frame(), fp = victim->sup.frame();
        uint8_t* pv = victim->payload;
        if (svpf_smooth_set_sources(session, idx, 1, &fs, &fp, &pv) != SVPF_OK) return false;
        victim->frame = idx;
        victim->used = ++tick;
        return true;
    }
    ~SmoothFilter() {
        if (session) svpf_smooth_destroy(session);
        if (ctx) {
            for (Slot& sl : slot) {
                sl.src.release(ctx);
                sl.sup.release(ctx);
                if (sl.payload) svpf_device_free(ctx, sl.payload);
            }
            out.release(ctx);
            svpf_destroy(ctx);
        }
    }
};

void smooth_init(Map*, Map*, void** instance, Node* node, Core*, const Api* api) {
    const Table t{api};
    if (auto set_vi = t.set_video_info()) set_vi(&static_cast<SmoothFilter*>(*instance)->vi, 1, node);
}

const FrameRef* smooth_get_frame(int k, int activation, void** instance, void**, FrameContext* fc, Core* core,
                                 const Api* api) {
    auto* f = static_cast<SmoothFilter*>(*instance);
    const Table t{api};
    svpf_smooth_frame fi{};
    if (svpf_smooth_frame_info(f->session, k, &fi) != SVPF_OK) return nullptr;
    if (activation == kArInitial) {
        if (auto request = t.request_frame_filter())
            for (int i = fi.first; i <= fi.last; i++) {
                request(i, f->clip, fc);
                request(i, f->super_clip, fc);
                request(i, f->vectors, fc);
            }
        return nullptr;
    }
    if (activation != kArAllFramesReady) return nullptr;
    const auto get = t.get_frame_filter();
    const auto make = t.new_video_frame();
    const auto release = t.free_frame();
    if (!get || !make || !t.get_stride() || !t.get_read_ptr() || !t.get_write_ptr()) return nullptr;
    const int cnt = fi.last - fi.first + 1;
    std::vector<const FrameRef*> fs((size_t)cnt), fp((size_t)cnt), fv((size_t)cnt);
    bool ok = true;
    for (int i = 0; i < cnt; i++) {
        fs[(size_t)i] = get(fi.first + i, f->clip, fc);
        fp[(size_t)i] = get(fi.first + i, f->super_clip, fc);
        fv[(size_t)i] = get(fi.first + i, f->vectors, fc);
        ok = ok && fs[(size_t)i] && fp[(size_t)i] && fv[(size_t)i];
    }
    const FrameRef* src = ok ? fs[(size_t)(fi.source - fi.first)] : nullptr;
    const FrameRef* next = ok ? fs[(size_t)(std::min(fi.source + 1, fi.last) - fi.first)] : nullptr;
    FrameRef* out = ok ? make(f->vi.format, f->vi.width, f->vi.height, src, core) : nullptr;
    ok = ok && out;
    if (ok) {
        std::lock_guard<std::mutex> lock(f->mu);
        for (int i = 0; i < cnt && ok; i++)
            ok = f->bind(t, fi.first + i, fs[(size_t)i], fp[(size_t)i], fv[(size_t)i], fi.first, fi.last);
        int os[3];
        svpvs::read_strides(t, out, os);
        ok = ok && f->out.fit(f->ctx, os, f->vi.height);
        const svpf_frame fo = f->out.frame();
        ok = ok && svpf_smooth_render(f->session, k, 1, &fo, nullptr) == SVPF_OK;
        ok = ok && svpvs::download(f->ctx, f->out, t, out);
    }
    if (ok) write_timing(t, src, next, out, fi.raw_phase, fi.ratio);
    if (release) {
        for (int i = 0; i < cnt; i++) {
            if (fs[(size_t)i]) release(fs[(size_t)i]);
            if (fp[(size_t)i]) release(fp[(size_t)i]);
            if (fv[(size_t)i]) release(fv[(size_t)i]);
        }
        if (out && !ok) release(out);
    }
    return ok ? out : nullptr;
}

void smooth_free(void* instance, Core*, const Api* api) {
    auto* f = static_cast<SmoothFilter*>(instance);
    const Table t{api};
    svpvs::free_node(t, f->clip);
    svpvs::free_node(t, f->super_clip);
    svpvs::free_node(t, f->vectors);
    svpvs::free_node(t, f->src);
    delete f;
}

void smooth_create(const Map* in, Map* out, void*, Core* core, const Api* api) {
    const Table t{api};
    const auto set_error = t.set_error();
    if (!set_error) return;
    const auto get_node = t.prop_get_node();
    const auto get_vi = t.get_video_info();
    const auto create = t.create_filter();
    const auto get_int = t.prop_get_int();
    const auto get_float = t.prop_get_float();
    if (!get_node || !get_vi || !create || !get_int || !get_float)
        return set_error(out, "SVSmoothFps: invalid VSAPI table");
    auto* f = new SmoothFilter;
    auto fail = [&](const char* msg) {
        smooth_free(f, core, api);
        set_error(out, msg);
    };
    int err = 0;
    f->clip = get_node(in, "clip", 0, &err);
    const VideoInfo* cvi = f->clip ? get_vi(f->clip) : nullptr;
    VideoInfo vi = cvi ? *cvi : VideoInfo{};
    if (vi.fps_den == 0) {  // source_info / fill_fps (filter.rs:225-272)
        err = 0;
        const double fps = get_float(in, "fps", 0, &err);
        if (err || !(fps >= 0.1)) return fail("SVSmoothFps: unable to determine source frame rate");
        vi.fps_num = (int64_t)(fps * 1000.0);
        vi.fps_den = 1000;
    }
    f->super_clip = get_node(in, "super", 0, &err);
    f->vectors = get_node(in, "vectors", 0, &err);
    err = 0;
    f->src = get_node(in, "src", 0, &err);
    err = 0;
    const int64_t sdata = get_int(in, "sdata", 0, &err);
    err = 0;
    const int64_t vdata_addr = get_int(in, "vdata", 0, &err);
    if (err || vdata_addr == 0 || !f->vectors) return fail("SVSmoothFps: invalid vectors stream [1]");
    if (!f->super_clip) return fail("SVSmoothFps: invalid super clip params");
    int32_t vdata[16];
    std::memcpy(vdata, reinterpret_cast<const void*>((intptr_t)vdata_addr), sizeof vdata);
    if (!svpvs::is_yuv420p8(vi.format) || vi.width <= 0 || vi.height <= 0)
        return fail("SVSmoothFps: source must be YV12 (8-bit 4:2:0) for CPU rendering");
    const std::string opt = svpvs::read_data(t, in, "opt");
    char msg[256] = {0};
    int32_t nout = 0;
    if (svpf_smooth_create(nullptr, opt.data(), opt.size(), vi.width, vi.height, vi.fps_num, vi.fps_den,
                           std::max(vi.num_frames, 1), sdata, vdata, nullptr, &nout, msg, sizeof msg) != SVPF_OK)
        return fail(msg[0] ? msg : svpf_last_error());
    const int rc = svpf_create(0, &f->ctx);
    if (rc != SVPF_OK) {
        f->ctx = nullptr;
        const std::string m = "SVSmoothFps: unable to init GPU-based renderer - code " + std::to_string(rc);
        return fail(m.c_str());
    }
    if (svpf_smooth_create(f->ctx, opt.data(), opt.size(), vi.width, vi.height, vi.fps_num, vi.fps_den,
                           std::max(vi.num_frames, 1), sdata, vdata, &f->session, &nout, msg, sizeof msg) != SVPF_OK)
        return fail(msg[0] ? msg : svpf_last_error());
    const VideoInfo* pvi = get_vi(f->super_clip);
    const VideoInfo* vvi = get_vi(f->vectors);
    f->super_rows = pvi ? pvi->height : 0;
    f->payload_len = vvi ? vvi->width : 0;
    f->src_frames = vi.num_frames;
    f->vi = vi;
    if (svpf_smooth_output_rate(f->session, &f->vi.fps_num, &f->vi.fps_den) != SVPF_OK)
        return fail(svpf_last_error());
    f->vi.num_frames = nout;  // output_info (core.rs:603-613), widened by the light.* padding
    if (svpf_smooth_output_size(f->session, &f->vi.width, &f->vi.height) != SVPF_OK) return fail(svpf_last_error());
    create(in, out, "SVSmoothFps", smooth_init, smooth_get_frame, smooth_free, kFmParallel, 0, f, core);
}

// SmoothFps_NVOF / SmoothFps_RIFE need NVIDIA optical flow or a RIFE model: out of this
// build's path, registered so that scripts naming them get the reference's message.
void unavailable_create(const Map*, Map* out, void*, Core*, const Api* api) {
    if (auto set_error = Table{api}.set_error()) set_error(out, "SVSmoothFps: CUDA is not availabe");
}

}  // namespace

SVPVS_EXPORT int64_t svpGetVersion() { return svpf_version(); }

SVPVS_EXPORT void VapourSynthPluginInit(ConfigPlugin config, RegisterFunction reg, Plugin* plugin) {
    if (!config) return;
    config("com.svp-team.flow2", "svp2", "SVPFlow2", kApiVersion, 1, plugin);
    if (!reg) return;
    reg("SmoothFps", "clip:clip;super:clip;sdata:int;vectors:clip;vdata:int;opt:data;src:clip:opt;fps:float:opt",
        smooth_create, nullptr, plugin);
    reg("SmoothFps_NVOF", "clip:clip;opt:data;vec_src:clip:opt;src:clip:opt;fps:float:opt", unavailable_create, nullptr,
        plugin);
    reg("SmoothFps_RIFE", "clip:clip;opt:data;rife_out:clip;vec_src:clip:opt;vdata:int:opt;src:clip:opt;fps:float:opt",
        unavailable_create, nullptr, plugin);
}
