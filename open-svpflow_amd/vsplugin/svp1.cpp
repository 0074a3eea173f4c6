// svp1 plugin shell (VapourSynth API 3): svpflow1's Super and Analyse filters over
// libsvpf.  It registers what the reference registers (crates/svpflow1/src/abi.rs:
// 19-46) and keeps its filters' create-time checks, error texts, output clips and
// frame requests (super_filter.rs:9-268, analyse_filter.rs:19-334); the frame work
// runs on the GPU through the C ABI (one context per filter instance, created on the
// first frame, calls on it serialised by a mutex).
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"

using namespace vs3;
using svpvs::DevFrame;

namespace {

// ------------------------------------------------------------------ Super

struct SuperFilter {
    Node* node = nullptr;
    VideoInfo vi{};  // output: the source's with height = super plane height
    svpf_super_params p{};
    int src_rows = 0;
    std::mutex mu;
    svpf_ctx* ctx = nullptr;
    DevFrame src, sup;

    bool build(const Table& t, const FrameRef* in, FrameRef* out) {
        std::lock_guard<std::mutex> lock(mu);
        if (!ctx && svpf_create(0, &ctx) != SVPF_OK) return false;
        int ds[3];
        svpvs::read_strides(t, out, ds);
        if (!svpvs::upload(ctx, src, t, in, src_rows) || !sup.fit(ctx, ds, vi.height)) return false;
        const svpf_frame fs = src.frame(), fd = sup.frame();
        if (svpf_super_build(ctx, &p, &fs, &fd) != SVPF_OK) return false;
        return svpvs::download(ctx, sup, t, out);
    }
    ~SuperFilter() {
        if (ctx) {
            src.release(ctx);
            sup.release(ctx);
            svpf_destroy(ctx);
        }
    }
};

void super_init(Map*, Map*, void** instance, Node* node, Core*, const Api* api) {
    const Table t{api};
    if (auto set_vi = t.set_video_info()) set_vi(&static_cast<SuperFilter*>(*instance)->vi, 1, node);
}

const FrameRef* super_get_frame(int n, int activation, void** instance, void**, FrameContext* fc, Core* core,
                                const Api* api) {
    auto* f = static_cast<SuperFilter*>(*instance);
    const Table t{api};
    if (activation == kArInitial) {
        if (auto request = t.request_frame_filter()) request(n, f->node, fc);
        return nullptr;
    }
    if (activation != kArAllFramesReady) return nullptr;
    const auto get = t.get_frame_filter();
    const auto make = t.new_video_frame();
    const auto release = t.free_frame();
    if (!get || !make || !t.get_stride() || !t.get_read_ptr() || !t.get_write_ptr()) return nullptr;
    const FrameRef* src = get(n, f->node, fc);
    if (!src) return nullptr;
    FrameRef* out = make(f->vi.format, f->vi.width, f->vi.height, src, core);
    const bool ok = out && f->build(t, src, out);
    if (release) {
        release(src);
        if (out && !ok) release(out);
    }
    return ok ? out : nullptr;
}

void super_free(void* instance, Core*, const Api* api) {
    auto* f = static_cast<SuperFilter*>(instance);
    svpvs::free_node(Table{api}, f->node);
    delete f;
}

void super_create(const Map* in, Map* out, void*, Core* core, const Api* api) {
    const Table t{api};
    const auto set_error = t.set_error();
    if (!set_error) return;
    const auto get_node = t.prop_get_node();
    const auto get_vi = t.get_video_info();
    const auto create = t.create_filter();
    const auto set_int = t.prop_set_int();
    if (!get_node || !get_vi || !create || !set_int) return set_error(out, "SVSuper: invalid VSAPI table");
    int err = 0;
    Node* node = get_node(in, "clip", 0, &err);
    if (!node) return set_error(out, "SVSuper: clip is required");
    const VideoInfo* vip = get_vi(node);
    if (!vip) {
        svpvs::free_node(t, node);
        return set_error(out, "SVSuper: invalid clip");
    }
    // the reference rejects only an empty clip here; the device path also needs
    // 8-bit 4:2:0, which the same message names
    if (vip->width <= 0 || vip->height <= 0 || !vip->format || !svpvs::is_yuv420p8(vip->format)) {
        svpvs::free_node(t, node);
        return set_error(out, "SVSuper: Clip must be YV12");
    }
    auto* f = new SuperFilter;
    f->node = node;
    f->vi = *vip;
    const std::string opt = svpvs::read_data(t, in, "opt");
    int64_t sdata = 0;
    char msg[256] = {0};
    if (svpf_super_params_from_opt(opt.data(), opt.size(), vip->width, vip->height, &f->p, &sdata, msg,
                                   sizeof msg) != SVPF_OK) {
        delete f;
        svpvs::free_node(t, node);
        return set_error(out, msg[0] ? msg : svpf_last_error());
    }
    set_int(out, "data", sdata, kPaReplace);
    f->src_rows = vip->height;
    f->vi.height = svpf_super_height(&f->p);
    create(in, out, "SVSuper", super_init, super_get_frame, super_free, kFmParallel, 0, f, core);
}

// ------------------------------------------------------------------ Analyse

struct AnalyseFilter {
    Node* super_node = nullptr;
    Node* src_node = nullptr;
    VideoInfo vi{};  // Gray8, payload_len x 1
    svpf_analyse_params p{};
    int32_t vdata[16] = {};  // SmoothFps reads this header through the `data` address
    int32_t payload_len = 0;
    int super_rows = 0;
    std::mutex mu;
    svpf_ctx* ctx = nullptr;
    DevFrame cur, ref;
    uint8_t* d_payload = nullptr;

    // The vector payload of the pair (f0, f1) into `payload` (the caller's: frames of one
    // instance may be requested from several threads at once).
    bool search(const Table& t, const FrameRef* f0, const FrameRef* f1, std::vector<uint8_t>& payload) {
        std::lock_guard<std::mutex> lock(mu);
        if (!ctx && svpf_create(0, &ctx) != SVPF_OK) return false;
        if (!d_payload) {
            void* d = nullptr;
            if (svpf_device_alloc(ctx, ((size_t)payload_len + 3) & ~(size_t)3, &d) != SVPF_OK) return false;
            d_payload = (uint8_t*)d;
        }
        if (!svpvs::upload(ctx, cur, t, f0, super_rows) || !svpvs::upload(ctx, ref, t, f1, super_rows)) return false;
        const svpf_frame a = cur.frame(), b = ref.frame();
        if (svpf_analyse_pair(ctx, &p, &a, &b, d_payload) != SVPF_OK) return false;
        payload.resize((size_t)payload_len);
        return svpf_copy_d2h(ctx, payload.data(), d_payload, payload.size()) == SVPF_OK;
    }
    ~AnalyseFilter() {
        if (ctx) {
            cur.release(ctx);
            ref.release(ctx);
            if (d_payload) svpf_device_free(ctx, d_payload);
            svpf_destroy(ctx);
        }
    }
};

void analyse_init(Map*, Map*, void** instance, Node* node, Core*, const Api* api) {
    const Table t{api};
    if (auto set_vi = t.set_video_info()) set_vi(&static_cast<AnalyseFilter*>(*instance)->vi, 1, node);
}

const FrameRef* analyse_get_frame(int n, int activation, void** instance, void**, FrameContext* fc, Core* core,
                                  const Api* api) {
    auto* f = static_cast<AnalyseFilter*>(*instance);
    const Table t{api};
    const int delta = std::max(f->p.delta, 1);
    if (activation == kArInitial) {
        if (auto request = t.request_frame_filter()) {
            request(n, f->super_node, fc);
            request(n + delta, f->super_node, fc);
        }
        return nullptr;
    }
    if (activation != kArAllFramesReady) return nullptr;
    const auto get = t.get_frame_filter();
    const auto make = t.new_video_frame();
    const auto release = t.free_frame();
    if (!get || !make || !t.get_stride() || !t.get_read_ptr() || !t.get_write_ptr()) return nullptr;
    const int n2 = std::min(n + delta, std::max(f->vi.num_frames - 1, 0));
    const FrameRef* f0 = get(n, f->super_node, fc);
    const FrameRef* f1 = get(n2, f->super_node, fc);
    FrameRef* out = nullptr;
    std::vector<uint8_t> payload;
    bool ok = f0 && f1 && f->search(t, f0, f1, payload);
    if (ok) {
        out = make(f->vi.format, f->payload_len, 1, f0, core);
        uint8_t* row = out ? t.get_write_ptr()(out, 0) : nullptr;
        const int stride = out ? t.get_stride()(out, 0) : 0;
        ok = row && stride >= f->payload_len;
        if (ok) {  // the payload, then zeros to the end of the row (analyse_filter.rs:312-322)
            std::memcpy(row, payload.data(), (size_t)f->payload_len);
            std::memset(row + f->payload_len, 0, (size_t)(stride - f->payload_len));
        }
    }
    if (release) {
        if (f0) release(f0);
        if (f1) release(f1);
        if (out && !ok) release(out);
    }
    return ok ? out : nullptr;
}

void analyse_free(void* instance, Core*, const Api* api) {
    auto* f = static_cast<AnalyseFilter*>(instance);
    const Table t{api};
    svpvs::free_node(t, f->super_node);
    svpvs::free_node(t, f->src_node);
    delete f;
}

void analyse_create(const Map* in, Map* out, void*, Core* core, const Api* api) {
    const Table t{api};
    const auto set_error = t.set_error();
    if (!set_error) return;
    const auto get_node = t.prop_get_node();
    const auto get_vi = t.get_video_info();
    const auto create = t.create_filter();
    const auto set_int = t.prop_set_int();
    const auto get_int = t.prop_get_int();
    if (!get_node || !get_vi || !create || !set_int || !get_int) return set_error(out, "SVAnalyse: invalid VSAPI");
    int err = 0;
    Node* super_node = get_node(in, "clip", 0, &err);
    if (!super_node) return set_error(out, "SVAnalyse: clip (super) required");
    Node* src_node = get_node(in, "src", 0, &err);
    if (!src_node) {
        svpvs::free_node(t, super_node);
        return set_error(out, "SVAnalyse: src required");
    }
    err = 0;
    const int64_t sdata = get_int(in, "sdata", 0, &err);
    auto fail = [&](const char* msg) {
        svpvs::free_node(t, super_node);
        svpvs::free_node(t, src_node);
        set_error(out, msg);
    };
    if (err) return fail("SVAnalyse: sdata required");
    const VideoInfo* svi = get_vi(src_node);
    const VideoInfo* pvi = get_vi(super_node);
    if (!svi || !pvi) return fail("SVAnalyse: clip (super) required");
    auto* f = new AnalyseFilter;
    const std::string opt = svpvs::read_data(t, in, "opt");
    char msg[256] = {0};
    if (svpf_analyse_params_from_opt(opt.data(), opt.size(), sdata, svi->width, svi->height, &f->p, f->vdata,
                                     &f->payload_len, msg, sizeof msg) != SVPF_OK) {
        delete f;
        return fail(msg[0] ? msg : svpf_last_error());
    }
    const Format* gray = t.get_format_preset() ? t.get_format_preset()(kPresetGray8, core) : nullptr;
    if (!gray) gray = pvi->format;
    if (!gray) {
        delete f;
        return fail("SVAnalyse: cannot get Gray8 format");
    }
    f->super_node = super_node;
    f->src_node = src_node;
    f->super_rows = pvi->height;
    f->vi = VideoInfo{gray, svi->fps_num, svi->fps_den, f->payload_len, 1, svi->num_frames, 0};
    if (pvi->num_frames > 0 && pvi->num_frames < f->vi.num_frames) f->vi.num_frames = pvi->num_frames;
    set_int(out, "data", (int64_t)(intptr_t)f->vdata, kPaReplace);
    create(in, out, "SVAnalyse", analyse_init, analyse_get_frame, analyse_free, kFmParallel, 0, f, core);
}

}  // namespace

SVPVS_EXPORT int64_t svpGetVersion() { return svpf_version(); }

SVPVS_EXPORT void VapourSynthPluginInit(ConfigPlugin config, RegisterFunction reg, Plugin* plugin) {
    if (!config) return;
    config("com.svp-team.flow1", "svp1", "SVPFlow1", kApiVersion, 1, plugin);
    if (!reg) return;
    reg("Super", "clip:clip;opt:data", super_create, nullptr, plugin);
    reg("Analyse", "clip:clip;sdata:int;src:clip;opt:data", analyse_create, nullptr, plugin);
}
