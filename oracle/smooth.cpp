// Oracle restatement of the SmoothFps render orchestration for 8-bit YUV420 sources:
// crates/svpflow2/src/options.rs (parsing, create-time adjustments), core.rs
// (source_frame/phase_256 :2054-2077, try_render :752-883, build_frame_prep
// :1095-1260, render_into_output :1262-1770) and metadata.rs scene classification
// (:275-362, 523-526).  The light border of a padded output is drawn by light.cpp
// after the interior is rendered; scene blending, HDR/DoVi and the debug overlays are
// not part of this path.  Test infrastructure only (see oracle.h).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "render.h"

namespace oracle {

namespace {

int32_t trunc_i32(double v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)std::trunc(v);
}
int32_t sat_i32(int64_t v) { return (int32_t)std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX); }

}  // namespace

uint8_t luma_byte(double v, double gamma) {  // metadata.rs:523-526
    int s = trunc_i32(std::pow(v, gamma) * 255.0);
    return (uint8_t)(std::max(s, 20) & 0xFF);
}

// analyze_scene (metadata.rs:349-362) -> class 0..3
int scene_class(const std::vector<DVec>* prev, const std::vector<DVec>* cur, const VectorData& vd,
                const SmoothOptions& o) {
    if (!prev && !cur) return 3;
    const std::vector<DVec>& p = prev ? *prev : *cur;
    const std::vector<DVec>& c = cur ? *cur : p;
    size_t count = std::min(p.size(), c.size());
    double denom = vd.marker == 3 ? 510.0 : 255.0;
    std::vector<uint8_t> map(count);
    for (size_t i = 0; i < count; i++) map[i] = luma_byte(((double)p[i].luma + (double)c[i].luma) / denom, o.scene_luma);
    // classify_scene (metadata.rs:300-347) over `current`
    int w = std::max(vd.grid_w, 0), h = std::max(vd.grid_h, 0);
    int scene_n = 0, m2 = 0, m1 = 0, considered = 0;
    int zero = sat_i32(o.lim_zero), lm1 = sat_i32(o.lim_m1), lm2 = sat_i32(o.lim_m2), lsc = sat_i32(o.lim_scene);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t idx = (size_t)y * w + x;
            if (idx >= c.size()) continue;
            int luma = std::max<int>(idx < map.size() ? map[idx] : 1, 1);
            int score = sat_i32((int64_t)c[idx].score * 255) / luma;
            if (score < zero) continue;
            considered++;
            if (score >= lsc) scene_n++;
            else if (score >= lm2) m2++;
            else if (score >= lm1) m1++;
        }
    int required = considered * sat_i32(o.lim_blocks) / 100;
    int high = scene_n + m2, mid = high + m1;
    int cls = scene_n >= required ? 3 : (high >= required ? 2 : (mid >= required ? 1 : 0));
    return std::max(cls, 0);
}

namespace {

std::vector<RVector> renderer_vectors(const std::vector<DVec>& v) {  // core.rs:3141-3150
    std::vector<RVector> r(v.size());
    for (size_t i = 0; i < v.size(); i++) r[i] = RVector{v[i].dx, v[i].dy, v[i].score};
    return r;
}

struct Decoded {
    bool ok = false;
    bool hp = false, hc = false;
    std::vector<DVec> prev, cur;
    int scene = 3;
    std::vector<RVector> vprev, vcur;  // previous-or-current / current-or-previous
};

Decoded decode(const SmoothClip& c, int frame) {  // cached_decode + decoded_entry, core.rs:962-1051
    Decoded d;
    frame = std::min(std::max(frame, 0), c.num_frames - 1);
    d.ok = decode_vectors(c.vec[(size_t)frame], c.vec_len, c.vd.flags, c.vd.grid_w, c.vd.grid_h, c.vd.flags, &d.hp,
                          &d.prev, &d.hc, &d.cur);
    if (!d.ok) return d;
    d.scene = scene_class(d.hp ? &d.prev : nullptr, d.hc ? &d.cur : nullptr, c.vd, c.opt);
    const std::vector<DVec>* p = d.hp ? &d.prev : (d.hc ? &d.cur : nullptr);
    const std::vector<DVec>* q = d.hc ? &d.cur : p;
    if (p) d.vprev = renderer_vectors(*p);
    if (q) d.vcur = renderer_vectors(*q);
    return d;
}

void copy_frame(const SmoothClip& c, int frame, uint8_t* dst[3]) {
    frame = std::min(std::max(frame, 0), c.num_frames - 1);
    for (int p = 0; p < 3; p++) {
        size_t n = p ? (size_t)(c.width / 2) * (c.height / 2) : (size_t)c.width * c.height;
        std::memcpy(dst[p], c.src[p][(size_t)frame], n);
    }
}

}  // namespace

Timing SmoothOptions::timing(int64_t snum, int64_t sden) const {  // options.rs:246-254
    int64_t num = has_num ? rate_num : (rate_abs ? 60 : 2);
    int64_t den = has_den ? rate_den : 1;
    return Timing::make(snum, sden, rate_abs, num == 0 ? 1 : num, den == 0 ? 1 : den);
}

bool smooth_options(const Value* opt, int64_t fps_num, int64_t fps_den, const VectorData& vd, SmoothOptions* o,
                    std::string* err) {
    int64_t iv;
    double fv;
    bool bv;
    auto I = [&](std::initializer_list<const char*> p, int64_t* v) { return opt && opt->int_at(p, v); };
    auto F = [&](std::initializer_list<const char*> p, double* v) { return opt && opt->float_at(p, v); };
    auto B = [&](std::initializer_list<const char*> p, bool* v) { return opt && opt->bool_at(p, v); };
    *o = SmoothOptions{};
    o->rate_abs = B({"rate", "abs"}, &bv) ? bv : false;
    o->has_num = I({"rate", "num"}, &o->rate_num);
    o->has_den = I({"rate", "den"}, &o->rate_den);
    o->scene_mode = I({"scene", "mode"}, &iv) ? iv : 3;  // Scene::from_value, options.rs:543-567
    o->scene_blend = B({"scene", "blend"}, &bv) ? bv : false;
    auto digit = [](int64_t v, int64_t d) { return (int)std::min<int64_t>((v / d) % 10, 3) - 1; };
    if (o->scene_mode == 3) {
        int64_t a = I({"scene", "adaptive"}, &iv) ? iv : 210;
        o->adaptive[0] = digit(a, 1); o->adaptive[1] = digit(a, 10); o->adaptive[2] = digit(a, 100);
    } else {
        o->adaptive[0] = -1; o->adaptive[1] = -1; o->adaptive[2] = 1;
    }
    o->force13 = B({"scene", "force13"}, &bv) ? bv : true;
    o->scene_luma = std::min(std::max(F({"scene", "luma"}, &fv) ? fv : 1.5, 0.0), 5.0);
    // SceneLimits::from_value (options.rs:655-687), mode 0 defaults
    int64_t blocks = I({"scene", "limits", "blocks"}, &iv) ? iv : 20;
    o->lim_blocks = (blocks >= 0 && blocks < 100) ? blocks : 100;
    o->lim_zero = I({"scene", "limits", "zero"}, &iv) ? iv : 200;
    o->lim_m1 = I({"scene", "limits", "m1"}, &iv) ? iv : 1600;
    o->lim_m2 = I({"scene", "limits", "m2"}, &iv) ? iv : 2800;
    o->lim_scene = I({"scene", "limits", "scene"}, &iv) ? iv : 4000;
    // Light::from_value (options.rs:147-156), Options::light_params (:279-286)
    o->light.sar = F({"light", "sar"}, &fv) ? fv : 1.0;
    o->light.aspect = F({"light", "aspect"}, &fv) ? fv : 0.0;
    o->light.zoom = F({"light", "zoom"}, &fv) ? fv : 0.0;
    auto sat32 = [](int64_t v) { return (int)std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX); };
    o->light_params.border = sat32(I({"light", "border"}, &iv) ? iv : 12);
    o->light_params.lights = std::max(sat32(I({"light", "lights"}, &iv) ? iv : 16), 1);
    o->light_params.length = sat32(I({"light", "length"}, &iv) ? iv : 100);
    o->light_params.cell = std::max(F({"light", "cell"}, &fv) ? fv : 1.0, 0.1);
    o->mask_area = I({"mask", "area"}, &iv) ? iv : 0;  // Mask::from_value, options.rs:569-581
    o->mask_cover = std::max<int64_t>(I({"mask", "cover"}, &iv) ? iv : 100, 0);
    o->mask_area_enabled = o->mask_area > 0;
    o->area_sharp = F({"mask", "area_sharp"}, &fv) ? fv : 1.0;
    o->algo = I({"algo"}, &iv) ? iv : 21;
    o->block = B({"block"}, &bv) ? bv : false;
    // validate (options.rs:169-177)
    bool valid_algo = o->algo == 1 || o->algo == 2 || o->algo == 11 || o->algo == 13 || o->algo == 21 ||
                      o->algo == 22 || o->algo == 23 || (o->algo >= 90 && o->algo <= 100);
    if (!valid_algo) { *err = "SVSmoothFps: incorrect 'algo' value"; return false; }
    if (o->scene_mode < 0 || o->scene_mode > 3) { *err = "SVSmoothFps: 'scene.mode' must be in [0;3]"; return false; }
    Timing t = o->timing(fps_num, fps_den);
    auto below = [&](int64_t lim) { return (__int128)t.frame_num < (__int128)lim * t.frame_den; };
    if ((o->scene_mode == 2 && below(2)) || (o->scene_mode == 1 && below(1))) {
        *err = "SVSmoothFps: 'scene.mode' value isn't supported for selected frame rate";
        return false;
    }
    if (o->scene_mode == 3 && below(1)) {  // normalize_scene_mode, options.rs:193-203
        o->scene_mode = 0;
        o->adaptive[0] = -1; o->adaptive[1] = -1; o->adaptive[2] = 1;
    }
    // scale_scene_limits by the vector block (options.rs:688-693, 768-772)
    auto scale = [&](int64_t v) { return (int64_t)trunc_i32((double)v * (double)vd.block_w * (double)vd.block_h / 32.0); };
    o->lim_zero = scale(o->lim_zero);
    o->lim_m1 = scale(o->lim_m1);
    o->lim_m2 = scale(o->lim_m2);
    o->lim_scene = scale(o->lim_scene);
    if (o->mask_area > 0) o->area_scale = (double)sat_i32(o->mask_area) / 15000.0;  // options.rs:220-225
    if (o->mask_cover == 0 && o->algo >= 21 && o->algo <= 31) o->algo = o->algo == 22 ? 13 : 11;  // :233-243
    return true;
}

int smooth_output_frames(const SmoothClip& c) {
    return c.opt.timing(c.fps_num, c.fps_den).scale_frame_count(c.num_frames);
}

int smooth_render_frame(const SmoothClip& c, int k, uint8_t* dst[3], int* algo_out) {
    const SmoothOptions& o = c.opt;
    Timing t = o.timing(c.fps_num, c.fps_den);
    *algo_out = 0;
    const int64_t sm = o.scene_mode;
    // source_frame / phase_256 (core.rs:2054-2077)
    bool nearest = k > 0 && !o.scene_blend && sm != 0 && sm != 2;
    int n = t.source_frame(k, nearest);
    int phase = 0;
    if (t.frame_num > 0) {
        if (sm == 1 || sm == 2) phase = t.scene_phase_256(t.raw_phase_256(k, n), sm == 1 ? 0 : 2);
        else phase = t.phase_256(k, n);
    }
    const int raw_phase = phase;
    auto fallback = [&]() {  // core.rs:266-299: next source when the phase is exactly 256
        if (raw_phase == 256) { copy_frame(c, n + 1, dst); return 1; }
        copy_frame(c, n, dst);
        return 0;
    };
    // try_render gates (core.rs:765-800)
    if (k == 0 || phase == 0 || phase == 256) return fallback();
    double ratio = (double)t.frame_num / (double)t.frame_den;
    if (!o.scene_blend && sm == 3 && t.frame_den != 0 && ratio < 2.0) return fallback();
    int algo_req = (int)o.algo;
    if (!(algo_req == 1 || algo_req == 2 || algo_req == 11 || algo_req == 13 || algo_req == 21 || algo_req == 22 ||
          algo_req == 23))
        return fallback();
    phase = std::min(std::max(phase, 0), 256);
    const VectorData& vd = c.vd;
    // prep (core.rs:1095-1260)
    Decoded self = decode(c, n);
    if (!self.ok) return fallback();
    int raw_class = self.scene;
    if (raw_class >= 3) {  // scene change: nearest source (core.rs:819-842)
        if (phase < 128) { copy_frame(c, n, dst); return 0; }
        copy_frame(c, n + 1, dst);
        return 1;
    }
    const size_t mw = (size_t)std::max(vd.motion_w(), 0), mh = (size_t)std::max(vd.motion_h(), 0);
    const size_t mc = mw * mh;
    VectorContext ctx{std::max(vd.block_w, 1), std::max(vd.block_h, 1), std::max(vd.marker, 1), c.width, c.height,
                      vd.overlap_x, vd.overlap_y, (size_t)std::max(vd.grid_w, 0), (size_t)std::max(vd.grid_h, 0), false,
                      &self.vprev, &self.vcur};
    std::vector<uint16_t> mx0(mc), my0(mc), mx1(mc), my1(mc), mx2(mc, 0), my2(mc, 0), mx3(mc, 0), my3(mc, 0);
    vector_planes(ctx, 0, mx0, my0, mw, mh);
    vector_planes(ctx, 1, mx1, my1, mw, mh);
    std::vector<uint8_t> mag0, mag1;
    if (o.mask_area_enabled) {
        mag0.assign(mc, 0);
        mag1.assign(mc, 0);
        magnitude_mask(ctx, 0, mag0, mw, mh, o.area_scale, o.area_sharp);
        magnitude_mask(ctx, 1, mag1, mw, mh, o.area_scale, o.area_sharp);
    }
    // neighbours for algo 23 (vector_neighbors, core.rs:2456-2506): n-1 and n+1
    bool side_ready = false;
    bool neighbours = algo_req == 23;
    if (algo_req == 23) {
        Decoded prev = decode(c, std::max(n - 1, 0));
        Decoded next = decode(c, n + 1);
        if (prev.ok && next.ok) {
            bool prev_side = prev.hc || prev.hp;
            bool next_side = next.hp || next.hc;
            if (next.scene < 3 && prev_side && next_side) {
                VectorContext next_ctx = ctx;
                next_ctx.a = &next.vprev;
                next_ctx.b = &self.vcur;
                VectorContext prev_ctx = ctx;
                prev_ctx.a = &self.vprev;
                prev_ctx.b = &prev.vcur;
                vector_planes(next_ctx, 0, mx2, my2, mw, mh);
                vector_planes(prev_ctx, 1, mx3, my3, mw, mh);
                side_ready = true;
            }
        }
    }
    // render_into_output (core.rs:1262-1770)
    int scene = (phase == 0 || phase == 256) ? 0 : raw_class;  // effective_scene
    if (sm == 3 && scene < 3) {  // adaptive_phase (core.rs:2132-2147)
        if (t.frame_den != 0) {
            if (ratio < 2.0) phase = 0;
            else if (o.adaptive[scene] >= 0) phase = t.scene_phase_256(t.raw_phase_256(k, n), o.adaptive[scene]);
        }
    }
    if (phase == 0 || phase == 256) return fallback();
    int algo = algo_req;
    if (scene >= 1 && scene < 3 && algo_req >= 11 && o.force13) algo = 13;  // effective_algorithm
    if (algo == 2 && phase > 128) algo = 1;
    *algo_out = algo;
    bool mode23_ready = side_ready && algo == 23 && scene < 3 && neighbours;
    bool interp = !o.block;
    // sources: the pel-expanded super frame when Analyse ran at pel > 1 (apply_request_flags, filter.rs:680-688)
    int render_mode = (std::fabs(ratio - 1.0) < 0.001) ? 0 : (vd.selector < 1 ? 1 : 2);
    bool request_super = (render_mode == 1 || render_mode == 2) && vd.marker != 1;
    uint64_t raw;
    std::memcpy(&raw, &c.sdata, 8);
    int scale = (int)((raw >> 32) & 0xFF);
    int n0 = std::min(std::max(n, 0), c.num_frames - 1), n1 = std::min(n + 1, c.num_frames - 1);
    RPlane s0[3], s1[3];
    bool super_ok = request_super && (scale == 1 || scale == 2 || scale == 4);
    for (int p = 0; p < 3; p++) {
        size_t pw = p ? (size_t)(c.width / 2) : (size_t)c.width, ph = p ? (size_t)(c.height / 2) : (size_t)c.height;
        if (super_ok) {
            uint32_t shift = scale == 1 ? 0 : (scale == 2 ? 1 : 2);
            size_t rows = p ? (size_t)c.super_rows / 2 : (size_t)c.super_rows;
            s0[p] = RPlane{c.sup[p][(size_t)n0], pw * rows, pw, pw * ph, shift};
            s1[p] = RPlane{c.sup[p][(size_t)n1], pw * rows, pw, pw * ph, shift};
        } else {
            s0[p] = RPlane{c.src[p][(size_t)n0], pw * ph, pw, 0, 0};
            s1[p] = RPlane{c.src[p][(size_t)n1], pw * ph, pw, 0, 0};
        }
    }
    int source_step = super_ok ? std::max(scale, 1) : 1;
    std::vector<uint8_t> cov0, cov1;
    if (algo >= 21) {
        cov0.assign(mc, 0);
        cov1.assign(mc, 0);
        coverage_mask(ctx, 0, cov0, mw, mh, (int)std::min<int64_t>(o.mask_cover, INT32_MAX), 256 - phase);
        coverage_mask(ctx, 1, cov1, mw, mh, (int)std::min<int64_t>(o.mask_cover, INT32_MAX), phase);
    }
    CpuConfig cfg{c.width, c.height, std::max(vd.block_w - vd.overlap_x, 1), std::max(vd.block_h - vd.overlap_y, 1),
                  vd.block_w / 2, vd.block_h / 2, mw, mh, 2, source_step, o.area_scale};
    CpuRenderer cpu(cfg);
    cpu.set_threshold(phase);
    RPlaneMut d[3];
    for (int p = 0; p < 3; p++) {
        size_t pw = p ? (size_t)(c.width / 2) : (size_t)c.width, ph = p ? (size_t)(c.height / 2) : (size_t)c.height;
        d[p] = RPlaneMut{dst[p], pw * ph, pw};
    }
    if (c.float_render && interp && !o.scene_blend) {
        FloatFrame ff{};
        ff.width = c.width;
        ff.height = c.height;
        for (int p = 0; p < 3; p++) {
            ff.s0[p] = c.src[p][(size_t)n0];
            ff.s1[p] = c.src[p][(size_t)n1];
        }
        ff.mw = (int)mw;
        ff.mh = (int)mh;
        const std::vector<uint16_t>* mp[8] = {&mx0, &my0, &mx1, &my1, &mx2, &my2, &mx3, &my3};
        for (int q = 0; q < 8; q++) ff.m[q] = mp[q]->data();
        ff.cov0 = cov0.data();
        ff.cov1 = cov1.data();
        ff.has_sad = o.mask_area_enabled;
        ff.mag0 = mag0.data();
        ff.mag1 = mag1.data();
        ff.algo = (algo == 23 && !mode23_ready) ? 21 : algo;
        ff.phase = phase;
        ff.block_w = cfg.block_w;
        ff.block_h = cfg.block_h;
        ff.origin_x = cfg.origin_x;
        ff.origin_y = cfg.origin_y;
        float_render(ff, dst);
        return 2;
    }
    Motion m0{&mx0, &my0}, m1{&mx1, &my1}, m2{&mx2, &my2}, m3{&mx3, &my3};
    std::vector<uint8_t> maxm;
    if (o.mask_area_enabled) maxm = max_mask(mag0, mag1);
    switch (algo) {
        case 1:
        case 2:
            cpu.mode1_or_2(algo == 1, interp, d, s0, s1, m0, m1, o.mask_area_enabled ? &mag0 : nullptr,
                           o.mask_area_enabled ? &mag1 : nullptr);
            break;
        case 11:
        case 13:
            cpu.mode11_or_13(algo == 13, interp, !interp, d, s0, s1, m0, m1, o.mask_area_enabled ? &maxm : nullptr);
            break;
        case 21:
        case 22:
            cpu.mode21_or_22(algo == 21, interp, d, s0, s1, m0, m1, cov0, cov1, o.mask_area_enabled ? &maxm : nullptr);
            break;
        case 23:
            if (mode23_ready) cpu.mode23(interp, d, s0, s1, m0, m1, m2, m3, cov0, cov1, o.mask_area_enabled ? &maxm : nullptr);
            else cpu.mode21_or_22(true, interp, d, s0, s1, m0, m1, cov0, cov1, o.mask_area_enabled ? &maxm : nullptr);
            break;
        default:
            return fallback();
    }
    return 2;
}

}  // namespace oracle
