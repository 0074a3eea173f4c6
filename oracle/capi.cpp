// extern "C" surface of the oracle for the Python test harness (ctypes).
// Test infrastructure only (see oracle.h).
#include <cstring>

#include "oracle.h"

using namespace oracle;

namespace {
bool parse_opt(const char* opt, size_t len, Value* v) {
    if (!opt || !len) return false;
    std::string err;
    return parse((const uint8_t*)opt, len, v, &err);
}
void put_err(char* err, size_t n, const std::string& m) {
    if (err && n) {
        std::strncpy(err, m.c_str(), n - 1);
        err[n - 1] = 0;
    }
}
}  // namespace

extern "C" {

// Option-string probe: returns 1 and fills the typed readings when the path exists.
// kind: 0 int, 1 float, 2 bool.  path: keys separated by '.'.
int oracle_opt_lookup(const char* opt, size_t len, const char* path, int kind, int* parsed, int64_t* iv,
                      double* fv) {
    Value v;
    std::string err;
    *parsed = parse((const uint8_t*)opt, len, &v, &err) ? 1 : 0;
    if (!*parsed) return 0;
    std::vector<std::string> keys;
    std::string cur;
    for (const char* p = path; *p; p++) {
        if (*p == '.') { keys.push_back(cur); cur.clear(); } else cur += *p;
    }
    keys.push_back(cur);
    const Value* at = &v;
    for (auto& k : keys) {
        at = at->get(k);
        if (!at) return 0;
    }
    Value root;
    root.kind = Value::Object;
    root.object.emplace_back("k", *at);
    if (kind == 0) return root.int_at({"k"}, iv) ? 1 : 0;
    if (kind == 1) return root.float_at({"k"}, fv) ? 1 : 0;
    bool b;
    if (!root.bool_at({"k"}, &b)) return 0;
    *iv = b;
    return 1;
}

// SuperOpts::from_opt; out = {width,height,pel,gpu,scale_down,scale_up,full,rc,levels}.
int oracle_super_opts(const char* opt, size_t len, int w, int h, int32_t* out, int64_t* sdata, int32_t* super_h,
                      char* err, size_t errlen) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    SuperOpts o;
    std::string e = super_opts_from(have ? &v : nullptr, w, h, &o);
    if (!e.empty()) {
        put_err(err, errlen, e);
        return -3;
    }
    out[0] = o.width; out[1] = o.height; out[2] = o.pel; out[3] = o.gpu; out[4] = o.scale_down;
    out[5] = o.scale_up; out[6] = o.full; out[7] = o.rc; out[8] = o.levels;
    *sdata = super_pack_data(o);
    *super_h = super_height(o);
    return 0;
}

// get_frame_super (super_filter.rs:143-268): zero the three planes, build_plane each.
int oracle_super_build(const char* opt, size_t len, int w, int h, const uint8_t* sy, int ssy, const uint8_t* su,
                       int ssu, const uint8_t* sv, int ssv, uint8_t* dy, int dsy, uint8_t* du, int dsu, uint8_t* dv,
                       int dsv) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    SuperOpts o;
    if (!super_opts_from(have ? &v : nullptr, w, h, &o).empty()) return -3;
    int sh = super_height(o);
    const uint8_t* src[3] = {sy, su, sv};
    int ss[3] = {ssy, ssu, ssv};
    uint8_t* dst[3] = {dy, du, dv};
    int ds[3] = {dsy, dsu, dsv};
    for (int p = 0; p < 3; p++) {
        size_t rows = p ? (size_t)sh / 2 : (size_t)sh;
        std::memset(dst[p], 0, (size_t)ds[p] * rows);
    }
    for (int p = 0; p < 3; p++) {
        size_t bw = p ? (size_t)w / 2 : (size_t)w, bh = p ? (size_t)h / 2 : (size_t)h;
        size_t rows = p ? (size_t)sh / 2 : (size_t)sh;
        build_plane(dst[p], (size_t)ds[p] * rows, (size_t)ds[p], src[p], (size_t)ss[p] * bh, (size_t)ss[p], bw, bh,
                    (size_t)w, (size_t)h, p != 0, o);
    }
    return 0;
}

}  // extern "C"

#include "analyse.h"

extern "C" {

// AnalyseOpts as seen by the vector clip: returns 0 or -3 with the error text.
// out = {vectors, block_w, block_h, overlap_x, overlap_y, pel, levels, super_levels, width, height,
//        lambda, lsad, search_type, coarse_search_type, distance, payload_len}
int oracle_analyse_opts(const char* opt, size_t len, int64_t sdata, int width, int height, int32_t* out,
                        int32_t* vdata16, char* err, size_t errlen) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    AnalyseOpts o;
    std::string e = analyse_opts_from(have ? &v : nullptr, sdata, nullptr, &o);
    if (!e.empty()) {
        put_err(err, errlen, e);
        return -3;
    }
    if (width > 0) o.width = width;
    if (height > 0) o.height = height;
    int bw, bh, ox, oy, gw, gh;
    o.output_block(&bw, &bh, &ox, &oy);
    o.grid(bw, bh, ox, oy, &gw, &gh);
    int nreg = (o.vectors == 1 || o.vectors == 2) ? 1 : 2;
    int32_t vals[16] = {o.vectors, o.block_w, o.block_h, o.overlap_x, o.overlap_y, o.pel, o.levels, o.super_levels,
                        o.width, o.height, o.lambda, o.lsad, o.search_type, o.coarse_search_type, o.distance,
                        0x40 + nreg * (4 + gw * gh * 8)};
    std::memcpy(out, vals, sizeof vals);
    auto h = pack_vdata_header(o);
    std::memcpy(vdata16, h.data(), 16 * 4);
    return 0;
}

// get_frame_analyse (analyse_filter.rs:229-334) on two super frames held in host
// memory: cur = super frame n, ref = super frame n+delta.  Writes the vector payload.
int oracle_analyse(const char* opt, size_t len, int64_t sdata, int width, int height, const uint8_t* cy, int cys,
                   const uint8_t* cu, int cus, const uint8_t* cv, int cvs, const uint8_t* ry, int rys,
                   const uint8_t* ru, int rus, const uint8_t* rv, int rvs, uint8_t* out, size_t out_cap,
                   size_t* out_len, uint64_t* evals16, uint64_t* pixels16, int threads) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    AnalyseOpts o;
    if (!analyse_opts_from(have ? &v : nullptr, sdata, nullptr, &o).empty()) return -3;
    if (width > 0) o.width = width;
    if (height > 0) o.height = height;
    size_t yh = (size_t)super_plane_height(o.height, o.pel, o.super_levels, true), uvh = yh / 2;
    SuperPlanes c{cy, (size_t)cys * yh, (size_t)cys, cu, (size_t)cus * uvh, (size_t)cus, cv, (size_t)cvs * uvh,
                  (size_t)cvs, (size_t)o.width, (size_t)o.height, o.pel, o.super_levels, true};
    SuperPlanes r{ry, (size_t)rys * yh, (size_t)rys, ru, (size_t)rus * uvh, (size_t)rus, rv, (size_t)rvs * uvh,
                  (size_t)rvs, (size_t)o.width, (size_t)o.height, o.pel, o.super_levels, true};
    bool hb, hf;
    std::vector<Vec8> bwd, fwd;
    SearchCounters sc;
    analyse_pair(c, r, o, &hb, &bwd, &hf, &fwd, &sc, threads != 0);
    auto payload = pack_vector_frame(o, hb ? &bwd : nullptr, hf ? &fwd : nullptr);
    *out_len = payload.size();
    if (payload.size() > out_cap) return -1;
    std::memcpy(out, payload.data(), payload.size());
    if (evals16) std::memcpy(evals16, sc.evals, sizeof sc.evals);
    if (pixels16) std::memcpy(pixels16, sc.pixels, sizeof sc.pixels);
    return 0;
}

// get_frame_analyse with a per-block trace of both directions (BlockTrace, 64 words
// each): trace0 receives direction 0 (cur -> ref), trace1 direction 1.  Runs the two
// directions one after the other.  Returns -1 when a trace buffer is too small (the
// counts needed are still written).
int oracle_analyse_trace(const char* opt, size_t len, int64_t sdata, int width, int height, const uint8_t* cy,
                         int cys, const uint8_t* cu, int cus, const uint8_t* cv, int cvs, const uint8_t* ry, int rys,
                         const uint8_t* ru, int rus, const uint8_t* rv, int rvs, uint8_t* out, size_t out_cap,
                         size_t* out_len, BlockTrace* trace0, size_t cap0, size_t* n0, BlockTrace* trace1,
                         size_t cap1, size_t* n1) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    AnalyseOpts o;
    if (!analyse_opts_from(have ? &v : nullptr, sdata, nullptr, &o).empty()) return -3;
    if (width > 0) o.width = width;
    if (height > 0) o.height = height;
    size_t yh = (size_t)super_plane_height(o.height, o.pel, o.super_levels, true), uvh = yh / 2;
    SuperPlanes c{cy, (size_t)cys * yh, (size_t)cys, cu, (size_t)cus * uvh, (size_t)cus, cv, (size_t)cvs * uvh,
                  (size_t)cvs, (size_t)o.width, (size_t)o.height, o.pel, o.super_levels, true};
    SuperPlanes r{ry, (size_t)rys * yh, (size_t)rys, ru, (size_t)rus * uvh, (size_t)rus, rv, (size_t)rvs * uvh,
                  (size_t)rvs, (size_t)o.width, (size_t)o.height, o.pel, o.super_levels, true};
    bool hb, hf;
    std::vector<Vec8> bwd, fwd;
    std::vector<BlockTrace> t0, t1;
    SearchCounters sc;
    sc.trace[0] = &t0;
    sc.trace[1] = &t1;
    analyse_pair(c, r, o, &hb, &bwd, &hf, &fwd, &sc, false);
    auto payload = pack_vector_frame(o, hb ? &bwd : nullptr, hf ? &fwd : nullptr);
    *out_len = payload.size();
    *n0 = t0.size();
    *n1 = t1.size();
    if (payload.size() > out_cap || t0.size() > cap0 || t1.size() > cap1) return -1;
    std::memcpy(out, payload.data(), payload.size());
    if (!t0.empty()) std::memcpy(trace0, t0.data(), t0.size() * sizeof(BlockTrace));
    if (!t1.empty()) std::memcpy(trace1, t1.data(), t1.size() * sizeof(BlockTrace));
    return 0;
}

// Single candidate cost for focused tests: (luma, chroma) by the reference's path selection.
int oracle_candidate_cost(int width, int height, int pel, int super_levels, const uint8_t* cy, int cys,
                          const uint8_t* cu, int cus, const uint8_t* cv, int cvs, const uint8_t* ry, int rys,
                          const uint8_t* ru, int rus, const uint8_t* rv, int rvs, int level, int px, int py, int bw,
                          int bh, int mvx, int mvy, int satd, int chroma, uint32_t* luma, uint32_t* colour) {
    size_t yh = (size_t)super_plane_height(height, pel, super_levels, true), uvh = yh / 2;
    SuperPlanes c{cy, (size_t)cys * yh, (size_t)cys, cu, (size_t)cus * uvh, (size_t)cus, cv, (size_t)cvs * uvh,
                  (size_t)cvs, (size_t)width, (size_t)height, pel, super_levels, true};
    SuperPlanes r{ry, (size_t)rys * yh, (size_t)rys, ru, (size_t)rus * uvh, (size_t)rus, rv, (size_t)rvs * uvh,
                  (size_t)rvs, (size_t)width, (size_t)height, pel, super_levels, true};
    candidate_cost(c, r, level, px, py, bw, bh, mvx, mvy, satd != 0, level == 0 ? pel : 1, chroma != 0, luma, colour);
    return 0;
}

uint32_t oracle_hadamard4(const int32_t* d16) {
    int32_t d[4][4];
    std::memcpy(d, d16, sizeof d);
    return hadamard4_satd(d);
}

int oracle_level_order(int gw, int gh, int level, int32_t* xy, int cap) {
    auto o = level_search_order(gw, gh, level);
    int n = (int)o.size();
    for (int i = 0; i < n && i < cap; i++) { xy[2 * i] = o[i].first; xy[2 * i + 1] = o[i].second; }
    return n;
}

int oracle_lambdas(int lambda, double plevel, int levels, int level, int pel, int lsad, uint32_t score,
                   int32_t* exact, int32_t* adapted) {
    AnalyseOpts o;
    o.lambda = lambda;
    o.plevel = plevel;
    o.levels = levels;
    *exact = exact_lambda(o, level, pel);
    *adapted = adapt_lambda(*exact, lsad, score);
    return 0;
}

int oracle_global(const int16_t* dxdy, int n, int32_t* gx, int32_t* gy) {
    std::vector<Vec8> f((size_t)n);
    for (int i = 0; i < n; i++) { f[(size_t)i].dx = dxdy[2 * i]; f[(size_t)i].dy = dxdy[2 * i + 1]; }
    int x, y;
    exact_global_doubled(f, &x, &y);
    *gx = x;
    *gy = y;
    return 0;
}

int oracle_rescale(int16_t* dxdy, uint32_t* score, int n, int width, int height, int blk) {
    std::vector<Vec8> f((size_t)n);
    for (int i = 0; i < n; i++) {
        f[(size_t)i].dx = dxdy[2 * i];
        f[(size_t)i].dy = dxdy[2 * i + 1];
        f[(size_t)i].score = score[i];
    }
    rescale_scores(f, width, height, blk);
    for (int i = 0; i < n; i++) score[i] = f[(size_t)i].score;
    return 0;
}

}  // extern "C"

#include "render.h"

extern "C" {

// SmoothFps create-time options: 0 ok, -3 with the reference's error text.
int oracle_smooth_opts(const char* opt, size_t len, int64_t fps_num, int64_t fps_den, const int32_t* vdata16,
                       int32_t nframes, int32_t* out_frames, int32_t* algo, int32_t* scene_mode, char* err,
                       size_t errlen) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    VectorData vd;
    if (!parse_vdata(vdata16, &vd)) return -1;
    SmoothOptions o;
    std::string e;
    if (!smooth_options(have ? &v : nullptr, fps_num, fps_den, vd, &o, &e)) {
        put_err(err, errlen, e);
        return -3;
    }
    *out_frames = o.timing(fps_num, fps_den).scale_frame_count(nframes);
    *algo = (int32_t)o.algo;
    *scene_mode = (int32_t)o.scene_mode;
    return 0;
}

// Render output frame k of a clip held in host memory (planes packed at stride = width).
int oracle_smooth_render(const char* opt, size_t len, int width, int height, int64_t fps_num, int64_t fps_den,
                         int nframes, int64_t sdata, const int32_t* vdata16, const uint8_t* const* src_y,
                         const uint8_t* const* src_u, const uint8_t* const* src_v, const uint8_t* const* sup_y,
                         const uint8_t* const* sup_u, const uint8_t* const* sup_v, int super_rows,
                         const uint8_t* const* vec, size_t vec_len, int k, uint8_t* dy, uint8_t* du, uint8_t* dv,
                         int32_t* algo_out, int float_render) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    SmoothClip c;
    c.width = width;
    c.height = height;
    c.num_frames = nframes;
    c.fps_num = fps_num;
    c.fps_den = fps_den;
    c.sdata = sdata;
    if (!parse_vdata(vdata16, &c.vd)) return -1;
    std::string e;
    if (!smooth_options(have ? &v : nullptr, fps_num, fps_den, c.vd, &c.opt, &e)) return -3;
    for (int i = 0; i < nframes; i++) {
        c.src[0].push_back(src_y[i]); c.src[1].push_back(src_u[i]); c.src[2].push_back(src_v[i]);
        c.sup[0].push_back(sup_y ? sup_y[i] : nullptr); c.sup[1].push_back(sup_u ? sup_u[i] : nullptr);
        c.sup[2].push_back(sup_v ? sup_v[i] : nullptr);
        c.vec.push_back(vec[i]);
    }
    c.super_rows = super_rows;
    c.vec_len = vec_len;
    c.float_render = float_render != 0;
    int px = 0, py = 0;
    light_padding(c.opt.light, width, height, &px, &py);
    int algo = 0;
    if (px == 0 && py == 0) {
        uint8_t* dst[3] = {dy, du, dv};
        int r = smooth_render_frame(c, k, dst, &algo);
        *algo_out = algo;
        return r;
    }
    // padded output (core.rs:560-613, 634-669, 1563-1590): the frame rendered (or
    // copied) at (pad_x, pad_y), then the light border of output frame k around it
    std::vector<uint8_t> ty((size_t)width * height), tu((size_t)(width / 2) * (height / 2)), tv(tu.size());
    uint8_t* tmp[3] = {ty.data(), tu.data(), tv.data()};
    int r = smooth_render_frame(c, k, tmp, &algo);
    *algo_out = algo;
    const int ow = width + 2 * px, oh = height + 2 * py;
    uint8_t* out[3] = {dy, du, dv};
    const int strides[3] = {ow, ow / 2, ow / 2}, rows[3] = {oh, oh / 2, oh / 2};
    for (int p = 0; p < 3; p++) {
        const int w = p ? width / 2 : width, h = p ? height / 2 : height;
        const int ox = p ? px / 2 : px, oy = p ? py / 2 : py;
        for (int y = 0; y < h; y++)
            std::memcpy(out[p] + (size_t)(y + oy) * strides[p] + ox, tmp[p] + (size_t)y * w, (size_t)w);
    }
    light_apply(width, height, px, py, c.opt.light_params, k, out, strides, rows);
    return r;
}

// Options::padding of an option string (0 ok, -3 bad options).
int oracle_smooth_padding(const char* opt, size_t len, int width, int height, int64_t fps_num, int64_t fps_den,
                          const int32_t* vdata16, int32_t* pad_x, int32_t* pad_y) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    VectorData vd;
    if (!parse_vdata(vdata16, &vd)) return -1;
    SmoothOptions o;
    std::string e;
    if (!smooth_options(have ? &v : nullptr, fps_num, fps_den, vd, &o, &e)) return -3;
    light_padding(o.light, width, height, pad_x, pad_y);
    return 0;
}

// The light border's parts, for the tests that pin them (light.rs).
void oracle_light_rng(uint32_t seed, int n, int32_t* out) {
    uint32_t st = seed;
    for (int i = 0; i < n; i++) {
        st = st * 214013u + 2531011u;
        out[i] = (int32_t)((st >> 16) & 0x7fff);
    }
}
void oracle_light_profile(const int64_t* accum, int dim, int border, int lights, double cell, int base, float* out) {
    std::vector<int64_t> a(accum, accum + std::max(dim, 0));
    std::vector<float> r = light_profile(a, dim, border, lights, cell, base);
    std::copy(r.begin(), r.end(), out);
}
void oracle_light_cubic(const int32_t* control, int lights, int dim, float* out) {
    std::vector<int> c(control, control + lights);
    std::vector<float> r = light_cubic_upsample(c, dim, lights);
    std::copy(r.begin(), r.end(), out);
}
void oracle_light_corner_geometry(int width, int height, int length, double* mix, double* weight) {
    std::vector<double> m, w;
    light_corner_geometry(width, height, length, &m, &w);
    std::copy(m.begin(), m.end(), mix);
    std::copy(w.begin(), w.end(), weight);
}
// The border of output frame `frame` around an interior already in the padded planes
// (strides width + 2 pad_x, half for chroma).
void oracle_light_apply(int width, int height, int pad_x, int pad_y, int border, int lights, int length, double cell,
                        int frame, uint8_t* y, uint8_t* u, uint8_t* v) {
    LightParams lp;
    lp.border = border;
    lp.lights = std::max(lights, 1);
    lp.length = length;
    lp.cell = std::max(cell, 0.1);
    const int ow = width + 2 * pad_x, oh = height + 2 * pad_y;
    uint8_t* planes[3] = {y, u, v};
    const int strides[3] = {ow, ow / 2, ow / 2}, rows[3] = {oh, oh / 2, oh / 2};
    light_apply(width, height, pad_x, pad_y, lp, frame, planes, strides, rows);
}

}  // extern "C"

extern "C" {

int oracle_truncdiv(int divisor, int value) { return truncdiv(divisor, value); }

// Vector sets as (dx, dy, magnitude) triples, grid_w * grid_h each.
static std::vector<RVector> rvecs(const int32_t* t, int n) {
    std::vector<RVector> v((size_t)n);
    for (int i = 0; i < n; i++) v[(size_t)i] = RVector{(int16_t)t[3 * i], (int16_t)t[3 * i + 1], (uint32_t)t[3 * i + 2]};
    return v;
}

// coverage_mask (renderer.rs:1738-1763) of set `which` (it splats the opposite set).
int oracle_coverage(int block_w, int block_h, int scale_shift_base, int origin_x, int origin_y, int grid_w, int grid_h,
                    const int32_t* a, const int32_t* b, int which, int width, int height, int strength, int threshold,
                    uint8_t* out) {
    auto va = rvecs(a, grid_w * grid_h), vb = rvecs(b, grid_w * grid_h);
    VectorContext ctx{block_w, block_h, scale_shift_base, 0, 0, origin_x, origin_y, (size_t)grid_w, (size_t)grid_h,
                      false, &va, &vb};
    std::vector<uint8_t> m((size_t)width * height);
    coverage_mask(ctx, which, m, (size_t)width, (size_t)height, strength, threshold);
    std::memcpy(out, m.data(), m.size());
    return 0;
}

// magnitude_mask (renderer.rs:1647-1678) of set `which`.
int oracle_magnitude(int block_w, int block_h, int grid_w, int grid_h, const int32_t* a, const int32_t* b, int which,
                     int width, int height, double scale, double exponent, uint8_t* out) {
    auto va = rvecs(a, grid_w * grid_h), vb = rvecs(b, grid_w * grid_h);
    VectorContext ctx{block_w, block_h, 1, 0, 0, 0, 0, (size_t)grid_w, (size_t)grid_h, false, &va, &vb};
    std::vector<uint8_t> m((size_t)width * height);
    magnitude_mask(ctx, which, m, (size_t)width, (size_t)height, scale, exponent);
    std::memcpy(out, m.data(), m.size());
    return 0;
}

int oracle_luma_byte(double v, double gamma) { return luma_byte(v, gamma); }

// analyze_scene over decoded vectors given as (score, luma) pairs; prev may be null.
int oracle_scene_class(const int32_t* prev, const int32_t* cur, int grid_w, int grid_h, int marker, double gamma,
                       int64_t zero, int64_t m1, int64_t m2, int64_t scene, int64_t blocks) {
    const int n = grid_w * grid_h;
    std::vector<DVec> p((size_t)n), c((size_t)n);
    for (int i = 0; i < n; i++) {
        c[(size_t)i] = DVec{0, 0, (uint32_t)cur[2 * i], (uint8_t)cur[2 * i + 1]};
        if (prev) p[(size_t)i] = DVec{0, 0, (uint32_t)prev[2 * i], (uint8_t)prev[2 * i + 1]};
    }
    VectorData vd{};
    vd.grid_w = grid_w;
    vd.grid_h = grid_h;
    vd.marker = marker;
    SmoothOptions o;
    o.scene_luma = gamma;
    o.lim_zero = zero; o.lim_m1 = m1; o.lim_m2 = m2; o.lim_scene = scene; o.lim_blocks = blocks;
    return scene_class(prev ? &p : nullptr, &c, vd, o);
}

}  // extern "C"
