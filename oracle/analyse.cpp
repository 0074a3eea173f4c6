// Oracle restatement of crates/svpflow1/src/analyse_opts.rs and analyse_search.rs.
// Test infrastructure only (see oracle.h).  Semantics follow the reference's x86-64
// build: where the SIMD kernels compute the same value as the scalar code they are
// folded into one scalar routine; where they differ (chroma cost: SATD only on the
// interior SIMD path, SAD on the clamped fallback) the x86-64 behaviour is kept.
#include "analyse.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace oracle {

namespace {

inline int32_t sat_mul_i32(int32_t a, int32_t b) {
    int64_t r = (int64_t)a * b;
    return (int32_t)std::min<int64_t>(std::max<int64_t>(r, INT32_MIN), INT32_MAX);
}
inline int32_t sat_add_i32(int32_t a, int32_t b) {
    int64_t r = (int64_t)a + b;
    return (int32_t)std::min<int64_t>(std::max<int64_t>(r, INT32_MIN), INT32_MAX);
}
inline uint32_t sat_add_u32(uint32_t a, uint32_t b) {
    uint64_t r = (uint64_t)a + b;
    return r > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)r;
}
inline uint32_t sat_mul_u32(uint32_t a, uint32_t b) {
    uint64_t r = (uint64_t)a * b;
    return r > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)r;
}
// Rust `f64 as i32`: saturating, NaN -> 0.
inline int32_t f2i(double v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline uint8_t at(const uint8_t* p, size_t len, size_t i, uint8_t fb) { return i < len ? p[i] : fb; }

// compiler-rt __powidf2: what Rust's f64::powi lowers to at run time.
double powi(double a, int b) {
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

struct Best {  // SearchBest, analyse_search.rs:228-233
    int mvx, mvy;
    uint32_t sad;
    int64_t cost;
};

}  // namespace

// ------------------------------------------------------------------ analyse_opts.rs

void AnalyseOpts::output_block(int* bw, int* bh, int* ox, int* oy) const {  // :221-232
    if (has_refine) {
        *bw = std::max(block_w / 2, 4);
        *bh = std::max(block_h / 2, 4);
        *ox = std::max(overlap_x / 2, 0);
        *oy = std::max(overlap_y / 2, 0);
    } else {
        *bw = block_w; *bh = block_h; *ox = overlap_x; *oy = overlap_y;
    }
}

void AnalyseOpts::grid(int bw, int bh, int ox, int oy, int* gw, int* gh) const {  // :234-240
    int sx = std::max(bw - ox, 1), sy = std::max(bh - oy, 1);
    *gw = std::max((width - ox) / sx, 1);
    *gh = std::max((height - oy) / sy, 1);
}

bool overlap_from_mode(int bw, int bh, int mode, int* ox, int* oy) {  // :259-269
    switch (mode) {
        case 0: *ox = 0; *oy = 0; return true;
        case 1: *ox = bw / 8; *oy = bh / 8; return true;
        case 2: *ox = bw / 4; *oy = bh / 4; return true;
        case 3: *ox = bw / 2; *oy = bh / 2; return true;
        default: return false;
    }
}

int auto_search_levels(int w, int h, int bw, int bh, int ox, int oy) {  // :271-295
    int sx = std::max(bw - ox, 1), sy = std::max(bh - oy, 1);
    int gw = std::max((w - ox) / sx, 0), gh = std::max((h - oy) / sy, 0);
    if (gw <= 0 || gh <= 0) return 1;
    int end_x = ox + sx * gw, end_y = oy + sy * gh;
    int levels = 0;
    for (;;) {
        int gwl = ((end_x >> levels) - ox) / sx, ghl = ((end_y >> levels) - oy) / sy;
        if (gwl <= 0 || ghl <= 0) break;
        levels++;
        if (levels > 12) break;
    }
    return std::max(levels, 1);
}

std::string analyse_opts_from(const Value* opt, int64_t sdata, const SuperOpts* hint, AnalyseOpts* o) {
    // unpack_sdata, analyse_opts.rs:243-257
    uint64_t raw;
    std::memcpy(&raw, &sdata, 8);
    int height = (int)(raw & 0xFFFF), width = (int)((raw >> 16) & 0xFFFF);
    int pel = (int)((raw >> 32) & 0xFF), slevels = (int)((raw >> 40) & 0xFF), selector = (int)((raw >> 48) & 0xFF);
    if (width <= 0 || height <= 0 || pel < 1 || pel > 4 || slevels <= 0) {
        if (!hint) return "SVAnalyse: invalid super clip params";
        width = hint->width; height = hint->height; pel = hint->pel; slevels = hint->levels; selector = hint->gpu;
    }
    int64_t iv;
    double fv;
    bool bv;
    auto I = [&](std::initializer_list<const char*> p, int64_t* v) { return opt && opt->int_at(p, v); };
    auto F = [&](std::initializer_list<const char*> p, double* v) { return opt && opt->float_at(p, v); };
    auto B = [&](std::initializer_list<const char*> p, bool* v) { return opt && opt->bool_at(p, v); };
    // main.X or top-level X (analyse_opts.rs:75-107)
    auto I2 = [&](std::initializer_list<const char*> a, std::initializer_list<const char*> b, int64_t dflt) {
        int64_t v;
        if (I(a, &v)) return v;
        if (I(b, &v)) return v;
        return dflt;
    };
    o->vectors = I({"vectors"}, &iv) ? (int)iv : 3;
    if (o->vectors < 1 || o->vectors > 3) return "SVAnalyse: vectors must be 1, 2 or 3";
    o->block_w = I({"block", "w"}, &iv) ? (int)iv : 16;
    o->block_h = I({"block", "h"}, &iv) ? (int)iv : o->block_w;
    if (!(o->block_w == 4 || o->block_w == 8 || o->block_w == 16 || o->block_w == 32))
        return "SVAnalyse: block.w must be 4/8/16/32";
    o->overlap_mode = I({"block", "overlap"}, &iv) ? (int)iv : 2;
    if (!overlap_from_mode(o->block_w, o->block_h, o->overlap_mode, &o->overlap_x, &o->overlap_y))
        return "SVAnalyse: overlap must be 0, 1 (=1/8*block), 2 (=1/4*block) or 3 (=1/2*block)";
    o->delta = (int)I2({"special", "delta"}, {"delta"}, 1);
    double lambda = 10.0;
    if (F({"main", "penalty", "lambda"}, &fv) || F({"penalty", "lambda"}, &fv)) lambda = fv;
    int lsad = (int)I2({"main", "penalty", "lsad"}, {"penalty", "lsad"}, 8000);
    double plevel = 1.5;
    if (F({"main", "penalty", "plevel"}, &fv) || F({"penalty", "plevel"}, &fv)) plevel = fv;
    o->plevel = std::fmax(plevel, 0.01);  // f64::max
    o->pnew = (int)I2({"main", "penalty", "pnew"}, {"penalty", "pnew"}, 50);
    o->pglobal = (int)I2({"main", "penalty", "pglobal"}, {"penalty", "pglobal"}, 50);
    o->pzero = (int)I2({"main", "penalty", "pzero"}, {"penalty", "pzero"}, 100);
    o->pnbour = (int)I2({"main", "penalty", "pnbour"}, {"penalty", "pnbour"}, 50);
    o->prev = (int)I2({"main", "penalty", "prev"}, {"penalty", "prev"}, 0);
    o->search_type = (int)I2({"main", "search", "type"}, {"search", "type"}, 4);
    o->distance = (int)I2({"main", "search", "distance"}, {"search", "distance"}, 2);
    o->coarse_search_type = (int)I2({"main", "search", "coarse", "type"}, {"search", "coarse", "type"}, o->search_type);
    o->coarse_distance = (int)I2({"main", "search", "coarse", "distance"}, {"search", "coarse", "distance"}, 0);
    o->coarse_trymany = false;
    if (B({"main", "search", "coarse", "trymany"}, &bv) || B({"search", "coarse", "trymany"}, &bv)) o->coarse_trymany = bv;
    int coarse_bad_sad = (int)I2({"main", "search", "coarse", "bad", "sad"}, {"search", "coarse", "bad", "sad"}, 1000);
    o->coarse_bad_range = I({"main", "search", "coarse", "bad", "range"}, &iv) ? (int)iv : -24;
    o->fine_satd = false;
    if (B({"main", "search", "satd"}, &bv) || B({"search", "satd"}, &bv)) o->fine_satd = bv;
    o->coarse_satd = B({"main", "search", "coarse", "satd"}, &bv) ? bv : true;
    o->sort = true;
    if (B({"main", "search", "sort"}, &bv) || B({"sort"}, &bv)) o->sort = bv;
    // refine[0] (analyse_opts.rs:150-164)
    const std::vector<Value>* refine_arr = opt ? opt->array_at({"refine"}) : nullptr;
    const Value* refine = (refine_arr && !refine_arr->empty()) ? &(*refine_arr)[0] : nullptr;
    o->has_refine = false;
    if (refine && refine->int_at({"thsad"}, &iv)) {
        o->has_refine = true;
        // `x as i32 * max((bw/2)*(bh/2), 1) / 32` in wrapping i32 arithmetic
        int32_t x = (int32_t)iv;
        int32_t m = std::max((o->block_w / 2) * (o->block_h / 2), 1);
        o->refine_thsad = (int32_t)((uint32_t)x * (uint32_t)m) / 32;
    }
    o->refine_search_type = (refine && refine->int_at({"search", "type"}, &iv)) ? (int)iv : o->search_type;
    o->refine_distance = (refine && refine->int_at({"search", "distance"}, &iv)) ? (int)iv : o->distance;
    o->refine_satd = (refine && refine->bool_at({"search", "satd"}, &bv)) ? bv : o->fine_satd;
    int levels = std::max(std::min(auto_search_levels(width, height, o->block_w, o->block_h, o->overlap_x, o->overlap_y),
                                   slevels), 1);
    if (I({"main", "levels"}, &iv) && iv > 0) levels = std::max(std::min((int)iv, slevels), 1);
    double area = (double)std::max(o->block_w * o->block_h, 1);
    o->lambda = f2i(2.0 * area * 1000.0 * 0.015625 * lambda);
    int area_scale = std::max(o->block_w * o->block_h, 1) / 32;
    o->lsad = sat_mul_i32(lsad, area_scale);
    o->coarse_bad_sad = sat_mul_i32(coarse_bad_sad, area_scale);
    o->pel = pel;
    o->selector = selector;
    o->levels = levels;
    o->super_levels = slevels;
    o->width = width;
    o->height = height;
    return "";
}

// ------------------------------------------------------------------ planes

size_t SuperPlanes::level_y_offset(int level) const {  // analyse_search.rs:49-58
    size_t y = 0;
    for (int lv = 0; lv < level; lv++) {
        size_t h = (size_t)reduce_dim((int)luma_h, lv);
        y += h * ((lv == 0 && full) ? (size_t)(pel * pel) : 1);
    }
    return y;
}

void SuperPlanes::level_size(int level, size_t* w, size_t* h) const {  // :60-65
    *w = (size_t)reduce_dim((int)luma_w, level);
    *h = (size_t)reduce_dim((int)luma_h, level);
}

static size_t chroma_level_offset(const SuperPlanes& s, int level) {  // :2704-2713
    size_t y = 0;
    for (int lv = 0; lv < level; lv++) {
        size_t h = (size_t)reduce_dim((int)s.luma_h, lv) / 2;
        y += std::max<size_t>(h, 1) * ((lv == 0 && s.full) ? (size_t)(s.pel * s.pel) : 1);
    }
    return y;
}

// ------------------------------------------------------------------ cost kernels

uint32_t hadamard4_satd(const int32_t d[4][4]) {  // :2681-2702
    int32_t t[4][4];
    for (int i = 0; i < 4; i++) {
        int32_t s0 = d[i][0] + d[i][1], s1 = d[i][0] - d[i][1], s2 = d[i][2] + d[i][3], s3 = d[i][2] - d[i][3];
        t[i][0] = s0 + s2; t[i][1] = s1 + s3; t[i][2] = s0 - s2; t[i][3] = s1 - s3;
    }
    int32_t s = 0;
    for (int j = 0; j < 4; j++) {
        int32_t s0 = t[0][j] + t[1][j], s1 = t[0][j] - t[1][j], s2 = t[2][j] + t[3][j], s3 = t[2][j] - t[3][j];
        s += std::abs(s0 + s2) + std::abs(s1 + s3) + std::abs(s0 - s2) + std::abs(s1 - s3);
    }
    return (uint32_t)std::max(s, 0);
}

static void mv_bounds(int px, int py, int bw, int bh, int lw, int lh, int pel, int b[4]) {  // :1620-1636
    pel = std::max(pel, 1);
    b[0] = -pel * px;
    b[1] = -pel * py;
    b[2] = pel * (lw - px - bw) + 1;
    b[3] = pel * (lh - py - bh) + 1;
}

static bool exact_ok(int mx, int my, const int b[4]) { return mx >= b[0] && my >= b[1] && mx < b[2] && my < b[3]; }

// :1638-1676
static bool edge_shift_origins(int px, int py, int mvx, int mvy, int bw, int bh, int lw, int lh, int pel, int* cx,
                               int* cy, int* rx, int* ry) {
    pel = std::max(pel, 1);
    int b[4];
    mv_bounds(px, py, bw, bh, lw, lh, pel, b);
    int ox = mvx < b[0] ? mvx - b[0] : ((b[2] > b[0] && mvx >= b[2]) ? mvx + 1 - b[2] : 0);
    int oy = mvy < b[1] ? mvy - b[1] : ((b[3] > b[1] && mvy >= b[3]) ? mvy + 1 - b[3] : 0);
    if (ox == 0 && oy == 0) return false;
    *cx = clampi(px - ox, 0, std::max(lw - bw, 0));
    *cy = clampi(py - oy, 0, std::max(lh - bh, 0));
    if (pel == 1) { *rx = *cx + mvx; *ry = *cy + mvy; }
    else if (pel == 4) { *rx = (4 * *cx + mvx) >> 2; *ry = (4 * *cy + mvy) >> 2; }
    else { *rx = (2 * *cx + mvx) >> 1; *ry = (2 * *cy + mvy) >> 1; }
    return true;
}

// satd_edge, :2305-2351: every coordinate clamped, buffer misses read 0.
static uint32_t satd_edge(const uint8_t* src, size_t slen, size_t sstride, size_t srow, int sx0, int sy0,
                          const uint8_t* ref, size_t rlen, size_t rstride, size_t rrow, int rx0, int ry0, size_t width,
                          size_t height, size_t n) {
    int w = (int)width, h = (int)height;
    uint32_t total = 0;
    size_t tiles = n / 4;
    for (size_t ty = 0; ty < tiles; ty++)
        for (size_t tx = 0; tx < tiles; tx++) {
            int32_t d[4][4];
            for (int r = 0; r < 4; r++) {
                size_t sy = (size_t)clampi(sy0 + (int)(ty * 4 + r), 0, std::max(h - 1, 0));
                size_t ry = (size_t)clampi(ry0 + (int)(ty * 4 + r), 0, std::max(h - 1, 0));
                for (int c = 0; c < 4; c++) {
                    size_t sx = (size_t)clampi(sx0 + (int)(tx * 4 + c), 0, std::max(w - 1, 0));
                    size_t rx = (size_t)clampi(rx0 + (int)(tx * 4 + c), 0, std::max(w - 1, 0));
                    d[r][c] = (int32_t)at(src, slen, (srow + sy) * sstride + sx, 0) -
                              (int32_t)at(ref, rlen, (rrow + ry) * rstride + rx, 0);
                }
            }
            total = sat_add_u32(total, hadamard4_satd(d));
        }
    return total >> 1;
}

// block_cost_edge_lc, :1996-2167 (luma + 4x chroma, chroma always evaluated here).
static bool block_cost_edge_lc(const SuperPlanes& src, const SuperPlanes& ref, int level, int px, int py, int bw,
                               int bh, int mvx, int mvy, bool use_satd, int pel, uint32_t* l_out, uint32_t* c_out) {
    size_t lw, lh;
    src.level_size(level, &lw, &lh);
    int cx, cy, rx0, ry0;
    if (!edge_shift_origins(px, py, mvx, mvy, bw, bh, (int)lw, (int)lh, pel, &cx, &cy, &rx0, &ry0)) return false;
    size_t y_off = src.level_y_offset(level);
    int sub_x = 0, sub_y = 0;
    if (level == 0 && pel >= 2) { sub_x = mvx & 1; sub_y = mvy & 1; }
    size_t sub_idx = (size_t)(sub_y * pel + sub_x);
    size_t cur_row0 = y_off, ref_row0 = y_off + sub_idx * lh;
    size_t bwu = (size_t)bw, bhu = (size_t)bh;
    int lwi = (int)lw, lhi = (int)lh;
    auto pow2 = [](size_t v) { return v == 4 || v == 8 || v == 16 || v == 32; };
    uint32_t sad_l;
    if (use_satd && pow2(bwu) && pow2(bhu) && bwu == bhu) {
        sad_l = satd_edge(src.y, src.y_len, src.y_stride, cur_row0, cx, cy, ref.y, ref.y_len, ref.y_stride, ref_row0,
                          rx0, ry0, lw, lh, bwu);
    } else {
        uint32_t sad = 0;
        for (size_t row = 0; row < bhu; row++) {
            size_t sy = (size_t)clampi(cy + (int)row, 0, std::max(lhi - 1, 0));
            size_t ry = (size_t)clampi(ry0 + (int)row, 0, std::max(lhi - 1, 0));
            for (size_t col = 0; col < bwu; col++) {
                size_t sx = (size_t)clampi(cx + (int)col, 0, std::max(lwi - 1, 0));
                size_t rx = (size_t)clampi(rx0 + (int)col, 0, std::max(lwi - 1, 0));
                int a = at(src.y, src.y_len, (cur_row0 + sy) * src.y_stride + sx, 0);
                int b = at(ref.y, ref.y_len, (ref_row0 + ry) * ref.y_stride + rx, 0);
                sad = sat_add_u32(sad, (uint32_t)std::abs(a - b));
            }
        }
        sad_l = sad;
    }
    size_t cw = bwu / 2, ch = bhu / 2, clw = lw / 2, clh = lh / 2;
    int ccx = (px >> 1) - ((px - cx) >> 1), ccy = (py >> 1) - ((py - cy) >> 1);  // :1679-1681
    int crx0 = ((mvx >> 1) + 2 * ccx) >> 1, cry0 = ((mvy >> 1) + 2 * ccy) >> 1;
    size_t c_off = chroma_level_offset(src, level);
    size_t c_sub = (level == 0 && pel >= 2) ? (size_t)((mvy & 1) * pel + (mvx & 1)) : 0;
    size_t c_ref = c_off + c_sub * clh;
    int clwi = (int)clw, clhi = (int)clh;
    uint32_t sad_uv;
    if (use_satd && cw == ch && (cw == 4 || cw == 8 || cw == 16)) {
        sad_uv = sat_add_u32(satd_edge(src.u, src.u_len, src.u_stride, c_off, ccx, ccy, ref.u, ref.u_len, ref.u_stride,
                                       c_ref, crx0, cry0, clw, clh, cw),
                             satd_edge(src.v, src.v_len, src.v_stride, c_off, ccx, ccy, ref.v, ref.v_len, ref.v_stride,
                                       c_ref, crx0, cry0, clw, clh, cw));
    } else {
        uint32_t sad = 0;
        for (size_t row = 0; row < ch; row++) {
            size_t sy = (size_t)clampi(ccy + (int)row, 0, std::max(clhi - 1, 0));
            size_t ry = (size_t)clampi(cry0 + (int)row, 0, std::max(clhi - 1, 0));
            for (size_t col = 0; col < cw; col++) {
                size_t sx = (size_t)clampi(ccx + (int)col, 0, std::max(clwi - 1, 0));
                size_t rx = (size_t)clampi(crx0 + (int)col, 0, std::max(clwi - 1, 0));
                int au = at(src.u, src.u_len, (c_off + sy) * src.u_stride + sx, 128);
                int bu = at(ref.u, ref.u_len, (c_ref + ry) * ref.u_stride + rx, 128);
                int av = at(src.v, src.v_len, (c_off + sy) * src.v_stride + sx, 128);
                int bv = at(ref.v, ref.v_len, (c_ref + ry) * ref.v_stride + rx, 128);
                sad = sat_add_u32(sat_add_u32(sad, (uint32_t)std::abs(au - bu)), (uint32_t)std::abs(av - bv));
            }
        }
        sad_uv = sad;
    }
    *l_out = sad_l;
    *c_out = sat_mul_u32(sad_uv, 4);
    return true;
}

// block_cost_luma_interior, :1719-1852.  The interior SIMD paths (sad_u8_simd, satd_interior)
// read exactly the bytes the clamped loops read when the block is interior, so the
// clamped loops below give the reference's value in every case.
static uint32_t block_cost_luma_interior(const SuperPlanes& src, const SuperPlanes& ref, int level, int px, int py,
                                         size_t bw, size_t bh, int mvx, int mvy, bool use_satd, int pel) {
    size_t lw, lh;
    src.level_size(level, &lw, &lh);
    px -= (level > 0 && px + (int)bw >= (int)lw) ? 1 : 0;
    py -= (level > 0 && py + (int)bh >= (int)lh) ? 1 : 0;
    size_t pxu = (size_t)(int64_t)px, pyu = (size_t)(int64_t)py;  // `as usize` (wraps for -1)
    size_t y_off = src.level_y_offset(level);
    int sub_x = 0, sub_y = 0, mvxp = mvx, mvyp = mvy;
    if (level == 0 && pel >= 2) { sub_x = mvx & 1; sub_y = mvy & 1; mvxp = mvx >> 1; mvyp = mvy >> 1; }
    size_t sub_idx = (size_t)(sub_y * pel + sub_x);
    size_t cur_row0 = y_off, ref_row0 = y_off + sub_idx * lh;
    auto pow2 = [](size_t v) { return v == 4 || v == 8 || v == 16 || v == 32; };
    if (use_satd && pow2(bw) && pow2(bh) && bw == bh) {
        // satd_luma_clamp, :2353-2436 (source outside the level reads 0, reference clamped)
        size_t n = bw;
        uint32_t total = 0;
        for (size_t ty = 0; ty < n / 4; ty++)
            for (size_t tx = 0; tx < n / 4; tx++) {
                int32_t d[4][4];
                for (size_t r = 0; r < 4; r++) {
                    size_t sy = pyu + ty * 4 + r;
                    size_t ry = (size_t)clampi((int)(int64_t)pyu + (int)(ty * 4 + r) + mvyp, 0, (int)lh - 1);
                    for (size_t c = 0; c < 4; c++) {
                        size_t sx = pxu + tx * 4 + c;
                        size_t rx = (size_t)clampi((int)(int64_t)pxu + (int)(tx * 4 + c) + mvxp, 0, (int)lw - 1);
                        int a = (sy < lh && sx < lw) ? at(src.y, src.y_len, (cur_row0 + sy) * src.y_stride + sx, 0) : 0;
                        int b = at(ref.y, ref.y_len, (ref_row0 + ry) * ref.y_stride + rx, 0);
                        d[r][c] = a - b;
                    }
                }
                total = sat_add_u32(total, hadamard4_satd(d));
            }
        return total >> 1;
    }
    uint32_t sad = 0;
    for (size_t row = 0; row < bh; row++) {
        size_t sy = pyu + row;
        size_t ry = (size_t)clampi(py + (int)row + mvyp, 0, (int)lh - 1);
        if (sy >= lh) break;
        for (size_t col = 0; col < bw; col++) {
            size_t sx = pxu + col;
            size_t rx = (size_t)clampi(px + (int)col + mvxp, 0, (int)lw - 1);
            if (sx >= lw) break;
            int a = at(src.y, src.y_len, (cur_row0 + sy) * src.y_stride + sx, 0);
            int b = at(ref.y, ref.y_len, (ref_row0 + ry) * ref.y_stride + rx, 0);
            sad = sat_add_u32(sad, (uint32_t)std::abs(a - b));
        }
    }
    return sad;
}

// chroma_sad_x4, :2169-2303 (x86-64 behaviour: the interior path uses SATD when asked,
// the clamped fallback is always SAD with out-of-slice reads of 128).
static uint32_t chroma_sad_x4(const SuperPlanes& src, const SuperPlanes& ref, int level, int pel, int mvx, int mvy,
                              size_t px, size_t py, size_t bw, size_t bh, size_t lw, size_t lh, bool use_satd) {
    size_t cw = bw / 2, ch = bh / 2;
    if (cw < 4 || ch < 4) return 0;
    size_t cpx = px / 2, cpy = py / 2, clw = lw / 2, clh = lh / 2;
    size_t c_off = chroma_level_offset(src, level);
    size_t c_sub_idx = 0;
    int mv_cx, mv_cy;
    if (level == 0 && pel >= 2) {
        c_sub_idx = (size_t)((mvy & 1) * pel + (mvx & 1));
        mv_cx = mvx >> 2;
        mv_cy = mvy >> 2;
    } else {
        mv_cx = mvx >> 1;
        mv_cy = mvy >> 1;
    }
    size_t c_cur = c_off;
    size_t c_ref = c_off + ((level == 0 && pel > 1) ? c_sub_idx * clh : 0);
    if (cw == 4 || cw == 8 || cw == 16) {
        int rx0 = (int)cpx + mv_cx, ry0 = (int)cpy + mv_cy;
        if (cpx + cw <= clw && cpy + ch <= clh && rx0 >= 0 && ry0 >= 0 && (size_t)rx0 + cw <= clw &&
            (size_t)ry0 + ch <= clh) {
            size_t su = (c_cur + cpy) * src.u_stride + cpx, sv = (c_cur + cpy) * src.v_stride + cpx;
            size_t ru = (c_ref + (size_t)ry0) * ref.u_stride + (size_t)rx0;
            size_t rv = (c_ref + (size_t)ry0) * ref.v_stride + (size_t)rx0;
            if (su + (ch - 1) * src.u_stride + cw <= src.u_len && sv + (ch - 1) * src.v_stride + cw <= src.v_len &&
                ru + (ch - 1) * ref.u_stride + cw <= ref.u_len && rv + (ch - 1) * ref.v_stride + cw <= ref.v_len) {
                auto cost = [&](const uint8_t* a, size_t as, const uint8_t* b, size_t bs) -> uint32_t {
                    if (use_satd) {  // satd_interior over cw x cw (reads cw rows even if ch < cw)
                        uint32_t total = 0;
                        for (size_t ty = 0; ty < cw / 4; ty++)
                            for (size_t tx = 0; tx < cw / 4; tx++) {
                                int32_t d[4][4];
                                for (size_t r = 0; r < 4; r++)
                                    for (size_t c = 0; c < 4; c++)
                                        d[r][c] = (int32_t)a[(ty * 4 + r) * as + tx * 4 + c] -
                                                  (int32_t)b[(ty * 4 + r) * bs + tx * 4 + c];
                                total = sat_add_u32(total, hadamard4_satd(d));
                            }
                        return total >> 1;
                    }
                    uint32_t s = 0;
                    for (size_t r = 0; r < ch; r++)
                        for (size_t c = 0; c < cw; c++) s += (uint32_t)std::abs((int)a[r * as + c] - (int)b[r * bs + c]);
                    return s;
                };
                return sat_mul_u32(sat_add_u32(cost(src.u + su, src.u_stride, ref.u + ru, ref.u_stride),
                                               cost(src.v + sv, src.v_stride, ref.v + rv, ref.v_stride)),
                                   4);
            }
        }
    }
    uint32_t sad_uv = 0;
    for (size_t row = 0; row < ch; row++) {
        size_t sy = cpy + row;
        size_t ry = (size_t)clampi((int)cpy + (int)row + mv_cy, 0, (int)clh - 1);
        if (sy >= clh) break;
        for (size_t col = 0; col < cw; col++) {
            size_t sx = cpx + col;
            size_t rx = (size_t)clampi((int)cpx + (int)col + mv_cx, 0, (int)clw - 1);
            if (sx >= clw) break;
            int au = at(src.u, src.u_len, (c_cur + sy) * src.u_stride + sx, 128);
            int av = at(src.v, src.v_len, (c_cur + sy) * src.v_stride + sx, 128);
            int bu = at(ref.u, ref.u_len, (c_ref + ry) * ref.u_stride + rx, 128);
            int bv = at(ref.v, ref.v_len, (c_ref + ry) * ref.v_stride + rx, 128);
            sad_uv = sat_add_u32(sat_add_u32(sad_uv, (uint32_t)std::abs(au - bu)), (uint32_t)std::abs(av - bv));
        }
    }
    return sat_mul_u32(sad_uv, 4);
}

// Path selection shared by exact_sad (:1223-1261) and exact_sad_if_better (:1158-1221):
// an mv outside mv_bounds takes block_cost_edge_lc, anything else the interior luma
// kernel plus chroma_sad_x4; chroma is dropped when `chroma` is false.
void candidate_cost(const SuperPlanes& src, const SuperPlanes& ref, int level, int px, int py, int bw, int bh,
                    int mvx, int mvy, bool satd, int pel, bool chroma, uint32_t* luma, uint32_t* colour) {
    pel = std::max(pel, 1);
    uint32_t l, c;
    if (block_cost_edge_lc(src, ref, level, px, py, bw, bh, mvx, mvy, satd, pel, &l, &c)) {
        *luma = l;
        *colour = chroma ? c : 0;
        return;
    }
    *luma = block_cost_luma_interior(src, ref, level, px, py, (size_t)bw, (size_t)bh, mvx, mvy, satd, pel);
    if (!chroma) { *colour = 0; return; }
    size_t lw, lh;
    src.level_size(level, &lw, &lh);
    *colour = chroma_sad_x4(src, ref, level, pel, mvx, mvy, (size_t)(int64_t)px, (size_t)(int64_t)py, (size_t)bw,
                            (size_t)bh, lw, lh, satd);
}

// ------------------------------------------------------------------ search

namespace {

struct Level {
    const SuperPlanes& src;
    const SuperPlanes& ref;
    int level, px, py, bw, bh, pel;
    bool satd, chroma;
    int pred_x, pred_y;  // motion-penalty predictor
    int bounds[4];
    int lambda, pnew;
    SearchCounters* counters;

    void count() const {
        if (!counters) return;
        int li = std::min(level, 15);
        counters->evals[li]++;
        counters->pixels[li] += (uint64_t)bw * bh + (chroma ? 2ull * (bw / 2) * (bh / 2) : 0);
    }
    uint32_t exact_sad(int mx, int my) const {  // :1223-1261
        count();
        uint32_t l, c;
        candidate_cost(src, ref, level, px, py, bw, bh, mx, my, satd, pel, chroma, &l, &c);
        return sat_add_u32(l, c);
    }
    int64_t motion_penalty(int mx, int my) const {  // :1263-1267
        int64_t dx = pred_x - mx, dy = pred_y - my;
        return ((int64_t)lambda * (dx * dx + dy * dy)) >> 8;
    }
    // exact_sad_if_better (:1158-1221); returns false when pruned.
    bool sad_if_better(int mx, int my, int64_t motion, int64_t best_cost, uint32_t* sad, int64_t* cost) const {
        int pelm = std::max(pel, 1);
        size_t lw, lh;
        src.level_size(level, &lw, &lh);
        int cx, cy, rx, ry;
        bool shifted = edge_shift_origins(px, py, mx, my, bw, bh, (int)lw, (int)lh, pelm, &cx, &cy, &rx, &ry);
        auto pen = [&](uint32_t v) { return (int64_t)v + (((int64_t)pnew * (int64_t)v) >> 8); };
        count();
        if (shifted || pnew < 0) {
            uint32_t l, c;
            candidate_cost(src, ref, level, px, py, bw, bh, mx, my, satd, pelm, chroma, &l, &c);
            *sad = sat_add_u32(l, c);
            *cost = motion + pen(l) + pen(c);
            return true;
        }
        uint32_t l = block_cost_luma_interior(src, ref, level, px, py, (size_t)bw, (size_t)bh, mx, my, satd, pelm);
        int64_t lower = motion + (int64_t)l + (((int64_t)pnew * (int64_t)l) >> 8);
        if (lower >= best_cost) return false;
        if (!chroma) { *sad = l; *cost = lower; return true; }
        uint32_t c = chroma_sad_x4(src, ref, level, pelm, mx, my, (size_t)(int64_t)px, (size_t)(int64_t)py, (size_t)bw,
                                   (size_t)bh, lw, lh, satd);
        *sad = sat_add_u32(l, c);
        *cost = lower + (int64_t)c + (((int64_t)pnew * (int64_t)c) >> 8);
        return true;
    }
    void check(int mx, int my, Best* best) const {  // exact_check, :1122-1155
        if (!exact_ok(mx, my, bounds)) return;
        int64_t motion = motion_penalty(mx, my);
        if (motion >= best->cost) return;
        uint32_t sad;
        int64_t cost;
        if (!sad_if_better(mx, my, motion, best->cost, &sad, &cost)) return;
        if (cost < best->cost) *best = Best{mx, my, sad, cost};
    }
    void check_dir(int mx, int my, int label, int* dir, Best* best) const {  // exact_check_dir, :902-939
        if (!exact_ok(mx, my, bounds)) return;
        int64_t motion = motion_penalty(mx, my);
        if (motion >= best->cost) return;
        uint32_t sad;
        int64_t cost;
        if (!sad_if_better(mx, my, motion, best->cost, &sad, &cost)) return;
        if (cost < best->cost) {
            best->sad = sad;
            best->cost = cost;
            *dir = label;
        }
    }
    void expanding(int cxx, int cyy, int radius, int step, Best* best) const {  // :829-886
        for (int i = -radius + step; i < radius; i += step) {
            check(cxx + i, cyy - radius, best);
            check(cxx + i, cyy + radius, best);
        }
        for (int j = -radius + step; j < radius; j += step) {
            check(cxx - radius, cyy + j, best);
            check(cxx + radius, cyy + j, best);
        }
        check(cxx - radius, cyy - radius, best);
        check(cxx - radius, cyy + radius, best);
        check(cxx + radius, cyy - radius, best);
        check(cxx + radius, cyy + radius, best);
    }
    void hex2(int radius, Best* best) const {  // exact_hex2, :941-1033
        static const int HEX2[8][2] = {{-1, -2}, {-2, 0}, {-1, 2}, {1, 2}, {2, 0}, {1, -2}, {-1, -2}, {-2, 0}};
        static const int MOD6M1[8] = {5, 0, 1, 2, 3, 4, 5, 0};
        int bmx = best->mvx, bmy = best->mvy;
        if (radius > 1) {
            int dir = -2;
            for (int i = 0; i < 6; i++) check_dir(bmx + HEX2[i + 1][0], bmy + HEX2[i + 1][1], i, &dir, best);
            if (dir != -2) {
                bmx += HEX2[dir + 1][0];
                bmy += HEX2[dir + 1][1];
                int i = 1;
                while (i < radius / 2 && exact_ok(bmx, bmy, bounds)) {
                    int odir = MOD6M1[dir + 1];
                    dir = -2;
                    for (int k = 0; k < 3; k++)
                        check_dir(bmx + HEX2[odir + k][0], bmy + HEX2[odir + k][1], odir + k - 1, &dir, best);
                    if (dir == -2) break;
                    bmx += HEX2[dir + 1][0];
                    bmy += HEX2[dir + 1][1];
                    i++;
                }
            }
            best->mvx = bmx;
            best->mvy = bmy;
        }
        expanding(bmx, bmy, 1, 1, best);
    }
    void umh(int cxx, int cyy, int radius, Best* best) const {  // exact_umh, :1036-1119
        for (int d = 1; d < radius; d += 2) {
            check(cxx - d, cyy, best);
            check(cxx + d, cyy, best);
        }
        for (int d = 1; d < radius; d += 2) {
            check(cxx, cyy - d, best);
            check(cxx, cyy + d, best);
        }
        static const int HEX4[16][2] = {{-4, 2}, {-4, 1}, {-4, 0}, {-4, -1}, {-4, -2}, {4, -2}, {4, -1}, {4, 0},
                                        {4, 1},  {4, 2},  {2, 3},  {0, 4},   {-2, 3},  {-2, -3}, {0, -4}, {2, -3}};
        for (int scale = 1; scale <= std::max(radius / 4, 1); scale++)
            for (auto& d : HEX4) check(cxx + d[0] * scale, cyy + d[1] * scale, best);
        hex2(radius, best);
    }
    void refine(int search_type, int radius, Best* best) const {  // exact_refine_search, :771-826
        switch (search_type) {
            case 2: hex2(radius, best); break;
            case 3: umh(best->mvx, best->mvy, radius, best); break;
            case 4: {
                int cxx = best->mvx, cyy = best->mvy;
                for (int y = cyy - radius; y <= cyy + radius; y++)
                    for (int x = cxx - radius; x <= cxx + radius; x++) check(x, y, best);
                break;
            }
            default: break;
        }
    }
};

int search_radius(const AnalyseOpts& o, bool fine, int level, int nlevels) {  // :1465-1476
    int d = fine ? o.distance : o.coarse_distance;
    if (d == 0) return (level == 0 && nlevels > 1) ? 0 : 10;
    return std::abs(d);
}

int adaptive_radius(int radius, uint32_t activity) {  // :1478-1484
    if (activity < 2) return 0;
    return ((radius * (int)std::min<uint32_t>(255 * activity / 150, 255)) >> 8) + 1;
}

uint32_t plane_activity(const uint8_t* data, size_t stride, size_t row, size_t x, size_t width, size_t height,
                        bool satd) {  // :1517-1553
    uint32_t area = (uint32_t)(width * height);
    uint32_t sum = 0;
    for (size_t y = 0; y < height; y++)
        for (size_t i = 0; i < width; i++) sum += data[(row + y) * stride + x + i];
    uint32_t mean = sum / area;
    if (!satd) {
        uint32_t s = 0;
        for (size_t y = 0; y < height; y++)
            for (size_t i = 0; i < width; i++) {
                uint32_t v = data[(row + y) * stride + x + i];
                s += v > mean ? v - mean : mean - v;
            }
        return s / area;
    }
    uint32_t total = 0;
    for (size_t y = 0; y < height; y += 4)
        for (size_t x0 = 0; x0 < width; x0 += 4) {
            int32_t d[4][4];
            for (int dy = 0; dy < 4; dy++)
                for (int dx = 0; dx < 4; dx++)
                    d[dy][dx] = (int32_t)data[(row + y + dy) * stride + x + x0 + dx] - (int32_t)mean;
            total += hadamard4_satd(d);
        }
    return (total >> 1) / area;
}

uint32_t block_activity(const SuperPlanes& src, int level, int px, int py, int bw, int bh, bool satd) {  // :1486-1515
    size_t lw, lh;
    src.level_size(level, &lw, &lh);
    int x = px - ((level > 0 && px + bw >= (int)lw) ? 1 : 0);
    int y = py - ((level > 0 && py + bh >= (int)lh) ? 1 : 0);
    uint32_t luma = plane_activity(src.y, src.y_stride, src.level_y_offset(level) + (size_t)y, (size_t)x, (size_t)bw,
                                   (size_t)bh, satd);
    size_t cw = (size_t)(bw / 2), ch = (size_t)(bh / 2);
    size_t off = chroma_level_offset(src, level);
    return luma + plane_activity(src.u, src.u_stride, off + (size_t)(py / 2), (size_t)(px / 2), cw, ch, satd) +
           plane_activity(src.v, src.v_stride, off + (size_t)(py / 2), (size_t)(px / 2), cw, ch, satd);
}

uint8_t block_luma_dc(const SuperPlanes& src, int level, int px, int py, int bw, int bh) {  // :1566-1617
    size_t lw, lh;
    src.level_size(level, &lw, &lh);
    size_t y_off = src.level_y_offset(level);
    uint32_t sum = 0, count = 0;
    for (int row = 0; row < bh; row++) {
        size_t y = (size_t)(int64_t)py + (size_t)row;
        if (y >= lh) break;
        for (int col = 0; col < bw; col++) {
            size_t x = (size_t)(int64_t)px + (size_t)col;
            if (x >= lw) break;
            sum += at(src.y, src.y_len, (y_off + y) * src.y_stride + x, 0);
            count++;
        }
    }
    return count == 0 ? 0 : (uint8_t)std::min<uint32_t>(sum / count, 255);
}

std::vector<std::pair<int, int>> initial_search_order(int gw, int gh) {  // :302-324
    std::vector<std::pair<int, int>> o;
    for (int y = 1; y < gh - 1; y++) {
        if ((y & 1) == 0) for (int x = 1; x < gw - 1; x++) o.emplace_back(x, y);
        else for (int x = gw - 2; x >= 1; x--) o.emplace_back(x, y);
    }
    for (int x = 0; x < gw; x++) {
        o.emplace_back(x, 0);
        if (gh > 1) o.emplace_back(x, gh - 1);
    }
    for (int y = 1; y < gh - 1; y++) {
        o.emplace_back(0, y);
        if (gw > 1) o.emplace_back(gw - 1, y);
    }
    return o;
}

std::vector<std::pair<int, int>> refine_search_order(const std::vector<std::pair<int, int>>& coarse, int cgw, int cgh,
                                                     int gw, int gh, int level) {  // :338-402
    if (level == 0) return initial_search_order(gw, gh);
    int sw = cgw - 1, sh = cgh - 1, dw = gw - 1, dh = gh - 1;
    int extra_x = std::max(dw - 2 * sw, 0), extra_y = std::max(dh - 2 * sh, 0);
    std::vector<std::pair<int, int>> interior, boundary;
    std::vector<char> seen((size_t)std::max(gw * gh, 0), 0);
    auto add = [&](int x, int y) {
        if (x < 0 || y < 0 || x >= dw || y >= dh) return;
        size_t i = (size_t)(y * gw + x);
        if (seen[i]) return;
        seen[i] = 1;
        if (x > 0 && y > 0 && x < dw - 1 && y < dh - 1) interior.emplace_back(x, y);
        else boundary.emplace_back(x, y);
    };
    static const int OFF[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (auto& p : coarse) {
        int x = p.first, y = p.second;
        if (x >= sw || y >= sh) continue;
        int left = 2 * x, top = 2 * y;
        for (auto& o : OFF) add(left + o[0], top + o[1]);
        if (x == sw - 1)
            for (int dx = 2; dx < 2 + extra_x; dx++)
                for (int dy = 0; dy < 2 + (y == sh - 1 ? 1 : 0) * extra_y; dy++) add(left + dx, top + dy);
        if (y == sh - 1)
            for (int dy = 2; dy < 2 + extra_y; dy++)
                for (int dx = 0; dx < 2; dx++) add(left + dx, top + dy);
    }
    interior.insert(interior.end(), boundary.begin(), boundary.end());
    for (int x = 0; x <= dw; x++) interior.emplace_back(x, dh);
    for (int y = 0; y < dh; y++) interior.emplace_back(dw, y);
    return interior;
}

void reorder_search_order(std::vector<std::pair<int, int>>& order, const std::vector<Vec8>& field, int gw, int gh,
                          int bw, int bh) {  // :404-419
    size_t count = (size_t)((gw - 1) * (gh - 1));
    uint32_t v = (uint32_t)std::max(bw * bh / 8, 1);
    int shift = 31 - __builtin_clz(v) + 1;
    std::vector<std::pair<int, int>> base;
    for (int y = 0; y < gh - 1; y++)
        for (int x = 0; x < gw - 1; x++) base.emplace_back(x, y);
    std::stable_sort(base.begin(), base.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return (field[(size_t)(a.second * gw + a.first)].score >> shift) <
               (field[(size_t)(b.second * gw + b.first)].score >> shift);
    });
    std::copy(base.begin(), base.end(), order.begin());
    (void)count;
}

void exact_interpolate_level(const std::vector<Vec8>& coarse, int cgw, int cgh, int gw, int gh, int bw, int bh,
                             int ox, int oy, int pel, std::vector<Vec8>* out_p,
                             std::vector<Vec8>* proj_p) {  // :421-478
    std::vector<Vec8>& out = *out_p;
    std::vector<Vec8>& projected = *proj_p;
    out.assign((size_t)(gw * gh), Vec8{});
    Vec8 uns{};
    uns.score = 0xFFFFFFFFu;
    projected.assign((size_t)(gw * gh), uns);
    int cw = cgw - 1, ch = cgh - 1;
    auto get = [&](int x, int y) -> Vec8 {
        if (x < 2 * cw && y < 2 * ch) return coarse[(size_t)((y >> 1) * cgw + (x >> 1))];
        int edge = ch + std::min(x >> 1, cw) - std::min(y >> 1, ch);
        int xx, yy;
        if (edge <= cw) { xx = edge; yy = ch; } else { xx = cw; yy = cw + ch - edge; }
        return coarse[(size_t)(yy * cgw + xx)];
    };
    for (int y = 0; y < gh - 1; y++)
        for (int x = 0; x < gw - 1; x++) {
            Vec8 p = get(x, y);
            int vx = (int)p.dx * 2 * pel, vy = (int)p.dy * 2 * pel;
            Vec8 v;
            v.dx = (int16_t)clampi(vx, -32768, 32767);
            v.dy = (int16_t)clampi(vy, -32768, 32767);
            v.score = p.score;
            v.luma = p.luma;
            out[(size_t)(y * gw + x)] = v;
            if (x >= 2 * cw || y >= 2 * ch) continue;
            int tx = clampi((vx + (bw >> 1) + (bw - ox) * x) / (bw - ox), 0, gw - 1);
            int ty = clampi((vy + (bh >> 1) + (bh - oy) * y) / (bh - oy), 0, gh - 1);
            Vec8& t = projected[(size_t)(ty * gw + tx)];
            if (v.score < t.score) {
                t.dx = (int16_t)(-v.dx);
                t.dy = (int16_t)(-v.dy);
                t.score = v.score;
                t.luma = 0;
            }
        }
}

void exact_search_level(const SuperPlanes& src, const SuperPlanes& ref, const AnalyseOpts& opts, int level, int gx,
                        int gy, const std::vector<Vec8>& prev, std::vector<Vec8>& field, int gw, int gh,
                        const std::vector<std::pair<int, int>>& order, SearchCounters* counters) {  // :480-681
    const int obw = opts.block_w, obh = opts.block_h;
    const int step_x = std::max(obw - opts.overlap_x, 1), step_y = std::max(obh - opts.overlap_y, 1);
    const int pel = level == 0 ? std::max(opts.pel, 1) : 1;
    const int base_radius = search_radius(opts, level == 0, level, std::max(opts.levels, 1));
    const int search_type = level == 0 ? opts.search_type : opts.coarse_search_type;
    const bool satd = level == 0 ? opts.fine_satd : opts.coarse_satd;
    int bad_count = 0;
    std::vector<char> processed((size_t)std::max((gw + 1) * (gh + 1), 0), 0);
    size_t lw, lh;
    src.level_size(level, &lw, &lh);
    for (auto& pos : order) {
        int bx = pos.first, by = pos.second;
        size_t idx = (size_t)(by * gw + bx);
        int px = (level > 0 && bx == gw - 1) ? (int)lw - obw : bx * step_x;
        int py = (level > 0 && by == gh - 1) ? (int)lh - obh : by * step_y;
        bool smallest = level == opts.levels - 1;
        int bw = obw, bh = obh;
        if (smallest) { px += obw >> 2; py += obh >> 2; bw = obw >> 1; bh = obh >> 1; }
        Level L{src, ref, level, px, py, bw, bh, pel, satd, !smallest, 0, 0, {}, 0, opts.pnew, counters};
        mv_bounds(px, py, bw, bh, (int)lw, (int)lh, pel, L.bounds);
        bool edge = level > 0 && (bx == gw - 1 || by == gh - 1);
        int sx = std::min(bx, gw - 2), sy = std::min(by, gh - 2);
        Vec8 interp = field[(size_t)(sy * gw + sx)];
        if (edge && !smallest) {
            Level E = L;
            E.px = sx * step_x;
            E.py = sy * step_y;
            E.chroma = true;
            interp.score = E.exact_sad(interp.dx, interp.dy);
        } else if (!edge) {
            interp = field[idx];
        }
        const Vec8 predictor = interp;
        int lambda_base = smallest ? 0 : exact_lambda(opts, level, pel);
        int lambda = adapt_lambda(lambda_base, smallest ? opts.lsad / 3 : opts.lsad, predictor.score);
        int radius = base_radius;
        if (level > 0 && opts.coarse_distance < 0)
            radius = adaptive_radius(base_radius, block_activity(src, level, px, py, bw, bh, satd));
        int seeds[8][2];
        int nseeds = 0;
        for (int ny = std::max(by - 1, 0); ny <= std::min(by + 1, gh - 1); ny++)
            for (int nx = std::max(bx - 1, 0); nx <= std::min(bx + 1, gw - 1); nx++)
                if (processed[(size_t)(ny * (gw + 1) + nx)]) {
                    const Vec8& n = field[(size_t)(ny * gw + nx)];
                    seeds[nseeds][0] = n.dx;
                    seeds[nseeds][1] = n.dy;
                    nseeds++;
                }
        L.lambda = lambda;
        // exact_search_block, :684-768
        int pmx = predictor.dx, pmy = predictor.dy;
        bool field_valid = predictor.score != 0xFFFFFFFFu;
        bool has_prev = idx < prev.size() && prev[idx].score != 0xFFFFFFFFu && prev[idx].score > 0;
        int stype = (edge && level == 1) ? 0 : search_type;
        auto special = [&](int mx, int my, int penalty) {
            uint32_t s = L.exact_sad(mx, my);
            return Best{mx, my, s, (int64_t)s + (((int64_t)penalty * (int64_t)s) >> 8)};
        };
        Best best;
        BlockTrace tr{};
        if (level == 0 && opts.distance == 0) {
            best = field_valid ? special(pmx, pmy, 0) : special(0, 0, 0);
            tr.stage_mvx = best.mvx;
            tr.stage_mvy = best.mvy;
            tr.stage_sad = (int32_t)best.sad;
        } else {
            Best fieldb{};
            if (field_valid) fieldb = special(pmx, pmy, 0);
            Best previous{};
            if (has_prev) previous = special(prev[idx].dx, prev[idx].dy, opts.prev);
            Best zero_best = special(0, 0, opts.pzero);
            Best global_best = special(gx, gy, opts.pglobal);
            best = field_valid ? fieldb : Best{0, 0, 0xFFFFFFFFu, std::numeric_limits<int64_t>::max()};
            if (has_prev && previous.cost < best.cost) {
                best = previous;
                pmx = previous.mvx;
                pmy = previous.mvy;
            }
            if (global_best.cost < best.cost) best = global_best;
            if (zero_best.cost < best.cost) best = zero_best;
            for (int k = 0; k < nseeds; k++) {
                uint32_t s = L.exact_sad(seeds[k][0], seeds[k][1]);
                int64_t c = (int64_t)s + (((int64_t)opts.pnbour * (int64_t)s) >> 8);
                if (c < best.cost) best = Best{seeds[k][0], seeds[k][1], s, c};
            }
            best.cost = (int64_t)best.sad;
            tr.stage_mvx = best.mvx;
            tr.stage_mvy = best.mvy;
            tr.stage_sad = (int32_t)best.sad;
            L.pred_x = pmx;
            L.pred_y = pmy;
            L.refine(stype, radius, &best);
        }
        tr.mpred_x = pmx;
        tr.mpred_y = pmy;
        // bad-block expansion (:607-672) uses the block's own predictor
        L.pred_x = predictor.dx;
        L.pred_y = predictor.dy;
        uint32_t found = best.sad;
        int32_t bad_limit = sat_add_i32(opts.coarse_bad_sad, sat_mul_i32(opts.coarse_bad_sad, bad_count) / 16);
        tr.found = (int32_t)found;
        tr.bad_before = bad_count;
        tr.bad_limit = bad_limit;
        if (level > 0 && idx > 1 && opts.coarse_bad_sad > 0 && (int32_t)best.sad > bad_limit) {
            tr.expanded = 1;
            bad_count++;
            if (opts.coarse_bad_range < 0) {
                int max_r = -opts.coarse_bad_range * pel;
                for (int r = pel; r < max_r; r += pel) {
                    L.expanding(0, 0, r, pel, &best);
                    if (best.sad < found / 4) break;
                }
            } else if (opts.coarse_bad_range > 0) {
                L.umh(0, 0, opts.coarse_bad_range * pel, &best);
            }
            L.expanding(best.mvx, best.mvy, 1, 1, &best);
        }
        Vec8 out;
        out.dx = (int16_t)best.mvx;
        out.dy = (int16_t)best.mvy;
        out.score = sat_mul_u32(best.sad, smallest ? 3u : 1u);
        out.luma = block_luma_dc(src, level, px, py, bw, bh);
        field[idx] = out;
        processed[(size_t)(by * (gw + 1) + bx)] = 1;
        if (counters && counters->trace[counters->dir]) {
            tr.level = level;
            tr.dir = counters->dir;
            tr.bx = bx; tr.by = by; tr.px = px; tr.py = py; tr.bw = bw; tr.bh = bh;
            tr.edge = edge; tr.smallest = smallest;
            tr.pred_dx = predictor.dx; tr.pred_dy = predictor.dy; tr.pred_score = (int32_t)predictor.score;
            tr.has_prev = has_prev;
            if (has_prev) { tr.prev_dx = prev[idx].dx; tr.prev_dy = prev[idx].dy; }
            tr.gx = gx; tr.gy = gy;
            tr.nseeds = nseeds;
            for (int k = 0; k < nseeds; k++) { tr.seeds[2 * k] = seeds[k][0]; tr.seeds[2 * k + 1] = seeds[k][1]; }
            tr.lambda = lambda; tr.radius = radius; tr.stype = stype;
            tr.out_dx = out.dx; tr.out_dy = out.dy; tr.out_score = (int32_t)out.score; tr.out_luma = out.luma;
            counters->trace[counters->dir]->push_back(tr);
        }
    }
}

std::vector<Vec8> exact_recalculate(const SuperPlanes& src, const SuperPlanes& ref, const AnalyseOpts& opts,
                                    std::vector<Vec8> field, int old_gw, int old_gh,
                                    SearchCounters* counters) {  // :1326-1451
    if (!opts.has_refine) return field;
    int thsad = opts.refine_thsad;
    int old_bw = opts.block_w, old_bh = opts.block_h, old_ox = opts.overlap_x, old_oy = opts.overlap_y;
    int bw, bh, ox, oy, gw, gh;
    opts.output_block(&bw, &bh, &ox, &oy);
    opts.grid(bw, bh, ox, oy, &gw, &gh);
    int old_sx = std::max(old_bw - old_ox, 1), old_sy = std::max(old_bh - old_oy, 1);
    int step_x = std::max(bw - ox, 1), step_y = std::max(bh - oy, 1);
    int pel = std::max(opts.pel, 1);
    size_t lw, lh;
    src.level_size(0, &lw, &lh);
    std::vector<Vec8> out((size_t)(gw * gh));
    auto get = [&](int x, int y) {
        return field[(size_t)(clampi(y, 0, old_gh - 1) * old_gw + clampi(x, 0, old_gw - 1))];
    };
    auto div_euclid = [](int64_t a, int64_t b) {
        int64_t q = a / b;
        if (a % b < 0) q += (b > 0) ? -1 : 1;
        return q;
    };
    for (int by = 0; by < gh; by++) {
        int dir = (opts.sort && (by & 1)) ? -1 : 1;
        for (int pos = 0; pos < gw; pos++) {
            int bx = dir > 0 ? pos : gw - 1 - pos;
            int center_x = bw / 2 + step_x * bx, center_y = bh / 2 + step_y * by;
            int old_x = (center_x - old_bw / 2) / old_sx, old_y = (center_y - old_bh / 2) / old_sy;
            int delta_x = std::max(center_x - (old_bw / 2 + old_sx * old_x), 0);
            int delta_y = std::max(center_y - (old_bh / 2 + old_sy * old_y), 0);
            Vec8 p00 = get(old_x, old_y), p10 = get(old_x + 1, old_y), p01 = get(old_x, old_y + 1),
                 p11 = get(old_x + 1, old_y + 1);
            auto interp = [&](int64_t a, int64_t b, int64_t c, int64_t d) {
                int64_t top = a * old_sx + (int64_t)delta_x * (b - a);
                int64_t bottom = c * old_sx + (int64_t)delta_x * (d - c);
                return div_euclid(top + div_euclid((int64_t)delta_y * (bottom - top), old_sy), old_sx);
            };
            int mvx = (int)interp(p00.dx, p10.dx, p01.dx, p11.dx);
            int mvy = (int)interp(p00.dy, p10.dy, p01.dy, p11.dy);
            int64_t parent_sad = interp(p00.score, p10.score, p01.score, p11.score) * (int64_t)(bw * bh) /
                                 (int64_t)(old_bw * old_bh);
            int px = bx * step_x, py = by * step_y;
            bool border = bx == 0 || by == 0 || bx == gw - 1 || by == gh - 1;
            int factor = (border && bw > 4) ? 3 : 1;
            int ebw = factor == 3 ? bw / 2 : bw, ebh = factor == 3 ? bh / 2 : bh;
            int epx = px + (factor == 3 ? bw / 4 : 0), epy = py + (factor == 3 ? bh / 4 : 0);
            Level L{src, ref, 0, epx, epy, ebw, ebh, pel, opts.refine_satd, true, 0, 0, {}, 0, opts.pnew / factor,
                    counters};
            mv_bounds(epx, epy, ebw, ebh, (int)lw, (int)lh, pel, L.bounds);
            int maxx = std::max(L.bounds[2] - 1, L.bounds[0]), maxy = std::max(L.bounds[3] - 1, L.bounds[1]);
            int prx = clampi(mvx, L.bounds[0], maxx), pry = clampi(mvy, L.bounds[1], maxy);  // exact_clip
            Best best{prx, pry, (uint32_t)std::min<int64_t>(std::max<int64_t>(parent_sad, 0), 0xFFFFFFFFll),
                      std::max<int64_t>(parent_sad, 0)};
            int threshold = thsad / factor;
            BlockTrace tr{};
            tr.pred_dx = mvx; tr.pred_dy = mvy;
            tr.pred_score = (int32_t)std::min<int64_t>(std::max<int64_t>(parent_sad, INT32_MIN), INT32_MAX);
            tr.stage_mvx = prx; tr.stage_mvy = pry; tr.stage_sad = (int32_t)best.sad;
            if (parent_sad > (int64_t)threshold) {
                best.sad = L.exact_sad(prx, pry);
                best.cost = (int64_t)best.sad;
                L.lambda = by == 0 ? 0 : (opts.lambda >> 2) / factor;
                L.pred_x = prx;
                L.pred_y = pry;
                tr.expanded = 1;   // searched
                tr.stage_sad = (int32_t)best.sad;
                tr.lambda = L.lambda;
                L.refine(opts.refine_search_type, std::abs(opts.refine_distance), &best);
            }
            Vec8 o;
            o.dx = (int16_t)best.mvx;
            o.dy = (int16_t)best.mvy;
            o.score = sat_mul_u32(best.sad, (uint32_t)factor);
            o.luma = block_luma_dc(src, 0, epx, epy, ebw, ebh);
            out[(size_t)(by * gw + bx)] = o;
            if (counters && counters->trace[counters->dir]) {
                tr.level = -1;
                tr.dir = counters->dir;
                tr.bx = bx; tr.by = by; tr.px = epx; tr.py = epy; tr.bw = ebw; tr.bh = ebh;
                tr.edge = factor;
                tr.radius = std::abs(opts.refine_distance);
                tr.stype = opts.refine_search_type;
                tr.mpred_x = prx; tr.mpred_y = pry;
                tr.found = (int32_t)best.sad;
                tr.out_dx = o.dx; tr.out_dy = o.dy; tr.out_score = (int32_t)o.score; tr.out_luma = o.luma;
                counters->trace[counters->dir]->push_back(tr);
            }
        }
    }
    return out;
}

// exact_search_hierarchy_bidir (:94-226) / exact_search_hierarchy (:235-293)
struct Dir {
    std::vector<Vec8> field;
    std::vector<std::pair<int, int>> order;
    int gx = 0, gy = 0;
};

}  // namespace

std::vector<std::pair<int, int>> level_search_order(int gw, int gh, int level) {  // :326-336
    int edge = level > 0 ? 1 : 0;
    int bw = gw - edge, bh = gh - edge;
    auto o = initial_search_order(bw, bh);
    if (edge) {
        for (int x = 0; x <= bw; x++) o.emplace_back(x, bh);
        for (int y = 0; y < bh; y++) o.emplace_back(bw, y);
    }
    return o;
}

void level_grid(const AnalyseOpts& opts, int level, int* gw, int* gh) {  // :1453-1463
    int sx = std::max(opts.block_w - opts.overlap_x, 1), sy = std::max(opts.block_h - opts.overlap_y, 1);
    int gw0 = std::max((opts.width - opts.overlap_x) / sx, 0), gh0 = std::max((opts.height - opts.overlap_y) / sy, 0);
    int ex = opts.overlap_x + sx * gw0, ey = opts.overlap_y + sy * gh0;
    *gw = std::max(((ex >> level) - opts.overlap_x) / sx, 1) + (level > 0 ? 1 : 0);
    *gh = std::max(((ey >> level) - opts.overlap_y) / sy, 1) + (level > 0 ? 1 : 0);
}

int exact_lambda(const AnalyseOpts& opts, int level, int pel) {  // :1279-1283
    return f2i((double)opts.lambda / powi(opts.plevel, opts.levels - level - 1) / (double)std::max(pel * pel, 1));
}

int adapt_lambda(int nlambda, int lsad, uint32_t pred_score) {  // :1555-1564
    double l = (double)std::max(lsad, 1);
    double half = pred_score == 0xFFFFFFFFu ? -1.0 : (double)(pred_score >> 1);
    double t = l / (l + half);
    return f2i(std::fmin((double)nlambda * t * t, 21474836.47));
}

void exact_global_doubled(const std::vector<Vec8>& field, int* gx, int* gy) {  // :1285-1324
    auto mode = [&](bool x) {
        std::vector<int> v;
        v.reserve(field.size());
        for (auto& f : field) v.push_back(x ? f.dx : f.dy);
        std::sort(v.begin(), v.end());
        int best = v.empty() ? 0 : v[0];
        size_t best_count = 0, i = 0;
        while (i < v.size()) {
            size_t j = i + 1;
            while (j < v.size() && v[j] == v[i]) j++;
            if (j - i > best_count) { best = v[i]; best_count = j - i; }
            i = j;
        }
        return best;
    };
    int mx = mode(true), my = mode(false);
    int64_t sx = 0, sy = 0, count = 0;
    for (auto& f : field)
        if (std::abs((int)f.dx - mx) < 6 && std::abs((int)f.dy - my) < 6) { sx += f.dx; sy += f.dy; count++; }
    if (count == 0) { *gx = 2 * mx; *gy = 2 * my; return; }
    auto mean = [&](int64_t s) {
        int64_t sg = s > 0 ? 1 : (s < 0 ? -1 : 0);
        return (s + sg * count / 2) / count;
    };
    *gx = 2 * (int)mean(sx);
    *gy = 2 * (int)mean(sy);
}

void rescale_scores(std::vector<Vec8>& field, int width, int height, int blk) {  // :2716-2738
    int diag = std::max(f2i(std::sqrt(powi((double)width, 2) + powi((double)height, 2))), 0);
    int denom = std::max(sat_mul_i32(diag, std::max(blk, 1)), 1);
    for (auto& v : field) {
        int mag = std::abs((int)v.dx) + std::abs((int)v.dy);
        int ratio = (100 * mag) / denom;
        int32_t score = (int32_t)v.score;
        if (ratio >= 51) {
            score = sat_mul_i32(score, 20);
        } else if (ratio >= 16) {
            uint64_t p = 3926827243ull * (uint64_t)(uint32_t)score;
            p *= (uint64_t)(uint32_t)(ratio - 15);
            int32_t v71 = (int32_t)(uint32_t)(p >> 32);
            score = sat_add_i32(score, 9 * ((v71 >> 31) + (v71 >> 5)));
        }
        v.score = (uint32_t)score & 0x00FFFFFFu;
    }
}

static std::vector<Vec8> search_hierarchy(const SuperPlanes& src, const SuperPlanes& ref, const AnalyseOpts& opts,
                                          SearchCounters* counters) {  // :235-293
    Dir d;
    int fgw = 0, fgh = 0;
    for (int level = std::max(opts.levels, 1) - 1; level >= 0; level--) {
        int gw, gh;
        level_grid(opts, level, &gw, &gh);
        int pel = level == 0 ? std::max(opts.pel, 1) : 1;
        d.order = d.order.empty() ? level_search_order(gw, gh, level) : refine_search_order(d.order, fgw, fgh, gw, gh, level);
        std::vector<Vec8> next, prev;
        if (d.field.empty()) next.assign((size_t)(gw * gh), Vec8{});
        else exact_interpolate_level(d.field, fgw, fgh, gw, gh, opts.block_w, opts.block_h, opts.overlap_x,
                                     opts.overlap_y, pel, &next, &prev);
        exact_search_level(src, ref, opts, level, d.gx, d.gy, prev, next, gw, gh, d.order, counters);
        if (level > 0) {
            exact_global_doubled(next, &d.gx, &d.gy);
            reorder_search_order(d.order, next, gw, gh, opts.block_w, opts.block_h);
        }
        d.field = std::move(next);
        fgw = gw;
        fgh = gh;
    }
    auto field = exact_recalculate(src, ref, opts, std::move(d.field), fgw, fgh, counters);
    int bw, bh, ox, oy;
    opts.output_block(&bw, &bh, &ox, &oy);
    rescale_scores(field, opts.width, opts.height, bw);
    return field;
}

void analyse_pair(const SuperPlanes& cur, const SuperPlanes& ref, const AnalyseOpts& opts, bool* has_bwd,
                  std::vector<Vec8>* bwd_out, bool* has_fwd, std::vector<Vec8>* fwd_out, SearchCounters* counters,
                  bool threads) {  // :68-92
    bool do_bwd = opts.vectors & 1, do_fwd = opts.vectors & 2;
    *has_bwd = do_bwd;
    *has_fwd = do_fwd;
    if (!(do_bwd && do_fwd)) {
        if (counters) counters->dir = 0;
        if (do_fwd) *fwd_out = search_hierarchy(cur, ref, opts, counters);
        if (counters) counters->dir = 1;
        if (do_bwd) *bwd_out = search_hierarchy(ref, cur, opts, counters);
        if (counters) counters->dir = 0;
        return;
    }
    // exact_search_hierarchy_bidir: each direction's level search takes the other
    // direction's projected coarse field as its `prev` candidate.
    Dir fw, bw;
    int fgw = 0, fgh = 0;
    SearchCounters cb;
    if (counters) {
        counters->dir = 0;
        cb.trace[0] = counters->trace[0];
        cb.trace[1] = counters->trace[1];
        cb.dir = 1;
    }
    for (int level = std::max(opts.levels, 1) - 1; level >= 0; level--) {
        int gw, gh;
        level_grid(opts, level, &gw, &gh);
        int pel = level == 0 ? std::max(opts.pel, 1) : 1;
        fw.order = fw.order.empty() ? level_search_order(gw, gh, level) : refine_search_order(fw.order, fgw, fgh, gw, gh, level);
        bw.order = bw.order.empty() ? level_search_order(gw, gh, level) : refine_search_order(bw.order, fgw, fgh, gw, gh, level);
        std::vector<Vec8> next_f, prev_f, next_b, prev_b;
        if (fw.field.empty()) next_f.assign((size_t)(gw * gh), Vec8{});
        else exact_interpolate_level(fw.field, fgw, fgh, gw, gh, opts.block_w, opts.block_h, opts.overlap_x,
                                     opts.overlap_y, pel, &next_f, &prev_f);
        if (bw.field.empty()) next_b.assign((size_t)(gw * gh), Vec8{});
        else exact_interpolate_level(bw.field, fgw, fgh, gw, gh, opts.block_w, opts.block_h, opts.overlap_x,
                                     opts.overlap_y, pel, &next_b, &prev_b);
        auto run_f = [&] { exact_search_level(cur, ref, opts, level, fw.gx, fw.gy, prev_b, next_f, gw, gh, fw.order, counters); };
        auto run_b = [&] { exact_search_level(ref, cur, opts, level, bw.gx, bw.gy, prev_f, next_b, gw, gh, bw.order,
                                              counters ? &cb : nullptr); };
        if (threads) {
            std::thread t(run_b);
            run_f();
            t.join();
        } else {
            run_f();
            run_b();
        }
        if (level > 0) {
            exact_global_doubled(next_f, &fw.gx, &fw.gy);
            exact_global_doubled(next_b, &bw.gx, &bw.gy);
            reorder_search_order(fw.order, next_f, gw, gh, opts.block_w, opts.block_h);
            reorder_search_order(bw.order, next_b, gw, gh, opts.block_w, opts.block_h);
        }
        fw.field = std::move(next_f);
        bw.field = std::move(next_b);
        fgw = gw;
        fgh = gh;
    }
    auto f = exact_recalculate(cur, ref, opts, std::move(fw.field), fgw, fgh, counters);
    auto b = exact_recalculate(ref, cur, opts, std::move(bw.field), fgw, fgh, counters ? &cb : nullptr);
    int obw, obh, oox, ooy;
    opts.output_block(&obw, &obh, &oox, &ooy);
    rescale_scores(f, opts.width, opts.height, obw);
    rescale_scores(b, opts.width, opts.height, obw);
    *fwd_out = std::move(f);
    *bwd_out = std::move(b);
    if (counters)
        for (int i = 0; i < 16; i++) { counters->evals[i] += cb.evals[i]; counters->pixels[i] += cb.pixels[i]; }
}

static void write_i32(std::vector<uint8_t>& b, size_t off, int32_t v) {
    if (off + 4 <= b.size()) std::memcpy(&b[off], &v, 4);
}

std::vector<uint8_t> pack_vector_frame(const AnalyseOpts& opts, const std::vector<Vec8>* previous,
                                       const std::vector<Vec8>* current) {  // :2740-2811
    int bw, bh, ox, oy, gw, gh;
    opts.output_block(&bw, &bh, &ox, &oy);
    opts.grid(bw, bh, ox, oy, &gw, &gh);
    size_t count = (size_t)(gw * gh);
    int flags = (previous ? 1 : 0) | (current ? 2 : 0);
    size_t region = 4 + count * 8;
    size_t nreg = (previous ? 1 : 0) + (current ? 1 : 0);
    std::vector<uint8_t> out(0x40 + nreg * region, 0);
    int32_t hdr[16] = {16, 0xA0, flags, bw, bh, opts.pel, opts.has_refine ? 1 : opts.levels, 0, opts.width,
                       opts.height, ox, oy, gw, gh, opts.selector, opts.delta};
    for (int i = 0; i < 16; i++) write_i32(out, (size_t)i * 4, hdr[i]);
    size_t cur = 0x40;
    int32_t marker = (int32_t)(2 * count + 1);
    auto write_vecs = [&](const std::vector<Vec8>& v) {
        write_i32(out, cur, marker);
        cur += 4;
        for (size_t i = 0; i < count; i++) {
            Vec8 x = i < v.size() ? v[i] : Vec8{};
            uint32_t d0 = ((uint32_t)(int32_t)x.dx << 16) | (uint32_t)(uint16_t)x.dy;
            uint32_t d1 = (x.score & 0x00FFFFFFu) | ((uint32_t)x.luma << 24);
            write_i32(out, cur, (int32_t)d0);
            write_i32(out, cur + 4, (int32_t)d1);
            cur += 8;
        }
    };
    if (previous) write_vecs(*previous);
    if (current) write_vecs(*current);
    return out;
}

std::vector<int32_t> pack_vdata_header(const AnalyseOpts& opts) {  // :2813-2841
    int bw, bh, ox, oy, gw, gh;
    opts.output_block(&bw, &bh, &ox, &oy);
    opts.grid(bw, bh, ox, oy, &gw, &gh);
    std::vector<int32_t> h(16, 0);
    h[0] = 0xA0;
    h[1] = opts.vectors == 1 ? 1 : (opts.vectors == 2 ? 2 : 3);
    h[2] = bw; h[3] = bh; h[4] = opts.pel;
    h[5] = opts.has_refine ? 1 : opts.levels;
    h[7] = opts.width; h[8] = opts.height; h[9] = ox; h[10] = oy; h[11] = gw; h[12] = gh;
    h[13] = opts.selector; h[14] = opts.delta;
    return h;
}

}  // namespace oracle
