// Oracle restatement of crates/svpflow1/src/super_opts.rs and super_build.rs.
// Test infrastructure only (see oracle.h).  Buffer reads follow the
// reference's slice semantics: `buf.get(i)` beyond the plane -> the fallback.
#include "oracle.h"

#include <algorithm>
#include <cstring>

namespace oracle {

int reduce_dim(int dim, int n) {  // super_opts.rs:70-75
    for (int i = 0; i < n; i++) dim = 2 * (dim / 4);
    return dim;
}

int auto_levels(int width, int height) {  // super_opts.rs:77-87
    int i = 0;
    for (;;) {
        int h = reduce_dim(height, i), w = reduce_dim(width, i);
        if (h < 4 || w < 4) return i;
        i++;
    }
}

int super_plane_height(int base_h, int pel, int levels, bool full) {  // super_opts.rs:89-95
    int total = full ? pel * pel * base_h : 0;
    for (int lv = 1; lv < levels; lv++) total += reduce_dim(base_h, lv);
    return total;
}

std::string super_opts_from(const Value* opt, int width, int height, SuperOpts* o) {  // super_opts.rs:17-49
    int64_t iv;
    bool bv;
    o->pel = (opt && opt->int_at({"pel"}, &iv)) ? (int)iv : 2;
    if (!(o->pel == 1 || o->pel == 2 || o->pel == 4)) return "SVSuper: pel has to be 1 or 2 or 4";
    o->gpu = (opt && opt->int_at({"gpu"}, &iv)) ? (int)iv : 1;
    if (o->gpu < 0 || o->gpu > 2) return "SVSuper: gpu has to be 0 or 1 or 2";
    o->scale_down = (opt && opt->int_at({"scale", "down"}, &iv)) ? (int)iv : 4;
    o->scale_up = (opt && opt->int_at({"scale", "up"}, &iv)) ? (int)iv : 2;
    if (o->gpu != 0) o->scale_up = 0;
    o->full = (opt && opt->bool_at({"full"}, &bv)) ? bv : true;
    if (o->pel >= 2 && !o->full) return "SVSuper: can't ignore finest level with high pel value";
    o->rc = (opt && opt->bool_at({"rc"}, &bv)) ? bv : false;
    o->levels = auto_levels(width, height);
    o->width = width;
    o->height = height;
    return "";
}

int64_t super_pack_data(const SuperOpts& o) {  // super_opts.rs:51-63
    uint8_t flags = o.rc ? 1 : 0;
    if (!o.full) flags |= 2;
    uint64_t raw = (uint64_t)(uint16_t)o.height | ((uint64_t)(uint16_t)o.width << 16) |
                   ((uint64_t)(uint8_t)o.pel << 32) | ((uint64_t)(uint8_t)o.levels << 40) |
                   ((uint64_t)(uint8_t)o.gpu << 48) | ((uint64_t)flags << 56);
    int64_t r;
    std::memcpy(&r, &raw, 8);
    return r;
}

namespace {

struct Buf {  // a mutable plane with the reference's `get`/index semantics
    uint8_t* p;
    size_t len;
    bool get(size_t i, uint8_t* v) const { if (i < len) { *v = p[i]; return true; } return false; }
};

inline uint8_t avg_epu8(uint8_t a, uint8_t b) { return (uint8_t)(((unsigned)a + (unsigned)b + 1) >> 1); }  // :245-248

// super_build.rs:250-252
inline uint8_t sample(const Buf& buf, size_t stride, size_t row0, size_t y, size_t x) {
    uint8_t v;
    return buf.get((row0 + y) * stride + x, &v) ? v : 0;
}

// super_build.rs:310-357
void horizontal_6tap_inplace(uint8_t* row, size_t row_len, size_t dst_w, std::vector<uint8_t>& scratch) {
    if (dst_w == 0 || row_len == 0) return;
    auto rget = [&](size_t i, uint8_t* v) { if (i < row_len) { *v = row[i]; return true; } return false; };
    if (dst_w <= 2) {
        for (size_t i = 0; i < dst_w; i++) {
            uint8_t a, b;
            if (!rget(2 * i, &a)) a = 0;
            if (!rget(2 * i + 1, &b)) b = a;
            row[i] = avg_epu8(a, b);
        }
        return;
    }
    if (scratch.size() < row_len) return;
    uint8_t* src = scratch.data();
    std::memcpy(src, row, row_len);
    size_t src_w = row_len;
    auto sget = [&](size_t i, uint8_t fb) { return i < src_w ? src[i] : fb; };
    row[0] = avg_epu8(src[0], sget(1, src[0]));
    uint8_t left = row[0];
    size_t interior = dst_w >= 2 ? dst_w - 2 : 0;
    for (size_t i = 0; i < interior; i++) {
        size_t k = 5 + 2 * i;
        uint32_t p_k = sget(k, src[src_w - 1]);
        uint32_t p_km3 = sget(k - 3, 0);
        uint32_t p_km2 = sget(k - 2, 0);
        uint32_t p_km4 = sget(k - 4, 0);
        uint32_t p_even = sget((k - 1) & ~(size_t)1, 0);
        uint32_t v = ((uint32_t)left + p_k + 10 * (p_km3 + p_km2) + 5 * (p_km4 + p_even) + 16) >> 5;
        row[i + 1] = (uint8_t)std::min<uint32_t>(v, 255);
        left = (uint8_t)p_km3;
    }
    size_t i = std::max<size_t>(dst_w >= 1 ? dst_w - 1 : 0, 1);
    while (i < dst_w) {
        uint8_t a = sget(2 * i, 0);
        uint8_t b = sget(2 * i + 1, a);
        row[i] = avg_epu8(a, b);
        i++;
    }
}

// super_build.rs:254-308
void reduce_6tap(Buf buf, size_t stride, size_t dst_row0, size_t dst_w, size_t dst_h, size_t src_row0,
                 size_t src_w, size_t src_h) {
    if (dst_w == 0 || dst_h == 0 || src_w == 0 || src_h == 0) return;
    size_t inter_w = src_w;
    std::vector<uint8_t> inter(inter_w * dst_h, 0);
    for (size_t dy = 0; dy < dst_h; dy++) {
        size_t sy = dy * 2;
        long win0 = (long)sy - 2;
        bool use_edge = dy == 0 || dy + 1 == dst_h || win0 < 0 || (sy + 3) >= src_h;
        if (use_edge) {
            size_t last = src_h ? src_h - 1 : 0;
            size_t y0 = std::min(sy, last), y1 = std::min(sy + 1, last);
            for (size_t x = 0; x < inter_w; x++)
                inter[dy * inter_w + x] = avg_epu8(sample(buf, stride, src_row0, y0, x), sample(buf, stride, src_row0, y1, x));
        } else {
            size_t s = (size_t)win0;
            for (size_t x = 0; x < inter_w; x++) {
                uint32_t p0 = sample(buf, stride, src_row0, s, x), p1 = sample(buf, stride, src_row0, s + 1, x);
                uint32_t p2 = sample(buf, stride, src_row0, s + 2, x), p3 = sample(buf, stride, src_row0, s + 3, x);
                uint32_t p4 = sample(buf, stride, src_row0, s + 4, x), p5 = sample(buf, stride, src_row0, s + 5, x);
                uint32_t v = (p0 + p5 + 5 * (p1 + p4) + 10 * (p2 + p3) + 16) >> 5;
                inter[dy * inter_w + x] = (uint8_t)std::min<uint32_t>(v, 255);
            }
        }
    }
    std::vector<uint8_t> scratch(inter_w, 0);
    for (size_t dy = 0; dy < dst_h; dy++) {
        uint8_t* row = &inter[dy * inter_w];
        horizontal_6tap_inplace(row, inter_w, dst_w, scratch);
        size_t base = (dst_row0 + dy) * stride;
        size_t n = std::min(inter_w, buf.len > base ? buf.len - base : 0);
        if (n > 0) std::memcpy(buf.p + base, row, n);
    }
}

// super_build.rs:230-243
void copy_row(Buf buf, size_t dst_off, size_t src_off, size_t n) {
    n = std::min(n, buf.len > dst_off ? buf.len - dst_off : 0);
    n = std::min(n, buf.len > src_off ? buf.len - src_off : 0);
    if (n == 0 || dst_off == src_off) return;
    if (dst_off < src_off) {
        for (size_t i = 0; i < n; i++) buf.p[dst_off + i] = buf.p[src_off + i];
    } else {
        std::vector<uint8_t> tmp(buf.p + src_off, buf.p + src_off + n);
        std::memcpy(buf.p + dst_off, tmp.data(), n);
    }
}

// super_build.rs:179-228: reads follow the buffer, so x+1 past the width reads the
// stride padding (or the next row), and row h reads the first x-half row.
void fill_bilinear_pel2(Buf dst, size_t stride, size_t row0, size_t w, size_t h) {
    size_t full = row0, h_plane = row0 + h, v_plane = row0 + 2 * h, hv_plane = row0 + 3 * h;
    for (size_t y = 0; y < h; y++) {
        size_t sb = (full + y) * stride, db = (h_plane + y) * stride;
        for (size_t x = 0; x < w; x++) {
            uint8_t a, b;
            if (!dst.get(sb + x, &a)) a = 0;
            if (!dst.get(sb + x + 1, &b)) b = a;
            if (db + x < dst.len) dst.p[db + x] = avg_epu8(a, b);
        }
    }
    for (size_t y = 0; y < h; y++) {
        size_t sb = (full + y) * stride, nb = (full + y + 1) * stride, db = (v_plane + y) * stride;
        for (size_t x = 0; x < w; x++) {
            uint8_t a, b;
            if (!dst.get(sb + x, &a)) a = 0;
            if (!dst.get(nb + x, &b)) b = a;
            if (db + x < dst.len) dst.p[db + x] = avg_epu8(a, b);
        }
    }
    copy_row(dst, v_plane * stride, full * stride, w);
    for (size_t y = 0; y < h; y++) {
        size_t sb = (full + y) * stride, gb = (full + y + 1) * stride + 1, db = (hv_plane + y) * stride;
        for (size_t x = 0; x < w; x++) {
            uint8_t a, b;
            if (!dst.get(sb + x, &a)) a = 0;
            if (!dst.get(gb + x, &b)) b = a;
            if (db + x < dst.len) dst.p[db + x] = avg_epu8(a, b);
        }
    }
    copy_row(dst, hv_plane * stride, full * stride + 1, w);
}

// super_build.rs:163-177 (scalar semantics of the SSE2 cubic8, which equals it lane by lane)
uint8_t half(const Buf& buf, size_t i, size_t step, size_t p, size_t n) {
    if (p < 2 || p + 4 >= n) {
        if (p + 1 == n) return buf.p[i];
        return avg_epu8(buf.p[i], buf.p[i + step]);
    }
    int v = (int)buf.p[i - 2 * step] + (int)buf.p[i + 3 * step] +
            5 * (4 * (int)buf.p[i] - (int)buf.p[i - step] + 4 * (int)buf.p[i + step] - (int)buf.p[i + 2 * step]) + 16;
    v >>= 5;
    return (uint8_t)std::min(255, std::max(0, v));
}

// super_build.rs:95-134.  The SSE2 cubic8 path computes in wrapping i16 lanes then
// packs with unsigned saturation; for u8 inputs the i16 sums stay in range
// (|v| <= 2*255 + 40*255 + 10*255 + 16 < 2^15) so it equals `half` per pixel.
void fill_bicubic_pel2(Buf dst, size_t stride, size_t row0, size_t w, size_t h) {
    size_t planes[4] = {row0, row0 + h, row0 + 2 * h, row0 + 3 * h};
    auto horiz = [&](size_t s, size_t d) {
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) dst.p[(d + y) * stride + x] = half(dst, (s + y) * stride + x, 1, x, w);
    };
    auto vert = [&](size_t s, size_t d) {
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) dst.p[(d + y) * stride + x] = half(dst, (s + y) * stride + x, stride, y, h);
    };
    horiz(planes[0], planes[1]);
    vert(planes[0], planes[2]);
    horiz(planes[2], planes[3]);
}

}  // namespace

// super_build.rs:8-93
void build_plane(uint8_t* dstp, size_t dst_len, size_t dst_stride, const uint8_t* src, size_t src_len,
                 size_t src_stride, size_t src_w, size_t src_h, size_t luma_w, size_t luma_h, bool chroma,
                 const SuperOpts& opts) {
    size_t levels = (size_t)opts.levels, pel = (size_t)opts.pel;
    if (levels == 0 || src_w == 0 || src_h == 0) return;
    Buf dst{dstp, dst_len};
    auto level_size = [&](size_t lv, size_t* w, size_t* h) {
        size_t yw = (size_t)reduce_dim((int)luma_w, (int)lv), yh = (size_t)reduce_dim((int)luma_h, (int)lv);
        *w = chroma ? yw / 2 : yw;
        *h = chroma ? yh / 2 : yh;
    };
    std::vector<size_t> level_y(levels);
    size_t y = 0;
    for (size_t lv = 0; lv < levels; lv++) {
        level_y[lv] = y;
        size_t w, h;
        level_size(lv, &w, &h);
        y += h * ((lv == 0 && opts.full) ? pel * pel : 1);
    }
    size_t base_w, base_h;
    level_size(0, &base_w, &base_h);
    for (size_t row = 0; row < std::min(base_h, src_h); row++) {  // :45-67
        size_t doff = (level_y[0] + row) * dst_stride, soff = row * src_stride;
        size_t n = std::min(std::min(base_w, src_w), std::min(dst_len > doff ? dst_len - doff : 0,
                                                               src_len > soff ? src_len - soff : 0));
        if (n > 0) std::memcpy(dstp + doff, src + soff, n);
        if (base_w > 0 && dst_stride > base_w) {
            uint8_t edge = dstp[doff + base_w - 1];
            for (size_t p = base_w; p < dst_stride; p++)
                if (doff + p < dst_len) dstp[doff + p] = edge;
        }
    }
    if (opts.scale_down == 4) {  // :69-84
        for (size_t lv = 0; lv + 1 < levels; lv++) {
            size_t sw, sh, dw, dh;
            level_size(lv, &sw, &sh);
            level_size(lv + 1, &dw, &dh);
            reduce_6tap(dst, dst_stride, level_y[lv + 1], dw, dh, level_y[lv], sw, sh);
        }
    }
    if (pel >= 2 && opts.gpu <= 1 && opts.full) {  // :86-92
        if (opts.scale_up == 0) fill_bilinear_pel2(dst, dst_stride, level_y[0], base_w, base_h);
        else if (opts.scale_up == 2) fill_bicubic_pel2(dst, dst_stride, level_y[0], base_w, base_h);
    }
}

}  // namespace oracle
