// Oracle restatement of crates/svpflow-core/src/renderer.rs (scalar semantics; the
// SSE2/SSE4.1 tile kernels of mode23_fast compute the same bytes), metadata.rs and
// frame_math.rs.  Test infrastructure only (see oracle.h).
#include "render.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace oracle {

namespace {

constexpr int MAX_VECTOR = 1023;
constexpr int LUT_CENTER = 1024;
constexpr int MAX_BLOCK = 32;
const int CPU_MAP[32] = {0, 1, 2, 2, -1, 3, 3, 3, -1, -1, -1, 4, -1, 4, -1, 4,
                         -1, -1, -1, -1, -1, -1, -1, -1, 5, -1, -1, -1, 5, -1, -1, 5};

int cpu_map(int v) {  // renderer.rs:3186-3193
    int i = v - 1;
    return (i >= 0 && i < 32) ? CPU_MAP[i] : -1;
}
int16_t trunc_div_256(int v) {  // renderer.rs:3195-3197
    int q = v / 256;
    if (q > 32767) return 32767;
    if (q < -32768) return -32768;
    return (int16_t)q;
}
int32_t trunc_f64(double v) {  // `as i32`
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}
inline uint8_t byte_from_i32(int v) { return (uint8_t)std::min(255, std::max(0, v)); }
inline uint8_t blend_255(uint8_t a, uint8_t b, int w) { return byte_from_i32(((int)a * (255 - w) + (int)b * w + 255) >> 8); }
inline uint8_t blend_256(uint8_t a, uint8_t b, int w) { return byte_from_i32(((int)a * (256 - w) + (int)b * w) >> 8); }
inline uint8_t clamp_between(uint8_t v, uint8_t a, uint8_t b) {
    uint8_t lo = std::min(a, b), hi = std::max(a, b);
    return v < lo ? lo : (v > hi ? hi : v);
}

uint16_t centered(int v) {  // renderer.rs:3237-3243
    if (std::abs(v) > MAX_VECTOR) return NEUTRAL;
    return (uint16_t)(v + NEUTRAL);
}

int clamp_to_frame(int value, int pos, int block, int frame) {  // renderer.rs:1738-1746
    if (value + pos < 0) return -pos;
    if (value + block + pos > frame) return std::max(frame - block - pos, 0);
    return value;
}

struct TruncDiv {  // renderer.rs:3149-3184
    int divisor;
    uint64_t magic = 0, bound = 0;
    explicit TruncDiv(int d) : divisor(d) {
        if (d > 0) {
            uint64_t du = (uint64_t)d;
            magic = ((1ull << 40) + du - 1) / du;
            bound = std::min<uint64_t>(std::min<uint64_t>((1ull << 40) / du, UINT64_MAX / magic), 1ull << 31);
        }
    }
    int div(int v) const {
        uint64_t a = (uint64_t)std::abs((int64_t)v);
        if (a < bound) {
            int q = (int)((a * magic) >> 40);
            return v < 0 ? -q : q;
        }
        if (divisor == 0) return 0;
        return v / divisor;
    }
};

struct PlaneParams {  // renderer.rs:1876-1892
    int width, height, block_w, block_h, origin_x, origin_y, x_shift, y_shift, x_div, y_div, source_step, max_x,
        max_y;
    int xw[MAX_BLOCK][2], yw[MAX_BLOCK][2];
};

struct Tile {  // renderer.rs:1894-1904
    int x, y, width, height, local_y;
    size_t i00, i01, i10, i11;
};

int coeff(int value, int total, int scale) {  // renderer.rs:2766-2770
    return trunc_f64(std::ceil((double)(value * scale) / (double)total - 0.5));
}

void interpolation_weights(int total, int shift, int out[MAX_BLOCK][2]) {  // renderer.rs:2740-2764
    for (int i = 0; i < MAX_BLOCK; i++) out[i][0] = out[i][1] = 0;
    if (total <= 0) return;
    int sh = std::max(shift, 0);
    int scale = sh < 32 ? (1 << sh) : 0;
    for (int p = 0; p < std::min(total, MAX_BLOCK); p++) {
        if (shift < 0) { out[p][0] = total - p; out[p][1] = p; }
        else if (scale == 0) { out[p][0] = 1; out[p][1] = 0; }
        else { out[p][0] = coeff(total - p, total, scale); out[p][1] = coeff(p, total, scale); }
    }
}

inline int interp_pair(int a, int b, int total, int shift, const int w[2]) {  // renderer.rs:2724-2738
    if (shift < 0) return (w[0] * a + w[1] * b) / total;
    return (w[0] * a + w[1] * b) >> shift;
}

// render_tiles, renderer.rs:1925-2010
template <class F>
void render_tiles(const PlaneParams& p, size_t grid_w, size_t grid_h, int row_start, int row_end, F render) {
    if (p.width <= 0 || p.height <= 0 || p.block_w <= 0 || p.block_h <= 0 || grid_w == 0 || grid_h == 0) return;
    int gw = (int)grid_w, gh = (int)grid_h;
    for (int gy = -1; gy < gh; gy++) {
        int base_y = gy < 0 ? 0 : p.origin_y + p.block_h * gy;
        int height = gy < 0 ? p.origin_y : p.block_h;
        if (base_y + p.block_h >= p.height || gy == gh - 1) height = p.height - base_y;
        int end_y = std::min(base_y + height, row_end);
        int clipped_y = std::max(base_y, row_start);
        height = end_y - clipped_y;
        if (height <= 0) continue;
        for (int gx = -1; gx < gw; gx++) {
            int base_x = gx < 0 ? 0 : p.origin_x + p.block_w * gx;
            int width = gx < 0 ? p.origin_x : p.block_w;
            if (base_x + p.block_w >= p.width || gx == gw - 1) width = p.width - base_x;
            if (width <= 0) continue;
            size_t x0 = (size_t)std::max(gx, 0), y0 = (size_t)std::max(gy, 0);
            size_t nx = (gx >= 0 && gx != gw - 1) ? 1 : 0;
            size_t ny = (gy >= 0 && gy != gh - 1) ? grid_w : 0;
            size_t i00 = y0 * grid_w + x0;
            render(Tile{base_x, clipped_y, width, height, clipped_y - base_y, i00, i00 + nx, i00 + ny, i00 + ny + nx});
        }
    }
}

struct Samples {
    int x[4], y[4];
};
struct Row2 {
    int x[2], y[2];
};

int motion_value(const int16_t* lut, const std::vector<uint16_t>& plane, size_t i) {  // renderer.rs:2577-2580
    size_t li = std::min<size_t>(i < plane.size() ? plane[i] : NEUTRAL, 2047);
    return lut[li];
}

Samples motion_samples(const Motion& m, const int16_t* lut, const PlaneParams& p, const Tile& t, bool interp) {
    Samples s;
    int x0 = motion_value(lut, *m.x, t.i00) / p.x_div, y0 = motion_value(lut, *m.y, t.i00) / p.y_div;
    if (!interp) {
        for (int k = 0; k < 4; k++) { s.x[k] = x0; s.y[k] = y0; }
        return s;
    }
    size_t idx[4] = {t.i00, t.i01, t.i10, t.i11};
    for (int k = 0; k < 4; k++) {
        s.x[k] = motion_value(lut, *m.x, idx[k]) / p.x_div;
        s.y[k] = motion_value(lut, *m.y, idx[k]) / p.y_div;
    }
    return s;
}

void interp_row(const int s[4], const PlaneParams& p, int y, bool interp, int out[2]) {  // renderer.rs:2582-2607
    if (!interp) { out[0] = out[1] = s[0]; return; }
    int yy = std::min(std::max(y, 0), MAX_BLOCK - 1);
    out[0] = interp_pair(s[0], s[2], p.block_h, p.y_shift, p.yw[yy]);
    out[1] = interp_pair(s[1], s[3], p.block_h, p.y_shift, p.yw[yy]);
}

Row2 motion_row(const Samples& s, const PlaneParams& p, int y, bool interp) {
    Row2 r;
    interp_row(s.x, p, y, interp, r.x);
    interp_row(s.y, p, y, interp, r.y);
    return r;
}

void interp_motion(const Row2& r, const PlaneParams& p, const int w[2], bool interp, int* dx, int* dy) {
    if (!interp) { *dx = r.x[0]; *dy = r.y[0]; return; }
    *dx = interp_pair(r.x[0], r.x[1], p.block_w, p.x_shift, w);
    *dy = interp_pair(r.y[0], r.y[1], p.block_w, p.x_shift, w);
}

void alpha_row(const std::vector<uint8_t>& m, const PlaneParams& p, const Tile& t, int y, bool interp, int out[2]) {
    auto g = [&](size_t i) { return i < m.size() ? (int)m[i] : 0; };
    int s[4] = {g(t.i00), g(t.i01), g(t.i10), g(t.i11)};
    interp_row(s, p, y, interp, out);
}

uint8_t alpha_at_row(const int row[2], const PlaneParams& p, const int w[2], bool interp) {
    if (!interp) return byte_from_i32(row[0]);
    return byte_from_i32(interp_pair(row[0], row[1], p.block_w, p.x_shift, w));
}

uint8_t mask_at(const std::vector<uint8_t>& m, size_t stride, int x, int y) {
    if (x < 0 || y < 0) return 0;
    size_t i = (size_t)y * stride + (size_t)x;
    return i < m.size() ? m[i] : 0;
}

uint8_t alpha_at_absolute(const std::vector<uint8_t>& m, const PlaneParams& p, size_t stride, int x, int y) {
    if (p.block_w <= 0 || p.block_h <= 0) return 0;
    int ix0 = 0, ix1 = 0, iy0 = 0, iy1 = 0;
    if (p.origin_x <= x) { ix0 = (x - p.origin_x) / p.block_w; ix1 = ix0 + 1; }
    if (p.origin_y <= y) { iy0 = (y - p.origin_y) / p.block_h; iy1 = iy0 + 1; }
    int fx = p.origin_x < x ? (x - p.origin_x) % p.block_w : 0;
    int fy = p.origin_y < y ? (y - p.origin_y) % p.block_h : 0;
    int tl = mask_at(m, stride, ix0, iy0), tr = mask_at(m, stride, ix1, iy0);
    int bl = mask_at(m, stride, ix0, iy1), br = mask_at(m, stride, ix1, iy1);
    int top = ((p.block_w - fx) * tl + fx * tr) / p.block_w;
    int bottom = ((p.block_w - fx) * bl + fx * br) / p.block_w;
    return byte_from_i32(((p.block_h - fy) * top + fy * bottom) / p.block_h);
}

size_t plane_index(const RPlane& pl, size_t x, size_t y) {  // renderer.rs:2782-2789
    if (pl.shift == 0) return y * pl.stride + x;
    size_t mask = (1ull << pl.shift) - 1;
    size_t sub = ((y & mask) << pl.shift) | (x & mask);
    return sub * pl.span + (y >> pl.shift) * pl.stride + (x >> pl.shift);
}

bool plane_covers(const RPlane& pl, const PlaneParams& p) {  // renderer.rs:2772-2780
    if (pl.shift != 0 && (1 << pl.shift) != p.source_step) return false;
    if (p.max_x < 0 || p.max_y < 0) return false;
    return plane_index(pl, (size_t)p.max_x, (size_t)p.max_y) < pl.len;
}

bool tile_interior(const PlaneParams& p, const Tile& t, const Samples* ms, int n) {  // renderer.rs:2808-2834
    if (p.source_step <= 0) return false;
    int mnx = INT32_MAX, mxx = INT32_MIN, mny = INT32_MAX, mxy = INT32_MIN;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) {
            mnx = std::min(mnx, ms[i].x[k]); mxx = std::max(mxx, ms[i].x[k]);
            mny = std::min(mny, ms[i].y[k]); mxy = std::max(mxy, ms[i].y[k]);
        }
    int64_t st = p.source_step;
    int64_t x0 = (int64_t)t.x * st, y0 = (int64_t)t.y * st;
    int64_t x1 = (int64_t)(t.x + t.width - 1) * st, y1 = (int64_t)(t.y + t.height - 1) * st;
    return x0 + mnx >= 0 && y0 + mny >= 0 && x1 + mxx <= p.max_x && y1 + mxy <= p.max_y;
}

uint8_t sample_plane(const RPlane& pl, const PlaneParams& p, int sx, int sy, int dx, int dy, bool direct) {
    int x = sx + dx, y = sy + dy;  // renderer.rs:2836-2856
    if (direct) return pl.data[plane_index(pl, (size_t)x, (size_t)y)];
    size_t cx = (size_t)std::min(std::max(x, 0), p.max_x), cy = (size_t)std::min(std::max(y, 0), p.max_y);
    size_t i = plane_index(pl, cx, cy);
    return i < pl.len ? pl.data[i] : 0;
}

uint8_t* output_row(const RPlaneMut& d, const Tile& t, int y, int row_start) {  // renderer.rs:2888-2899
    int64_t row = (int64_t)t.y + y - row_start;
    if (row < 0 || t.x < 0 || t.width < 0) return nullptr;
    size_t start = (size_t)row * d.stride + (size_t)t.x;
    if (start + (size_t)t.width > d.len) return nullptr;
    return d.data + start;
}

}  // namespace

// ---------------------------------------------------------------- motion planes / masks

void vector_planes(const VectorContext& ctx, int which, std::vector<uint16_t>& dst_x, std::vector<uint16_t>& dst_y,
                   size_t width, size_t height) {  // renderer.rs:1680-1736
    const std::vector<RVector>& v = which == 0 ? *ctx.a : *ctx.b;
    int step_x = ctx.block_w - ctx.origin_x, step_y = ctx.block_h - ctx.origin_y;
    for (size_t y = 0; y < height; y++) {
        int pos_y = (int)y * step_y;
        for (size_t x = 0; x < width; x++) {
            size_t out = y * width + x;
            if (out >= dst_x.size() || out >= dst_y.size()) return;
            size_t sx = std::min(x, ctx.grid_w ? ctx.grid_w - 1 : 0), sy = std::min(y, ctx.grid_h ? ctx.grid_h - 1 : 0);
            size_t si = sy * ctx.grid_w + sx;
            RVector vec = si < v.size() ? v[si] : RVector{};
            int dx = vec.dx, dy = vec.dy;
            if (std::abs(dx) > MAX_VECTOR || std::abs(dy) > MAX_VECTOR) {
                dst_x[out] = NEUTRAL;
                dst_y[out] = NEUTRAL;
                continue;
            }
            if (!ctx.raw) {
                int pos_x = (int)x * step_x;
                dx = clamp_to_frame(dx, pos_x, ctx.block_w, ctx.frame_w);
                dy = clamp_to_frame(dy, pos_y, ctx.block_h, ctx.frame_h);
            }
            dst_x[out] = centered(dx);
            dst_y[out] = centered(dy);
        }
    }
}

void magnitude_mask(const VectorContext& ctx, int which, std::vector<uint8_t>& out, size_t width, size_t height,
                    double scale, double exponent) {  // renderer.rs:1647-1678, 3245-3259
    const std::vector<RVector>& v = which == 0 ? *ctx.a : *ctx.b;
    int area = std::max(ctx.block_w * ctx.block_h, 1);
    for (size_t y = 0; y < height; y++)
        for (size_t x = 0; x < width; x++) {
            size_t i = y * width + x;
            if (i >= out.size()) return;
            if (x < ctx.grid_w && y < ctx.grid_h) {
                size_t si = y * ctx.grid_w + x;
                uint32_t mag = si < v.size() ? v[si].magnitude : 0;
                double val = std::pow(4.0 * (double)mag * scale / (double)area, exponent) * 255.0;
                uint8_t b;
                if (!std::isfinite(val) || val <= 0.0) b = 0;
                else if (val >= 255.0) b = 255;
                else b = (uint8_t)val;
                out[i] = b;
            } else if (x > 0) {
                out[i] = out[i - 1];
            } else if (y > 0) {
                out[i] = out[i - width];
            } else {
                out[i] = 0;
            }
        }
}

void coverage_mask(const VectorContext& ctx, int which, std::vector<uint8_t>& out, size_t width, size_t height,
                   int strength, int threshold) {  // renderer.rs:1738-1763, 2945-3147
    size_t stride = width + 2;
    size_t len = stride * (height + 2);
    std::vector<int32_t> acc(len, 0), sums(len, 0);
    const std::vector<RVector>& vecs = which == 0 ? *ctx.b : *ctx.a;  // opposite set
    auto add_point = [&](int x, int y, int w) {
        if (x < 0 || y < 0) return;
        size_t i = (size_t)y * stride + (size_t)x;
        if (i < acc.size()) {
            int64_t r = (int64_t)acc[i] + w;
            acc[i] = (int32_t)std::min<int64_t>(std::max<int64_t>(r, INT32_MIN), INT32_MAX);
        }
    };
    // splat_coverage
    int grid_w = (int)ctx.grid_w, grid_h = (int)ctx.grid_h, out_w = (int)width, out_h = (int)height;
    int x_step = ctx.block_w - ctx.origin_x, y_step = ctx.block_h - ctx.origin_y;
    int denom = ctx.scale_shift_base << 8;
    if (grid_w > 0 && grid_h > 0 && out_w > 0 && out_h > 0 && x_step != 0 && y_step != 0 && denom != 0) {
        int x_bias = 1 - x_step, y_bias = 1 - y_step;
        TruncDiv dd(denom), xd(x_step), yd(y_step);
        for (int y = 0; y < grid_h; y++) {
            int x_base = 0;
            for (int x = 0; x < grid_w; x++) {
                size_t vi = (size_t)y * ctx.grid_w + (size_t)x;
                RVector v = vi < vecs.size() ? vecs[vi] : RVector{};
                int dx = dd.div(threshold * (int)v.dx), dy = dd.div(threshold * (int)v.dy);
                int sx = x_base + dx, sy = y_step * y + dy;
                int left = xd.div(sx + (x_bias & (sx >> 31)));
                int top = yd.div(sy + (y_bias & (sy >> 31)));
                int right = left + 1, bottom = top + 1;
                int next_x = right * x_step;
                int lw = next_x - sx, tw = y_step * bottom - sy;
                int rw = sx + ctx.block_w - next_x, bw = ctx.block_h - tw;
                bool interior = left >= 0 && right < out_w && top >= 0 && bottom < out_h;
                size_t base = interior ? (size_t)top * stride + (size_t)left : 0;
                if (interior && base + stride + 2 <= acc.size()) {
                    acc[base] += lw * tw;
                    acc[base + 1] += tw * rw;
                    acc[base + stride] += lw * bw;
                    acc[base + stride + 1] += rw * bw;
                } else {
                    if (top >= 0 && top < out_h && left >= 0 && left < out_w) add_point(left, top, lw * tw);
                    if (left >= -1 && top >= 0 && top < out_h && right < out_w) add_point(right, top, tw * rw);
                    if (top >= -1 && bottom < out_h && left >= -1 && right < out_w) add_point(right, bottom, rw * bw);
                    if (left >= 0 && top >= -1 && bottom < out_h && left < out_w) add_point(left, bottom, lw * bw);
                }
                x_base += x_step;
            }
        }
    }
    // window_3x3
    if (stride < 2) {
        sums = acc;
    } else {
        size_t rows = len / stride;
        for (size_t r = 0; r < rows; r++) {
            const int32_t* s = &acc[r * stride];
            int32_t* d = &sums[r * stride];
            d[0] = s[0];
            d[1] = s[1] + s[0];
            for (size_t c = 2; c < stride; c++) d[c] = s[c] + s[c - 1] + s[c - 2];
        }
        for (size_t y = rows - 1; y >= 1; y--) {
            int32_t* cur = &sums[y * stride];
            const int32_t* r1 = &sums[(y - 1) * stride];
            if (y >= 2) {
                const int32_t* r2 = &sums[(y - 2) * stride];
                for (size_t c = 0; c < stride; c++) cur[c] += r1[c] + r2[c];
            } else {
                for (size_t c = 0; c < stride; c++) cur[c] += r1[c];
            }
        }
    }
    // finish_coverage
    int area = std::max(ctx.block_w * ctx.block_h, 1);
    double st = (double)strength / 100.0, af = (double)area;
    auto cbyte = [&](int remaining) -> uint8_t {
        double v = (double)remaining * st * 256.0 / af;
        if (!std::isfinite(v)) return 0;
        int iv = trunc_f64(v);
        return iv >= 255 ? 255 : (uint8_t)(iv & 0xFF);
    };
    for (size_t y = 0; y < height; y++) {
        size_t os = y * width, as = (y + 1) * stride + 1;
        if (os + width > out.size() || as + width > sums.size()) return;
        for (size_t x = 0; x < width; x++) {
            int covered = sums[as + x] >> 3;
            int remaining = area <= covered ? 0 : area - covered;
            out[os + x] = cbyte(remaining);
        }
    }
}

std::vector<uint8_t> max_mask(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
    std::vector<uint8_t> m(std::min(a.size(), b.size()));
    for (size_t i = 0; i < m.size(); i++) m[i] = std::max(a[i], b[i]);
    return m;
}

// ---------------------------------------------------------------- CpuRenderer

CpuRenderer::CpuRenderer(const CpuConfig& c) : cfg_(c) {  // renderer.rs:211-233
    wmap_ = cpu_map(c.block_w);
    hmap_ = cpu_map(c.block_h);
    hwmap_ = cpu_map(c.block_w / 2);
    hhmap_ = cpu_map(c.block_h / 2);
    std::memset(inverse_, 0, sizeof inverse_);
    std::memset(tlut_, 0, sizeof tlut_);
    set_threshold(0);
}

void CpuRenderer::set_threshold(int t) {  // renderer.rs:235-264
    threshold_ = t;
    int inv = 256 - t;
    limit_ = t > 126 ? trunc_f64(256.0 - (double)inv * cfg_.scale) : trunc_f64((double)t * cfg_.scale);
    for (int i = 0; i < 2048; i++) {
        int s = i - LUT_CENTER;
        inverse_[i] = trunc_div_256(s * inv);
        tlut_[i] = trunc_div_256(s * t);
    }
}

namespace {

PlaneParams make_params(const CpuConfig& c, int wmap, int hmap, int hwmap, int hhmap, bool chroma, bool zero_origin) {
    PlaneParams p;  // renderer.rs:1539-1606
    int ox, oy;
    if (zero_origin) { ox = 0; oy = 0; }
    else if (chroma) { ox = c.origin_x / 2; oy = c.origin_y / c.chroma_y_div; }
    else { ox = c.origin_x; oy = c.origin_y; }
    if (chroma) {
        p.width = c.width / 2; p.height = c.height / c.chroma_y_div;
        p.block_w = c.block_w / 2; p.block_h = c.block_h / c.chroma_y_div;
        p.x_shift = hwmap; p.y_shift = hhmap; p.x_div = 2; p.y_div = c.chroma_y_div;
    } else {
        p.width = c.width; p.height = c.height; p.block_w = c.block_w; p.block_h = c.block_h;
        p.x_shift = wmap; p.y_shift = hmap; p.x_div = 1; p.y_div = 1;
    }
    p.origin_x = ox; p.origin_y = oy;
    p.source_step = c.source_step;
    p.max_x = std::max(p.width * c.source_step - 1, 0);
    p.max_y = std::max(p.height * c.source_step - 1, 0);
    interpolation_weights(p.block_w, p.x_shift, p.xw);
    interpolation_weights(p.block_h, p.y_shift, p.yw);
    return p;
}

// Shared per-pixel walk of render_selected_warp_rows / render_dual_warp_rows /
// render_mode23_plane_rows_inner (renderer.rs:881-982, 1030-1171, 1294-1537).
struct Walk {
    const PlaneParams& p;
    const PlaneParams* mp;  // mask params when zero_origin
    size_t grid_w;
    bool interp;
};

}  // namespace

void CpuRenderer::mode1_or_2(bool mode1, bool interp, RPlaneMut dst[3], const RPlane s0[3], const RPlane s1[3],
                             Motion m0, Motion m1, const std::vector<uint8_t>* mask_a,
                             const std::vector<uint8_t>* mask_b) {
    const RPlane* src = mode1 ? s1 : s0;  // renderer.rs:319-323
    Motion m = mode1 ? m1 : m0;
    const int16_t* lut = mode1 ? inverse_ : tlut_;
    const std::vector<uint8_t>* mask = mode1 ? mask_a : mask_b;
    for (int pi = 0; pi < 3; pi++) {
        bool chroma = pi != 0;
        bool zero = !interp;
        PlaneParams p = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, zero);
        PlaneParams mp = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, false);
        int height = chroma ? cfg_.height / cfg_.chroma_y_div : cfg_.height;
        const RPlane& s = src[pi];
        render_tiles(p, cfg_.grid_w, cfg_.grid_h, 0, height, [&](const Tile& t) {
            Samples ms = motion_samples(m, lut, p, t, interp);
            bool direct = plane_covers(s, p) && tile_interior(p, t, &ms, 1);
            for (int y = 0; y < t.height; y++) {
                int ly = t.local_y + y;
                int sy = (t.y + y) * p.source_step;
                uint8_t* out = output_row(dst[pi], t, y, 0);
                if (!out) continue;
                Row2 mr = motion_row(ms, p, ly, interp);
                int arow[2] = {0, 0};
                if (mask) alpha_row(*mask, p, t, ly, interp, arow);
                for (int x = 0; x < t.width; x++) {
                    int px = t.x + x, py = t.y + y, sx = px * p.source_step;
                    const int* w = p.xw[std::min(x, MAX_BLOCK - 1)];
                    int dx, dy;
                    interp_motion(mr, p, w, interp, &dx, &dy);
                    uint8_t base = sample_plane(s, p, sx, sy, 0, 0, direct);
                    uint8_t warped = sample_plane(s, p, sx, sy, dx, dy, direct);
                    uint8_t v = warped;
                    if (mask) {
                        uint8_t a = zero ? alpha_at_absolute(*mask, mp, cfg_.grid_w, px, py) : alpha_at_row(arow, p, w, interp);
                        v = blend_255(warped, base, a);  // mode1_pixel
                    }
                    out[x] = v;
                }
            }
        });
    }
}

namespace {

// render_dual_warp_rows with a pixel functor (cur0, cur1, mv0, mv1, a0, a1, a2 flags).
template <class Pixel>
void dual_warp(const CpuConfig& cfg, const PlaneParams& p, const PlaneParams& mp, bool zero, bool interp,
               RPlaneMut dst, const RPlane& s0, const RPlane& s1, Motion m0, Motion m1, const int16_t* tlut,
               const int16_t* ilut, const std::vector<uint8_t>* k0, const std::vector<uint8_t>* k1,
               const std::vector<uint8_t>* k2, int height, Pixel pixel) {
    render_tiles(p, cfg.grid_w, cfg.grid_h, 0, height, [&](const Tile& t) {
        Samples ms[2] = {motion_samples(m0, tlut, p, t, interp), motion_samples(m1, ilut, p, t, interp)};
        bool direct = plane_covers(s0, p) && plane_covers(s1, p) && tile_interior(p, t, ms, 2);
        for (int y = 0; y < t.height; y++) {
            int ly = t.local_y + y;
            int sy = (t.y + y) * p.source_step;
            uint8_t* out = output_row(dst, t, y, 0);
            if (!out) continue;
            Row2 r0 = motion_row(ms[0], p, ly, interp), r1 = motion_row(ms[1], p, ly, interp);
            int a0r[2] = {0, 0}, a1r[2] = {0, 0}, a2r[2] = {0, 0};
            if (k0) alpha_row(*k0, p, t, ly, interp, a0r);
            if (k1) alpha_row(*k1, p, t, ly, interp, a1r);
            if (k2) alpha_row(*k2, p, t, ly, interp, a2r);
            for (int x = 0; x < t.width; x++) {
                int px = t.x + x, py = t.y + y, sx = px * p.source_step;
                const int* w = p.xw[std::min(x, MAX_BLOCK - 1)];
                int dx0, dy0, dx1, dy1;
                interp_motion(r0, p, w, interp, &dx0, &dy0);
                interp_motion(r1, p, w, interp, &dx1, &dy1);
                uint8_t cur0 = sample_plane(s0, p, sx, sy, 0, 0, direct);
                uint8_t cur1 = sample_plane(s1, p, sx, sy, 0, 0, direct);
                uint8_t mv0 = sample_plane(s0, p, sx, sy, dx0, dy0, direct);
                uint8_t mv1 = sample_plane(s1, p, sx, sy, dx1, dy1, direct);
                auto alpha = [&](const std::vector<uint8_t>* k, const int row[2]) -> int {
                    if (!k) return -1;
                    return zero ? alpha_at_absolute(*k, mp, cfg.grid_w, px, py) : alpha_at_row(row, p, w, interp);
                };
                out[x] = pixel(cur0, cur1, mv0, mv1, alpha(k0, a0r), alpha(k1, a1r), alpha(k2, a2r));
            }
        }
    });
}

}  // namespace

void CpuRenderer::mode11_or_13(bool mode13, bool interp, bool zero_origin, RPlaneMut dst[3], const RPlane s0[3],
                               const RPlane s1[3], Motion m0, Motion m1, const std::vector<uint8_t>* max) {
    const int t = threshold_, lim = limit_;
    (void)lim;
    auto pixel = [&](uint8_t c0, uint8_t c1, uint8_t v0, uint8_t v1, int a, int, int) -> uint8_t {
        // mode11_or_13_pixel (renderer.rs:2918-2942)
        uint8_t motion = mode13 ? clamp_between(blend_256(c0, c1, t), v0, v1) : blend_256(v0, v1, t);
        if (a >= 0) return blend_255(motion, blend_256(c0, c1, t), a);
        return motion;
    };
    for (int pi = 0; pi < 3; pi++) {
        bool chroma = pi != 0;
        PlaneParams p = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, zero_origin);
        PlaneParams mp = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, false);
        int height = chroma ? cfg_.height / cfg_.chroma_y_div : cfg_.height;
        dual_warp(cfg_, p, mp, zero_origin, interp, dst[pi], s0[pi], s1[pi], m0, m1, tlut_, inverse_, max, nullptr,
                  nullptr, height, pixel);
    }
}

void CpuRenderer::mode21_or_22(bool mode21, bool interp, RPlaneMut dst[3], const RPlane s0[3], const RPlane s1[3],
                               Motion m0, Motion m1, const std::vector<uint8_t>& ca, const std::vector<uint8_t>& cb,
                               const std::vector<uint8_t>* final_mask) {
    const int t = threshold_, lim = limit_;
    auto pixel = [&](uint8_t c0, uint8_t c1, uint8_t v0, uint8_t v1, int a0, int a1, int af) -> uint8_t {
        // mode21_or_22_pixel (renderer.rs:617-641) with mode21_pixel / mode22_pixel (:1830-1848)
        int al0 = a0 < 0 ? 0 : a0, al1 = a1 < 0 ? 0 : a1;
        uint8_t a = blend_255(v0, v1, al0), b = blend_255(v1, v0, al1);
        uint8_t motion = mode21 ? blend_256(a, b, t) : clamp_between(blend_256(c0, c1, t), a, b);
        if (af >= 0) return blend_255(motion, blend_256(c0, c1, lim), af);
        return motion;
    };
    bool zero = !interp;
    for (int pi = 0; pi < 3; pi++) {
        bool chroma = pi != 0;
        PlaneParams p = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, zero);
        PlaneParams mp = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, false);
        int height = chroma ? cfg_.height / cfg_.chroma_y_div : cfg_.height;
        dual_warp(cfg_, p, mp, zero, interp, dst[pi], s0[pi], s1[pi], m0, m1, tlut_, inverse_, &ca, &cb, final_mask,
                  height, pixel);
    }
}

void CpuRenderer::mode23(bool interp, RPlaneMut dst[3], const RPlane s0[3], const RPlane s1[3], Motion m0, Motion m1,
                         Motion m2, Motion m3, const std::vector<uint8_t>& ca, const std::vector<uint8_t>& cb,
                         const std::vector<uint8_t>* final_mask) {
    const int q0 = threshold_;
    bool zero = !interp;
    for (int pi = 0; pi < 3; pi++) {
        bool chroma = pi != 0;
        PlaneParams p = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, zero);
        PlaneParams mp = make_params(cfg_, wmap_, hmap_, hwmap_, hhmap_, chroma, false);
        int height = chroma ? cfg_.height / cfg_.chroma_y_div : cfg_.height;
        const RPlane& a0 = s0[pi];
        const RPlane& a1 = s1[pi];
        render_tiles(p, cfg_.grid_w, cfg_.grid_h, 0, height, [&](const Tile& t) {
            Samples ms[4] = {motion_samples(m0, tlut_, p, t, interp), motion_samples(m1, inverse_, p, t, interp),
                             motion_samples(m2, tlut_, p, t, interp), motion_samples(m3, inverse_, p, t, interp)};
            bool direct = plane_covers(a0, p) && plane_covers(a1, p) && tile_interior(p, t, ms, 4);
            for (int y = 0; y < t.height; y++) {
                int ly = t.local_y + y;
                int sy = (t.y + y) * p.source_step;
                uint8_t* out = output_row(dst[pi], t, y, 0);
                if (!out) continue;
                Row2 r[4];
                for (int k = 0; k < 4; k++) r[k] = motion_row(ms[k], p, ly, interp);
                int c0r[2], c1r[2], fr[2] = {0, 0};
                alpha_row(ca, p, t, ly, interp, c0r);
                alpha_row(cb, p, t, ly, interp, c1r);
                if (final_mask) alpha_row(*final_mask, p, t, ly, interp, fr);
                for (int x = 0; x < t.width; x++) {
                    int px = t.x + x, py = t.y + y, sx = px * p.source_step;
                    const int* w = p.xw[std::min(x, MAX_BLOCK - 1)];
                    int d[4][2];
                    for (int k = 0; k < 4; k++) interp_motion(r[k], p, w, interp, &d[k][0], &d[k][1]);
                    auto alpha = [&](const std::vector<uint8_t>& k, const int row[2]) -> uint8_t {
                        // alpha_for_origin (renderer.rs:2680-2694)
                        return zero ? alpha_at_absolute(k, mp, cfg_.grid_w, px, py) : alpha_at_row(row, p, w, true);
                    };
                    uint8_t A = sample_plane(a0, p, sx, sy, d[0][0], d[0][1], direct);
                    uint8_t B = sample_plane(a1, p, sx, sy, d[1][0], d[1][1], direct);
                    uint8_t Cc = sample_plane(a0, p, sx, sy, d[2][0], d[2][1], direct);
                    uint8_t D = sample_plane(a1, p, sx, sy, d[3][0], d[3][1], direct);
                    uint8_t t0 = alpha(ca, c0r), t1 = alpha(cb, c1r);
                    // mode23_pixel (renderer.rs:1850-1868)
                    uint8_t i0 = blend_255(A, clamp_between(D, A, B), t0);
                    uint8_t i1 = blend_255(B, clamp_between(Cc, A, B), t1);
                    uint8_t motion = blend_256(i0, i1, q0);
                    if (final_mask) {
                        uint8_t fm = alpha(*final_mask, fr);
                        uint8_t ba = sample_plane(a0, p, sx, sy, 0, 0, direct);
                        uint8_t bb = sample_plane(a1, p, sx, sy, 0, 0, direct);
                        motion = blend_255(motion, blend_256(ba, bb, q0), fm);
                    }
                    out[x] = motion;
                }
            }
        });
    }
}

// ---------------------------------------------------------------- metadata

int VectorData::motion_w() const {  // metadata.rs:211-244 extended_grid
    int span = (block_w - overlap_x) * grid_w + overlap_x;
    return grid_w + (span < width ? 1 : 0);
}
int VectorData::motion_h() const {
    int span = (block_h - overlap_y) * grid_h + overlap_y;
    return grid_h + (span < height ? 1 : 0);
}

bool parse_vdata(const int32_t h[16], VectorData* o) {  // metadata.rs:75-114 (byte offsets / 4)
    if (h[0] != 0xA0) return false;
    o->flags = h[1]; o->block_w = h[2]; o->block_h = h[3]; o->marker = h[4];
    o->width = h[7]; o->height = h[8]; o->overlap_x = h[9]; o->overlap_y = h[10];
    o->grid_w = h[11]; o->grid_h = h[12]; o->selector = h[13]; o->delta = h[14];
    return true;
}

bool decode_vectors(const uint8_t* payload, size_t len, int mode_flags, int gw, int gh, int flags, bool* has_prev,
                    std::vector<DVec>* prev, bool* has_cur, std::vector<DVec>* cur) {  // metadata.rs:246-273, 460-481
    auto rd = [&](size_t off, int32_t* v) {
        if (off + 4 > len) return false;
        std::memcpy(v, payload + off, 4);
        return true;
    };
    int32_t a, b, c;
    if (!rd(0, &a) || !rd(4, &b) || !rd(8, &c) || a != 16 || b != 0xA0 || c != mode_flags) return false;
    size_t count = (size_t)std::max(gw, 0) * (size_t)std::max(gh, 0);
    size_t first = 0x40, second = first + 4 + count * 8;
    auto region = [&](size_t off, std::vector<DVec>* out) {
        int32_t m;
        if (!rd(off, &m) || m != (int32_t)(2 * count + 1)) return false;
        out->resize(count);
        for (size_t i = 0; i < count; i++) {
            uint32_t r0, r1;
            if (off + 4 + i * 8 + 8 > len) return false;
            std::memcpy(&r0, payload + off + 4 + i * 8, 4);
            std::memcpy(&r1, payload + off + 8 + i * 8, 4);
            uint32_t rot = (r0 >> 16) | (r0 << 16);
            (*out)[i].dx = (int16_t)(rot & 0xFFFF);
            (*out)[i].dy = (int16_t)(rot >> 16);
            (*out)[i].score = r1 & 0x00FFFFFF;
            (*out)[i].luma = (uint8_t)(r1 >> 24);
        }
        return true;
    };
    *has_prev = (flags & 1) != 0;
    *has_cur = (flags & 2) != 0;
    if (*has_prev && !region(first, prev)) return false;
    if (*has_cur && !region(second, cur)) return false;
    return true;
}

// ---------------------------------------------------------------- Timing (frame_math.rs)

namespace {
int64_t gcd64(int64_t a, int64_t b) {
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a;
}
int64_t sat_mul64(int64_t a, int64_t b) {
    __int128 r = (__int128)a * b;
    if (r > INT64_MAX) return INT64_MAX;
    if (r < INT64_MIN) return INT64_MIN;
    return (int64_t)r;
}
int clamp_i32_64(int64_t v) { return (int)std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX); }
}  // namespace

Timing Timing::make(int64_t snum, int64_t sden, bool absolute, int64_t rnum, int64_t rden) {
    auto nz = [](int64_t v) { return v == 0 ? 1 : v; };
    auto normalize = [](int64_t n, int64_t d, int64_t* on, int64_t* od) {  // frame_math.rs:142-151
        if (n == 0 || d == 0) { *on = 0; *od = 1; return; }
        int64_t g = gcd64(n, d);
        *on = n / g;
        *od = d / g;
    };
    rnum = nz(rnum);
    rden = nz(rden);
    Timing t;
    if (absolute) normalize(rnum, rden, &t.fps_num, &t.fps_den);
    else normalize(sat_mul64(snum, rnum), sat_mul64(sden, rden), &t.fps_num, &t.fps_den);
    normalize(sat_mul64(sden, t.fps_num), sat_mul64(snum, t.fps_den), &t.frame_num, &t.frame_den);
    return t;
}

int Timing::source_frame(int frame, bool nearest) const {  // frame_math.rs:39-49
    if (frame_num == 0) return frame;
    int64_t scaled = sat_mul64(frame, frame_den);
    if (nearest) return clamp_i32_64((scaled + frame_num / 2) / frame_num);
    return clamp_i32_64(scaled / frame_num);
}

int Timing::phase_256(int frame, int sf) const {  // frame_math.rs:51-67
    if (frame_num <= 0) return 0;
    __int128 scaled = (__int128)frame * frame_den;
    __int128 whole = (__int128)sf * frame_num;
    __int128 rem = scaled - whole;
    if (rem <= 0) return 0;
    __int128 v = (rem * 512 + frame_num) / ((__int128)frame_num * 2);
    return (int)std::min<__int128>(std::max<__int128>(v, INT32_MIN), INT32_MAX);
}

double Timing::raw_phase_256(int frame, int sf) const {  // frame_math.rs:69-73
    return ((double)frame * (double)frame_den / (double)frame_num - (double)sf) * 256.0;
}

int Timing::scene_phase_256(double raw, int64_t mode) const {  // frame_math.rs:88-126
    double step = (double)frame_den * 256.0 / (double)frame_num;
    if (!std::isfinite(raw) || !std::isfinite(step) || step <= 0.0) return 0;
    auto tr = [](double v) { return trunc_f64(v); };
    double lf = std::floor(raw / step);
    double lrem = raw - lf * step;
    int left_floor = tr(lf);
    int left;
    if (mode > 1) left = tr(lrem) > 0 ? left_floor : std::max(left_floor - 1, 0);
    else left = left_floor + (tr(lrem) >= tr(std::fabs(lrem - step)) ? 1 : 0);
    double rf = std::floor((256.0 - raw - 0.001) / step);
    double rrem = 256.0 - (rf * step + raw);
    int right = tr(rf) + (mode > 1 ? (std::fabs(rrem - step) < 0.1 ? 1 : 0) : (tr(rrem) > tr(std::fabs(rrem - step)) ? 1 : 0));
    int total = left + right;
    if (total == 0) return 0;
    int phase = left * 256 / total;
    if ((mode & ~2) == 0) return phase;
    if (left > right) return 256 - (right * 256 / total) / 2;
    return phase / 2;
}

int Timing::scale_frame_count(int frames) const {  // frame_math.rs:128-135
    if (frames <= 0 || frame_den == 0) return frames;
    return clamp_i32_64(sat_mul64(frames - 1, frame_num) / frame_den + 1);
}

int truncdiv(int divisor, int value) { return TruncDiv(divisor).div(value); }  // renderer.rs:3149-3184

}  // namespace oracle
