// SmoothFps render for gfx950 (replaces the CPU render path of crates/svpflow2/src/core.rs
// :886-1770 over crates/svpflow-core/src/renderer.rs and metadata.rs).
//
// Per source frame (prep_kernel + prep_finish_kernel): decode the vector payload,
// build the two motion planes (vector_planes), the magnitude masks and the scene
// class (luma_map + classify_scene) -- all integer, computed once per source frame.
// Per output frame (plan_kernel): resolve the scene class, adaptive phase and the
// effective algorithm.  Then coverage masks (splat with integer atomics, 3x3 box,
// LUT) and one pixel-parallel render kernel that gathers from the pel-expanded super
// planes.  Every value is integer except the per-frame threshold limit (IEEE double,
// as the reference) and the magnitude-mask power (exact for the default exponent 1).
#include "smooth_dev.h"
#include "svpf_internal.h"

namespace svpf {
namespace {

constexpr uint16_t kNeutral = 1024;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ uint8_t byte_of(int v) { return (uint8_t)clampi(v, 0, 255); }
__device__ __forceinline__ uint8_t blend255(int a, int b, int w) { return byte_of((a * (255 - w) + b * w + 255) >> 8); }
__device__ __forceinline__ uint8_t blend256(int a, int b, int w) { return byte_of((a * (256 - w) + b * w) >> 8); }
__device__ __forceinline__ uint8_t clamp_between(int v, int a, int b) {
    const int lo = min(a, b), hi = max(a, b);
    return (uint8_t)(v < lo ? lo : (v > hi ? hi : v));
}
__device__ __forceinline__ int trunc_div_256(int v) { return clampi(v / 256, -32768, 32767); }
__device__ __forceinline__ int f64_to_i32(double v) {
    if (isnan(v)) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int)v;
}
__device__ __forceinline__ uint32_t rd32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// ------------------------------------------------------------------ prep

// decode_vectors (metadata.rs:246-273) validity: header words and region markers.
__device__ bool payload_valid(const uint8_t* pay, int len, int flags, int count) {
    if (len < 12) return false;
    if ((int)rd32(pay) != 16 || (int)rd32(pay + 4) != 0xA0 || (int)rd32(pay + 8) != flags) return false;
    const int first = 0x40, second = first + 4 + count * 8;
    if (flags & 1) {
        if (first + 4 + count * 8 > len || (int)rd32(pay + first) != 2 * count + 1) return false;
    }
    if (flags & 2) {
        if (second + 4 + count * 8 > len || (int)rd32(pay + second) != 2 * count + 1) return false;
    }
    return true;
}

struct Dec {
    int dx, dy;
    uint32_t score;
    int luma;
};

__device__ __forceinline__ Dec decode_at(const uint8_t* pay, int off, int i) {  // decode_region, metadata.rs:460-481
    const uint32_t r0 = rd32(pay + off + 4 + 8 * i), r1 = rd32(pay + off + 8 + 8 * i);
    Dec d;
    d.dx = (int16_t)(r0 >> 16);
    d.dy = (int16_t)(r0 & 0xFFFF);
    d.score = r1 & 0x00FFFFFFu;
    d.luma = (int)(r1 >> 24);
    return d;
}

// set a = previous-or-current, set b = current-or-previous (core.rs:968-970).
__device__ __forceinline__ void decode_sets(const SmoothGeom& g, const uint8_t* pay, int i, Dec& a, Dec& b) {
    const int first = 0x40, second = first + 4 + g.count * 8;
    const bool hp = g.flags & 1, hc = g.flags & 2;
    Dec p{}, c{};
    if (hp) p = decode_at(pay, first, i);
    if (hc) c = decode_at(pay, second, i);
    a = hp ? p : c;
    b = hc ? c : p;
}

__global__ void prep_clear_kernel(const PrepJob* __restrict__ jobs, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (int k = 0; k < 6; k++) jobs[t].stats[k] = 0;
}

// One thread per vector block: decoded vectors and the scene counters (classify_scene,
// metadata.rs:300-347, over set b with the luma map of sets a and b).
__global__ void prep_vectors_kernel(const PrepJob* __restrict__ jobs, const SmoothGeom* __restrict__ gp) {
    const PrepJob j = jobs[blockIdx.y];
    const SmoothGeom& g = *gp;
    const bool valid = payload_valid(j.payload, g.payload_len, g.flags, g.count);
    if (blockIdx.x == 0 && threadIdx.x == 0 && valid) j.stats[0] = 1;
    if (!valid) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.count; i += gridDim.x * blockDim.x) {
        Dec a, b;
        decode_sets(g, j.payload, i, a, b);
        j.vec[0][2 * i] = (int16_t)a.dx;
        j.vec[0][2 * i + 1] = (int16_t)a.dy;
        j.vec[1][2 * i] = (int16_t)b.dx;
        j.vec[1][2 * i + 1] = (int16_t)b.dy;
        const int luma = max((int)g.luma_lut[a.luma + b.luma], 1);
        const long long s255 = (long long)b.score * 255;
        const int score = (int)(s255 > 2147483647LL ? 2147483647LL : s255) / luma;
        // the frame's class counts (classify_scene): one add per counter per wave, the
        // wave's cells counted by ballot (every cell adding to one word would serialise)
        const bool considered = score >= g.lim_zero;
        const bool scene = considered && score >= g.lim_scene;
        const bool m2 = considered && !scene && score >= g.lim_m2;
        const bool m1 = considered && !scene && !m2 && score >= g.lim_m1;
        const uint64_t bc = __ballot(considered), bs = __ballot(scene), b2 = __ballot(m2), b1 = __ballot(m1);
        if (bc != 0 && (threadIdx.x & 63) == (unsigned)__builtin_ctzll(bc)) {
            atomicAdd(&j.stats[2], (int)__popcll(bc));
            if (bs) atomicAdd(&j.stats[3], (int)__popcll(bs));
            if (b2) atomicAdd(&j.stats[4], (int)__popcll(b2));
            if (b1) atomicAdd(&j.stats[5], (int)__popcll(b1));
        }
    }
}

__device__ __forceinline__ uint16_t centered(int v) { return abs(v) > 1023 ? kNeutral : (uint16_t)(v + kNeutral); }
__device__ __forceinline__ int clamp_to_frame(int v, int pos, int block, int frame) {  // renderer.rs:1738-1746
    if (v + pos < 0) return -pos;
    if (v + block + pos > frame) return max(frame - block - pos, 0);
    return v;
}

// scale_magnitude (renderer.rs:3245-3259).  With the exponent 1 the power is the base
// itself (pow(x, 1) == x), so the IEEE products and quotient below are the whole
// computation; any other exponent reads the host-built table (magnitudes are 24-bit).
__device__ uint8_t scale_magnitude(uint32_t mag, int area, const SmoothGeom& g) {
    if (g.mag_lut) return g.mag_lut[mag & 0xFFFFFFu];
    const double v = __dmul_rn(__ddiv_rn(__dmul_rn(__dmul_rn(4.0, (double)mag), g.area_scale), (double)area), 255.0);
    if (!isfinite(v) || v <= 0.0) return 0;
    if (v >= 255.0) return 255;
    return (uint8_t)v;
}

// One thread per motion-grid cell: vector_planes (renderer.rs:1680-1736) for sets a and b
// and magnitude_mask (:1647-1678), whose cells past the grid repeat the left (or, in
// column 0, the upper) neighbour.
__global__ void prep_planes_kernel(const PrepJob* __restrict__ jobs, const SmoothGeom* __restrict__ gp) {
    const PrepJob j = jobs[blockIdx.y];
    const SmoothGeom& g = *gp;
    const int mc = g.mw * g.mh;
    const bool valid = payload_valid(j.payload, g.payload_len, g.flags, g.count);
    if (!valid) return;
    const int step_x = g.vbw - g.vox, step_y = g.vbh - g.voy;
    const int area = max(g.vbw * g.vbh, 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < mc; i += gridDim.x * blockDim.x) {
        const int x = i % g.mw, y = i / g.mw;
        const int sx = min(x, g.gw - 1), sy = min(y, g.gh - 1);
        Dec a, b;
        decode_sets(g, j.payload, sy * g.gw + sx, a, b);
        for (int set = 0; set < 2; set++) {
            const Dec& v = set ? b : a;
            uint16_t ox, oy;
            if (abs(v.dx) > 1023 || abs(v.dy) > 1023) {
                ox = kNeutral;
                oy = kNeutral;
            } else {
                ox = centered(clamp_to_frame(v.dx, x * step_x, g.vbw, g.width));
                oy = centered(clamp_to_frame(v.dy, y * step_y, g.vbh, g.height));
            }
            j.mv[2 * set][i] = ox;
            j.mv[2 * set + 1][i] = oy;
        }
        if (g.mask_area) {
            int cx, cy;
            if (y < g.gh) { cx = min(x, g.gw - 1); cy = y; }
            else { cx = 0; cy = g.gh - 1; }
            Dec ma, mb;
            decode_sets(g, j.payload, cy * g.gw + cx, ma, mb);
            j.mag[0][i] = scale_magnitude(ma.score, area, g);
            j.mag[1][i] = scale_magnitude(mb.score, area, g);
        }
    }
}

// classify_scene's decision from the counters.
__global__ void prep_finish_kernel(const PrepJob* __restrict__ jobs, int n, const SmoothGeom* __restrict__ gp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int32_t* s = jobs[t].stats;
    if (!s[0]) { s[1] = 3; return; }
    const int required = s[2] * gp->lim_blocks / 100;
    const int high = s[3] + s[4], mid = high + s[5];
    s[1] = s[3] >= required ? 3 : (high >= required ? 2 : (mid >= required ? 1 : 0));
}

// ------------------------------------------------------------------ per-frame plan

// try_render / render_into_output decisions (core.rs:765-842, 1338-1383).
__global__ void plan_kernel(const FrameJob* __restrict__ jobs, FramePlan* __restrict__ plans, int n,
                            const SmoothGeom* __restrict__ gp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const SmoothGeom& g = *gp;
    FramePlan* p = plans + t;
    const int32_t* sn = jobs[t].stats_n;
    p->decision = p->raw_phase == 256 ? 1 : 0;
    p->algo = 0;
    p->mode23_ready = 0;
    if (p->early || !sn[0]) return;
    const int cls = sn[1];
    const int clamped = clampi(p->raw_phase, 0, 256);
    if (cls >= 3) {  // scene change: nearest source
        p->decision = clamped < 128 ? 0 : 1;
        return;
    }
    const int phase = p->phase_for[cls];
    if (phase == 0 || phase == 256) return;
    int algo = g.algo_req;
    if (cls >= 1 && cls < 3 && g.algo_req >= 11 && g.force13) algo = 13;
    if (algo == 2 && phase > 128) algo = 1;
    const int32_t* sp = jobs[t].stats_prev;
    const int32_t* sx = jobs[t].stats_next;
    const bool side_ready = g.algo_req == 23 && sp[0] && sx[0] && sx[1] < 3;
    p->mode23_ready = side_ready && algo == 23 && cls < 3;
    p->decision = 2;
    p->algo = algo;
    p->phase = phase;
    const int inv = 256 - phase;
    p->limit = phase > 126 ? f64_to_i32(256.0 - (double)inv * g.area_scale) : f64_to_i32((double)phase * g.area_scale);
}

// ------------------------------------------------------------------ coverage masks

__global__ void cov_clear_kernel(const FrameJob* __restrict__ jobs, const SmoothGeom* __restrict__ gp) {
    const FrameJob& j = jobs[blockIdx.y];
    const int len = (gp->mw + 2) * (gp->mh + 2);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        j.cov_acc[0][i] = 0;
        j.cov_acc[1][i] = 0;
    }
}

// TruncDiv (renderer.rs:3149-3184).  Its 40-bit reciprocal m = ceil(2^40/d) is used only
// for |v| < 2^40/d, where |v|*m/2^40 = |v|/d + e with 0 <= e < |v|/2^40 < 1/d, so the
// floor of that product is the floor of |v|/d; every other case is Rust's `/`.  So it
// is truncating division for every divisor other than 0 (which gives 0), and the
// device divides directly.
__device__ __forceinline__ int trunc_div(int v, int d) { return d == 0 ? 0 : v / d; }

// splat_coverage (renderer.rs:2945-3049): mask 0 splats set b at threshold 256-phase,
// mask 1 splats set a at threshold phase.  Integer atomics: the sums do not depend on order.
__global__ void cov_splat_kernel(const FrameJob* __restrict__ jobs, const FramePlan* __restrict__ plans,
                                 const SmoothGeom* __restrict__ gp) {
    const FrameJob& j = jobs[blockIdx.y];
    const FramePlan p = plans[blockIdx.y];
    const SmoothGeom& g = *gp;
    if (p.decision != 2 || p.algo < 21) return;
    const int x_step = g.vbw - g.vox, y_step = g.vbh - g.voy;
    const int denom = max(g.marker, 1) << 8;
    if (x_step == 0 || y_step == 0 || denom == 0 || g.mw <= 0 || g.mh <= 0) return;
    const int x_bias = 1 - x_step, y_bias = 1 - y_step;
    const int stride = g.mw + 2;
    const int total = 2 * g.count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int m = t / g.count, i = t % g.count;
        const int x = i % g.gw, y = i / g.gw;
        const int16_t* vec = m == 0 ? j.vec_b : j.vec_a;
        const int threshold = m == 0 ? 256 - p.phase : p.phase;
        int32_t* acc = m == 0 ? j.cov_acc[0] : j.cov_acc[1];  // (an index by lane would use scratch)
        const int dx = trunc_div(threshold * (int)vec[2 * i], denom), dy = trunc_div(threshold * (int)vec[2 * i + 1], denom);
        const int sx = x * x_step + dx, sy = y_step * y + dy;
        const int left = trunc_div(sx + (x_bias & (sx >> 31)), x_step);
        const int top = trunc_div(sy + (y_bias & (sy >> 31)), y_step);
        const int right = left + 1, bottom = top + 1;
        const int next_x = right * x_step;
        const int lw = next_x - sx, tw = y_step * bottom - sy;
        const int rw = sx + g.vbw - next_x, bw = g.vbh - tw;
        auto add = [&](int px, int py, int w) {
            if (px < 0 || py < 0) return;
            atomicAdd(&acc[py * stride + px], w);
        };
        if (left >= 0 && right < g.mw && top >= 0 && bottom < g.mh) {
            add(left, top, lw * tw);
            add(left + 1, top, tw * rw);
            add(left, top + 1, lw * bw);
            add(left + 1, top + 1, rw * bw);
        } else {
            if (top >= 0 && top < g.mh && left >= 0 && left < g.mw) add(left, top, lw * tw);
            if (left >= -1 && top >= 0 && top < g.mh && right < g.mw) add(right, top, tw * rw);
            if (top >= -1 && bottom < g.mh && left >= -1 && right < g.mw) add(right, bottom, rw * bw);
            if (left >= 0 && top >= -1 && bottom < g.mh && left < g.mw) add(left, bottom, lw * bw);
        }
    }
}

// window_3x3 + finish_coverage (renderer.rs:3051-3130) and the max of the magnitude masks.
__global__ void cov_finish_kernel(const FrameJob* __restrict__ jobs, const FramePlan* __restrict__ plans,
                                 const SmoothGeom* __restrict__ gp) {
    const FrameJob& j = jobs[blockIdx.y];
    const FramePlan p = plans[blockIdx.y];
    const SmoothGeom& g = *gp;
    if (p.decision != 2) return;
    const int mc = g.mw * g.mh;
    const int stride = g.mw + 2;
    const int area = max(g.vbw * g.vbh, 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < mc; i += gridDim.x * blockDim.x) {
        const int x = i % g.mw, y = i / g.mw;
        if (g.mask_area) j.maxmag[i] = max(j.mag[0][i], j.mag[1][i]);
        if (p.algo < 21) continue;
        for (int m = 0; m < 2; m++) {
            const int32_t* acc = j.cov_acc[m];
            int s = 0;  // 3x3 window centred on (x, y), clipped at the top and left edges
            for (int yy = max(y - 1, 0); yy <= y + 1; yy++)
                for (int xx = max(x - 1, 0); xx <= x + 1; xx++) s += acc[yy * stride + xx];
            const int covered = s >> 3;
            const int remaining = area <= covered ? 0 : area - covered;
            j.cov[m][i] = g.cover_lut[remaining];
        }
    }
}

// ------------------------------------------------------------------ render

struct PP {  // PlaneParams (renderer.rs:1876-1892)
    int width, height, bw, bh, ox, oy, xs, ys, xdiv, ydiv, step, max_x, max_y;
    const int (*xw)[2];
    const int (*yw)[2];
};

__device__ __forceinline__ int interp_pair(int a, int b, int total, int shift, const int w[2]) {
    if (shift < 0) return (w[0] * a + w[1] * b) / total;
    return (w[0] * a + w[1] * b) >> shift;
}

// v / 2 (truncating, as C's division) when `half`, else v: the chroma planes halve the
// luma motion.  A runtime divisor of 1 or 2 would compile to a full signed division.
__device__ __forceinline__ int div12(int v, bool half) {
    return half ? (v + (int)((uint32_t)v >> 31)) >> 1 : v;
}

// Reads of the frame and motion planes as global memory (the descriptor's pointers are
// generic, which compiles to flat loads).
typedef __attribute__((address_space(1))) const uint8_t gm_u8;
typedef __attribute__((address_space(1))) const uint16_t gm_u16;
typedef __attribute__((address_space(1))) const uint16_t gm_u16;

__device__ __forceinline__ int mval(const uint16_t* plane, int idx, int count, int t) {  // motion_value + LUT
    const int li = min(idx < count ? (int)((gm_u16*)plane)[idx] : (int)kNeutral, 2047);
    return trunc_div_256((li - 1024) * t);
}

__device__ __forceinline__ int alpha_abs(const uint8_t* m, int count, const PP& p, int stride, int x, int y) {
    int ix0 = 0, ix1 = 0, iy0 = 0, iy1 = 0;  // alpha_at_absolute, renderer.rs:2696-2722
    if (p.ox <= x) { ix0 = (x - p.ox) / p.bw; ix1 = ix0 + 1; }
    if (p.oy <= y) { iy0 = (y - p.oy) / p.bh; iy1 = iy0 + 1; }
    const int fx = p.ox < x ? (x - p.ox) % p.bw : 0, fy = p.oy < y ? (y - p.oy) % p.bh : 0;
    auto at = [&](int xx, int yy) { const int i = yy * stride + xx; return i < count ? (int)((gm_u8*)m)[i] : 0; };
    const int top = ((p.bw - fx) * at(ix0, iy0) + fx * at(ix1, iy0)) / p.bw;
    const int bottom = ((p.bw - fx) * at(ix0, iy1) + fx * at(ix1, iy1)) / p.bw;
    return byte_of(((p.bh - fy) * top + fy * bottom) / p.bh);
}

struct Src {
    const uint8_t* d;
    int stride, span, shift;
};

// A plane (or the pel^2 subplanes of a super plane) is below 2^32 bytes and its rows and
// strides below 2^16, so the offset is 32-bit arithmetic from the plane's base.
__device__ __forceinline__ int sample(const Src& s, const PP& p, int sx, int sy, int dx, int dy) {
    const int x = clampi(sx + dx, 0, p.max_x), y = clampi(sy + dy, 0, p.max_y);
    gm_u8* d = (gm_u8*)s.d;
    if (s.shift == 0) return d[__umul24((uint32_t)y, (uint32_t)s.stride) + (uint32_t)x];
    const int mask = (1 << s.shift) - 1;
    const uint32_t sub = (uint32_t)(((y & mask) << s.shift) | (x & mask));
    return d[sub * (uint32_t)s.span + __umul24((uint32_t)(y >> s.shift), (uint32_t)s.stride) + (uint32_t)(x >> s.shift)];
}

// One thread per (plane row, tile) of one frame: the pixels of a tile row share their
// motion samples (the 4 grid corners of every vector set), the row weights and the
// vertical interpolations, so those are computed once; each pixel then takes its column
// weight, its source taps and the blends.  Tiles (render_tiles, renderer.rs:1925-2010)
// are drawn in order -1 (the origin strip), 0, 1, ...; the first tile reaching the plane
// end takes the rest of the plane, and a later tile overwrites an earlier one, so a
// pixel at or past the origin belongs to tile min((x - origin) / block, last).
// Thread x-index k covers tile k - 1; a frame copied whole (decision != 2) is copied in
// runs of 8 columns per thread instead.
__global__ void __launch_bounds__(256) render_kernel(const FrameJob* __restrict__ jobs, const FramePlan* __restrict__ plans,
                                                     const SmoothGeom* __restrict__ gp) {
    const int plane = blockIdx.z % 3;
    const FrameJob& j = jobs[blockIdx.z / 3];
    const SmoothGeom& g = *gp;
    const bool chroma = plane != 0;
    const int pw = chroma ? g.width / 2 : g.width, ph = chroma ? g.height / 2 : g.height;
    const int k = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
    if (py >= ph) return;
    const FramePlan plan = plans[blockIdx.z / 3];
    uint8_t* orow = j.dst[plane] + (size_t)py * j.dstride[plane];
    if (plan.decision != 2) {  // copy of source n or n+1 (padded_output with no padding)
        const uint8_t* srow = (plan.decision == 1 ? j.src1[plane] : j.src0[plane]) + (size_t)py * j.sstride[plane];
        for (int x = 8 * k; x < min(8 * k + 8, pw); x++) orow[x] = ((gm_u8*)srow)[x];
        return;
    }
    const bool interp = g.interp != 0;
    const bool zero = !interp;
    PP p;
    p.width = pw; p.height = ph;
    p.bw = chroma ? g.cbw / 2 : g.cbw;
    p.bh = chroma ? g.cbh / 2 : g.cbh;
    p.ox = zero ? 0 : (chroma ? g.cox / 2 : g.cox);
    p.oy = zero ? 0 : (chroma ? g.coy / 2 : g.coy);
    p.xs = g.xshift[chroma]; p.ys = g.yshift[chroma];
    p.xdiv = chroma ? 2 : 1; p.ydiv = chroma ? 2 : 1;
    p.step = g.source_step;
    p.max_x = max(pw * g.source_step - 1, 0);
    p.max_y = max(ph * g.source_step - 1, 0);
    p.xw = g.xw[chroma]; p.yw = g.yw[chroma];
    PP mp = p;  // mask params with the regular origin (zero-origin renders)
    mp.ox = chroma ? g.cox / 2 : g.cox;
    mp.oy = chroma ? g.coy / 2 : g.coy;
    if (p.bw <= 0 || p.bh <= 0) return;
    const int mw = g.mw, mh = g.mh;
    // this thread's columns [xb, xe) and tile gx with base bx
    const int gx = k - 1;
    int xb, xe, bx;
    if (gx < 0) {
        xb = 0; xe = min(p.ox, pw); bx = 0;
    } else {
        const int last = g.tile_last_x[chroma];  // the first tile reaching the plane end (-1: none)
        if (gx > last) return;
        bx = p.ox + p.bw * gx;
        xb = bx;
        xe = gx == last ? pw : min(bx + p.bw, pw);
    }
    if (xb >= xe) return;
    int gy, by;
    if (py < p.oy) {
        gy = -1; by = 0;
    } else {  // (py - oy) / bh as a multiply-high: exact for operands below 2^16
        const uint32_t q = g.rcp_bh[chroma] ? __umulhi((uint32_t)(py - p.oy), g.rcp_bh[chroma]) : (uint32_t)(py - p.oy);
        gy = min((int)q, g.tile_last_y[chroma]);
        by = p.oy + p.bh * gy;
    }
    const int x0 = max(gx, 0), y0 = max(gy, 0);
    const int nx = (gx >= 0 && gx != mw - 1) ? 1 : 0, ny = (gy >= 0 && gy != mh - 1) ? mw : 0;
    const int i00 = y0 * mw + x0, i01 = i00 + nx, i10 = i00 + ny, i11 = i00 + ny + nx;
    const int* wy = p.yw[min(py - by, 31)];
    const int mc = mw * mh;
    const int t = plan.phase, inv = 256 - plan.phase;
    // a vector set at this tile row (motion_samples + the vertical interpolation);
    // its value at a pixel is mv_at (interpolate_* with the column weight)
    struct MV {
        int rx0, rx1, ry0, ry1;
    };
    auto mv_row = [&](const uint16_t* mx, const uint16_t* my, int tt) -> MV {
        const bool hx = p.xdiv == 2, hy = p.ydiv == 2;
        const int a0 = div12(mval(mx, i00, mc, tt), hx), b0 = div12(mval(my, i00, mc, tt), hy);
        if (!interp) return MV{a0, a0, b0, b0};
        const int a1 = div12(mval(mx, i01, mc, tt), hx), a2 = div12(mval(mx, i10, mc, tt), hx);
        const int a3 = div12(mval(mx, i11, mc, tt), hx);
        const int b1 = div12(mval(my, i01, mc, tt), hy), b2 = div12(mval(my, i10, mc, tt), hy);
        const int b3 = div12(mval(my, i11, mc, tt), hy);
        return MV{interp_pair(a0, a2, p.bh, p.ys, wy), interp_pair(a1, a3, p.bh, p.ys, wy),
                  interp_pair(b0, b2, p.bh, p.ys, wy), interp_pair(b1, b3, p.bh, p.ys, wy)};
    };
    auto mv_at = [&](const MV& m, const int* wx, int& dx, int& dy) {
        if (!interp) { dx = m.rx0; dy = m.ry0; return; }
        dx = interp_pair(m.rx0, m.rx1, p.bw, p.xs, wx);
        dy = interp_pair(m.ry0, m.ry1, p.bw, p.xs, wx);
    };
    // a mask at this tile row (alpha: interpolated, or alpha_at_absolute for zero-origin)
    struct AR {
        int r0, r1;
    };
    auto alpha_row = [&](const uint8_t* m) -> AR {
        if (zero) return AR{0, 0};
        auto at = [&](int i) { return i < mc ? (int)((gm_u8*)m)[i] : 0; };
        return AR{interp_pair(at(i00), at(i10), p.bh, p.ys, wy), interp_pair(at(i01), at(i11), p.bh, p.ys, wy)};
    };
    auto alpha_at = [&](const uint8_t* m, const AR& r, const int* wx, int px) -> int {
        if (zero) return alpha_abs(m, mc, mp, mw, px, py);
        return byte_of(interp_pair(r.r0, r.r1, p.bw, p.xs, wx));
    };
    Src s0, s1;
    if (g.use_super) {
        s0 = Src{j.sup0[plane], j.supstride[plane], j.supstride[plane] * ph, g.super_shift};
        s1 = Src{j.sup1[plane], j.supstride[plane], j.supstride[plane] * ph, g.super_shift};
    } else {
        s0 = Src{j.src0[plane], j.sstride[plane], 0, 0};
        s1 = Src{j.src1[plane], j.sstride[plane], 0, 0};
    }
    const int sy = py * p.step;
    const int algo = plan.algo;
    if (algo == 1 || algo == 2) {  // render_mode1_or_2 -> render_selected_warp (renderer.rs:308-327, 881-982)
        const bool m1 = algo == 1;
        const Src& s = m1 ? s1 : s0;
        const MV m = m1 ? mv_row(j.mv_n[2], j.mv_n[3], inv) : mv_row(j.mv_n[0], j.mv_n[1], t);
        const uint8_t* mk = m1 ? j.mag[0] : j.mag[1];
        const AR ar = g.mask_area ? alpha_row(mk) : AR{0, 0};
        for (int px = xb; px < xe; px++) {
            const int* wx = p.xw[min(px - bx, 31)];
            const int sx = px * p.step;
            int dx, dy;
            mv_at(m, wx, dx, dy);
            const int warped = sample(s, p, sx, sy, dx, dy);
            int v = warped;
            if (g.mask_area) v = blend255(warped, sample(s, p, sx, sy, 0, 0), alpha_at(mk, ar, wx, px));
            orow[px] = (uint8_t)v;
        }
    } else if (algo == 11 || algo == 13) {  // render_mode11_or_13 (renderer.rs:330-437, 2918-2942)
        const MV m0 = mv_row(j.mv_n[0], j.mv_n[1], t), m1 = mv_row(j.mv_n[2], j.mv_n[3], inv);
        const AR am = g.mask_area ? alpha_row(j.maxmag) : AR{0, 0};
        for (int px = xb; px < xe; px++) {
            const int* wx = p.xw[min(px - bx, 31)];
            const int sx = px * p.step;
            int dx0, dy0, dx1, dy1;
            mv_at(m0, wx, dx0, dy0);
            mv_at(m1, wx, dx1, dy1);
            const int c0 = sample(s0, p, sx, sy, 0, 0), c1 = sample(s1, p, sx, sy, 0, 0);
            const int v0 = sample(s0, p, sx, sy, dx0, dy0), v1 = sample(s1, p, sx, sy, dx1, dy1);
            const int mo = algo == 13 ? clamp_between(blend256(c0, c1, t), v0, v1) : blend256(v0, v1, t);
            orow[px] = (uint8_t)(g.mask_area ? blend255(mo, blend256(c0, c1, t), alpha_at(j.maxmag, am, wx, px)) : mo);
        }
    } else if ((algo == 21 || algo == 22) || (algo == 23 && !plan.mode23_ready)) {  // renderer.rs:617-641, 1830-1848
        const bool m21 = algo != 22;
        const MV m0 = mv_row(j.mv_n[0], j.mv_n[1], t), m1 = mv_row(j.mv_n[2], j.mv_n[3], inv);
        const AR a0r = alpha_row(j.cov[0]), a1r = alpha_row(j.cov[1]);
        const AR am = g.mask_area ? alpha_row(j.maxmag) : AR{0, 0};
        for (int px = xb; px < xe; px++) {
            const int* wx = p.xw[min(px - bx, 31)];
            const int sx = px * p.step;
            int dx0, dy0, dx1, dy1;
            mv_at(m0, wx, dx0, dy0);
            mv_at(m1, wx, dx1, dy1);
            const int c0 = sample(s0, p, sx, sy, 0, 0), c1 = sample(s1, p, sx, sy, 0, 0);
            const int v0 = sample(s0, p, sx, sy, dx0, dy0), v1 = sample(s1, p, sx, sy, dx1, dy1);
            const int a0 = alpha_at(j.cov[0], a0r, wx, px), a1 = alpha_at(j.cov[1], a1r, wx, px);
            const int a = blend255(v0, v1, a0), b = blend255(v1, v0, a1);
            const int mo = m21 ? blend256(a, b, t) : clamp_between(blend256(c0, c1, t), a, b);
            orow[px] = (uint8_t)(g.mask_area ? blend255(mo, blend256(c0, c1, plan.limit), alpha_at(j.maxmag, am, wx, px))
                                             : mo);
        }
    } else {  // algo 23 (renderer.rs:1294-1537, 1850-1868)
        const MV m0 = mv_row(j.mv_n[0], j.mv_n[1], t), m1 = mv_row(j.mv_n[2], j.mv_n[3], inv);
        const MV m2 = mv_row(j.mv2[0], j.mv2[1], t), m3 = mv_row(j.mv3[0], j.mv3[1], inv);
        const AR a0r = alpha_row(j.cov[0]), a1r = alpha_row(j.cov[1]);
        const AR am = g.mask_area ? alpha_row(j.maxmag) : AR{0, 0};
        for (int px = xb; px < xe; px++) {
            const int* wx = p.xw[min(px - bx, 31)];
            const int sx = px * p.step;
            int d[4][2];
            mv_at(m0, wx, d[0][0], d[0][1]);
            mv_at(m1, wx, d[1][0], d[1][1]);
            mv_at(m2, wx, d[2][0], d[2][1]);
            mv_at(m3, wx, d[3][0], d[3][1]);
            const int A = sample(s0, p, sx, sy, d[0][0], d[0][1]), B = sample(s1, p, sx, sy, d[1][0], d[1][1]);
            const int Cc = sample(s0, p, sx, sy, d[2][0], d[2][1]), D = sample(s1, p, sx, sy, d[3][0], d[3][1]);
            const int t0 = alpha_at(j.cov[0], a0r, wx, px), t1 = alpha_at(j.cov[1], a1r, wx, px);
            const int i0 = blend255(A, clamp_between(D, A, B), t0);
            const int i1 = blend255(B, clamp_between(Cc, A, B), t1);
            int mo = blend256(i0, i1, t);
            if (g.mask_area) {
                const int ba = sample(s0, p, sx, sy, 0, 0), bb = sample(s1, p, sx, sy, 0, 0);
                mo = blend255(mo, blend256(ba, bb, t), alpha_at(j.maxmag, am, wx, px));
            }
            orow[px] = (uint8_t)mo;
        }
    }
}

}  // namespace

// ------------------------------------------------------------------ float renderer
// The reference's float GPU renderer (crates/svpflow2/src/gpu.rs:1039-1168, fed as
// gpu.rs:754-957 and core.rs:3235-3330 feed it): vectors sampled with a B-spline cubic
// made of four bilinear fetches, source taps bilinear with clamp-to-edge, luma blended in
// linear light.  Image filtering is computed in fp32 with the weights the OpenCL
// specification writes (its precision is the device's choice there), division is IEEE.
// One thread per pixel of one plane of one frame.

struct F4 {
    float x, y, z, w;
};
__device__ __forceinline__ float fmix(float x, float y, float a) { return x + (y - x) * a; }
__device__ __forceinline__ F4 fmix4(const F4& x, const F4& y, float a) {
    return F4{fmix(x.x, y.x, a), fmix(x.y, y.y, a), fmix(x.z, y.z, a), fmix(x.w, y.w, a)};
}
__device__ __forceinline__ float median3f(float a, float b, float c) {
    const float lo = fminf(a, b);
    return fmaxf(lo, fminf(a + b - lo, c));
}

// read_imagef, CLK_FILTER_LINEAR | CLK_ADDRESS_CLAMP_TO_EDGE, unnormalized (u, v)
template <class Tex>
__device__ __forceinline__ float fbilinear(const Tex& tex, int w, int h, float u, float v) {
    const float x = u - 0.5f, y = v - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    const int xa = clampi(i0, 0, w - 1), xb = clampi(i0 + 1, 0, w - 1);
    const int ya = clampi(j0, 0, h - 1), yb = clampi(j0 + 1, 0, h - 1);
    const float one_a = 1.0f - a, one_b = 1.0f - b;
    return one_a * one_b * tex(xa, ya) + a * one_b * tex(xb, ya) + one_a * b * tex(xa, yb) + a * b * tex(xb, yb);
}

__device__ __forceinline__ void cubic_axis(float p, float& g0, float& h0, float& h1) {
    const float grid = p - 0.5f;
    const float index = floorf(grid);
    const float f = grid - index, r = 1.0f - f;
    const float w0 = r * r * r / 6.0f;
    const float w1 = 2.0f / 3.0f - 0.5f * f * f * (2.0f - f);
    const float w2 = 2.0f / 3.0f - 0.5f * r * r * (2.0f - r);
    const float w3 = f * f * f / 6.0f;
    g0 = w0 + w1;
    const float g1 = w2 + w3;
    h0 = w1 / g0 - 0.5f + index;
    h1 = w3 / g1 + 1.5f + index;
}

// cubic_sample of a 4-channel image given as four planes of the motion grid.
template <class T, class Conv>
__device__ __forceinline__ F4 fcubic(const T* c0, const T* c1, const T* c2, const T* c3, int w, int h, float px,
                                     float py, const Conv& conv) {
    float gx, h0x, h1x, gy, h0y, h1y;
    cubic_axis(px, gx, h0x, h1x);
    cubic_axis(py, gy, h0y, h1y);
    auto fetch = [&](float u, float v) {
        return F4{fbilinear([&](int x, int y) { return conv(c0[y * w + x]); }, w, h, u, v),
                  fbilinear([&](int x, int y) { return conv(c1[y * w + x]); }, w, h, u, v),
                  fbilinear([&](int x, int y) { return conv(c2[y * w + x]); }, w, h, u, v),
                  fbilinear([&](int x, int y) { return conv(c3[y * w + x]); }, w, h, u, v)};
    };
    const F4 a = fetch(h0x, h0y), b = fetch(h1x, h0y), c = fetch(h0x, h1y), d = fetch(h1x, h1y);
    return fmix4(fmix4(d, b, gy), fmix4(c, a, gy), gx);
}

__global__ void __launch_bounds__(256) render_float_kernel(const FrameJob* __restrict__ jobs,
                                                           const FramePlan* __restrict__ plans,
                                                           const SmoothGeom* __restrict__ gp) {
    const int plane = blockIdx.z % 3;
    const FrameJob& j = jobs[blockIdx.z / 3];
    const SmoothGeom& g = *gp;
    const bool chroma = plane != 0;
    const int W = chroma ? g.width / 2 : g.width, H = chroma ? g.height / 2 : g.height;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const FramePlan plan = plans[blockIdx.z / 3];
    uint8_t* out = j.dst[plane] + (size_t)y * j.dstride[plane] + x;
    if (plan.decision != 2) {  // copies, as the integer path
        const uint8_t* s = plan.decision == 1 ? j.src1[plane] : j.src0[plane];
        *out = s[(size_t)y * j.sstride[plane] + x];
        return;
    }
    const int mw = g.mw, mh = g.mh;
    const int algo = (plan.algo == 23 && !plan.mode23_ready) ? 21 : plan.algo;
    const bool has_sad = g.mask_area != 0;
    const int xr = chroma ? 2 : 1, yr = chroma ? 2 : 1;
    const gm_u8* s0 = (const gm_u8*)j.src0[plane];
    const gm_u8* s1 = (const gm_u8*)j.src1[plane];
    const int ss = j.sstride[plane];
    auto tex0 = [&](int tx, int ty) {
        const uint8_t t = s0[ty * ss + tx];
        return chroma ? (float)t / 255.0f : g.lin_lut[t];
    };
    auto tex1 = [&](int tx, int ty) {
        const uint8_t t = s1[ty * ss + tx];
        return chroma ? (float)t / 255.0f : g.lin_lut[t];
    };
    const float maxx = (float)(W - 1), maxy = (float)(H - 1);
    const float denx = (float)xr * 256.0f, deny = (float)yr * 256.0f;  // pel = source_step = 1
    auto source = [&](bool second, float vx, float vy, int time) {
        float dx = vx * 65535.0f - 1024.0f, dy = vy * 65535.0f - 1024.0f;
        dx = dx * (float)time / denx;
        dy = dy * (float)time / deny;
        const float px = fminf(fmaxf((float)x + dx, 0.0f), maxx);
        const float py = fminf(fmaxf((float)y + dy, 0.0f), maxy);
        return 255.0f * (second ? fbilinear(tex1, W, H, px + 0.5f, py + 0.5f)
                                : fbilinear(tex0, W, H, px + 0.5f, py + 0.5f));
    };
    auto u16n = [](uint16_t v) { return (float)v / 65535.0f; };
    auto u8n = [](uint8_t v) { return (float)v / 255.0f; };
    const float vpx = (float)(x * xr - g.cox) / (float)g.cbw;
    const float vpy = (float)(y * yr - g.coy) / (float)g.cbh;
    const gm_u16* m0x = (const gm_u16*)j.mv_n[0];
    const gm_u16* m0y = (const gm_u16*)j.mv_n[1];
    const gm_u16* m1x = (const gm_u16*)j.mv_n[2];
    const gm_u16* m1y = (const gm_u16*)j.mv_n[3];
    const F4 v = fcubic(m0x, m1y, m1x, m0y, mw, mh, vpx, vpy, u16n);
    const float ref_f = source(false, v.x, v.w, plan.phase);
    const float ref_b = source(true, v.z, v.y, 256 - plan.phase);
    const float ref_f0 = 255.0f * fbilinear(tex0, W, H, (float)x + 0.5f, (float)y + 0.5f);
    const float ref_b0 = 255.0f * fbilinear(tex1, W, H, (float)x + 0.5f, (float)y + 0.5f);
    const float time = (float)plan.phase / 256.0f;
    F4 mask{0.0f, 0.0f, 0.0f, 0.0f};
    if (algo >= 21 || has_sad) {
        // (area b, coverage a, coverage b, area a); a plane the frame lacks reads as 0
        const gm_u8* c0 = (const gm_u8*)(has_sad ? j.mag[1] : j.cov[0]);
        const gm_u8* c1 = (const gm_u8*)(algo >= 21 ? j.cov[0] : j.mag[0]);
        const gm_u8* c2 = (const gm_u8*)(algo >= 21 ? j.cov[1] : j.mag[0]);
        const gm_u8* c3 = (const gm_u8*)(has_sad ? j.mag[0] : j.cov[0]);
        mask = fcubic(c0, c1, c2, c3, mw, mh, vpx, vpy, u8n);
        if (!has_sad) { mask.x = 0.0f; mask.w = 0.0f; }
        if (algo < 21) { mask.y = 0.0f; mask.z = 0.0f; }
    }
    float result;
    if (algo == 1) {
        result = ref_b;
    } else if (algo == 2) {
        result = ref_f;
    } else if (algo == 11) {
        result = fmix(ref_f, ref_b, time);
    } else if (algo == 13) {
        result = median3f(ref_f, ref_b, fmix(ref_f0, ref_b0, time));
    } else if (algo == 21) {
        result = fmix(fmix(ref_f, ref_b, mask.y), fmix(ref_b, ref_f, mask.z), time);
    } else if (algo == 22) {
        result = median3f(fmix(ref_f, ref_b, mask.y), fmix(ref_b, ref_f, mask.z), fmix(ref_f0, ref_b0, time));
    } else {
        const F4 e = fcubic((const gm_u16*)j.mv2[0], (const gm_u16*)j.mv3[1], (const gm_u16*)j.mv3[0],
                            (const gm_u16*)j.mv2[1], mw, mh, vpx, vpy, u16n);
        const float ref_ff = source(false, e.x, e.w, plan.phase);
        const float ref_bb = source(true, e.z, e.y, 256 - plan.phase);
        result = fmix(fmix(ref_f, median3f(ref_b, ref_bb, ref_f), mask.y),
                      fmix(ref_b, median3f(ref_b, ref_ff, ref_f), mask.z), time);
    }
    if (has_sad) {
        if (algo == 1) result = fmix(result, ref_b0, mask.x);
        else if (algo == 2) result = fmix(result, ref_f0, mask.w);
        else result = fmix(result, fmix(ref_f0, ref_b0, time), fmix(mask.x, mask.w, time));
    }
    if (!chroma) {  // linear light back to the gamma curve
        float val = result / 255.0f;
        val = val < 0.018f ? val * 4.5f : 1.099f * powf(val, 0.45f) - 0.099f;
        result = val * 255.0f;
    }
    *out = (uint8_t)fminf(fmaxf(roundf(result), 0.0f), 255.0f);
}

hipError_t launch_smooth_prep(const PrepJob* d_jobs, int njobs, const SmoothGeom* d_geom, int count, int mc,
                              hipStream_t s) {
    dim3 g1((count + 255) / 256, njobs), g2((mc + 255) / 256, njobs);
    hipLaunchKernelGGL(prep_clear_kernel, dim3((njobs + 63) / 64), dim3(64), 0, s, d_jobs, njobs);
    hipLaunchKernelGGL(prep_vectors_kernel, g1, dim3(256), 0, s, d_jobs, d_geom);
    hipLaunchKernelGGL(prep_planes_kernel, g2, dim3(256), 0, s, d_jobs, d_geom);
    hipLaunchKernelGGL(prep_finish_kernel, dim3((njobs + 63) / 64), dim3(64), 0, s, d_jobs, njobs, d_geom);
    return hipGetLastError();
}

hipError_t launch_smooth_masks(const FrameJob* d_jobs, FramePlan* d_plans, int njobs, const SmoothGeom* d_geom, int count,
                               int mw, int mh, bool coverage, hipStream_t s) {
    hipLaunchKernelGGL(plan_kernel, dim3((njobs + 63) / 64), dim3(64), 0, s, d_jobs, d_plans, njobs, d_geom);
    const int acc = (mw + 2) * (mh + 2), mc = mw * mh;
    if (coverage) {
        hipLaunchKernelGGL(cov_clear_kernel, dim3((acc + 255) / 256, njobs), dim3(256), 0, s, d_jobs, d_geom);
        hipLaunchKernelGGL(cov_splat_kernel, dim3((2 * count + 255) / 256, njobs), dim3(256), 0, s, d_jobs, d_plans, d_geom);
    }
    hipLaunchKernelGGL(cov_finish_kernel, dim3((mc + 255) / 256, njobs), dim3(256), 0, s, d_jobs, d_plans, d_geom);
    return hipGetLastError();
}

hipError_t launch_smooth_render(const FrameJob* d_jobs, const FramePlan* d_plans, int njobs, const SmoothGeom* d_geom,
                                int width, int height, int mw, hipStream_t s) {
    // threads per row: the origin strip and the tiles (at most mw of them), or the runs
    // of 8 columns of a copied frame
    const int per_row = std::max(mw + 1, (width + 7) / 8);
    dim3 grid((per_row + 255) / 256, height, 3 * njobs);
    hipLaunchKernelGGL(render_kernel, grid, dim3(256), 0, s, d_jobs, d_plans, d_geom);
    return hipGetLastError();
}

hipError_t launch_smooth_render_float(const FrameJob* d_jobs, const FramePlan* d_plans, int njobs,
                                      const SmoothGeom* d_geom, int width, int height, hipStream_t s) {
    dim3 grid((width + 255) / 256, height, 3 * njobs);
    hipLaunchKernelGGL(render_float_kernel, grid, dim3(256), 0, s, d_jobs, d_plans, d_geom);
    return hipGetLastError();
}

}  // namespace svpf
