// Oracle restatement of the reference's float GPU renderer: the OpenCL kernels
// `linear_luma` and `render_frame` of crates/svpflow2/src/gpu.rs:1039-1168, fed as
// GpuContext::render_frame (gpu.rs:754-957) and gpu_render_frame (core.rs:3235-3330) feed
// them.  Test infrastructure only (see oracle.h).
//
// OpenCL leaves the precision of image filtering and of native_divide / native_powr to
// the device.  This restatement fixes them: read_imagef's bilinear weights in fp32 as
// the OpenCL specification writes them, IEEE division and libm powf, every expression
// in the kernel's operation order without contraction (-ffp-contract=off).
#include <algorithm>
#include <cmath>

#include "render.h"

namespace oracle {

namespace {

struct F4 {
    float x, y, z, w;
};

inline float mixf(float x, float y, float a) { return x + (y - x) * a; }  // OpenCL mix
inline F4 mix4(const F4& x, const F4& y, float a) {
    return F4{mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a), mixf(x.w, y.w, a)};
}
inline float median3(float a, float b, float c) {  // gpu.rs median3
    const float lo = std::fmin(a, b);
    return std::fmax(lo, std::fmin(a + b - lo, c));
}
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// read_imagef with CLK_FILTER_LINEAR | CLK_ADDRESS_CLAMP_TO_EDGE at unnormalized (u, v):
// texels floor(u - 0.5) and its right / lower neighbours, weights the fractions.
template <class Tex>
float bilinear(const Tex& tex, int w, int h, float u, float v) {
    const float x = u - 0.5f, y = v - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    const int xa = clampi(i0, 0, w - 1), xb = clampi(i0 + 1, 0, w - 1);
    const int ya = clampi(j0, 0, h - 1), yb = clampi(j0 + 1, 0, h - 1);
    const float one_a = 1.0f - a, one_b = 1.0f - b;
    return one_a * one_b * tex(xa, ya) + a * one_b * tex(xb, ya) + one_a * b * tex(xa, yb) + a * b * tex(xb, yb);
}

// cubic_sample (gpu.rs): B-spline weights folded into four bilinear fetches.
template <class Tex4>
F4 cubic(const Tex4& tex, int w, int h, float px, float py) {
    auto axis = [](float p, float* g0, float* h0, float* h1) {
        const float grid = p - 0.5f;
        const float index = std::floor(grid);
        const float f = grid - index, r = 1.0f - f;
        const float w0 = r * r * r / 6.0f;
        const float w1 = 2.0f / 3.0f - 0.5f * f * f * (2.0f - f);
        const float w2 = 2.0f / 3.0f - 0.5f * r * r * (2.0f - r);
        const float w3 = f * f * f / 6.0f;
        *g0 = w0 + w1;
        const float g1 = w2 + w3;
        *h0 = w1 / *g0 - 0.5f + index;
        *h1 = w3 / g1 + 1.5f + index;
    };
    float gx, h0x, h1x, gy, h0y, h1y;
    axis(px, &gx, &h0x, &h1x);
    axis(py, &gy, &h0y, &h1y);
    auto fetch = [&](float u, float v) {
        return F4{bilinear([&](int x, int y) { return tex(x, y).x; }, w, h, u, v),
                  bilinear([&](int x, int y) { return tex(x, y).y; }, w, h, u, v),
                  bilinear([&](int x, int y) { return tex(x, y).z; }, w, h, u, v),
                  bilinear([&](int x, int y) { return tex(x, y).w; }, w, h, u, v)};
    };
    const F4 a = fetch(h0x, h0y), b = fetch(h1x, h0y), c = fetch(h0x, h1y), d = fetch(h1x, h1y);
    return mix4(mix4(d, b, gy), mix4(c, a, gy), gx);
}

}  // namespace

// linear_luma (gpu.rs): the float image of a luma plane, one value per 8-bit sample.
void float_linear_lut(float lut[256]) {
    for (int t = 0; t < 256; t++) {
        const float v = (float)t / 255.0f;
        lut[t] = v < 0.081f ? v / 4.5f : std::pow((v + 0.099f) / 1.099f, 1.0f / 0.45f);
    }
}

void float_render(const FloatFrame& f, uint8_t* dst[3]) {
    float lut[256];
    float_linear_lut(lut);
    const int mw = f.mw, mh = f.mh;
    auto u16 = [](uint16_t v) { return (float)v / 65535.0f; };   // CL_UNORM_INT16
    auto u8n = [](uint8_t v) { return (float)v / 255.0f; };      // CL_UNORM_INT8
    // the vectors image (gpu.rs:793-806): (motion0.x, motion1.y, motion1.x, motion0.y)
    auto base_tex = [&](int x, int y) {
        const size_t i = (size_t)y * mw + x;
        return F4{u16(f.m[0][i]), u16(f.m[3][i]), u16(f.m[2][i]), u16(f.m[1][i])};
    };
    // vectors_ext (:812-823): (motion2.x, motion3.y, motion3.x, motion2.y)
    auto ext_tex = [&](int x, int y) {
        const size_t i = (size_t)y * mw + x;
        return F4{u16(f.m[4][i]), u16(f.m[7][i]), u16(f.m[6][i]), u16(f.m[5][i])};
    };
    // masks (:848-858): (area b, coverage a, coverage b, area a)
    auto mask_tex = [&](int x, int y) {
        const size_t i = (size_t)y * mw + x;
        return F4{f.has_sad ? u8n(f.mag1[i]) : 0.0f, f.algo >= 21 ? u8n(f.cov0[i]) : 0.0f,
                  f.algo >= 21 ? u8n(f.cov1[i]) : 0.0f, f.has_sad ? u8n(f.mag0[i]) : 0.0f};
    };
    for (int plane = 0; plane < 3; plane++) {
        const bool chroma = plane != 0;
        const int W = chroma ? f.width / 2 : f.width, H = chroma ? f.height / 2 : f.height;
        const int xr = chroma ? 2 : 1, yr = chroma ? 2 : 1;
        const uint8_t* s0 = f.s0[plane];
        const uint8_t* s1 = f.s1[plane];
        auto tex0 = [&](int x, int y) {
            const uint8_t t = s0[(size_t)y * W + x];
            return chroma ? u8n(t) : lut[t];
        };
        auto tex1 = [&](int x, int y) {
            const uint8_t t = s1[(size_t)y * W + x];
            return chroma ? u8n(t) : lut[t];
        };
        const float maxx = (float)(W - 1), maxy = (float)(H - 1);
        const float denx = (float)(xr * 1) * 256.0f, deny = (float)(yr * 1) * 256.0f;  // pel = source_step = 1
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                auto source = [&](bool second, float vx, float vy, int time) {
                    float dx = vx * 65535.0f - 1024.0f, dy = vy * 65535.0f - 1024.0f;
                    dx = dx * (float)time / denx;
                    dy = dy * (float)time / deny;
                    const float px = std::fmin(std::fmax((float)x + dx, 0.0f), maxx);
                    const float py = std::fmin(std::fmax((float)y + dy, 0.0f), maxy);
                    return 255.0f * (second ? bilinear(tex1, W, H, px + 0.5f, py + 0.5f)
                                            : bilinear(tex0, W, H, px + 0.5f, py + 0.5f));
                };
                const float vpx = (float)(x * xr - f.origin_x) / (float)f.block_w;
                const float vpy = (float)(y * yr - f.origin_y) / (float)f.block_h;
                const F4 v = cubic(base_tex, mw, mh, vpx, vpy);
                const float ref_f = source(false, v.x, v.w, f.phase);
                const float ref_b = source(true, v.z, v.y, 256 - f.phase);
                const float ref_f0 = 255.0f * bilinear(tex0, W, H, (float)x + 0.5f, (float)y + 0.5f);
                const float ref_b0 = 255.0f * bilinear(tex1, W, H, (float)x + 0.5f, (float)y + 0.5f);
                const float time = (float)f.phase / 256.0f;
                F4 mask{0.0f, 0.0f, 0.0f, 0.0f};
                if (f.algo >= 21 || f.has_sad) mask = cubic(mask_tex, mw, mh, vpx, vpy);
                float result;
                switch (f.algo) {
                    case 1: result = ref_b; break;
                    case 2: result = ref_f; break;
                    case 11: result = mixf(ref_f, ref_b, time); break;
                    case 13: result = median3(ref_f, ref_b, mixf(ref_f0, ref_b0, time)); break;
                    case 21: result = mixf(mixf(ref_f, ref_b, mask.y), mixf(ref_b, ref_f, mask.z), time); break;
                    case 22:
                        result = median3(mixf(ref_f, ref_b, mask.y), mixf(ref_b, ref_f, mask.z),
                                         mixf(ref_f0, ref_b0, time));
                        break;
                    default: {
                        const F4 e = cubic(ext_tex, mw, mh, vpx, vpy);
                        const float ref_ff = source(false, e.x, e.w, f.phase);
                        const float ref_bb = source(true, e.z, e.y, 256 - f.phase);
                        result = mixf(mixf(ref_f, median3(ref_b, ref_bb, ref_f), mask.y),
                                      mixf(ref_b, median3(ref_b, ref_ff, ref_f), mask.z), time);
                    }
                }
                if (f.has_sad) {
                    if (f.algo == 1) result = mixf(result, ref_b0, mask.x);
                    else if (f.algo == 2) result = mixf(result, ref_f0, mask.w);
                    else result = mixf(result, mixf(ref_f0, ref_b0, time), mixf(mask.x, mask.w, time));
                }
                if (!chroma) {  // linear_luma: back to the gamma curve
                    float val = result / 255.0f;
                    val = val < 0.018f ? val * 4.5f : 1.099f * std::pow(val, 0.45f) - 0.099f;
                    result = val * 255.0f;
                }
                const float r = std::fmin(std::fmax(std::round(result), 0.0f), 255.0f);
                dst[plane][(size_t)y * W + x] = (uint8_t)r;
            }
    }
}

}  // namespace oracle
