// SmoothFps light border for gfx950 (replaces crates/svpflow2/src/light.rs, drawn by
// core.rs:634-669, 1771, 1866-1897 on every output frame of a padded clip).
//
// Options::padding (options.rs:256-277) widens the output by a border on each side;
// the render writes the frame at (pad_x, pad_y) and this pass draws the border from it,
// for a batch of output frames:
//   * light_draws_kernel: the frame's random draws (the reference's srand/rand
//     sequence seeded by the output frame number), one per thread by LCG jump-ahead;
//   * light_profile_kernel: one workgroup per (frame, plane, side): the sums of
//     `border` samples inward from each edge position, their window means at `lights`
//     control points (prefix sums in LDS) and the cubic curve through them, evaluated
//     at every edge position in f32 exactly as the reference's running loop does (the
//     loop's segment index is a prefix minimum, so every position is computed at once);
//   * light_fill_kernel: one thread per border pixel: the strip value (profile faded
//     towards the outer edge, plus table noise) or the corner value (a blend of the two
//     adjoining strips' end values, faded by distance, plus noise).
// The arithmetic and its order are the reference's: f32 for profiles and strips, f64 for
// the corners (their geometry table is built on the host with the same formulas).
#include "smooth_dev.h"
#include "svpf_internal.h"

namespace svpf {
namespace {

constexpr int kLightThreads = 256;
constexpr int kTableLen = 128;  // light.rs:12
constexpr int kSkipEven = 15;   // draws discarded on even frames (light.rs:67-72)

enum { kTop = 0, kBottom = 1, kLeft = 2, kRight = 3 };

// Geometry of plane p (0 luma, 1-2 chroma at half size) of a padded frame.
struct PlaneGeo {
    int w, h, px, py;
};
__device__ __forceinline__ PlaneGeo plane_geo(const SmoothGeom& g, int p) {
    return p == 0 ? PlaneGeo{g.width, g.height, g.pad_x, g.pad_y}
                  : PlaneGeo{g.width / 2, g.height / 2, g.pad_x / 2, g.pad_y / 2};
}
__device__ __forceinline__ bool side_exists(const PlaneGeo& q, int o) { return o < 2 ? q.py > 0 : q.px > 0; }
__device__ __forceinline__ int side_dim(const PlaneGeo& q, int o) { return o < 2 ? q.w : q.h; }
__device__ __forceinline__ int side_pad(const PlaneGeo& q, int o) { return o < 2 ? q.py : q.px; }

// Offset of side o of plane p in the frame's profiles (planes in order, sides in the
// order the reference builds them: top, bottom, left, right).
__device__ __forceinline__ int prof_off(const SmoothGeom& g, int p, int o) {
    int off = 0;
    for (int pp = 0; pp < 3; pp++) {
        const PlaneGeo q = plane_geo(g, pp);
        for (int oo = 0; oo < 4; oo++) {
            if (!side_exists(q, oo)) continue;
            if (pp == p && oo == o) return off;
            off += side_dim(q, oo);
        }
    }
    return off;
}

// Random draws a plane takes: one per strip row (or column), then one per corner row.
__device__ __forceinline__ int plane_strip_draws(const PlaneGeo& q) {
    return (q.py > 0 ? 2 * q.py : 0) + (q.px > 0 ? 2 * q.px : 0);
}
__device__ __forceinline__ int plane_draws(const PlaneGeo& q) {
    return plane_strip_draws(q) + ((q.px > 0 && q.py > 0) ? 4 * q.py : 0);
}

// The LCG state after n steps from s (x -> 214013 x + 2531011 mod 2^32, light.rs:27-31),
// by squaring the affine map.
__device__ __forceinline__ uint32_t lcg_after(uint32_t s, uint32_t n) {
    uint32_t A = 1u, C = 0u, a = 214013u, c = 2531011u;
    while (n) {
        if (n & 1u) {
            A *= a;
            C = C * a + c;
        }
        c *= a + 1u;
        a *= a;
        n >>= 1;
    }
    return A * s + C;
}

__device__ __forceinline__ int out_frame(const FramePlan& pl) { return pl.k > 0 ? pl.k : 0; }

// Draw n (0-based) of output frame k's sequence: srand((k + 1) / 2), then rand() n + 1
// times (light.rs:62-72).
__global__ void __launch_bounds__(kLightThreads) light_draws_kernel(const FrameJob* __restrict__ jobs,
                                                                     const FramePlan* __restrict__ plans,
                                                                     const SmoothGeom* __restrict__ gp) {
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n >= gp->light_draws) return;
    const uint32_t seed = (uint32_t)((out_frame(plans[blockIdx.y]) + 1) / 2);
    jobs[blockIdx.y].draws[n] = (int32_t)((lcg_after(seed, (uint32_t)n + 1u) >> 16) & 0x7fffu);
}

// T(o): the first segment index i >= -1 at which the running loop of cubic_upsample
// (light.rs:524-541) stops advancing at position o, i.e. !(nx > ny(i)), ny(i) =
// (i + 0.5) / lights, which grows with i.
__device__ __forceinline__ int advance_limit(float nx, float lights_f, int lights) {
    int lo = -1, hi = lights;  // ny(lights) > 1 >= nx
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(nx > ((float)mid + 0.5f) / lights_f)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// build_profile + control_points + cubic_upsample (light.rs:450-558) of one side of one
// plane of one frame.
__global__ void __launch_bounds__(kLightThreads) light_profile_kernel(const FrameJob* __restrict__ jobs,
                                                                       const SmoothGeom* __restrict__ gp) {
    __shared__ uint32_t pre[kLightMaxDim + 1];  // prefix sums of the edge sums
    __shared__ int ctrl[kLightMaxLights];
    __shared__ int part[kLightThreads];
    const SmoothGeom& g = *gp;
    const FrameJob& j = jobs[blockIdx.y];
    // side blockIdx.x of the frame: planes in order, sides top, bottom, left, right
    int p = 0, o = 0;
    for (int q = (int)blockIdx.x, pp = 0; pp < 3; pp++) {
        const PlaneGeo G = plane_geo(g, pp);
        for (int oo = 0; oo < 4; oo++) {
            if (!side_exists(G, oo)) continue;
            if (q-- == 0) {
                p = pp;
                o = oo;
                pp = 3;
                break;
            }
        }
    }
    const PlaneGeo G = plane_geo(g, p);
    const int dim = side_dim(G, o);
    const int base = p == 0 ? 0 : 127;
    const int border = max(p == 0 ? g.light_border : g.light_border / 2, 1);
    // lights across the frame's width; a vertical side scales them by height / width
    const int lights = o < 2 ? g.light_lights : max(g.light_lights * G.h / G.w, 1);
    const uint8_t* plane = j.obase[p];
    const int stride = j.dstride[p];
    const int tid = (int)threadIdx.x, T = (int)blockDim.x;
    // the sum of `border` samples inward from each edge position (every one inside the frame)
    for (int c = tid; c < dim; c += T) {
        uint32_t sum = 0;
        for (int r = 0; r < border; r++) {
            int x, y;
            if (o == kTop) { x = G.px + c; y = G.py + r; }
            else if (o == kBottom) { x = G.px + c; y = G.py + G.h - 1 - r; }
            else if (o == kLeft) { x = G.px + r; y = G.py + c; }
            else { x = G.px + G.w - 1 - r; y = G.py + c; }
            sum += plane[(size_t)y * stride + x];
        }
        pre[c + 1] = sum;
    }
    if (tid == 0) pre[0] = 0;
    __syncthreads();
    const int chunk = (dim + T - 1) / T;
    const int lo = min(tid * chunk, dim), hi = min(lo + chunk, dim);
    {  // inclusive prefix sums (exact: a window sum stays below 2^32 with border <= 255)
        uint32_t acc = 0;
        for (int c = lo; c < hi; c++) acc += pre[c + 1];
        part[tid] = (int)acc;
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int t = 0; t < T; t++) {
                const uint32_t v = (uint32_t)part[t];
                part[t] = (int)run;
                run += v;
            }
        }
        __syncthreads();
        uint32_t run = (uint32_t)part[tid];
        for (int c = lo; c < hi; c++) {
            run += pre[c + 1];
            pre[c + 1] = run;
        }
        __syncthreads();
    }
    // control_points (light.rs:479-511)
    const int half = (int)((double)dim * g.light_cell / (double)(2 * lights));
    const int window = 2 * half + 3;
    for (int k = tid; k < lights; k += T) {
        const double center = ((double)k + 0.5) * (double)dim / (double)lights;
        int a = (int)(center - (double)(half + 1));
        if (a < 0) a = 0;
        int b = window + a;
        if (b >= dim) b = dim;
        const int count = b - a;
        ctrl[k] = count > 0 ? (int)((pre[b] - pre[a]) / (uint32_t)(border * count)) : base;
    }
    __syncthreads();
    // cubic_upsample (light.rs:513-558).  The loop's segment index obeys
    // idx(o) = min(idx(o - 1) + 1, T(o)) from idx(-1) = -1, so
    // idx(o) = min(o, o + min over j <= o of (T(j) - j)): a prefix minimum.
    const float width_f = (float)(dim - 1);
    const float lights_f = (float)lights;
    const int last = lights - 1;
    {
        int m = INT32_MAX;
        for (int c = lo; c < hi; c++) m = min(m, advance_limit((float)c / width_f, lights_f, lights) - c);
        part[tid] = m;
        __syncthreads();
        if (tid == 0) {
            int run = INT32_MAX;
            for (int t = 0; t < T; t++) {
                const int v = part[t];
                part[t] = run;
                run = min(run, v);
            }
        }
        __syncthreads();
    }
    float* out = j.prof + prof_off(g, p, o);
    int run = part[tid];
    for (int c = lo; c < hi; c++) {
        const float nx = (float)c / width_f;
        run = min(run, advance_limit(nx, lights_f, lights) - c);
        const int idx = min(c, c + run);
        float p1, pr, m0, m1;
        if (idx < 0) {  // before the first advance
            p1 = pr = (float)ctrl[0];
            m0 = m1 = 0.0f;
        } else {
            const int right = min(max(idx, 0), last), left = max(idx - 1, 0);
            const int lidx = max(idx - 1, 1) - 1, rr = min(max(idx + 1, 0), last);
            p1 = (float)ctrl[left];
            pr = (float)ctrl[right];
            const float pl = (float)ctrl[lidx], prr = (float)ctrl[rr];
            const float slope = (pr - p1) * 0.5f;
            m0 = (p1 - pl) * 0.5f + slope;
            m1 = (prr - pr) * 0.5f + slope;
        }
        const float t = nx * lights_f - (float)idx + 0.5f;
        const float tt = t * t;
        const float omt = 1.0f - t;
        const float omt2 = omt * omt;
        const float t2 = t + t;
        const float term1 = ((t - 1.0f) * tt) * m1;
        const float term2 = ((3.0f - t2) * tt) * pr;
        const float term3 = (omt2 * t) * m0;
        const float term4 = ((t2 + 1.0f) * omt2) * p1;
        const float v = term1 + (term2 + (term3 + term4));
        out[c] = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
    }
}

__device__ __forceinline__ int trunc_f(float v) { return (int)v; }

// A strip end value as Endpoints::from_profiles keeps it (light.rs:386-430).
__device__ __forceinline__ int strip_end(const SmoothGeom& g, const float* prof, const PlaneGeo& G, int p, int o,
                                         bool first, int base) {
    const int pad = side_pad(G, o);
    const float edge = first ? prof[prof_off(g, p, o)] : prof[prof_off(g, p, o) + side_dim(G, o) - 1];
    const bool far = o == kBottom || o == kRight;
    const bool faded = (long long)pad * (long long)g.light_length >= 200;
    const float factor = (far && faded && pad > 1) ? (float)pad / (float)(pad - 1) : 1.0f;
    const float base_f = (float)base;
    return trunc_f((edge - base_f) * factor + base_f);
}

// fill_strip and fill_corners (light.rs:237-342): one thread per border pixel.
__global__ void __launch_bounds__(kLightThreads) light_fill_kernel(const FrameJob* __restrict__ jobs,
                                                                    const FramePlan* __restrict__ plans,
                                                                    const SmoothGeom* __restrict__ gp) {
    const SmoothGeom& g = *gp;
    int rem = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (rem >= g.light_pixels) return;
    const FrameJob& j = jobs[blockIdx.y];
    const int frame = out_frame(plans[blockIdx.y]);
    const int skip = frame % 2 == 0 ? kSkipEven : 0;
    // plane, then the pixel: the top band, the middle rows' left and right parts, the bottom band
    int p = 0, draw0 = skip + 6 * kTableLen;
    PlaneGeo G = plane_geo(g, 0);
    for (;;) {
        G = plane_geo(g, p);
        const int ow = G.w + 2 * G.px;
        const int n = 2 * G.py * ow + 2 * G.px * G.h;
        if (rem < n || p == 2) break;
        rem -= n;
        draw0 += plane_draws(G);
        p++;
    }
    const int ow = G.w + 2 * G.px;
    int x, y;
    if (rem < G.py * ow) {
        y = rem / ow;
        x = rem - y * ow;
    } else if ((rem -= G.py * ow) < 2 * G.px * G.h) {
        const int r = rem / (2 * G.px), c = rem - r * 2 * G.px;
        y = G.py + r;
        x = c < G.px ? c : c + G.w;
    } else {
        rem -= 2 * G.px * G.h;
        const int r = rem / ow;
        y = G.py + G.h + r;
        x = rem - r * ow;
    }
    const int base = p == 0 ? 0 : 127;
    const int* draws = j.draws;
    auto noise = [&](int table, int k) { return draws[skip + table * kTableLen + (k & 0x7f)] % 3 - 1; };
    const bool in_x = x >= G.px && x < G.px + G.w, in_y = y >= G.py && y < G.py + G.h;
    int value;
    if (in_x || in_y) {  // a strip
        int o, d, i;
        if (in_x) {
            i = x - G.px;
            if (y < G.py) { o = kTop; d = y; }
            else { o = kBottom; d = y - G.py - G.h; }
        } else {
            i = y - G.py;
            if (x < G.px) { o = kLeft; d = x; }
            else { o = kRight; d = x - G.px - G.w; }
        }
        const int pad = side_pad(G, o);
        // the row's draw: strips in the order top, bottom, left, right, one draw per row
        int k = draw0 + d;
        if (o == kBottom) k += G.py;
        if (o >= kLeft) k += (G.py > 0 ? 2 * G.py : 0) + (o == kRight ? G.px : 0);
        const int base_rand = draws[k];
        const int noise_base = base_rand - (base_rand & 0x80);
        const bool faded = (long long)pad * (long long)g.light_length >= 200;
        double f = 1.0;
        if (faded && pad > 1)
            f = (o == kTop || o == kLeft) ? (double)d / (double)(pad - 1) : (double)(pad - d) / (double)(pad - 1);
        const float factor = (float)f;
        const float base_f = (float)base;
        value = trunc_f((j.prof[prof_off(g, p, o) + i] - base_f) * factor + base_f);
        const int table = p == 0 ? (o < 2 ? 0 : 2) : (o < 2 ? 1 : 3);
        if ((value & 0xFF) >= 1 && (value & 0xFF) <= 254) value += noise(table, noise_base + i);
    } else {  // a corner: 0 top-left, 1 top-right, 2 bottom-left, 3 bottom-right
        const bool right = x >= G.px + G.w, bottom = y >= G.py + G.h;
        const int c = (bottom ? 2 : 0) + (right ? 1 : 0);
        const int col = right ? x - G.px - G.w : x, row = bottom ? y - G.py - G.h : y;
        const int gi = bottom ? row : G.py - 1 - row;   // mirrored towards the frame
        const int gj = right ? col : G.px - 1 - col;
        const double* cell = g.corner_geo[p == 0 ? 0 : 1] + 2 * ((size_t)gi * G.px + gj);
        // a5: the vertical strip's end, a6: the horizontal strip's end (light.rs:293-327)
        const int a5 = strip_end(g, j.prof, G, p, right ? kRight : kLeft, !bottom, base);
        const int a6 = strip_end(g, j.prof, G, p, bottom ? kBottom : kTop, !right, base);
        const int base_rand = draws[draw0 + plane_strip_draws(G) + c * G.py + row];
        const double mix = cell[0], weight = cell[1];
        const double blend = (double)a6 * mix + (double)a5 * (1.0 - mix);
        value = base + (int)((blend - (double)base) * weight);
        if ((value & 0xFFFF) >= 1 && (value & 0xFFFF) <= 254) value += noise(p == 0 ? 4 : 5, base_rand + col);
    }
    j.obase[p][(size_t)y * j.dstride[p] + x] = (uint8_t)value;
}

}  // namespace

hipError_t launch_smooth_light(const FrameJob* d_jobs, const FramePlan* d_plans, int njobs,
                               const SmoothGeom* d_geom, const SmoothGeom& g, hipStream_t s) {
    if (njobs <= 0 || (g.pad_x == 0 && g.pad_y == 0)) return hipSuccess;
    hipLaunchKernelGGL(light_draws_kernel, dim3((g.light_draws + kLightThreads - 1) / kLightThreads, njobs),
                       dim3(kLightThreads), 0, s, d_jobs, d_plans, d_geom);
    hipLaunchKernelGGL(light_profile_kernel, dim3(g.light_slots, njobs), dim3(kLightThreads), 0, s, d_jobs, d_geom);
    hipLaunchKernelGGL(light_fill_kernel, dim3((g.light_pixels + kLightThreads - 1) / kLightThreads, njobs),
                       dim3(kLightThreads), 0, s, d_jobs, d_plans, d_geom);
    return hipGetLastError();
}

}  // namespace svpf
