// Oracle restatement of the SmoothFps light border (crates/svpflow2/src/light.rs) and
// the output padding it fills (Options::padding, options.rs:256-277, 788-796).  The
// border is drawn on every output frame of a padded clip, after the interior was
// rendered or copied (core.rs:634-669, 1771, 1866-1897).  Test infrastructure only
// (see oracle.h).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "render.h"

namespace oracle {

namespace {

// srand/rand as the reference seeds them (light.rs:18-31): a 32-bit LCG, 15 bits out.
struct Rng {
    uint32_t state;
    int next() {
        state = state * 214013u + 2531011u;
        return (int)((state >> 16) & 0x7fff);
    }
};

constexpr int kTableLen = 128;    // light.rs:12
constexpr int kStripTables = 6;   // light.rs:14

struct PlaneCfg {  // light.rs:33-39, 86-99
    int base, border, table_h, table_v, table_corner;
};

enum Orient { kTop, kBottom, kLeft, kRight };  // light.rs:41-47

struct Geo {  // light.rs:432-437
    int width, height, pad_x, pad_y;
};

// Orient::coord (light.rs:439-448)
void coord(Orient o, const Geo& g, int d, int i, int* x, int* y) {
    switch (o) {
        case kTop: *x = g.pad_x + i; *y = d; break;
        case kBottom: *x = g.pad_x + i; *y = g.pad_y + g.height + d; break;
        case kLeft: *x = d; *y = g.pad_y + i; break;
        default: *x = g.pad_x + g.width + d; *y = g.pad_y + i; break;
    }
}

int trunc_i(double v) { return (int)v; }        // light.rs:589-592 (values in range)
int trunc_f(float v) { return (int)v; }         // light.rs:594-597
int noise_index(int base, int off) { return (int)((uint32_t)(base + off) & 0x7f); }  // :584-587

struct PlaneView {
    uint8_t* data;
    size_t len;
    int stride;
    // read / write (light.rs:560-582): outside the plane's bytes reads nothing (0) and
    // writes are dropped
    int read(int x, int y) const {
        if (x < 0 || y < 0) return 0;
        size_t i = (size_t)y * (size_t)stride + (size_t)x;
        return i < len ? data[i] : 0;
    }
    void write(int x, int y, int v) {
        if (x < 0 || y < 0) return;
        size_t i = (size_t)y * (size_t)stride + (size_t)x;
        if (i < len) data[i] = (uint8_t)(v & 0xFF);
    }
};

}  // namespace

// control_points (light.rs:479-511): per light k the mean of `accum` over a window of
// 2*half+3 positions starting at trunc(centre - half - 1), divided by border*count.
std::vector<int> light_control_points(const std::vector<int64_t>& accum, int dim, int lights, int border, int half,
                                      int window, int base) {
    std::vector<int> control((size_t)std::max(lights, 0), base);
    for (int k = 0; k < lights; k++) {
        const double center = ((double)k + 0.5) * (double)dim / (double)lights;
        int lo = trunc_i(center - (double)(half + 1));
        if (lo < 0) lo = 0;
        int hi = window + lo;
        if (hi >= dim) hi = dim;
        const int count = hi - lo;
        if (count > 0 && border > 0) {
            int64_t sum = 0;
            for (int i = lo; i < hi; i++) sum += accum[(size_t)i];
            const int64_t q = sum / ((int64_t)border * (int64_t)count);
            control[(size_t)k] = (q >= INT32_MIN && q <= INT32_MAX) ? (int)q : base;
        }
    }
    return control;
}

// cubic_upsample (light.rs:513-558): a cubic Hermite curve through the control points
// (tangents from the neighbours), evaluated at dim positions in f32.
std::vector<float> light_cubic_upsample(const std::vector<int>& control, int dim, int lights) {
    std::vector<float> out((size_t)std::max(dim, 0), 0.0f);
    if (control.empty()) return out;
    const float width_f = (float)(dim - 1);
    const float lights_f = (float)lights;
    const int last = lights - 1;
    int idx = -1;
    float p1 = (float)control[0], pr = (float)control[0], m0 = 0.0f, m1 = 0.0f;
    auto clampi = [](int v, int lo, int hi) { return std::min(std::max(v, lo), hi); };
    for (int o = 0; o < dim; o++) {
        const float nx = (float)o / width_f;
        const float ny = ((float)idx + 0.5f) / lights_f;
        if (nx > ny) {
            const int adv = idx + 1;
            const int right = clampi(adv, 0, last);
            const int left = std::max(idx, 0);
            const int lidx = std::max(idx, 1) - 1;
            const int rr = clampi(idx + 2, 0, last);
            p1 = (float)control[(size_t)left];
            pr = (float)control[(size_t)right];
            const float pl = (float)control[(size_t)lidx];
            const float prr = (float)control[(size_t)rr];
            const float slope = (pr - p1) * 0.5f;
            m0 = (p1 - pl) * 0.5f + slope;
            m1 = (prr - pr) * 0.5f + slope;
            idx = adv;
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
        const float value = term1 + (term2 + (term3 + term4));
        out[(size_t)o] = std::min(std::max(value, 0.0f), 255.0f);
    }
    return out;
}

// build_profile (light.rs:450-477): accum[col] = sum of `border` samples inward from the
// edge (line(r, col) for r < border), then the control points and their upsampling.
std::vector<float> light_profile(const std::vector<int64_t>& accum, int dim, int border, int lights, double cell,
                                 int base) {
    if (dim <= 0) return {};
    border = std::max(border, 1);
    lights = std::max(lights, 1);
    const int half = trunc_i((double)dim * cell / (double)(2 * lights));
    const int window = 2 * half + 3;
    const std::vector<int> control = light_control_points(accum, dim, lights, border, half, window, base);
    return light_cubic_upsample(control, dim, lights);
}

// corner_geometry (light.rs:350-373): per corner cell the blend between the two
// adjoining strips' end values and the fade weight.
void light_corner_geometry(int width, int height, int length, std::vector<double>* mix, std::vector<double>* weight) {
    mix->clear();
    weight->clear();
    for (int i = 0; i < height; i++)
        for (int j = 0; j < width; j++) {
            const double x = (double)(j + 1) / (double)width;
            const double y = (double)(i + 1) / (double)height;
            const double distance = std::sqrt(x * x + y * y) * 100.0;
            const double m = x > y ? y / x * 0.5 : (y >= 0.0001 ? 1.0 - x / y * 0.5 : 1.0);
            const double w = length > 0 ? std::max(1.0 - distance / (double)length, 0.0) : 0.0;
            mix->push_back(m);
            weight->push_back(w);
        }
}

// LightState::apply (light.rs:54-129) on one padded output frame (8-bit YUV420): the
// planes hold the rendered interior at (pad_x, pad_y) (chroma at half), the border
// around it is drawn here.  width/height: the source size.
void light_apply(int width, int height, int pad_x, int pad_y, const LightParams& lp, int frame, uint8_t* planes[3],
                 const int strides[3], const int rows[3]) {
    if (pad_x == 0 && pad_y == 0) return;
    frame = std::max(frame, 0);
    Rng rng{(uint32_t)((frame + 1) / 2)};
    if (frame % 2 == 0)
        for (int k = 0; k < 15; k++) rng.next();
    int noise[kStripTables][kTableLen];
    for (auto& t : noise)
        for (int& v : t) v = rng.next() % 3 - 1;
    const PlaneCfg luma{0, lp.border, 0, 2, 4};
    const PlaneCfg chroma{127, lp.border / 2, 1, 3, 5};
    for (int p = 0; p < 3; p++) {
        const PlaneCfg& cfg = p == 0 ? luma : chroma;
        PlaneView pv{planes[p], (size_t)strides[p] * (size_t)rows[p], strides[p]};
        const int w = p == 0 ? width : width / 2, h = p == 0 ? height : height / 2;
        const int px = p == 0 ? pad_x : pad_x / 2, py = p == 0 ? pad_y : pad_y / 2;
        // LightState::plane (light.rs:131-235)
        if (w <= 0 || h <= 0) continue;
        const Geo geo{w, h, px, py};
        const int lights_h = lp.lights;
        const int lights_v = std::max(lp.lights * h / w, 1);
        struct Prof {
            Orient o;
            std::vector<float> v;
        };
        std::vector<Prof> profs;
        auto prof = [&](Orient o) {
            const bool horiz = o == kTop || o == kBottom;
            const int dim = horiz ? w : h;
            const int border = std::max(cfg.border, 1);
            std::vector<int64_t> accum((size_t)std::max(dim, 0), 0);
            for (int c = 0; c < dim; c++) {
                int64_t s = 0;
                for (int r = 0; r < border; r++) {
                    int x, y;
                    if (o == kTop) { x = px + c; y = py + r; }
                    else if (o == kBottom) { x = px + c; y = py + h - 1 - r; }
                    else if (o == kLeft) { x = px + r; y = py + c; }
                    else { x = px + w - 1 - r; y = py + c; }
                    s += pv.read(x, y);
                }
                accum[(size_t)c] = s;
            }
            profs.push_back({o, light_profile(accum, dim, cfg.border, horiz ? lights_h : lights_v, lp.cell, cfg.base)});
        };
        if (py > 0) { prof(kTop); prof(kBottom); }
        if (px > 0) { prof(kLeft); prof(kRight); }
        // Endpoints::from_profiles (light.rs:386-430)
        auto saved = [&](Orient o, int pad, bool far, bool first) {
            for (const Prof& q : profs) {
                if (q.o != o) continue;
                if (q.v.empty()) return cfg.base;
                const float edge = first ? q.v.front() : q.v.back();
                const bool faded = (int64_t)pad * (int64_t)lp.length >= 200;
                const float factor = (far && faded && pad > 1) ? (float)pad / (float)(pad - 1) : 1.0f;
                const float base_f = (float)cfg.base;
                return trunc_f((edge - base_f) * factor + base_f);
            }
            return cfg.base;
        };
        const int top_left = saved(kTop, py, false, true), top_right = saved(kTop, py, false, false);
        const int bottom_left = saved(kBottom, py, true, true), bottom_right = saved(kBottom, py, true, false);
        const int left_top = saved(kLeft, px, false, true), left_bottom = saved(kLeft, px, false, false);
        const int right_top = saved(kRight, px, true, true), right_bottom = saved(kRight, px, true, false);
        // fill_strip (light.rs:237-279), profile by profile in the order they were built
        for (const Prof& q : profs) {
            const int* table = noise[(q.o == kTop || q.o == kBottom) ? cfg.table_h : cfg.table_v];
            const int dim = (int)q.v.size();
            const int pad = (q.o == kTop || q.o == kBottom) ? py : px;
            const bool faded = (int64_t)pad * (int64_t)lp.length >= 200;
            for (int d = 0; d < pad; d++) {
                double f = 1.0;
                if (faded && pad > 1)
                    f = (q.o == kTop || q.o == kLeft) ? (double)d / (double)(pad - 1) : (double)(pad - d) / (double)(pad - 1);
                const int base_rand = rng.next();
                const int noise_base = base_rand - (base_rand & 0x80);
                const float factor = (float)f;
                const float base_f = (float)cfg.base;
                for (int i = 0; i < dim; i++) {
                    int v = trunc_f((q.v[(size_t)i] - base_f) * factor + base_f);
                    if ((v & 0xFF) >= 1 && (v & 0xFF) <= 254) v += table[noise_index(noise_base, i)];
                    int x, y;
                    coord(q.o, geo, d, i, &x, &y);
                    pv.write(x, y, v);
                }
            }
        }
        // fill_corners (light.rs:281-342)
        if (px <= 0 || py <= 0) continue;
        std::vector<double> mix, wt;
        light_corner_geometry(px, py, lp.length, &mix, &wt);
        const int* table = noise[cfg.table_corner];
        struct Corner {
            int ox, oy, a5, a6;
            bool mx, my;
        };
        const Corner corners[4] = {{0, 0, left_top, top_left, true, true},
                                   {px + w, 0, right_top, top_right, false, true},
                                   {0, py + h, left_bottom, bottom_left, true, false},
                                   {px + w, py + h, right_bottom, bottom_right, false, false}};
        for (const Corner& c : corners)
            for (int row = 0; row < py; row++) {
                const int base_rand = rng.next();
                const int gi = c.my ? py - 1 - row : row;
                for (int col = 0; col < px; col++) {
                    const int gj = c.mx ? px - 1 - col : col;
                    const size_t ci = (size_t)gi * (size_t)px + (size_t)gj;
                    const double blend = (double)c.a6 * mix[ci] + (double)c.a5 * (1.0 - mix[ci]);
                    int v = cfg.base + trunc_i((blend - (double)cfg.base) * wt[ci]);
                    if ((v & 0xFFFF) >= 1 && (v & 0xFFFF) <= 254) v += table[noise_index(base_rand + col, 0)];
                    pv.write(c.ox + col, c.oy + row, v);
                }
            }
    }
}

// Options::padding (options.rs:256-277) with pad_x / pad_y (:788-796): the border the
// light.* options add on each side of the output.
void light_padding(const LightOpts& l, int width, int height, int* pad_x, int* pad_y) {
    *pad_x = *pad_y = 0;
    if (width <= 0 || height <= 0) return;
    const double w = (double)width, h = (double)height, src_aspect = w / h;
    double aspect = l.aspect / l.sar;
    if (aspect < 0.01) aspect = src_aspect;
    const double zoom = std::max(l.zoom, 0.0);
    int tw = width, th = height;
    auto tr = [](double v) -> int {
        if (std::isnan(v)) return 0;
        if (v >= 2147483647.0) return INT32_MAX;
        if (v <= -2147483648.0) return INT32_MIN;
        return (int)v;
    };
    if (std::fabs(aspect - src_aspect) <= 0.01) {
    } else if (src_aspect <= aspect) {
        tw = tr(aspect * h);
    } else {
        th = tr(w / aspect);
    }
    const int x = (int)((int64_t)tr((double)tw * zoom / 100.0) + tw - width);
    const int y = (int)((int64_t)tr((double)th * zoom / 100.0) + th - height);
    const int hx = x / 2, hy = y / 2;
    *pad_x = (hx + 3 * (hx + 1 < 0 ? 1 : 0) + 1) & ~3;
    *pad_y = (hy + (hy + 1 < 0 ? 1 : 0) + 1) & ~1;
}

}  // namespace oracle
