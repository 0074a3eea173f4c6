// Oracle restatement of the SmoothFps render path: crates/svpflow-core/src/renderer.rs,
// metadata.rs, frame_math.rs and the render orchestration of crates/svpflow2/src/core.rs.
// Test infrastructure only (see oracle.h).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "oracle.h"

namespace oracle {

constexpr uint16_t NEUTRAL = 1024;

struct RVector {  // renderer::Vector (renderer.rs:11-16)
    int16_t dx = 0, dy = 0;
    uint32_t magnitude = 0;
};

struct VectorContext {  // renderer.rs:18-32
    int block_w, block_h, scale_shift_base, frame_w, frame_h, origin_x, origin_y;
    size_t grid_w, grid_h;
    bool raw;
    const std::vector<RVector>* a;
    const std::vector<RVector>* b;
};

// A source plane: linear (stride) or a pel^2 super plane (stride, span, shift), renderer.rs:90-128.
struct RPlane {
    const uint8_t* data = nullptr;
    size_t len = 0, stride = 0, span = 0;
    uint32_t shift = 0;
};

struct RPlaneMut {
    uint8_t* data = nullptr;
    size_t len = 0, stride = 0;
};

struct CpuConfig {  // renderer.rs:34-46
    int width, height, block_w, block_h, origin_x, origin_y;
    size_t grid_w, grid_h;
    int chroma_y_div, source_step;
    double scale;
};

struct Motion {
    const std::vector<uint16_t>* x;
    const std::vector<uint16_t>* y;
};

void vector_planes(const VectorContext& ctx, int which, std::vector<uint16_t>& dx, std::vector<uint16_t>& dy,
                   size_t w, size_t h);  // renderer.rs:1680-1736
void coverage_mask(const VectorContext& ctx, int which, std::vector<uint8_t>& out, size_t w, size_t h,
                   int strength, int threshold);  // renderer.rs:1738-1763
void magnitude_mask(const VectorContext& ctx, int which, std::vector<uint8_t>& out, size_t w, size_t h, double scale,
                    double exponent);  // renderer.rs:1647-1678
std::vector<uint8_t> max_mask(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b);
int truncdiv(int divisor, int value);  // TruncDiv::new(divisor).div(value), renderer.rs:3149-3184

class CpuRenderer {  // renderer.rs:63-1638
public:
    explicit CpuRenderer(const CpuConfig& c);
    void set_threshold(int t);
    int threshold() const { return threshold_; }
    int threshold_limit() const { return limit_; }
    // render_mode1_or_2 (renderer.rs:308-327)
    void mode1_or_2(bool mode1, bool interp, RPlaneMut dst[3], const RPlane s0[3], const RPlane s1[3], Motion m0,
                    Motion m1, const std::vector<uint8_t>* mask_a, const std::vector<uint8_t>* mask_b);
    // render_mode11_or_13 (renderer.rs:330-437); `max` = max_mask(area masks) or null
    void mode11_or_13(bool mode13, bool interp, bool zero_origin, RPlaneMut dst[3], const RPlane s0[3],
                      const RPlane s1[3], Motion m0, Motion m1, const std::vector<uint8_t>* max);
    // render_mode21_or_22 via render_plane_rows / _uv_rows (renderer.rs:571-641, 801-829)
    void mode21_or_22(bool mode21, bool interp, RPlaneMut dst[3], const RPlane s0[3], const RPlane s1[3], Motion m0,
                      Motion m1, const std::vector<uint8_t>& ca, const std::vector<uint8_t>& cb,
                      const std::vector<uint8_t>* final_mask);
    // render_mode23 (renderer.rs:644-666, 1175-1537)
    void mode23(bool interp, RPlaneMut dst[3], const RPlane s0[3], const RPlane s1[3], Motion m0, Motion m1,
                Motion m2, Motion m3, const std::vector<uint8_t>& ca, const std::vector<uint8_t>& cb,
                const std::vector<uint8_t>* final_mask);

private:
    CpuConfig cfg_;
    int wmap_, hmap_, hwmap_, hhmap_;
    int threshold_ = 0, limit_ = 0;
    int16_t inverse_[2048], tlut_[2048];
};

// The float GPU renderer's inputs for one output frame (gpu.rs:754-957, core.rs:3235-3330):
// the two source frames (planes at stride = plane width), the motion planes of the
// motion grid (m[0..3] = motion0.x/y, motion1.x/y; m[4..7] = motion2.x/y, motion3.x/y for
// algorithm 23), the coverage masks (algorithms >= 21) and the area (magnitude) masks.
struct FloatFrame {
    int width, height;
    const uint8_t* s0[3];
    const uint8_t* s1[3];
    int mw, mh;
    const uint16_t* m[8];
    const uint8_t *cov0, *cov1, *mag0, *mag1;
    bool has_sad;
    int algo;                      // 21 for algorithm 23 without its neighbours
    int phase;                     // the render threshold
    int block_w, block_h;          // effective block (block - overlap)
    int origin_x, origin_y;        // vector block / 2
};
void float_linear_lut(float lut[256]);                  // linear_luma kernel (gpu.rs)
void float_render(const FloatFrame& f, uint8_t* dst[3]);  // render_frame kernel (gpu.rs:1102-1167)

// ---------------------------------------------------------------- metadata / timing

struct VectorData {  // metadata.rs:1-12 (vdata header)
    int flags, block_w, block_h, width, height, grid_w, grid_h, marker, selector, delta, overlap_x, overlap_y;
    int motion_w() const;
    int motion_h() const;
};
bool parse_vdata(const int32_t h[16], VectorData* out);  // metadata.rs:90-114

struct DVec {
    int16_t dx, dy;
    uint32_t score;
    uint8_t luma;
};
// decode_vectors (metadata.rs:246-273): false on a malformed payload.
bool decode_vectors(const uint8_t* payload, size_t len, int mode_flags, int gw, int gh, int flags, bool* has_prev,
                    std::vector<DVec>* prev, bool* has_cur, std::vector<DVec>* cur);

struct Timing {  // frame_math.rs:1-136
    int64_t fps_num, fps_den, frame_num, frame_den;
    static Timing make(int64_t snum, int64_t sden, bool absolute, int64_t rnum, int64_t rden);
    int source_frame(int frame, bool nearest) const;
    int phase_256(int frame, int source_frame) const;
    double raw_phase_256(int frame, int source_frame) const;
    int scene_phase_256(double raw, int64_t mode) const;
    int scale_frame_count(int frames) const;
};

// ---------------------------------------------------------------- SmoothFps

// Light::from_value (options.rs:147-156) and LightParams (options.rs:279-286).
struct LightOpts {
    double sar = 1.0, aspect = 0.0, zoom = 0.0;
};
struct LightParams {
    int border = 12, lights = 16, length = 100;
    double cell = 1.0;
};

struct SmoothOptions {  // the options.rs subset the render path reads
    bool rate_abs = false;
    bool has_num = false, has_den = false;
    int64_t rate_num = 0, rate_den = 0;
    int64_t algo = 21;
    int64_t scene_mode = 3;
    bool scene_blend = false;
    int adaptive[3] = {-1, -1, 1};
    bool force13 = true;
    double scene_luma = 1.5;
    int64_t lim_blocks = 20, lim_zero = 200, lim_m1 = 1600, lim_m2 = 2800, lim_scene = 4000;
    int64_t mask_cover = 100, mask_area = 0;
    bool mask_area_enabled = false;
    double area_scale = 1.0, area_sharp = 1.0;
    bool block = false;
    LightOpts light;
    LightParams light_params;
    Timing timing(int64_t snum, int64_t sden) const;
};
// Options::from_value + create-time adjustments (options.rs:130-240, filter.rs:44-139).
bool smooth_options(const Value* opt, int64_t fps_num, int64_t fps_den, const VectorData& vd, SmoothOptions* out,
                    std::string* err);

// Everything the render of one output frame reads.
struct SmoothClip {
    int width, height, num_frames;
    int64_t fps_num, fps_den;
    int64_t sdata;
    VectorData vd;
    SmoothOptions opt;
    // per source frame: YUV planes (stride = width / width/2), super planes (stride = width / width/2,
    // rows = super height), vector payloads
    std::vector<const uint8_t*> src[3];
    std::vector<const uint8_t*> sup[3];
    int super_rows;
    std::vector<const uint8_t*> vec;
    size_t vec_len;
    // Rendered frames go through the float GPU renderer (gpu_render_frame, core.rs:1429,
    // 1630-1653) where the reference would: interpolated vectors, no scene blending.
    bool float_render = false;
};

int smooth_output_frames(const SmoothClip& c);
uint8_t luma_byte(double v, double gamma);  // metadata.rs:523-526
// analyze_scene (metadata.rs:349-362): class 0..3 of `cur` (luma from prev + cur).
int scene_class(const std::vector<DVec>* prev, const std::vector<DVec>* cur, const VectorData& vd,
                const SmoothOptions& o);
// Renders output frame k into dst (3 planes, strides width, width/2).  Returns a code of
// what happened: 0 copied source n, 1 copied source n+1, 2 rendered; *algo_out = algorithm used.
int smooth_render_frame(const SmoothClip& c, int k, uint8_t* dst[3], int* algo_out);

// Light border and output padding (light.rs, options.rs:256-277, 788-796).
void light_padding(const LightOpts& l, int width, int height, int* pad_x, int* pad_y);
void light_apply(int width, int height, int pad_x, int pad_y, const LightParams& lp, int frame, uint8_t* planes[3],
                 const int strides[3], const int rows[3]);
std::vector<int> light_control_points(const std::vector<int64_t>& accum, int dim, int lights, int border, int half,
                                      int window, int base);
std::vector<float> light_cubic_upsample(const std::vector<int>& control, int dim, int lights);
std::vector<float> light_profile(const std::vector<int64_t>& accum, int dim, int border, int lights, double cell,
                                 int base);
void light_corner_geometry(int width, int height, int length, std::vector<double>* mix, std::vector<double>* weight);

}  // namespace oracle
