// Device-side layouts of the SmoothFps render (shared by smooth.hip and smooth_host.cpp).
#pragma once
#include <cstdint>

namespace svpf {

// Per-clip constants of the render (renderer::CpuConfig + plane parameters + the
// options the per-frame decisions read).  Lives in device memory.
struct SmoothGeom {
    int32_t width, height;           // source frame
    int32_t vbw, vbh, vox, voy;      // vector block and overlap (vdata)
    int32_t gw, gh;                  // vector grid
    int32_t mw, mh;                  // motion grid (extended_grid)
    int32_t marker, flags;           // vdata pel marker and region flags
    int32_t count;                   // gw * gh
    int32_t payload_len;             // bytes of a vector payload (0x40 + 4 + 8 * count per direction)
    // CpuConfig (core.rs:1506-1518)
    int32_t cbw, cbh;                // effective block (block - overlap)
    int32_t cox, coy;                // origin = vector block / 2
    int32_t source_step;             // pel when the super frame is the source, else 1
    int32_t use_super, super_shift;  // source = pel^2 super plane (shift = log2 pel)
    int32_t super_rows;              // luma rows of a super frame
    int32_t xshift[2], yshift[2];    // cpu_map of the block (luma) / half block (chroma)
    int32_t xw[2][32][2], yw[2][32][2];  // interpolation weights, luma / chroma
    int32_t interp;                  // !block
    // render tiles per plane kind (luma, chroma): the last tile column and row (the first
    // tile reaching the plane end; -1: the origin lies past the plane, no tiles), and
    // ceil(2^32 / tile height) so that a row's tile is a multiply-high (0: height 1)
    int32_t tile_last_x[2], tile_last_y[2];
    uint32_t rcp_bh[2];
    // options
    int32_t algo_req, force13, mask_area, mask_cover;
    double area_scale, area_sharp;
    // scale_magnitude for every 24-bit magnitude, built on the host with the host libm's
    // pow (as the reference computes it), when area_sharp != 1; null otherwise.
    const uint8_t* mag_lut;
    int32_t lim_zero, lim_m1, lim_m2, lim_scene, lim_blocks;
    int32_t nframes;
    int32_t float_render;            // rendered frames through the float renderer (svpf_smooth_set_float_render)
    // light border (Options::padding and LightParams, options.rs:256-286; light.rs):
    // pads per side of the luma plane (chroma: half), zero without a border
    int32_t pad_x, pad_y;
    int32_t light_border, light_lights, light_length;
    double light_cell;
    int32_t light_slots;             // border profiles per frame over the three planes
    int32_t light_prof_len;          // floats of one frame's profiles
    int32_t light_draws;             // random draws of an even frame (the most a frame takes)
    int32_t light_pixels;            // border pixels per frame over the three planes
    const double* corner_geo[2];     // (mix, weight) per corner cell: luma pad_x x pad_y, chroma half
    float lin_lut[256];              // linear-light luma of each 8-bit sample (the float renderer's luma image)
    uint8_t luma_lut[512];           // luma_byte((a + b) / denom, gamma) for a + b in 0..510
    uint8_t cover_lut[4097];         // coverage_byte for remaining = 0..area (area <= 4096)
};

// Per-source-frame prep (decode + motion planes + scene class), core.rs:962-1051.
struct PrepJob {
    const uint8_t* payload;
    int16_t* vec[2];   // decoded (dx, dy) of set a (previous-or-current) and set b (current-or-previous)
    uint16_t* mv[4];   // mvx0, mvy0 (set a = previous), mvx1, mvy1 (set b = current)
    uint8_t* mag[2];   // magnitude masks (set a, set b), when mask.area > 0
    int32_t* stats;    // [valid, class, considered, scene, m2, m1]
};

// Per-output-frame plan, resolved on the device once the classes are known.
struct FramePlan {
    int32_t k, n, n1;          // output frame, source frame and its successor (clamped)
    int32_t raw_phase;         // Timing::phase_256 (or scene phase)
    int32_t phase_for[3];      // phase after adaptive_phase for scene classes 0..2
    int32_t early;             // 1: the frame is a copy whatever the vectors say
    // resolved on the device
    int32_t decision;          // 0 copy n, 1 copy n+1, 2 render
    int32_t algo;              // effective algorithm
    int32_t phase, limit;      // threshold and threshold_limit
    int32_t mode23_ready;
};

// Buffers of one output frame.  The frame's FramePlan sits at the same index of the
// plan table that follows the job table in the uploaded descriptor.
struct FrameJob {
    const int32_t* stats_n;    // prep stats of n, n-1, n+1
    const int32_t* stats_prev;
    const int32_t* stats_next;
    const uint16_t* mv_n[4];   // motion planes of n (a: mv0, b: mv1)
    const uint16_t* mv2[2];    // mvx0/mvy0 of n+1 (next frame's set a)
    const uint16_t* mv3[2];    // mvx1/mvy1 of n-1 (previous frame's set b)
    const uint8_t* mag[2];     // magnitude masks of n
    int32_t* cov_acc[2];       // (mw+2)*(mh+2) accumulators
    uint8_t* cov[2];           // coverage masks (mw*mh)
    uint8_t* maxmag;           // max of magnitude masks (mw*mh)
    const int16_t* vec_a;      // decoded vectors of n (set a / set b) as dx, dy pairs
    const int16_t* vec_b;
    const uint8_t* src0[3];    // source planes of n / n+1 (raw frames)
    const uint8_t* src1[3];
    int32_t sstride[3];
    const uint8_t* sup0[3];    // super planes of n / n+1
    const uint8_t* sup1[3];
    int32_t supstride[3];
    uint8_t* dst[3];           // the output frame's interior (the whole frame without padding)
    int32_t dstride[3];
    uint8_t* obase[3];         // the padded output frame (light border)
    float* prof;               // its border profiles
    int32_t* draws;            // its random draws
};

// Limits of the light border: profile length (frame side) and control points.
constexpr int kLightMaxDim = 8192;
constexpr int kLightMaxLights = 4096;

}  // namespace svpf
