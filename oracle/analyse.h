// Oracle restatement of the Analyse block search (crates/svpflow1/src/analyse_opts.rs,
// analyse_search.rs, analyse_filter.rs).  Test infrastructure only (see oracle.h).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "oracle.h"

namespace oracle {

struct AnalyseOpts {  // analyse_opts.rs:4-40
    int vectors = 3, block_w = 16, block_h = 16, overlap_mode = 2, overlap_x = 0, overlap_y = 0, delta = 1;
    int pel = 2, selector = 0, levels = 1, super_levels = 1, width = 0, height = 0;
    int lambda = 0, lsad = 0;
    double plevel = 1.5;
    int pnew = 50, pglobal = 50, pzero = 100, pnbour = 50, prev = 0;
    int search_type = 4, coarse_search_type = 4, distance = 2, coarse_distance = 0;
    bool coarse_trymany = false;
    int coarse_bad_sad = 0, coarse_bad_range = -24;
    bool fine_satd = false, coarse_satd = true;
    bool has_refine = false;
    int refine_thsad = 0, refine_search_type = 4, refine_distance = 2;
    bool refine_satd = false, sort = true;

    void output_block(int* bw, int* bh, int* ox, int* oy) const;  // :221-232
    void grid(int bw, int bh, int ox, int oy, int* gw, int* gh) const;  // :234-240
};

// analyse_opts.rs:45-219 (+ unpack_sdata :243-257).  `hint` replaces invalid sdata.
std::string analyse_opts_from(const Value* opt, int64_t sdata, const SuperOpts* hint, AnalyseOpts* out);
bool overlap_from_mode(int bw, int bh, int mode, int* ox, int* oy);  // :259-269
int auto_search_levels(int w, int h, int bw, int bh, int ox, int oy);  // :271-295

struct Vec8 {  // analyse_search.rs:25-31
    int16_t dx = 0, dy = 0;
    uint32_t score = 0;
    uint8_t luma = 0;
};

struct SuperPlanes {  // analyse_search.rs:33-66
    const uint8_t* y; size_t y_len, y_stride;
    const uint8_t* u; size_t u_len, u_stride;
    const uint8_t* v; size_t v_len, v_stride;
    size_t luma_w, luma_h;
    int pel, levels;
    bool full;
    size_t level_y_offset(int level) const;
    void level_size(int level, size_t* w, size_t* h) const;
};

// One record per block visit of a level search (level >= 0) or of the refine pass
// (level == -1): the inputs the reference's rules give the block and what it stored.
// The tests recompute each field from those rules (analyse_search.rs:480-768,
// :1326-1451) and compare.  All fields are 32-bit; scores are u32 bit patterns.
struct BlockTrace {
    int32_t level, dir, bx, by, px, py, bw, bh, edge, smallest;   // dir 0 = cur->ref, 1 = ref->cur
    int32_t pred_dx, pred_dy, pred_score;                         // predictor as read (score: u32)
    int32_t has_prev, prev_dx, prev_dy;                           // projected opposite candidate
    int32_t gx, gy;                                               // global predictor
    int32_t nseeds, seeds[16];                                    // seeds[2k], seeds[2k+1]
    int32_t lambda, radius, stype;
    int32_t stage_mvx, stage_mvy, stage_sad;                      // predictor-stage winner
    int32_t mpred_x, mpred_y;                                     // motion-penalty centre of the refinement
    int32_t found, bad_before, bad_limit, expanded;               // bad-block expansion (coarse levels)
    int32_t out_dx, out_dy, out_score, out_luma;                  // stored vector
    int32_t pad[13];
};
static_assert(sizeof(BlockTrace) == 64 * 4, "BlockTrace is 64 words");

// Instrumentation: candidate cost evaluations per level (index 0..15), both directions,
// and the pixels each touched (luma + chroma), for the algorithmic-work figure.
// `trace[dir]`, when set, receives a BlockTrace per block visit of that direction.
struct SearchCounters {
    uint64_t evals[16] = {};
    uint64_t pixels[16] = {};
    std::vector<BlockTrace>* trace[2] = {nullptr, nullptr};
    int dir = 0;
};

// analyse_pair (analyse_search.rs:68-92).  Fields are gw*gh raster vectors.
void analyse_pair(const SuperPlanes& cur, const SuperPlanes& ref, const AnalyseOpts& opts, bool* has_bwd,
                  std::vector<Vec8>* bwd, bool* has_fwd, std::vector<Vec8>* fwd, SearchCounters* counters,
                  bool threads);

std::vector<uint8_t> pack_vector_frame(const AnalyseOpts& opts, const std::vector<Vec8>* previous,
                                       const std::vector<Vec8>* current);  // :2740-2798
std::vector<int32_t> pack_vdata_header(const AnalyseOpts& opts);           // :2813-2841

// Exposed pieces for focused tests.
uint32_t hadamard4_satd(const int32_t diff[4][4]);                           // :2681-2702
std::vector<std::pair<int, int>> level_search_order(int gw, int gh, int level);  // :326-336
void level_grid(const AnalyseOpts& opts, int level, int* gw, int* gh);        // :1453-1463
int exact_lambda(const AnalyseOpts& opts, int level, int pel);                // :1279-1283
int adapt_lambda(int nlambda, int lsad, uint32_t pred_score);                 // :1555-1564
void exact_global_doubled(const std::vector<Vec8>& field, int* gx, int* gy); // :1285-1324
void rescale_scores(std::vector<Vec8>& field, int width, int height, int blk);  // :2716-2738
// Cost of one candidate (luma, chroma) by the reference's path selection.
void candidate_cost(const SuperPlanes& src, const SuperPlanes& ref, int level, int px, int py, int bw, int bh,
                    int mvx, int mvy, bool satd, int pel, bool chroma, uint32_t* luma, uint32_t* colour);

}  // namespace oracle
