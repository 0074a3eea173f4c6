// Device-side data layouts of the Analyse search (shared by analyse.hip and the
// host orchestration in analyse_host.cpp).
#pragma once
#include <cstddef>
#include <cstdint>

namespace svpf {

// One block vector of a level field (analyse_search.rs:25-31 Vec8), padded to 16 bytes.
struct alignas(16) DVec {
    int16_t dx, dy;
    uint32_t score;
    uint32_t luma;
    uint32_t done;  // unused by the search (visit positions decide which neighbours seed)
};

// Planes of one super frame (device pointers, byte strides, byte lengths).
struct SPlanes {
    const uint8_t* p[3];
    int32_t stride[3];
    int64_t len[3];
};

// Per-level geometry and options shared by all chains of one clip.
struct LevelGeom {
    int32_t level, gw, gh;
    int32_t lw, lh;        // luma level size
    int32_t y_off;         // first luma row of the level (level_y_offset)
    int32_t c_off;         // first chroma row of the level (chroma_level_offset)
    int32_t clw, clh;      // chroma level size
    int32_t pel;           // pel of this level (opts.pel at level 0, else 1)
    int32_t bw, bh;        // opts block size
    int32_t step_x, step_y;
    int32_t smallest;      // level == levels-1 (half-size centred block, no chroma)
    int32_t satd;
    int32_t search_type;
    int32_t radius;
    int32_t lambda_base;   // exact_lambda(level) (0 at the smallest level)
    int32_t lsad;          // opts.lsad (lsad/3 at the smallest level)
    int32_t pnew, pglobal, pzero, pnbour, pprev;
    int32_t early;         // level == 0 && distance == 0: predictor only
    int32_t bad_sad, bad_range;
    int32_t adaptive;      // coarse distance < 0: adaptive radius from block activity
};

// One (pair, direction) search chain at one level.
struct Chain {
    SPlanes src;           // frame whose blocks are searched
    SPlanes ref;           // reference frame
    const DVec* in;        // interpolated predictors of the level (read only)
    DVec* res;             // searched vectors (the level's result)
    const DVec* prev;      // projected opposite-direction field (nullptr at the top level)
    const int32_t* order;  // visit order, packed x | y << 16
    int32_t* opos;         // position of each block in the visit order (raster indexed)
    LevelGeom geom;        // the level's geometry (a copy per chain: one descriptor read gives both)
    const int32_t* global; // (gx, gy) global predictor from the coarser level; nullptr = (0, 0)
    uint32_t base;         // first index of this chain's blocks in the level's flag arrays
};

// Result of a block's refinement search, keyed by the predictor winner it started
// from.  For a block off the level's edge row and column, the refinement depends only
// on that start vector (its predictor, lambda, radius and bounds come from the level's
// fixed inputs), so a re-evaluation whose predictor list picks the same winner reuses
// the stored result instead of searching again.
struct alignas(8) Memo {
    int16_t kx, ky;   // predictor winner (start of the refinement)
    int16_t rx, ry;   // refined vector
    uint32_t rsad;    // its sad
    uint32_t valid;
    int64_t rcost;    // its cost (the bad-block expansion compares against it)
};

// What a block's last evaluation read from the level: the vector of each of its 3x3
// neighbours visited before it (slot 4, the block itself, unused) and the running
// bad-block count.  The other inputs of a block off the edge row and column are fixed
// for the level, so an evaluation whose neighbours and count read the same values
// again would reproduce the stored result; sweeps skip it.
struct alignas(16) SeedRec {
    uint32_t seed[9];  // (dy << 16) | (uint16_t)dx of present neighbours
    int32_t bad_count;
    uint32_t valid;
    uint32_t pad;
};
// validate_kernel reads a record as three 16-byte words (valid: word 10)
static_assert(sizeof(SeedRec) == 48 && offsetof(SeedRec, valid) == 40, "SeedRec layout");

// Worklists of the parallel sweeps of one level (all chains together).  Sweep s
// processes wl[s & 1][0 .. counts[s]) whose members are flagged in flags[s & 1]; a
// block whose result changes queues its later-visited 3x3 neighbours for sweep s + 1.
// Coarse levels with the bad-block pass also keep, per block, the pre-expansion cost
// (`found`), whether the expansion ran (`badf`) and the running count of bad blocks
// visited before it (`cpre`, rebuilt after every sweep from the first visit position
// whose found/bad changed, `minpos` per chain).  The rebuild reads copies of the three
// in visit order (`fseq`, `bseq`, `cseq`), which the evaluations and the rebuild keep
// equal to the per-block words, so its reads are contiguous and need no order lookup.
constexpr int kCntQueued = 3, kCntSkipped = 4, kCntWords = 8;
constexpr int kDcDeferred = 0, kDcChased = 1;

struct SweepState {
    unsigned long long* wl[2];  // items: chain index << 32 | block raster index
    uint32_t* flags[2];         // per block (chain base + raster index): queued for that parity
    // counts[s % 3] = items queued for sweep s >= 1 (a ring: sweep s reads its slot,
    // queues into the next and clears the one after, which sweep s - 1 read), so every
    // sweep with the same s mod 6 launches with the same arguments and parity;
    // counts[kCntQueued] = items of all later sweeps, counts[kCntSkipped] = items not
    // evaluated (statistics)
    uint32_t* counts;
    uint32_t* found;            // per block
    uint32_t* badf;             // per block
    int32_t* cpre;              // per block
    uint32_t* minpos;           // per chain
    Memo* memo;                 // per block
    SeedRec* seeds;             // per block
    unsigned long long* dl;     // blocks the tile-plan kernels passed on in the current sweep
    uint32_t* dcount;           // dcount[kDcDeferred]: their number; dcount[kDcChased]: chased evaluations
    uint32_t* fseq;             // per visit position: found, 0 for raster blocks 0 and 1 (never counted)
    uint32_t* bseq;             // per visit position: badf
    int32_t* cseq;              // per visit position: cpre
    uint32_t* stamp;            // per block: the last sweep (>= 1) whose evaluation claimed it
    int max_sweeps;
    int chase;                  // visit-order successors a later sweep evaluates right after a change
};


// Level-transition and packing job descriptors (one per chain).
struct Transition {
    const DVec* coarse;     // searched field of level l+1 (cgw x cgh)
    DVec* next;             // field of level l (gw x gh), written
    unsigned long long* proj_key;  // gw*gh projection keys (score << 32 | source raster index)
    DVec* proj;             // projected field for the opposite direction of level l
    int cgw, cgh, gw, gh;
    int bw, bh, ox, oy, pel;
};

struct GlobalJob {
    const DVec* field;
    int n;
    int lw, lh;     // bounds of |dx|, |dy| (level size)
    int* out;       // (gx, gy)
};

struct SortJob {
    const DVec* field;     // level l+1 field (gw x gh)
    uint32_t* keys;        // (gw-1)*(gh-1)
    uint32_t* vals;
    int gw, gh, shift;
    int seg_begin;         // offset of this chain's segment in the concatenated arrays
};

struct OrderJob {
    const uint32_t* parents;  // sorted parent positions (x | y << 16), count = (cgw-1)*(cgh-1)
    int32_t* order;           // order of level l (gw*gh entries)
    int cgw, cgh, gw, gh;
};

// The refine pass of one chain (exact_recalculate, analyse_search.rs:1326-1451): the
// chain's planes; `in` = its searched finest-level field (old grid), `res` = the refined
// field; `geom` describes the output grid at the finest level (output block and step,
// satd = the refine pass's).
struct RefineJob {
    Chain ch;
    int32_t old_gw, old_gh, old_bw, old_bh, old_sx, old_sy;
    int32_t thsad, lambda, pnew, type, distance;
};

struct PackJob {
    const DVec* field;      // final level-0 field
    uint8_t* out;           // payload region start (marker position)
    uint8_t* payload;       // payload start: receives the 64-byte header
    const int32_t* header;  // 16 header words
    int count;
};

}  // namespace svpf
