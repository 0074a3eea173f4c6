// Analyse block search for gfx950 (replaces crates/svpflow1/src/analyse_search.rs).
//
// Execution model
//   * A pyramid level of every (frame pair, direction) chain of the batch is searched
//     as the fixed point of its dependency graph.  The reference visits the blocks in
//     one order (analyse_search.rs:480-681), each block seeding from the 3x3 neighbours
//     visited before it and, at the coarse levels, comparing with the running count of
//     bad blocks; a block's result depends only on blocks earlier in that order, so the
//     fixed point is unique and equals the sequential result.
//   * Sweep 0 evaluates every block once (sweep0_kernel), dispatched in visit order over
//     interleaved segments so that most blocks already see their predecessors' results;
//     validate_kernel then queues every block whose recorded seeds differ from its
//     neighbours' results.  Each later sweep (sweep_kernel) re-evaluates the queued
//     blocks and queues the later-visited neighbours of every block whose vector
//     changed, until a sweep changes nothing.  A block whose inputs are unchanged since
//     its last evaluation is skipped (seed record), and a refinement that starts from
//     the same predictor winner is reused (memo).  Coarse levels rebuild the running
//     bad-block counts after every sweep (sweep_scan_kernel).  A level that has not
//     converged within the sweep limit is searched sequentially (search_level_kernel).
//   * One wavefront evaluates one block.  Its inputs (neighbours, predictors, memo,
//     seed record, previous result) arrive in one round of loads, one word per lane.
//     The candidates of a search step go 16 per pass, 4 lanes each: lane `part` owns a
//     4x4 luma tile (and parts 0-1 the chroma tiles), costs are packed 16-bit SATD or
//     v_sad_u8, and the step's winner is the lowest (cost, list index) across groups
//     (DPP / permlane argmin), which equals the reference's sequential strict-`<`
//     running minimum (exact_check, :1122-1155); candidates that cannot win (out of the
//     bounds, or a motion penalty already at the best cost) are dropped before a pass.
//   * sweep0_kernel's tile-plan instantiation (three waves per SIMD) handles the blocks
//     whose tiles lie inside the level and passes the rest to deferred_kernel; the
//     finest level of a multi-level search has its own instantiation of both sweep
//     kernels, without the coarse levels' state.
//   * Level transitions (interpolation + opposite-field projection, global predictor,
//     score sort, child visit order) run as separate kernels over all chains, so a
//     batch of pairs never returns to the host between levels.
#include <cmath>
#include <type_traits>
#include <hipcub/hipcub.hpp>

#include "analyse_cost.h"
#include "analyse_dev.h"
#include "svpf_internal.h"

namespace svpf {
namespace {

constexpr int kWave = 64;
using namespace cost;

// Lane of the calling thread within its wavefront (kernels here run one or more whole
// wavefronts per workgroup).
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
// The lane index as a value the compiler cannot carry across block evaluations: the lane
// masks an evaluation derives from it are computed there, not kept (and spilled) across
// a kernel's loop.
__device__ __forceinline__ int lane_here() {
    int l = lane_id();
    asm volatile("" : "+v"(l));
    return l;
}

// One of three values by a per-lane index, selected as values: a conditional over
// struct members is an lvalue, which the compiler turns into a selected address and a
// load, placing the whole struct in scratch memory.
template <class T>
__device__ __forceinline__ T pick3(int i, T a, T b, T c) { return i == 0 ? a : (i == 1 ? b : c); }

// Packed 16-bit SATD of one 4x4 tile (same value as had4_rows): the differences and
// the first three butterfly stages run two columns per instruction; the last stage
// uses |a + b| + |a - b| = 2 max(|a|, |b|), so it needs no butterflies.  Every
// intermediate stays within +-2040.
typedef short __attribute__((ext_vector_type(2))) s16x2;

__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t pk_bytes01(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c010c00u); }
__device__ __forceinline__ uint32_t pk_bytes23(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c030c02u); }

__device__ __forceinline__ uint32_t had4_rows_pk(const uint32_t a[4], const uint32_t b[4]) {
    s16x2 P[4], M[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const s16x2 A = as_s16x2(pk_bytes01(a[r])) - as_s16x2(pk_bytes01(b[r]));  // columns 0, 1
        const s16x2 B = as_s16x2(pk_bytes23(a[r])) - as_s16x2(pk_bytes23(b[r]));  // columns 2, 3
        P[r] = A + B;
        M[r] = A - B;
    }
    s16x2 q[8];
    {
        const s16x2 p01s = P[0] + P[1], p01d = P[0] - P[1], p23s = P[2] + P[3], p23d = P[2] - P[3];
        const s16x2 m01s = M[0] + M[1], m01d = M[0] - M[1], m23s = M[2] + M[3], m23d = M[2] - M[3];
        q[0] = p01s + p23s; q[1] = p01s - p23s; q[2] = p01d + p23d; q[3] = p01d - p23d;
        q[4] = m01s + m23s; q[5] = m01s - m23s; q[6] = m01d + m23d; q[7] = m01d - m23d;
    }
    s16x2 acc = {0, 0};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const s16x2 v = __builtin_elementwise_max(q[k], -q[k]);  // |.|
        acc += __builtin_elementwise_max(v, v.yx);                // max of the pair, in both halves
    }
    return 2u * (uint32_t)(uint16_t)acc.x;
}

// Interior block cost: every source and reference byte lies inside its level, so rows
// are read as 32-bit words (unaligned) with no per-byte checks.  N x M block.
template <int N, int M>
__device__ __forceinline__ uint32_t block_fast(const uint8_t* s, int ss, const uint8_t* r, int rs, bool satd) {
    uint32_t total = 0;
    if (satd) {
#pragma unroll
        for (int ty = 0; ty < M / 4; ty++)
#pragma unroll
            for (int tx = 0; tx < N / 4; tx++) {
                uint32_t a[4], b[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    a[k] = ld32(s + (ty * 4 + k) * ss + tx * 4);
                    b[k] = ld32(r + (ty * 4 + k) * rs + tx * 4);
                }
                total += had4_rows(a, b);
            }
        return total >> 1;
    }
#pragma unroll
    for (int y = 0; y < M; y++)
#pragma unroll
        for (int x = 0; x < N; x += 4) total = __builtin_amdgcn_sad_u8(ld32(s + y * ss + x), ld32(r + y * rs + x), total);
    return total;
}

// Dispatch of block_fast over the block sizes the search uses; false = not handled.
__device__ __forceinline__ bool block_fast_any(int n, int m, const uint8_t* s, int ss, const uint8_t* r, int rs,
                                               bool satd, uint32_t& out) {
    if (n == 8 && m == 8) { out = block_fast<8, 8>(s, ss, r, rs, satd); return true; }
    if (n == 4 && m == 4) { out = block_fast<4, 4>(s, ss, r, rs, satd); return true; }
    if (n == 16 && m == 16) { out = block_fast<16, 16>(s, ss, r, rs, satd); return true; }
    if (n == 32 && m == 32) { out = block_fast<32, 32>(s, ss, r, rs, satd); return true; }
    return false;
}

// ------------------------------------------------------------------ cooperative cost

constexpr int kGroup = 4;                 // lanes per candidate
constexpr int kPerPass = kWave / kGroup;  // candidates per pass

// Wave-uniform reads of one lane's value (scalar results keep the search state in SGPRs).
__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ uint32_t rlu(uint32_t v, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)v, lane); }
__device__ __forceinline__ int64_t rl64(int64_t v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Cross-lane moves without the LDS crossbar: DPP within rows of 16 lanes, the gfx950
// permlane swaps across rows and halves.
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, 0xF, 0xF, false);
}
constexpr int kDppQuadXor1 = 0xB1;   // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadXor2 = 0x4E;   // quad_perm [2, 3, 0, 1]
constexpr int kDppRowRor4 = 0x124;     // lane i <- lane (i + 4) mod 16 within each 16 lanes
constexpr int kDppRowRor8 = 0x128;     // lane i <- lane i ^ 8 within each 16 lanes

// Sum over the 4 lanes of a group (a quad; every lane of the group receives it).
__device__ __forceinline__ uint32_t group_sum(uint32_t v) {
    v += dpp<kDppQuadXor1>(v);
    v += dpp<kDppQuadXor2>(v);
    return v;
}

// Half of a 4x4 tile's SATD, computed by a pair of lanes (quad lanes 2j and 2j + 1, both
// active): this lane holds rows 0-1 (`upper`) or rows 2-3 of the tile.  The horizontal
// butterflies and the vertical one within the two rows run in-lane; the vertical stage
// joining the halves takes the partner's row sums (upper) or row differences (lower)
// by DPP, so each lane completes half of the coefficients, and the last stage is
// had4_rows_pk's 2 max(|a|, |b|).  The pair's two results add up to
// had4_rows_pk(tile) / 2.
__device__ __forceinline__ uint32_t had4_half_pk(const uint32_t a[2], const uint32_t b[2], bool upper) {
    s16x2 P[2], M[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const s16x2 A = as_s16x2(pk_bytes01(a[r])) - as_s16x2(pk_bytes01(b[r]));  // columns 0, 1
        const s16x2 B = as_s16x2(pk_bytes23(a[r])) - as_s16x2(pk_bytes23(b[r]));  // columns 2, 3
        P[r] = A + B;
        M[r] = A - B;
    }
    const s16x2 ps = P[0] + P[1], pd = P[0] - P[1], ms = M[0] + M[1], md = M[0] - M[1];
    // the upper lane keeps its sums and sends its differences, the lower one the reverse
    const s16x2 rp = as_s16x2(dpp<kDppQuadXor1>(__builtin_bit_cast(uint32_t, upper ? pd : ps)));
    const s16x2 rm = as_s16x2(dpp<kDppQuadXor1>(__builtin_bit_cast(uint32_t, upper ? md : ms)));
    const s16x2 ap = upper ? ps : rp, bp = upper ? rp : pd;  // (rows 01 + rows 23) or (rows 01 - rows 23)
    const s16x2 am = upper ? ms : rm, bm = upper ? rm : md;
    const s16x2 q[4] = {ap + bp, ap - bp, am + bm, am - bm};
    s16x2 acc = {0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const s16x2 v = __builtin_elementwise_max(q[k], -q[k]);
        acc += __builtin_elementwise_max(v, v.yx);
    }
    return (uint32_t)(uint16_t)acc.x;
}

// One candidate's cost computed by the 4 lanes of a group (part = lane & 3); every
// lane of the group receives (L, C).  Blocks whose luma and chroma lie inside their
// levels (the common case) split into units -- 4x4 tiles for SATD, rows for SAD --
// spread over the group; any other path is computed whole by part 0 (cand_cost_full).
__device__ __forceinline__ void group_cost(const Blk& b, int mvx, int mvy, bool chroma, int part, uint32_t& Lout,
                                           uint32_t& Cout) {
    const int pel = b.pel;
    const int b0 = -pel * b.px, b1 = -pel * b.py;
    const int b2 = pel * (b.lw - b.px - b.bw) + 1, b3 = pel * (b.lh - b.py - b.bh) + 1;
    const bool edge = mvx < b0 || (b2 > b0 && mvx >= b2) || mvy < b1 || (b3 > b1 && mvy >= b3);
    const bool sq = b.satd && pow2_4_32(b.bw) && pow2_4_32(b.bh) && b.bw == b.bh;
    const bool subpel = b.level == 0 && pel >= 2;
    // interior luma geometry (block_cost_luma_interior)
    const int ipx = b.px - ((b.level > 0 && b.px + b.bw >= b.lw) ? 1 : 0);
    const int ipy = b.py - ((b.level > 0 && b.py + b.bh >= b.lh) ? 1 : 0);
    const int mxp = subpel ? (mvx >> 1) : mvx, myp = subpel ? (mvy >> 1) : mvy;
    const int sub = subpel ? (mvy & 1) * pel + (mvx & 1) : 0;
    const int frx = ipx + mxp, fry = ipy + myp;
    const bool lfast = !edge && pow2_4_32(b.bw) && b.bw == b.bh && ipx >= 0 && ipy >= 0 && ipx + b.bw <= b.lw &&
                       ipy + b.bh <= b.lh && frx >= 0 && fry >= 0 && frx + b.bw <= b.lw && fry + b.bh <= b.lh;
    // interior chroma geometry (chroma_sad_x4)
    const int cw = b.bw / 2, ch = b.bh / 2;
    const bool cuse = chroma && cw >= 4 && ch >= 4;
    const int cpx = b.px / 2, cpy = b.py / 2;
    const int mcx = subpel ? (mvx >> 2) : (mvx >> 1), mcy = subpel ? (mvy >> 2) : (mvy >> 1);
    const int crx = cpx + mcx, cry = cpy + mcy;
    const bool cfast = cw == ch && (cw == 4 || cw == 8 || cw == 16) && cpx + cw <= b.clw && cpy + ch <= b.clh &&
                       crx >= 0 && cry >= 0 && crx + cw <= b.clw && cry + ch <= b.clh;
    const bool fast = lfast && (!cuse || cfast);
    uint32_t a0 = 0, a1 = 0, a2 = 0;  // luma, u, v partial sums (fast) or (L, C, 0) from part 0
    if (fast) {
        const int rrow = b.y_off + sub * b.lh;
        const int cref = b.c_off + ((b.level == 0 && pel > 1) ? sub * b.clh : 0);
        const uint8_t* ls = b.sp0 + (int64_t)(b.y_off + ipy) * b.ss0 + ipx;
        const uint8_t* lr = b.rp0 + (int64_t)(rrow + fry) * b.rs0 + frx;
        const uint8_t* us = b.sp1 + (int64_t)(b.c_off + cpy) * b.ss1 + cpx;
        const uint8_t* ur = b.rp1 + (int64_t)(cref + cry) * b.rs1 + crx;
        const uint8_t* vs = b.sp2 + (int64_t)(b.c_off + cpy) * b.ss2 + cpx;
        const uint8_t* vr = b.rp2 + (int64_t)(cref + cry) * b.rs2 + crx;
        if (sq) {  // luma 4x4 tiles, then u tiles, then v tiles (chroma SATD whenever luma is square SATD)
            const int lt = b.bw >> 2, ct = cw >> 2;
            const int nl = lt * lt, nc = cuse ? ct * ct : 0;
            for (int u = part; u < nl + 2 * nc; u += kGroup) {
                const int plane = u < nl ? 0 : (u < nl + nc ? 1 : 2);
                const int t = plane == 0 ? u : (plane == 1 ? u - nl : u - nl - nc);
                const int per = plane == 0 ? lt : ct;
                const int ty = t / per, tx = t - ty * per;
                const uint8_t* sp = plane == 0 ? ls : (plane == 1 ? us : vs);
                const uint8_t* rp = plane == 0 ? lr : (plane == 1 ? ur : vr);
                const int ss = pick3(plane, b.ss0, b.ss1, b.ss2);
                const int rs = pick3(plane, b.rs0, b.rs1, b.rs2);
                uint32_t x[4], y[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    x[k] = ld32(sp + (int64_t)(ty * 4 + k) * ss + tx * 4);
                    y[k] = ld32(rp + (int64_t)(ty * 4 + k) * rs + tx * 4);
                }
                const uint32_t h = had4_rows_pk(x, y);
                if (plane == 0) a0 += h; else if (plane == 1) a1 += h; else a2 += h;
            }
        } else if (!b.satd) {  // SAD rows: luma rows, then u rows, then v rows
            const int nl = b.bh, nc = cuse ? ch : 0;
            for (int u = part; u < nl + 2 * nc; u += kGroup) {
                const int plane = u < nl ? 0 : (u < nl + nc ? 1 : 2);
                const int r = plane == 0 ? u : (plane == 1 ? u - nl : u - nl - nc);
                const uint8_t* sp = plane == 0 ? ls : (plane == 1 ? us : vs);
                const uint8_t* rp = plane == 0 ? lr : (plane == 1 ? ur : vr);
                const int ss = pick3(plane, b.ss0, b.ss1, b.ss2);
                const int rs = pick3(plane, b.rs0, b.rs1, b.rs2);
                const int words = (plane == 0 ? (int)b.bw : cw) >> 2;
                uint32_t acc = 0;
                for (int w = 0; w < words; w++)
                    acc = __builtin_amdgcn_sad_u8(ld32(sp + (int64_t)r * ss + 4 * w), ld32(rp + (int64_t)r * rs + 4 * w), acc);
                if (plane == 0) a0 += acc; else if (plane == 1) a1 += acc; else a2 += acc;
            }
        } else {  // SATD requested on a square non-power-of-two size: SAD luma, SATD chroma
            // (not reached: lfast requires a power-of-two square block, so sq holds whenever satd does)
        }
    } else if (part == 0) {
        cand_cost_full(b, mvx, mvy, chroma, a0, a1);
    }
    a0 = group_sum(a0);
    a1 = group_sum(a1);
    a2 = group_sum(a2);
    if (fast) {
        Lout = sq ? (a0 >> 1) : a0;
        const uint32_t U = b.satd ? (a1 >> 1) : a1, V = b.satd ? (a2 >> 1) : a2;
        Cout = cuse ? sat_mul(sat_add(U, V), 4) : 0;
    } else {
        Lout = a0;
        Cout = a1;
    }
}

// ---- compile-time tile path of 4x4 and 8x8 blocks

typedef __attribute__((address_space(1))) const uint8_t g_u8;
// Chain descriptors are written by the host before a launch and only read by kernels.
typedef __attribute__((address_space(4))) const struct Chain const_chain;
typedef __attribute__((address_space(1))) const uint32_t __attribute__((aligned(1))) g_u32u;
typedef __attribute__((address_space(4))) const struct SweepState const_sweep;
typedef __attribute__((address_space(4))) const char const_byte;
// The leading arguments of the kernels that read the sweep state this way (the chain
// descriptors, then the state): its place in the kernel-argument segment.
struct SweepArgsHead {
    const Chain* chains;
    SweepState st;
};
// The kernel's sweep-state argument as seen by one block evaluation: read in place in
// the kernel-argument segment through a pointer the compiler cannot follow across
// evaluations, so that the fields load (scalar loads) where an evaluation uses them
// instead of living across the kernel's loop.  (The address of the by-value parameter
// itself would be a private copy's.)
__device__ __forceinline__ const SweepState& state_here() {
    const_sweep* p = (const_sweep*)((const_byte*)__builtin_amdgcn_kernarg_segment_ptr() + offsetof(SweepArgsHead, st));
    asm volatile("" : "+s"(p));
    return *(const SweepState*)p;
}
__device__ __forceinline__ uint32_t gld32(const g_u8* p) { return *reinterpret_cast<const g_u32u*>(p); }

// Each lane of a group (part = lane % 4) owns the luma 4x4 tile `part` of an 8x8 block
// (part 0 alone for 4x4) and two rows of a chroma tile: rows 0-1 (even parts) or 2-3
// (odd parts) of the u tile (parts 0-1) or the v tile (parts 2-3), whose SATD a pair of
// lanes computes together (had4_half_pk).  Their source rows are loaded once per block.  Candidates whose luma and chroma blocks lie inside their
// levels and off the edge rule (the vectors of one rectangle, fixed per block: the
// conditions of group_cost's interior path) take this path; the rest take group_cost.
struct TilePlan {
    int on;                       // uniform
    int xlo, xhi, ylo, yhi;       // uniform: the rectangle of vectors
    int subpel, pel, lh, clh;     // uniform
    const g_u8* lbase;            // uniform: reference luma plane
    int lrs, satd;                // uniform: its stride; SATD (else SAD)
    int dc, n;                    // uniform: the tiles are the block's own pixels; block size
    int lrow0, lcol0, luse;       // lane: luma tile
    int crow0, ccol0, cuse, crs;  // lane: two rows of a chroma tile (u: parts 0-1, v: parts 2-3)
    const g_u8* cbase;
    uint32_t lsrc[4], csrc[2];
};

__device__ __forceinline__ void make_plan(const Blk& b, bool chroma, int part, TilePlan& P) {
    P.on = 0;
    const int N = b.bw;
    if (b.bw != b.bh || (N != 4 && N != 8)) return;
    const bool subpel = b.level == 0 && b.pel >= 2;
    const int ipx = b.px - ((b.level > 0 && b.px + b.bw >= b.lw) ? 1 : 0);
    const int ipy = b.py - ((b.level > 0 && b.py + b.bh >= b.lh) ? 1 : 0);
    const int cw = N / 2;
    const bool cuse = chroma && cw >= 4;
    const int cpx = b.px / 2, cpy = b.py / 2;
    if (ipx < 0 || ipy < 0 || ipx + N > b.lw || ipy + N > b.lh) return;
    if (cuse && (cpx + cw > b.clw || cpy + cw > b.clh)) return;
    const int sh = subpel ? 1 : 0, csh = subpel ? 2 : 1;
    const int pel = b.pel;
    const int b0 = -pel * b.px, b1 = -pel * b.py;
    const int b2 = pel * (b.lw - b.px - b.bw) + 1, b3 = pel * (b.lh - b.py - b.bh) + 1;
    // (m >> s) >= a  <=>  m >= a * 2^s;  (m >> s) <= a  <=>  m <= (a + 1) * 2^s - 1
    int xlo = max(b0, -ipx * (1 << sh)), ylo = max(b1, -ipy * (1 << sh));
    int xhi = (b.lw - N - ipx + 1) * (1 << sh) - 1, yhi = (b.lh - N - ipy + 1) * (1 << sh) - 1;
    if (b2 > b0) xhi = min(xhi, b2 - 1);
    if (b3 > b1) yhi = min(yhi, b3 - 1);
    if (cuse) {
        xlo = max(xlo, -cpx * (1 << csh));
        ylo = max(ylo, -cpy * (1 << csh));
        xhi = min(xhi, (b.clw - cw - cpx + 1) * (1 << csh) - 1);
        yhi = min(yhi, (b.clh - cw - cpy + 1) * (1 << csh) - 1);
    }
    P.on = 1;
    P.xlo = xlo; P.xhi = xhi; P.ylo = ylo; P.yhi = yhi;
    P.lbase = (const g_u8*)b.rp0;
    P.lrs = b.rs0;
    P.satd = b.satd ? 1 : 0;
    P.dc = ipx == b.px && ipy == b.py;
    P.n = N;
    P.subpel = subpel ? 1 : 0;
    P.pel = pel;
    P.lh = b.lh;
    P.clh = subpel ? b.clh : 0;
    const int tx = part & 1, ty = (part >> 1) & 1;
    P.luse = N == 8 || part == 0;
    P.lrow0 = b.y_off + ipy + 4 * ty;
    P.lcol0 = ipx + 4 * tx;
    P.cuse = cuse;
    const int chalf = 2 * (part & 1);  // this lane's first row of its chroma tile
    P.crow0 = b.c_off + cpy + chalf;
    P.ccol0 = cpx;
    P.crs = part < 2 ? b.rs1 : b.rs2;
    P.cbase = (const g_u8*)(part < 2 ? b.rp1 : b.rp2);
    const int css = part < 2 ? b.ss1 : b.ss2;
    const g_u8* ls = (const g_u8*)b.sp0 + (P.lrow0 * b.ss0 + P.lcol0);
    const g_u8* cs = (const g_u8*)(part < 2 ? b.sp1 : b.sp2) + (P.crow0 * css + P.ccol0);
#pragma unroll
    for (int k = 0; k < 4; k++) P.lsrc[k] = P.luse ? gld32(ls + k * b.ss0) : 0u;
#pragma unroll
    for (int k = 0; k < 2; k++) P.csrc[k] = P.cuse ? gld32(cs + k * css) : 0u;
}

__device__ __forceinline__ bool plan_covers(const TilePlan& P, int mvx, int mvy) {
    return P.on && mvx >= P.xlo && mvx <= P.xhi && mvy >= P.ylo && mvy <= P.yhi;
}

// Cost of one 4x4 tile against the reference rows at r (stride rs).
__device__ __forceinline__ uint32_t tile_cost(const uint32_t src[4], const g_u8* r, int rs, bool satd) {
    uint32_t y[4];
#pragma unroll
    for (int k = 0; k < 4; k++) y[k] = gld32(r + k * rs);
    if (satd) return had4_rows_pk(src, y) >> 1;
    uint32_t h = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) h = __builtin_amdgcn_sad_u8(src[k], y[k], h);
    return h;
}

// The candidate's (luma, 4 x chroma) cost through the plan; every lane of the group
// receives it (the same values as group_cost).  Rows, subplane indexes and strides are
// below 2^16 and a covered candidate's row is not negative, so their products are
// exact as 24-bit multiplies (full rate; a 32-bit multiply is quarter rate).
__device__ __forceinline__ void plan_cost(const TilePlan& P, int mvx, int mvy, uint32_t& L, uint32_t& C) {
    const int sub = P.subpel ? __mul24(mvy & 1, P.pel) + (mvx & 1) : 0;
    uint32_t hl = 0, hc = 0;
    if (P.luse) {
        const int sh = P.subpel;  // luma vectors are in half pels when sub-pel
        const int row = P.lrow0 + __mul24(sub, P.lh) + (mvy >> sh), col = P.lcol0 + (mvx >> sh);
        hl = tile_cost(P.lsrc, P.lbase + (__mul24(row, P.lrs) + col), P.lrs, P.satd != 0);
    }
    if (P.cuse) {  // (uniform: the whole group takes part in the pair exchange)
        const int csh = P.subpel + 1;
        const int row = P.crow0 + __mul24(sub, P.clh) + (mvy >> csh), col = P.ccol0 + (mvx >> csh);
        const g_u8* r = P.cbase + (__mul24(row, P.crs) + col);
        const uint32_t y[2] = {gld32(r), gld32(r + P.crs)};
        if (P.satd) {
            hc = had4_half_pk(P.csrc, y, (threadIdx.x & 1) == 0);
        } else {
            hc = __builtin_amdgcn_sad_u8(P.csrc[0], y[0], 0u);
            hc = __builtin_amdgcn_sad_u8(P.csrc[1], y[1], hc);
        }
    }
    L = group_sum(hl);
    C = 4u * group_sum(hc);
}

struct Best {
    int mvx, mvy;
    uint32_t sad;
    int64_t cost;
};

// Motion-penalty context of the current search (predictor, bounds, lambda, pnew).
struct Search {
    int pmx, pmy;
    int bnd[4];
    int lambda, pnew;
    bool chroma;
    bool serial;  // pnew < -256: replay the reference's skip rules candidate by candidate
};

__device__ __forceinline__ bool in_bounds(const Search& s, int x, int y) {
    return x >= s.bnd[0] && y >= s.bnd[1] && x < s.bnd[2] && y < s.bnd[3];
}

// motion_penalty (:1263-1267) of vector (x, y) against the search's predictor.  Vector
// components are 16-bit, so |dx|, |dy| < 2^16 and each square (below 2^32, as an
// unsigned sum below 2^33) is exact from a 24-bit multiply's 48-bit product.
__device__ __forceinline__ int64_t motion_of(const Search& s, int x, int y) {
    const int dx = s.pmx - x, dy = s.pmy - y;
    const uint64_t d2 = (uint64_t)(uint32_t)__mul24(dx, dx) + (uint64_t)(uint32_t)__mul24(dy, dy);
    return ((int64_t)s.lambda * (int64_t)d2) >> 8;
}

// exact_check's first two tests (:1122-1155): inside the vector bounds, and a motion
// penalty below the best cost (with pnew >= -256 a candidate's cost is at least its
// motion penalty, so a point failing this cannot win).
__device__ __forceinline__ bool survives(const Search& s, int64_t best_cost, int x, int y) {
    return in_bounds(s, x, y) && motion_of(s, x, y) < best_cost;
}

// Number of set bits of the wave mask m below this lane.
__device__ __forceinline__ int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Forward permute: lane `dst` receives v (ds_permute_b32; the whole wave must be active).
__device__ __forceinline__ int fwd_permute(int dst, int v) { return __builtin_amdgcn_ds_permute(dst << 2, v); }

// Lane 0 of each quad to the whole quad (DPP quad_perm [0, 0, 0, 0]).
__device__ __forceinline__ uint32_t quad_bcast0(uint32_t v) { return dpp<0x00>(v); }

__device__ __forceinline__ void take_lower(int64_t& c, int& k, int64_t oc, int ok) {
    if (oc < c || (oc == c && ok < k)) { c = oc; k = ok; }
}

__device__ __forceinline__ int64_t join64(uint32_t lo, uint32_t hi) { return (int64_t)(((uint64_t)hi << 32) | lo); }

// Lowest (cost, index) across the groups of the wave (values are uniform per group;
// every lane receives the result; needs the whole wave active).
__device__ __forceinline__ void group_argmin(int64_t& c, int& k) {
    {  // the 4 groups of a row of 16 lanes: rotations by 4, then by 8
        const uint32_t lo = dpp<kDppRowRor4>((uint32_t)c), hi = dpp<kDppRowRor4>((uint32_t)((uint64_t)c >> 32));
        take_lower(c, k, join64(lo, hi), (int)dpp<kDppRowRor4>((uint32_t)k));
    }
    {
        const uint32_t lo = dpp<kDppRowRor8>((uint32_t)c), hi = dpp<kDppRowRor8>((uint32_t)((uint64_t)c >> 32));
        take_lower(c, k, join64(lo, hi), (int)dpp<kDppRowRor8>((uint32_t)k));
    }
    {  // rows r and r ^ 1: the swap leaves one row's values in each result
        const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)c, (uint32_t)c, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)((uint64_t)c >> 32),
                                                         (uint32_t)((uint64_t)c >> 32), false, false);
        const auto kk = __builtin_amdgcn_permlane16_swap((uint32_t)k, (uint32_t)k, false, false);
        c = join64(lo[0], hi[0]);
        k = (int)kk[0];
        take_lower(c, k, join64(lo[1], hi[1]), (int)kk[1]);
    }
    {  // halves
        const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)c, (uint32_t)c, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)((uint64_t)c >> 32),
                                                         (uint32_t)((uint64_t)c >> 32), false, false);
        const auto kk = __builtin_amdgcn_permlane32_swap((uint32_t)k, (uint32_t)k, false, false);
        c = join64(lo[0], hi[0]);
        k = (int)kk[0];
        take_lower(c, k, join64(lo[1], hi[1]), (int)kk[1]);
    }
}

// Costs below this pack with the group index (kPerPass = 16 groups) into 32 bits.
constexpr int64_t kPackedCost = (int64_t)1 << 27;
static_assert(kPerPass <= 16, "4 bits of group index");

// Minimum of v over the wave (values uniform per group; needs the whole wave active);
// every lane receives it, as a scalar.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, dpp<kDppRowRor4>(v));
    v = min(v, dpp<kDppRowRor8>(v));
    const auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    v = min((uint32_t)a[0], (uint32_t)a[1]);
    const auto b2 = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    v = min((uint32_t)b2[0], (uint32_t)b2[1]);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__constant__ int kHex2[8][2] = {{-1, -2}, {-2, 0}, {-1, 2}, {1, 2}, {2, 0}, {1, -2}, {-1, -2}, {-2, 0}};
__constant__ int kMod6m1[8] = {5, 0, 1, 2, 3, 4, 5, 0};

// Candidate lists of the reference's search steps, in its evaluation order.
enum PatKind : int {
    kPatPred = 0,   // predictor list (lanes hold the candidates)
    kPatExpand,     // exact_expanding (:829-886): square ring of radius r, step
    kPatHexFirst,   // exact_hex2 first step (:941-1033): 6 points
    kPatHexStep,    // exact_hex2 iteration: 3 points from odir
    kPatUmh,        // exact_umh (:1036-1119): cross, then 16-point hexagons
    kPatExh,        // exhaustive (2r+1)^2 window, raster order (:771-826)
    kPatHexExp,     // kPatHexFirst's 6 points, then expanding(1, 1) around the same centre
};

struct Pattern {
    int kind, cx, cy, r, step, n;
};

__device__ __forceinline__ int pat_count(int kind, int r, int step) {
    if (kind == kPatExpand) return 4 * ((2 * r - 1) / step) + 4;
    if (kind == kPatHexFirst) return 6;
    if (kind == kPatHexExp) return 6 + 8;
    if (kind == kPatHexStep) return 3;
    if (kind == kPatUmh) return 4 * (r / 2) + 16 * max(r / 4, 1);
    return (2 * r + 1) * (2 * r + 1);  // kPatExh
}

// kHex2 and kHex4 as 4-bit fields (x + 2, y + 2 and x + 4, y + 4), so a lane's entry
// is a shift of a constant rather than a memory read.
constexpr uint32_t kHex2X = 0x01343101u, kHex2Y = 0x20024420u;
constexpr uint64_t kHex4X = 0x6422468888800000ull, kHex4Y = 0x1017876543223456ull;

// Point k of expanding(cx, cy, r, step) (:829-886): the top and bottom rows, the left
// and right columns, then the four corners.
__device__ __forceinline__ void pat_expand(int cx, int cy, int r, int step, int k, int& x, int& y) {
    const unsigned uk = (unsigned)k;
    const int m = (2 * r - 1) / step;  // values of i in -r+step, -r+2step, ... < r (uniform)
    if (k < 2 * m) {
        x = cx - r + step * (1 + (int)(uk >> 1));
        y = (k & 1) ? cy + r : cy - r;
    } else if (k < 4 * m) {
        const unsigned j = uk - 2 * m;
        y = cy - r + step * (1 + (int)(j >> 1));
        x = (j & 1) ? cx + r : cx - r;
    } else {
        const int c = k - 4 * m;  // (-r,-r), (-r,+r), (+r,-r), (+r,+r)
        x = (c < 2) ? cx - r : cx + r;
        y = (c & 1) ? cy + r : cy - r;
    }
}

// kHex2[e] around (cx, cy).
__device__ __forceinline__ void pat_hex(int cx, int cy, unsigned e, int& x, int& y) {
    x = cx + (int)((kHex2X >> (4 * e)) & 15u) - 2;
    y = cy + (int)((kHex2Y >> (4 * e)) & 15u) - 2;
}

__device__ __forceinline__ void pat_gen_kind(int kind, const Pattern& p, int k, int& x, int& y) {
    const int cx = p.cx, cy = p.cy, r = p.r;
    const unsigned uk = (unsigned)k;
    if (kind == kPatExpand) {
        pat_expand(cx, cy, r, p.step, k, x, y);
    } else if (kind == kPatHexFirst || kind == kPatHexStep) {  // kHex2[k + 1], or kHex2[odir + k] (p.step = odir)
        pat_hex(cx, cy, kind == kPatHexFirst ? uk + 1 : (unsigned)p.step + uk, x, y);
    } else if (kind == kPatHexExp) {
        if (k < 6) pat_hex(cx, cy, uk + 1, x, y);
        else pat_expand(cx, cy, 1, 1, k - 6, x, y);
    } else if (kind == kPatUmh) {
        const int nd = r / 2;
        if (k < 2 * nd) {
            const int d = 1 + 2 * (int)(uk >> 1);
            x = (k & 1) ? cx + d : cx - d;
            y = cy;
        } else if (k < 4 * nd) {
            const unsigned j = uk - 2 * nd;
            const int d = 1 + 2 * (int)(j >> 1);
            x = cx;
            y = (j & 1) ? cy + d : cy - d;
        } else {
            const unsigned j = uk - 4 * nd, h = j & 15u;
            const int sc = 1 + (int)(j >> 4);
            x = cx + __mul24((int)((kHex4X >> (4 * h)) & 15u) - 4, sc);  // |.| <= 4 times a small scale
            y = cy + __mul24((int)((kHex4Y >> (4 * h)) & 15u) - 4, sc);
        }
    } else {  // kPatExh: k = row * w + column, the quotient by a multiply (k, w < 2^16)
        const unsigned w = 2u * (unsigned)r + 1u;
        unsigned q = w == 1u ? uk : __umulhi(uk, 0xFFFFFFFFu / w + 1u);
        if (q * w > uk) q--;
        if (uk - q * w >= w) q++;
        x = cx - r + (int)(uk - q * w);
        y = cy - r + (int)q;
    }
}
__device__ __forceinline__ void pat_gen(const Pattern& p, int k, int& x, int& y) { pat_gen_kind(p.kind, p, k, x, y); }

// adapt_lambda (:1555-1564), IEEE double as the reference.
__device__ __attribute__((always_inline)) int adapt_lambda(int nl, int lsad, uint32_t score) {
    const double l = (double)max(lsad, 1);
    const double half = score == 0xFFFFFFFFu ? -1.0 : (double)(score >> 1);
    const double t = __ddiv_rn(l, __dadd_rn(l, half));
    const double v = fmin(__dmul_rn(__dmul_rn((double)nl, t), t), 21474836.47);
    if (isnan(v)) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int)v;
}

// Per-lane byte sum helper for block_luma_dc (:1566-1617) and plane activity.
__device__ __attribute__((always_inline)) uint8_t luma_dc(const Blk& b) {
    const int lane = lane_id();
    uint32_t sum = 0, cnt = 0;
    // the reference stops a row at the first column past the level, and stops at the first row past it
    const int rows = min(b.bh, max(b.lh - b.py, 0));
    const int cols = min(b.bw, max(b.lw - b.px, 0));
    for (int i = lane; i < rows * cols; i += kWave) {
        const int r = i / cols, c = i % cols;
        sum += (uint32_t)rd(b.sp0, b.sl0, (int64_t)(b.y_off + b.py + r) * b.ss0 + b.px + c, 0);
        cnt++;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sum += __shfl_xor(sum, off);
        cnt += __shfl_xor(cnt, off);
    }
    return cnt == 0 ? 0 : (uint8_t)min(sum / cnt, 255u);
}

// plane_activity (:1517-1553) of one plane region, wave-parallel.
__device__ __attribute__((always_inline)) uint32_t plane_activity(const uint8_t* p, int stride, int row, int x, int w, int h, bool satd) {
    const int lane = lane_id();
    const uint32_t area = (uint32_t)(w * h);
    uint32_t sum = 0;
    for (int i = lane; i < w * h; i += kWave) sum += (uint32_t)rd(p, INT64_MAX, (int64_t)(row + i / w) * stride + x + i % w, 0);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
    const uint32_t mean = sum / area;
    uint32_t acc = 0;
    if (!satd) {
        for (int i = lane; i < w * h; i += kWave) {
            const uint32_t v = (uint32_t)rd(p, INT64_MAX, (int64_t)(row + i / w) * stride + x + i % w, 0);
            acc += v > mean ? v - mean : mean - v;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
        return acc / area;
    }
    const int tw = (w + 3) / 4, th = (h + 3) / 4;
    for (int t = lane; t < tw * th; t += kWave) {
        const int y0 = (t / tw) * 4, x0 = (t % tw) * 4;
        int d[16];
        for (int dy = 0; dy < 4; dy++)
            for (int dx = 0; dx < 4; dx++)
                d[dy * 4 + dx] = rd(p, INT64_MAX, (int64_t)(row + y0 + dy) * stride + x + x0 + dx, 0) - (int)mean;
        acc += had4(d);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    return (acc >> 1) / area;
}

// A wave-uniform struct read into scalar registers: every lane loads the same words
// and readfirstlane marks them uniform (the kernels also store to global memory, so
// the compiler cannot prove the descriptors unclobbered and would keep them in VGPRs).
template <class T>
__device__ __forceinline__ T uniform_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized");
    T v;
    const int* src = reinterpret_cast<const int*>(p);
    int* dst = reinterpret_cast<int*>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return v;
}
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ void init_blk(const Chain& ch, const LevelGeom& g, Blk& b) {
    b.sp0 = ch.src.p[0]; b.sp1 = ch.src.p[1]; b.sp2 = ch.src.p[2];
    b.rp0 = ch.ref.p[0]; b.rp1 = ch.ref.p[1]; b.rp2 = ch.ref.p[2];
    b.ss0 = ch.src.stride[0]; b.ss1 = ch.src.stride[1]; b.ss2 = ch.src.stride[2];
    b.rs0 = ch.ref.stride[0]; b.rs1 = ch.ref.stride[1]; b.rs2 = ch.ref.stride[2];
    b.sl0 = (int)ch.src.len[0]; b.sl1 = (int)ch.src.len[1]; b.sl2 = (int)ch.src.len[2];
    b.rl0 = (int)ch.ref.len[0]; b.rl1 = (int)ch.ref.len[1]; b.rl2 = (int)ch.ref.len[2];
    b.level = g.level; b.lw = g.lw; b.lh = g.lh; b.y_off = g.y_off; b.c_off = g.c_off;
    b.clw = g.clw; b.clh = g.clh; b.pel = g.pel; b.satd = g.satd != 0;
    b.px = 0; b.py = 0; b.bw = g.bw; b.bh = g.bh;
}

// Steps of one block's search (exact_search_block :684-768, the refine searches
// :771-1119, the coarse bad-block pass :607-672).  A step names what to do with the
// result of the candidate list just evaluated; every list goes through the single
// evaluation site of search_block.
enum Step : int {
    kStEnd = 0,
    kStEdgeDone,      // edge row/column: the predictor's score measured at its source block
    kStPredDone,      // predictor candidates evaluated
    kStUmhDone,       // exact_umh's cross + hexagons evaluated: continue with exact_hex2
    kStHexBegin,      // exact_hex2 entry (no list)
    kStHexFirstDone,
    kStHexLoop,       // (no list)
    kStHexStepDone,
    kStExp1Done,      // exact_hex2's closing expanding(1, 1) evaluated
    kStExhDone,
    kStRefineDone,    // (no list) bad-block decision
    kStBadRing,       // (no list)
    kStBadRingDone,
    kStBadFinal,      // (no list)
    kStBadFinalDone,
};

// One block of exact_search_level (:512-680).  `in` holds the level's interpolated
// predictors and `res` the results; the 3x3 neighbours visited before this block
// (opos below its own) seed the search and, on the edge row/column, give the
// predictor -- exactly what the reference's `processed` grid selects in its visit
// order.  Returns the block's vector (uniform across the wave), *bad = whether the
// coarse-level bad-block expansion ran and *found_out = the cost before it.
// A search of the refine pass (exact_recalculate, :1326-1451): the rectangle searched,
// the clipped interpolated start vector (also the motion predictor) and the pass's
// options; `factor` is 3 on the border (half-size rectangle, score x 3), else 1.
struct RefineStart {
    int px, py, bw, bh;
    int mvx, mvy;
    int lambda, pnew;  // (lambda >> 2) / factor (0 on the first row), pnew / factor
    int type, radius;
    int factor;
};

// kFast: the tile-plan instantiation, without the general cost path (group_cost,
// luma_dc), so that it fits three waves per SIMD.  *defer tells the caller what the
// general instantiation must redo: kDeferWhole -- an edge block (its predictor is
// measured at the neighbour's position) or a block without a plan; it returns at once,
// having stored nothing; kDeferExact -- a candidate fell outside the plan's rectangle
// and was skipped: the result is provisional (close, so that the blocks visited next
// seed from it), no memo entry is stored and the seed record is invalidated.
constexpr int kDeferNone = 0, kDeferWhole = 1, kDeferExact = 2;

// The sweep state's words of one block (set by a sweep's evaluation; nullptr-free) and,
// on return, what the evaluation read of them and of the block's previous result.
struct BlockIO {
    const int32_t* cpre;   // running bad-block count before the block
    const uint32_t* found;
    const uint32_t* badf;
    bool fresh;            // sweep 0: the memo and seed record are known invalid
    DVec old;              // out: the block's stored result before this evaluation
    uint32_t oldf, oldb;   // out: its found / badf words
    int me;                // out: its visit position
};

// Lanes of a block's gathered inputs; lanes 0-8 read the 3x3 neighbours.
enum : int {
    kInMe = 9, kInPred = 10, kInPrev = 12, kInGlob = 14, kInMemo = 16, kInRecBad = 22, kInRecValid = 23,
    kInOld = 24, kInFound = 28, kInBadf = 29, kInCpre = 30,
};

// kL0: the finest level of a multi-level search with pnew >= -256, as compile-time
// facts (no edge row/column, no half-size block, no bad-block pass, no adaptive radius,
// no serial replay), so that the instantiation carries none of their state.
// kLean: a coarser level whose chains all have pnew >= -256 (no serial replay).  With
// kFast, both run the search's steps as straight-line code instead of the step machine.
template <bool kRefine, bool kFast = false, bool kL0 = false, bool kLean = false>
__device__ __forceinline__ DVec search_block(const Chain& ch, const LevelGeom& g, Blk& b, int idx,
                                             int bad_count, bool* bad, uint32_t* found_out, Memo* memo,
                                             SeedRec* rec, bool* unchanged, const RefineStart& rs,
                                             int* defer = nullptr, BlockIO* io = nullptr) {
    const int lane = lane_here();
    const int part = lane & (kGroup - 1);
    const int gw = g.gw, gh = g.gh, level = kL0 ? 0 : g.level;
    const bool smallest = !kRefine && !kL0 && g.smallest != 0;
    if constexpr (kL0) b.level = 0;
    const int bx = idx % gw, by = idx / gw;
    int px, py, bw, bh;
    if constexpr (kRefine) {
        px = rs.px; py = rs.py; bw = rs.bw; bh = rs.bh;
    } else {
        px = (level > 0 && bx == gw - 1) ? g.lw - g.bw : bx * g.step_x;
        py = (level > 0 && by == gh - 1) ? g.lh - g.bh : by * g.step_y;
        bw = g.bw; bh = g.bh;
        if (smallest) { px += g.bw >> 2; py += g.bh >> 2; bw = g.bw >> 1; bh = g.bh >> 1; }
    }
    b.px = px; b.py = py; b.bw = bw; b.bh = bh;
    const bool edge = !kRefine && level > 0 && (bx == gw - 1 || by == gh - 1);
    if constexpr (kFast) {
        if (edge && !smallest) {
            *defer = kDeferWhole;
            return DVec{};
        }
    }
    // the tile plan of the block at its own position (its source rows load here, in
    // flight with the inputs below); the edge predictor's measure uses group_cost
    TilePlan tp;
    tp.on = 0;
    if (!edge || smallest) {
        make_plan(b, !smallest, part, tp);
        if constexpr (kFast) {
            if (!(tp.on && tp.dc)) {
                *defer = kDeferWhole;
                return DVec{};
            }
        }
    }
    // ---- the block's inputs in one round of loads: lane k < 9 reads its 3x3 neighbour
    // (visit position, vector) and the seed record's copy of it; lanes 9-30 read one
    // word each of the block's own inputs, taken below with readlane.
    const bool fresh = io != nullptr && io->fresh;
    uint32_t word = 0, nvec = 0;
    int npos = INT32_MAX;
    if constexpr (!kRefine) {
        const uint32_t* own = reinterpret_cast<const uint32_t*>(ch.opos + idx);  // stand-in for absent words
        const bool rec_ok = rec != nullptr && !fresh, memo_ok = memo != nullptr && !fresh;
        const uint32_t* w = own;
        if (lane < 9) w = rec_ok ? rec->seed + lane : own;
        else if (lane < kInPred) w = own;
        else if (lane < kInPrev) w = reinterpret_cast<const uint32_t*>(ch.in + idx) + (lane - kInPred);
        else if (lane < kInGlob) w = ch.prev ? reinterpret_cast<const uint32_t*>(ch.prev + idx) + (lane - kInPrev) : own;
        else if (lane < kInMemo) w = ch.global ? reinterpret_cast<const uint32_t*>(ch.global) + (lane - kInGlob) : own;
        else if (lane < kInRecBad) w = memo_ok ? reinterpret_cast<const uint32_t*>(memo) + (lane - kInMemo) : own;
        else if (lane == kInRecBad) w = rec_ok ? reinterpret_cast<const uint32_t*>(&rec->bad_count) : own;
        else if (lane == kInRecValid) w = rec_ok ? &rec->valid : own;
        else if (lane < kInFound) w = io ? reinterpret_cast<const uint32_t*>(ch.res + idx) + (lane - kInOld) : own;
        else if (lane == kInFound) w = io ? io->found : own;
        else if (lane == kInBadf) w = io ? io->badf : own;
        else if (lane == kInCpre) w = io ? reinterpret_cast<const uint32_t*>(io->cpre) : own;
        word = *w;
        const int nx = bx - 1 + lane % 3, ny = by - 1 + lane / 3;
        if (lane < 9 && nx >= 0 && ny >= 0 && nx < gw && ny < gh) {
            const int n = ny * gw + nx;
            npos = ch.opos[n];
            // device-scope load: the neighbour may have been stored in this launch by
            // another CU (its vector is the first word of the record)
            nvec = __hip_atomic_load(reinterpret_cast<const uint32_t*>(&ch.res[n]), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    auto in_word = [&](int k) { return rlu(word, k); };
    int gx = 0, gy = 0, me = 0;
    if constexpr (kRefine) {
        if (ch.global) { gx = uniform_int(ch.global[0]); gy = uniform_int(ch.global[1]); }
    } else {
        if (ch.global) { gx = (int)in_word(kInGlob); gy = (int)in_word(kInGlob + 1); }
        me = (int)in_word(kInMe);
        if (io) {
            bad_count = (int)in_word(kInCpre);
            io->me = me;
            const uint32_t o0 = in_word(kInOld);
            io->old = DVec{(int16_t)(o0 & 0xFFFFu), (int16_t)(o0 >> 16), in_word(kInOld + 1), in_word(kInOld + 2),
                           in_word(kInOld + 3)};
            io->oldf = in_word(kInFound);
            io->oldb = in_word(kInBadf);
        }
    }
    DVec pred;
    int epx = 0, epy = 0;
    if constexpr (kRefine) {
        pred = DVec{(int16_t)rs.mvx, (int16_t)rs.mvy, 0u, 0u, 0u};
    } else if (edge) {
        const int si = min(by, gh - 2) * gw + min(bx, gw - 2);
        if (uniform_int(ch.opos[si]) < me) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            pred = uniform_load(&ch.res[si]);
        } else {
            pred = uniform_load(&ch.in[si]);
        }
        epx = (si % gw) * g.step_x;
        epy = (si / gw) * g.step_y;
    } else {
        const uint32_t p0 = in_word(kInPred);
        pred = DVec{(int16_t)(p0 & 0xFFFFu), (int16_t)(p0 >> 16), in_word(kInPred + 1), 0u, 0u};
    }
    int radius = kRefine ? rs.radius : g.radius;
    if (!kRefine && level > 0 && g.adaptive) {
        const int ax = px - ((px + bw >= g.lw) ? 1 : 0), ay = py - ((py + bh >= g.lh) ? 1 : 0);
        uint32_t act = plane_activity(b.sp0, b.ss0, g.y_off + ay, ax, bw, bh, b.satd) +
                       plane_activity(b.sp1, b.ss1, g.c_off + py / 2, px / 2, bw / 2, bh / 2, b.satd) +
                       plane_activity(b.sp2, b.ss2, g.c_off + py / 2, px / 2, bw / 2, bh / 2, b.satd);
        radius = act < 2 ? 0 : ((radius * (int)min(255u * act / 150u, 255u)) >> 8) + 1;
    }
    // seeds: vectors of the 3x3 neighbours visited earlier, row-major (:574-584)
    int nseed_lane = 0;
    int sdx = 0, sdy = 0;
    uint32_t seed_word = 0;
    if (!kRefine && lane < 9 && npos < me) {
        seed_word = nvec;
        nseed_lane = 1; sdx = (int16_t)(seed_word & 0xFFFFu); sdy = (int16_t)(seed_word >> 16);
    }
    // the same neighbours and bad-block count as the last evaluation: the same result
    if (edge) rec = nullptr;
    if (!kRefine && rec && !fresh) {
        const bool differs = lane < 9 && nseed_lane && word != seed_word;  // word: the record's copy
        if (__ballot(differs) == 0 && in_word(kInRecValid) != 0 && (int)in_word(kInRecBad) == bad_count) {
            *unchanged = true;
            return DVec{};
        }
    }
    const uint64_t smask = __ballot(nseed_lane);
    const int nseeds = __popcll(smask);
    // seed j (the j-th present neighbour) to lane 4 + j; lanes without one send to lane 63
    const int sdst = nseed_lane ? 4 + lane_rank(smask) : kWave - 1;
    const int seed_x = fwd_permute(sdst, sdx), seed_y = fwd_permute(sdst, sdy);
    bool has_prev = false;
    int prx = 0, pry = 0;
    if (ch.prev) {
        if constexpr (kRefine) {
            const DVec pv = uniform_load(&ch.prev[idx]);
            has_prev = pv.score != 0xFFFFFFFFu && pv.score > 0;
            prx = pv.dx; pry = pv.dy;
        } else {
            const uint32_t v0 = in_word(kInPrev), sc = in_word(kInPrev + 1);
            has_prev = sc != 0xFFFFFFFFu && sc > 0;
            prx = (int16_t)(v0 & 0xFFFFu); pry = (int16_t)(v0 >> 16);
        }
    }

    // the search context; its predictor is the block's own (pred) except where the
    // refinement takes the previous-field vector
    Search cur;
    cur.pnew = kRefine ? rs.pnew : g.pnew;
    cur.chroma = !smallest;
    cur.serial = !kL0 && !kLean && cur.pnew < -256;
    cur.bnd[0] = -g.pel * px; cur.bnd[1] = -g.pel * py;
    cur.bnd[2] = g.pel * (g.lw - px - bw) + 1; cur.bnd[3] = g.pel * (g.lh - py - bh) + 1;
    cur.pmx = pred.dx; cur.pmy = pred.dy;
    cur.lambda = 0;

    // candidate list held in lanes (lane k = candidate k) for the list steps
    int lx = 0, ly = 0, lpen = 0, lact = 0;
    uint32_t lsad = 0;  // list candidate `lane`'s sad after evaluation
    int64_t lcost = 0;
    bool field_valid = false;
    auto set_pred_list = [&]() {
        // the reference's comparison order (:700-740): field, prev, global, zero, seeds
        field_valid = pred.score != 0xFFFFFFFFu;
        lact = 0; lx = 0; ly = 0; lpen = 0;
        if (lane == 0) { lact = field_valid; lx = pred.dx; ly = pred.dy; }
        else if (lane == 1) { lact = has_prev; lx = prx; ly = pry; lpen = g.pprev; }
        else if (lane == 2) { lact = 1; lx = gx; ly = gy; lpen = g.pglobal; }
        else if (lane == 3) { lact = 1; lpen = g.pzero; }
        else if (lane < 4 + nseeds) { lact = 1; lx = seed_x; ly = seed_y; lpen = g.pnbour; }
    };

    Best best{0, 0, 0xFFFFFFFFu, INT64_MAX};
    Pattern P;
    bool pchroma;
    int step;
    if (edge && !smallest) {
        b.px = epx; b.py = epy;
        lact = lane == 0; lx = pred.dx; ly = pred.dy; lpen = 0;
        P = Pattern{kPatPred, 0, 0, 0, 0, 1};
        pchroma = true;
        step = kStEdgeDone;
    } else {
        set_pred_list();
        // the refine pass measures only its start vector (lane 0's field entry)
        P = Pattern{kPatPred, 0, 0, 0, 0, kRefine ? 1 : 4 + nseeds};
        pchroma = !smallest;
        step = kStPredDone;
    }
    int bmx = 0, bmy = 0, dir = -2, hi = 0, odir = 0, hex_r = 0, exp_next = kStRefineDone, ring = 0;
    uint32_t found = 0;
    bool is_bad = false;
    // memo of the refinement (blocks off the edge row and column only)
    if (edge) memo = nullptr;
    bool memo_hit = false, provisional = false;
    int key_x = 0, key_y = 0;
    // ---------------- the evaluation site: the current list P, kPerPass candidates per
    // pass, one group each; returns the index of the list's winner (-1: none improved).
    // Kind is the list's kind as a compile-time constant (its branches fold), or -1 to
    // read it from P.
    auto eval_site = [&](auto kind_c) -> int {
        constexpr int kKind = decltype(kind_c)::value;
        int win = -1;
        {
            const int kind = kKind >= 0 ? kKind : (int)uniform_int(P.kind);
            // the pattern is wave-uniform: keep it in scalar registers so the pattern
            // branches below are scalar branches
            P = uniform_load(&P);
            const bool list = kind == kPatPred;
            // Long lists (UMH, exhaustive, wide rings) go in windows of 2 * kWave points:
            // one lane per point drops the points out of bounds or whose motion penalty
            // already reaches the best cost, and the survivors, in list order, fill the
            // passes densely.  The pass index below is then a survivor's rank, which
            // orders them as their list index does.
            const bool dense = !list && !cur.serial && P.n > kPerPass && kind != kPatHexExp;
            const int wlen = dense ? 2 * kWave : max(P.n, 1);
            for (int w0 = 0; w0 < P.n; w0 += wlen) {
                int cnt = P.n, n0 = 0;
                uint64_t m0 = 0, m1 = 0;  // survivors: points w0 + lane, w0 + kWave + lane
                if (dense) {
                    bool a0 = false, a1 = false;
                    int x, y;
                    if (w0 + lane < P.n) {
                        pat_gen_kind(kind, P, w0 + lane, x, y);
                        a0 = survives(cur, best.cost, x, y);
                    }
                    if (w0 + kWave + lane < P.n) {
                        pat_gen_kind(kind, P, w0 + kWave + lane, x, y);
                        a1 = survives(cur, best.cost, x, y);
                    }
                    m0 = __ballot(a0);
                    m1 = __ballot(a1);
                    n0 = __popcll(m0);
                    cnt = n0 + __popcll(m1);
                }
            for (int base = 0; base < cnt; base += kPerPass) {
                const int k = base + lane / kGroup;
                int mx = 0, my = 0, tpen = 0;
                int64_t motion = INT64_MAX;
                bool valid = k < cnt;
                if (list) {
                    // every lane takes part in the shuffles: a lane past the list end
                    // still holds a candidate another group reads
                    const int kl = min(k, kWave - 1);
                    mx = __shfl(lx, kl);
                    my = __shfl(ly, kl);
                    tpen = __shfl(lpen, kl);
                    const int tact = __shfl(lact, kl);
                    valid = valid && tact != 0;
                } else if (dense) {
                    // the list index of the survivor of rank base + g to lane g * kGroup, then
                    // to its whole group; the other lanes send to odd lanes, which the
                    // broadcast overwrites
                    const int r0 = lane_rank(m0), r1 = n0 + lane_rank(m1);
                    const bool a0 = (m0 >> lane) & 1, a1 = (m1 >> lane) & 1;
                    const int d0 = (a0 && r0 >= base && r0 < base + kPerPass) ? (r0 - base) * kGroup : (lane | 1);
                    const int d1 = (a1 && r1 >= base && r1 < base + kPerPass) ? (r1 - base) * kGroup : (lane | 1);
                    const int i0 = fwd_permute(d0, w0 + lane), i1 = fwd_permute(d1, w0 + kWave + lane);
                    const int li = (int)quad_bcast0((uint32_t)(k < n0 ? i0 : i1));
                    if (valid) {
                        pat_gen_kind(kind, P, li, mx, my);
                        // (inside the bounds) the best may have improved in an earlier pass
                        motion = motion_of(cur, mx, my);
                        valid = motion < best.cost;
                    }
                } else if (valid) {
                    pat_gen_kind(kind, P, k, mx, my);
                    valid = in_bounds(cur, mx, my);
                    if (valid) {
                        motion = motion_of(cur, mx, my);
                        // cost >= motion (pnew >= -256): a candidate whose motion penalty
                        // reaches the best so far cannot win, as in exact_check (:1122-1155)
                        if (!cur.serial) valid = motion < best.cost;
                    }
                }
                if (!list && __ballot(valid) == 0) continue;  // no candidate of this pass can win
                uint32_t L = 0, C = 0;
                bool off_plan = false;
                if (valid) {
                    if (plan_covers(tp, mx, my)) plan_cost(tp, mx, my, L, C);
                    else if constexpr (!kFast) group_cost(b, mx, my, pchroma, part, L, C);
                    else off_plan = true;
                }
                if constexpr (kFast) {
                    if (off_plan) valid = false;
                    if (__ballot(off_plan) != 0) provisional = true;
                }
                const uint32_t sad = sat_add(L, C);
                int64_t cost = INT64_MAX;
                if (valid) {
                    if (list) {
                        cost = (int64_t)sad + (((int64_t)tpen * (int64_t)sad) >> 8);
                    } else {
                        cost = motion + (int64_t)L + (((int64_t)cur.pnew * (int64_t)L) >> 8) + (int64_t)C +
                               (((int64_t)cur.pnew * (int64_t)C) >> 8);
                    }
                }
                if (list) {  // each list candidate's sad and cost back into its own lane
                    const int from = ((lane - base) & (kPerPass - 1)) * kGroup;
                    const uint32_t t_sad = __shfl(sad, from);
                    const int64_t t_cost = __shfl(cost, from);
                    if (lane >= base && lane < base + kPerPass) { lsad = t_sad; lcost = t_cost; }
                }
                if (!cur.serial || list) {
                    // lowest (cost, index) == the reference's running strict-`<` minimum;
                    // kPatHexExp: the hexagon's 6 points first, and the ring only when none
                    // of them improves (exact_hex2 then expands around the same centre)
                    const bool two = kind == kPatHexExp;
                    const bool in1 = valid && (!two || k < 6);
                    int64_t mc;
                    int mk;
                    if (best.cost < kPackedCost) {
                        // every improving cost is below 2^27: (cost, group) packed in 32 bits
                        const int g = lane / kGroup;
                        const bool imp = valid && cost < best.cost;
                        uint32_t key = wave_min_u32((imp && in1) ? ((uint32_t)cost << 4) | (uint32_t)g : ~0u);
                        if (two && key == ~0u) key = wave_min_u32((imp && k >= 6) ? ((uint32_t)cost << 4) | (uint32_t)g : ~0u);
                        mc = key == ~0u ? INT64_MAX : (int64_t)(key >> 4);
                        mk = key == ~0u ? 0x7FFFFFFF : base + (int)(key & 15u);
                    } else {
                        mc = in1 ? cost : INT64_MAX;
                        mk = in1 ? k : 0x7FFFFFFF;
                        group_argmin(mc, mk);
                        mc = rl64(mc, 0);
                        mk = rl(mk, 0);
                        if (two && !(mk != 0x7FFFFFFF && mc < best.cost)) {
                            const bool in2 = valid && k >= 6;
                            mc = in2 ? cost : INT64_MAX;
                            mk = in2 ? k : 0x7FFFFFFF;
                            group_argmin(mc, mk);
                            mc = rl64(mc, 0);
                            mk = rl(mk, 0);
                        }
                    }
                    if (mk != 0x7FFFFFFF && mc < best.cost) {
                        const int src = (mk - base) * kGroup;
                        best.mvx = rl(mx, src);
                        best.mvy = rl(my, src);
                        best.sad = rlu(sad, src);
                        best.cost = mc;
                        win = mk;
                    }
                } else {  // pnew < -256: exact_check's skip rules, candidate by candidate
                    const int cnt1 = min(kPerPass, P.n - base);
                    for (int j = 0; j < cnt1; j++) {
                        const int src = j * kGroup;
                        if (!rl((int)valid, src)) continue;
                        if (rl64(motion, src) >= best.cost) continue;
                        const int64_t cj = rl64(cost, src);
                        if (cj < best.cost) {
                            best.mvx = rl(mx, src);
                            best.mvy = rl(my, src);
                            best.sad = rlu(sad, src);
                            best.cost = cj;
                            win = base + j;
                        }
                    }
                }
            }
            }
        }
        return win;
    };
    constexpr bool kLine = kFast && (kL0 || kLean);
    if constexpr (kLine) {
        // The search's steps as straight-line code (the step machine below with its
        // transitions fixed: predictor list, then UMH's lists and hex2's, hex2's alone or
        // the exhaustive window, the memo, and at coarse levels the bad-block pass), each
        // list evaluated by the evaluation site specialised for its kind.  The tile-plan
        // instantiations only, without the serial replay (the general cost path,
        // specialised five times, would multiply the later sweeps' code).
        using std::integral_constant;
        // exact_hex2 around the best vector with radius r, closing with expanding(1, 1)
        auto hex2 = [&](int r) {
            bmx = best.mvx;
            bmy = best.mvy;
            if (r > 1) {
                P = Pattern{kPatHexExp, bmx, bmy, 1, 1, pat_count(kPatHexExp, 1, 1)};
                int w = eval_site(integral_constant<int, kPatHexExp>{});
                // a winner among the first 6 points moves the centre; otherwise the
                // closing expanding(1, 1) was this pass's second list
                if (w >= 0 && w < 6) {
                    dir = w;
                    bmx += kHex2[dir + 1][0];
                    bmy += kHex2[dir + 1][1];
                    hi = 1;
                    while (hi < r / 2 && in_bounds(cur, bmx, bmy)) {
                        odir = kMod6m1[dir + 1];
                        P = Pattern{kPatHexStep, bmx, bmy, 0, odir, 3};
                        w = eval_site(integral_constant<int, kPatHexStep>{});
                        if (w < 0) break;
                        dir = odir + w - 1;
                        bmx += kHex2[dir + 1][0];
                        bmy += kHex2[dir + 1][1];
                        hi++;
                    }
                    best.mvx = bmx;
                    best.mvy = bmy;
                    P = Pattern{kPatExpand, bmx, bmy, 1, 1, pat_count(kPatExpand, 1, 1)};
                    (void)eval_site(integral_constant<int, kPatExpand>{});
                }
            } else {
                P = Pattern{kPatExpand, bmx, bmy, 1, 1, pat_count(kPatExpand, 1, 1)};
                (void)eval_site(integral_constant<int, kPatExpand>{});
            }
        };
        (void)eval_site(integral_constant<int, kPatPred>{});
        cur.lambda = adapt_lambda(smallest ? 0 : g.lambda_base, g.lsad, pred.score);
        if (g.early) {  // level 0 with distance 0: the predictor (or zero) is the answer
            const int src = field_valid ? 0 : 3;
            best.mvx = rl(lx, src);
            best.mvy = rl(ly, src);
            best.sad = rlu(lsad, src);
            best.cost = (int64_t)best.sad;
            found = best.sad;
        } else {
            if (has_prev) {
                const int64_t fc = rl64(lcost, 0), pc = rl64(lcost, 1);
                if (pc < (field_valid ? fc : INT64_MAX)) { cur.pmx = prx; cur.pmy = pry; }
            }
            best.cost = (int64_t)best.sad;
            key_x = best.mvx;
            key_y = best.mvy;
            if (memo && !fresh) {  // (gathered with the block's inputs)
                const uint32_t k0 = in_word(kInMemo), r0 = in_word(kInMemo + 1);
                if (in_word(kInMemo + 3) != 0 && (int16_t)(k0 & 0xFFFFu) == key_x && (int16_t)(k0 >> 16) == key_y) {
                    best = Best{(int16_t)(r0 & 0xFFFFu), (int16_t)(r0 >> 16), in_word(kInMemo + 2),
                                join64(in_word(kInMemo + 4), in_word(kInMemo + 5))};
                    memo_hit = true;
                }
            }
            if (!memo_hit) {
                // (the edge blocks that reach here are the smallest level's)
                const int stype = (edge && level == 1) ? 0 : g.search_type;
                if (stype == 3) {
                    P = Pattern{kPatUmh, best.mvx, best.mvy, radius, 0, pat_count(kPatUmh, radius, 0)};
                    (void)eval_site(integral_constant<int, kPatUmh>{});
                }
                if (stype == 2 || stype == 3) {
                    hex_r = radius;
                    hex2(hex_r);
                } else if (stype == 4) {
                    P = Pattern{kPatExh, best.mvx, best.mvy, radius, 0, pat_count(kPatExh, radius, 0)};
                    (void)eval_site(integral_constant<int, kPatExh>{});
                }
            }
            if (memo && !memo_hit && !provisional && lane == 0)
                *memo = Memo{(int16_t)key_x, (int16_t)key_y, (int16_t)best.mvx, (int16_t)best.mvy, best.sad, 1u,
                             best.cost};
            found = best.sad;
            if constexpr (!kL0) {
                // bad-block expansion at coarse levels (:607-672), around the zero vector
                const int bad_limit = sat_add_i32(g.bad_sad, sat_mul_i32(g.bad_sad, bad_count) / 16);
                if (level > 0 && idx > 1 && g.bad_sad > 0 && (int)best.sad > bad_limit) {
                    is_bad = true;
                    cur.pmx = pred.dx;  // the block's own predictor
                    cur.pmy = pred.dy;
                    if (g.bad_range < 0) {  // rings of growing radius until one quarters the sad
                        for (ring = g.pel; ring < -g.bad_range * g.pel; ring += g.pel) {
                            P = Pattern{kPatExpand, 0, 0, ring, g.pel, pat_count(kPatExpand, ring, g.pel)};
                            (void)eval_site(integral_constant<int, kPatExpand>{});
                            if (best.sad < found / 4) break;
                        }
                    } else if (g.bad_range > 0) {
                        hex_r = g.bad_range * g.pel;
                        P = Pattern{kPatUmh, 0, 0, hex_r, 0, pat_count(kPatUmh, hex_r, 0)};
                        (void)eval_site(integral_constant<int, kPatUmh>{});
                        hex2(hex_r);
                    }
                    P = Pattern{kPatExpand, best.mvx, best.mvy, 1, 1, pat_count(kPatExpand, 1, 1)};
                    (void)eval_site(integral_constant<int, kPatExpand>{});
                }
            }
        }
    }
    for (bool generic = !kLine; generic;) {
        // ---------------- the evaluation site: kPerPass candidates per pass, one group each
        const int win = eval_site(std::integral_constant<int, -1>{});
        // ---------------- act on the result; pick the next list (or finish)
        bool next = false;
        while (!next && step != kStEnd) {
            switch (step) {
                case kStEdgeDone:
                    pred.score = rlu(lsad, 0);
                    b.px = px; b.py = py;
                    make_plan(b, !smallest, part, tp);
                    set_pred_list();
                    best = Best{0, 0, 0xFFFFFFFFu, INT64_MAX};
                    P = Pattern{kPatPred, 0, 0, 0, 0, 4 + nseeds};
                    pchroma = !smallest;
                    step = kStPredDone;
                    next = true;
                    break;
                case kStPredDone: {
                    cur.lambda = kRefine ? rs.lambda : adapt_lambda(smallest ? 0 : g.lambda_base, g.lsad, pred.score);
                    if (!kRefine && g.early) {
                        // level 0 with distance 0: the predictor (or zero) is the answer
                        const int src = field_valid ? 0 : 3;
                        best.mvx = rl(lx, src);
                        best.mvy = rl(ly, src);
                        best.sad = rlu(lsad, src);
                        best.cost = (int64_t)best.sad;
                        found = best.sad;
                        step = kStEnd;
                        break;
                    }
                    if (has_prev) {
                        const int64_t fc = rl64(lcost, 0), pc = rl64(lcost, 1);
                        if (pc < (field_valid ? fc : INT64_MAX)) { cur.pmx = prx; cur.pmy = pry; }
                    }
                    best.cost = (int64_t)best.sad;
                    key_x = best.mvx;
                    key_y = best.mvy;
                    if (memo && !fresh) {  // (gathered with the block's inputs)
                        const uint32_t k0 = in_word(kInMemo), r0 = in_word(kInMemo + 1);
                        if (in_word(kInMemo + 3) != 0 && (int16_t)(k0 & 0xFFFFu) == key_x &&
                            (int16_t)(k0 >> 16) == key_y) {
                            best = Best{(int16_t)(r0 & 0xFFFFu), (int16_t)(r0 >> 16), in_word(kInMemo + 2),
                                        join64(in_word(kInMemo + 4), in_word(kInMemo + 5))};
                            memo_hit = true;
                            step = kStRefineDone;
                            break;
                        }
                    }
                    const int stype = kRefine ? rs.type : ((edge && level == 1) ? 0 : g.search_type);
                    exp_next = kStRefineDone;
                    if (stype == 2) {
                        hex_r = radius;
                        step = kStHexBegin;
                    } else if (stype == 3) {
                        hex_r = radius;
                        P = Pattern{kPatUmh, best.mvx, best.mvy, radius, 0, pat_count(kPatUmh, radius, 0)};
                        step = kStUmhDone;
                        next = true;
                    } else if (stype == 4) {
                        P = Pattern{kPatExh, best.mvx, best.mvy, radius, 0, pat_count(kPatExh, radius, 0)};
                        step = kStExhDone;
                        next = true;
                    } else {
                        step = kStRefineDone;
                    }
                    break;
                }
                case kStUmhDone:
                    step = kStHexBegin;
                    break;
                case kStHexBegin:
                    bmx = best.mvx;
                    bmy = best.mvy;
                    if (hex_r > 1) {
                        dir = -2;
                        // without a winner among the 6 points exact_hex2 expands around the
                        // same centre: both lists in one pass (not in the serial replay)
                        P = cur.serial ? Pattern{kPatHexFirst, bmx, bmy, 0, 0, 6}
                                       : Pattern{kPatHexExp, bmx, bmy, 1, 1, pat_count(kPatHexExp, 1, 1)};
                        step = kStHexFirstDone;
                    } else {
                        P = Pattern{kPatExpand, bmx, bmy, 1, 1, pat_count(kPatExpand, 1, 1)};
                        step = kStExp1Done;
                    }
                    next = true;
                    break;
                case kStHexFirstDone:
                    if (win >= 0 && win < 6) {
                        dir = win;
                        bmx += kHex2[dir + 1][0];
                        bmy += kHex2[dir + 1][1];
                        hi = 1;
                        step = kStHexLoop;
                    } else if (P.kind == kPatHexExp) {
                        // expanding(1, 1) around (bmx, bmy) was evaluated in the same pass,
                        // and `best` holds its result
                        step = kStExp1Done;
                    } else {
                        best.mvx = bmx;
                        best.mvy = bmy;
                        P = Pattern{kPatExpand, bmx, bmy, 1, 1, pat_count(kPatExpand, 1, 1)};
                        step = kStExp1Done;
                        next = true;
                    }
                    break;
                case kStHexLoop:
                    if (hi < hex_r / 2 && in_bounds(cur, bmx, bmy)) {
                        odir = kMod6m1[dir + 1];
                        dir = -2;
                        P = Pattern{kPatHexStep, bmx, bmy, 0, odir, 3};
                        step = kStHexStepDone;
                    } else {
                        best.mvx = bmx;
                        best.mvy = bmy;
                        P = Pattern{kPatExpand, bmx, bmy, 1, 1, pat_count(kPatExpand, 1, 1)};
                        step = kStExp1Done;
                    }
                    next = true;
                    break;
                case kStHexStepDone:
                    if (win >= 0) {
                        dir = odir + win - 1;
                        bmx += kHex2[dir + 1][0];
                        bmy += kHex2[dir + 1][1];
                        hi++;
                        step = kStHexLoop;
                    } else {
                        best.mvx = bmx;
                        best.mvy = bmy;
                        P = Pattern{kPatExpand, bmx, bmy, 1, 1, pat_count(kPatExpand, 1, 1)};
                        step = kStExp1Done;
                        next = true;
                    }
                    break;
                case kStExp1Done:
                    step = exp_next;
                    break;
                case kStExhDone:
                    step = kStRefineDone;
                    break;
                case kStRefineDone: {
                    if (memo && !memo_hit && !provisional && lane == 0)
                        *memo = Memo{(int16_t)key_x, (int16_t)key_y, (int16_t)best.mvx, (int16_t)best.mvy, best.sad, 1u,
                                     best.cost};
                    // bad-block expansion at coarse levels (:607-672), around the zero vector
                    found = best.sad;
                    const int bad_limit = sat_add_i32(g.bad_sad, sat_mul_i32(g.bad_sad, bad_count) / 16);
                    if (level > 0 && idx > 1 && g.bad_sad > 0 && (int)best.sad > bad_limit) {
                        is_bad = true;
                        cur.pmx = pred.dx;  // the block's own predictor
                        cur.pmy = pred.dy;
                        if (g.bad_range < 0) {
                            ring = g.pel;
                            step = kStBadRing;
                        } else if (g.bad_range > 0) {
                            hex_r = g.bad_range * g.pel;
                            exp_next = kStBadFinal;
                            P = Pattern{kPatUmh, 0, 0, hex_r, 0, pat_count(kPatUmh, hex_r, 0)};
                            step = kStUmhDone;
                            next = true;
                        } else {
                            step = kStBadFinal;
                        }
                    } else {
                        step = kStEnd;
                    }
                    break;
                }
                case kStBadRing:
                    if (ring < -g.bad_range * g.pel) {
                        P = Pattern{kPatExpand, 0, 0, ring, g.pel, pat_count(kPatExpand, ring, g.pel)};
                        step = kStBadRingDone;
                        next = true;
                    } else {
                        step = kStBadFinal;
                    }
                    break;
                case kStBadRingDone:
                    if (best.sad < found / 4) {
                        step = kStBadFinal;
                    } else {
                        ring += g.pel;
                        step = kStBadRing;
                    }
                    break;
                case kStBadFinal:
                    P = Pattern{kPatExpand, best.mvx, best.mvy, 1, 1, pat_count(kPatExpand, 1, 1)};
                    step = kStBadFinalDone;
                    next = true;
                    break;
                case kStBadFinalDone:
                default:
                    step = kStEnd;
                    break;
            }
        }
        if (!next) break;
    }
    *bad = is_bad;
    *found_out = found;
    if (provisional) {
        *defer = kDeferExact;
        if (rec && lane == 0) rec->valid = 0;
    } else if (rec) {
        if (lane < 9) rec->seed[lane] = seed_word;
        if (lane == 0) {
            rec->bad_count = bad_count;
            rec->valid = 1;
        }
    }
    DVec o;
    o.dx = (int16_t)best.mvx;
    o.dy = (int16_t)best.mvy;
    o.score = sat_mul(best.sad, kRefine ? (uint32_t)rs.factor : (smallest ? 3u : 1u));
    if (tp.on && tp.dc) {
        // the plan's source tiles are the block's pixels, all inside the level (the
        // N x N mean of block_luma_dc)
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) sum = __builtin_amdgcn_sad_u8(tp.lsrc[k], 0u, sum);
        sum = group_sum(sum);
        o.luma = (uint8_t)(sum >> (tp.n == 8 ? 6 : 4));
    } else if constexpr (!kFast) {  // (kFast: every block it finishes has a plan with dc)
        o.luma = luma_dc(b);
    }
    o.done = 1;
    return o;
}

// exact_search_level (:480-681) in the reference's visit order: one wavefront per
// chain walks the whole level, keeping the bad-block count in a register.  The
// fallback for a level whose parallel sweeps have not converged within the sweep
// limit (svpf_analyse_set_max_sweeps); it runs the general search (2 waves per SIMD).
__global__ void __launch_bounds__(kWave, 2) search_level_kernel(const Chain* __restrict__ chains) {
    const Chain ch = uniform_load(&chains[blockIdx.x]);
    const LevelGeom& g = ch.geom;
    Blk b;
    init_blk(ch, g, b);
    int bad_count = 0;
    const int nblocks = g.gw * g.gh;
    for (int it = 0; it < nblocks; it++) {
        const int packed = uniform_int(ch.order[it]);
        const int bx = packed & 0xFFFF, by = packed >> 16;
        if (bx >= g.gw || by >= g.gh) continue;  // defensive: orders always hold in-grid positions
        const int idx = by * g.gw + bx;
        bool bad;
        uint32_t found;
        const DVec o = search_block<false>(ch, g, b, idx, bad_count, &bad, &found, nullptr, nullptr, nullptr,
                                           RefineStart{});
        bad_count += bad ? 1 : 0;
        if (threadIdx.x == 0) ch.res[idx] = o;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
}

// Per level: visit positions from the order, results initialised to the predictors,
// sweep flags cleared, sweep counters zeroed.
__global__ void sweep_init_kernel(const Chain* __restrict__ chains, SweepState st) {
    const Chain ch = chains[blockIdx.y];
    const LevelGeom& g = ch.geom;
    const int n = g.gw * g.gh;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int packed = ch.order[i];
        const int bx = packed & 0xFFFF, by = packed >> 16;
        if (bx < g.gw && by < g.gh) ch.opos[by * g.gw + bx] = i;
        ch.res[i] = ch.in[i];
        st.flags[0][ch.base + i] = 0;
        st.flags[1][ch.base + i] = 0;
        st.found[ch.base + i] = 0;
        st.badf[ch.base + i] = 0;
        st.cpre[ch.base + i] = 1 << 30;  // sweep 0 runs no bad-block pass: the limit saturates
        st.fseq[ch.base + i] = 0;         // (position i: the same three words)
        st.bseq[ch.base + i] = 0;
        st.cseq[ch.base + i] = 1 << 30;
        st.memo[ch.base + i].valid = 0;
        st.seeds[ch.base + i].valid = 0;
        st.stamp[ch.base + i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.minpos[blockIdx.y] = 0;  // the first rebuild covers the whole order
    }
    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (int k = threadIdx.x; k < kCntWords; k += blockDim.x) {
            st.counts[k] = 0;
            st.dcount[k] = 0;
        }
}

// Block n of chain ci into sweep s + 1's list wl[q], once (flags[q]), for every calling
// lane whose `want` is set.  The list length is one counter for the whole level, so the
// lanes that win their block's flag take consecutive slots with one counter update per
// wave (the same counter updated by every block would serialise the queueing).
__device__ __forceinline__ void queue_wave(const SweepState& st, int ci, const Chain& ch, int n, bool want, int s) {
    const int q = (s + 1) & 1;
    const bool take = want && atomicExch(&st.flags[q][ch.base + n], 1u) == 0u;
    const uint64_t m = __ballot(take);
    if (m == 0) return;
    const int leader = __builtin_amdgcn_readfirstlane(__builtin_ctzll(m));
    uint32_t first = 0;
    if (lane_id() == leader) first = atomicAdd(&st.counts[(s + 1) % 3], (uint32_t)__popcll(m));
    first = (uint32_t)__builtin_amdgcn_readlane((int)first, leader);
    if (take) st.wl[q][first + lane_rank(m)] = ((unsigned long long)ci << 32) | (unsigned)n;
}

constexpr int kScanThreads = 1024;  // threads of a workgroup that rebuilds one chain's count
constexpr int kScanRegs = 32;       // a chunk of at most this many positions is held in registers

// Exclusive prefix sum of v over the workgroup's threads (wsum: one int per wave).
__device__ __forceinline__ int block_excl_scan(int v, int* wsum) {
    const int lane = lane_id(), w = (int)(threadIdx.x / kWave);
    int x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    int base = 0;
    for (int j = 0; j < w; j++) base += wsum[j];
    __syncthreads();
    return base + x - v;
}

// The running bad-block count after positions [lo, hi) of the visit order, from input
// c.  fseq holds each position's pre-expansion cost, 0 for blocks that never count
// (the first two raster blocks): with bad_sad > 0 the limit is at least bad_sad.  The
// costs are loaded kScanBatch at a time (independent loads, one wait per batch).
constexpr int kScanBatch = 8;
__device__ __forceinline__ int bad_limit(int bad_sad, int c) { return sat_add_i32(bad_sad, sat_mul_i32(bad_sad, c) / 16); }
__device__ __forceinline__ int scan_walk(const uint32_t* fseq, int lo, int hi, int bad_sad, int c) {
    for (int k0 = lo; k0 < hi; k0 += kScanBatch) {
        uint32_t f[kScanBatch];
#pragma unroll
        for (int j = 0; j < kScanBatch; j++) f[j] = k0 + j < hi ? fseq[k0 + j] : 0u;
#pragma unroll
        for (int j = 0; j < kScanBatch; j++)
            if ((int)f[j] > bad_limit(bad_sad, c)) c++;
    }
    return c;
}

// Rebuild, after sweep s, of a chain's running bad-block count over the visit order
// (coarse levels with the bad-block pass, :607-611): from the first position whose
// found/bad changed, cpre is recomputed and every block whose bad flag no longer
// matches its count is queued.  Run by a whole workgroup of kScanThreads threads
// (sh: kScanThreads + kScanThreads / kWave ints of LDS).
//
// The recurrence c' = c + [f > limit(c)] is split into one chunk of positions per
// thread.  A chunk's input count is found by iterating
//     in[t] = in[0] + sum over u < t of (walk(u, in[u]) - in[u])
// from the previous rebuild's counts: chunk 0's input is exact, so chunk t is exact
// after t iterations and a repeat with no change is the sequential result (usually
// after one or two iterations, since a sweep changes few costs).  Every read is of the
// visit-order copies; a block's raster index is looked up only where its count or its
// bad flag changes.
__device__ void scan_chain(const Chain* __restrict__ chains, int ci, const SweepState& st, int s,
                           int* sh) {
    const Chain ch = uniform_load(&chains[ci]);
    const LevelGeom& g = ch.geom;
    if (g.level == 0 || g.bad_sad <= 0) return;
    const uint32_t start = __hip_atomic_load(&st.minpos[ci], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();  // every thread has read minpos before it is cleared below
    if (start == 0xFFFFFFFFu) return;
    const int T = (int)blockDim.x, tid = (int)threadIdx.x;
    const int n = g.gw * g.gh;
    const int m = (n - (int)start + T - 1) / T;
    const int lo = min((int)start + tid * m, n), hi = min(lo + m, n);
    const uint32_t* fs = st.fseq + ch.base;
    const uint32_t* bs = st.bseq + ch.base;
    int32_t* cs = st.cseq + ch.base;
    const int c0 = start > 0 ? cs[start] : 0;
    int guess = c0;
    if (tid > 0 && lo < hi) {
        const int old = cs[lo];
        guess = old < (1 << 30) ? old : c0;  // before the first rebuild of the level: no guess
    }
    int* wsum = sh + T;
    int in = guess;
    // Chunks of at most kScanRegs positions (every level of up to kScanThreads *
    // kScanRegs blocks) are walked from registers, loaded once; longer ones from memory.
    // (m is the same for every thread, so is the branch.)
    if (m <= kScanRegs) {
        uint32_t fr[kScanRegs];
#pragma unroll
        for (int j = 0; j < kScanRegs; j++) fr[j] = lo + j < hi ? fs[lo + j] : 0u;  // 0: never counts
        auto walk = [&](int c) {
#pragma unroll
            for (int j = 0; j < kScanRegs; j++)
                if ((int)fr[j] > bad_limit(g.bad_sad, c)) c++;
            return c;
        };
        for (int it = 0; it <= T; it++) {  // exact after T iterations at the latest
            const int d = walk(in) - in;
            const int v = c0 + block_excl_scan(d, wsum);
            const int changed = v != in;
            in = v;
            if (!__syncthreads_or(changed)) break;
        }
        int c = in;
#pragma unroll
        for (int j0 = 0; j0 < kScanRegs; j0 += kScanBatch) {
            const int k0 = lo + j0;
            if (k0 >= hi) break;
            uint32_t b[kScanBatch];
            int cc[kScanBatch];
#pragma unroll
            for (int j = 0; j < kScanBatch; j++) {
                const bool v = k0 + j < hi;
                b[j] = v ? bs[k0 + j] : 0u;
                cc[j] = v ? cs[k0 + j] : c;
            }
#pragma unroll
            for (int j = 0; j < kScanBatch; j++) {
                const int k = k0 + j;
                const bool v = k < hi;
                const int bj = (int)fr[j0 + j] > bad_limit(g.bad_sad, c) ? 1 : 0;
                const bool cchg = v && cc[j] != c, want = v && (uint32_t)bj != b[j];
                int idx = 0;
                if (cchg || want) {
                    const int pk = ch.order[k];
                    idx = (pk >> 16) * g.gw + (pk & 0xFFFF);
                }
                if (cchg) {
                    st.cpre[ch.base + idx] = c;
                    cs[k] = c;
                }
                queue_wave(st, ci, ch, idx, want, s);
                if (v) c += bj;
            }
        }
        if (tid == 0) st.minpos[ci] = 0xFFFFFFFFu;
        return;
    }
    for (int it = 0; it <= T; it++) {  // exact after T iterations at the latest
        const int d = scan_walk(fs, lo, hi, g.bad_sad, in) - in;
        const int v = c0 + block_excl_scan(d, wsum);
        const int changed = v != in;
        in = v;
        if (!__syncthreads_or(changed)) break;
    }
    int c = in;
    for (int k0 = lo; k0 < hi; k0 += kScanBatch) {
        uint32_t f[kScanBatch], b[kScanBatch];
        int cc[kScanBatch];
#pragma unroll
        for (int j = 0; j < kScanBatch; j++) {
            const bool v = k0 + j < hi;
            f[j] = v ? fs[k0 + j] : 0u;
            b[j] = v ? bs[k0 + j] : 0u;
            cc[j] = v ? cs[k0 + j] : c;
        }
#pragma unroll
        for (int j = 0; j < kScanBatch; j++) {
            const int k = k0 + j;
            const bool v = k < hi;
            const int bj = (int)f[j] > bad_limit(g.bad_sad, c) ? 1 : 0;
            const bool cchg = v && cc[j] != c, want = v && (uint32_t)bj != b[j];
            int idx = 0;
            if (cchg || want) {
                const int pk = ch.order[k];
                idx = (pk >> 16) * g.gw + (pk & 0xFFFF);
            }
            if (cchg) {
                st.cpre[ch.base + idx] = c;
                cs[k] = c;
            }
            queue_wave(st, ci, ch, idx, want, s);
            if (v) c += bj;
        }
    }
    if (tid == 0) st.minpos[ci] = 0xFFFFFFFFu;
}

__global__ void __launch_bounds__(kScanThreads) sweep_scan_kernel(const Chain* __restrict__ chains, SweepState st,
                                                                   int s) {
    __shared__ int sh[kScanThreads + kScanThreads / kWave];
    scan_chain(chains, blockIdx.x, st, s, sh);
}

// Evaluates block idx of chain ci in sweep s; when its vector (or score) changes, the
// result is stored and every 3x3 neighbour visited after it is queued for sweep s + 1.
// kFast: the tile-plan evaluation; a block it passes on (whole, or with a provisional
// result stored as any other) goes to the list of sweep s that deferred_kernel
// completes before sweep s + 1.
// Returns whether the block's stored vector changed; *me_out = its visit position.
template <bool kFast, bool kL0 = false, bool kLean = false>
__device__ __forceinline__ bool sweep_item(const Chain* __restrict__ chains, int ci, int idx, const SweepState& st,
                                           int s, int* me_out = nullptr, int clear_parity = -1) {
    // the descriptor read in place through the constant address space: scalar loads the
    // compiler may repeat instead of keeping (and spilling) every field
    const Chain& ch = *(const Chain*)((const_chain*)chains + ci);
    const LevelGeom& g = ch.geom;
    // the block leaves list clear_parity (it may be queued there again from sweep s + 1 on)
    if (clear_parity >= 0 && lane_id() == 0) st.flags[clear_parity][ch.base + idx] = 0;
    Blk b;
    init_blk(ch, g, b);
    bool bad;
    uint32_t found;
    bool unchanged = false;
    int defer = kDeferNone;
    BlockIO io;
    io.cpre = st.cpre + ch.base + idx;
    io.found = st.found + ch.base + idx;
    io.badf = st.badf + ch.base + idx;
    io.fresh = s == 0;  // sweep_init cleared every memo and seed record of the level
    const DVec o = search_block<false, kFast, kL0, kLean>(ch, g, b, idx, 0, &bad, &found, st.memo + ch.base + idx,
                                              st.seeds + ch.base + idx, &unchanged, RefineStart{}, &defer, &io);
    if (kFast && defer != kDeferNone) {
        if (lane_id() == 0) st.dl[atomicAdd(&st.dcount[kDcDeferred], 1u)] = ((unsigned long long)ci << 32) | (unsigned)idx;
        // in a later sweep a provisional result is not stored either: storing it would queue
        // the block's neighbours twice (now and when the deferred search changes it again)
        if (defer == kDeferWhole || s > 0) return false;
    }
    if (unchanged) {  // nothing it reads changed since its last evaluation
        if (lane_id() == 0) atomicAdd(&st.counts[kCntSkipped], 1u);  // skip count (statistics)
        return false;
    }
    // (the block's own words: no other evaluation stores them while this one runs)
    const DVec old = io.old;
    const bool moved = o.dx != old.dx || o.dy != old.dy || o.score != old.score;
    const uint32_t oldf = io.oldf, oldb = io.oldb;
    const int me = io.me;
    if (me_out) *me_out = me;
    const int lane = lane_id();
    if (lane == 0) {
        if (moved || o.luma != old.luma || o.done != old.done) ch.res[idx] = o;
        if (found != oldf || (uint32_t)bad != oldb) {
            st.found[ch.base + idx] = found;
            st.badf[ch.base + idx] = bad ? 1u : 0u;
            if (!kL0 && g.level > 0 && g.bad_sad > 0) {  // the count rebuild's copies, by visit position
                st.fseq[ch.base + me] = idx > 1 ? found : 0u;
                st.bseq[ch.base + me] = bad ? 1u : 0u;
                atomicMin(&st.minpos[ci], (uint32_t)me);
            }
        }
    }
    // sweep 0 queues nothing: validate_kernel compares every block's seed record with
    // its neighbours once the sweep is complete
    if (!moved || s == 0) return moved;
    const int bx = idx % g.gw, by = idx / g.gw;
    const int nx = bx - 1 + lane % 3, ny = by - 1 + lane / 3;
    const bool in = lane < 9 && lane != 4 && nx >= 0 && ny >= 0 && nx < g.gw && ny < g.gh;
    const int n = in ? ny * g.gw + nx : idx;
    queue_wave(st, ci, ch, n, in && ch.opos[n] > me, s);
    return true;
}

// Position pos of the finest level's visit order, initial_search_order (:302-324): the
// interior rows 1 .. gh-2 as a meander (even rows left to right), then the top and
// bottom rows column by column, then the left and right columns row by row.
__device__ __forceinline__ void level0_order(int pos, int gw, int gh, int& x, int& y) {
    const int iw = gw - 2, ni = iw * (gh - 2);
    if (pos < ni) {
        const int r = pos / iw, t = pos - r * iw;
        y = 1 + r;
        x = (y & 1) == 0 ? 1 + t : gw - 2 - t;
    } else if (pos < ni + 2 * gw) {
        const int j = pos - ni;
        x = j >> 1;
        y = (j & 1) ? gh - 1 : 0;
    } else {
        const int j = pos - ni - 2 * gw;
        y = 1 + (j >> 1);
        x = (j & 1) ? gw - 1 : 0;
    }
}

// Sweep 0: every block of every chain, dispatched in visit order.  Each chain's visit
// order is cut into `nseg` segments, and the (chain, segment) slots v = ci * nseg + seg
// are interleaved: item u = k * nslots + v takes position k of slot v.  With nslots a
// multiple of 8 (and the grid too), the hardware's round-robin of workgroups over the
// 8 XCDs sends every item of a slot to the same XCD, whose L2 then holds the results
// its earlier items stored in place.  So a block dispatched after its earlier-visited
// neighbours have finished already seeds from their searched vectors, and
// validate_kernel finds its seed record up to date.
// kFast: the tile-plan instantiation (three waves per SIMD), which passes the blocks it
// cannot finish to deferred_kernel.  run: positions per item, searched in order by one
// wave.  The coarse levels' order lists each sorted parent's children together, so with
// runs of 8 or more a block's earlier siblings are mostly stored before it reads them;
// at the finest level consecutive meander positions keep the reference rows of a
// block's window in the caches for the next.
template <bool kFast, bool kL0 = false, bool kLean = false>
__global__ void __launch_bounds__(kWave, kFast ? (kLean ? 5 : 3) : 2) sweep0_kernel(const Chain* __restrict__ chains,
                                                                     SweepState st, int nslots, int nseg,
                                                                     int run, long long total) {
    for (long long u = blockIdx.x; u < total; u += gridDim.x) {
        const int v = (int)(u % nslots), k0 = (int)(u / nslots);
        const int ci = v / nseg, seg = v - ci * nseg;
        const LevelGeom* g = &chains[ci].geom;
        const int gw = g->gw, gh = g->gh;
        const int n = gw * gh, len = (n + nseg - 1) / nseg;
        for (int j = 0; j < run; j++) {
            const int k = k0 * run + j;
            const int pos = seg * len + k;
            if (k >= len || pos >= n) break;
            int bx, by;
            if (g->level == 0) {
                level0_order(pos, gw, gh, bx, by);
            } else {
                const int pk = chains[ci].order[pos];
                bx = pk & 0xFFFF;
                by = pk >> 16;
            }
            if (bx >= gw || by >= gh) continue;
            sweep_item<kFast, kL0, kLean>(chains, ci, by * g->gw + bx, state_here(), 0);
            // the next position's seed reads see this block's result
            if (run > 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
    }
}

// After sweep 0: queues for sweep 1 every block that must be evaluated again -- one
// whose recorded neighbour vectors differ from the neighbours' results now (every
// other input of a block off the edge row and column is fixed for the level; a changed
// bad-block count is handled by the count rebuild), and every edge block (no record).
// One thread per block (grid y: the chain), blocks in raster order; a thread issues all its loads at once
// (the record as three 16-byte words, each neighbour's visit position and vector).
__global__ void __launch_bounds__(256) validate_kernel(const Chain* __restrict__ chains, SweepState st) {
    const int ci = blockIdx.y, idx0 = blockIdx.x * blockDim.x + threadIdx.x;  // (chain, raster index)
    const Chain& ch = chains[ci];
    const LevelGeom& g = ch.geom;
    const int gw = g.gw, gh = g.gh;
    const bool here = idx0 < gw * gh;
    const int idx = here ? idx0 : 0;
    const int bx = idx % gw, by = idx / gw;
    const bool edge = g.level > 0 && (bx == gw - 1 || by == gh - 1);
    const uint4* rw = reinterpret_cast<const uint4*>(&st.seeds[ch.base + idx]);
    const uint4 r0 = rw[0], r1 = rw[1], r2 = rw[2];
    const uint32_t seed[9] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x};
    const int me = ch.opos[idx];
    int pos[9];
    uint32_t vec[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const int nx = bx - 1 + q % 3, ny = by - 1 + q / 3;
        const bool in = q != 4 && nx >= 0 && ny >= 0 && nx < gw && ny < gh;
        const int n = in ? ny * gw + nx : idx;
        pos[q] = in ? ch.opos[n] : INT32_MAX;
        vec[q] = *reinterpret_cast<const uint32_t*>(&ch.res[n]);
    }
    bool stale = edge || r2.z == 0;  // no record at the edge; SeedRec::valid
#pragma unroll
    for (int q = 0; q < 9; q++) stale = stale || (pos[q] < me && vec[q] != seed[q]);
    queue_wave(st, ci, ch, idx, here && stale, 0);
}

// Claims block idx for an evaluation in sweep s (s >= 1): false when another wave of
// this sweep already has it.  Two waves never evaluate one block in the same sweep.
__device__ __forceinline__ bool claim_block(const SweepState& st, uint32_t base, int idx, int s) {
    uint32_t old = 0;
    if (lane_id() == 0) old = atomicMax(&st.stamp[base + idx], (uint32_t)s);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)old) < (uint32_t)s;
}

// The block after visit position me, when it is a 3x3 neighbour of block idx (the one
// block whose inputs a change of idx alters and that comes right after it); else -1.
__device__ __forceinline__ int order_successor(const Chain* __restrict__ chains, int ci, int idx, int me) {
    const LevelGeom* g = &chains[ci].geom;
    const int gw = g->gw, gh = g->gh;
    if (me + 1 >= gw * gh) return -1;
    const int pk = uniform_int(chains[ci].order[me + 1]);
    const int bx = pk & 0xFFFF, by = pk >> 16;
    const int cx = idx % gw, cy = idx / gw;
    if (bx >= gw || by >= gh || abs(bx - cx) > 1 || abs(by - cy) > 1) return -1;
    return by * gw + bx;
}

// Sweep s >= 1 over the blocks queued by sweep s - 1.  A block whose vector changed has
// its later neighbours queued for sweep s + 1; the one visited right after it (its
// order successor) is also evaluated at once by the same wave, and so on along the
// order while vectors change, up to st.chase blocks.  A change then travels along the
// visit order within one sweep instead of one block per sweep.  The fixed point is the
// same: a chased block read its predecessors' latest values, and any predecessor that
// changes after that queues it for the next sweep.  (The lists are short after the first
// few sweeps and each sweep waits on its slowest wave, so a second, deferred pass per
// sweep would cost more than three waves per SIMD gain.)
//
// kFast (with kL0 or kLean): the straight-line tile-plan instantiation, for the long
// lists of the first sweeps after sweep 0; a block it cannot finish stores nothing and
// goes to the deferred list, which deferred_kernel completes before sweep s + 1 (the
// host clears that list's length first).  No chase.
template <bool kL0, bool kFast = false, bool kLean = false>
__global__ void __launch_bounds__(kWave, kFast ? (kLean ? 5 : 3) : 2) sweep_kernel(const Chain* __restrict__ chains, SweepState st,
                                                                     int s) {
    const int p = s & 1;
    // this workgroup's first item, loaded with the list length rather than after it (the
    // grid is at most one workgroup per block of the level, so the slot exists)
    const unsigned long long first = st.wl[p][blockIdx.x];
    const uint32_t n = (uint32_t)uniform_int((int)__atomic_load_n(&st.counts[s % 3], __ATOMIC_RELAXED));
    const int chase = st.chase;
    if (blockIdx.x == 0 && lane_id() == 0) {
        st.counts[(s + 2) % 3] = 0;  // read by sweep s - 1; sweep s + 1 queues into it
        atomicAdd(&st.counts[kCntQueued], n);
    }
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const unsigned long long item = i == blockIdx.x ? first : st.wl[p][i];
        // uniform across the wave: keeps the chain's descriptors in scalar registers
        const int ci = __builtin_amdgcn_readfirstlane((int)(item >> 32));
        int idx = __builtin_amdgcn_readfirstlane((int)(item & 0xFFFFFFFFu));
        if (kFast || chase == 0) {  // every queued block once (its flag), no claim needed
            sweep_item<kFast, kL0, kLean>(chains, ci, idx, state_here(), s, nullptr, p);
            continue;
        }
        const uint32_t base = (uint32_t)uniform_int((int)chains[ci].base);
        if (lane_id() == 0) st.flags[p][base + idx] = 0;
        if (!claim_block(st, base, idx, s)) {  // a chase of this sweep evaluated it
            if (lane_id() == 0) atomicAdd(&st.counts[kCntSkipped], 1u);
            continue;
        }
        for (int k = 0;; k++) {
            int me = 0;
            if (!sweep_item<false, kL0>(chains, ci, idx, state_here(), s, &me) || k >= chase) break;
            const int next = order_successor(chains, ci, idx, me);
            if (next < 0 || !claim_block(st, base, next, s)) break;
            // this wave's store of idx before the successor reads it (the same wave: a
            // workgroup-scope fence orders it; waves elsewhere that read the old value are
            // queued again by idx's change)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane_id() == 0) atomicAdd(&st.dcount[kDcChased], 1u);  // chased (statistics)
            idx = next;
        }
    }
}

// The blocks of sweep s that the tile-plan instantiation passed on, evaluated with the
// general cost path before sweep s + 1 starts.  A block passed on whole stored nothing;
// one with a provisional result has no valid seed record, so it is searched in full,
// and a result that differs from the provisional one is stored and propagated as any
// change is.  So the sweep's invariant -- every block whose inputs changed is
// evaluated again -- holds as for the others.
// kL0: the finest level's instantiation (its general search fits three waves per SIMD).
template <bool kL0>
__global__ void __launch_bounds__(kWave, 2) deferred_kernel(const Chain* __restrict__ chains, SweepState st,
                                                            int s) {
    const uint32_t n = (uint32_t)uniform_int((int)__atomic_load_n(&st.dcount[kDcDeferred], __ATOMIC_RELAXED));
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const unsigned long long item = st.dl[i];
        const int ci = __builtin_amdgcn_readfirstlane((int)(item >> 32));
        const int idx = __builtin_amdgcn_readfirstlane((int)(item & 0xFFFFFFFFu));
        sweep_item<false, kL0>(chains, ci, idx, state_here(), s);
    }
}

// Statistics of one converged level run (profiling): out4 += {1 run, `sweeps` launched,
// list items not evaluated -- nothing the block reads had changed, or a chase of the
// same sweep had evaluated it (counts[kCntSkipped]) --, evaluations run (every block in
// sweep 0 plus the items of the later sweeps (counts[kCntQueued]) and the chased blocks
// (dcount[kDcChased]), less those not evaluated)}.  Launches on one stream are ordered,
// so plain adds suffice.
__global__ void __launch_bounds__(64) sweep_stats_kernel(SweepState st, int sweeps, long long total_blocks,
                                                         unsigned long long* __restrict__ out4) {
    if (threadIdx.x == 0) {
        const unsigned long long queued = st.counts[kCntQueued];
        const unsigned long long skipped = st.counts[kCntSkipped];
        const unsigned long long chased = st.dcount[kDcChased];
        out4[0] += 1;
        out4[1] += (unsigned long long)sweeps;
        out4[2] += skipped;
        out4[3] += (unsigned long long)total_blocks + queued + chased - skipped;
    }
}

// The refine pass (exact_recalculate, :1326-1451): one wavefront per output block of
// every job (blockIdx.y).  Each block's start vector and parent score are the bilinear
// interpolation of the searched field at its centre (Euclidean division); blocks on
// the border search a half-size rectangle with a third of the penalties and threshold.
// Blocks read nothing the pass writes, so the raster order of the reference's loop is
// not needed.
__device__ __forceinline__ int64_t div_euclid(int64_t a, int64_t b) {
    int64_t q = a / b;
    if (a % b < 0) q += b > 0 ? -1 : 1;
    return q;
}

__global__ void __launch_bounds__(kWave, 2) refine_kernel(const RefineJob* __restrict__ jobs) {
    const Chain ch = uniform_load(&jobs[blockIdx.y].ch);
    const LevelGeom& g = ch.geom;
    const int idx = (int)blockIdx.x;
    if (idx >= g.gw * g.gh) return;
    const int ogw = uniform_int(jobs[blockIdx.y].old_gw), ogh = uniform_int(jobs[blockIdx.y].old_gh);
    const int obw = uniform_int(jobs[blockIdx.y].old_bw), obh = uniform_int(jobs[blockIdx.y].old_bh);
    const int osx = uniform_int(jobs[blockIdx.y].old_sx), osy = uniform_int(jobs[blockIdx.y].old_sy);
    const int thsad = uniform_int(jobs[blockIdx.y].thsad), lam = uniform_int(jobs[blockIdx.y].lambda);
    const int pnew = uniform_int(jobs[blockIdx.y].pnew), type = uniform_int(jobs[blockIdx.y].type);
    const int distance = uniform_int(jobs[blockIdx.y].distance);
    const int bx = idx % g.gw, by = idx / g.gw;
    const int bw = g.bw, bh = g.bh;
    const int cx = bw / 2 + g.step_x * bx, cy = bh / 2 + g.step_y * by;
    const int ox = (cx - obw / 2) / osx, oy = (cy - obh / 2) / osy;
    const int64_t dx = max(cx - (obw / 2 + osx * ox), 0), dy = max(cy - (obh / 2 + osy * oy), 0);
    auto at = [&](int x, int y) { return uniform_load(&ch.in[min(max(y, 0), ogh - 1) * ogw + min(max(x, 0), ogw - 1)]); };
    const DVec p00 = at(ox, oy), p10 = at(ox + 1, oy), p01 = at(ox, oy + 1), p11 = at(ox + 1, oy + 1);
    auto interp = [&](int64_t a, int64_t b2, int64_t c, int64_t d) {
        const int64_t top = a * osx + dx * (b2 - a), bottom = c * osx + dx * (d - c);
        return div_euclid(top + div_euclid(dy * (bottom - top), osy), osx);
    };
    const int mvx = (int)interp(p00.dx, p10.dx, p01.dx, p11.dx);
    const int mvy = (int)interp(p00.dy, p10.dy, p01.dy, p11.dy);
    const int64_t parent_sad = interp(p00.score, p10.score, p01.score, p11.score) * (int64_t)(bw * bh) /
                               (int64_t)(obw * obh);
    const bool border = bx == 0 || by == 0 || bx == g.gw - 1 || by == g.gh - 1;
    const int factor = (border && bw > 4) ? 3 : 1;
    const int ebw = factor == 3 ? bw / 2 : bw, ebh = factor == 3 ? bh / 2 : bh;
    const int epx = bx * g.step_x + (factor == 3 ? bw / 4 : 0), epy = by * g.step_y + (factor == 3 ? bh / 4 : 0);
    // mv_bounds (:1620-1636), then exact_clip (:1269-1277) of the start vector
    const int b0 = -g.pel * epx, b1 = -g.pel * epy;
    const int b2 = g.pel * (g.lw - epx - ebw) + 1, b3 = g.pel * (g.lh - epy - ebh) + 1;
    const int prx = min(max(mvx, b0), max(b2 - 1, b0)), pry = min(max(mvy, b1), max(b3 - 1, b1));
    Blk b;
    init_blk(ch, g, b);
    DVec o;
    if (parent_sad > (int64_t)(thsad / factor)) {
        const RefineStart rs{epx, epy, ebw, ebh, prx, pry, by == 0 ? 0 : (lam >> 2) / factor, pnew / factor,
                             type, abs(distance), factor};
        bool bad;
        uint32_t found;
        bool unchanged = false;
        o = search_block<true>(ch, g, b, idx, 0, &bad, &found, nullptr, nullptr, &unchanged, rs);
    } else {
        b.px = epx; b.py = epy; b.bw = ebw; b.bh = ebh;
        const uint32_t sad = (uint32_t)min(max(parent_sad, (int64_t)0), (int64_t)0xFFFFFFFFll);
        o.dx = (int16_t)prx;
        o.dy = (int16_t)pry;
        o.score = sat_mul(sad, (uint32_t)factor);
        o.luma = luma_dc(b);
        o.done = 1;
    }
    if (lane_id() == 0) ch.res[idx] = o;
}

// ------------------------------------------------------------------ level transitions


// exact_interpolate_level (:421-478), part 1: child predictors and projection keys.
__global__ void interp_kernel(const Transition* __restrict__ tr) {
    const Transition t = tr[blockIdx.y];
    const int n = t.gw * t.gh;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = i % t.gw, y = i / t.gw;
        DVec o{};
        o.dx = 0; o.dy = 0; o.score = 0; o.luma = 0; o.done = 0;
        t.proj_key[i] = ~0ull;
        // top level (no coarser field): every block starts at Vec8::default()
        if (t.coarse && x < t.gw - 1 && y < t.gh - 1) {
            const int cw = t.cgw - 1, chh = t.cgh - 1;
            DVec p;
            if (x < 2 * cw && y < 2 * chh) {
                p = t.coarse[(y >> 1) * t.cgw + (x >> 1)];
            } else {
                const int e = chh + min(x >> 1, cw) - min(y >> 1, chh);
                const int xx = e <= cw ? e : cw, yy = e <= cw ? chh : cw + chh - e;
                p = t.coarse[yy * t.cgw + xx];
            }
            const int vx = (int)p.dx * 2 * t.pel, vy = (int)p.dy * 2 * t.pel;
            o.dx = (int16_t)clampi(vx, -32768, 32767);
            o.dy = (int16_t)clampi(vy, -32768, 32767);
            o.score = p.score;
            o.luma = p.luma;
        }
        t.next[i] = o;
    }
}

// part 2: scatter the minimum-score key to each projection target (raster order breaks ties).
__global__ void project_kernel(const Transition* __restrict__ tr) {
    const Transition t = tr[blockIdx.y];
    if (!t.coarse) return;
    const int cw = t.cgw - 1, chh = t.cgh - 1;
    const int n = (t.gw - 1) * (t.gh - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = i % (t.gw - 1), y = i / (t.gw - 1);
        if (x >= 2 * cw || y >= 2 * chh) continue;
        const DVec p = t.coarse[(y >> 1) * t.cgw + (x >> 1)];
        const int vx = (int)p.dx * 2 * t.pel, vy = (int)p.dy * 2 * t.pel;
        const int sxs = t.bw - t.ox, sys = t.bh - t.oy;
        const int tx = clampi((vx + (t.bw >> 1) + sxs * x) / sxs, 0, t.gw - 1);
        const int ty = clampi((vy + (t.bh >> 1) + sys * y) / sys, 0, t.gh - 1);
        const unsigned long long key = ((unsigned long long)p.score << 32) | (unsigned)(y * t.gw + x);
        atomicMin(&t.proj_key[ty * t.gw + tx], key);
    }
}

// part 3: materialise the projected vectors (negated) from the winning keys.
__global__ void project_finish_kernel(const Transition* __restrict__ tr) {
    const Transition t = tr[blockIdx.y];
    if (!t.coarse) return;
    const int n = t.gw * t.gh;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long key = t.proj_key[i];
        DVec o{};
        o.dx = 0; o.dy = 0; o.score = 0xFFFFFFFFu; o.luma = 0; o.done = 0;
        if (key != ~0ull) {
            const int si = (int)(key & 0xFFFFFFFFu);
            const int x = si % t.gw, y = si / t.gw;
            const DVec p = t.coarse[(y >> 1) * t.cgw + (x >> 1)];
            const int vx = (int)p.dx * 2 * t.pel, vy = (int)p.dy * 2 * t.pel;
            o.dx = (int16_t)(-(int)(int16_t)clampi(vx, -32768, 32767));
            o.dy = (int16_t)(-(int)(int16_t)clampi(vy, -32768, 32767));
            o.score = p.score;
        }
        t.proj[i] = o;
    }
}


constexpr int kHist = 8192;

// exact_global_doubled (:1285-1324): modes of dx and dy (smallest value among the most
// frequent), then twice the rounded mean of the vectors within 5 of both modes.
__global__ void __launch_bounds__(256) global_kernel(const GlobalJob* __restrict__ jobs) {
    const GlobalJob j = jobs[blockIdx.x];
    __shared__ uint32_t hist[kHist];
    __shared__ int red_i[256];
    __shared__ uint32_t red_c[256];
    __shared__ long long red_l[3][256];
    int mode[2];
    for (int comp = 0; comp < 2; comp++) {
        // value range: searched coarse vectors stay within the level size
        int lo = INT32_MAX, hi = INT32_MIN;
        for (int i = threadIdx.x; i < j.n; i += 256) {
            const int v = comp ? j.field[i].dy : j.field[i].dx;
            lo = min(lo, v); hi = max(hi, v);
        }
        red_i[threadIdx.x] = lo;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red_i[threadIdx.x] = min(red_i[threadIdx.x], red_i[threadIdx.x + s]);
            __syncthreads();
        }
        lo = red_i[0];
        __syncthreads();
        red_i[threadIdx.x] = hi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red_i[threadIdx.x] = max(red_i[threadIdx.x], red_i[threadIdx.x + s]);
            __syncthreads();
        }
        hi = red_i[0];
        __syncthreads();
        uint32_t best_c = 0;
        int best_v = lo;
        for (int base = lo; base <= hi; base += kHist) {
            for (int i = threadIdx.x; i < kHist; i += 256) hist[i] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < j.n; i += 256) {
                const int v = (comp ? j.field[i].dy : j.field[i].dx) - base;
                if (v >= 0 && v < kHist) atomicAdd(&hist[v], 1u);
            }
            __syncthreads();
            // first maximum in value order
            uint32_t c = 0;
            int v = 0x7FFFFFFF;
            for (int i = threadIdx.x; i < kHist; i += 256)
                if (hist[i] > c) { c = hist[i]; v = i; }
            red_c[threadIdx.x] = c;
            red_i[threadIdx.x] = v;
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if (threadIdx.x < s) {
                    const uint32_t oc = red_c[threadIdx.x + s];
                    const int ov = red_i[threadIdx.x + s];
                    if (oc > red_c[threadIdx.x] || (oc == red_c[threadIdx.x] && ov < red_i[threadIdx.x])) {
                        red_c[threadIdx.x] = oc;
                        red_i[threadIdx.x] = ov;
                    }
                }
                __syncthreads();
            }
            if (red_c[0] > best_c) { best_c = red_c[0]; best_v = base + red_i[0]; }
            __syncthreads();
        }
        mode[comp] = best_v;
    }
    long long sx = 0, sy = 0, cnt = 0;
    for (int i = threadIdx.x; i < j.n; i += 256) {
        const DVec f = j.field[i];
        if (abs((int)f.dx - mode[0]) < 6 && abs((int)f.dy - mode[1]) < 6) { sx += f.dx; sy += f.dy; cnt++; }
    }
    red_l[0][threadIdx.x] = sx; red_l[1][threadIdx.x] = sy; red_l[2][threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 3; q++) red_l[q][threadIdx.x] += red_l[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sx = red_l[0][0]; sy = red_l[1][0]; cnt = red_l[2][0];
        if (j.n == 0) { mode[0] = 0; mode[1] = 0; }
        if (cnt == 0) {
            j.out[0] = 2 * mode[0];
            j.out[1] = 2 * mode[1];
        } else {
            auto mean = [&](long long s) { const long long sg = s > 0 ? 1 : (s < 0 ? -1 : 0); return (s + sg * cnt / 2) / cnt; };
            j.out[0] = 2 * (int)mean(sx);
            j.out[1] = 2 * (int)mean(sy);
        }
    }
}


// reorder_search_order (:404-419), key preparation: score >> shift over the raster of
// the (gw-1) x (gh-1) interior; the stable radix sort then gives the reference's order.
__global__ void sort_keys_kernel(const SortJob* __restrict__ jobs) {
    const SortJob j = jobs[blockIdx.y];
    const int n = (j.gw - 1) * (j.gh - 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = i % (j.gw - 1), y = i / (j.gw - 1);
        j.keys[j.seg_begin + i] = j.field[y * j.gw + x].score >> j.shift;
        j.vals[j.seg_begin + i] = (uint32_t)(x | (y << 16));
    }
}


// Children of one parent in the reference's generation order (refine_search_order
// :338-402), reported through `emit(x, y)`; children outside the level are skipped.
template <class F>
__device__ void children(int px, int py, int sw, int sh, int dw, int dh, int ex, int ey, F emit) {
    const int left = 2 * px, top = 2 * py;
    const int off[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int k = 0; k < 4; k++) emit(left + off[k][0], top + off[k][1]);
    if (px == sw - 1)
        for (int dx = 2; dx < 2 + ex; dx++)
            for (int dy = 0; dy < 2 + (py == sh - 1 ? 1 : 0) * ey; dy++) emit(left + dx, top + dy);
    if (py == sh - 1)
        for (int dy = 2; dy < 2 + ey; dy++)
            for (int dx = 0; dx < 2; dx++) emit(left + dx, top + dy);
}

// refine_search_order for level l >= 1: interior children of the sorted parents, then
// their boundary children, then the edge row and column.  One workgroup per chain.
__global__ void __launch_bounds__(256) child_order_kernel(const OrderJob* __restrict__ jobs) {
    const OrderJob j = jobs[blockIdx.x];
    const int sw = j.cgw - 1, sh = j.cgh - 1, dw = j.gw - 1, dh = j.gh - 1;
    const int ex = max(dw - 2 * sw, 0), ey = max(dh - 2 * sh, 0);
    const int np = sw * sh;
    __shared__ int scan_a[256], scan_b[256];
    __shared__ int total_int;
    // pass 0: number of interior children overall
    for (int pass = 0; pass < 2; pass++) {
        int run_a = 0, run_b = 0;
        for (int base = 0; base < np; base += 256) {
            const int k = base + threadIdx.x;
            int na = 0, nb = 0;
            int ppx = 0, ppy = 0;
            if (k < np) {
                const uint32_t pp = j.parents[k];
                ppx = (int)(pp & 0xFFFF); ppy = (int)(pp >> 16);
                children(ppx, ppy, sw, sh, dw, dh, ex, ey, [&](int x, int y) {
                    if (x < 0 || y < 0 || x >= dw || y >= dh) return;
                    if (x > 0 && y > 0 && x < dw - 1 && y < dh - 1) na++; else nb++;
                });
            }
            scan_a[threadIdx.x] = na;
            scan_b[threadIdx.x] = nb;
            __syncthreads();
            for (int off = 1; off < 256; off <<= 1) {  // inclusive Hillis-Steele scan
                const int va = threadIdx.x >= off ? scan_a[threadIdx.x - off] : 0;
                const int vb = threadIdx.x >= off ? scan_b[threadIdx.x - off] : 0;
                __syncthreads();
                scan_a[threadIdx.x] += va;
                scan_b[threadIdx.x] += vb;
                __syncthreads();
            }
            if (pass == 1 && k < np) {
                int pa = run_a + scan_a[threadIdx.x] - na;
                int pb = total_int + run_b + scan_b[threadIdx.x] - nb;
                children(ppx, ppy, sw, sh, dw, dh, ex, ey, [&](int x, int y) {
                    if (x < 0 || y < 0 || x >= dw || y >= dh) return;
                    if (x > 0 && y > 0 && x < dw - 1 && y < dh - 1) {
                        if (pa < dw * dh) j.order[pa] = x | (y << 16);
                        pa++;
                    } else {
                        if (pb < dw * dh) j.order[pb] = x | (y << 16);
                        pb++;
                    }
                });
            }
            run_a += scan_a[255];
            run_b += scan_b[255];
            __syncthreads();
        }
        if (pass == 0) {
            if (threadIdx.x == 0) total_int = run_a;
            __syncthreads();
        } else {
            const int start = run_a + run_b;  // == dw * dh
            for (int x = threadIdx.x; x <= dw; x += 256) j.order[start + x] = x | (dh << 16);
            for (int y = threadIdx.x; y < dh; y += 256) j.order[start + dw + 1 + y] = dw | (y << 16);
        }
    }
}


// rescale_scores (:2716-2738) + write_vecs (:2800-2811) for one region.
__global__ void pack_kernel(const PackJob* __restrict__ jobs, int denom) {
    const PackJob j = jobs[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < j.count; i += gridDim.x * blockDim.x) {
        const DVec v = j.field[i];
        const int mag = abs((int)v.dx) + abs((int)v.dy);
        const int ratio = (100 * mag) / denom;
        int score = (int)v.score;
        if (ratio >= 51) {
            const long long r = (long long)score * 20;
            score = (int)(r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : r));
        } else if (ratio >= 16) {
            const unsigned long long p = 3926827243ull * (unsigned long long)(uint32_t)score * (unsigned long long)(uint32_t)(ratio - 15);
            const int v71 = (int)(uint32_t)(p >> 32);
            const long long r = (long long)score + 9LL * ((v71 >> 31) + (v71 >> 5));
            score = (int)(r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : r));
        }
        const uint32_t d0 = ((uint32_t)(int32_t)v.dx << 16) | (uint32_t)(uint16_t)v.dy;
        const uint32_t d1 = ((uint32_t)score & 0x00FFFFFFu) | ((uint32_t)(v.luma & 0xFF) << 24);
        uint8_t* o = j.out + 4 + 8 * (size_t)i;
        // payload offsets are 4-byte aligned when the payload buffer is
        reinterpret_cast<uint32_t*>(o)[0] = d0;
        reinterpret_cast<uint32_t*>(o)[1] = d1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) reinterpret_cast<int32_t*>(j.out)[0] = 2 * j.count + 1;
    if (blockIdx.x == 0 && threadIdx.x < 16 && j.header) reinterpret_cast<int32_t*>(j.payload)[threadIdx.x] = j.header[threadIdx.x];
}

}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_search_level(const Chain* d_chains, int nchains, hipStream_t s) {
    hipLaunchKernelGGL(search_level_kernel, dim3(nchains), dim3(kWave), 0, s, d_chains);
    return hipGetLastError();
}

hipError_t launch_sweep_init(const Chain* d_chains, int nchains, int max_n, const SweepState& st, hipStream_t s) {
    hipLaunchKernelGGL(sweep_init_kernel, dim3((max_n + 255) / 256, nchains), dim3(256), 0, s, d_chains, st);
    return hipGetLastError();
}

hipError_t launch_sweep0(const Chain* d_chains, int nchains, int max_n, const SweepState& st, hipStream_t s,
                         int level, bool fast, bool l0, bool lean, double slot_scale, int coarse_slots, int fine_run,
                         int coarse_run) {
    // segments per chain: a multiple of q so that the slot count is a multiple of the 8
    // XCDs, and no segment much shorter than seg_min blocks.  Fewer slots put more items
    // of a slot in flight at once (read before their predecessor is stored); more slots
    // put more blocks at a segment start (read before the previous segment reaches
    // them).  The slot count that needed the fewest evaluations at the finest level
    // (1080p, 8x8) fell with the batch.  Two waves per SIMD (the general kernel): about
    // 2000 slots at 32 chains, 1080 at 120, 720 at 240, i.e. at most about
    // 2048 * sqrt(32 / chains), rounded down to whole segments per chain.  Three (the
    // tile-plan kernel): about 3600 slots at 32 chains and again 720 at 240 (3
    // segments beat 2 and 4), i.e. about 3584 * (32 / chains)^0.8, with segments of
    // 1024 blocks or more (the finest level's factor slot_scale and run fine_run are
    // tuned per workload).  The coarser levels' orders are sorted by parent score, so a
    // block's earlier-visited neighbours lie anywhere before it: one segment per chain
    // (coarse_slots 256 at 240 chains) keeps in flight only the positions just before
    // the ones being searched (1080p, 120 pairs: 6% faster than 4096 slots), and items of
    // coarse_run positions (16: four parents' children; 2% faster than 8 there) search
    // siblings in order.
    const double r = 32.0 / nchains;
    const double t0 = fast ? 3584.0 * std::pow(r, 0.8) : 2048.0 * std::sqrt(r);
    const int target = level == 0 ? (int)std::min(8192.0, std::max(256.0, t0 * slot_scale)) : coarse_slots;
    // runs at the finest level: consecutive positions searched by one wave keep the
    // reference rows of a block's window in the caches for the next (fine_run: 1 to 32)
    const int run = level == 0 ? ((fast && l0) ? fine_run : 1) : (fast ? coarse_run : 8);
    const int seg_min = fast ? 1024 : 2048;
    int g8 = 8;
    for (int a = nchains % 8; a != 0;) {
        const int t = g8 % a;
        g8 = a;
        a = t;
    }
    const int q = 8 / g8;
    int nseg = q * std::max(target / (nchains * q), 1);
    const int cap = std::max(q, max_n / seg_min / q * q);
    nseg = std::min(nseg, cap);
    const int nslots = nchains * nseg;
    const long long total = (long long)nslots * (((max_n + nseg - 1) / nseg + run - 1) / run);
    const unsigned grid = (unsigned)std::min<long long>(total, 1 << 22);
    if (fast && l0)
        hipLaunchKernelGGL((sweep0_kernel<true, true>), dim3(grid), dim3(kWave), 0, s, d_chains, st, nslots, nseg, run,
                           total);
    else if (fast && lean)
        hipLaunchKernelGGL((sweep0_kernel<true, false, true>), dim3(grid), dim3(kWave), 0, s, d_chains, st, nslots,
                           nseg, run, total);
    else if (fast)
        hipLaunchKernelGGL((sweep0_kernel<true>), dim3(grid), dim3(kWave), 0, s, d_chains, st, nslots, nseg, run,
                           total);
    else
        hipLaunchKernelGGL((sweep0_kernel<false>), dim3(grid), dim3(kWave), 0, s, d_chains, st, nslots, nseg, run,
                           total);
    return hipGetLastError();
}

hipError_t launch_refine(const void* d_jobs, int njobs, int max_n, hipStream_t s) {
    hipLaunchKernelGGL(refine_kernel, dim3(max_n, njobs), dim3(kWave), 0, s, (const RefineJob*)d_jobs);
    return hipGetLastError();
}

hipError_t launch_sweep(const Chain* d_chains, int grid, const SweepState& st, int sweep, hipStream_t s, bool l0) {
    if (l0)
        hipLaunchKernelGGL(sweep_kernel<true>, dim3(grid), dim3(kWave), 0, s, d_chains, st, sweep);
    else
        hipLaunchKernelGGL(sweep_kernel<false>, dim3(grid), dim3(kWave), 0, s, d_chains, st, sweep);
    return hipGetLastError();
}

hipError_t launch_sweep_fast(const Chain* d_chains, int grid, const SweepState& st, int sweep, hipStream_t s, bool l0) {
    if (l0)
        hipLaunchKernelGGL((sweep_kernel<true, true, false>), dim3(grid), dim3(kWave), 0, s, d_chains, st, sweep);
    else
        hipLaunchKernelGGL((sweep_kernel<false, true, true>), dim3(grid), dim3(kWave), 0, s, d_chains, st, sweep);
    return hipGetLastError();
}

hipError_t launch_sweep_stats(const SweepState& st, int sweeps, long long total_blocks, unsigned long long* out4,
                              hipStream_t s) {
    hipLaunchKernelGGL(sweep_stats_kernel, dim3(1), dim3(64), 0, s, st, sweeps, total_blocks, out4);
    return hipGetLastError();
}

hipError_t launch_deferred(const Chain* d_chains, int grid, const SweepState& st, int sweep, hipStream_t s, bool l0) {
    if (l0)
        hipLaunchKernelGGL(deferred_kernel<true>, dim3(grid), dim3(kWave), 0, s, d_chains, st, sweep);
    else
        hipLaunchKernelGGL(deferred_kernel<false>, dim3(grid), dim3(kWave), 0, s, d_chains, st, sweep);
    return hipGetLastError();
}


hipError_t launch_validate(const Chain* d_chains, int nchains, int max_n, const SweepState& st, hipStream_t s) {
    hipLaunchKernelGGL(validate_kernel, dim3((max_n + 255) / 256, nchains), dim3(256), 0, s, d_chains, st);
    return hipGetLastError();
}

hipError_t launch_sweep_scan(const Chain* d_chains, int nchains, const SweepState& st, int sweep, hipStream_t s) {
    hipLaunchKernelGGL(sweep_scan_kernel, dim3(nchains), dim3(kScanThreads), 0, s, d_chains, st, sweep);
    return hipGetLastError();
}

hipError_t launch_interp(const void* d_tr, int ntr, int max_n, hipStream_t s) {
    const Transition* t = (const Transition*)d_tr;
    dim3 grid((max_n + 255) / 256, ntr);
    hipLaunchKernelGGL(interp_kernel, grid, dim3(256), 0, s, t);
    hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, s, t);
    hipLaunchKernelGGL(project_finish_kernel, grid, dim3(256), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_global(const void* d_jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(global_kernel, dim3(njobs), dim3(256), 0, s, (const GlobalJob*)d_jobs);
    return hipGetLastError();
}

hipError_t launch_sort_keys(const void* d_jobs, int njobs, int max_n, hipStream_t s) {
    dim3 grid((max_n + 255) / 256, njobs);
    hipLaunchKernelGGL(sort_keys_kernel, grid, dim3(256), 0, s, (const SortJob*)d_jobs);
    return hipGetLastError();
}

hipError_t segmented_sort(void* d_temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, int n, int nseg, const int* d_begin,
                          const int* d_end, hipStream_t s) {
    return hipcub::DeviceSegmentedRadixSort::SortPairs(d_temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, n,
                                                       nseg, d_begin, d_end, 0, 32, s);
}

hipError_t launch_child_order(const void* d_jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(child_order_kernel, dim3(njobs), dim3(256), 0, s, (const OrderJob*)d_jobs);
    return hipGetLastError();
}

hipError_t launch_pack(const void* d_jobs, int njobs, int max_count, int denom, hipStream_t s) {
    dim3 grid((max_count + 255) / 256, njobs);
    hipLaunchKernelGGL(pack_kernel, grid, dim3(256), 0, s, (const PackJob*)d_jobs, denom);
    return hipGetLastError();
}

size_t transition_size() { return sizeof(Transition); }
size_t global_job_size() { return sizeof(GlobalJob); }
size_t sort_job_size() { return sizeof(SortJob); }
size_t order_job_size() { return sizeof(OrderJob); }
size_t pack_job_size() { return sizeof(PackJob); }

}  // namespace svpf
