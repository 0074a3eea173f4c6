// Analyse block search for gfx950 (replaces crates/svpflow1/src/analyse_search.rs).
//
// Execution model
//   * One wavefront per (frame pair, direction) chain walks one pyramid level in the
//     reference's exact visit order (analyse_search.rs:512-680).  The order is a single
//     dependency chain (each block seeds from already-searched neighbours), so blocks
//     are never reordered; parallelism comes from candidates and from chains.
//   * Within a block, the candidates of one search step go across the 64 lanes: each
//     lane computes one candidate's full cost (luma SAD/SATD + 4x chroma).  The step's
//     winner is the lane argmin with the lowest list index on ties, which equals the
//     reference's sequential strict-`<` running minimum (exact_check, :1122-1155): the
//     early-outs of exact_check / exact_sad_if_better only skip candidates whose full
//     cost cannot be below the running best (penalties are monotone for pnew >= -256;
//     below that a lane-serial scan reproduces the skip rules literally).
//   * Level transitions (interpolation + opposite-field projection, global predictor,
//     score sort, child visit order) run as separate kernels over all chains, so a
//     batch of pairs never returns to the host between levels.
#include <hipcub/hipcub.hpp>

#include "analyse_dev.h"
#include "svpf_internal.h"

namespace svpf {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r < a ? 0xFFFFFFFFu : r;
}
__device__ __forceinline__ uint32_t sat_mul(uint32_t a, uint32_t m) {
    uint64_t r = (uint64_t)a * m;
    return r > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)r;
}
__device__ __forceinline__ int sat_mul_i32(int a, int b) {
    const long long r = (long long)a * b;
    return (int)(r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : r));
}
__device__ __forceinline__ int sat_add_i32(int a, int b) {
    const long long r = (long long)a + b;
    return (int)(r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : r));
}
__device__ __forceinline__ bool pow2_4_32(int v) { return v == 4 || v == 8 || v == 16 || v == 32; }
// A byte of a plane with the reference's slice semantics (`get(i).unwrap_or(fb)`).
__device__ __forceinline__ int rd(const uint8_t* p, int64_t len, int64_t i, int fb) {
    return (i >= 0 && i < len) ? (int)p[i] : fb;
}

// Sum of |H4 D H4^T| over one 4x4 tile (analyse_search.rs:2681-2702).
__device__ __forceinline__ uint32_t had4(const int d[16]) {
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int s0 = d[i * 4 + 0] + d[i * 4 + 1], s1 = d[i * 4 + 0] - d[i * 4 + 1];
        int s2 = d[i * 4 + 2] + d[i * 4 + 3], s3 = d[i * 4 + 2] - d[i * 4 + 3];
        t[i * 4 + 0] = s0 + s2; t[i * 4 + 1] = s1 + s3; t[i * 4 + 2] = s0 - s2; t[i * 4 + 3] = s1 - s3;
    }
    int s = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int s0 = t[j] + t[4 + j], s1 = t[j] - t[4 + j], s2 = t[8 + j] + t[12 + j], s3 = t[8 + j] - t[12 + j];
        s += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3);
    }
    return (uint32_t)max(s, 0);
}


typedef uint32_t __attribute__((aligned(1))) u32u;
__device__ __forceinline__ uint32_t ld32(const uint8_t* p) { return *reinterpret_cast<const u32u*>(p); }

// Sum of |H4 D H4^T| for one 4x4 tile given four packed rows of source and reference.
__device__ __forceinline__ uint32_t had4_rows(const uint32_t a[4], const uint32_t b[4]) {
    int d[16];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) d[r * 4 + c] = (int)((a[r] >> (8 * c)) & 255) - (int)((b[r] >> (8 * c)) & 255);
    return had4(d);
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

// The block being searched (uniform across the wave).
struct Blk {
    const uint8_t* sp[3];
    const uint8_t* rp[3];
    int ss[3], rs[3];
    int64_t sl[3], rl[3];
    int level, lw, lh, y_off, c_off, clw, clh, pel;
    int px, py, bw, bh;
    bool satd;
};

// satd_edge (analyse_search.rs:2305-2351): all coordinates clamped into the plane.
__device__ __attribute__((always_inline)) uint32_t satd_edge(const uint8_t* s, int64_t sl, int ss, int srow, int sx0, int sy0, const uint8_t* r,
                              int64_t rl, int rs, int rrow, int rx0, int ry0, int w, int h, int n) {
    uint32_t total = 0;
    const int mw = max(w - 1, 0), mh = max(h - 1, 0);
    for (int ty = 0; ty < n / 4; ty++)
        for (int tx = 0; tx < n / 4; tx++) {
            int d[16];
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                int sy = clampi(sy0 + ty * 4 + rr, 0, mh), ry = clampi(ry0 + ty * 4 + rr, 0, mh);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    int sx = clampi(sx0 + tx * 4 + c, 0, mw), rx = clampi(rx0 + tx * 4 + c, 0, mw);
                    d[rr * 4 + c] = rd(s, sl, (int64_t)(srow + sy) * ss + sx, 0) - rd(r, rl, (int64_t)(rrow + ry) * rs + rx, 0);
                }
            }
            total = sat_add(total, had4(d));
        }
    return total >> 1;
}

// Candidate cost (luma, 4x chroma) by the reference's path selection:
// out-of-bounds mv -> block_cost_edge_lc (:1996-2167); otherwise
// block_cost_luma_interior (:1719-1852) + chroma_sad_x4 (:2169-2303).
__device__ __attribute__((noinline)) void cand_cost(const Blk& b, int mvx, int mvy, bool chroma, uint32_t& L, uint32_t& C) {
    const int pel = b.pel;
    const int b0 = -pel * b.px, b1 = -pel * b.py;
    const int b2 = pel * (b.lw - b.px - b.bw) + 1, b3 = pel * (b.lh - b.py - b.bh) + 1;
    const int ox = mvx < b0 ? mvx - b0 : ((b2 > b0 && mvx >= b2) ? mvx + 1 - b2 : 0);
    const int oy = mvy < b1 ? mvy - b1 : ((b3 > b1 && mvy >= b3) ? mvy + 1 - b3 : 0);
    const bool sq = b.satd && pow2_4_32(b.bw) && pow2_4_32(b.bh) && b.bw == b.bh;
    const bool subpel = b.level == 0 && pel >= 2;
    if (ox != 0 || oy != 0) {
        // ---- edge path (edge_shift_origins :1638-1676 + block_cost_edge_lc)
        const int cx = clampi(b.px - ox, 0, max(b.lw - b.bw, 0));
        const int cy = clampi(b.py - oy, 0, max(b.lh - b.bh, 0));
        int rx0, ry0;
        if (pel == 1) { rx0 = cx + mvx; ry0 = cy + mvy; }
        else if (pel == 4) { rx0 = (4 * cx + mvx) >> 2; ry0 = (4 * cy + mvy) >> 2; }
        else { rx0 = (2 * cx + mvx) >> 1; ry0 = (2 * cy + mvy) >> 1; }
        const int sub = subpel ? (mvy & 1) * pel + (mvx & 1) : 0;
        const int rrow = b.y_off + sub * b.lh;
        if (sq) {
            L = satd_edge(b.sp[0], b.sl[0], b.ss[0], b.y_off, cx, cy, b.rp[0], b.rl[0], b.rs[0], rrow, rx0, ry0, b.lw,
                          b.lh, b.bw);
        } else {
            uint32_t sad = 0;
            const int mw = max(b.lw - 1, 0), mh = max(b.lh - 1, 0);
            for (int row = 0; row < b.bh; row++) {
                const int sy = clampi(cy + row, 0, mh), ry = clampi(ry0 + row, 0, mh);
                for (int col = 0; col < b.bw; col++) {
                    const int sx = clampi(cx + col, 0, mw), rx = clampi(rx0 + col, 0, mw);
                    sad += (uint32_t)abs(rd(b.sp[0], b.sl[0], (int64_t)(b.y_off + sy) * b.ss[0] + sx, 0) -
                                         rd(b.rp[0], b.rl[0], (int64_t)(rrow + ry) * b.rs[0] + rx, 0));
                }
            }
            L = sad;
        }
        C = 0;
        if (chroma) {
            const int cw = b.bw / 2, ch = b.bh / 2;
            const int ccx = (b.px >> 1) - ((b.px - cx) >> 1), ccy = (b.py >> 1) - ((b.py - cy) >> 1);
            const int crx0 = ((mvx >> 1) + 2 * ccx) >> 1, cry0 = ((mvy >> 1) + 2 * ccy) >> 1;
            const int cref = b.c_off + sub * b.clh;
            uint32_t uv;
            if (b.satd && cw == ch && (cw == 4 || cw == 8 || cw == 16)) {
                uv = sat_add(satd_edge(b.sp[1], b.sl[1], b.ss[1], b.c_off, ccx, ccy, b.rp[1], b.rl[1], b.rs[1], cref, crx0,
                                       cry0, b.clw, b.clh, cw),
                             satd_edge(b.sp[2], b.sl[2], b.ss[2], b.c_off, ccx, ccy, b.rp[2], b.rl[2], b.rs[2], cref, crx0,
                                       cry0, b.clw, b.clh, cw));
            } else {
                uv = 0;
                const int mw = max(b.clw - 1, 0), mh = max(b.clh - 1, 0);
                for (int row = 0; row < ch; row++) {
                    const int sy = clampi(ccy + row, 0, mh), ry = clampi(cry0 + row, 0, mh);
                    for (int col = 0; col < cw; col++) {
                        const int sx = clampi(ccx + col, 0, mw), rx = clampi(crx0 + col, 0, mw);
                        const int64_t si = (int64_t)(b.c_off + sy), ri = (int64_t)(cref + ry);
                        uv += (uint32_t)abs(rd(b.sp[1], b.sl[1], si * b.ss[1] + sx, 128) - rd(b.rp[1], b.rl[1], ri * b.rs[1] + rx, 128));
                        uv += (uint32_t)abs(rd(b.sp[2], b.sl[2], si * b.ss[2] + sx, 128) - rd(b.rp[2], b.rl[2], ri * b.rs[2] + rx, 128));
                    }
                }
            }
            C = sat_mul(uv, 4);
        }
        return;
    }
    // ---- interior path, luma (block_cost_luma_interior)
    const int ipx = b.px - ((b.level > 0 && b.px + b.bw >= b.lw) ? 1 : 0);
    const int ipy = b.py - ((b.level > 0 && b.py + b.bh >= b.lh) ? 1 : 0);
    const int mxp = subpel ? (mvx >> 1) : mvx, myp = subpel ? (mvy >> 1) : mvy;
    const int sub = subpel ? (mvy & 1) * pel + (mvx & 1) : 0;
    const int rrow = b.y_off + sub * b.lh;
    const uint8_t* S0 = b.sp[0];
    const uint8_t* R0 = b.rp[0];
    const int frx = ipx + mxp, fry = ipy + myp;
    bool done_luma = false;
    if (ipx >= 0 && ipy >= 0 && ipx + b.bw <= b.lw && ipy + b.bh <= b.lh && frx >= 0 && fry >= 0 &&
        frx + b.bw <= b.lw && fry + b.bh <= b.lh && (sq || (b.bw == b.bh))) {
        // interior: satd_interior / sad_u8_simd read exactly these rows
        done_luma = block_fast_any(b.bw, b.bh, S0 + (int64_t)(b.y_off + ipy) * b.ss[0] + ipx, b.ss[0],
                                   R0 + (int64_t)(rrow + fry) * b.rs[0] + frx, b.rs[0], sq, L);
    }
    if (done_luma) {
    } else if (sq) {  // satd_luma_clamp (:2353-2436): source outside the level reads 0, reference clamped
        uint32_t total = 0;
        const int n = b.bw;
        for (int ty = 0; ty < n / 4; ty++)
            for (int tx = 0; tx < n / 4; tx++) {
                int d[16];
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int sy = ipy + ty * 4 + rr;
                    const int ry = clampi(sy + myp, 0, b.lh - 1);
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const int sx = ipx + tx * 4 + c;
                        const int rx = clampi(sx + mxp, 0, b.lw - 1);
                        const int a = (sy >= 0 && sy < b.lh && sx >= 0 && sx < b.lw)
                                          ? rd(S0, b.sl[0], (int64_t)(b.y_off + sy) * b.ss[0] + sx, 0) : 0;
                        d[rr * 4 + c] = a - rd(R0, b.rl[0], (int64_t)(rrow + ry) * b.rs[0] + rx, 0);
                    }
                }
                total = sat_add(total, had4(d));
            }
        L = total >> 1;
    } else {  // SAD with the scalar loop's row/column breaks
        uint32_t sad = 0;
        for (int row = 0; row < b.bh; row++) {
            const int sy = ipy + row;
            if (sy < 0 || sy >= b.lh) break;
            const int ry = clampi(ipy + row + myp, 0, b.lh - 1);
            for (int col = 0; col < b.bw; col++) {
                const int sx = ipx + col;
                if (sx < 0 || sx >= b.lw) break;
                const int rx = clampi(ipx + col + mxp, 0, b.lw - 1);
                sad += (uint32_t)abs(rd(S0, b.sl[0], (int64_t)(b.y_off + sy) * b.ss[0] + sx, 0) -
                                     rd(R0, b.rl[0], (int64_t)(rrow + ry) * b.rs[0] + rx, 0));
            }
        }
        L = sad;
    }
    C = 0;
    if (!chroma) return;
    // ---- interior path, chroma (chroma_sad_x4; x86-64 semantics)
    const int cw = b.bw / 2, ch = b.bh / 2;
    if (cw < 4 || ch < 4) return;
    const int cpx = b.px / 2, cpy = b.py / 2;
    const int mcx = subpel ? (mvx >> 2) : (mvx >> 1), mcy = subpel ? (mvy >> 2) : (mvy >> 1);
    const int cref = b.c_off + ((b.level == 0 && pel > 1) ? sub * b.clh : 0);
    if (cw == 4 || cw == 8 || cw == 16) {
        const int rx0 = cpx + mcx, ry0 = cpy + mcy;
        if (cpx + cw <= b.clw && cpy + ch <= b.clh && rx0 >= 0 && ry0 >= 0 && rx0 + cw <= b.clw && ry0 + ch <= b.clh) {
            uint32_t acc = 0;
            if (cw == ch) {
                uint32_t vu, vv;
                const bool fu = block_fast_any(cw, ch, b.sp[1] + (int64_t)(b.c_off + cpy) * b.ss[1] + cpx, b.ss[1],
                                               b.rp[1] + (int64_t)(cref + ry0) * b.rs[1] + rx0, b.rs[1], b.satd, vu);
                const bool fv = block_fast_any(cw, ch, b.sp[2] + (int64_t)(b.c_off + cpy) * b.ss[2] + cpx, b.ss[2],
                                               b.rp[2] + (int64_t)(cref + ry0) * b.rs[2] + rx0, b.rs[2], b.satd, vv);
                if (fu && fv) {
                    C = sat_mul(sat_add(vu, vv), 4);
                    return;
                }
            }
            for (int pl = 1; pl < 3; pl++) {
                const int64_t sbase = (int64_t)(b.c_off + cpy) * b.ss[pl] + cpx;
                const int64_t rbase = (int64_t)(cref + ry0) * b.rs[pl] + rx0;
                uint32_t v = 0;
                if (b.satd) {  // satd_interior over cw x cw
                    for (int ty = 0; ty < cw / 4; ty++)
                        for (int tx = 0; tx < cw / 4; tx++) {
                            int d[16];
#pragma unroll
                            for (int rr = 0; rr < 4; rr++)
#pragma unroll
                                for (int c = 0; c < 4; c++)
                                    d[rr * 4 + c] = rd(b.sp[pl], b.sl[pl], sbase + (int64_t)(ty * 4 + rr) * b.ss[pl] + tx * 4 + c, 0) -
                                                    rd(b.rp[pl], b.rl[pl], rbase + (int64_t)(ty * 4 + rr) * b.rs[pl] + tx * 4 + c, 0);
                            v = sat_add(v, had4(d));
                        }
                    v >>= 1;
                } else {
                    for (int rr = 0; rr < ch; rr++)
                        for (int c = 0; c < cw; c++)
                            v += (uint32_t)abs(rd(b.sp[pl], b.sl[pl], sbase + (int64_t)rr * b.ss[pl] + c, 0) -
                                               rd(b.rp[pl], b.rl[pl], rbase + (int64_t)rr * b.rs[pl] + c, 0));
                }
                acc = sat_add(acc, v);
            }
            C = sat_mul(acc, 4);
            return;
        }
    }
    uint32_t uv = 0;
    for (int row = 0; row < ch; row++) {
        const int sy = cpy + row;
        if (sy >= b.clh) break;
        const int ry = clampi(cpy + row + mcy, 0, b.clh - 1);
        for (int col = 0; col < cw; col++) {
            const int sx = cpx + col;
            if (sx >= b.clw) break;
            const int rx = clampi(cpx + col + mcx, 0, b.clw - 1);
            const int64_t si = (int64_t)(b.c_off + sy), ri = (int64_t)(cref + ry);
            uv += (uint32_t)abs(rd(b.sp[1], b.sl[1], si * b.ss[1] + sx, 128) - rd(b.rp[1], b.rl[1], ri * b.rs[1] + rx, 128));
            uv += (uint32_t)abs(rd(b.sp[2], b.sl[2], si * b.ss[2] + sx, 128) - rd(b.rp[2], b.rl[2], ri * b.rs[2] + rx, 128));
        }
    }
    C = sat_mul(uv, 4);
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
    bool serial;  // pnew < -256: replay the reference's skip rules lane by lane
};

__device__ __forceinline__ bool in_bounds(const Search& s, int x, int y) {
    return x >= s.bnd[0] && y >= s.bnd[1] && x < s.bnd[2] && y < s.bnd[3];
}

__device__ __forceinline__ void wave_argmin(int64_t& c, int& k) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int64_t oc = __shfl_xor(c, off);
        const int ok = __shfl_xor(k, off);
        if (oc < c || (oc == c && ok < k)) { c = oc; k = ok; }
    }
}

// exact_check over candidates 0..n-1 in list order (gen(k, &x, &y) yields candidate k).
// Updates `best` as the reference's sequence of exact_check calls would and returns
// the index of the candidate that last improved it, or -1.
template <class Gen>
__device__ __attribute__((always_inline)) int check_list(const Blk& b, const Search& s, int n, Gen gen, Best& best) {
    const int lane = threadIdx.x;
    int win = -1;
    for (int base = 0; base < n; base += kWave) {
        const int k = base + lane;
        int64_t cost = INT64_MAX, motion = INT64_MAX;
        uint32_t sad = 0;
        int mx = 0, my = 0;
        bool valid = false;
        if (k < n) {
            gen(k, mx, my);
            if (in_bounds(s, mx, my)) {
                valid = true;
                const int64_t dx = s.pmx - mx, dy = s.pmy - my;
                motion = ((int64_t)s.lambda * (dx * dx + dy * dy)) >> 8;
                uint32_t l, c;
                cand_cost(b, mx, my, s.chroma, l, c);
                sad = sat_add(l, c);
                cost = motion + (int64_t)l + (((int64_t)s.pnew * (int64_t)l) >> 8) + (int64_t)c +
                       (((int64_t)s.pnew * (int64_t)c) >> 8);
            }
        }
        if (!s.serial) {
            int64_t mc = valid ? cost : INT64_MAX;
            int mk = valid ? k : 0x7FFFFFFF;
            wave_argmin(mc, mk);
            if (mk != 0x7FFFFFFF && mc < best.cost) {
                const int src = mk - base;
                best.mvx = __shfl(mx, src);
                best.mvy = __shfl(my, src);
                best.sad = __shfl(sad, src);
                best.cost = mc;
                win = mk;
            }
        } else {
            const int cnt = min(kWave, n - base);
            for (int j = 0; j < cnt; j++) {
                if (!__shfl((int)valid, j)) continue;
                if (__shfl(motion, j) >= best.cost) continue;
                const int64_t cj = __shfl(cost, j);
                if (cj < best.cost) {
                    best.mvx = __shfl(mx, j);
                    best.mvy = __shfl(my, j);
                    best.sad = __shfl(sad, j);
                    best.cost = cj;
                    win = base + j;
                }
            }
        }
    }
    return win;
}

// exact_expanding (:829-886): square ring of `radius` around (cx, cy) with `step`.
__device__ __attribute__((always_inline)) void expanding(const Blk& b, const Search& s, int cx, int cy, int r, int step, Best& best) {
    const int m = (2 * r - 1) / step;  // values of i in -r+step, -r+2step, ... < r
    const int n = 4 * m + 4;
    check_list(b, s, n,
               [&](int k, int& x, int& y) {
                   if (k < 2 * m) {
                       x = cx - r + step * (1 + k / 2);
                       y = (k & 1) ? cy + r : cy - r;
                   } else if (k < 4 * m) {
                       const int j = k - 2 * m;
                       y = cy - r + step * (1 + j / 2);
                       x = (j & 1) ? cx + r : cx - r;
                   } else {
                       const int c = k - 4 * m;  // (-r,-r), (-r,+r), (+r,-r), (+r,+r)
                       x = (c < 2) ? cx - r : cx + r;
                       y = (c & 1) ? cy + r : cy - r;
                   }
               },
               best);
}

__constant__ int kHex2[8][2] = {{-1, -2}, {-2, 0}, {-1, 2}, {1, 2}, {2, 0}, {1, -2}, {-1, -2}, {-2, 0}};
__constant__ int kMod6m1[8] = {5, 0, 1, 2, 3, 4, 5, 0};
__constant__ int kHex4[16][2] = {{-4, 2}, {-4, 1}, {-4, 0}, {-4, -1}, {-4, -2}, {4, -2}, {4, -1}, {4, 0},
                                 {4, 1},  {4, 2},  {2, 3},  {0, 4},   {-2, 3},  {-2, -3}, {0, -4}, {2, -3}};

// exact_hex2 (:941-1033).
__device__ __attribute__((always_inline)) void hex2(const Blk& b, const Search& s, int radius, Best& best) {
    int bmx = best.mvx, bmy = best.mvy;
    if (radius > 1) {
        int dir = -2;
        int w = check_list(b, s, 6, [&](int k, int& x, int& y) { x = bmx + kHex2[k + 1][0]; y = bmy + kHex2[k + 1][1]; }, best);
        if (w >= 0) dir = w;
        if (dir != -2) {
            bmx += kHex2[dir + 1][0];
            bmy += kHex2[dir + 1][1];
            int i = 1;
            while (i < radius / 2 && in_bounds(s, bmx, bmy)) {
                const int odir = kMod6m1[dir + 1];
                dir = -2;
                w = check_list(b, s, 3, [&](int k, int& x, int& y) { x = bmx + kHex2[odir + k][0]; y = bmy + kHex2[odir + k][1]; }, best);
                if (w >= 0) dir = odir + w - 1;
                if (dir == -2) break;
                bmx += kHex2[dir + 1][0];
                bmy += kHex2[dir + 1][1];
                i++;
            }
        }
        best.mvx = bmx;
        best.mvy = bmy;
    }
    expanding(b, s, bmx, bmy, 1, 1, best);
}

// exact_umh (:1036-1119): cross, multi-scale 16-point hexagon, then hex2.
__device__ __attribute__((always_inline)) void umh(const Blk& b, const Search& s, int cx, int cy, int radius, Best& best) {
    const int nd = radius / 2;                   // odd d in [1, radius)
    const int ns = max(radius / 4, 1);
    const int n = 4 * nd + 16 * ns;
    check_list(b, s, n,
               [&](int k, int& x, int& y) {
                   if (k < 2 * nd) {
                       const int d = 1 + 2 * (k / 2);
                       x = (k & 1) ? cx + d : cx - d;
                       y = cy;
                   } else if (k < 4 * nd) {
                       const int j = k - 2 * nd, d = 1 + 2 * (j / 2);
                       x = cx;
                       y = (j & 1) ? cy + d : cy - d;
                   } else {
                       const int j = k - 4 * nd, sc = 1 + j / 16, h = j % 16;
                       x = cx + kHex4[h][0] * sc;
                       y = cy + kHex4[h][1] * sc;
                   }
               },
               best);
    hex2(b, s, radius, best);
}

// exact_refine_search (:771-826).
__device__ __attribute__((always_inline)) void refine(const Blk& b, const Search& s, int type, int radius, Best& best) {
    if (type == 2) {
        hex2(b, s, radius, best);
    } else if (type == 3) {
        umh(b, s, best.mvx, best.mvy, radius, best);
    } else if (type == 4) {
        const int cx = best.mvx, cy = best.mvy, w = 2 * radius + 1;
        check_list(b, s, w * w, [&](int k, int& x, int& y) { x = cx - radius + k % w; y = cy - radius + k / w; }, best);
    }
}

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
    const int lane = threadIdx.x;
    uint32_t sum = 0, cnt = 0;
    // the reference stops a row at the first column past the level, and stops at the first row past it
    const int rows = min(b.bh, max(b.lh - b.py, 0));
    const int cols = min(b.bw, max(b.lw - b.px, 0));
    for (int i = lane; i < rows * cols; i += kWave) {
        const int r = i / cols, c = i % cols;
        sum += (uint32_t)rd(b.sp[0], b.sl[0], (int64_t)(b.y_off + b.py + r) * b.ss[0] + b.px + c, 0);
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
    const int lane = threadIdx.x;
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

__device__ __forceinline__ DVec load_vec(const DVec* p) {
    // Field entries written earlier by this wave: make them visible before reading.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return *p;
}

// exact_search_level (:480-681) for one chain: one workgroup = one wavefront.
__global__ void __launch_bounds__(kWave) search_level_kernel(const Chain* __restrict__ chains) {
    const Chain ch = chains[blockIdx.x];
    const LevelGeom g = *ch.geom;
    const int lane = threadIdx.x;
    const int gx = ch.global ? ch.global[0] : 0, gy = ch.global ? ch.global[1] : 0;
    Blk b;
    for (int i = 0; i < 3; i++) {
        b.sp[i] = ch.src.p[i]; b.rp[i] = ch.ref.p[i];
        b.ss[i] = ch.src.stride[i]; b.rs[i] = ch.ref.stride[i];
        b.sl[i] = ch.src.len[i]; b.rl[i] = ch.ref.len[i];
    }
    b.level = g.level; b.lw = g.lw; b.lh = g.lh; b.y_off = g.y_off; b.c_off = g.c_off;
    b.clw = g.clw; b.clh = g.clh; b.pel = g.pel; b.satd = g.satd != 0;
    const int gw = g.gw, gh = g.gh, level = g.level;
    const bool smallest = g.smallest != 0;
    int bad_count = 0;
    const int nblocks = gw * gh;
    for (int it = 0; it < nblocks; it++) {
        const int packed = ch.order[it];
        const int bx = packed & 0xFFFF, by = packed >> 16;
        if (bx >= gw || by >= gh) continue;  // defensive: orders always hold in-grid positions
        const int idx = by * gw + bx;
        int px = (level > 0 && bx == gw - 1) ? g.lw - g.bw : bx * g.step_x;
        int py = (level > 0 && by == gh - 1) ? g.lh - g.bh : by * g.step_y;
        int bw = g.bw, bh = g.bh;
        if (smallest) { px += g.bw >> 2; py += g.bh >> 2; bw = g.bw >> 1; bh = g.bh >> 1; }
        b.px = px; b.py = py; b.bw = bw; b.bh = bh;
        const bool edge = level > 0 && (bx == gw - 1 || by == gh - 1);
        const int sx = min(bx, gw - 2), sy = min(by, gh - 2);
        DVec pred = load_vec(&ch.field[sy * gw + sx]);
        if (edge && !smallest) {
            // predictor score re-measured at the source block's own position
            Blk e = b;
            e.px = sx * g.step_x;
            e.py = sy * g.step_y;
            uint32_t l, c;
            cand_cost(e, pred.dx, pred.dy, true, l, c);
            pred.score = sat_add(l, c);
        } else if (!edge) {
            pred = load_vec(&ch.field[idx]);
        }
        const int lambda = adapt_lambda(smallest ? 0 : g.lambda_base, g.lsad, pred.score);
        int radius = g.radius;
        if (level > 0 && g.adaptive) {
            const int ax = px - ((px + bw >= g.lw) ? 1 : 0), ay = py - ((py + bh >= g.lh) ? 1 : 0);
            uint32_t act = plane_activity(b.sp[0], b.ss[0], g.y_off + ay, ax, bw, bh, b.satd) +
                           plane_activity(b.sp[1], b.ss[1], g.c_off + py / 2, px / 2, bw / 2, bh / 2, b.satd) +
                           plane_activity(b.sp[2], b.ss[2], g.c_off + py / 2, px / 2, bw / 2, bh / 2, b.satd);
            radius = act < 2 ? 0 : ((radius * (int)min(255u * act / 150u, 255u)) >> 8) + 1;
        }
        // seeds: vectors of already-searched 3x3 neighbours, row-major (:574-584)
        int nseed_lane = 0;
        int sdx = 0, sdy = 0;
        if (lane < 9) {
            const int nx = bx - 1 + lane % 3, ny = by - 1 + lane / 3;
            if (nx >= 0 && ny >= 0 && nx < gw && ny < gh) {
                const DVec nb = load_vec(&ch.field[ny * gw + nx]);
                if (nb.done) { nseed_lane = 1; sdx = nb.dx; sdy = nb.dy; }
            }
        }
        const uint64_t smask = __ballot(nseed_lane);
        const int nseeds = __popcll(smask);

        Search s;
        s.pmx = pred.dx;
        s.pmy = pred.dy;
        s.pnew = g.pnew;
        s.chroma = !smallest;
        s.serial = g.pnew < -256;
        s.bnd[0] = -g.pel * px; s.bnd[1] = -g.pel * py;
        s.bnd[2] = g.pel * (g.lw - px - bw) + 1; s.bnd[3] = g.pel * (g.lh - py - bh) + 1;
        s.lambda = lambda;

        // ---- exact_search_block (:684-768): predictor candidates across lanes
        const bool field_valid = pred.score != 0xFFFFFFFFu;
        bool has_prev = false;
        int prx = 0, pry = 0;
        if (ch.prev) {
            const DVec pv = ch.prev[idx];
            has_prev = pv.score != 0xFFFFFFFFu && pv.score > 0;
            prx = pv.dx; pry = pv.dy;
        }
        // lane 0 field, 1 prev, 2 zero, 3 global, 4.. seeds
        int cmx = 0, cmy = 0, pen = 0;
        bool active = false;
        if (lane == 0) { active = field_valid; cmx = pred.dx; cmy = pred.dy; pen = 0; }
        else if (lane == 1) { active = has_prev; cmx = prx; cmy = pry; pen = g.pprev; }
        else if (lane == 2) { active = true; pen = g.pzero; }
        else if (lane == 3) { active = true; cmx = gx; cmy = gy; pen = g.pglobal; }
        else if (lane < 4 + nseeds) { active = true; pen = g.pnbour; }
        {  // seed j lives in the lane of the j-th set bit of the ballot
            const int j = lane - 4;
            int src = 0, cnt = 0;
            for (int q = 0; q < 9; q++)
                if ((smask >> q) & 1) { if (cnt == j) src = q; cnt++; }
            const int tdx = __shfl(sdx, src), tdy = __shfl(sdy, src);
            if (lane >= 4 && lane < 4 + nseeds) { cmx = tdx; cmy = tdy; }
        }
        uint32_t csad = 0;
        int64_t ccost = 0;
        if (active) {
            uint32_t l, c;
            cand_cost(b, cmx, cmy, !smallest, l, c);
            csad = sat_add(l, c);
            ccost = (int64_t)csad + (((int64_t)pen * (int64_t)csad) >> 8);
        }
        Best best;
        Search r = s;
        r.pmx = pred.dx;
        r.pmy = pred.dy;
        if (g.early) {
            // level 0 with distance 0: the predictor (or zero) is the answer
            const int src = field_valid ? 0 : 2;
            best.mvx = __shfl(cmx, src);
            best.mvy = __shfl(cmy, src);
            best.sad = __shfl(csad, src);
            best.cost = (int64_t)best.sad;
        } else {
            if (field_valid) {
                best = Best{pred.dx, pred.dy, __shfl(csad, 0), __shfl(ccost, 0)};
            } else {
                best = Best{0, 0, 0xFFFFFFFFu, INT64_MAX};
            }
            if (has_prev) {
                const int64_t pc = __shfl(ccost, 1);
                if (pc < best.cost) {
                    best = Best{prx, pry, __shfl(csad, 1), pc};
                    r.pmx = prx;
                    r.pmy = pry;
                }
            }
            {
                const int64_t c3 = __shfl(ccost, 3);
                if (c3 < best.cost) best = Best{gx, gy, __shfl(csad, 3), c3};
                const int64_t c2 = __shfl(ccost, 2);
                if (c2 < best.cost) best = Best{0, 0, __shfl(csad, 2), c2};
            }
            for (int j = 0; j < nseeds; j++) {
                const int64_t cj = __shfl(ccost, 4 + j);
                if (cj < best.cost) best = Best{__shfl(cmx, 4 + j), __shfl(cmy, 4 + j), __shfl(csad, 4 + j), cj};
            }
            best.cost = (int64_t)best.sad;
            const int stype = (edge && level == 1) ? 0 : g.search_type;
            refine(b, r, stype, radius, best);
        }
        // ---- bad-block expansion at coarse levels (:607-672), around the zero vector
        const uint32_t found = best.sad;
        const int bad_limit = sat_add_i32(g.bad_sad, sat_mul_i32(g.bad_sad, bad_count) / 16);
        if (level > 0 && idx > 1 && g.bad_sad > 0 && (int)best.sad > bad_limit) {
            bad_count++;
            Search e = s;
            e.pmx = pred.dx;
            e.pmy = pred.dy;
            if (g.bad_range < 0) {
                const int max_r = -g.bad_range * g.pel;
                for (int rr = g.pel; rr < max_r; rr += g.pel) {
                    expanding(b, e, 0, 0, rr, g.pel, best);
                    if (best.sad < found / 4) break;
                }
            } else if (g.bad_range > 0) {
                umh(b, e, 0, 0, g.bad_range * g.pel, best);
            }
            expanding(b, e, best.mvx, best.mvy, 1, 1, best);
        }
        const uint8_t dc = luma_dc(b);
        if (lane == 0) {
            DVec o;
            o.dx = (int16_t)best.mvx;
            o.dy = (int16_t)best.mvy;
            o.score = sat_mul(best.sad, smallest ? 3u : 1u);
            o.luma = dc;
            o.done = 1;
            ch.field[idx] = o;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
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
