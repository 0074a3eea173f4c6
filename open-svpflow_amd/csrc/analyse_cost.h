// Candidate cost of the Analyse search (the reference's per-candidate path
// selection: block_cost_luma_interior, block_cost_edge_lc, chroma_sad_x4,
// satd_* of crates/svpflow1/src/analyse_search.rs), shared by the HIP search and a
// host build that checks it against the oracle on the CPU.
#pragma once
#include <cstdint>
#include <cstdlib>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SVPF_HD __host__ __device__
#else
#define SVPF_HD
#endif

#if defined(__HIPCC__)
#define SVPF_UNROLL _Pragma("unroll")
#define SVPF_NOUNROLL _Pragma("unroll 1")
#else
#define SVPF_UNROLL
#define SVPF_NOUNROLL
#endif

#if !defined(__HIPCC__)
#include <algorithm>
#endif

namespace svpf {
namespace cost {

#if !defined(__HIPCC__)
using std::max;
using std::min;
#endif

SVPF_HD inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
SVPF_HD inline uint32_t sat_add(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r < a ? 0xFFFFFFFFu : r;
}
SVPF_HD inline uint32_t sat_mul(uint32_t a, uint32_t m) {
    uint64_t r = (uint64_t)a * m;
    return r > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)r;
}
SVPF_HD inline int sat_mul_i32(int a, int b) {
    const long long r = (long long)a * b;
    return (int)(r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : r));
}
SVPF_HD inline int sat_add_i32(int a, int b) {
    const long long r = (long long)a + b;
    return (int)(r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : r));
}
SVPF_HD inline bool pow2_4_32(int v) { return v == 4 || v == 8 || v == 16 || v == 32; }
// A byte of a plane with the reference's slice semantics (`get(i).unwrap_or(fb)`).
SVPF_HD inline int rd(const uint8_t* p, int64_t len, int64_t i, int fb) {
    return (i >= 0 && i < len) ? (int)p[i] : fb;
}

// Sum of |H4 D H4^T| over one 4x4 tile (analyse_search.rs:2681-2702).
SVPF_HD inline uint32_t had4(const int d[16]) {
    int t[16];
SVPF_UNROLL
    for (int i = 0; i < 4; i++) {
        int s0 = d[i * 4 + 0] + d[i * 4 + 1], s1 = d[i * 4 + 0] - d[i * 4 + 1];
        int s2 = d[i * 4 + 2] + d[i * 4 + 3], s3 = d[i * 4 + 2] - d[i * 4 + 3];
        t[i * 4 + 0] = s0 + s2; t[i * 4 + 1] = s1 + s3; t[i * 4 + 2] = s0 - s2; t[i * 4 + 3] = s1 - s3;
    }
    int s = 0;
SVPF_UNROLL
    for (int j = 0; j < 4; j++) {
        int s0 = t[j] + t[4 + j], s1 = t[j] - t[4 + j], s2 = t[8 + j] + t[12 + j], s3 = t[8 + j] - t[12 + j];
        s += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3);
    }
    return (uint32_t)max(s, 0);
}


typedef uint32_t __attribute__((aligned(1))) u32u;
SVPF_HD inline uint32_t ld32(const uint8_t* p) { return *reinterpret_cast<const u32u*>(p); }

// Sum of |H4 D H4^T| for one 4x4 tile given four packed rows of source and reference.
SVPF_HD inline uint32_t had4_rows(const uint32_t a[4], const uint32_t b[4]) {
    int d[16];
SVPF_UNROLL
    for (int r = 0; r < 4; r++)
SVPF_UNROLL
        for (int c = 0; c < 4; c++) d[r * 4 + c] = (int)((a[r] >> (8 * c)) & 255) - (int)((b[r] >> (8 * c)) & 255);
    return had4(d);
}

// The block being searched (uniform across the wave).
struct Blk {
    // one field per plane, no arrays: a plane chosen per lane must compile to selects
    // (an array indexed that way would place the descriptor in scratch memory)
    const uint8_t *sp0, *sp1, *sp2;
    const uint8_t *rp0, *rp1, *rp2;
    int ss0, ss1, ss2, rs0, rs1, rs2;
    int sl0, sl1, sl2, rl0, rl1, rl2;  // plane byte lengths (< 2^31)
    int level, lw, lh, y_off, c_off, clw, clh, pel;
    int px, py, bw, bh;
    bool satd;
};

// ---- whole-candidate cost for the paths near the level borders (rare, so compact:
// byte reads with 32-bit offsets; plane lengths stay below 2^31, which the host checks)

// A byte of a plane with the reference's slice semantics (`get(i).unwrap_or(fb)`).
SVPF_HD inline int rdb(const uint8_t* p, int len, int i, int fb) {
    return (unsigned)i < (unsigned)len ? (int)p[i] : fb;
}

// Sum of |H4 D H4^T| over the n x n tiles of a block whose pixel (x, y) difference
// is diff(x, y) (satd_edge :2305-2351, satd_luma_clamp :2353-2436, satd_interior).
template <class Diff>
SVPF_HD inline uint32_t satd_tiles(int n, Diff diff) {
    uint32_t total = 0;
    const int nt = n >> 2;
SVPF_NOUNROLL
    for (int t = 0; t < nt * nt; t++) {
        const int ty = t / nt, tx = t - ty * nt;
        int d[16];
SVPF_UNROLL
        for (int rr = 0; rr < 4; rr++)
SVPF_UNROLL
            for (int c = 0; c < 4; c++) d[rr * 4 + c] = diff(tx * 4 + c, ty * 4 + rr);
        total = sat_add(total, had4(d));
    }
    return total >> 1;
}

// Candidate cost (luma, 4x chroma) by the reference's path selection, computed whole
// by one lane: out-of-bounds mv -> block_cost_edge_lc (:1996-2167); otherwise
// block_cost_luma_interior (:1719-1852) + chroma_sad_x4 (:2169-2303).  group_cost
// uses it for the candidates its split path does not cover.
SVPF_HD inline void cand_cost_full(const Blk& b, int mvx, int mvy, bool chroma, uint32_t& L, uint32_t& C) {
    const int pel = b.pel;
    const int b0 = -pel * b.px, b1 = -pel * b.py;
    const int b2 = pel * (b.lw - b.px - b.bw) + 1, b3 = pel * (b.lh - b.py - b.bh) + 1;
    const int ox = mvx < b0 ? mvx - b0 : ((b2 > b0 && mvx >= b2) ? mvx + 1 - b2 : 0);
    const int oy = mvy < b1 ? mvy - b1 : ((b3 > b1 && mvy >= b3) ? mvy + 1 - b3 : 0);
    const bool sq = b.satd && pow2_4_32(b.bw) && pow2_4_32(b.bh) && b.bw == b.bh;
    const bool subpel = b.level == 0 && pel >= 2;
    const int sub = subpel ? (mvy & 1) * pel + (mvx & 1) : 0;
    const uint8_t *S0 = b.sp0, *R0 = b.rp0, *S1 = b.sp1, *R1 = b.rp1, *S2 = b.sp2, *R2 = b.rp2;
    const int sl0 = b.sl0, rl0 = b.rl0, sl1 = b.sl1, rl1 = b.rl1, sl2 = b.sl2, rl2 = b.rl2;
    const int ss0 = b.ss0, rs0 = b.rs0, ss1 = b.ss1, rs1 = b.rs1, ss2 = b.ss2, rs2 = b.rs2;
    const int rrow = b.y_off + sub * b.lh;
    if (ox != 0 || oy != 0) {
        // ---- edge path (edge_shift_origins :1638-1676 + block_cost_edge_lc): every
        // coordinate clamped into its plane
        const int cx = clampi(b.px - ox, 0, max(b.lw - b.bw, 0));
        const int cy = clampi(b.py - oy, 0, max(b.lh - b.bh, 0));
        int rx0, ry0;
        if (pel == 1) { rx0 = cx + mvx; ry0 = cy + mvy; }
        else if (pel == 4) { rx0 = (4 * cx + mvx) >> 2; ry0 = (4 * cy + mvy) >> 2; }
        else { rx0 = (2 * cx + mvx) >> 1; ry0 = (2 * cy + mvy) >> 1; }
        const int mw = max(b.lw - 1, 0), mh = max(b.lh - 1, 0);
        auto ldiff = [&](int x, int y) {
            return rdb(S0, sl0, (b.y_off + clampi(cy + y, 0, mh)) * ss0 + clampi(cx + x, 0, mw), 0) -
                   rdb(R0, rl0, (rrow + clampi(ry0 + y, 0, mh)) * rs0 + clampi(rx0 + x, 0, mw), 0);
        };
        if (sq) {
            L = satd_tiles(b.bw, ldiff);
        } else {
            uint32_t sad = 0;
SVPF_NOUNROLL
            for (int i = 0; i < b.bw * b.bh; i++) sad += (uint32_t)abs(ldiff(i % b.bw, i / b.bw));
            L = sad;
        }
        C = 0;
        if (chroma) {
            const int cw = b.bw / 2, ch = b.bh / 2;
            const int ccx = (b.px >> 1) - ((b.px - cx) >> 1), ccy = (b.py >> 1) - ((b.py - cy) >> 1);
            const int crx0 = ((mvx >> 1) + 2 * ccx) >> 1, cry0 = ((mvy >> 1) + 2 * ccy) >> 1;
            const int cref = b.c_off + sub * b.clh;
            const int cmw = max(b.clw - 1, 0), cmh = max(b.clh - 1, 0);
            uint32_t uv = 0;
            if (b.satd && cw == ch && (cw == 4 || cw == 8 || cw == 16)) {
                auto udiff = [&](int x, int y) {
                    return rdb(S1, sl1, (b.c_off + clampi(ccy + y, 0, cmh)) * ss1 + clampi(ccx + x, 0, cmw), 0) -
                           rdb(R1, rl1, (cref + clampi(cry0 + y, 0, cmh)) * rs1 + clampi(crx0 + x, 0, cmw), 0);
                };
                auto vdiff = [&](int x, int y) {
                    return rdb(S2, sl2, (b.c_off + clampi(ccy + y, 0, cmh)) * ss2 + clampi(ccx + x, 0, cmw), 0) -
                           rdb(R2, rl2, (cref + clampi(cry0 + y, 0, cmh)) * rs2 + clampi(crx0 + x, 0, cmw), 0);
                };
                uv = sat_add(satd_tiles(cw, udiff), satd_tiles(cw, vdiff));
            } else {
SVPF_NOUNROLL
                for (int i = 0; i < cw * ch; i++) {
                    const int col = i % cw, row = i / cw;
                    const int so = b.c_off + clampi(ccy + row, 0, cmh), ro = cref + clampi(cry0 + row, 0, cmh);
                    const int sx = clampi(ccx + col, 0, cmw), rx = clampi(crx0 + col, 0, cmw);
                    uv += (uint32_t)abs(rdb(S1, sl1, so * ss1 + sx, 128) - rdb(R1, rl1, ro * rs1 + rx, 128));
                    uv += (uint32_t)abs(rdb(S2, sl2, so * ss2 + sx, 128) - rdb(R2, rl2, ro * rs2 + rx, 128));
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
    const int frx = ipx + mxp, fry = ipy + myp;
    if (ipx >= 0 && ipy >= 0 && ipx + b.bw <= b.lw && ipy + b.bh <= b.lh && frx >= 0 && fry >= 0 &&
        frx + b.bw <= b.lw && fry + b.bh <= b.lh && pow2_4_32(b.bw) && b.bw == b.bh) {
        // satd_interior / sad_u8_simd read exactly these rows
        auto ddiff = [&](int x, int y) {
            return rdb(S0, sl0, (b.y_off + ipy + y) * ss0 + ipx + x, 0) - rdb(R0, rl0, (rrow + fry + y) * rs0 + frx + x, 0);
        };
        if (sq) {
            L = satd_tiles(b.bw, ddiff);
        } else {
            uint32_t sad = 0;
SVPF_NOUNROLL
            for (int i = 0; i < b.bw * b.bh; i++) sad += (uint32_t)abs(ddiff(i % b.bw, i / b.bw));
            L = sad;
        }
    } else if (sq) {  // satd_luma_clamp (:2353-2436): source outside the level reads 0, reference clamped
        L = satd_tiles(b.bw, [&](int x, int y) {
            const int sy = ipy + y, sx = ipx + x;
            const int a = (sy >= 0 && sy < b.lh && sx >= 0 && sx < b.lw) ? rdb(S0, sl0, (b.y_off + sy) * ss0 + sx, 0) : 0;
            return a - rdb(R0, rl0, (rrow + clampi(sy + myp, 0, b.lh - 1)) * rs0 + clampi(sx + mxp, 0, b.lw - 1), 0);
        });
    } else {  // SAD with the scalar loop's row/column breaks
        uint32_t sad = 0;
SVPF_NOUNROLL
        for (int row = 0; row < b.bh; row++) {
            const int sy = ipy + row;
            if (sy < 0 || sy >= b.lh) break;
            const int ry = clampi(ipy + row + myp, 0, b.lh - 1);
SVPF_NOUNROLL
            for (int col = 0; col < b.bw; col++) {
                const int sx = ipx + col;
                if (sx < 0 || sx >= b.lw) break;
                const int rx = clampi(ipx + col + mxp, 0, b.lw - 1);
                sad += (uint32_t)abs(rdb(S0, sl0, (b.y_off + sy) * ss0 + sx, 0) - rdb(R0, rl0, (rrow + ry) * rs0 + rx, 0));
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
            const int so = b.c_off + cpy, ro = cref + ry0;
            uint32_t vu, vv;
            if (b.satd) {  // satd_interior over cw x cw
                vu = satd_tiles(cw, [&](int x, int y) {
                    return rdb(S1, sl1, (so + y) * ss1 + cpx + x, 0) - rdb(R1, rl1, (ro + y) * rs1 + rx0 + x, 0);
                });
                vv = satd_tiles(cw, [&](int x, int y) {
                    return rdb(S2, sl2, (so + y) * ss2 + cpx + x, 0) - rdb(R2, rl2, (ro + y) * rs2 + rx0 + x, 0);
                });
            } else {
                vu = 0;
                vv = 0;
SVPF_NOUNROLL
                for (int i = 0; i < cw * ch; i++) {
                    const int x = i % cw, y = i / cw;
                    vu += (uint32_t)abs(rdb(S1, sl1, (so + y) * ss1 + cpx + x, 0) - rdb(R1, rl1, (ro + y) * rs1 + rx0 + x, 0));
                    vv += (uint32_t)abs(rdb(S2, sl2, (so + y) * ss2 + cpx + x, 0) - rdb(R2, rl2, (ro + y) * rs2 + rx0 + x, 0));
                }
            }
            C = sat_mul(sat_add(vu, vv), 4);
            return;
        }
    }
    uint32_t uv = 0;
SVPF_NOUNROLL
    for (int row = 0; row < ch; row++) {
        const int sy = cpy + row;
        if (sy >= b.clh) break;
        const int ry = clampi(cpy + row + mcy, 0, b.clh - 1);
SVPF_NOUNROLL
        for (int col = 0; col < cw; col++) {
            const int sx = cpx + col;
            if (sx >= b.clw) break;
            const int rx = clampi(cpx + col + mcx, 0, b.clw - 1);
            const int si = b.c_off + sy, ri = cref + ry;
            uv += (uint32_t)abs(rdb(S1, sl1, si * ss1 + sx, 128) - rdb(R1, rl1, ri * rs1 + rx, 128));
            uv += (uint32_t)abs(rdb(S2, sl2, si * ss2 + sx, 128) - rdb(R2, rl2, ri * rs2 + rx, 128));
        }
    }
    C = sat_mul(uv, 4);
}


}  // namespace cost
}  // namespace svpf
