// Super pyramid build for gfx950 (replaces crates/svpflow1/src/super_build.rs:8-362).
//
// Two kernels, both writing every byte of their rows (so the super frame needs no
// clearing pass; the reference zeroes the plane first, super_filter.rs:210-223):
//   super_level0_kernel  one workgroup per (up to kL0Rows source rows, plane, frame):
//                        stages the full-pel rows y-2..y+3 of each (edge-padded to the
//                        stride) in LDS with 4-byte loads and writes those rows of every
//                        level-0 subplane (bilinear pel-2 interiors four pixels at a time).
//                        Fewer rows where their staging would exceed 64 KiB of LDS.
//   super_reduce_kernel  one workgroup per (destination row, plane, frame) of level
//                        l+1: vertical 6-tap into LDS (four columns per thread), then the
//                        horizontal 6-tap with the reference's first-tap rule, written
//                        over the full stride.
// Rows are written as 32-bit words when the row start is 4-byte aligned, bytes otherwise.
#include <algorithm>

#include "svpf_internal.h"

namespace svpf {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint8_t avg8(uint32_t a, uint32_t b) { return (uint8_t)((a + b + 1) >> 1); }
// avg8 of each of the four byte lanes of a and b: a + b = 2(a & b) + (a ^ b), so the
// rounded-up half is (a | b) - ((a ^ b) >> 1) per byte (the shifted xor masked to 7 bits).
__device__ __forceinline__ uint32_t avg8x4(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7F7F7F7Fu); }

typedef __attribute__((address_space(1))) const uint8_t g_u8;
typedef __attribute__((address_space(1))) const uint32_t __attribute__((aligned(1))) g_u32u;

struct L0Args {
    SuperGeom g;
    const Planes* src;
    const Planes* dst;
    int rows;  // source rows per workgroup, 1 .. kL0Rows (their staging fits 64 KiB of LDS)
};

// Full-pel region F(r, c) = src[r][min(c, w-1)] for c < stride (super_build.rs:45-67:
// copy then pad with the edge byte).  Rows y-2..y+3 come from LDS, others (only the
// first rows, read where the pel fill runs past the last row) from global memory.
// LDS bytes by address space: a generic pointer here let the compiler merge the LDS and
// global reads of F into one load through a selected pointer, held in scratch memory.
typedef __attribute__((address_space(3))) const uint8_t lds_u8;

struct Rows {
    lds_u8* lds;         // 6 rows of `pitch` bytes
    const uint8_t* sp;   // source plane
    int ss, pitch, y, h, w, ds;
    __device__ __forceinline__ uint8_t F(int r, int c) const {
        int k = r - y + 2;
        if (k >= 0 && k < 6) return lds[k * pitch + c];
        return sp[(size_t)r * ss + min(c, w - 1)];
    }
    // x-half subplane byte (super_build.rs:183-192); beyond w the row keeps its zero.
    // The read of x+1 follows the buffer: past the stride it is the next row's first
    // byte, and past the last row it is the first x-half byte.
    __device__ __forceinline__ uint8_t hp(int yy, int x) const {
        if (x >= w) return 0;
        uint8_t b;
        if (x + 1 < ds) b = F(yy, x + 1);
        else if (yy + 1 < h) b = F(yy + 1, 0);
        else b = avg8(F(0, 0), ds > 1 ? F(0, 1) : F(1, 0));
        return avg8(F(yy, x), b);
    }
};

constexpr int kL0Rows = 4;  // most source rows per workgroup of super_level0_kernel

__global__ void __launch_bounds__(kThreads) super_level0_kernel(L0Args a) {
    const int y0 = blockIdx.x * a.rows;
    const int plane = blockIdx.y;
    const int frame = blockIdx.z;
    const SuperGeom& g = a.g;
    const int w = g.level_w[plane][0], h = g.level_h[plane][0];
    if (y0 >= h) return;
    const int ny = min(a.rows, h - y0);
    // references into the descriptor array: a local copy indexed by `plane` would be
    // placed in scratch memory
    const Planes& src = a.src[frame];
    const Planes& dst = a.dst[frame];
    const int ds = dst.stride[plane];
    const int ss = src.stride[plane];
    const uint8_t* sp = src.p[plane];
    uint8_t* dp = dst.p[plane];

    extern __shared__ uint8_t lds[];
    const int pitch = ds + 4;
    // rows y0-2 .. y0+ny+2 (clamped), edge-padded to the stride: the words inside the
    // row with one (unaligned) 4-byte load per row, all rows' loads in flight, then the
    // padded tail by bytes.  Output row y0 + j reads LDS rows j .. j+5.
    const int nst = ny + 5;
    const int wfull = w & ~3;
    for (int c = threadIdx.x * 4; c < wfull; c += kThreads * 4) {
        uint32_t v[kL0Rows + 5];
#pragma unroll
        for (int k = 0; k < kL0Rows + 5; k++) {
            const int r = min(max(y0 - 2 + k, 0), h - 1);
            if (k < nst) v[k] = *(const g_u32u*)((const g_u8*)sp + ((size_t)r * ss + c));
        }
#pragma unroll
        for (int k = 0; k < kL0Rows + 5; k++) {
            if (k >= nst) break;
            uint8_t* L = lds + k * pitch + c;
            L[0] = (uint8_t)v[k];
            L[1] = (uint8_t)(v[k] >> 8);
            L[2] = (uint8_t)(v[k] >> 16);
            L[3] = (uint8_t)(v[k] >> 24);
        }
    }
    for (int k = 0; k < nst; k++) {
        const int r = min(max(y0 - 2 + k, 0), h - 1);
        const uint8_t* s = sp + (size_t)r * ss;
        for (int c = wfull + threadIdx.x; c < pitch; c += kThreads) lds[k * pitch + c] = s[min(c, w - 1)];
    }
    __syncthreads();
    const int pel = g.pel;
    const int nsub = g.full ? pel * pel : 1;
    const bool fill = pel >= 2 && g.gpu <= 1 && g.full;
    const bool bilinear = fill && g.scale_up == 0;
    const bool bicubic = fill && g.scale_up == 2;
    const bool aligned4 = ((((uintptr_t)dp) | (uintptr_t)ds | (uintptr_t)pitch) & 3) == 0;
    for (int j = 0; j < ny; j++) {
    const int y = y0 + j;
    const Rows R{(lds_u8*)(lds + j * pitch), sp, ss, pitch, y, h, w, ds};

    // Bilinear pel 2 away from the edges (rows 1 .. h-2, pixels whose right neighbour
    // lies inside the row): the four subplanes four pixels at a time from words of rows y
    // and y+1 -- full F(y, x), x-half avg(F(y, x), F(y, x+1)), y-half avg(F(y, x),
    // F(y+1, x)), diagonal avg(F(y, x), F(y+1, x+1)) (super_build.rs:183-223 off its
    // first-row and past-the-end rules).  The rest of each row takes the loop below.
    int done = 0;  // columns [0, done) written here
    if (bilinear && pel == 2 && aligned4 && y >= 1 && y + 1 < h) {
        done = (w - 1) & ~3;  // x + 4 < w for every x of a word below `done`
        const int r0 = g.level_row[plane][0];
        const uint32_t* L0 = (const uint32_t*)(lds + (j + 2) * pitch);  // row y
        const uint32_t* L1 = (const uint32_t*)(lds + (j + 3) * pitch);  // row y + 1
        for (int c = threadIdx.x * 4; c < done; c += kThreads * 4) {
            const uint32_t a = L0[c >> 2], an = L0[(c >> 2) + 1], b = L1[c >> 2], bn = L1[(c >> 2) + 1];
            const uint32_t a1 = __builtin_amdgcn_alignbyte(an, a, 1), b1 = __builtin_amdgcn_alignbyte(bn, b, 1);
            *reinterpret_cast<uint32_t*>(dp + (size_t)(r0 + y) * ds + c) = a;
            *reinterpret_cast<uint32_t*>(dp + (size_t)(r0 + h + y) * ds + c) = avg8x4(a, a1);
            *reinterpret_cast<uint32_t*>(dp + (size_t)(r0 + 2 * h + y) * ds + c) = avg8x4(a, b);
            *reinterpret_cast<uint32_t*>(dp + (size_t)(r0 + 3 * h + y) * ds + c) = avg8x4(a, b1);
        }
    }
    for (int sub = 0; sub < nsub; sub++) {
        uint8_t* row = dp + (size_t)(g.level_row[plane][0] + sub * h + y) * ds;
        auto value = [&](int x) __attribute__((always_inline)) -> uint8_t {
            if (sub == 0) return R.F(y, x);
            if (sub > 3 || x >= w) return 0;
            if (bilinear) {
                if (sub == 1) return R.hp(y, x);
                if (sub == 2) {  // super_build.rs:194-209: row 0 is then replaced by the full row 0
                    if (y == 0) return R.F(0, x);
                    return avg8(R.F(y, x), (y + 1 < h) ? R.F(y + 1, x) : R.hp(0, x));
                }
                // sub == 3, super_build.rs:211-223: row 0 is the full row 0 shifted by one byte
                if (y == 0) return (x + 1 < ds) ? R.F(0, x + 1) : (h > 1 ? R.F(1, 0) : R.hp(0, 0));
                uint8_t b;
                if (x + 1 < ds) {
                    b = (y + 1 < h) ? R.F(y + 1, x + 1) : R.hp(0, x + 1);
                } else {  // stride == w: the diagonal read wraps to row y+2, column 0
                    b = (y + 2 < h) ? R.F(y + 2, 0) : R.hp(y + 2 - h, 0);
                }
                return avg8(R.F(y, x), b);
            }
            if (bicubic) {  // super_build.rs:95-177 (`half` per pixel)
                auto vert = [&](int xx) __attribute__((always_inline)) -> uint8_t {
                    if (y < 2 || y + 4 >= h) return (y + 1 == h) ? R.F(y, xx) : avg8(R.F(y, xx), R.F(y + 1, xx));
                    int v = (int)R.F(y - 2, xx) + (int)R.F(y + 3, xx) +
                            5 * (4 * (int)R.F(y, xx) - (int)R.F(y - 1, xx) + 4 * (int)R.F(y + 1, xx) -
                                 (int)R.F(y + 2, xx)) + 16;
                    return (uint8_t)min(255, max(0, v >> 5));
                };
                auto horiz = [&](auto get, int xx) __attribute__((always_inline)) -> uint8_t {
                    if (xx < 2 || xx + 4 >= w) return (xx + 1 == w) ? get(xx) : avg8(get(xx), get(xx + 1));
                    int v = (int)get(xx - 2) + (int)get(xx + 3) +
                            5 * (4 * (int)get(xx) - (int)get(xx - 1) + 4 * (int)get(xx + 1) - (int)get(xx + 2)) + 16;
                    return (uint8_t)min(255, max(0, v >> 5));
                };
                if (sub == 1) return horiz([&](int xx) __attribute__((always_inline)) { return R.F(y, xx); }, x);
                if (sub == 2) return vert(x);
                return horiz(vert, x);
            }
            return 0;
        };
        if ((((uintptr_t)row) & 3) == 0) {
            for (int c = done + threadIdx.x * 4; c < ds; c += kThreads * 4) {
                if (c + 4 <= ds) {
                    uint32_t wv = (uint32_t)value(c) | ((uint32_t)value(c + 1) << 8) | ((uint32_t)value(c + 2) << 16) |
                                  ((uint32_t)value(c + 3) << 24);
                    *reinterpret_cast<uint32_t*>(row + c) = wv;
                } else {
                    for (int k = c; k < ds; k++) row[k] = value(k);
                }
            }
        } else {
            for (int c = done + threadIdx.x; c < ds; c += kThreads) row[c] = value(c);
        }
    }
    }
}

struct ReduceArgs {
    SuperGeom g;
    const Planes* dst;
    int level;  // source level; writes level+1
};

// reduce_6tap + horizontal_6tap_inplace (super_build.rs:254-357) for one destination row.
__global__ void __launch_bounds__(kThreads) super_reduce_kernel(ReduceArgs a) {
    const int dy = blockIdx.x;
    const int plane = blockIdx.y;
    const int frame = blockIdx.z;
    const SuperGeom& g = a.g;
    const int l = a.level;
    const int src_w = g.level_w[plane][l], src_h = g.level_h[plane][l];
    const int dst_w = g.level_w[plane][l + 1], dst_h = g.level_h[plane][l + 1];
    if (dy >= dst_h) return;
    const Planes& P = a.dst[frame];
    const int ds = P.stride[plane];
    uint8_t* base = P.p[plane];
    const uint8_t* srow0 = base + (size_t)g.level_row[plane][l] * ds;
    uint8_t* drow = base + (size_t)(g.level_row[plane][l + 1] + dy) * ds;

    extern __shared__ uint8_t inter[];  // src_w bytes: the vertically filtered row
    const int sy = dy * 2;
    const bool use_edge = dy == 0 || dy + 1 == dst_h || sy - 2 < 0 || sy + 3 >= src_h;
    // four columns per thread from 4-byte loads: the 6-tap in two 16-bit lanes (even and
    // odd bytes; a lane's sum is at most 255 * 32 + 16 = 8176, so the reference's
    // min(255, .) never applies), the edge rows' average bytewise; the rest below
    const int wq = src_w & ~3;
    for (int x = threadIdx.x * 4; x < wq; x += kThreads * 4) {
        uint32_t v;
        if (use_edge) {
            const int y0 = min(sy, src_h - 1), y1 = min(sy + 1, src_h - 1);
            v = avg8x4(*(const g_u32u*)((const g_u8*)srow0 + ((size_t)y0 * ds + x)),
                       *(const g_u32u*)((const g_u8*)srow0 + ((size_t)y1 * ds + x)));
        } else {
            const g_u8* c = (const g_u8*)srow0 + ((size_t)(sy - 2) * ds + x);
            uint32_t q[6];
#pragma unroll
            for (int k = 0; k < 6; k++) q[k] = *(const g_u32u*)(c + (size_t)k * ds);
            auto tap = [&](int sh) {
                uint32_t e[6];
#pragma unroll
                for (int k = 0; k < 6; k++) e[k] = (q[k] >> sh) & 0x00FF00FFu;
                const uint32_t t = e[0] + e[5] + 5u * (e[1] + e[4]) + 10u * (e[2] + e[3]) + 0x00100010u;
                return ((t >> 5) & 0x00FF00FFu) << sh;
            };
            v = tap(0) | tap(8);
        }
        *reinterpret_cast<uint32_t*>(inter + x) = v;
    }
    for (int x = wq + threadIdx.x; x < src_w; x += kThreads) {
        uint8_t v;
        if (use_edge) {
            int y0 = min(sy, src_h - 1), y1 = min(sy + 1, src_h - 1);
            v = avg8(srow0[(size_t)y0 * ds + x], srow0[(size_t)y1 * ds + x]);
        } else {
            const uint8_t* c = srow0 + (size_t)(sy - 2) * ds + x;
            uint32_t p0 = c[0], p1 = c[ds], p2 = c[2 * ds], p3 = c[3 * ds], p4 = c[4 * ds], p5 = c[5 * ds];
            v = (uint8_t)min(255u, (p0 + p5 + 5 * (p1 + p4) + 10 * (p2 + p3) + 16) >> 5);
        }
        inter[x] = v;
    }
    __syncthreads();
    auto s = [&](int i) -> uint32_t { return inter[i]; };
    auto out = [&](int j) -> uint8_t {
        if (j >= src_w) return 0;           // beyond the reduced row: never written (zero)
        if (j >= dst_w) return inter[j];    // vertical-only bytes (super_build.rs:303-305)
        if (dst_w <= 2) {                   // super_build.rs:314-320
            uint32_t x0 = (2 * j < src_w) ? s(2 * j) : 0;
            uint32_t x1 = (2 * j + 1 < src_w) ? s(2 * j + 1) : x0;
            return avg8(x0, x1);
        }
        uint32_t first = avg8(s(0), src_w > 1 ? s(1) : s(0));
        if (j == 0) return (uint8_t)first;
        if (j == dst_w - 1) {               // super_build.rs:350-356
            uint32_t x0 = (2 * j < src_w) ? s(2 * j) : 0;
            uint32_t x1 = (2 * j + 1 < src_w) ? s(2 * j + 1) : x0;
            return avg8(x0, x1);
        }
        // interior j = i + 1, k = 2j + 3; left = out[0] for j == 1 else s[2j-2] (super_build.rs:327-348)
        int k = 2 * j + 3;
        uint32_t left = (j == 1) ? first : s(2 * j - 2);
        uint32_t pk = (k < src_w) ? s(k) : s(src_w - 1);
        uint32_t v = (left + pk + 10 * (s(k - 3) + s(k - 2)) + 5 * (s(k - 4) + s(k - 1)) + 16) >> 5;
        return (uint8_t)min(255u, v);
    };
    if ((((uintptr_t)drow) & 3) == 0) {
        for (int c = threadIdx.x * 4; c < ds; c += kThreads * 4) {
            if (c + 4 <= ds) {
                uint32_t wv = (uint32_t)out(c) | ((uint32_t)out(c + 1) << 8) | ((uint32_t)out(c + 2) << 16) |
                              ((uint32_t)out(c + 3) << 24);
                *reinterpret_cast<uint32_t*>(drow + c) = wv;
            } else {
                for (int k = c; k < ds; k++) drow[k] = out(k);
            }
        }
    } else {
        for (int c = threadIdx.x; c < ds; c += kThreads) drow[c] = out(c);
    }
}

// Levels that the reduce never reaches (scale_down != 4) stay zero, like the
// reference's pre-cleared plane.
__global__ void __launch_bounds__(kThreads) super_zero_rows_kernel(SuperGeom g, const Planes* dst, int level) {
    const int dy = blockIdx.x, plane = blockIdx.y, frame = blockIdx.z;
    if (dy >= g.level_h[plane][level]) return;
    const Planes P = dst[frame];
    uint8_t* row = P.p[plane] + (size_t)(g.level_row[plane][level] + dy) * P.stride[plane];
    for (int c = threadIdx.x; c < P.stride[plane]; c += kThreads) row[c] = 0;
}

}  // namespace

hipError_t launch_super_level0(const SuperGeom& g, const Planes* src, const Planes* dst, int nframes,
                               int max_stride, hipStream_t s) {
    const int pitch = max_stride + 4;
    const int rows = std::max(1, std::min(kL0Rows, 65536 / pitch - 5));
    L0Args a{g, src, dst, rows};
    dim3 grid((g.level_h[0][0] + rows - 1) / rows, 3, nframes);
    size_t lds = (size_t)(rows + 5) * pitch;
    hipLaunchKernelGGL(super_level0_kernel, grid, dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_super_reduce(const SuperGeom& g, const Planes* dst, int nframes, int level,
                                       hipStream_t s) {
    dim3 grid(g.level_h[0][level + 1], 3, nframes);
    if (g.scale_down == 4) {
        ReduceArgs a{g, dst, level};
        size_t lds = (size_t)g.level_w[0][level] + 4;
        hipLaunchKernelGGL(super_reduce_kernel, grid, dim3(kThreads), lds, s, a);
    } else {
        hipLaunchKernelGGL(super_zero_rows_kernel, grid, dim3(kThreads), 0, s, g, dst, level + 1);
    }
    return hipGetLastError();
}


}  // namespace svpf
