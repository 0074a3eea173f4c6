"""Closed-form pins of the oracle's small arithmetic pieces, each restated here from
the reference's definition in Python (big integers, exact fractions, the host libm's
pow through math.pow) and compared on edge values and random inputs:

  * TruncDiv (crates/svpflow-core/src/renderer.rs:3149-3184): truncating division for
    every divisor the render builds and at the reciprocal's range edges;
  * exact_global_doubled (crates/svpflow1/src/analyse_search.rs:1285-1324): brute force
    with ties, negative sums and an empty window;
  * rescale_scores (:2716-2738) at the ratio boundaries 15/16/50/51 and at saturation;
  * exact_lambda (:1279-1283) and adapt_lambda (:1555-1564) including u32::MAX scores
    and the 21,474,836.47 cap;
  * luma_byte and classify_scene (crates/svpflow-core/src/metadata.rs:300-347,
    523-526) with score maps built at every threshold and one either side;
  * the coverage mask (renderer.rs:1738-1763, 2945-3147) against a naive splat of each
    block's moved footprint onto the grid and a 3x3 box sum;
  * the magnitude mask (renderer.rs:1647-1678, 3245-3259) for exponents other than 1.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from test_oracle_search_rules import (MAXU, adapt_lambda, as_i32, exact_lambda, f2i, global_doubled, rescale,
                                      tdiv)


@pytest.fixture(scope="module")
def L(oracle):
    lib = oracle.L
    i32p = C.POINTER(C.c_int32)
    lib.oracle_truncdiv.argtypes = [C.c_int, C.c_int]
    lib.oracle_coverage.argtypes = [C.c_int] * 7 + [i32p, i32p] + [C.c_int] * 5 + [C.POINTER(C.c_uint8)]
    lib.oracle_magnitude.argtypes = [C.c_int] * 4 + [i32p, i32p] + [C.c_int] * 3 + [C.c_double, C.c_double,
                                                                                   C.POINTER(C.c_uint8)]
    lib.oracle_luma_byte.argtypes = [C.c_double, C.c_double]
    lib.oracle_scene_class.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_double] + [C.c_int64] * 5
    return lib


# ------------------------------------------------------------------------------ TruncDiv

def test_truncdiv_is_truncating_division(L):
    rng = np.random.default_rng(1)
    divisors = list(range(1, 41)) + [256, 512, 768, 1024, 2048, 4096, 3 << 8, 7 << 8, 1 << 20, 999_983,
                                     (1 << 31) - 1] + [int(x) for x in rng.integers(1, 1 << 31, 40)]
    for d in divisors:
        magic = -(-(1 << 40) // d)
        bound = min((1 << 40) // d, (2 ** 64 - 1) // magic, 1 << 31)
        vals = {0, 1, -1, 2 ** 31 - 1, -2 ** 31, d, -d, d - 1, -(d - 1), d + 1, -(d + 1)}
        for b in (bound - 1, bound, bound + 1):
            vals |= {b, -b}
        vals |= {int(x) for x in rng.integers(-2 ** 31, 2 ** 31, 200)}
        vals |= {int(x) for x in rng.integers(-4096 * 1024, 4096 * 1024, 200)}
        for v in vals:
            if -2 ** 31 <= v < 2 ** 31:
                assert L.oracle_truncdiv(d, v) == tdiv(v, d), (d, v)
    for v in (0, 5, -5, 2 ** 31 - 1):
        assert L.oracle_truncdiv(0, v) == 0
        assert L.oracle_truncdiv(-7, v) == tdiv(v, -7)


# ------------------------------------------------------------------------------ global predictor

def _global(L, vecs):
    a = np.array(vecs, np.int16).reshape(-1)
    gx, gy = C.c_int32(), C.c_int32()
    L.oracle_global(a.ctypes.data_as(C.POINTER(C.c_int16)), len(vecs), C.byref(gx), C.byref(gy))
    return gx.value, gy.value


def test_global_predictor_brute_force(L):
    cases = [
        [(3, 4)],
        [(0, 0), (0, 0), (10, 10), (20, 10), (30, 10)],              # modes (0, 10): window empty
        [(1, 1), (2, 2), (1, 2), (2, 1)],                            # ties: smallest value wins
        [(-3, -7), (-3, -7), (-4, -8), (-9, 5), (-2, -7)],           # negative sums round away from zero
        [(5, 5), (10, 5), (11, 5), (-1, 5), (0, 5)],                 # window edges |d| = 5 in, 6 out
        [(-1, 1), (-2, 2), (-1, 1), (-2, 2), (-3, 3)],               # rounded mean at .5 for negatives
        [(7, -7)] * 3 + [(8, -8)] * 3,
    ]
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = int(rng.integers(1, 60))
        spread = int(rng.choice([2, 6, 20, 300]))
        cases.append([(int(rng.integers(-spread, spread + 1)), int(rng.integers(-spread, spread + 1)))
                      for _ in range(n)])
    for vecs in cases:
        assert _global(L, vecs) == global_doubled(vecs), vecs


# ------------------------------------------------------------------------------ rescale_scores

def test_rescale_scores_boundaries(L):
    w, h, blk = 1920, 1080, 8
    diag = int(math.sqrt(w * w + h * h))                  # 2202
    denom = diag * blk
    rows = []
    for ratio in (0, 14, 15, 16, 17, 30, 49, 50, 51, 52, 200):
        mag = -(-ratio * denom // 100)                    # smallest |dx|+|dy| with this ratio
        for m in (mag - 1, mag, mag + 1):
            m = max(m, 0)
            for dx in (m // 2, -(m - m // 2)):
                dy = m - abs(dx)
                if abs(dx) > 32767 or abs(dy) > 32767:
                    continue
                for score in (0, 1, 31, 32, 1000, 0xFFFFFF, 0x1000000, 2 ** 31 - 1, 2 ** 31, MAXU, 107374183,
                              107374182):
                    rows.append((dx, dy, score))
    rng = np.random.default_rng(3)
    for _ in range(2000):
        rows.append((int(rng.integers(-3000, 3000)), int(rng.integers(-3000, 3000)), int(rng.integers(0, 2 ** 32))))
    dxdy = np.array([(r[0], r[1]) for r in rows], np.int16).reshape(-1)
    sc = np.array([r[2] for r in rows], np.uint32)
    L.oracle_rescale(dxdy.ctypes.data_as(C.POINTER(C.c_int16)), sc.ctypes.data_as(C.POINTER(C.c_uint32)), len(rows),
                     w, h, blk)
    for r, got in zip(rows, sc):
        assert int(got) == rescale(r[2], r[0], r[1], w, h, blk), r
    # the 2/3 small frame: diag from the f64 square root
    rows = [(dx, 0, 5000) for dx in range(0, 40)]
    dxdy = np.array([(r[0], r[1]) for r in rows], np.int16).reshape(-1)
    sc = np.array([r[2] for r in rows], np.uint32)
    L.oracle_rescale(dxdy.ctypes.data_as(C.POINTER(C.c_int16)), sc.ctypes.data_as(C.POINTER(C.c_uint32)), len(rows),
                     64, 48, 4)
    assert [int(x) for x in sc] == [rescale(r[2], r[0], r[1], 64, 48, 4) for r in rows]


# ------------------------------------------------------------------------------ lambdas

def test_lambda_closed_forms(L):
    ex, ad = C.c_int32(), C.c_int32()
    rng = np.random.default_rng(4)
    cases = []
    for lam in (0, 1, 20000, 64000, 2 ** 31 - 1, -20000):
        for plevel in (1.5, 1.0, 0.01, 2.0, 1.3, 0.7, 2.5, 1.1):
            for levels, level in ((4, 0), (4, 1), (4, 3), (8, 0), (1, 0), (5, 2)):
                for pel in (1, 2, 4):
                    for lsad in (0, 1, 2, 16000, 64000):
                        for score in (0, 1, 2, 3, 16000, 0xFFFFFF, MAXU - 1, MAXU):
                            cases.append((lam, plevel, levels, level, pel, lsad, score))
    for _ in range(3000):
        cases.append((int(rng.integers(-2 ** 20, 2 ** 31)), float(rng.uniform(0.01, 4)), int(rng.integers(1, 9)), 0,
                      int(rng.choice([1, 2, 4])), int(rng.integers(0, 2 ** 20)), int(rng.integers(0, 2 ** 32))))
    for lam, plevel, levels, level, pel, lsad, score in cases:
        L.oracle_lambdas(lam, plevel, levels, level, pel, lsad, score, C.byref(ex), C.byref(ad))
        e = exact_lambda(lam, plevel, levels, level, pel)
        assert ex.value == e, (lam, plevel, levels, level, pel)
        assert ad.value == adapt_lambda(e, lsad, score), (e, lsad, score)
    # the cap: a huge lambda with a zero score saturates at 21,474,836
    assert adapt_lambda(2 ** 31 - 1, 8000, 0) == 21474836
    # lsad 1 and an unsearched score divide by zero: inf is capped, NaN (0 * inf) too
    assert adapt_lambda(5, 1, MAXU) == 21474836 and adapt_lambda(0, 1, MAXU) == 21474836


# ------------------------------------------------------------------------------ scene classification

def _luma_byte(v, gamma):
    s = f2i(math.pow(v, gamma) * 255.0)
    return max(s, 20) & 0xFF


def test_luma_byte_closed_form(L):
    for gamma in (1.5, 0.0, 1.0, 2.2, 5.0, 0.37):
        for denom in (255.0, 510.0):
            for i in range(511):
                v = i / denom
                assert L.oracle_luma_byte(v, gamma) == _luma_byte(v, gamma), (gamma, denom, i)


def _classify(scores, lumas, zero, m1, m2, scene, blocks):
    considered = sc = m2c = m1c = 0
    for s, l in zip(scores, lumas):
        v = min(s * 255, 2 ** 31 - 1) // max(l, 1)
        if v < zero:
            continue
        considered += 1
        if v >= scene:
            sc += 1
        elif v >= m2:
            m2c += 1
        elif v >= m1:
            m1c += 1
    req = considered * blocks // 100
    if sc >= req:
        return 3
    if sc + m2c >= req:
        return 2
    return 1 if sc + m2c + m1c >= req else 0


def test_scene_class_at_thresholds(L):
    gamma = 1.5
    zero, m1, m2, scene = 400, 3200, 5600, 8000        # default limits scaled for 8x8 blocks (x2)
    gw, gh = 10, 10
    n = gw * gh
    rng = np.random.default_rng(6)
    lb = {}
    for a in range(256):                                 # block luma pairs -> luma_byte for both markers
        lb[a] = _luma_byte((a + a) / 255.0, gamma)

    def run(scores, luma, blocks, marker=2, prev_luma=None):
        cur = np.array([(s, l) for s, l in zip(scores, luma)], np.int32).reshape(-1)
        pl = luma if prev_luma is None else prev_luma
        prev = np.array([(0, l) for l in pl], np.int32).reshape(-1)
        got = L.oracle_scene_class(prev.ctypes.data_as(C.POINTER(C.c_int32)),
                                   cur.ctypes.data_as(C.POINTER(C.c_int32)), gw, gh, marker, gamma, zero, m1, m2,
                                   scene, blocks)
        den = 510.0 if marker == 3 else 255.0
        lm = [_luma_byte((a + b) / den, gamma) for a, b in zip(pl, luma)]
        return got, _classify(scores, lm, zero, m1, m2, scene, blocks)

    checked = 0
    for target in (zero, m1, m2, scene):
        for delta in (-1, 0, 1):
            for frac in (0, 1, 19, 20, 21, 50, 99, 100):
                # block lumas summing to 255 give luma_byte 255, so a block's weighted
                # score is its score: the blocks sit exactly at target + delta
                luma = [100] * n
                prev_luma = [155] * n
                assert _luma_byte(1.0, gamma) == 255
                scores = [target + delta if i < frac else int(rng.integers(0, zero)) for i in range(n)]
                for blocks in (0, 1, 20, 21, 50, 99, 100):
                    got, want_c = run(scores, luma, blocks, prev_luma=prev_luma)
                    assert got == want_c, (target, delta, frac, blocks)
                    checked += 1
    for _ in range(300):
        luma = [int(x) for x in rng.integers(0, 256, n)]
        scores = [int(x) for x in rng.integers(0, 20000, n)]
        got, want_c = run(scores, luma, int(rng.integers(0, 101)), marker=int(rng.choice([2, 3])))
        assert got == want_c
    assert checked > 500


# ------------------------------------------------------------------------------ coverage mask

def _naive_coverage(vecs, gw, gh, bw, bh, ox, oy, base, w, h, strength, threshold):
    """Every vector moves its block by threshold*v/(base*256) (truncated) and lays the
    block's area on the step grid: the cell holding the moved corner and the cells right,
    below and diagonal get the overlaps of the block with them; cells outside the w x h
    grid are dropped.  Then a 3x3 box sum, covered = sum >> 3, and the byte of the
    remaining area."""
    xs, ys = bw - ox, bh - oy
    acc = np.zeros((h, w), np.int64)
    for y in range(gh):
        for x in range(gw):
            dx, dy = vecs[y * gw + x][:2]
            sx = x * xs + tdiv(threshold * dx, base * 256)
            sy = y * ys + tdiv(threshold * dy, base * 256)
            left, top = sx // xs, sy // ys                       # floor division: the cell holding the corner
            lw = (left + 1) * xs - sx
            tw = (top + 1) * ys - sy
            for cx, wx in ((left, lw), (left + 1, bw - lw)):
                for cy, wy in ((top, tw), (top + 1, bh - tw)):
                    if 0 <= cx < w and 0 <= cy < h:
                        acc[cy, cx] += wx * wy
    pad = np.pad(acc, 1)
    box = sum(pad[1 + a:1 + a + h, 1 + b:1 + b + w] for a in (-1, 0, 1) for b in (-1, 0, 1))
    area = max(bw * bh, 1)
    out = np.zeros((h, w), np.uint8)
    for yy in range(h):
        for xx in range(w):
            rem = max(area - (int(box[yy, xx]) >> 3), 0)
            v = rem * (strength / 100.0) * 256.0 / area
            iv = f2i(v)
            out[yy, xx] = 255 if iv >= 255 else (iv & 0xFF)
    return out


@pytest.mark.parametrize("geom", [(8, 8, 4, 4, 2), (8, 8, 2, 2, 2), (16, 16, 8, 8, 1), (16, 8, 4, 2, 4),
                                  (4, 4, 0, 0, 2), (32, 32, 16, 16, 2)])
def test_coverage_mask_naive_splat(L, geom):
    bw, bh, ox, oy, base = geom
    rng = np.random.default_rng(bw * 100 + bh + ox + base)
    gw, gh = 9, 7
    for trial in range(12):
        w, h = gw + int(rng.integers(0, 2)), gh + int(rng.integers(0, 2))
        span = int(rng.choice([4, 40, 300]))
        a = [(int(rng.integers(-span, span + 1)), int(rng.integers(-span, span + 1)), 0) for _ in range(gw * gh)]
        b = [(int(rng.integers(-span, span + 1)), int(rng.integers(-span, span + 1)), 0) for _ in range(gw * gh)]
        threshold = int(rng.integers(0, 257))
        strength = int(rng.choice([100, 50, 150, 0]))
        for which in (0, 1):
            out = np.zeros((h, w), np.uint8)
            A = np.array(a, np.int32).reshape(-1)
            B = np.array(b, np.int32).reshape(-1)
            L.oracle_coverage(bw, bh, base, ox, oy, gw, gh, A.ctypes.data_as(C.POINTER(C.c_int32)),
                              B.ctypes.data_as(C.POINTER(C.c_int32)), which, w, h, strength, threshold,
                              out.ctypes.data_as(C.POINTER(C.c_uint8)))
            want = _naive_coverage(b if which == 0 else a, gw, gh, bw, bh, ox, oy, base, w, h, strength, threshold)
            assert np.array_equal(out, want), (geom, trial, which)


# ------------------------------------------------------------------------------ magnitude mask

@pytest.mark.parametrize("expo", [1.0, 0.5, 2.0, 0.37, 3.3])
def test_magnitude_mask_closed_form(L, expo):
    rng = np.random.default_rng(int(expo * 100))
    bw, bh, gw, gh, w, h = 8, 8, 7, 5, 8, 6
    area = bw * bh
    for scale in (100 / 15000.0, 1.0, 0.001, 37 / 15000.0):
        mags = [int(x) for x in rng.integers(0, 1 << 24, gw * gh)]
        # include the magnitudes around every output byte step
        mags[:12] = [0, 1, 2, 3, 15, 16, 17, 100, 1000, 2399, 2400, 2401]
        trip = np.array([(0, 0, m) for m in mags], np.int32).reshape(-1)
        out = np.zeros((h, w), np.uint8)
        L.oracle_magnitude(bw, bh, gw, gh, trip.ctypes.data_as(C.POINTER(C.c_int32)),
                           trip.ctypes.data_as(C.POINTER(C.c_int32)), 0, w, h, scale, expo,
                           out.ctypes.data_as(C.POINTER(C.c_uint8)))

        def byte(m):
            v = math.pow(4.0 * m * scale / area, expo) * 255.0
            if not math.isfinite(v) or v <= 0.0:
                return 0
            return 255 if v >= 255.0 else int(v)

        for y in range(h):
            for x in range(w):
                # cells past the grid repeat the left neighbour (or, in column 0, the one above)
                sx, sy = (min(x, gw - 1), y) if y < gh else (0, gh - 1)
                assert out[y, x] == byte(mags[sy * gw + sx]), (expo, scale, x, y)
