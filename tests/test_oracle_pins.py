"""Pins of the CPU oracle's Analyse and SmoothFps restatements.

The reference ships no tests, fixtures or golden vectors for this path, so the
oracle is pinned by facts the mathematics fixes:
  * block costs: textbook SAD and 4x4 Hadamard SATD (sum |H D H^T| / 2, summed per
    plane) computed here in numpy on random interior candidates, with the chroma
    cost 4 * (U + V) at the luma vector's chroma position, for pel 1 and pel 2;
    identical blocks cost 0;
  * the exhaustive search: a frame that is a pure integer translation of the
    other is matched by the exact shift, at cost 0, at every interior block, and
    the forward field is the negated backward one;
  * the renderer: identical sources with zero vectors give back the source for
    every algorithm (all blends of equal values are the value), and a static
    scene with the default limits is classified as a scene change (copies).
A dropped term, a wrong sign, a transposed operand or an off-by-one in the
sub-pel plane selection fails one of these.
"""
import numpy as np
import pytest

import synth

H4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]], dtype=np.int64)


def _satd4(d):
    return int(np.abs(H4 @ d @ H4.T).sum())


def _block_cost(a, b, satd):
    """Textbook cost of a block difference: SAD, or the sum over 4x4 tiles of
    |H D H^T| halved."""
    d = a.astype(np.int64) - b.astype(np.int64)
    if not satd:
        return int(np.abs(d).sum())
    h, w = d.shape
    tot = sum(_satd4(d[y:y + 4, x:x + 4]) for y in range(0, h, 4) for x in range(0, w, 4))
    return tot >> 1


def _frames(seed, w, h):
    rng = np.random.default_rng(seed)
    return tuple(synth.texture(int(rng.integers(1 << 30)), hh, ww) for hh, ww in ((h, w), (h // 2, w // 2),
                                                                                   (h // 2, w // 2)))


@pytest.mark.parametrize("pel", [1, 2])
@pytest.mark.parametrize("satd", [False, True])
def test_block_cost_is_textbook_sad_satd(oracle, pel, satd):
    w, h = 128, 96
    sopt = "{pel:%d}" % pel
    so, _ = oracle.super_opts(sopt, w, h)
    cur = oracle.super_build(sopt, _frames(1, w, h))
    ref = oracle.super_build(sopt, _frames(2, w, h))
    rng = np.random.default_rng(7 + pel + 2 * satd)
    checked = 0
    for _ in range(300):
        bw = int(rng.choice([8, 16]))
        px = int(rng.integers(0, (w - bw) // 4 + 1)) * 4
        py = int(rng.integers(0, (h - bw) // 4 + 1)) * 4
        mvx = int(rng.integers(-12 * pel, 12 * pel + 1))
        mvy = int(rng.integers(-12 * pel, 12 * pel + 1))
        # luma reference block: sub-pel plane (mvy & 1) * pel + (mvx & 1) at the integer part (pel 2)
        ix, iy = (mvx >> 1, mvy >> 1) if pel == 2 else (mvx, mvy)
        sub = ((mvy & 1) * pel + (mvx & 1)) if pel == 2 else 0
        rx, ry = px + ix, py + iy
        cw = bw // 2
        cx, cy = (mvx >> 2, mvy >> 2) if pel == 2 else (mvx >> 1, mvy >> 1)
        crx, cry = px // 2 + cx, py // 2 + cy
        if not (0 <= rx and rx + bw <= w and 0 <= ry and ry + bw <= h):
            continue
        # mv_bounds (analyse_search.rs:1620-1636): [-pel*p, pel*(size - p - b) + 1) keeps the interior path
        if not (-pel * px <= mvx < pel * (w - px - bw) + 1 and -pel * py <= mvy < pel * (h - py - bw) + 1):
            continue
        if not (0 <= crx and crx + cw <= w // 2 and 0 <= cry and cry + cw <= h // 2):
            continue
        L, C = oracle.candidate_cost(w, h, pel, so["levels"], cur, ref, 0, px, py, bw, bw, mvx, mvy, satd, True)
        want_l = _block_cost(cur[0][py:py + bw, px:px + bw],
                             ref[0][sub * h + ry:sub * h + ry + bw, rx:rx + bw], satd)
        csub = sub * (h // 2)
        u = _block_cost(cur[1][py // 2:py // 2 + cw, px // 2:px // 2 + cw],
                        ref[1][csub + cry:csub + cry + cw, crx:crx + cw], satd)
        v = _block_cost(cur[2][py // 2:py // 2 + cw, px // 2:px // 2 + cw],
                        ref[2][csub + cry:csub + cry + cw, crx:crx + cw], satd)
        assert (L, C) == (want_l, 4 * (u + v)), (px, py, bw, mvx, mvy)
        checked += 1
    assert checked > 100
    # identical frames: zero cost at the zero vector
    L, C = oracle.candidate_cost(w, h, pel, so["levels"], cur, cur, 0, 32, 32, 16, 16, 0, 0, satd, True)
    assert (L, C) == (0, 0)


@pytest.mark.parametrize("pel", [1, 2])
def test_exhaustive_search_finds_integer_translation(oracle, pel):
    w, h, sx, sy = 160, 128, 4, -2          # even shifts: chroma moves by whole pixels too
    big = _frames(11, w + 32, h + 32)
    a = tuple(p[8:8 + (h if i == 0 else h // 2), 8:8 + (w if i == 0 else w // 2)] for i, p in enumerate(big))
    b = tuple(p[8 + (sy if i == 0 else sy // 2):8 + (sy if i == 0 else sy // 2) + (h if i == 0 else h // 2),
                8 + (sx if i == 0 else sx // 2):8 + (sx if i == 0 else sx // 2) + (w if i == 0 else w // 2)]
              for i, p in enumerate(big))
    sopt = "{pel:%d}" % pel
    so, _ = oracle.super_opts(sopt, w, h)
    sa, sb = oracle.super_build(sopt, a), oracle.super_build(sopt, b)
    aopt = "{block:{w:16,h:16,overlap:0},main:{search:{type:4,distance:8,coarse:{type:4,distance:6}}," \
           "lambda:0,penalty:{pnew:0,pzero:0,pglobal:0,pnbour:0}}}"
    pay = oracle.analyse(aopt, so["sdata"], w, h, sa, sb)
    hdr, regions = oracle.decode_payload(pay)
    gw, gh = hdr[12], hdr[13]
    assert len(regions) == 2
    vecs = [r.reshape(gh, gw, -1) for r in regions]   # (dx, dy, score, luma) per block
    inner = (slice(1, gh - 1), slice(1, gw - 1))
    # one region holds +shift and the other -shift (which is which is the reference's
    # direction convention, not fixed by the translation)
    s0 = 1 if vecs[0][inner][..., 0].flat[0] == pel * sx else -1
    for r, sign in ((vecs[0], s0), (vecs[1], -s0)):
        blk = r[inner]
        assert np.all(blk[..., 0] == sign * pel * sx), np.unique(blk[..., 0])
        assert np.all(blk[..., 1] == sign * pel * sy), np.unique(blk[..., 1])
        assert np.all(blk[..., 2] == 0)


@pytest.mark.parametrize("algo", [1, 2, 11, 13, 21, 22, 23])
def test_render_of_static_scene_is_the_source(oracle, algo):
    w, h = 128, 96
    f = _frames(21, w, h)
    frames = [f, f, f, f]
    sopt = "{pel:2}"
    so, _ = oracle.super_opts(sopt, w, h)
    sups = [oracle.super_build(sopt, x) for x in frames]
    aopt = "{block:{w:8,h:8,overlap:2},main:{search:{type:3,distance:8,satd:true}}}"
    ao, _ = oracle.analyse_opts(aopt, so["sdata"], w, h)
    pays = [oracle.analyse(aopt, so["sdata"], w, h, sups[i], sups[min(i + 1, 3)]) for i in range(4)]
    # a static scene has every block score below the default zero limit, which
    # classify_scene (metadata.rs:300-347) counts as a scene change (no block
    # considered, none required): with limit 0 every block counts and the frames
    # are interpolated
    ropt = "{rate:{num:5,den:2},algo:%d,mask:{area:100},scene:{mode:0,limits:{zero:0}}}" % algo
    clip = oracle.SmoothClip(ropt, frames, sups, pays, so["sdata"], ao["vdata"], fps=(24, 1))
    nout = oracle.smooth_opts(ropt, 24, 1, ao["vdata"], 4)[0]["out_frames"]
    interp = 0
    for k in range(nout):
        out, kind, _ = clip.render(k)
        interp += kind == 2
        for got, want in zip(out, f):
            assert np.array_equal(got, want), (algo, k, kind)
    assert interp > 0
    # default limits: no block reaches the zero limit, so the scene class is 3 and every frame is a copy
    clip = oracle.SmoothClip(ropt.replace(",limits:{zero:0}", ""), frames, sups, pays, so["sdata"], ao["vdata"],
                             fps=(24, 1))
    assert all(clip.render(k)[1] in (0, 1) for k in range(nout))


@pytest.mark.parametrize("search", ["type:3,distance:16,satd:true", "type:2,distance:8"])
@pytest.mark.parametrize("pel", [1, 2])
def test_hierarchy_finds_integer_translation(oracle, search, pel):
    """The whole hierarchy with the default penalties (lambda 10, pnew 50, pzero 100, ...):
    a frame that is a pure integer translation of the other is matched by pel * shift at
    every block away from the border, at cost 0, and the two directions are negatives."""
    w, h, sx, sy = 192, 128, 6, -4
    big = _frames(31 + pel, w + 32, h + 32)
    a = tuple(p[8:8 + (h if i == 0 else h // 2), 8:8 + (w if i == 0 else w // 2)] for i, p in enumerate(big))
    b = tuple(p[8 + (sy if i == 0 else sy // 2):8 + (sy if i == 0 else sy // 2) + (h if i == 0 else h // 2),
                8 + (sx if i == 0 else sx // 2):8 + (sx if i == 0 else sx // 2) + (w if i == 0 else w // 2)]
              for i, p in enumerate(big))
    sopt = "{pel:%d}" % pel
    so, _ = oracle.super_opts(sopt, w, h)
    sa, sb = oracle.super_build(sopt, a), oracle.super_build(sopt, b)
    aopt = "{block:{w:8,h:8,overlap:2},main:{search:{%s}}}" % search
    hdr, regions = oracle.decode_payload(oracle.analyse(aopt, so["sdata"], w, h, sa, sb))
    gw, gh = hdr[12], hdr[13]
    vecs = [r.reshape(gh, gw, -1) for r in regions]
    m = 2                                          # blocks whose shifted block leaves the frame
    inner = (slice(m, gh - m), slice(m, gw - m))
    s0 = 1 if vecs[0][inner][..., 0].flat[0] == pel * sx else -1
    for r, sign in ((vecs[0], s0), (vecs[1], -s0)):
        blk = r[inner]
        assert np.all(blk[..., 0] == sign * pel * sx), np.unique(blk[..., 0], return_counts=True)
        assert np.all(blk[..., 1] == sign * pel * sy), np.unique(blk[..., 1], return_counts=True)
        assert np.all(blk[..., 2] == 0)


@pytest.mark.parametrize("algo", [1, 2, 11, 13, 21, 22, 23])
def test_float_render_of_static_scene_is_the_source(oracle, algo):
    """The float renderer (the reference's OpenCL kernel) on identical sources with zero
    vectors: every blend of equal values is the value, and the luma's trip through linear
    light and back rounds to the sample it started from."""
    w, h = 128, 96
    f = _frames(23, w, h)
    frames = [f, f, f, f]
    sopt = "{pel:2}"
    so, _ = oracle.super_opts(sopt, w, h)
    sups = [oracle.super_build(sopt, x) for x in frames]
    aopt = "{block:{w:8,h:8,overlap:2},main:{search:{type:3,distance:8,satd:true}}}"
    ao, _ = oracle.analyse_opts(aopt, so["sdata"], w, h)
    pays = [oracle.analyse(aopt, so["sdata"], w, h, sups[i], sups[min(i + 1, 3)]) for i in range(4)]
    ropt = "{rate:{num:5,den:2},algo:%d,mask:{area:100},scene:{mode:0,limits:{zero:0}}}" % algo
    clip = oracle.SmoothClip(ropt, frames, sups, pays, so["sdata"], ao["vdata"], fps=(24, 1), float_render=True)
    nout = oracle.smooth_opts(ropt, 24, 1, ao["vdata"], 4)[0]["out_frames"]
    interp = 0
    for k in range(nout):
        out, kind, _ = clip.render(k)
        interp += kind == 2
        for got, want in zip(out, f):
            assert np.array_equal(got, want), (algo, k, kind)
    assert interp > 0


def test_float_render_translation_moves_the_source(oracle):
    """Algorithm 2 of the float renderer displaces source n by phase/256 of each vector
    (in vector units; the reference's kernel takes pel as 1): a whole-pixel translation
    at pel 1 renders, away from the borders, the source shifted by the fraction of the
    shift the phase gives -- exact when that is a whole number of pixels."""
    w, h, sx, sy = 160, 128, 8, -4
    big = _frames(41, w + 32, h + 32)
    a = tuple(p[8:8 + (h if i == 0 else h // 2), 8:8 + (w if i == 0 else w // 2)] for i, p in enumerate(big))
    b = tuple(p[8 + (sy if i == 0 else sy // 2):8 + (sy if i == 0 else sy // 2) + (h if i == 0 else h // 2),
                8 + (sx if i == 0 else sx // 2):8 + (sx if i == 0 else sx // 2) + (w if i == 0 else w // 2)]
              for i, p in enumerate(big))
    frames = [a, b, b, b]
    sopt = "{pel:1}"
    so, _ = oracle.super_opts(sopt, w, h)
    sups = [oracle.super_build(sopt, x) for x in frames]
    aopt = "{block:{w:16,h:16,overlap:0},main:{search:{type:4,distance:8,coarse:{type:4,distance:6}}," \
           "lambda:0,penalty:{pnew:0,pzero:0,pglobal:0,pnbour:0}}}"
    ao, _ = oracle.analyse_opts(aopt, so["sdata"], w, h)
    pays = [oracle.analyse(aopt, so["sdata"], w, h, sups[i], sups[min(i + 1, 3)]) for i in range(4)]
    # 2x: output 1 is at phase 128 of pair (0, 1); algorithm 2 samples frame 0 at x + v/2
    ropt = "{rate:{num:2,den:1},algo:2,scene:{mode:0,limits:{zero:0}}}"
    clip = oracle.SmoothClip(ropt, frames, sups, pays, so["sdata"], ao["vdata"], fps=(24, 1), float_render=True)
    out, kind, algo = clip.render(1)
    assert kind == 2 and algo == 2
    hdr, regions = oracle.decode_payload(pays[0])
    # set a (the backward region) holds (sx, sy) away from the borders: frame 0 at
    # x + (sx, sy) / 2; the cubic vector sampling reaches two blocks, so the border
    # blocks' vectors stay out of the compared area
    m = 40
    got = out[0][m:h - m, m:w - m]
    want = a[0][m + sy // 2:h - m + sy // 2, m + sx // 2:w - m + sx // 2]
    # luma goes through linear light and back (exact after rounding); chroma moves by
    # half the luma shift
    assert np.array_equal(got, want)
    cm = m // 2
    for p in (1, 2):
        g = out[p][cm:h // 2 - cm, cm:w // 2 - cm]
        q = a[p][cm + sy // 4:h // 2 - cm + sy // 4, cm + sx // 4:w // 2 - cm + sx // 4]
        assert np.array_equal(g, q), p
