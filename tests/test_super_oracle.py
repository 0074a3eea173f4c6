"""Pins of the CPU oracle's Super restatement (super_build.rs / super_opts.rs).

The reference ships no tests or fixtures for Super, so the oracle is pinned by
closed forms (constant image, copied level 0, pel averages), an impulse response
that fixes the 6-tap weights and their placement, and a naive per-pixel
restatement in Python run on small random planes (stride == width and padded).
"""
import numpy as np
import pytest


def test_super_opts_defaults_and_errors(oracle):
    o, err = oracle.super_opts(None, 1920, 1080)
    assert err is None
    assert (o["pel"], o["gpu"], o["scale_down"], o["scale_up"], o["full"], o["levels"]) == (2, 1, 4, 0, 1, 9)
    assert o["super_h"] == 5390                      # SURVEY shape legend, 1080p pel 2
    assert oracle.super_opts(None, 640, 360)[0]["super_h"] == 1790
    assert oracle.super_opts("{pel:1}", 3840, 2160)[0]["super_h"] == 4310
    assert oracle.super_opts("{pel:3}", 64, 64)[1] == "SVSuper: pel has to be 1 or 2 or 4"
    assert oracle.super_opts("{gpu:5}", 64, 64)[1] == "SVSuper: gpu has to be 0 or 1 or 2"
    assert oracle.super_opts("{full:false}", 64, 64)[1] == "SVSuper: can't ignore finest level with high pel value"
    o, _ = oracle.super_opts("{gpu:0}", 64, 64)
    assert o["scale_up"] == 2
    # pack_data bit layout: H | W<<16 | pel<<32 | levels<<40 | gpu<<48 | flags<<56
    o, _ = oracle.super_opts("{pel:4, rc:true}", 100, 50)
    assert o["sdata"] == 50 | (100 << 16) | (4 << 32) | (o["levels"] << 40) | (1 << 48) | (1 << 56)


def test_params_parser_quirks(oracle):
    assert oracle.opt_lookup("{a:{b:3}}", "a.b", 0) == (True, 3)
    assert oracle.opt_lookup("{a:1, a:2}", "a", 0) == (True, 2)          # last duplicate wins
    assert oracle.opt_lookup("{a:-}", "a", 1) == (True, 0.0)              # lone '-' is 0
    assert oracle.opt_lookup("{a:-.5}", "a", 1) == (True, -0.5)
    assert oracle.opt_lookup("{a:2.9}", "a", 0) == (True, 2)               # truncation
    assert oracle.opt_lookup("{a:-2.9}", "a", 0) == (True, -2)
    assert oracle.opt_lookup("{a:1e300}", "a", 0) == (True, None)          # out of i64 range
    assert oracle.opt_lookup("{a:true}", "a", 0) == (True, 1)
    assert oracle.opt_lookup("{a:null}", "a", 2) == (True, False)
    assert oracle.opt_lookup('{a:"x"}', "a", 2) == (True, True)
    assert oracle.opt_lookup("{a:[]}", "a", 2) == (True, False)
    assert oracle.opt_lookup("/*c*/ {a: /*c*/ 4 // end\n}", "a", 0) == (True, 4)
    assert oracle.opt_lookup("{/*c*/ a:4}", "a", 0) == (True, 4)           # comment after '{' is skipped
    assert oracle.opt_lookup("{a:1, /*c*/ b:4}", "b", 0)[0] is False        # ...but not before a later member
    assert oracle.opt_lookup("{a:4} trailing", "a", 0) == (True, 4)        # trailing data after an object
    assert oracle.opt_lookup("[1] x", "a", 0)[0] is False


def _rand_frame(rng, w, h):
    return (rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
            rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))


def _reduce_dim(d, n):
    for _ in range(n):
        d = 2 * (d // 4)
    return d


def test_constant_image_gives_constant_pyramid(oracle):
    w, h = 96, 64
    planes = tuple(np.full(s, 77, np.uint8) for s in ((h, w), (h // 2, w // 2), (h // 2, w // 2)))
    for stride_pad in (0, 8):
        strides = (w + stride_pad, w // 2 + stride_pad, w // 2 + stride_pad)
        out = oracle.super_build("{pel:2}", planes, strides)
        o, _ = oracle.super_opts("{pel:2}", w, h)
        for pi, plane in enumerate(out):
            pw, ph = (w, h) if pi == 0 else (w // 2, h // 2)
            st = strides[pi]
            # level 0: 4 subplanes
            assert (plane[:ph, :st] == 77).all()                       # copy + edge padding
            avgs = plane[ph:4 * ph, :pw].copy()
            if stride_pad:
                # last diagonal byte reads the zero padding of the first x-half row
                assert avgs[-1, -1] == (77 + 0 + 1) >> 1
                avgs[-1, -1] = 77
            assert (avgs == 77).all()                                  # averages of a constant
            assert (plane[ph:4 * ph, pw:] == 0).all()                  # subplane padding untouched
            row = 4 * ph
            for lv in range(1, o["levels"]):
                lw = _reduce_dim(w, lv) // (2 if pi else 1)
                lh = _reduce_dim(h, lv) // (2 if pi else 1)
                sw = _reduce_dim(w, lv - 1) // (2 if pi else 1)
                blk = plane[row:row + lh]
                assert (blk[:, :sw] == 77).all()                       # filtered + vertical-only bytes
                assert (blk[:, sw:] == 0).all()
                row += lh
            assert row == plane.shape[0]


def test_level0_copy_and_pel_averages(oracle):
    rng = np.random.default_rng(5)
    w, h = 64, 40
    planes = _rand_frame(rng, w, h)
    out = oracle.super_build("{pel:2}", planes, (w + 16, w // 2 + 16, w // 2 + 16))
    Y = planes[0].astype(np.int32)
    S = out[0].astype(np.int32)
    assert (S[:h, :w] == Y).all()
    assert (S[:h, w:] == Y[:, -1:]).all()
    avg = lambda a, b: (a + b + 1) >> 1
    assert (S[h:2 * h, :w - 1] == avg(Y[:, :-1], Y[:, 1:])).all()
    assert (S[h:2 * h, w - 1] == Y[:, -1]).all()                      # x+1 reads the edge padding
    assert (S[2 * h + 1:3 * h - 1, :w] == avg(Y[1:-1], Y[2:])).all()
    assert (S[2 * h, :w] == Y[0]).all()                                # row 0 replaced by the full row
    assert (S[3 * h + 1:4 * h - 1, :w - 1] == avg(Y[1:-1, :-1], Y[2:, 1:])).all()
    assert (S[3 * h, :w - 1] == Y[0, 1:]).all()                        # row 0 = full row 0 shifted


def test_impulse_fixes_6tap_weights(oracle):
    """A single 255 at (y0, x0) of level 0 gives the separable [1,5,10,10,5,1]/32 taps at level 1."""
    w, h = 64, 64
    taps = {-2: 1, -1: 5, 0: 10, 1: 10, 2: 5, 3: 1}
    for y0, x0 in ((30, 30), (31, 33), (28, 29)):
        Y = np.zeros((h, w), np.uint8)
        Y[y0, x0] = 255
        planes = (Y, np.zeros((h // 2, w // 2), np.uint8), np.zeros((h // 2, w // 2), np.uint8))
        out = oracle.super_build("{pel:1}", planes)
        L1 = out[0][h:h + 32, :32].astype(np.int64)
        exp = np.zeros((32, 32), np.int64)
        for dy in range(2, 30):
            wv = taps.get(y0 - 2 * dy, 0)
            inter = (255 * wv + 16) >> 5
            for dx in range(2, 30):
                wh = taps.get(x0 - 2 * dx, 0)
                exp[dy, dx] = (inter * wh + 16) >> 5
        assert (L1[2:30, 2:30] == exp[2:30, 2:30]).all()


# ---- naive per-pixel restatement (independent of the oracle's code shape)

def _naive_super(planes, strides, opt_pel):
    Y = planes[0]
    h, w = Y.shape
    levels = 0
    while _reduce_dim(h, levels) >= 4 and _reduce_dim(w, levels) >= 4:
        levels += 1
    res = []
    for pi, P in enumerate(planes):
        st = strides[pi]
        ch = pi != 0
        lw = [(_reduce_dim(w, l) // 2 if ch else _reduce_dim(w, l)) for l in range(levels)]
        lh = [(_reduce_dim(h, l) // 2 if ch else _reduce_dim(h, l)) for l in range(levels)]
        rows = opt_pel * opt_pel * lh[0] + sum(lh[1:])
        buf = np.zeros(rows * st, np.int64)
        bw, bh = lw[0], lh[0]
        for y in range(bh):
            for x in range(st):
                buf[y * st + x] = P[y, min(x, bw - 1)]
        get = lambda i, fb: buf[i] if i < len(buf) else fb
        row = opt_pel * opt_pel * bh
        for l in range(levels - 1):
            sw, sh, dw, dh = lw[l], lh[l], lw[l + 1], lh[l + 1]
            src0 = 0 if l == 0 else row_prev
            inter = np.zeros((dh, sw), np.int64)
            for dy in range(dh):
                sy = 2 * dy
                edge = dy == 0 or dy + 1 == dh or sy - 2 < 0 or sy + 3 >= sh
                for x in range(sw):
                    if edge:
                        a = buf[(src0 + min(sy, sh - 1)) * st + x]
                        b = buf[(src0 + min(sy + 1, sh - 1)) * st + x]
                        inter[dy, x] = (a + b + 1) >> 1
                    else:
                        p = [buf[(src0 + sy - 2 + k) * st + x] for k in range(6)]
                        inter[dy, x] = min(255, (p[0] + p[5] + 5 * (p[1] + p[4]) + 10 * (p[2] + p[3]) + 16) >> 5)
            for dy in range(dh):
                s = list(inter[dy])
                o = list(s)
                if dw <= 2:
                    for i in range(dw):
                        o[i] = (s[2 * i] + s[2 * i + 1] + 1) >> 1
                else:
                    o[0] = (s[0] + s[1] + 1) >> 1
                    for j in range(1, dw - 1):
                        left = o[0] if j == 1 else s[2 * j - 2]
                        v = (left + s[2 * j + 3] + 10 * (s[2 * j] + s[2 * j + 1]) + 5 * (s[2 * j - 1] + s[2 * j + 2]) + 16) >> 5
                        o[j] = min(255, v)
                    j = dw - 1
                    o[j] = (s[2 * j] + s[2 * j + 1] + 1) >> 1
                for x in range(sw):
                    buf[(row + dy) * st + x] = o[x]
            row_prev = row
            row += dh
        if opt_pel == 2:
            for y in range(bh):  # x-half
                for x in range(bw):
                    a = buf[y * st + x]
                    buf[(bh + y) * st + x] = (a + get(y * st + x + 1, a) + 1) >> 1
            for y in range(bh):  # y-half
                for x in range(bw):
                    a = buf[y * st + x]
                    buf[(2 * bh + y) * st + x] = (a + get((y + 1) * st + x, a) + 1) >> 1
            for x in range(bw):
                buf[2 * bh * st + x] = buf[x]
            for y in range(bh):  # diagonal
                for x in range(bw):
                    a = buf[y * st + x]
                    buf[(3 * bh + y) * st + x] = (a + get((y + 1) * st + 1 + x, a) + 1) >> 1
            for x in range(bw):
                buf[3 * bh * st + x] = buf[1 + x]
        res.append(buf.reshape(rows, st).astype(np.uint8))
    return res


@pytest.mark.parametrize("w,h,pad,pel", [(24, 16, 0, 2), (24, 16, 8, 2), (40, 24, 0, 1), (36, 20, 4, 2)])
def test_oracle_matches_naive_restatement(oracle, w, h, pad, pel):
    rng = np.random.default_rng(w * 100 + h + pad)
    planes = _rand_frame(rng, w, h)
    strides = (w + pad, w // 2 + pad, w // 2 + pad)
    out = oracle.super_build(f"{{pel:{pel}}}", planes, strides)
    ref = _naive_super(planes, strides, pel)
    for a, b in zip(out, ref):
        assert a.shape == b.shape
        assert (a == b).all()
