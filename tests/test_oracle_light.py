"""Pins of the oracle's light border and output padding (oracle/light.cpp, restating
crates/svpflow2/src/light.rs and Options::padding, options.rs:256-277, 788-796).

Each part is checked against something fixed outside the oracle's code:
  * the random draws: the published srand/rand sequence (the 32-bit LCG 214013 x +
    2531011, 15 bits out) for seed 1;
  * the edge curve: the textbook uniform Catmull-Rom spline through control points at
    the lights' centres, with the end points repeated, in float64 (the oracle's f32
    evaluation agrees within 0.02: its segment parameter comes from nx * lights in f32,
    while a wrong term, tangent or segment is off by the control points' differences);
  * the control points: brute-force window means in Python integers;
  * the corner geometry: its closed forms (diagonal blend 1/2, the blend of a cell and
    of its transpose summing to 1, the fade 1 - distance / length clamped at 0);
  * the padding: sizes worked out by hand from the aspect / zoom rule;
  * the strip and corner layout: a numpy restatement written from light.rs (whole
    rows and columns at a time, float32 and float64 as there), on random interiors,
    with the draws from the pinned generator and the curves from the pinned profile;
  * invariants: an interior at the planes' base values (luma 0, chroma 127) gives a
    luma border of exactly 0 and a chroma border within 125..128; the outermost faded
    row and column of a luma strip are 0.
CPU only (no GPU)."""
import numpy as np
import pytest

import oracle_lib as O


def lcg(seed, n):
    """The srand/rand draws: state = 214013 state + 2531011 (mod 2^32), 15 bits out."""
    out, st = [], seed & 0xFFFFFFFF
    for _ in range(n):
        st = (st * 214013 + 2531011) & 0xFFFFFFFF
        out.append((st >> 16) & 0x7FFF)
    return out


def test_rng_known_answers():
    # the first ten rand() values after srand(1) of the C runtime this generator comes from
    want = [41, 18467, 6334, 26500, 19169, 15724, 11478, 29358, 26962, 24464]
    assert list(O.light_rng(1, 10)) == want
    assert lcg(1, 10) == want
    assert list(O.light_rng(12345, 500)) == lcg(12345, 500)


def catmull_rom(control, dim):
    """Textbook uniform Catmull-Rom through points at u_k = (k + 0.5) / L (k = 0..L-1),
    the end points repeated, evaluated at u = o / (dim - 1), clamped to [0, 255]."""
    P = np.asarray(control, np.float64)
    L = len(P)
    out = np.empty(dim)
    for o in range(dim):
        u = o / (dim - 1)
        i = int(np.sum((np.arange(-1, L + 1) + 0.5) / L < u)) - 1  # segment: u in (u_{i-1}, u_i]
        def p(k):
            return P[min(max(k, 0), L - 1)]
        t = (u - (i - 1 + 0.5) / L) * L
        p0, p1 = p(i - 1), p(i)
        m0, m1 = (p(i) - p(i - 2)) / 2, (p(i + 1) - p(i - 1)) / 2
        h00, h10, h01, h11 = 2 * t**3 - 3 * t**2 + 1, t**3 - 2 * t**2 + t, -2 * t**3 + 3 * t**2, t**3 - t**2
        out[o] = min(max(h00 * p0 + h10 * m0 + h01 * p1 + h11 * m1, 0.0), 255.0)
    return out


@pytest.mark.parametrize("lights,dim,seed", [(1, 40, 1), (2, 50, 2), (5, 300, 3), (16, 1920, 4), (33, 97, 5),
                                             (100, 1080, 6)])
def test_edge_curve_is_catmull_rom(lights, dim, seed):
    rng = np.random.default_rng(seed)
    control = rng.integers(0, 256, lights)
    got = O.light_cubic(control, dim)
    want = catmull_rom(control, dim)
    assert np.abs(got - want).max() < 0.02


def window_means(accum, lights, border, cell, base):
    dim = len(accum)
    half = int(dim * cell / (2 * lights))
    window = 2 * half + 3
    out = []
    for k in range(lights):
        lo = max(int((k + 0.5) * dim / lights - (half + 1)), 0)
        hi = min(window + lo, dim)
        out.append(sum(int(a) for a in accum[lo:hi]) // (border * (hi - lo)) if hi > lo else base)
    return out


@pytest.mark.parametrize("dim,border,lights,cell,seed", [(200, 12, 16, 1.0, 1), (97, 3, 7, 2.5, 2),
                                                         (640, 1, 40, 0.1, 3), (300, 6, 9, 40.0, 4)])
def test_control_points_are_window_means(dim, border, lights, cell, seed):
    rng = np.random.default_rng(seed)
    accum = rng.integers(0, 255 * border + 1, dim)
    got = O.light_profile(accum, border, lights, cell, 127)
    want = O.light_cubic(window_means(accum, lights, border, cell, 127), dim)
    assert np.array_equal(got, want)


def test_corner_geometry_closed_forms():
    for (w, h, length) in ((7, 7, 100), (12, 5, 60), (3, 9, 0), (16, 16, 150)):
        mix, wt = O.light_corner_geometry(w, h, length)
        x = (np.arange(w) + 1) / w
        y = ((np.arange(h) + 1) / h)[:, None]
        dist = np.sqrt(x * x + y * y) * 100
        want_w = np.maximum(1 - dist / length, 0) if length > 0 else np.zeros((h, w))
        assert np.abs(wt - want_w).max() < 1e-12
        assert np.all(mix > 0) and np.all(mix < 1)
        below = x[None, :] > y  # x > y: blend y / 2x, under 1/2
        assert np.all(mix[below] < 0.5) and np.all(mix[~below] >= 0.5)
        if w == h:
            assert np.allclose(np.diag(mix), 0.5)
            assert np.abs(mix + mix.T - 1).max() < 1e-15


@pytest.mark.parametrize("w,h,light,want", [
    (1920, 1080, "light:{aspect:2.35}", (308, 0)),    # tw = 2538: extra 618, half 309, to a multiple of 4
    (640, 360, "light:{zoom:10}", (32, 18)),          # extra 64 x 36
    (640, 360, "light:{aspect:1.33}", (0, 60)),       # th = 481: extra 121, half 60
    (1920, 1080, "light:{aspect:1.78}", (0, 0)),      # within 0.01 of the source aspect
    (640, 360, "light:{aspect:1.0,sar:0.5}", (40, 0)),  # aspect 2 / sar: tw = 720
    (640, 360, "light:{aspect:2.35,zoom:5}", (124, 10)),  # tw 846: extra 42 + 206; th 360: extra 18
])
def test_padding_worked_examples(w, h, light, want):
    so, _ = O.super_opts("{pel:2}", w, h)
    ao, _ = O.analyse_opts("{block:{w:16,h:16}}", so["sdata"], w, h)
    vdata = ao["vdata"]
    frames = [[np.zeros((h, w), np.uint8), np.zeros((h // 2, w // 2), np.uint8), np.zeros((h // 2, w // 2), np.uint8)]]
    clip = O.SmoothClip("{" + light + "}", frames, None, [np.zeros(64, np.uint8)], 0, vdata, fps=(30, 1))
    assert clip.padding() == want


def border_numpy(planes, width, height, px, py, frame, border=12, lights=16, length=100, cell=1.0):
    """light.rs's apply / plane / fill_strip / fill_corners written with numpy rows."""
    frame = max(frame, 0)
    draws = lcg((frame + 1) // 2, 100000)
    pos = 15 if frame % 2 == 0 else 0
    noise = []
    for _ in range(6):
        noise.append(np.array([d % 3 - 1 for d in draws[pos:pos + 128]]))
        pos += 128
    for p in range(3):
        P = planes[p]
        w, h = (width, height) if p == 0 else (width // 2, height // 2)
        ppx, ppy = (px, py) if p == 0 else (px // 2, py // 2)
        base = 0 if p == 0 else 127
        b = max(border if p == 0 else border // 2, 1)
        th, tv, tc = (0, 2, 4) if p == 0 else (1, 3, 5)
        inner = P[ppy:ppy + h, ppx:ppx + w].astype(np.int64)
        lv = max(lights * h // w, 1)
        sides = []
        if ppy > 0:
            sides += [("top", inner[:b].sum(0), lights), ("bottom", inner[::-1][:b].sum(0), lights)]
        if ppx > 0:
            sides += [("left", inner[:, :b].sum(1), lv), ("right", inner[:, ::-1][:, :b].sum(1), lv)]
        prof = {name: O.light_profile(acc, b, ls, max(cell, 0.1), base) for name, acc, ls in sides}
        f32 = np.float32
        ends = {}
        for name, pr in prof.items():
            pad = ppy if name in ("top", "bottom") else ppx
            far = name in ("bottom", "right")
            fac = f32(pad / (pad - 1)) if far and pad * length >= 200 and pad > 1 else f32(1)
            ends[name] = [int((f32(e) - f32(base)) * fac + f32(base)) for e in (pr[0], pr[-1])]
        for name, _, _ in sides:
            pr = prof[name].astype(np.float32)
            pad = ppy if name in ("top", "bottom") else ppx
            faded = pad * length >= 200
            table = noise[th if name in ("top", "bottom") else tv]
            for d in range(pad):
                f = 1.0
                if faded and pad > 1:
                    f = d / (pad - 1) if name in ("top", "left") else (pad - d) / (pad - 1)
                r = draws[pos]
                pos += 1
                v = ((pr - f32(base)) * f32(f) + f32(base)).astype(np.int32)
                lo = v & 0xFF
                i = np.arange(len(pr))
                v = np.where((lo >= 1) & (lo <= 254), v + table[(r - (r & 0x80) + i) & 0x7F], v)
                v = (v & 0xFF).astype(np.uint8)
                if name == "top":
                    P[d, ppx:ppx + w] = v
                elif name == "bottom":
                    P[ppy + h + d, ppx:ppx + w] = v
                elif name == "left":
                    P[ppy:ppy + h, d] = v
                else:
                    P[ppy:ppy + h, ppx + w + d] = v
        if ppx <= 0 or ppy <= 0:
            continue
        mix, wt = O.light_corner_geometry(ppx, ppy, length)
        corners = [(0, 0, ends["left"][0], ends["top"][0], True, True),
                   (ppx + w, 0, ends["right"][0], ends["top"][1], False, True),
                   (0, ppy + h, ends["left"][1], ends["bottom"][0], True, False),
                   (ppx + w, ppy + h, ends["right"][1], ends["bottom"][1], False, False)]
        for ox, oy, a5, a6, mx, my in corners:
            for row in range(ppy):
                r = draws[pos]
                pos += 1
                gi = ppy - 1 - row if my else row
                gj = np.arange(ppx)[::-1] if mx else np.arange(ppx)
                m, wg = mix[gi, gj], wt[gi, gj]
                v = base + ((a6 * m + a5 * (1.0 - m) - base) * wg).astype(np.int64)
                lo = v & 0xFFFF
                v = np.where((lo >= 1) & (lo <= 254), v + noise[tc][(r + np.arange(ppx)) & 0x7F], v)
                P[oy + row, ox:ox + ppx] = (v & 0xFF).astype(np.uint8)
    return planes


def padded_random(width, height, px, py, seed):
    rng = np.random.default_rng(seed)
    ow, oh = width + 2 * px, height + 2 * py
    planes = [rng.integers(0, 256, (oh, ow), dtype=np.uint8),
              rng.integers(0, 256, (oh // 2, ow // 2), dtype=np.uint8),
              rng.integers(0, 256, (oh // 2, ow // 2), dtype=np.uint8)]
    # smooth interiors (a blurred field), so the curves are not flat noise
    for p, P in enumerate(planes):
        a = P.astype(np.float64)
        for _ in range(3):
            a = (a + np.roll(a, 1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 0) + np.roll(a, -1, 1)) / 5
        planes[p] = a.astype(np.uint8)
    return planes


@pytest.mark.parametrize("w,h,px,py,frame,kw", [
    (160, 96, 32, 18, 0, {}),
    (160, 96, 32, 18, 7, {"border": 4, "lights": 5, "length": 20, "cell": 2.5}),
    (200, 120, 0, 24, 3, {}),
    (200, 120, 40, 0, 10, {"lights": 64, "cell": 40.0}),
    (96, 64, 8, 6, 1, {"length": 0, "border": -3}),
    (120, 80, 12, 10, 2, {"lights": 1, "length": 150}),
])
def test_border_layout_matches_numpy_restatement(w, h, px, py, frame, kw):
    a = padded_random(w, h, px, py, seed=w + px + frame)
    b = [x.copy() for x in a]
    O.light_apply(a, w, h, px, py, frame, **kw)
    border_numpy(b, w, h, px, py, frame, **kw)
    for p in range(3):
        assert np.array_equal(a[p], b[p]), f"plane {p}: {np.argwhere(a[p] != b[p])[:5]}"


def test_border_of_a_base_valued_frame():
    w, h, px, py = 160, 96, 32, 18
    ow, oh = w + 2 * px, h + 2 * py
    planes = [np.zeros((oh, ow), np.uint8), np.full((oh // 2, ow // 2), 127, np.uint8),
              np.full((oh // 2, ow // 2), 127, np.uint8)]
    for frame in (0, 1, 2, 9):
        P = [x.copy() for x in planes]
        O.light_apply(P, w, h, px, py, frame)
        assert not P[0].any()
        # a flat 127 curve is 127 within f32 rounding (the Hermite weights sum to 1 only
        # to a few ulp), so it truncates to 126 or 127 before the noise of -1, 0 or 1
        for c in (P[1], P[2]):
            assert c.min() >= 125 and c.max() <= 128 and (c != 127).any()


def test_outer_edge_of_faded_strips_is_black():
    w, h, px, py = 160, 96, 32, 18
    P = padded_random(w, h, px, py, seed=5)
    O.light_apply(P, w, h, px, py, 4)
    Y = P[0]
    assert not Y[0, px:px + w].any()  # top strip, d = 0: factor 0, no noise on 0
    assert not Y[py:py + h, 0].any()  # left strip, d = 0
