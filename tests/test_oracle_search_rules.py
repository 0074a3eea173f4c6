"""Pins of the oracle's Analyse hierarchy, block by block, from its search trace.

The reference ships no fixtures for this path, and its visit orders, seeds,
predictors, lambdas, bad-block count and global predictor have no closed form over
a whole frame.  What does have one is each block's decision given its inputs.  The
oracle records, for every block visit, the inputs it used and what it stored
(`oracle_analyse_trace`).  These tests recompute every recorded input from the
reference's rules (crates/svpflow1/src/analyse_search.rs), using only values they
derive themselves -- the previous level's stored field, the visit positions, and the
candidate cost that test_oracle_pins.py pins to textbook SAD/SATD -- and compare:

  * visit orders (:302-419): the meander and border interleave of the finest level,
    and for coarse levels the 2x2 child groups of the parents in stable score-key
    order, interior before boundary, then the edge row and column;
  * block geometry (:513-530) and the search bounds (:1620-1636);
  * the predictor (:532-552): the coarser field with each child reading its parent
    clamped to the coarse grid, times 2 (times pel at level 0); an edge block reads
    its inner neighbour's current vector and re-scores it;
  * the projected opposite candidate (:455-474): the block holding a coarse child's
    moved centre keeps the lowest-score child, negated;
  * the global predictor (:1285-1324), brute force: per-axis mode (smallest value
    among the most frequent), rounded mean of the vectors within 5 of it, doubled;
  * exact_lambda (:1279-1283) and adapt_lambda (:1555-1564) as float closed forms;
  * seeds (:574-584): the stored vectors of the 3x3 neighbours visited earlier;
  * the predictor stage (:684-768): first strict minimum of sad + (penalty*sad >> 8)
    over field, previous, global, zero and seeds;
  * exhaustive refinement (:797-823): brute-force argmin over the window;
  * the bad-block running count (:607-612);
  * the stored vector (:673-678): score x3 at the coarsest level, luma DC the floor
    mean of the block clipped to the level;
  * rescale_scores (:2716-2738) from the final field to the payload, in Python
    big integers;
  * the refine pass (:1326-1451): start vector and parent score as the floor of the
    exact bilinear interpolation, border blocks at half size with a third of the
    penalties, and an exhaustive brute force.
A wrong neighbour, a penalty on the wrong candidate, a swapped direction, a dropped
clamp or an off-by-one in the running count fails one of these.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

import synth

MAXU = 0xFFFFFFFF
I32_MIN, I32_MAX = -2 ** 31, 2 ** 31 - 1


def tdiv(a: int, b: int) -> int:
    """Integer division truncating toward zero (Rust `/`)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def sat32(v: int) -> int:
    return max(I32_MIN, min(I32_MAX, v))


def as_i32(v: int) -> int:
    v &= MAXU
    return v - 2 ** 32 if v >= 2 ** 31 else v


def f2i(v: float) -> int:
    """Rust `f64 as i32`: saturating, NaN -> 0."""
    if math.isnan(v):
        return 0
    if v >= I32_MAX:
        return I32_MAX
    if v <= I32_MIN:
        return I32_MIN
    return int(v)


def clamp16(v: int) -> int:
    return max(-32768, min(32767, v))


def reduce_dim(d: int, n: int) -> int:
    for _ in range(n):
        d = 2 * (d // 4)
    return d


def powi(a: float, n: int) -> float:
    """f64::powi as compiled: repeated squaring (compiler-rt __powidf2)."""
    recip = n < 0
    r = 1.0
    while True:
        if n & 1:
            r *= a
        n = tdiv(n, 2)
        if n == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def exact_lambda(lam: int, plevel: float, levels: int, level: int, pel: int) -> int:
    return f2i(lam / powi(plevel, levels - level - 1) / max(pel * pel, 1))


def adapt_lambda(nl: int, lsad: int, score: int) -> int:
    l = float(max(lsad, 1))
    half = -1.0 if score == MAXU else float(score >> 1)
    den = l + half
    t = l / den if den != 0.0 else math.inf        # IEEE: a positive over +0 is +inf
    v = nl * t * t
    v = 21474836.47 if math.isnan(v) else min(v, 21474836.47)   # f64::min ignores a NaN operand
    return f2i(v)


def global_doubled(vecs):
    """Brute force: per-axis mode (smallest value among the most frequent), rounded
    mean (half away from zero) of the vectors within 5 of both modes, doubled."""
    if not vecs:
        return 0, 0

    def mode(vals):
        cnt = {}
        for v in vals:
            cnt[v] = cnt.get(v, 0) + 1
        top = max(cnt.values())
        return min(v for v, c in cnt.items() if c == top)

    mx, my = mode([v[0] for v in vecs]), mode([v[1] for v in vecs])
    sel = [v for v in vecs if abs(v[0] - mx) <= 5 and abs(v[1] - my) <= 5]
    if not sel:
        return 2 * mx, 2 * my

    def rmean(vals):
        q = Fraction(sum(vals), len(vals))
        r = math.floor(abs(q) + Fraction(1, 2))
        return r if q >= 0 else -r

    return 2 * rmean([v[0] for v in sel]), 2 * rmean([v[1] for v in sel])


def rescale(score: int, dx: int, dy: int, width: int, height: int, blk: int) -> int:
    diag = max(f2i(math.sqrt(float(width) * width + float(height) * height)), 0)
    denom = max(sat32(diag * max(blk, 1)), 1)
    ratio = tdiv(100 * (abs(dx) + abs(dy)), denom)
    s = as_i32(score)
    if ratio >= 51:
        s = sat32(s * 20)
    elif ratio >= 16:
        v71 = as_i32(((3926827243 * (s & MAXU) * (ratio - 15)) % 2 ** 64) >> 32)
        s = sat32(s + 9 * ((v71 >> 31) + (v71 >> 5)))
    return (s & MAXU) & 0xFFFFFF


# ------------------------------------------------------------------------------ options

OVERLAP = {0: lambda b: 0, 1: lambda b: b // 8, 2: lambda b: b // 4, 3: lambda b: b // 2}


class Case:
    def __init__(self, name, w, h, pel, bw=8, bh=8, ov=2, levels=0, stype=3, dist=8, ctype=None, cdist=0,
                 satd=True, csatd=True, lam=10.0, lsad=8000, plevel=1.5, pnew=50, pglobal=50, pzero=100, pnbour=50,
                 prev=0, bad=1000, brange=-24, refine=None, seed=5, vectors=3):
        self.__dict__.update(locals())
        del self.__dict__["self"]
        self.ctype = stype if ctype is None else ctype

    def opt(self) -> str:
        ref = ""
        if self.refine:
            th, rt, rd = self.refine
            ref = ",refine:[{thsad:%d,search:{type:%d,distance:%d}}]" % (th, rt, rd)
        return ("{vectors:%d,block:{w:%d,h:%d,overlap:%d},main:{levels:%d,search:{type:%d,distance:%d,satd:%s,sort:true,"
                "coarse:{type:%d,distance:%d,satd:%s,bad:{sad:%d,range:%d}}},penalty:{lambda:%r,lsad:%d,plevel:%r,"
                "pnew:%d,pglobal:%d,pzero:%d,pnbour:%d,prev:%d}}%s}") % (
            self.vectors, self.bw, self.bh, self.ov, self.levels, self.stype, self.dist, str(self.satd).lower(),
            self.ctype, self.cdist, str(self.csatd).lower(), self.bad, self.brange, self.lam, self.lsad, self.plevel,
            self.pnew, self.pglobal, self.pzero, self.pnbour, self.prev, ref)


CASES = [
    Case("umh_satd_8x8", 128, 96, 2),
    Case("hex2_16x16_sad", 192, 128, 2, bw=16, bh=16, ov=0, stype=2, dist=8, satd=False, csatd=False, seed=9),
    Case("exhaustive_pel1", 128, 96, 1, ov=1, stype=4, dist=3, cdist=2, pnew=20, bad=40, seed=13),
    Case("exhaustive_pel2_prev_badumh", 112, 80, 2, ov=3, stype=4, dist=2, ctype=4, cdist=2, prev=30, pnew=0,
         pzero=40, pglobal=10, pnbour=0, bad=30, brange=6, lam=4.0, plevel=2.0, seed=17),
    Case("dist0_nonsquare", 128, 96, 2, bw=16, bh=8, ov=2, stype=3, dist=0, levels=3, bad=20, brange=-6, seed=19),
    Case("refine_exhaustive", 128, 96, 2, bw=16, bh=16, ov=2, stype=2, dist=4, refine=(100, 4, 2), seed=23),
    Case("fwd_only_4x4", 64, 48, 2, bw=4, bh=4, ov=0, stype=4, dist=2, satd=False, vectors=2, seed=29),
]


def _auto_levels(w, h, bw, bh, ox, oy):
    sx, sy = max(bw - ox, 1), max(bh - oy, 1)
    gw, gh = max(tdiv(w - ox, sx), 0), max(tdiv(h - oy, sy), 0)
    if gw <= 0 or gh <= 0:
        return 1
    ex, ey = ox + sx * gw, oy + sy * gh
    n = 0
    while True:
        if tdiv((ex >> n) - ox, sx) <= 0 or tdiv((ey >> n) - oy, sy) <= 0:
            break
        n += 1
        if n > 12:
            break
    return max(n, 1)


def _derive(c: Case, oracle, sdata):
    """Every option value the search uses, derived here from the option values."""
    o, err = oracle.analyse_opts(c.opt(), sdata, c.w, c.h)
    assert err is None, err
    ox, oy = OVERLAP[c.ov](c.bw), OVERLAP[c.ov](c.bh)
    so, _ = oracle.super_opts("{pel:%d}" % c.pel, c.w, c.h)
    levels = max(min(_auto_levels(c.w, c.h, c.bw, c.bh, ox, oy), so["levels"]), 1)
    if c.levels > 0:
        levels = max(min(c.levels, so["levels"]), 1)
    area = max(c.bw * c.bh, 1)
    d = dict(ox=ox, oy=oy, levels=levels, super_levels=so["levels"],
             lam=f2i(2.0 * float(area) * 1000.0 * 0.015625 * c.lam),
             lsad=sat32(c.lsad * (area // 32)), bad=sat32(c.bad * (area // 32)))
    assert (o["overlap_x"], o["overlap_y"], o["levels"], o["lambda"], o["lsad"]) == \
        (ox, oy, levels, d["lam"], d["lsad"])
    return d


def _level_grid(c, d, level):
    sx, sy = max(c.bw - d["ox"], 1), max(c.bh - d["oy"], 1)
    gw0, gh0 = max(tdiv(c.w - d["ox"], sx), 0), max(tdiv(c.h - d["oy"], sy), 0)
    ex, ey = d["ox"] + sx * gw0, d["oy"] + sy * gh0
    gw = max(tdiv((ex >> level) - d["ox"], sx), 1)
    gh = max(tdiv((ey >> level) - d["oy"], sy), 1)
    return gw + (level > 0), gh + (level > 0)


# ------------------------------------------------------------------------------ orders

def _meander(gw, gh):
    """initial_search_order: interior rows alternating direction, then rows 0 and
    gh-1 interleaved by x, then columns 0 and gw-1 interleaved by y."""
    out = []
    for y in range(1, gh - 1):
        xs = range(1, gw - 1) if y % 2 == 0 else range(gw - 2, 0, -1)
        out += [(x, y) for x in xs]
    for x in range(gw):
        out.append((x, 0))
        if gh > 1:
            out.append((x, gh - 1))
    for y in range(1, gh - 1):
        out.append((0, y))
        if gw > 1:
            out.append((gw - 1, y))
    return out


def _check_meander_properties(order, gw, gh):
    n_in = max(gw - 2, 0) * max(gh - 2, 0)
    inner = order[:n_in]
    for a, b in zip(inner, inner[1:]):
        assert abs(a[0] - b[0]) + abs(a[1] - b[1]) == 1            # a meander: each step to a neighbour
    for x, y in inner:
        assert 0 < x < gw - 1 and 0 < y < gh - 1
    rows = {}
    for x, y in inner:
        rows.setdefault(y, []).append(x)
    for y, xs in rows.items():
        assert xs == sorted(xs, reverse=(y % 2 == 1))              # even rows left to right


def _check_child_order(order, gw, gh, parents_sorted, cgw, cgh):
    """refine_search_order as properties: interior children, then boundary children,
    each grouped by parent in the parents' rank order, the four children of a parent
    in the 2x2 order (0,0),(1,0),(1,1),(0,1); then the edge row and edge column."""
    dw, dh = gw - 1, gh - 1
    sw, sh = cgw - 1, cgh - 1
    tail = [(x, dh) for x in range(dw + 1)] + [(dw, y) for y in range(dh)]
    assert order[-len(tail):] == tail
    head = order[:-len(tail)]
    assert sorted(head) == sorted((x, y) for y in range(dh) for x in range(dw))
    interior = [p for p in head if 0 < p[0] < dw - 1 and 0 < p[1] < dh - 1]
    assert head[:len(interior)] == interior                         # interior first, then the boundary
    rank = {p: i for i, p in enumerate(parents_sorted)}

    def parent(p):
        return min(p[0] // 2, sw - 1), min(p[1] // 2, sh - 1)

    for part in (interior, head[len(interior):]):
        ranks = [rank[parent(p)] for p in part]
        assert ranks == sorted(ranks)
        groups = {}
        for p in part:
            groups.setdefault(parent(p), []).append(p)
        quad = [(0, 0), (1, 0), (1, 1), (0, 1)]
        for (qx, qy), kids in groups.items():
            offs = [(x - 2 * qx, y - 2 * qy) for x, y in kids if x < 2 * qx + 2 and y < 2 * qy + 2]
            assert offs == [q for q in quad if q in offs]


# ------------------------------------------------------------------------------ the check

class Planes:
    def __init__(self, sup):
        self.sup = sup


def _luma_dc(sup_y, c, d, level, px, py, bw, bh):
    lw, lh = reduce_dim(c.w, level), reduce_dim(c.h, level)
    off = 0 if level == 0 else c.pel * c.pel * c.h + sum(reduce_dim(c.h, k) for k in range(1, level))
    blk = sup_y[off + py:off + min(py + bh, lh), px:min(px + bw, lw)]
    if blk.size == 0:
        return 0
    return min(int(blk.astype(np.int64).sum()) // blk.size, 255)


def _run_case(c: Case, oracle):
    frames = synth.clip(c.w, c.h, 2, seed=c.seed, vmax=4.0, nobj=(2, 2))
    sopt = "{pel:%d}" % c.pel
    so, _ = oracle.super_opts(sopt, c.w, c.h)
    sa, sb = oracle.super_build(sopt, frames[0]), oracle.super_build(sopt, frames[1])
    d = _derive(c, oracle, so["sdata"])
    pay, t0, t1 = oracle.analyse_trace(c.opt(), so["sdata"], c.w, c.h, sa, sb)
    assert np.array_equal(pay, oracle.analyse(c.opt(), so["sdata"], c.w, c.h, sa, sb))
    return frames, so, sa, sb, d, pay, {0: t0, 1: t1}


def _cost_fn(oracle, c, d, src, ref):
    cache = {}

    def cost(level, px, py, bw, bh, mvx, mvy, satd, chroma):
        k = (level, px, py, bw, bh, mvx, mvy, satd, chroma)
        if k not in cache:
            cache[k] = oracle.candidate_cost(c.w, c.h, c.pel, d["super_levels"], src, ref, level, px, py, bw, bh,
                                             mvx, mvy, satd, chroma)
        return cache[k]

    return cost


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_search_rules_block_by_block(oracle, case):
    c = case
    frames, so, sa, sb, d, pay, traces = _run_case(c, oracle)
    L = d["levels"]
    dirs = [k for k in (0, 1) if len(traces[k])]
    assert dirs == ([0, 1] if c.vectors == 3 else [0] if c.vectors == 2 else [1])
    planes = {0: (sa, sb), 1: (sb, sa)}                       # direction 0 searches cur -> ref
    costs = {k: _cost_fn(oracle, c, d, *planes[k]) for k in dirs}
    sx, sy = max(c.bw - d["ox"], 1), max(c.bh - d["oy"], 1)
    fields = {k: {} for k in dirs}                            # level -> {(x, y): (dx, dy, score, luma)}
    orders = {k: {} for k in dirs}
    counted = dict(blocks=0, exhaustive=0, bad=0, prev=0, edge=0, seeds=0, stage_moves=0)
    for level in range(L - 1, -1, -1):
        gw, gh = _level_grid(c, d, level)
        lw, lh = reduce_dim(c.w, level), reduce_dim(c.h, level)
        pel = c.pel if level == 0 else 1
        smallest = level == L - 1
        satd = c.satd if level == 0 else c.csatd
        for k in dirs:
            T = traces[k].level(level)
            order = [(int(x), int(y)) for x, y in zip(T["bx"], T["by"])]
            assert sorted(order) == sorted((x, y) for y in range(gh) for x in range(gw)), (level, k)
            pos = {p: i for i, p in enumerate(order)}
            # ---- visit order
            if level == L - 1:
                e = int(level > 0)
                assert order == _meander(gw - e, gh - e) + ([(x, gh - 1) for x in range(gw)] +
                                                           [(gw - 1, y) for y in range(gh - 1)] if e else [])
            elif level == 0:
                assert order == _meander(gw, gh)
                _check_meander_properties(order, gw, gh)
            else:
                cgw, cgh = _level_grid(c, d, level + 1)
                shift = max(c.bw * c.bh // 8, 1).bit_length()          # ilog2 + 1
                F = fields[k][level + 1]
                base = [(x, y) for y in range(cgh - 1) for x in range(cgw - 1)]
                parents = sorted(base, key=lambda p: ((F[p][2] >> shift), p[1] * cgw + p[0]))
                _check_child_order(order, gw, gh, parents, cgw, cgh)
            orders[k][level] = order
            # ---- predictors of this level
            if smallest:
                interp = lambda x, y: (0, 0, 0, 0)                     # noqa: E731
                proj = {}
                g = (0, 0)
            else:
                cgw, cgh = _level_grid(c, d, level + 1)
                F = fields[k][level + 1]

                def interp(x, y, F=F, cgw=cgw, cgh=cgh):
                    if x >= gw - 1 or y >= gh - 1:
                        return (0, 0, 0, 0)
                    p = F[(min(x >> 1, cgw - 1), min(y >> 1, cgh - 1))]
                    return (clamp16(p[0] * 2 * pel), clamp16(p[1] * 2 * pel), p[2], p[3])

                # projected candidate from the other direction (bidirectional) or this one
                G = fields[1 - k][level + 1] if c.vectors == 3 else F
                proj = {}
                for y in range(min(2 * (cgh - 1), gh - 1)):          # the last row and column
                    for x in range(min(2 * (cgw - 1), gw - 1)):      # of the finer grid project nothing
                        p = G[(x >> 1, y >> 1)]
                        vx, vy = p[0] * 2 * pel, p[1] * 2 * pel
                        tx = min(max(tdiv(vx + (c.bw >> 1) + sx * x, sx), 0), gw - 1)
                        ty = min(max(tdiv(vy + (c.bh >> 1) + sy * y, sy), 0), gh - 1)
                        if p[2] < proj.get((tx, ty), (0, 0, MAXU))[2]:
                            proj[(tx, ty)] = (-clamp16(vx), -clamp16(vy), p[2])
                g = global_doubled([v[:2] for v in F.values()])
            cur = {}
            for i, (bx, by) in enumerate(order):
                r = T.rec(i)
                idx = by * gw + bx
                edge = level > 0 and (bx == gw - 1 or by == gh - 1)
                px = lw - c.bw if (level > 0 and bx == gw - 1) else bx * sx
                py = lh - c.bh if (level > 0 and by == gh - 1) else by * sy
                bw, bh = c.bw, c.bh
                if smallest:
                    px, py, bw, bh = px + (c.bw >> 2), py + (c.bh >> 2), c.bw >> 1, c.bh >> 1
                assert (r["px"], r["py"], r["bw"], r["bh"], r["edge"], r["smallest"]) == \
                    (px, py, bw, bh, int(edge), int(smallest)), (level, bx, by)
                cost = lambda mx, my, ch=not smallest: costs[k](level, px, py, bw, bh, mx, my, satd, ch)  # noqa
                sad = lambda mx, my: min(sum(cost(mx, my)), MAXU)  # noqa: E731
                # predictor
                if edge:
                    s = (min(bx, gw - 2), min(by, gh - 2))
                    pv = cur[s] if pos[s] < i else interp(*s)
                    if not smallest:
                        l_, c_ = costs[k](level, s[0] * sx, s[1] * sy, bw, bh, pv[0], pv[1], satd, True)
                        pv = (pv[0], pv[1], min(l_ + c_, MAXU), pv[3])
                    counted["edge"] += 1
                else:
                    pv = interp(bx, by)
                assert (r["pred_dx"], r["pred_dy"], r["pred_score"]) == pv[:3], (level, k, bx, by)
                # previous (projected) candidate
                pp = proj.get((bx, by))
                has_prev = pp is not None and pp[2] != MAXU and pp[2] > 0
                assert r["has_prev"] == int(has_prev)
                if has_prev:
                    assert (r["prev_dx"], r["prev_dy"]) == pp[:2]
                    counted["prev"] += 1
                assert (r["gx"], r["gy"]) == g
                # lambda, radius, search type
                base = 0 if smallest else exact_lambda(d["lam"], c.plevel, L, level, pel)
                lam = adapt_lambda(base, tdiv(d["lsad"], 3) if smallest else d["lsad"], pv[2])
                assert r["lambda"] == lam, (level, bx, by)
                dd = c.dist if level == 0 else c.cdist
                radius = ((0 if (level == 0 and L > 1) else 10) if dd == 0 else abs(dd))
                assert r["radius"] == radius
                stype = 0 if (edge and level == 1) else (c.stype if level == 0 else c.ctype)
                assert r["stype"] == stype
                # seeds: stored vectors of the 3x3 neighbours visited before
                seeds = [cur[(nx, ny)][:2] for ny in range(max(by - 1, 0), min(by + 1, gh - 1) + 1)
                         for nx in range(max(bx - 1, 0), min(bx + 1, gw - 1) + 1) if pos[(nx, ny)] < i]
                got = [(r[f"seed{2 * j}"], r[f"seed{2 * j + 1}"]) for j in range(r["nseeds"])]
                assert got == seeds, (level, bx, by)
                counted["seeds"] += len(seeds)
                # predictor stage
                valid = pv[2] != MAXU
                mpred = pv[:2]
                if level == 0 and c.dist == 0:
                    st = (pv[:2], sad(*pv[:2])) if valid else ((0, 0), sad(0, 0))
                else:
                    best = (pv[:2], sad(*pv[:2]), sad(*pv[:2])) if valid else ((0, 0), MAXU, 2 ** 63 - 1)
                    cands = []
                    if has_prev:
                        cands.append((pp[:2], c.prev, True))
                    cands += [(g, c.pglobal, False), ((0, 0), c.pzero, False)]
                    cands += [(s_, c.pnbour, False) for s_ in seeds]
                    for mv, pen, is_prev in cands:
                        s_ = sad(*mv)
                        cc = s_ + ((pen * s_) >> 8)
                        if cc < best[2]:
                            best = (mv, s_, cc)
                            if is_prev:
                                mpred = mv
                    st = best[:2]
                assert ((r["stage_mvx"], r["stage_mvy"]), r["stage_sad"]) == st, (level, k, bx, by)
                assert (r["mpred_x"], r["mpred_y"]) == mpred
                counted["stage_moves"] += st[0] != pv[:2]
                # refinement
                if not (level == 0 and c.dist == 0) and c.pnew >= 0 and lam >= 0:
                    assert r["found"] <= st[1]                      # a move needs a strictly lower cost
                if stype == 4 and c.pnew >= 0 and not (level == 0 and c.dist == 0):
                    b0, b1 = -pel * px, -pel * py
                    b2, b3 = pel * (lw - px - bw) + 1, pel * (lh - py - bh) + 1
                    bmv, bsad, bcost = st[0], st[1], st[1]
                    for yy in range(st[0][1] - radius, st[0][1] + radius + 1):
                        for xx in range(st[0][0] - radius, st[0][0] + radius + 1):
                            if not (b0 <= xx < b2 and b1 <= yy < b3):
                                continue
                            motion = (lam * ((mpred[0] - xx) ** 2 + (mpred[1] - yy) ** 2)) >> 8
                            l_, c_ = cost(xx, yy)
                            cc = motion + l_ + ((c.pnew * l_) >> 8) + c_ + ((c.pnew * c_) >> 8)
                            if cc < bcost:
                                bmv, bsad, bcost = (xx, yy), min(l_ + c_, MAXU), cc
                    assert (r["found"]) == bsad, (level, bx, by)
                    if not r["expanded"]:
                        assert (r["out_dx"], r["out_dy"]) == bmv
                    counted["exhaustive"] += 1
                # stored vector
                if not r["expanded"]:
                    assert r["out_score"] == min(r["found"] * (3 if smallest else 1), MAXU)
                src_y = planes[k][0][0]
                assert r["out_luma"] == _luma_dc(src_y, c, d, level, px, py, bw, bh)
                cur[(bx, by)] = (r["out_dx"], r["out_dy"], r["out_score"], r["out_luma"])
                counted["blocks"] += 1
            # bad-block running count, in visit order
            n = 0
            for i, (bx, by) in enumerate(order):
                r = T.rec(i)
                lim = sat32(d["bad"] + tdiv(sat32(d["bad"] * n), 16))
                assert (r["bad_before"], r["bad_limit"]) == (n, lim)
                bad = level > 0 and by * gw + bx > 1 and d["bad"] > 0 and as_i32(r["found"]) > lim
                assert r["expanded"] == int(bad), (level, bx, by)
                n += bad
                counted["bad"] += bad
            fields[k][level] = cur
    # refine pass, then rescale, against the payload
    finals = {}
    for k in dirs:
        F0 = fields[k][0]
        if c.refine:
            F0 = _check_refine_pass(c, d, traces[k], F0, costs[k], planes[k][0][0], counted)
        finals[k] = F0
    hdr, regions = oracle.decode_payload(pay)
    gw, gh, obw = int(hdr[12]), int(hdr[13]), int(hdr[3])
    region_of = {1: 0, 0: 1} if c.vectors == 3 else {dirs[0]: 0}    # payload: backward region first
    for k in dirs:
        reg = regions[region_of[k]]
        for y in range(gh):
            for x in range(gw):
                v = finals[k][(x, y)]
                want = (v[0], v[1], rescale(v[2], v[0], v[1], c.w, c.h, obw), v[3])
                assert tuple(int(q) for q in reg[y * gw + x]) == want, (k, x, y)
    # the case exercised what it is meant to
    assert counted["blocks"] > 100 and counted["seeds"] > 0
    if not smallest or L > 1:
        assert counted["prev"] > 0 or c.vectors != 3
    if c.stype == 4 or c.ctype == 4:
        assert counted["exhaustive"] > 0
    if c.name.startswith(("exhaustive_pel1", "exhaustive_pel2", "dist0")):
        assert counted["bad"] > 0


def _check_refine_pass(c, d, T, F0, cost, src_y, counted):
    """exact_recalculate: the finer output grid's start vectors from the finest field."""
    R = T.level(-1)
    obw, obh = max(c.bw // 2, 4), max(c.bh // 2, 4)
    oox, ooy = max(d["ox"] // 2, 0), max(d["oy"] // 2, 0)
    ssx, ssy = max(obw - oox, 1), max(obh - ooy, 1)
    gw, gh = max(tdiv(c.w - oox, ssx), 1), max(tdiv(c.h - ooy, ssy), 1)
    ogw, ogh = _level_grid(c, d, 0)
    osx, osy = max(c.bw - d["ox"], 1), max(c.bh - d["oy"], 1)
    lw, lh = c.w, c.h
    th = tdiv(as_i32(c.refine[0] * max((c.bw // 2) * (c.bh // 2), 1)), 32)
    assert len(R) == gw * gh
    out = {}
    for i in range(len(R)):
        r = R.rec(i)
        bx, by = r["bx"], r["by"]
        cx, cy = obw // 2 + ssx * bx, obh // 2 + ssy * by
        ox_, oy_ = tdiv(cx - c.bw // 2, osx), tdiv(cy - c.bh // 2, osy)
        fx = Fraction(max(cx - (c.bw // 2 + osx * ox_), 0), osx)
        fy = Fraction(max(cy - (c.bh // 2 + osy * oy_), 0), osy)

        def at(x, y):
            return F0[(min(max(x, 0), ogw - 1), min(max(y, 0), ogh - 1))]

        p = [at(ox_, oy_), at(ox_ + 1, oy_), at(ox_, oy_ + 1), at(ox_ + 1, oy_ + 1)]

        def bil(j):       # floor of the exact bilinear interpolation
            v = (p[0][j] * (1 - fx) * (1 - fy) + p[1][j] * fx * (1 - fy) + p[2][j] * (1 - fx) * fy +
                 p[3][j] * fx * fy)
            return math.floor(v)

        mvx, mvy = bil(0), bil(1)
        parent = tdiv(bil(2) * (obw * obh), c.bw * c.bh)
        assert (r["pred_dx"], r["pred_dy"]) == (mvx, mvy), (bx, by)
        assert r["pred_score"] == max(min(parent, I32_MAX), I32_MIN)
        border = bx == 0 or by == 0 or bx == gw - 1 or by == gh - 1
        factor = 3 if (border and obw > 4) else 1
        ebw, ebh = (obw // 2, obh // 2) if factor == 3 else (obw, obh)
        epx, epy = bx * ssx + (obw // 4 if factor == 3 else 0), by * ssy + (obh // 4 if factor == 3 else 0)
        assert (r["px"], r["py"], r["bw"], r["bh"], r["edge"]) == (epx, epy, ebw, ebh, factor)
        pel = c.pel
        b0, b1, b2, b3 = -pel * epx, -pel * epy, pel * (lw - epx - ebw) + 1, pel * (lh - epy - ebh) + 1
        start = (min(max(mvx, b0), max(b2 - 1, b0)), min(max(mvy, b1), max(b3 - 1, b1)))
        assert (r["stage_mvx"], r["stage_mvy"]) == start
        searched = parent > tdiv(th, factor)
        assert r["expanded"] == int(searched)
        if searched:
            lam = 0 if by == 0 else tdiv(d["lam"] >> 2, factor)
            assert r["lambda"] == lam
            pnew = tdiv(c.pnew, factor)
            cc = lambda mx, my: cost(0, epx, epy, ebw, ebh, mx, my, c.satd, True)  # noqa: E731
            s0 = min(sum(cc(*start)), MAXU)
            assert r["stage_sad"] == s0
            if c.refine[1] == 4 and pnew >= 0:
                rad = abs(c.refine[2])
                bmv, bsad, bcost = start, s0, s0
                for yy in range(start[1] - rad, start[1] + rad + 1):
                    for xx in range(start[0] - rad, start[0] + rad + 1):
                        if not (b0 <= xx < b2 and b1 <= yy < b3):
                            continue
                        motion = (lam * ((start[0] - xx) ** 2 + (start[1] - yy) ** 2)) >> 8
                        l_, c_ = cc(xx, yy)
                        q = motion + l_ + ((pnew * l_) >> 8) + c_ + ((pnew * c_) >> 8)
                        if q < bcost:
                            bmv, bsad, bcost = (xx, yy), min(l_ + c_, MAXU), q
                assert (r["out_dx"], r["out_dy"], r["found"]) == (bmv[0], bmv[1], bsad), (bx, by)
                counted["exhaustive"] += 1
        else:
            assert (r["out_dx"], r["out_dy"]) == start
            assert r["found"] == max(min(parent, MAXU), 0)
        assert r["out_score"] == min(r["found"] * factor, MAXU)
        assert r["out_luma"] == _luma_dc(src_y, c, d, 0, epx, epy, ebw, ebh)
        out[(bx, by)] = (r["out_dx"], r["out_dy"], r["out_score"], r["out_luma"])
    # the refine pass visits rows in order, alternating direction with `sort`
    rows = [R.rec(i)["by"] for i in range(len(R))]
    assert rows == sorted(rows)
    return out
