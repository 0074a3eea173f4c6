#!/usr/bin/env python3
"""Benchmark of the open-svpflow motion path on one MI355X.

Workload (1080p, config 2): a 120-frame 1920x1080 YUV420P8 layered synthetic clip
(blurred-noise pan up to 12 px/frame, 6-12 rotating textured rectangles, a scene cut at
frame 60), Super {pel:2} -> Analyse 8x8 blocks, overlap 4, UMH range 16, SATD + chroma,
4 levels -> SmoothFps algorithm 23 with coverage and magnitude masks at 24->60 fps (298
output frames, 238 interpolated).

One step = the batched path over the whole clip with the source frames already in
HBM: one Super launch set for all frames, one Analyse batch for all 120 frame pairs,
one SmoothFps render batch for all output frames.  `value` is output frames per second
of that step.  Beside it, measured in the same run:
  * plain: the per-frame path (one Super frame, one Analyse pair, one output frame per
    call, as a VapourSynth host pulls them) over the first 24 source frames, with its
    SmoothFps session created before the clock starts;
  * roofline of the dominant kernel set, the Analyse level search, in its own unit:
    candidate-pixel integer ops (the oracle's candidate counter over the workload's
    pairs, 7 ops per pixel for SATD) per second against the VALU lane-op peak, timed
    with HIP events around each level's launches on the library's stream; `traffic` is
    its measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-
    calibrated) from profiles/search_traffic.json;
  * config4: 64 clips of 480p/720p/1080p/1440p with 16-96 frames each, every pair of
    every clip through svpf_analyse_multi in chunks, every clip rendered (algorithm 23);
  * serving (config 5): p50/p99 output-frame latency under Poisson arrivals;
  * cpu_baseline: the C++ oracle restatement on the host cores (frame pairs in
    parallel threads, fwd/bwd in 2 threads each), on a mini-clip of the same content.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "open-svpflow_amd"))

import svpf  # noqa: E402
import autotune  # noqa: E402
import synth  # noqa: E402

SUPER = "{pel:2}"
ANALYSE = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
SMOOTH = "{rate:{num:5,den:2},algo:23,mask:{area:100},scene:{mode:0}}"
FPS = (24, 1)
HBM_PEAK_GBS = 8000.0
# VALU peak in 32-bit lane operations: 256 CUs x 64 lanes x 2.4 GHz.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# The oracle's candidate counter over the 120 pairs of this workload, both directions
# (python bench.py --count-candidates): pixels compared by every candidate cost the
# reference evaluates (luma + chroma), per frame pair.  Every level uses SATD here, at
# 7 integer ops per pixel (16 subtractions, 64 butterfly adds, 16 absolute values and
# 16 adds per 4x4 tile).
SEARCH_CANDIDATE_PIXELS_PER_PAIR = 1.4017490e9  # levels 0-3: 1.133e9, 2.192e8, 4.536e7, 4.437e6
SATD_OPS_PER_PIXEL = 7


class Clip:
    """Device-resident sources, Super frames, vector payloads and outputs of one clip."""

    def __init__(self, ctx: svpf.Context, frames):
        self.ctx = ctx
        self.n = len(frames)
        h, w = frames[0][0].shape
        self.w, self.h = w, h
        self.sp, self.sdata = svpf.super_params(SUPER, w, h)
        sh = svpf.super_height(self.sp)
        self.super_rows = sh
        strides = (w, w // 2, w // 2)
        self.src = [svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2)) for _ in frames]
        for d, f in zip(self.src, frames):
            d.upload(f)
        self.sup = [svpf.DeviceFrame.alloc(ctx, strides, (sh, sh // 2, sh // 2)) for _ in frames]
        self.ap, self.vdata, self.plen = svpf.analyse_params(ANALYSE, self.sdata, w, h)
        self.pay = [ctx.alloc(self.plen) for _ in frames]
        self.refs = [self.sup[min(i + 1, self.n - 1)] for i in range(self.n)]
        self.smooth = svpf.Smooth(ctx, SMOOTH, w, h, FPS[0], FPS[1], self.n, self.sdata, self.vdata)
        self.nout = self.smooth.output_frames
        self.out = [svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2)) for _ in range(self.nout)]
        ctx.sync()

    def batched_step(self):
        svpf.super_build_batch(self.ctx, self.sp, self.src, self.sup)
        svpf.analyse_batch(self.ctx, self.ap, self.sup, self.refs, self.pay)
        self.smooth.set_sources(0, self.src, self.sup, self.pay)
        self.smooth.render(0, self.out)

    def plain_session(self, frames: int):
        """The per-frame path's SmoothFps session over source frames 0 .. frames-1."""
        return svpf.Smooth(self.ctx, SMOOTH, self.w, self.h, FPS[0], FPS[1], frames, self.sdata, self.vdata)

    def plain_pass(self, sm, frames: int):
        """Per-frame calls over source frames 0 .. frames-1: a Super frame per call
        (frames + 1 of them: the last pair's reference), an Analyse pair per call (each
        against the next source frame, no clamped self-pair), an output frame per call."""
        for i in range(frames + 1):
            svpf.super_build(self.ctx, self.sp, self.src[i], self.sup[i])
        for i in range(frames):
            svpf.analyse_pair(self.ctx, self.ap, self.sup[i], self.sup[i + 1], self.pay[i])
        sm.set_sources(0, self.src[:frames], self.sup[:frames], self.pay[:frames])
        for k in range(sm.output_frames):
            sm.render(k, self.out[k:k + 1])
        self.ctx.sync()
        return sm.output_frames


def search_bytes_per_step(clip: Clip) -> float:
    """Algorithmic bytes of the Analyse search: every level plane of the current and the
    reference super frame read once per frame pair (luma + 2 chroma; the pel^2 subplanes
    exist at level 0 only)."""
    pel = clip.sp.pel
    total = 0.0
    w, h = clip.w, clip.h
    for lv in range(clip.ap.levels):
        lw, lh = w, h
        for _ in range(lv):
            lw, lh = 2 * (lw // 4), 2 * (lh // 4)
        total += (pel * pel if lv == 0 else 1) * (lw * lh + 2 * (lw // 2) * (lh // 2))
    return 2.0 * total * clip.n


def render_bytes_per_step(clip: Clip, rendered: int, copies: int) -> float:
    """Rendered frame: output + two sources once + 8 motion planes (u16) + 3 masks (u8)
    over the motion grid; a copied frame: output + one source."""
    fb = clip.w * clip.h * 1.5
    gw, gh = clip.vdata[11], clip.vdata[12]
    grid = (gw + 1) * (gh + 1)
    return rendered * (3 * fb + 2 * 8 * grid + 3 * grid) + copies * 2 * fb


def count_candidates(frames, threads: int = 8):
    """The oracle's candidate counters over every frame pair of the workload: per level,
    the candidate evaluations and the pixels they compare (luma + chroma), both
    directions.  Host only (no GPU); it gives SEARCH_OPS_PER_PAIR below."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O
    from concurrent.futures import ThreadPoolExecutor
    h, w = frames[0][0].shape
    so, _ = O.super_opts(SUPER, w, h)
    with ThreadPoolExecutor(threads) as ex:
        sups = list(ex.map(lambda f: O.super_build(SUPER, f), frames))
        n = len(frames)
        res = list(ex.map(lambda i: O.analyse(ANALYSE, so["sdata"], w, h, sups[i], sups[min(i + 1, n - 1)],
                                              counters=True)[1:], range(n)))
    evals = [sum(r[0][lv] for r in res) for lv in range(16)]
    pixels = [sum(r[1][lv] for r in res) for lv in range(16)]
    return evals, pixels


def cpu_baseline(frames, pairs: int):
    """The oracle on the host: a mini-clip of `pairs`+1 frames, Analyse pairs in parallel
    threads (2 threads each for fwd/bwd), then every output frame of the mini-clip."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O
    sub = frames[:pairs + 1]
    h, w = sub[0][0].shape
    so, _ = O.super_opts(SUPER, w, h)
    t0 = time.perf_counter()
    sups = [None] * len(sub)

    def do_super(i):
        sups[i] = O.super_build(SUPER, sub[i])

    th = [threading.Thread(target=do_super, args=(i,)) for i in range(len(sub))]
    [t.start() for t in th]
    [t.join() for t in th]
    pays = [None] * len(sub)

    def do_analyse(i):
        pays[i] = O.analyse(ANALYSE, so["sdata"], w, h, sups[i], sups[min(i + 1, len(sub) - 1)], threads=True)

    th = [threading.Thread(target=do_analyse, args=(i,)) for i in range(len(sub))]
    [t.start() for t in th]
    [t.join() for t in th]
    ao, _ = O.analyse_opts(ANALYSE, so["sdata"], w, h)
    clip = O.SmoothClip(SMOOTH, sub, sups, pays, so["sdata"], ao["vdata"], fps=FPS)
    nout = O.smooth_opts(SMOOTH, FPS[0], FPS[1], ao["vdata"], len(sub))[0]["out_frames"]
    outs = list(range(nout))
    nthreads = min(len(sub), 8)

    def do_render(part):
        for k in part:
            clip.render(k)

    th = [threading.Thread(target=do_render, args=(outs[i::nthreads],)) for i in range(nthreads)]
    [t.start() for t in th]
    [t.join() for t in th]
    dt = time.perf_counter() - t0
    return nout, dt, 2 * len(sub)


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with its family/model numbers (a virtual
    machine may report only a generic model name)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # the first processor's entry
                k, _, v = line.partition(":")
                info[k.strip()] = v.strip()
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    if "cpu family" in info and "model" in info:
        name += f" (family {info['cpu family']}, model {info['model']})"
    return name


# ------------------------------------------------------------------------- config 4

C4_RES = ((854, 480), (1280, 720), (1920, 1080), (2560, 1440))
C4_ANALYSE = "{block:{w:8,h:8,overlap:3},main:{search:{type:3,distance:16,satd:true}}}"  # levels auto


class C4Clip:
    """One clip of the batched workload: its source frames (from a device pool of the
    layered generator, played back and forth from a per-clip offset), and its own super
    frames, vector payloads, SmoothFps session and output frames."""

    def __init__(self, ctx, w, h, n, pool, offset):
        self.w, self.h, self.n = w, h, n
        order = list(range(len(pool))) + list(range(len(pool) - 2, 0, -1))
        self.src = [pool[order[(offset + i) % len(order)]] for i in range(n)]
        self.sp, self.sdata = svpf.super_params(SUPER, w, h)
        sh = svpf.super_height(self.sp)
        strides = (w, w // 2, w // 2)
        self.sup = [svpf.DeviceFrame.alloc(ctx, strides, (sh, sh // 2, sh // 2)) for _ in range(n)]
        self.ap, self.vdata, self.plen = svpf.analyse_params(C4_ANALYSE, self.sdata, w, h)
        self.pay = [ctx.alloc(self.plen) for _ in range(n)]
        self.refs = [self.sup[min(i + 1, n - 1)] for i in range(n)]
        self.sm = svpf.Smooth(ctx, SMOOTH, w, h, FPS[0], FPS[1], n, self.sdata, self.vdata)
        self.nout = self.sm.output_frames
        self.out = [svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2)) for _ in range(self.nout)]


def config4_setup(ctx, clips: int = 64, seed: int = 4000, pool_frames: int = 10):
    rng = np.random.default_rng(seed)
    pools = {}
    for i, (w, h) in enumerate(C4_RES):  # two seeds (content and velocities) per size
        for k in range(2):
            frames = synth.clip(w, h, pool_frames, seed=seed + 10 * i + k, vmax=16.0, nobj=(6, 12), rot_deg=2.0)
            dev = []
            for f in frames:
                d = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
                d.upload(f)
                dev.append(d)
            pools[(w, h, k)] = dev
    out = []
    for c in range(clips):
        w, h = C4_RES[int(rng.integers(len(C4_RES)))]
        n = int(rng.integers(16, 97))
        out.append(C4Clip(ctx, w, h, n, pools[(w, h, int(rng.integers(2)))], int(rng.integers(2 * pool_frames))))
    ctx.sync()
    return out


def config4_pass(ctx, clips, pairs_per_call: int):
    """Super for every frame (one batched call per clip), Analyse for every pair of every
    clip (svpf_analyse_multi over chunks of whole clips, mixing sizes and pyramid
    depths), then each clip's SmoothFps render of all its output frames."""
    for c in clips:
        svpf.super_build_batch(ctx, c.sp, c.src, c.sup)
    chunk = []
    for c in clips + [None]:
        if c is None or (chunk and sum(x.n for x in chunk) + c.n > pairs_per_call):
            svpf.analyse_multi(ctx, [x.ap for x in chunk for _ in range(x.n)], [f for x in chunk for f in x.sup],
                               [f for x in chunk for f in x.refs], [p for x in chunk for p in x.pay])
            chunk = []
        if c is not None:
            chunk.append(c)
    for c in clips:
        c.sm.set_sources(0, c.src, c.sup, c.pay)
        c.sm.render(0, c.out)


SERVE_RES = ((854, 480), (1280, 720), (1920, 1080), (2560, 1440))
SERVE_WEIGHTS = (0.2, 0.3, 0.4, 0.1)


def serving_run(ctx, seconds: float, pixel_rate_budget: float, seed: int = 5000, arrivals_per_s: float = 40.0,
                budget_gib: float = 48.0):
    """Config-5 serving: Poisson arrivals of streams (480p/720p/1080p/1440p, 60-600
    frames, 2x at 30 fps or 2.5x at 24 fps) over `seconds`, each delivering frames in
    real time; admission by an HBM budget and the throughput budget; the config-2
    Analyse / SmoothFps settings.  Content: a device pool of 6 frames per size of the
    layered generator, played back and forth so that motion stays continuous."""
    import serve

    rng = np.random.default_rng(seed)
    pools = {}
    for i, (w, h) in enumerate(SERVE_RES):
        frames = synth.clip(w, h, 6, seed=seed + i, vmax=16.0, nobj=(6, 12), rot_deg=2.0)
        dev = []
        for f in frames:
            d = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
            d.upload(f)
            dev.append(d)
        pools[(w, h)] = dev

    def source(dev, off):
        order = list(range(len(dev))) + list(range(len(dev) - 2, 0, -1))
        return lambda i: dev[order[(off + i) % len(order)]]

    specs, t, sid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / arrivals_per_s)
        if t >= seconds:
            break
        w, h = SERVE_RES[rng.choice(len(SERVE_RES), p=SERVE_WEIGHTS)]
        two = rng.random() < 0.5
        specs.append(serve.StreamSpec(sid, w, h, int(rng.integers(60, 601)), (2, 1) if two else (5, 2),
                                      (30, 1) if two else (24, 1), t, source(pools[(w, h)], int(rng.integers(10)))))
        sid += 1
    srv = serve.Server(ctx, budget_bytes=int(budget_gib * (1 << 30)), max_wait=1.0,
                       pixel_rate_budget=pixel_rate_budget)
    stats = srv.run(specs, duration=seconds)
    srv.close()
    stats.update({"arrivals": len(specs), "arrival_rate_per_s": arrivals_per_s, "hbm_budget_gib": budget_gib,
                  "pixel_rate_budget": round(pixel_rate_budget), "window_s": seconds})
    return stats


def search_traffic():
    """Measured HBM bytes of the Analyse search per step (profiles/search_traffic.json,
    written from two rocprofv3 counter passes over this workload), or None."""
    path = os.path.join(HERE, "profiles", "search_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--plain-frames", type=int, default=24, help="source frames of the per-frame path sample")
    ap.add_argument("--cpu-pairs", type=int, default=64, help="frame pairs of the CPU baseline sample (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="length of the config-5 serving run that gives the p50/p99 frame latency (0: skip)")
    ap.add_argument("--serve-budget", type=float, default=0.5,
                    help="serving admission: fraction of the batched source pixel rate the admitted streams may deliver")
    ap.add_argument("--config4-clips", type=int, default=64, help="clips of the config-4 batched run (0: skip)")
    ap.add_argument("--config4-pairs-per-call", type=int, default=480,
                    help="frame pairs per svpf_analyse_multi call of the config-4 run (whole clips)")
    ap.add_argument("--tune-seconds", type=float, default=45.0,
                    help="wall-time budget of the Analyse schedule autotuner on the headline clip (0: defaults)")
    ap.add_argument("--tuning", type=str, default="", metavar="K=V,...",
                    help="a fixed Analyse schedule instead of the autotuner, e.g. slot_scale=1.4,sweep_grid=2048")
    ap.add_argument("--config4-tune-seconds", type=float, default=30.0,
                    help="the same for the config-4 batch, on its first svpf_analyse_multi call (0: keep)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run 4 device-to-device copies of MIB MiB each (the known-size reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE scaling)")
    ap.add_argument("--count-candidates", action="store_true",
                    help="host only: print the oracle's candidate counters over the workload's frame pairs and exit")
    args = ap.parse_args()
    if args.count_candidates:
        frames = synth.config_clip(2, frames=args.frames)
        evals, pixels = count_candidates(frames)
        print(json.dumps({"pairs": len(frames), "evaluations_per_level": evals[:8], "pixels_per_level": pixels[:8],
                          "candidate_pixels_per_pair": sum(pixels) / len(frames)}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the motion path shards by clip / frame range with no collective "
              "(see the sharding notes); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    print(f"[bench] generating {args.frames} frames 1920x1080", file=sys.stderr, flush=True)
    frames = synth.config_clip(2, frames=args.frames)
    ctx = svpf.Context(0)
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.memset(a, 1, nb)
        for _ in range(4):
            ctx.copy(b, a, nb)
        ctx.sync()
        ctx.free(a)
        ctx.free(b)
        print(f"[bench] copy calibration: 4 x {nb} bytes device to device", file=sys.stderr, flush=True)
    clip = Clip(ctx, frames)
    tuning = None
    if args.tuning:
        fixed = {k: float(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))}
        autotune.apply(ctx, fixed)
        tuning = {"fixed": fixed}
    elif args.tune_seconds > 0:
        # the search schedule for this batch, measured (untimed; every setting gives the same vectors)
        print(f"[bench] autotune of the Analyse schedule, {args.tune_seconds:.0f} s budget", file=sys.stderr, flush=True)
        svpf.super_build_batch(ctx, clip.sp, clip.src, clip.sup)
        res = autotune.tune(ctx, lambda: svpf.analyse_batch(ctx, clip.ap, clip.sup, clip.refs, clip.pay),
                            [(p, clip.plen) for p in clip.pay], budget_s=args.tune_seconds)
        tuning = res.as_dict()
        print(f"[bench] autotune: {tuning}", file=sys.stderr, flush=True)
    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        clip.batched_step()
    ctx.sync()

    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    ctx.profile(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    for s in range(args.steps):
        clip.batched_step()
    gpu_ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    prof = {k: ctx.profile_read(k) for k in ("super", "search", "smooth_prep", "render")}
    sweeps = {lv: ctx.analyse_stats(lv) for lv in range(clip.ap.levels)}
    ctx.profile(False)
    dec = clip.smooth.render(0, clip.out, decisions=True)
    rendered = sum(1 for d in dec if d[0] == 2)
    copies = len(dec) - rendered
    ms_per_step = 1000.0 * wall / args.steps
    value = clip.nout * args.steps / wall

    # plain per-frame path: the first plain_frames source frames, session made beforehand
    plain = None
    if args.plain_frames > 1:
        nf = min(args.plain_frames, clip.n - 1)
        print(f"[bench] plain per-frame path, {nf} source frames", file=sys.stderr, flush=True)
        autotune.apply(ctx, autotune.DEFAULT)  # one pair per call: the library's schedule
        sm = clip.plain_session(nf)
        clip.plain_pass(sm, nf)  # one untimed pass: the per-pair workspace and descriptors exist
        ctx.sync()
        t0 = time.perf_counter()
        plain_out = clip.plain_pass(sm, nf)
        plain_dt = time.perf_counter() - t0
        sm.close()
        plain = {"value": round(plain_out / plain_dt, 3), "unit": "fps", "source_frames": nf,
                 "output_frames": plain_out, "seconds": round(plain_dt, 3),
                 "ms_per_output_frame": round(1000.0 * plain_dt / plain_out, 3),
                 "ms_per_source_frame": round(1000.0 * plain_dt / nf, 3)}

    # roofline of the Analyse search: candidate-pixel integer ops against the VALU peak
    search_ms, search_n = prof["search"]
    pixels_per_pair = SEARCH_CANDIDATE_PIXELS_PER_PAIR
    ops = pixels_per_pair * SATD_OPS_PER_PIXEL * clip.n * args.steps
    search_tops = ops / (search_ms / 1000.0) / 1e12 if search_ms > 0 else 0.0
    sb = search_bytes_per_step(clip) * args.steps
    tr = search_traffic()
    traffic = None
    if tr:
        traffic = (tr["fetch_bytes_per_step"] + tr["write_bytes_per_step"]) / tr["level_launches_per_step"]
    render_ms, render_n = prof["render"]
    rb = render_bytes_per_step(clip, rendered, copies) * args.steps
    render_gbs = rb / (render_ms / 1000.0) / 1e9 if render_ms > 0 else 0.0
    super_ms, super_n = prof["super"]
    fb = clip.w * clip.h * 1.5
    sbytes = (fb + clip.super_rows * clip.w * 1.5) * clip.n * args.steps
    super_gbs = sbytes / (super_ms / 1000.0) / 1e9 if super_ms > 0 else 0.0

    c4 = None
    if args.config4_clips > 0:
        print(f"[bench] config 4: {args.config4_clips} clips", file=sys.stderr, flush=True)
        clips = config4_setup(ctx, args.config4_clips)
        c4_tuning = None
        if args.config4_tune_seconds > 0:
            for c in clips:
                svpf.super_build_batch(ctx, c.sp, c.src, c.sup)
            first, npairs = [], 0
            for c in clips:
                if first and npairs + c.n > args.config4_pairs_per_call:
                    break
                first.append(c)
                npairs += c.n
            print(f"[bench] config 4: autotune on its first call ({npairs} pairs), "
                  f"{args.config4_tune_seconds:.0f} s budget", file=sys.stderr, flush=True)
            res = autotune.tune(
                ctx, lambda: svpf.analyse_multi(ctx, [x.ap for x in first for _ in range(x.n)],
                                                [f for x in first for f in x.sup], [f for x in first for f in x.refs],
                                                [q for x in first for q in x.pay]),
                [(q, x.plen) for x in first for q in x.pay], budget_s=args.config4_tune_seconds)
            c4_tuning = res.as_dict()
            print(f"[bench] config 4 autotune: {c4_tuning}", file=sys.stderr, flush=True)
        config4_pass(ctx, clips, args.config4_pairs_per_call)  # warm-up (workspace, sessions' prep buffers)
        ctx.sync()
        t0 = time.perf_counter()
        config4_pass(ctx, clips, args.config4_pairs_per_call)
        ctx.sync()
        dt = time.perf_counter() - t0
        nout = sum(c.nout for c in clips)
        by_res = {f"{w}x{h}": sum(1 for c in clips if (c.w, c.h) == (w, h)) for (w, h) in C4_RES}
        c4 = {"value": round(nout / dt, 3), "unit": "fps", "seconds": round(dt, 3), "clips": len(clips),
              "source_frames": sum(c.n for c in clips), "output_frames": nout, "clips_by_size": by_res,
              "pairs_per_analyse_call": args.config4_pairs_per_call, "analyse_tuning": c4_tuning,
              "source_megapixels_per_s": round(sum(c.n * c.w * c.h for c in clips) / dt / 1e6, 1)}
        for c in clips:
            c.sm.close()
            for f in c.sup + c.out:
                for ptr in (f.y, f.u, f.v):
                    ctx.free(ptr)
            for ptr in c.pay:
                ctx.free(ptr)
        for dev in {id(f): f for c in clips for f in c.src}.values():
            for ptr in (dev.y, dev.u, dev.v):
                ctx.free(ptr)
        del clips

    serving = None
    if args.serve_seconds > 0:
        # throughput budget: 40% of the source pixel rate the batched step sustained (serving
        # steps are small batches, whose level searches are bound by latency, not
        # throughput; at 75% the admitted load outruns them and latency grows without bound)
        src_rate = clip.n * clip.w * clip.h * args.steps / wall
        print(f"[bench] serving run, {args.serve_seconds:.0f} s of arrivals", file=sys.stderr, flush=True)
        autotune.apply(ctx, autotune.DEFAULT)  # small batches: the library's schedule
        serving = serving_run(ctx, args.serve_seconds, args.serve_budget * src_rate)

    cpu = None
    if args.cpu_pairs > 0:
        print(f"[bench] cpu baseline, {args.cpu_pairs} pairs", file=sys.stderr, flush=True)
        nout, dt, threads = cpu_baseline(frames, args.cpu_pairs)
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cpu = {"value": round(nout / dt, 3), "unit": "fps", "cores": min(cores, threads), "threads": threads,
               "cores_available": cores, "cpu_model": cpu_model(), "kind": "port",
               "sample": f"oracle (C++ restatement, g++ -O3) on a {args.cpu_pairs + 1}-frame 1080p mini-clip of "
                         f"the same content: Super, Analyse ({args.cpu_pairs + 1} pairs in parallel threads, fwd/bwd "
                         f"in 2 threads each), SmoothFps of all {nout} output frames; {dt:.1f} s wall"}

    line = {
        "metric": "output fps at 1080p (bit-exact MVs), batched vs plain GPU; p99 frame latency",
        "value": round(value, 3),
        "unit": "fps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"1080p YUV420P8 layered clip, {clip.n} frames -> {clip.nout} output frames "
                               f"({rendered} interpolated): Super {SUPER}, Analyse {ANALYSE}, SmoothFps {SMOOTH} "
                               f"at {FPS[0]}/{FPS[1]} fps",
                   "path": "batched (whole clip per launch set)", "frames": clip.n, "output_frames": clip.nout},
        "plain": plain,
        "batched_over_plain": round(value / plain["value"], 2) if plain else None,
        "gpu_ms_per_step_events": round(gpu_ms / args.steps, 2),
        "roofline": {"bound": "valu",
                     "kernel": "Analyse level search (sweep0/deferred/validate/sweep/scan kernels of one pyramid "
                               "level, HIP events around each level)",
                     "achieved": round(search_tops, 4), "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s",
                     "frac": search_tops / VALU_PEAK_TOPS,
                     "traffic": traffic,
                     "traffic_source": tr.get("source") if tr else None,
                     "launches": search_n, "avg_launch_ms": round(search_ms / max(search_n, 1), 3),
                     "algorithmic_ops_per_launch": ops / max(search_n, 1),
                     "algorithmic_work": f"{pixels_per_pair:.4g} candidate pixels per frame pair (oracle counter, "
                                         f"both directions, all levels) x {SATD_OPS_PER_PIXEL} ops (SATD)",
                     "algorithmic_bytes_per_launch": sb / max(search_n, 1)},
        "roofline_render": {"bound": "hbm", "achieved": round(render_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": render_gbs / HBM_PEAK_GBS, "launches": render_n,
                            "avg_launch_ms": round(render_ms / max(render_n, 1), 3)},
        "roofline_super": {"bound": "hbm", "achieved": round(super_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": super_gbs / HBM_PEAK_GBS, "launches": super_n,
                           "avg_launch_ms": round(super_ms / max(super_n, 1), 3)},
        "stage_ms_per_step": {k: round(v[0] / args.steps, 2) for k, v in prof.items()},
        "analyse_tuning": tuning,
        "analyse_sweeps": {f"level{lv}": v for lv, v in sweeps.items()},
        "config4": c4,
        "p99_frame_latency_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
