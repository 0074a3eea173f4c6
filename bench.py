#!/usr/bin/env python3
"""Benchmark of the open-svpflow motion path on one MI355X.

Workload (1080p): a 120-frame 1920x1080 YUV420P8 layered synthetic clip (blurred-noise
pan up to 12 px/frame, 6-12 rotating textured rectangles, a scene cut at frame 60),
Super {pel:2} -> Analyse 8x8 blocks, overlap 4, UMH range 16, SATD + chroma, 4 levels
-> SmoothFps algorithm 23 with coverage and magnitude masks at 24->60 fps (298 output
frames, 238 interpolated).

One step = the batched path over the whole clip with the source frames already in
HBM: one Super launch set for all frames, one Analyse batch for all 120 frame pairs,
one SmoothFps render batch for all output frames.  `value` is output frames per second
of that step.  Beside it, measured in the same run:
  * plain: the per-frame path (one Super frame, one Analyse pair, one output frame per
    call, as a VapourSynth host pulls them) on a mini-clip of the first frames;
  * roofline of the dominant kernel (the Analyse block search), with HIP events around
    each of its launches on the library's stream;
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
import synth  # noqa: E402

SUPER = "{pel:2}"
ANALYSE = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
SMOOTH = "{rate:{num:5,den:2},algo:23,mask:{area:100},scene:{mode:0}}"
FPS = (24, 1)
HBM_PEAK_GBS = 8000.0


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

    def plain_pass(self, frames: int):
        """Per-frame calls over the mini-clip of the first `frames` source frames."""
        sm = svpf.Smooth(self.ctx, SMOOTH, self.w, self.h, FPS[0], FPS[1], frames, self.sdata, self.vdata)
        for i in range(frames):
            svpf.super_build(self.ctx, self.sp, self.src[i], self.sup[i])
        for i in range(frames):
            svpf.analyse_pair(self.ctx, self.ap, self.sup[i], self.sup[min(i + 1, frames - 1)], self.pay[i])
        sm.set_sources(0, self.src[:frames], self.sup[:frames], self.pay[:frames])
        for k in range(sm.output_frames):
            sm.render(k, self.out[k:k + 1])
        nout = sm.output_frames
        self.ctx.sync()
        sm.close()
        return nout


def search_bytes_per_step(clip: Clip) -> float:
    """Algorithmic bytes of the Analyse search: every level plane of the current and the
    reference super frame read once per frame pair (pel^2 subplanes of luma + 2 chroma)."""
    pel = clip.sp.pel
    total = 0.0
    w, h = clip.w, clip.h
    for lv in range(clip.ap.levels):
        lw, lh = w, h
        for _ in range(lv):
            lw, lh = 2 * (lw // 4), 2 * (lh // 4)
        total += pel * pel * (lw * lh + 2 * (lw // 2) * (lh // 2))
    return 2.0 * total * clip.n


def render_bytes_per_step(clip: Clip, rendered: int, copies: int) -> float:
    """Rendered frame: output + two sources once + 8 motion planes (u16) + 3 masks (u8)
    over the motion grid; a copied frame: output + one source."""
    fb = clip.w * clip.h * 1.5
    gw, gh = clip.vdata[11], clip.vdata[12]
    grid = (gw + 1) * (gh + 1)
    return rendered * (3 * fb + 2 * 8 * grid + 3 * grid) + copies * 2 * fb


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--plain-frames", type=int, default=3, help="mini-clip length of the per-frame path sample")
    ap.add_argument("--cpu-pairs", type=int, default=8, help="frame pairs of the CPU baseline sample (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="length of the config-5 serving run that gives the p50/p99 frame latency (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run 4 device-to-device copies of MIB MiB each (the known-size reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE scaling)")
    args = ap.parse_args()
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

    # plain per-frame path on a mini-clip
    plain_out, plain_dt, plain_fps = 0, 0.0, 0.0
    if args.plain_frames > 1:
        print(f"[bench] plain per-frame path, {args.plain_frames} frames", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        plain_out = clip.plain_pass(args.plain_frames)
        plain_dt = time.perf_counter() - t0
        plain_fps = plain_out / plain_dt

    search_ms, search_n = prof["search"]
    sb = search_bytes_per_step(clip) * args.steps
    search_gbs = sb / (search_ms / 1000.0) / 1e9 if search_ms > 0 else 0.0
    render_ms, render_n = prof["render"]
    rb = render_bytes_per_step(clip, rendered, copies) * args.steps
    render_gbs = rb / (render_ms / 1000.0) / 1e9 if render_ms > 0 else 0.0
    super_ms, super_n = prof["super"]
    fb = clip.w * clip.h * 1.5
    sbytes = (fb + clip.super_rows * clip.w * 1.5) * clip.n * args.steps
    super_gbs = sbytes / (super_ms / 1000.0) / 1e9 if super_ms > 0 else 0.0

    serving = None
    if args.serve_seconds > 0:
        # throughput budget: 25% of the source pixel rate the batched step sustained (serving
        # steps are small batches, whose level searches are bound by latency, not throughput)
        src_rate = clip.n * clip.w * clip.h * args.steps / wall
        print(f"[bench] serving run, {args.serve_seconds:.0f} s of arrivals", file=sys.stderr, flush=True)
        serving = serving_run(ctx, args.serve_seconds, 0.25 * src_rate)

    cpu = None
    if args.cpu_pairs > 0:
        print(f"[bench] cpu baseline, {args.cpu_pairs} pairs", file=sys.stderr, flush=True)
        nout, dt, threads = cpu_baseline(frames, args.cpu_pairs)
        cpu = {"value": round(nout / dt, 3), "unit": "fps", "cores": threads, "kind": "port",
               "sample": f"oracle (C++ restatement) on a {args.cpu_pairs + 1}-frame 1080p mini-clip of the same "
                         f"content: Super, Analyse ({args.cpu_pairs + 1} pairs in parallel threads, fwd/bwd in 2 "
                         f"threads each), SmoothFps of all {nout} output frames; {dt:.1f} s wall, "
                         f"os.cpu_count()={os.cpu_count()}"}

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
        "plain": {"value": round(plain_fps, 3), "unit": "fps", "frames": args.plain_frames,
                  "output_frames": plain_out, "seconds": round(plain_dt, 3)},
        "batched_over_plain": round(value / plain_fps, 2) if plain_fps > 0 else None,
        "gpu_ms_per_step_events": round(gpu_ms / args.steps, 2),
        "roofline": {"bound": "hbm", "kernel": "Analyse level search (sweep0/deferred/validate/sweep kernels of one "
                                                "pyramid level; integer-VALU bound)",
                     "achieved": round(search_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": search_gbs / HBM_PEAK_GBS, "traffic": None,
                     "launches": search_n, "avg_launch_ms": round(search_ms / max(search_n, 1), 3),
                     "algorithmic_bytes_per_launch": sb / max(search_n, 1)},
        "roofline_render": {"bound": "hbm", "achieved": round(render_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": render_gbs / HBM_PEAK_GBS, "launches": render_n,
                            "avg_launch_ms": round(render_ms / max(render_n, 1), 3)},
        "roofline_super": {"bound": "hbm", "achieved": round(super_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": super_gbs / HBM_PEAK_GBS, "launches": super_n,
                           "avg_launch_ms": round(super_ms / max(super_n, 1), 3)},
        "stage_ms_per_step": {k: round(v[0] / args.steps, 2) for k, v in prof.items()},
        "analyse_sweeps": {f"level{lv}": v for lv, v in sweeps.items()},
        "p99_frame_latency_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
