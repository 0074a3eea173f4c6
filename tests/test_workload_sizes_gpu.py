"""GPU parity at the frame sizes of the batched and serving workloads, and the config-3
render at full size.

* One svpf_analyse_multi call over clips of 854x480 (with 32-byte aligned strides, as a
  VapourSynth frame has them: 864 / 448), 1280x720 and 2560x1440 with automatic pyramid
  depths (6, 7 and 8 levels); every payload equals the oracle's, and SmoothFps
  algorithm 23 renders every output frame of each clip as the oracle does.  The level-0
  sweep count of the 1440p pair is printed (the level converges under the sweep limit).
* Serving: a 480p and a 1440p stream give the same output frames alone and together,
  equal to the oracle's render.
* Config 3's SmoothFps (algorithm 2, 30->60) over 3840x2160 pel-1 Super frames.
"""
import numpy as np
import pytest

import serve
import svpf
import synth

pytestmark = pytest.mark.gpu

AUTO = "{block:{w:8,h:8,overlap:3},main:{search:{type:3,distance:16,satd:true}}}"
SMOOTH = "{rate:{num:5,den:2},algo:23,mask:{area:100},scene:{mode:0}}"
CFG3 = "{block:{w:16,h:16,overlap:3},main:{levels:5,search:{type:4,distance:8,satd:false,coarse:{satd:false}}}}"


def _align(v, a=32):
    return (v + a - 1) // a * a


class Clip:
    def __init__(self, ctx, frames, sopt, aopt, align=False):
        self.frames = frames
        h, w = frames[0][0].shape
        self.w, self.h = w, h
        self.strides = (_align(w), _align(w // 2), _align(w // 2)) if align else (w, w // 2, w // 2)
        sp, self.sdata = svpf.super_params(sopt, w, h)
        sh = svpf.super_height(sp)
        self.src, self.sup = [], []
        for f in frames:
            s = svpf.DeviceFrame.alloc(ctx, self.strides, (h, h // 2, h // 2))
            s.upload(f)
            self.src.append(s)
            self.sup.append(svpf.DeviceFrame.alloc(ctx, self.strides, (sh, sh // 2, sh // 2)))
        svpf.super_build_batch(ctx, sp, self.src, self.sup)
        self.ap, self.vdata, self.plen = svpf.analyse_params(aopt, self.sdata, w, h)
        self.pay = [ctx.alloc(self.plen) for _ in frames]
        n = len(frames)
        self.refs = [self.sup[min(i + 1, n - 1)] for i in range(n)]


def _render_check(ctx, oracle, clip, opt, fps, host_sup, host_pay):
    w, h = clip.w, clip.h
    sm = svpf.Smooth(ctx, opt, w, h, fps[0], fps[1], len(clip.frames), clip.sdata, clip.vdata)
    sm.set_sources(0, clip.src, clip.sup, clip.pay)
    outs = [svpf.DeviceFrame.alloc(ctx, clip.strides, (h, h // 2, h // 2)) for _ in range(sm.output_frames)]
    dec = sm.render(0, outs, decisions=True)
    ctx.sync()
    ref = oracle.SmoothClip(opt, clip.frames, host_sup, host_pay, clip.sdata, clip.vdata, fps=fps)
    rendered = 0
    for k, o in enumerate(outs):
        want, kind, algo = ref.render(k)
        assert dec[k] == (kind, algo if kind == 2 else 0), f"{w}x{h}: decision of output {k}"
        rendered += kind == 2
        for p, (g, q) in enumerate(zip(o.download(), want)):
            assert np.array_equal(g[:q.shape[0], :q.shape[1]], q), f"{w}x{h}: output {k} plane {p}"
    sm.close()
    for o in outs:
        for ptr in (o.y, o.u, o.v):
            ctx.free(ptr)
    return rendered


def test_multi_clip_480p_720p_1440p(gpu_ctx, oracle):
    ctx = gpu_ctx
    specs = [(854, 480, 3, 4101, True), (1280, 720, 3, 4102, False), (2560, 1440, 2, 4103, False)]
    clips = []
    for w, h, n, seed, align in specs:
        print(f"workload sizes: clip {w}x{h}, {n} frames, aligned strides {align}", flush=True)
        frames = synth.clip(w, h, n, seed=seed, vmax=16.0, nobj=(6, 12), rot_deg=2.0)
        clips.append(Clip(ctx, frames, "{pel:2}", AUTO, align))
    assert [c.ap.levels for c in clips] == [6, 7, 8]
    params, curs, refs, pays = [], [], [], []
    for c in clips:
        params += [c.ap] * len(c.frames)
        curs += c.sup
        refs += c.refs
        pays += c.pay
    print("workload sizes: one svpf_analyse_multi over 8 pairs", flush=True)
    svpf.analyse_multi(ctx, params, curs, refs, pays)
    ctx.sync()
    for c in clips:
        hs = [d.download() for d in c.sup]
        want_sup = [oracle.super_build("{pel:2}", f, c.strides) for f in c.frames]
        for i in range(len(c.frames)):
            assert all(np.array_equal(a, b) for a, b in zip(hs[i], want_sup[i])), f"{c.w}x{c.h} super {i}"
        host_pay = []
        for i in range(len(c.frames)):
            j = min(i + 1, len(c.frames) - 1)
            print(f"workload sizes: oracle pair {i} of {c.w}x{c.h}", flush=True)
            want = oracle.analyse(AUTO, c.sdata, c.w, c.h, hs[i], hs[j], threads=True)
            got = ctx.download(c.pay[i], c.plen)
            assert np.array_equal(got, want), f"{c.w}x{c.h}: vector payload {i} differs from the oracle"
            host_pay.append(got)
        print(f"workload sizes: render {c.w}x{c.h} algo 23", flush=True)
        sups = [[p[:, :(c.w if k == 0 else c.w // 2)] for k, p in enumerate(s)] for s in hs]
        assert _render_check(ctx, oracle, c, SMOOTH, (24, 1), sups, host_pay) > 0
    # level-0 sweeps of the 1440p pair alone (statistics are kept for converged levels)
    big = clips[2]
    ctx.profile(True)
    try:
        svpf.analyse_pair(ctx, big.ap, big.sup[0], big.sup[1], big.pay[0])
        ctx.sync()
        st = ctx.analyse_stats(0)
    finally:
        ctx.profile(False)
    print(f"workload sizes: 2560x1440 level 0: {st}", flush=True)
    assert st["runs"] == 1 and 0 < st["sweeps"] < 4096


def _pool(ctx, frames):
    h, w = frames[0][0].shape
    devs = []
    for f in frames:
        d = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
        d.upload(f)
        devs.append(d)
    return lambda i: devs[i]


def _serve(ctx, specs, keep):
    got = {}

    def on_output(sid, k, frame):
        if sid in keep:
            got[(sid, k)] = frame.download()

    srv = serve.Server(ctx, budget_bytes=16 << 30, out_per_step=4, on_output=on_output)
    stats = srv.run(specs)
    srv.close()
    return got, stats


def test_serving_480p_and_1440p_streams(gpu_ctx, oracle):
    ctx = gpu_ctx
    fa = synth.clip(854, 480, 6, seed=4201, vmax=12.0, nobj=(4, 8), rot_deg=2.0)
    fb = synth.clip(2560, 1440, 4, seed=4202, vmax=16.0, nobj=(6, 12), rot_deg=2.0)
    a = serve.StreamSpec(0, 854, 480, 6, (5, 2), (24, 1), 0.0, _pool(ctx, fa))
    b = serve.StreamSpec(1, 2560, 1440, 4, (2, 1), (30, 1), 0.0, _pool(ctx, fb))
    print("serve sizes: 480p alone", flush=True)
    alone_a, _ = _serve(ctx, [a], {0})
    print("serve sizes: 1440p alone", flush=True)
    alone_b, _ = _serve(ctx, [b], {1})
    print("serve sizes: both together", flush=True)
    mixed, st = _serve(ctx, [a, b], {0, 1})
    assert st["finished"] == 2 and st["dropped"] == 0
    for alone in (alone_a, alone_b):
        assert alone and all(k in mixed for k in alone)
        for key in alone:
            for x, y in zip(alone[key], mixed[key]):
                assert np.array_equal(x, y), f"output {key} differs alone vs in the mix"
    for sid, frames, rate, fps in ((0, fa, (5, 2), (24, 1)), (1, fb, (2, 1), (30, 1))):
        h, w = frames[0][0].shape
        n = len(frames)
        print(f"serve sizes: oracle of stream {sid} ({w}x{h})", flush=True)
        so, _ = oracle.super_opts(serve.SUPER_OPT, w, h)
        hs = [oracle.super_build(serve.SUPER_OPT, f) for f in frames]
        pays = [oracle.analyse(serve.ANALYSE_OPT, so["sdata"], w, h, hs[i], hs[min(i + 1, n - 1)], threads=True)
                for i in range(n)]
        ao, _ = oracle.analyse_opts(serve.ANALYSE_OPT, so["sdata"], w, h)
        opt = serve.smooth_opt(rate)
        clip = oracle.SmoothClip(opt, frames, hs, pays, so["sdata"], ao["vdata"], fps=fps)
        keys = sorted(k for k in mixed if k[0] == sid)
        assert len(keys) == oracle.smooth_opts(opt, fps[0], fps[1], ao["vdata"], n)[0]["out_frames"]
        for (_, k) in keys:
            want, _, _ = clip.render(k)
            for g, q in zip(mixed[(sid, k)], want):
                assert np.array_equal(g[:q.shape[0], :q.shape[1]], q), f"stream {sid} output {k} vs the oracle"


def test_config3_render_full_size(gpu_ctx, oracle):
    ctx = gpu_ctx
    print("full size: config 3 render, 3840x2160 pel 1, algorithm 2", flush=True)
    frames = synth.config_clip(3, frames=3)
    c = Clip(ctx, frames, "{pel:1}", CFG3)
    svpf.analyse_batch(ctx, c.ap, c.sup, c.refs, c.pay)
    ctx.sync()
    hs = [d.download() for d in c.sup]
    host_pay = [ctx.download(p, c.plen) for p in c.pay]
    assert _render_check(ctx, oracle, c, "{rate:{num:2,den:1},algo:2,scene:{mode:0}}", (30, 1), hs, host_pay) > 0
