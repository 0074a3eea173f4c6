"""GPU parity at the benchmark's full frame sizes.

Config 2 (1920x1080, Super pel 2, Analyse 8x8 overlap 3 UMH r16 SATD at 4 levels,
SmoothFps algorithm 23 at 24->60 with masks): three frames of the layered generator
go through the batched HIP path and the CPU oracle; the super frames, both vector
payloads and every output frame are byte-identical.  Config 3 (3840x2160, pel 1,
16x16 overlap 3 exhaustive r8 SAD at 5 levels): one frame pair, Super and Analyse.
"""
import numpy as np
import pytest

import svpf
import synth

pytestmark = pytest.mark.gpu

CFG2 = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
CFG2_SMOOTH = "{rate:{num:5,den:2},algo:23,mask:{area:100},scene:{mode:0}}"
CFG3 = "{block:{w:16,h:16,overlap:3},main:{levels:5,search:{type:4,distance:8,satd:false,coarse:{satd:false}}}}"


def _run(ctx, frames, sopt, aopt):
    h, w = frames[0][0].shape
    sp, sdata = svpf.super_params(sopt, w, h)
    sh = svpf.super_height(sp)
    strides = (w, w // 2, w // 2)
    src = [svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2)) for _ in frames]
    sup = [svpf.DeviceFrame.alloc(ctx, strides, (sh, sh // 2, sh // 2)) for _ in frames]
    for d, f in zip(src, frames):
        d.upload(f)
    svpf.super_build_batch(ctx, sp, src, sup)
    ap, vdata, plen = svpf.analyse_params(aopt, sdata, w, h)
    n = len(frames)
    pay = [ctx.alloc(plen) for _ in range(n)]
    svpf.analyse_batch(ctx, ap, sup, [sup[min(i + 1, n - 1)] for i in range(n)], pay)
    ctx.sync()
    return src, sup, pay, sdata, vdata, plen


def test_config2_full_size(gpu_ctx, oracle):
    ctx = gpu_ctx
    print("full size: config 2, 1920x1080, 3 frames", flush=True)
    frames = synth.config_clip(2, frames=3)
    h, w = frames[0][0].shape
    src, sup, pay, sdata, vdata, plen = _run(ctx, frames, "{pel:2}", CFG2)
    hs = [oracle.super_build("{pel:2}", f) for f in frames]
    for i, d in enumerate(sup):
        assert all(np.array_equal(a, b) for a, b in zip(d.download(), hs[i])), f"super frame {i}"
    hp = []
    for i in range(3):
        want = oracle.analyse(CFG2, sdata, w, h, hs[i], hs[min(i + 1, 2)], threads=True)
        got = ctx.download(pay[i], plen)
        assert np.array_equal(got, want), f"vector payload {i}"
        hp.append(got)
    sm = svpf.Smooth(ctx, CFG2_SMOOTH, w, h, 24, 1, 3, sdata, vdata)
    sm.set_sources(0, src, sup, pay)
    outs = [svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2)) for _ in range(sm.output_frames)]
    dec = sm.render(0, outs, decisions=True)
    clip = oracle.SmoothClip(CFG2_SMOOTH, frames, hs, hp, sdata, vdata, fps=(24, 1))
    rendered = 0
    for k, o in enumerate(outs):
        want, kind, algo = clip.render(k)
        assert dec[k] == (kind, algo if kind == 2 else 0), f"decision of output {k}"
        rendered += kind == 2
        for g, q in zip(o.download(), want):
            assert np.array_equal(g[:q.shape[0], :q.shape[1]], q), f"output frame {k}"
    sm.close()
    assert rendered > 0


def test_config3_full_size_analyse(gpu_ctx, oracle):
    ctx = gpu_ctx
    print("full size: config 3, 3840x2160, 1 pair", flush=True)
    frames = synth.config_clip(3, frames=2)
    h, w = frames[0][0].shape
    src, sup, pay, sdata, vdata, plen = _run(ctx, frames, "{pel:1}", CFG3)
    hs = [oracle.super_build("{pel:1}", f) for f in frames]
    for i, d in enumerate(sup):
        assert all(np.array_equal(a, b) for a, b in zip(d.download(), hs[i])), f"super frame {i}"
    want = oracle.analyse(CFG3, sdata, w, h, hs[0], hs[1], threads=True)
    assert np.array_equal(ctx.download(pay[0], plen), want), "vector payload 0"
