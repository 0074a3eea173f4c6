"""GPU tests of the multi-clip batch and the serving loop.

* svpf_analyse_multi: frame pairs of clips with different sizes (and so different
  grids and pyramid depths) searched in one launch set give, pair by pair, the
  oracle's vector payload.
* Serving: a stream's output frames are byte-identical whether it is served alone
  or together with streams of other sizes and rates that share every step, and
  they equal the oracle's SmoothFps render of the same clip.
"""
import numpy as np
import pytest

import serve
import svpf
import synth

pytestmark = pytest.mark.gpu


def _upload(ctx, frames):
    h, w = frames[0][0].shape
    out = []
    for f in frames:
        d = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
        d.upload(f)
        out.append(d)
    return out


def test_analyse_multi_clips_match_oracle(gpu_ctx, oracle):
    ctx = gpu_ctx
    clips = [synth.clip(320, 192, 3, seed=501, vmax=8.0, nobj=(2, 3)),
             synth.clip(256, 144, 2, seed=502, vmax=6.0, nobj=(1, 2)),
             synth.clip(480, 272, 2, seed=503, vmax=10.0, nobj=(2, 4), rot_deg=2.0)]
    params, curs, refs, pays, want = [], [], [], [], []
    for frames in clips:
        h, w = frames[0][0].shape
        print(f"analyse_multi: clip {w}x{h}, {len(frames)} frames", flush=True)
        sp, sdata = svpf.super_params("{pel:2}", w, h)
        sh = svpf.super_height(sp)
        src = _upload(ctx, frames)
        sup = [svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (sh, sh // 2, sh // 2)) for _ in frames]
        svpf.super_build_batch(ctx, sp, src, sup)
        ap, vdata, plen = svpf.analyse_params(serve.ANALYSE_OPT, sdata, w, h)
        hs = [oracle.super_build("{pel:2}", f) for f in frames]
        for i in range(len(frames)):
            j = min(i + 1, len(frames) - 1)
            params.append(ap)
            curs.append(sup[i])
            refs.append(sup[j])
            pays.append(ctx.alloc(plen))
            want.append((plen, oracle.analyse(serve.ANALYSE_OPT, sdata, w, h, hs[i], hs[j])))
    svpf.analyse_multi(ctx, params, curs, refs, pays)
    ctx.sync()
    for i, (d, (plen, ref)) in enumerate(zip(pays, want)):
        got = ctx.download(d, plen)
        assert np.array_equal(got, ref), f"multi-clip pair {i} differs from the oracle"


class _Pool:
    def __init__(self, ctx, frames):
        self.frames = frames
        self.dev = _upload(ctx, frames)

    def __call__(self, i):
        return self.dev[i]


def _serve(ctx, specs, keep):
    got = {}

    def on_output(sid, k, frame):
        if sid in keep:
            got[(sid, k)] = frame.download()

    srv = serve.Server(ctx, budget_bytes=8 << 30, out_per_step=4, on_output=on_output)
    stats = srv.run(specs)
    srv.close()
    return got, stats


def test_stream_alone_equals_in_mix_and_oracle(gpu_ctx, oracle):
    ctx = gpu_ctx
    fa = synth.clip(320, 192, 8, seed=601, vmax=8.0, nobj=(2, 3))
    fb = synth.clip(256, 144, 7, seed=602, vmax=6.0, nobj=(1, 2))
    fc = synth.clip(320, 192, 6, seed=603, vmax=10.0, nobj=(2, 2), rot_deg=2.0)
    pa, pb, pc = _Pool(ctx, fa), _Pool(ctx, fb), _Pool(ctx, fc)
    a = serve.StreamSpec(0, 320, 192, 8, (5, 2), (24, 1), 0.0, pa)
    b = serve.StreamSpec(1, 256, 144, 7, (2, 1), (30, 1), 0.0, pb)
    c = serve.StreamSpec(2, 320, 192, 6, (5, 2), (24, 1), 0.05, pc)
    print("serve: stream 0 alone", flush=True)
    alone, st1 = _serve(ctx, [a], {0})
    print("serve: streams 0, 1, 2 together", flush=True)
    mixed, st2 = _serve(ctx, [a, b, c], {0})
    assert st1["finished"] == 1 and st2["finished"] == 3 and st2["dropped"] == 0
    keys = sorted(alone)
    assert keys == sorted(mixed) and len(keys) > 0
    for key in keys:
        for x, y in zip(alone[key], mixed[key]):
            assert np.array_equal(x, y), f"output {key} differs alone vs in the mix"
    # the oracle's render of the same clip
    sopt = serve.SUPER_OPT
    so, _ = oracle.super_opts(sopt, 320, 192)
    hs = [oracle.super_build(sopt, f) for f in fa]
    pays = [oracle.analyse(serve.ANALYSE_OPT, so["sdata"], 320, 192, hs[i], hs[min(i + 1, 7)]) for i in range(8)]
    ao, _ = oracle.analyse_opts(serve.ANALYSE_OPT, so["sdata"], 320, 192)
    clip = oracle.SmoothClip(serve.smooth_opt((5, 2)), fa, hs, pays, so["sdata"], ao["vdata"], fps=(24, 1))
    assert len(keys) == oracle.smooth_opts(serve.smooth_opt((5, 2)), 24, 1, ao["vdata"], 8)[0]["out_frames"]
    for (_, k) in keys:
        want, _, _ = clip.render(k)
        for g, q in zip(alone[(0, k)], want):
            assert np.array_equal(g[:q.shape[0], :q.shape[1]], q), f"served output {k} differs from the oracle"
