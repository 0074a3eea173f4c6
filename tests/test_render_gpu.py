"""GPU parity of svpflow2.SmoothFps: libsvpf's HIP render vs the CPU oracle.

Every output frame of a short clip is rendered by both from the same sources,
Super frames and vector payloads (from the GPU Analyse, itself checked
byte-exact against the oracle in test_analyse_gpu.py); the three planes, the
copy/render decision and the effective algorithm must be identical.  The cases
cover the benchmark settings (algorithm 13 at 30->60, algorithm 23 with masks
and scene detection at 24->60 with a scene cut, algorithm 2 at pel 1) and the
option paths of the render: every algorithm, block vs interpolated vectors,
scene modes 0-3, magnitude masks, coverage strength 0 and the 2x/2.5x/1.5x
rates, rendered as one batch and frame by frame.
"""
import numpy as np
import pytest

import svpf
import synth

pytestmark = pytest.mark.gpu

A1 = "{block:{w:16,h:16,overlap:0},main:{search:{type:2,distance:8,satd:false,coarse:{satd:false}}}}"
A2 = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
A3 = "{block:{w:16,h:16,overlap:3},main:{levels:5,search:{type:4,distance:8,satd:false,coarse:{satd:false}}}}"


class Pipeline:
    """Sources, Super frames and vector payloads of one clip on the device."""

    def __init__(self, ctx, frames, super_opt, analyse_opt):
        self.ctx = ctx
        self.frames = frames
        h, w = frames[0][0].shape
        self.w, self.h = w, h
        sp, self.sdata = svpf.super_params(super_opt, w, h)
        sh = svpf.super_height(sp)
        self.src, self.sup = [], []
        for f in frames:
            s = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
            s.upload(f)
            d = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (sh, sh // 2, sh // 2))
            self.src.append(s)
            self.sup.append(d)
        svpf.super_build_batch(ctx, sp, self.src, self.sup)
        ap, self.vdata, plen = svpf.analyse_params(analyse_opt, self.sdata, w, h)
        n = len(frames)
        self.pay = [ctx.alloc(plen) for _ in range(n)]
        refs = [self.sup[min(i + 1, n - 1)] for i in range(n)]
        svpf.analyse_batch(ctx, ap, self.sup, refs, self.pay)
        ctx.sync()
        self.host_sup = [d.download() for d in self.sup]
        self.host_pay = [ctx.download(p, plen) for p in self.pay]

    def outputs(self, n, size=None):
        w, h = size or (self.w, self.h)
        return [svpf.DeviceFrame.alloc(self.ctx, (w, w // 2, w // 2), (h, h // 2, h // 2)) for _ in range(n)]


def _check(ctx, oracle, pipe, opt, fps=(30, 1), per_frame=False, expect_render=True):
    print(f"render parity: {pipe.w}x{pipe.h} {opt} fps={fps} per_frame={per_frame}", flush=True)
    n = len(pipe.frames)
    sm = svpf.Smooth(ctx, opt, pipe.w, pipe.h, fps[0], fps[1], n, pipe.sdata, pipe.vdata)
    outs = []
    try:
        sm.set_sources(0, pipe.src, pipe.sup, pipe.pay)
        nout = sm.output_frames
        outs = pipe.outputs(nout, sm.output_size)
        if per_frame:
            dec = []
            for k in range(nout):
                dec += sm.render(k, outs[k:k + 1], decisions=True)
        else:
            dec = sm.render(0, outs, decisions=True)
        ctx.sync()
        clip = oracle.SmoothClip(opt, pipe.frames, pipe.host_sup, pipe.host_pay, pipe.sdata, pipe.vdata, fps=fps)
        assert nout == oracle.smooth_opts(opt, fps[0], fps[1], pipe.vdata, n)[0]["out_frames"]
        rendered = 0
        px, py = clip.padding()
        assert sm.output_size == (pipe.w + 2 * px, pipe.h + 2 * py)
        for k in range(nout):
            want, rc, algo = clip.render(k)
            got = outs[k].download()
            assert got[0].shape == want[0].shape
            assert dec[k] == (rc, algo if rc == 2 else 0), f"frame {k}: decision {dec[k]} want {(rc, algo)}"
            rendered += rc == 2
            for p in range(3):
                g = got[p][:want[p].shape[0], :want[p].shape[1]]
                if not np.array_equal(g, want[p]):
                    bad = np.argwhere(g != want[p])
                    y, x = bad[0]
                    raise AssertionError(f"frame {k} plane {p} (algo {algo}): {len(bad)} samples differ, first at "
                                         f"({x},{y}) got {g[y, x]} want {want[p][y, x]}")
        if expect_render:
            assert rendered > 0, "no frame was interpolated"
    finally:
        sm.close()
        for o in outs:
            for ptr in (o.y, o.u, o.v):
                ctx.free(ptr)
    return dec


@pytest.fixture(scope="module")
def pipe1(gpu_ctx):
    return Pipeline(gpu_ctx, synth.config_clip(1, frames=5), "{pel:2}", A1)


@pytest.fixture(scope="module")
def pipe2(gpu_ctx):
    frames = synth.clip(480, 272, 6, seed=2024, vmax=12.0, nobj=(3, 5), rot_deg=2.0, cut_every=4)
    return Pipeline(gpu_ctx, frames, "{pel:2}", A2)


def test_render_cfg1_settings(gpu_ctx, oracle, pipe1):
    # With the default scene.mode 3 the source frame is rounded to the nearest one, so at
    # exactly 2x every odd output frame lands on phase 0 and is a copy (core.rs:2054-2077).
    dec = _check(gpu_ctx, oracle, pipe1, "{rate:{num:2,den:1},algo:13}", expect_render=False)
    assert all(d[0] in (0, 1) for d in dec)
    _check(gpu_ctx, oracle, pipe1, "{rate:{num:2,den:1},algo:13,scene:{mode:0}}")


def test_render_cfg2_settings(gpu_ctx, oracle, pipe2):
    dec = _check(gpu_ctx, oracle, pipe2, "{rate:{num:5,den:2},algo:23,mask:{area:100}}", fps=(24, 1))
    assert any(d == (2, 23) for d in dec)


def test_render_cfg2_settings_per_frame(gpu_ctx, oracle, pipe2):
    _check(gpu_ctx, oracle, pipe2, "{rate:{num:5,den:2},algo:23,mask:{area:100}}", fps=(24, 1), per_frame=True)


def test_render_cfg3_settings_small(gpu_ctx, oracle):
    frames = synth.clip(512, 288, 4, seed=3030, vmax=12.0, nobj=(3, 5), rot_deg=2.0)
    pipe = Pipeline(gpu_ctx, frames, "{pel:1}", A3)
    _check(gpu_ctx, oracle, pipe, "{rate:{num:2,den:1},algo:2,scene:{mode:0}}")


@pytest.mark.parametrize("opt", [
    "{rate:{num:2,den:1},algo:1,scene:{mode:0}}",
    "{rate:{num:5,den:2},algo:2,scene:{mode:0}}",
    "{rate:{num:2,den:1},algo:11,scene:{mode:0}}",
    "{rate:{num:5,den:2},algo:13,block:true,scene:{mode:0}}",
    "{rate:{num:2,den:1},algo:21,scene:{mode:0}}",
    "{rate:{num:5,den:2},algo:22,scene:{mode:0},mask:{area:50,area_sharp:1.0}}",
    "{rate:{num:5,den:2},algo:23,block:true,scene:{mode:0}}",
    "{rate:{num:5,den:2},algo:23,scene:{mode:0},mask:{cover:60,area:200}}",
    "{rate:{num:2,den:1},algo:21,scene:{mode:0},mask:{cover:0}}",
    "{rate:{num:5,den:2},algo:13,mask:{area:100}}",
    "{rate:{num:5,den:2},algo:23,scene:{mode:3,adaptive:111,force13:false}}",
    "{rate:{num:3,den:1},algo:21,scene:{mode:1}}",
    "{rate:{num:3,den:1},algo:22,scene:{mode:2,blend:true}}",
    "{rate:{num:5,den:2},algo:21,scene:{mode:0,limits:{zero:0,m1:10,m2:20,scene:40000,blocks:50}}}",
    "{rate:{num:3,den:2},algo:13,scene:{mode:0}}",
    "{rate:{num:5,den:2},algo:23,scene:{mode:0},mask:{area:100,area_sharp:0.5}}",
    "{rate:{num:5,den:2},algo:13,scene:{mode:0},mask:{area:200,area_sharp:2.0}}",
    "{rate:{num:2,den:1},algo:21,scene:{mode:0},mask:{area:40,area_sharp:0.37}}",
])
def test_render_options(gpu_ctx, oracle, pipe1, opt):
    _check(gpu_ctx, oracle, pipe1, opt)


@pytest.mark.parametrize("light", [
    "light:{aspect:2.35}",                                   # left and right strips only
    "light:{aspect:1.33}",                                   # top and bottom strips only
    "light:{zoom:10}",                                       # all four strips and the corners
    "light:{zoom:10,border:4,lights:5,length:20,cell:2.5}",  # chroma's 9-row strips unfaded
    "light:{aspect:2.35,zoom:5,length:0,border:-3}",         # corners of weight 0, border at least 1
    "light:{zoom:3,lights:1,cell:0.05}",                     # 8 x 6 pads, one control point
    "light:{aspect:1.0,sar:0.5,lights:64,cell:40}",          # windows wider than the frame
])
def test_render_light_border(gpu_ctx, oracle, pipe1, light):
    """light.* padding: the output is widened by Options::padding (options.rs:256-277), the
    frame rendered (or copied) at the pad offset and the light border drawn around it on
    every output frame (light.rs), as the oracle restates it."""
    _check(gpu_ctx, oracle, pipe1, "{rate:{num:2,den:1},algo:13,scene:{mode:0}," + light + "}")


def test_render_light_border_algo23_per_frame(gpu_ctx, oracle, pipe2):
    for per_frame in (False, True):
        _check(gpu_ctx, oracle, pipe2, "{rate:{num:5,den:2},algo:23,mask:{area:100},light:{zoom:6,aspect:2.0}}",
               fps=(24, 1), per_frame=per_frame)


def test_render_cfg2_magnitude_exponents(gpu_ctx, oracle, pipe2):
    # mask.area_sharp != 1: the magnitude mask reads the host-built pow table
    for sharp in (0.5, 2.0):
        _check(gpu_ctx, oracle, pipe2, "{rate:{num:5,den:2},algo:23,mask:{area:100,area_sharp:%s},scene:{mode:0}}"
               % sharp, fps=(24, 1))


def test_render_cfg2_vectors_all_algos(gpu_ctx, oracle, pipe2):
    for a in (1, 2, 11, 13, 21, 22):
        _check(gpu_ctx, oracle, pipe2, f"{{rate:{{num:5,den:2}},algo:{a},scene:{{mode:0}}}}", fps=(24, 1))


def test_render_rejects_options(gpu_ctx):
    """Create-time rejections with the reference's messages (options.rs:169-203, filter.rs:578-709)."""
    w, h = 640, 360
    _, sdata = svpf.super_params("{pel:2}", w, h)
    _, vdata, _ = svpf.analyse_params(A1, sdata, w, h)
    _, sdata0 = svpf.super_params("{pel:2,gpu:0}", w, h)
    _, vdata0, _ = svpf.analyse_params(A1, sdata0, w, h)
    assert vdata[13] == 1 and vdata0[13] == 0

    def make(opt, sd=sdata, vd=vdata, size=(w, h)):
        return svpf.Smooth(gpu_ctx, opt, size[0], size[1], 30, 1, 5, sd, vd)

    def rejects(msg, *args, **kw):
        with pytest.raises(svpf.SvpfError, match=msg.replace("[", r"\[").replace("]", r"\]")):
            make(*args, **kw)

    rejects("SVSmoothFps: incorrect 'algo' value", "{algo:5}")
    rejects("SVSmoothFps: 'scene.mode' must be in [0;3]", "{scene:{mode:4}}")
    rejects("SVSmoothFps: 'scene.mode' value isn't supported for selected frame rate", "{rate:{num:1,den:2},scene:{mode:2}}")
    rejects("SVSmoothFps: 'cubic' mode isn't available with CPU rendering", "{cubic:1}", sdata0, vdata0)
    make("{cubic:1}").close()  # selector 1: accepted (the render keeps the integer CPU semantics)
    rejects("SVSmoothFps: invalid super clip params", "{}", sdata0, vdata)
    rejects("SVSmoothFps: source and vectors frame sizes are different", "{}", size=(320, 180))
    bad = list(vdata)
    bad[14] = 2
    rejects("SVSmoothFps: vectors with delta>1 are not allowed.", "{}", vd=bad)
    bad = list(vdata)
    bad[0] = 0
    rejects("SVSmoothFps: invalid vectors stream [2]", "{}", vd=bad)
    sm = make("{rate:{num:5,den:2}}")
    assert sm.output_frames == 11  # (5 - 1) * 5 / 2 + 1 (frame_math.rs:128-135)
    sm.close()
