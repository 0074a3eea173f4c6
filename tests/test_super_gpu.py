"""GPU parity of svpflow1.Super: libsvpf's HIP kernels vs the CPU oracle, byte for byte.

Covers stride == width (the portable layout) and padded strides (the VapourSynth
layout), widths that are not multiples of 4, every pel value, the bicubic pel-2
fill (gpu=0), a disabled pyramid (scale.down != 4) and the batched entry point.
"""
import numpy as np
import pytest

import svpf
import synth

pytestmark = pytest.mark.gpu


def _rand_frame(rng, w, h):
    return (rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
            rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))


def _run(ctx, opt, planes, strides, dst_strides):
    h, w = planes[0].shape
    p, sdata = svpf.super_params(opt, w, h)
    sh = svpf.super_height(p)
    src = svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2))
    dst = svpf.DeviceFrame.alloc(ctx, dst_strides, (sh, sh // 2, sh // 2))
    src.upload(planes)
    for ptr, s, r in zip((dst.y, dst.u, dst.v), dst.strides, dst.rows):
        ctx.memset(ptr, 0xA5, s * r)  # poison: every byte must be written
    svpf.super_build(ctx, p, src, dst)
    ctx.sync()
    out = dst.download()
    for ptr in (src.y, src.u, src.v, dst.y, dst.u, dst.v):
        ctx.free(ptr)
    return out, sdata


CASES = [
    ("{pel:2}", 640, 360, 0, 0),
    ("{pel:2}", 640, 360, 32, 32),
    ("{pel:2}", 854, 480, 0, 0),       # width not a multiple of 4, stride == width
    ("{pel:2}", 854, 480, 10, 10),
    ("{pel:1}", 320, 180, 0, 16),
    ("{pel:4}", 96, 64, 0, 0),
    ("{pel:2, gpu:0}", 200, 120, 0, 8),  # bicubic pel-2 fill
    ("{pel:2, gpu:2}", 128, 72, 0, 0),  # no pel fill
    ("{pel:2, scale:{down:2}}", 128, 72, 0, 0),  # no pyramid
    ("{pel:2}", 34, 18, 0, 0),          # tiny, few levels
    ("{pel:2}", 1920, 1080, 0, 0),
]


@pytest.mark.parametrize("opt,w,h,spad,dpad", CASES)
def test_super_matches_oracle(gpu_ctx, oracle, opt, w, h, spad, dpad):
    print(f"super parity: opt={opt} {w}x{h} spad={spad} dpad={dpad}", flush=True)
    rng = np.random.default_rng(w * 7 + h + spad + dpad)
    planes = _rand_frame(rng, w, h)
    strides = (w + spad, w // 2 + spad, w // 2 + spad)
    dstrides = (w + dpad, w // 2 + dpad, w // 2 + dpad)
    got, sdata = _run(gpu_ctx, opt, planes, strides, dstrides)
    ref = oracle.super_build(opt, planes, strides, dstrides)
    o, _ = oracle.super_opts(opt, w, h)
    assert sdata == o["sdata"]
    for pi in range(3):
        assert got[pi].shape == ref[pi].shape
        bad = np.argwhere(got[pi] != ref[pi])
        assert bad.size == 0, f"plane {pi}: {len(bad)} bytes differ, first at {bad[:3].tolist()}"


def test_super_batch_on_synthetic_clip(gpu_ctx, oracle):
    frames = synth.config_clip(1, frames=4)
    h, w = frames[0][0].shape
    p, _ = svpf.super_params("{pel:2}", w, h)
    sh = svpf.super_height(p)
    strides = (w, w // 2, w // 2)
    srcs, dsts = [], []
    for f in frames:
        s = svpf.DeviceFrame.alloc(gpu_ctx, strides, (h, h // 2, h // 2))
        s.upload(f)
        srcs.append(s)
        dsts.append(svpf.DeviceFrame.alloc(gpu_ctx, strides, (sh, sh // 2, sh // 2)))
    svpf.super_build_batch(gpu_ctx, p, srcs, dsts)
    gpu_ctx.sync()
    for f, d in zip(frames, dsts):
        ref = oracle.super_build("{pel:2}", f)
        got = d.download()
        for pi in range(3):
            assert (got[pi] == ref[pi]).all()
