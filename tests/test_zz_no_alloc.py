"""svpf_analyse_pair_noalloc (the allocation-free form of svpf_analyse_pair) through the
C ABI: its payload equals svpf_analyse_pair's byte for byte on calls in a row, each on a
smaller frame pair than the last (the shared scratch array keeps the previous call's
contents), starting at the largest documented size, and the smallest equals the oracle; a pair larger than the scratch array
is rejected with an error that says so."""
import ctypes as C

import numpy as np
import pytest

import svpf
import synth

pytestmark = pytest.mark.gpu

CFG2 = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
SMALL = "{block:{w:8,h:8,overlap:2},main:{search:{type:3,distance:8,satd:true,coarse:{bad:{sad:200}}}}}"


def _noalloc():
    L = svpf.lib()
    f = L.svpf_analyse_pair_noalloc
    f.argtypes = [C.c_void_p, C.POINTER(svpf.AnalyseParams), C.POINTER(svpf.Frame), C.POINTER(svpf.Frame),
                  C.c_void_p]
    f.restype = C.c_int
    return f


def _supers(ctx, frames, sopt):
    h, w = frames[0][0].shape
    p, sdata = svpf.super_params(sopt, w, h)
    sh = svpf.super_height(p)
    out = []
    for f in frames:
        src = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
        src.upload(f)
        dst = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (sh, sh // 2, sh // 2))
        svpf.super_build(ctx, p, src, dst)
        out.append(dst)
    return sdata, out


def test_noalloc_equals_analyse_pair(gpu_ctx, oracle):
    ctx = gpu_ctx
    fn = _noalloc()
    # the largest pair the scratch array is documented to hold (2560x1440, 8x8 blocks,
    # overlap 4), then smaller ones
    for w, h, aopt, seed in ((2560, 1440, CFG2, 70), (1280, 720, CFG2, 71), (320, 192, SMALL, 72)):
        print(f"noalloc: {w}x{h} {aopt}", flush=True)
        frames = synth.clip(w, h, 2, seed=seed, vmax=8.0, nobj=(2, 3), rot_deg=1.0)
        sdata, sup = _supers(ctx, frames, "{pel:2}")
        p, _, plen = svpf.analyse_params(aopt, sdata, w, h)
        want_d, got_d = ctx.alloc(plen), ctx.alloc(plen)
        ctx.memset(want_d, 0x11, plen)
        ctx.memset(got_d, 0x77, plen)
        svpf.analyse_pair(ctx, p, sup[0], sup[1], want_d)
        c, r = sup[0].desc(), sup[1].desc()
        svpf.check(fn(ctx.h, C.byref(p), C.byref(c), C.byref(r), C.c_void_p(got_d)))
        ctx.sync()
        want, got = ctx.download(want_d, plen), ctx.download(got_d, plen)
        assert np.array_equal(got, want), f"{w}x{h}: {int((got != want).sum())} payload bytes differ"
        if w == 320:
            hosts = [d.download() for d in sup]
            assert np.array_equal(got, oracle.analyse(aopt, sdata, w, h, hosts[0], hosts[1]))


def test_noalloc_rejects_a_pair_beyond_its_scratch(gpu_ctx):
    ctx = gpu_ctx
    fn = _noalloc()
    w, h = 3840, 2160
    p, sdata = svpf.super_params("{pel:2}", w, h)
    sh = svpf.super_height(p)
    sup = [svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (sh, sh // 2, sh // 2)) for _ in range(2)]
    ap, _, plen = svpf.analyse_params("{block:{w:4,h:4,overlap:0},main:{search:{type:3,distance:8}}}", sdata, w, h)
    pay = ctx.alloc(plen)
    c, r = sup[0].desc(), sup[1].desc()
    rc = fn(ctx.h, C.byref(ap), C.byref(c), C.byref(r), C.c_void_p(pay))
    assert rc == -1
    msg = svpf.lib().svpf_last_error().decode()
    assert "scratch memory" in msg and "maximum of 134217728" in msg, msg
