"""The VapourSynth plugin shells (libsvpflow1_vs.so, libsvpflow2_vs.so) driven by a
fake VapourSynth API 3 core (tests/vsfake.py).

CPU: registration (identifiers, namespaces, function names and argument strings, API
version, svpGetVersion; crates/svpflow1/src/abi.rs:3-46, svpflow2/src/abi.rs,
strings.rs:3-17), create-time checks and messages, and the output clips' video info.
GPU: a clip through Super -> Analyse -> SmoothFps frame by frame, each stage
byte-identical to the oracle, with the outputs' duration properties.
"""
import os

import numpy as np
import pytest

import svpf
import synth
import vsfake

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(HERE, "..", "open-svpflow_amd")
LIB1 = os.path.join(PKG, "libsvpflow1_vs.so")
LIB2 = os.path.join(PKG, "libsvpflow2_vs.so")

AOPT = b"{block:{w:8,h:8,overlap:3},main:{search:{type:3,distance:8,satd:true}}}"
ROPT = b"{rate:{num:5,den:2},algo:23,mask:{area:100},scene:{mode:0}}"


@pytest.fixture
def core():
    if not (os.path.exists(LIB1) and os.path.exists(LIB2)):
        pytest.fail("plugin libraries are missing: run __graft_entry__.build()")
    c = vsfake.Core()
    c.load(LIB1)
    c.load(LIB2)
    yield c
    c.close()


def _clip(core, w=160, h=96, n=4, seed=11):
    frames = synth.clip(w, h, n, seed=seed, vmax=3.0, nobj=(1, 1))
    props = lambda i: {"_DurationNum": 1, "_DurationDen": 24, "_PTS": 1000 + 41667 * i}
    return frames, core.source(frames, fps=(24, 1), props=props)


def test_registration(core):
    p1, p2 = core.plugins["svp1"], core.plugins["svp2"]
    assert (p1["id"], p1["name"], p1["api"], p1["readonly"]) == ("com.svp-team.flow1", "SVPFlow1", 0x30002, 1)
    assert (p2["id"], p2["name"], p2["api"]) == ("com.svp-team.flow2", "SVPFlow2", 0x30002)
    assert p1["version"] == p2["version"] == 1174405393 == svpf.lib().svpf_version()
    assert {k: v[0] for k, v in p1["functions"].items()} == {
        "Super": "clip:clip;opt:data", "Analyse": "clip:clip;sdata:int;src:clip;opt:data"}
    f2 = {k: v[0] for k, v in p2["functions"].items()}
    assert f2["SmoothFps"] == ("clip:clip;super:clip;sdata:int;vectors:clip;vdata:int;opt:data;src:clip:opt;"
                               "fps:float:opt")
    assert set(f2) == {"SmoothFps", "SmoothFps_NVOF", "SmoothFps_RIFE"}


def test_create_checks_and_clip_info(core):
    frames, src = _clip(core)
    with pytest.raises(vsfake.VSError, match="SVSuper: clip is required"):
        core.invoke("svp1", "Super", opt=b"{pel:2}")
    with pytest.raises(vsfake.VSError) as e:  # the reference's option message, as the C ABI gives it
        core.invoke("svp1", "Super", clip=src, opt=b"{pel:3}")
    with pytest.raises(svpf.SvpfError) as e2:
        svpf.super_params("{pel:3}", 160, 96)
    assert str(e.value) == str(e2.value)
    out = core.invoke("svp1", "Super", clip=src, opt=b"{pel:2}")
    sup, sdata = out["clip"][0][1], out["data"][0][1]
    p, want_sdata = svpf.super_params("{pel:2}", 160, 96)
    vi = core.video_info(sup)
    assert sdata == want_sdata and (vi.width, vi.height, vi.num_frames) == (160, svpf.super_height(p), 4)
    with pytest.raises(vsfake.VSError, match="SVAnalyse: src required"):
        core.invoke("svp1", "Analyse", clip=sup, sdata=sdata, opt=AOPT)
    with pytest.raises(vsfake.VSError, match="SVAnalyse: sdata required"):
        core.invoke("svp1", "Analyse", clip=sup, src=src, opt=AOPT)
    a = core.invoke("svp1", "Analyse", clip=sup, sdata=sdata, src=src, opt=AOPT)
    vec, vdata_addr = a["clip"][0][1], a["data"][0][1]
    _, vdata, plen = svpf.analyse_params(AOPT.decode(), sdata, 160, 96)
    vvi = core.video_info(vec)
    assert (vvi.width, vvi.height, vvi.num_frames) == (plen, 1, 4)
    assert vsfake.FORMATS[vsfake.Format.from_address(vvi.format).id] is vsfake.GRAY8
    # the `data` property is the address of the 16-word vdata header (analyse_filter.rs:117-120)
    import ctypes
    assert list((ctypes.c_int32 * 16).from_address(vdata_addr)) == list(vdata)
    with pytest.raises(vsfake.VSError, match="SVSmoothFps: incorrect 'algo' value"):
        core.invoke("svp2", "SmoothFps", clip=src, super=sup, sdata=sdata, vectors=vec, vdata=vdata_addr,
                    opt=b"{rate:{num:2,den:1},algo:7}")
    with pytest.raises(vsfake.VSError, match="SVSmoothFps: invalid vectors stream"):
        core.invoke("svp2", "SmoothFps", clip=src, super=sup, sdata=sdata, vectors=vec, opt=ROPT)
    # light.* padding (Options::padding, options.rs:256-277) is drawn within stated limits;
    # settings beyond them are rejected rather than drawn differently
    for light, msg in ((b"light:{zoom:10,border:300}", "light.border above 255 is not supported"),
                       (b"light:{zoom:10,border:97}", "light.border deeper than the frame"),
                       (b"light:{zoom:10,lights:5000}", "light.lights above 4096 a side is not supported"),
                       (b"light:{aspect:1.0,sar:0.5,cell:2000}", "light.cell above 1000 is not supported")):
        with pytest.raises(vsfake.VSError, match=msg):
            core.invoke("svp2", "SmoothFps", clip=src, super=sup, sdata=sdata, vectors=vec, vdata=vdata_addr,
                        opt=b"{rate:{num:2,den:1}," + light + b"}")
    with pytest.raises(vsfake.VSError, match="SVSmoothFps: CUDA is not availabe"):
        core.invoke("svp2", "SmoothFps_NVOF", clip=src, opt=b"{}")


@pytest.mark.gpu
def test_plugin_chain_matches_oracle(core, oracle):
    frames, src = _clip(core, n=4)
    print("vs plugins: 160x96, 4 frames through Super / Analyse / SmoothFps", flush=True)
    out = core.invoke("svp1", "Super", clip=src, opt=b"{pel:2}")
    sup, sdata = out["clip"][0][1], out["data"][0][1]
    a = core.invoke("svp1", "Analyse", clip=sup, sdata=sdata, src=src, opt=AOPT)
    vec, vdata_addr = a["clip"][0][1], a["data"][0][1]
    s = core.invoke("svp2", "SmoothFps", clip=src, super=sup, sdata=sdata, vectors=vec, vdata=vdata_addr, opt=ROPT)
    smooth = s["clip"][0][1]
    # a Super frame depends on its row stride (the pel fill reads one byte past the
    # width, super_build.rs:190,217): the oracle builds with the plugin frames' strides
    st = tuple(core.frame(sup, 0).planes[pl].shape[1] for pl in range(3))
    hs = [oracle.super_build("{pel:2}", f, strides=st, dst_strides=st) for f in frames]
    for i in range(4):
        got = core.frame(sup, i)
        for pl in range(3):
            assert np.array_equal(got.planes[pl], hs[i][pl]), f"super frame {i} plane {pl}"
    so, _ = oracle.super_opts("{pel:2}", 160, 96)
    pays = []
    for i in range(4):
        want = oracle.analyse(AOPT.decode(), so["sdata"], 160, 96, hs[i], hs[min(i + 1, 3)])
        row = core.frame(vec, i).planes[0][0]
        assert np.array_equal(row[:len(want)], want) and not row[len(want):].any(), f"vector frame {i}"
        pays.append(want)
    ao, _ = oracle.analyse_opts(AOPT.decode(), so["sdata"], 160, 96)
    clip = oracle.SmoothClip(ROPT.decode(), frames, hs, pays, so["sdata"], ao["vdata"], fps=(24, 1))
    svi = core.video_info(smooth)
    assert (svi.fps_num, svi.fps_den) == (60, 1)
    for k in range(svi.num_frames):
        got = core.frame(smooth, k)
        want, _, _ = clip.render(k)
        for pl in range(3):
            assert np.array_equal(got.view(pl), want[pl][:got.view(pl).shape[0], :got.view(pl).shape[1]]), \
                f"output frame {k} plane {pl}"
        # every output lasts 1/60 s: (1/24 s) / 2.5, in units of 1e-7 s (frame.rs:111-138)
        assert got.props["_DurationNum"][0][1] == 10_000_000 // 60 and got.props["_DurationDen"][0][1] == 10_000_000


@pytest.mark.gpu
def test_plugin_light_padding_matches_oracle(core, oracle):
    """SmoothFps with light.* padding: the output clip is widened (output_info, core.rs
    :603-613) and every frame carries the light border (light.rs), as the oracle draws it."""
    frames, src = _clip(core, n=4)
    print("vs plugins: 160x96 with light.zoom padding", flush=True)
    out = core.invoke("svp1", "Super", clip=src, opt=b"{pel:2}")
    sup, sdata = out["clip"][0][1], out["data"][0][1]
    a = core.invoke("svp1", "Analyse", clip=sup, sdata=sdata, src=src, opt=AOPT)
    vec, vdata_addr = a["clip"][0][1], a["data"][0][1]
    ropt = b"{rate:{num:5,den:2},algo:23,mask:{area:100},scene:{mode:0},light:{zoom:12,aspect:2.0}}"
    s = core.invoke("svp2", "SmoothFps", clip=src, super=sup, sdata=sdata, vectors=vec, vdata=vdata_addr, opt=ropt)
    smooth = s["clip"][0][1]
    st = tuple(core.frame(sup, 0).planes[pl].shape[1] for pl in range(3))
    hs = [oracle.super_build("{pel:2}", f, strides=st, dst_strides=st) for f in frames]
    so, _ = oracle.super_opts("{pel:2}", 160, 96)
    pays = [oracle.analyse(AOPT.decode(), so["sdata"], 160, 96, hs[i], hs[min(i + 1, 3)]) for i in range(4)]
    ao, _ = oracle.analyse_opts(AOPT.decode(), so["sdata"], 160, 96)
    clip = oracle.SmoothClip(ropt.decode(), frames, hs, pays, so["sdata"], ao["vdata"], fps=(24, 1))
    px, py = clip.padding()
    assert px > 0 and py > 0
    svi = core.video_info(smooth)
    assert (svi.width, svi.height) == (160 + 2 * px, 96 + 2 * py)
    for k in range(svi.num_frames):
        got = core.frame(smooth, k)
        want, _, _ = clip.render(k)
        for pl in range(3):
            assert np.array_equal(got.view(pl), want[pl]), f"output frame {k} plane {pl}"
