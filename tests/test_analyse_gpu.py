"""GPU parity of svpflow1.Analyse: libsvpf's HIP search vs the CPU oracle.

Vectors, costs and luma DC of every block (the whole vector payload) must be
byte-identical, for the benchmark settings (hex2/SAD 16x16, UMH/SATD 8x8 with
overlap, exhaustive pel 1) and for option paths that change the search
(single direction, distance 0, non-square blocks, negative pnew, positive
bad-block range, adaptive coarse radius, padded strides).
"""
import numpy as np
import pytest

import svpf
import synth

pytestmark = pytest.mark.gpu

CFG1 = "{block:{w:16,h:16,overlap:0},main:{search:{type:2,distance:8,satd:false,coarse:{satd:false}}}}"
CFG2 = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
CFG3 = "{block:{w:16,h:16,overlap:3},main:{levels:5,search:{type:4,distance:8,satd:false,coarse:{satd:false}}}}"


def _supers(ctx, oracle, frames, sopt, pad=0):
    h, w = frames[0][0].shape
    p, sdata = svpf.super_params(sopt, w, h)
    sh = svpf.super_height(p)
    strides = (w + pad, w // 2 + pad, w // 2 + pad)
    devs, hosts = [], []
    for f in frames:
        src = svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2))
        src.upload(f)
        dst = svpf.DeviceFrame.alloc(ctx, strides, (sh, sh // 2, sh // 2))
        svpf.super_build(ctx, p, src, dst)
        devs.append(dst)
    ctx.sync()
    for d in devs:
        hosts.append(d.download())
    return sdata, devs, hosts


def _check(ctx, oracle, frames, sopt, aopt, pad=0, batch=False):
    h, w = frames[0][0].shape
    sdata, devs, hosts = _supers(ctx, oracle, frames, sopt, pad)
    p, vdata, plen = svpf.analyse_params(aopt, sdata, w, h)
    pays = [ctx.alloc(plen) for _ in range(len(frames) - 1)]
    if batch:
        svpf.analyse_batch(ctx, p, devs[:-1], devs[1:], pays)
    else:
        for i in range(len(frames) - 1):
            svpf.analyse_pair(ctx, p, devs[i], devs[i + 1], pays[i])
    ctx.sync()
    for i in range(len(frames) - 1):
        got = ctx.download(pays[i], plen)
        ref = oracle.analyse(aopt, sdata, w, h, hosts[i], hosts[i + 1])
        assert got.size == ref.size
        if not np.array_equal(got, ref):
            hg, rg = oracle.decode_payload(got)
            hr, rr = oracle.decode_payload(ref)
            assert (hg == hr).all(), f"header differs: {hg} vs {hr}"
            for k, (a, b) in enumerate(zip(rg, rr)):
                bad = np.argwhere((a != b).any(1))[:, 0]
                assert bad.size == 0, (f"pair {i} region {k}: {bad.size} blocks differ; first {bad[:4].tolist()} "
                                       f"got {a[bad[:4]].tolist()} want {b[bad[:4]].tolist()}")
    for d in pays:
        ctx.free(d)


@pytest.mark.parametrize("aopt", [CFG1, None])
def test_analyse_cfg1_settings(gpu_ctx, oracle, aopt):
    print(f"analyse parity: cfg1 {aopt}", flush=True)
    frames = synth.config_clip(1, frames=3)
    _check(gpu_ctx, oracle, frames, "{pel:2}", aopt)


def test_analyse_cfg2_settings_small(gpu_ctx, oracle):
    print("analyse parity: cfg2 settings 480x272", flush=True)
    frames = synth.clip(480, 272, 3, seed=2024, vmax=12.0, nobj=(3, 5), rot_deg=2.0)
    _check(gpu_ctx, oracle, frames, "{pel:2}", CFG2)


def test_analyse_cfg3_settings_small(gpu_ctx, oracle):
    print("analyse parity: cfg3 settings 512x288 pel1", flush=True)
    frames = synth.clip(512, 288, 2, seed=3033, vmax=10.0, nobj=(3, 5), rot_deg=2.0)
    _check(gpu_ctx, oracle, frames, "{pel:1}", CFG3)


@pytest.mark.parametrize("aopt", [
    "{vectors:1, block:{w:8,h:8,overlap:2}, main:{search:{type:3,distance:8,satd:true}}}",
    "{vectors:2, block:{w:16,h:8,overlap:1}, main:{search:{type:2,distance:6}}}",
    "{block:{w:8,h:8,overlap:0}, main:{search:{type:4,distance:0}}}",
    "{block:{w:8,h:8,overlap:3}, main:{penalty:{pnew:-300, pzero:10}, search:{type:2,distance:6}}}",
    "{block:{w:16,h:16,overlap:2}, main:{search:{type:3,distance:12,coarse:{bad:{range:6, sad:200}}}}}",
    # tile-plan coarse levels: bad-block UMH + hex2 (range > 0), rings (range < 0) with a low
    # threshold, and two levels (the smallest level is level 1: its edge blocks skip refinement)
    "{block:{w:8,h:8,overlap:2}, main:{search:{type:3,distance:12,coarse:{bad:{range:6, sad:150}}}}}",
    "{block:{w:4,h:4,overlap:0}, main:{search:{type:2,distance:6,coarse:{bad:{range:-3, sad:40}}}}}",
    "{block:{w:8,h:8,overlap:2}, main:{levels:2, search:{type:3,distance:8,coarse:{bad:{sad:200}}}}}",
    "{block:{w:8,h:8,overlap:3}, main:{search:{type:2,distance:8,coarse:{distance:-6, satd:false}}}}",
    "{block:{w:4,h:4,overlap:0}, main:{search:{type:2,distance:4,satd:true}}}",
    "{block:{w:32,h:32,overlap:3}, main:{search:{type:3,distance:8,satd:true}}}",
])
def test_analyse_option_paths(gpu_ctx, oracle, aopt):
    print(f"analyse parity: {aopt}", flush=True)
    frames = synth.clip(320, 192, 2, seed=77, vmax=8.0, nobj=(2, 3), rot_deg=1.0)
    _check(gpu_ctx, oracle, frames, "{pel:2}", aopt)


@pytest.mark.parametrize("aopt", [
    # refine pass (exact_recalculate): the main search's type and distance, then explicit
    # ones; thsad 0 re-searches every block, a huge thsad none; 16x16 parents give 8x8
    # output blocks whose border blocks search half-size (factor 3)
    "{block:{w:8,h:8,overlap:2}, main:{search:{type:3,distance:8,satd:true}}, refine:[{thsad:100}]}",
    "{block:{w:16,h:16,overlap:2}, main:{search:{type:2,distance:6}}, "
    "refine:[{thsad:50, search:{type:4,distance:2,satd:true}}]}",
    "{block:{w:16,h:16,overlap:0}, main:{search:{type:2,distance:6}}, refine:[{thsad:0, search:{type:2,distance:4}}]}",
    "{block:{w:16,h:16,overlap:2}, main:{search:{type:3,distance:8}}, refine:[{thsad:1000000}]}",
])
def test_analyse_refine_pass(gpu_ctx, oracle, aopt):
    print(f"analyse parity (refine): {aopt}", flush=True)
    frames = synth.clip(320, 192, 3, seed=78, vmax=8.0, nobj=(2, 3), rot_deg=1.0)
    _check(gpu_ctx, oracle, frames, "{pel:2}", aopt)


def test_analyse_pel4(gpu_ctx, oracle):
    # quarter-pel Super frames: level 0 reads the subplane (mvy & 1) * pel + (mvx & 1) at mv >> 1,
    # as the reference does for every pel >= 2 (analyse_search.rs:1740-1747)
    print("analyse parity: pel 4 super, 8x8 UMH SATD", flush=True)
    frames = synth.clip(256, 144, 2, seed=79, vmax=6.0, nobj=(1, 2), rot_deg=0.0)
    _check(gpu_ctx, oracle, frames, "{pel:4, gpu:0}", "{block:{w:8,h:8,overlap:2}, main:{search:{type:3,distance:8,satd:true}}}")


def test_analyse_batched_padded(gpu_ctx, oracle):
    print("analyse parity: batch of 4 pairs, padded stride", flush=True)
    frames = synth.clip(352, 200, 5, seed=91, vmax=6.0, nobj=(2, 2), rot_deg=0.0)
    _check(gpu_ctx, oracle, frames, "{pel:2}", CFG2, pad=24, batch=True)


@pytest.mark.parametrize("max_sweeps", [8, 24])
def test_analyse_sequential_fallback(gpu_ctx, oracle, max_sweeps):
    """A level whose parallel sweeps do not converge within the limit is searched again by
    the sequential kernel (one wavefront walking the visit order); the payload is the same."""
    print(f"analyse parity: sequential fallback after {max_sweeps} sweeps", flush=True)
    frames = synth.clip(320, 192, 3, seed=81, vmax=8.0, nobj=(2, 3), rot_deg=1.0)
    aopt = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true,coarse:{bad:{sad:300}}}}}"
    gpu_ctx.profile(True)
    gpu_ctx.set_max_sweeps(max_sweeps)
    try:
        _check(gpu_ctx, oracle, frames, "{pel:2}", aopt, batch=True)
        # sweep statistics are recorded only for levels that converged: level 0 of a
        # 52x31 grid needs far more than 24 sweeps, so it fell back
        assert gpu_ctx.analyse_stats(0)["runs"] == 0
    finally:
        gpu_ctx.set_max_sweeps(0)
        gpu_ctx.profile(False)
    with pytest.raises(svpf.SvpfError):
        gpu_ctx.set_max_sweeps(5000)


@pytest.mark.parametrize("tuning", [dict(chase=16), dict(chase=1), dict(chase=4096, sweep_grid=256),
                                    dict(fine_run=2), dict(fine_run=8, slot_scale=0.5), dict(fine_run=4, sweep_grid=512),
                                    dict(coarse_run=4, coarse_slots=64), dict(coarse_run=16, coarse_slots=4096),
                                    dict(fine_run=32, slot_scale=2.8),
                                    dict(slot_scale=0.25, coarse_slots=512), dict(slot_scale=4.0, sweep_grid=64),
                                    dict(graphs=-1), dict(graphs=-1, sweep_grid=128),
                                    dict(fast_list_min=1), dict(fast_list_min=1, graphs=-1, sweep_grid=64),
                                    dict(fast_list_min=-1)])
def test_analyse_tuning_keeps_vectors(gpu_ctx, oracle, tuning):
    """The search's schedule settings (chase along the visit order, slot counts, sweep
    grid) change only how the level's fixed point is reached, not the vectors."""
    print(f"analyse parity: tuning {tuning}", flush=True)
    frames = synth.clip(480, 272, 4, seed=2025, vmax=12.0, nobj=(3, 5), rot_deg=2.0)
    gpu_ctx.set_tuning(**tuning)
    try:
        _check(gpu_ctx, oracle, frames, "{pel:2}", CFG2, batch=True)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("aopt", [
    CFG2,
    "{block:{w:8,h:8,overlap:2}, main:{levels:3, search:{type:3,distance:8,coarse:{bad:{sad:200}}}}}",
    "{block:{w:4,h:4,overlap:0}, main:{search:{type:2,distance:6,coarse:{bad:{range:-3, sad:40}}}}}",
    "{block:{w:8,h:8,overlap:3}, main:{search:{type:2,distance:8,coarse:{distance:-6, satd:false}}}}",
])
def test_analyse_fast_later_sweeps(gpu_ctx, oracle, aopt):
    """Every later sweep, finest and coarse levels, through the tile-plan kernel and its
    deferred pass (fast_list_min 1), with and without bad-block passes and count
    rebuilds: the same payloads as the oracle."""
    print(f"analyse parity: fast later sweeps {aopt}", flush=True)
    frames = synth.clip(416, 240, 4, seed=909, vmax=12.0, nobj=(3, 5), rot_deg=2.0)
    gpu_ctx.set_tuning(fast_list_min=1, coarse_slots=64, coarse_run=4)
    try:
        _check(gpu_ctx, oracle, frames, "{pel:2}", aopt, batch=True)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("bad", [dict(fine_run=3), dict(coarse_run=2), dict(slot_scale=-1.0), dict(graphs=2),
                                 dict(fast_list_min=-2), dict(sweep_grid=-1)])
def test_analyse_tuning_rejects_out_of_range(gpu_ctx, bad):
    """svpf_analyse_set_tuning refuses settings outside their documented ranges with an
    error, and leaves the context usable."""
    with pytest.raises(svpf.SvpfError, match="svpf_analyse_set_tuning"):
        gpu_ctx.set_tuning(**bad)
    gpu_ctx.set_tuning()


def test_analyse_graph_replays_match_launches(gpu_ctx, oracle):
    """The later sweeps replayed from a captured hipGraph (cached per launch
    configuration, so later calls replay the graph captured by the first) give the
    payloads of the same sweeps launched one by one, call after call, with other work
    and synchronisations on the stream between the replays."""
    print("analyse parity: sweep graph replays vs launches", flush=True)
    frames = synth.clip(416, 240, 4, seed=4242, vmax=10.0, nobj=(2, 4), rot_deg=1.5)
    h, w = frames[0][0].shape
    sdata, devs, hosts = _supers(gpu_ctx, oracle, frames, "{pel:2}")
    p, vdata, plen = svpf.analyse_params(CFG2, sdata, w, h)
    pays = [gpu_ctx.alloc(plen) for _ in range(len(frames) - 1)]
    gpu_ctx.profile(True)
    try:
        gpu_ctx.set_tuning(graphs=-1)
        svpf.analyse_batch(gpu_ctx, p, devs[:-1], devs[1:], pays)
        gpu_ctx.sync()
        eager = [gpu_ctx.download(q, plen) for q in pays]
        for i in range(len(frames) - 1):
            ref = oracle.analyse(CFG2, sdata, w, h, hosts[i], hosts[i + 1])
            assert np.array_equal(eager[i], ref), f"pair {i}: launches differ from the oracle"
        gpu_ctx.set_tuning()
        for call in range(3):
            for q in pays:  # other work on the stream between the calls
                gpu_ctx.memset(q, 0xA5, plen)
            svpf.analyse_pair(gpu_ctx, p, devs[0], devs[1], pays[0])  # its own graphs
            gpu_ctx.sync()
            svpf.analyse_batch(gpu_ctx, p, devs[:-1], devs[1:], pays)
            gpu_ctx.sync()
            for i, q in enumerate(pays):
                assert np.array_equal(gpu_ctx.download(q, plen), eager[i]), f"call {call}, pair {i}: graph replay differs"
        st = gpu_ctx.analyse_stats(0)
        assert st["runs"] >= 4 and st["sweeps"] > st["runs"], st
    finally:
        gpu_ctx.set_tuning()
        gpu_ctx.profile(False)
