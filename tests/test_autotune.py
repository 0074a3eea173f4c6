"""The Analyse schedule autotuner: its coordinate descent on the host (a synthetic cost
surface), and on the GPU over a small clip (every trial keeps the payloads; the best
setting is left applied)."""
import time

import pytest

import autotune
import svpf
import synth


def _surface(best):
    """A separable cost with its minimum at `best` (values as in autotune.SPACE)."""
    calls = []

    def measure(s):
        calls.append(dict(s))
        c = 100.0
        for k, v in s.items():
            vals = autotune.SPACE[k]
            c += 3.0 * abs(vals.index(v) - vals.index(best[k]))
        return c
    return measure, calls


def test_descend_finds_separable_minimum():
    best = {"slot_scale": 1.4, "sweep_grid": 2048, "coarse_slots": 1024, "fine_run": 8, "coarse_run": 4}
    measure, calls = _surface(best)
    got, cost, trials = autotune.descend(measure)
    assert got == best and cost == 100.0
    # no setting is measured twice
    keys = [tuple(sorted(c.items())) for c in calls]
    assert len(keys) == len(set(keys)) == len(trials)
    # the first trial is the library default
    assert trials[0][0] == autotune.DEFAULT


def test_descend_stops_when_nothing_improves():
    measure, calls = _surface(dict(autotune.DEFAULT))
    got, cost, trials = autotune.descend(measure, rounds=5)
    assert got == autotune.DEFAULT
    # one round over every value of every knob (the default counted once)
    assert len(trials) == 1 + sum(len(v) - 1 for v in autotune.SPACE.values())


def test_descend_respects_deadline():
    measure, calls = _surface({"slot_scale": 2.8, "sweep_grid": 1024, "coarse_slots": 1024, "fine_run": 2,
                                "coarse_run": 32})
    got, cost, trials = autotune.descend(measure, deadline=time.monotonic() - 1.0)
    assert len(trials) == 1 and got == autotune.DEFAULT


@pytest.mark.gpu
def test_tune_on_gpu_keeps_payloads(gpu_ctx):
    ctx = gpu_ctx
    frames = synth.clip(480, 272, 6, seed=77, vmax=12.0, nobj=(3, 5), rot_deg=2.0)
    h, w = frames[0][0].shape
    sp, sdata = svpf.super_params("{pel:2}", w, h)
    sh = svpf.super_height(sp)
    src, sup = [], []
    for f in frames:
        s = svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (h, h // 2, h // 2))
        s.upload(f)
        src.append(s)
        sup.append(svpf.DeviceFrame.alloc(ctx, (w, w // 2, w // 2), (sh, sh // 2, sh // 2)))
    svpf.super_build_batch(ctx, sp, src, sup)
    ap, vdata, plen = svpf.analyse_params("{block:{w:8,h:8,overlap:3},main:{search:{type:3,distance:16,satd:true}}}",
                                          sdata, w, h)
    pay = [ctx.alloc(plen) for _ in frames]
    refs = [sup[min(i + 1, len(frames) - 1)] for i in range(len(frames))]
    print("autotune: 480x272, 6 frames", flush=True)
    try:
        res = autotune.tune(ctx, lambda: svpf.analyse_batch(ctx, ap, sup, refs, pay), [(p, plen) for p in pay],
                            rounds=1, budget_s=60)
        print("autotune result:", res.as_dict(), flush=True)
        assert all(res.best[k] in autotune.SPACE[k] for k in autotune.SPACE)
        assert res.best_ms <= res.default_ms
        assert len(res.trials) >= 2
    finally:
        ctx.set_tuning()
