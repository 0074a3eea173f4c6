"""Serving loop over libsvpf: SmoothFps streams admitted by an HBM budget and
batched continuously.

A stream (one SmoothFps request) delivers its source frames at its own frame rate
from its arrival time.  It waits in a FIFO queue until its device footprint fits in
the budget (and, when one is given, its pixel rate fits in a throughput budget), and
is dropped if it waits longer than `max_wait` seconds.  Each step
handles every admitted stream together:
  1. Super for the source frames that have arrived, one batched call per resolution;
  2. Analyse for every frame pair whose two super frames exist, every stream in one
     multi-clip call (svpf_analyse_multi);
  3. per stream, SmoothFps for every output frame whose inputs are ready.
Super frames and vector payloads stay on the device in per-stream rings.  A ring slot
is reused only when no pending output can need its frame, and the session's derived
state of older frames is released.  For every output frame the loop records the time
its last input frame arrived and the time its render finished (after the step's
synchronization); their difference is the frame latency.  It includes the time the
stream waited for admission; the latency counted from the later of that arrival and
the stream's admission is recorded beside it.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable

import numpy as np

import svpf

SUPER_OPT = "{pel:2}"
ANALYSE_OPT = "{block:{w:8,h:8,overlap:3},main:{levels:4,search:{type:3,distance:16,satd:true}}}"
RING = 12       # super frames / payloads kept per stream
LOOKAHEAD = 3   # source frames past an output's source frame n that its render may read (vectors to n + 2)


def smooth_opt(rate) -> str:
    return "{rate:{num:%d,den:%d},algo:23,mask:{area:100},scene:{mode:0}}" % rate


@dataclasses.dataclass
class StreamSpec:
    sid: int
    width: int
    height: int
    frames: int
    rate: tuple            # output / source frame-rate factor (num, den)
    fps: tuple             # source frame rate (num, den)
    arrival: float         # seconds from the start of the run
    source: Callable       # i -> svpf.DeviceFrame of source frame i (device resident, read only)


class _Stream:
    def __init__(self, srv: "Server", spec: StreamSpec):
        ctx = srv.ctx
        self.spec = spec
        w, h = spec.width, spec.height
        self.n = spec.frames
        self.fps = spec.fps[0] / spec.fps[1]
        self.sp, self.sdata = srv.super_params(w, h)
        self.ap, self.vdata, self.plen = srv.analyse_params(w, h, self.sdata)
        sh = svpf.super_height(self.sp)
        strides = (w, w // 2, w // 2)
        self.sup = [svpf.DeviceFrame.alloc(ctx, strides, (sh, sh // 2, sh // 2)) for _ in range(RING)]
        self.pay = [ctx.alloc(self.plen) for _ in range(RING)]
        self.out = [svpf.DeviceFrame.alloc(ctx, strides, (h, h // 2, h // 2)) for _ in range(srv.out_per_step)]
        self.sm = svpf.Smooth(ctx, smooth_opt(spec.rate), w, h, spec.fps[0], spec.fps[1], self.n, self.sdata,
                              self.vdata)
        self.nout = self.sm.output_frames
        self.ingested = 0   # super frames built for source frames 0 .. ingested-1
        self.paired = 0     # vector payloads computed for source frames 0 .. paired-1
        self.bound = 0      # source frames bound to the SmoothFps session
        self.released = 0
        self.next_out = 0
        self.bytes = srv.footprint(spec)
        self.admitted_at = srv.clock()

    def src_of(self, k: int) -> int:
        """Source frame of output k at scene mode 0 (floor of the output time)."""
        return min(k * self.spec.rate[1] // self.spec.rate[0], self.n - 1)

    def keep_from(self) -> int:
        """Lowest source frame a pending output can still read."""
        if self.next_out >= self.nout:
            return self.n
        return max(self.src_of(self.next_out) - 2, 0)

    def arrived(self, t: float) -> int:
        if t < self.spec.arrival:
            return 0
        return min(self.n, int((t - self.spec.arrival) * self.fps) + 1)

    def arrival_of(self, i: int) -> float:
        return self.spec.arrival + i / self.fps

    def close(self, ctx) -> None:
        self.sm.close()
        for f in self.sup + self.out:
            for p in (f.y, f.u, f.v):
                ctx.free(p)
        for p in self.pay:
            ctx.free(p)


class Server:
    """Continuous batching of SmoothFps streams on one context."""

    def __init__(self, ctx: svpf.Context, budget_bytes: int, max_wait: float = 2.0, out_per_step: int = 8,
                 on_output: Callable | None = None, pixel_rate_budget: float | None = None):
        self.ctx = ctx
        self.budget = budget_bytes
        # optional throughput budget: source pixels per second the admitted streams may
        # deliver in total (their frame rates times frame sizes)
        self.pixel_budget = pixel_rate_budget
        self.pixel_rate = 0.0
        self.max_wait = max_wait
        self.out_per_step = out_per_step
        self.on_output = on_output
        self.queue: list[StreamSpec] = []
        self.active: list[_Stream] = []
        self.used = 0
        self.latency: list[float] = []
        self.latency_admitted: list[float] = []
        self.t0 = time.perf_counter()
        self.outputs = 0
        self.admitted = 0
        self.dropped = 0
        self.finished = 0
        self.steps = 0
        self._sp, self._ap = {}, {}

    # -- per-resolution options (SuperOpts / AnalyseOpts are the same for every stream)
    def super_params(self, w, h):
        if (w, h) not in self._sp:
            self._sp[(w, h)] = svpf.super_params(SUPER_OPT, w, h)
        return self._sp[(w, h)]

    def analyse_params(self, w, h, sdata):
        if (w, h) not in self._ap:
            self._ap[(w, h)] = svpf.analyse_params(ANALYSE_OPT, sdata, w, h)
        return self._ap[(w, h)]

    def footprint(self, spec: StreamSpec) -> int:
        """Device bytes a stream holds: its rings, output frames and per-frame derived state."""
        w, h = spec.width, spec.height
        sp, sdata = self.super_params(w, h)
        _, vdata, plen = self.analyse_params(w, h, sdata)
        sh = svpf.super_height(sp)
        grid = (vdata[12] + 1) * (vdata[13] + 1)
        prep = 8 * grid + 8 * grid + 2 * grid + 4096
        return RING * (sh * w * 3 // 2 + plen + prep) + self.out_per_step * (w * h * 3 // 2 + 12 * grid)

    def submit(self, spec: StreamSpec) -> None:
        self.queue.append(spec)

    def _admit(self, now: float) -> None:
        while self.queue and self.queue[0].arrival <= now:
            spec = self.queue[0]
            need = self.footprint(spec)
            rate = spec.width * spec.height * spec.fps[0] / spec.fps[1]
            fits = self.used + need <= self.budget
            if self.pixel_budget is not None:
                fits = fits and (self.pixel_rate + rate <= self.pixel_budget or not self.active)
            if fits:
                self.queue.pop(0)
                self.active.append(_Stream(self, spec))
                self.used += need
                self.pixel_rate += rate
                self.admitted += 1
            elif now - spec.arrival > self.max_wait:
                self.queue.pop(0)
                self.dropped += 1
            else:
                break

    def step(self, now: float) -> int:
        """One batched step at time `now` (seconds since the start); returns the outputs rendered."""
        self._admit(now)
        if not self.active:
            return 0
        self.steps += 1
        ctx = self.ctx
        # 1. Super for arrived frames (bounded by the ring)
        groups: dict = {}
        for s in self.active:
            lim = min(s.arrived(now), s.keep_from() + RING)
            for j in range(s.ingested, lim):
                g = groups.setdefault((s.spec.width, s.spec.height), (s.sp, [], []))
                g[1].append(s.spec.source(j))
                g[2].append(s.sup[j % RING])
            s.ingested = max(s.ingested, lim)
        for p, srcs, sups in groups.values():
            svpf.super_build_batch(ctx, p, srcs, sups)
        # 2. Analyse every ready pair of every stream in one multi-clip call
        params, curs, refs, pays = [], [], [], []
        for s in self.active:
            while s.paired < s.n and s.paired < s.keep_from() + RING:
                j = s.paired
                nxt = min(j + 1, s.n - 1)
                if nxt >= s.ingested:
                    break
                params.append(s.ap)
                curs.append(s.sup[j % RING])
                refs.append(s.sup[nxt % RING])
                pays.append(s.pay[j % RING])
                s.paired += 1
        if params:
            svpf.analyse_multi(ctx, params, curs, refs, pays)
        # 3. SmoothFps: outputs whose inputs are all computed
        rendered = []
        for s in self.active:
            if s.bound < s.paired:
                idx = range(s.bound, s.paired)
                s.sm.set_sources(s.bound, [s.spec.source(i) for i in idx], [s.sup[i % RING] for i in idx],
                                 [s.pay[i % RING] for i in idx])
                s.bound = s.paired
            k0 = s.next_out
            k = k0
            while k < s.nout and k - k0 < self.out_per_step:
                last = min(s.src_of(k) + LOOKAHEAD, s.n - 1)
                if last >= s.paired:
                    break
                k += 1
            if k > k0:
                s.sm.render(k0, s.out[:k - k0])
                rendered.append((s, k0, k))
                s.next_out = k
        ctx.sync()
        done = self.clock()
        count = 0
        for s, k0, k1 in rendered:
            for k in range(k0, k1):
                ready = s.arrival_of(min(s.src_of(k) + LOOKAHEAD, s.n - 1))
                self.latency.append(max(done - ready, 0.0))
                self.latency_admitted.append(max(done - max(ready, s.admitted_at), 0.0))
                if self.on_output:
                    self.on_output(s.spec.sid, k, s.out[k - k0])
            count += k1 - k0
        self.outputs += count
        # release what no pending output reads; retire finished streams
        for s in list(self.active):
            kf = min(s.keep_from(), s.bound)
            if kf > s.released:
                s.sm.release(s.released, kf - s.released)
                s.released = kf
            if s.next_out >= s.nout:
                s.close(ctx)
                self.active.remove(s)
                self.used -= s.bytes
                self.pixel_rate -= s.spec.width * s.spec.height * s.spec.fps[0] / s.spec.fps[1]
                self.finished += 1
        return count

    def clock(self) -> float:
        return time.perf_counter() - self.t0

    def run(self, specs, duration: float | None = None, drain: bool = True) -> dict:
        """Serves `specs` (by arrival time) until every stream is done (or `duration`
        seconds pass); returns the statistics."""
        for sp in sorted(specs, key=lambda x: x.arrival):
            self.submit(sp)
        self.t0 = time.perf_counter()
        last_progress = 0.0
        while self.queue or self.active:
            now = self.clock()
            if duration is not None and now >= duration:
                break
            if not self.active and self.queue and self.queue[0].arrival > now:
                time.sleep(min(self.queue[0].arrival - now, 0.01))
                continue
            if self.step(now):
                last_progress = now
            elif self.active:
                # nothing ready: frames not arrived yet (a stream whose frames have all
                # arrived always has work until it finishes)
                if all(s.arrived(now) >= s.n for s in self.active) and now - last_progress > 5.0:
                    raise RuntimeError("serving loop made no progress with every frame arrived")
                time.sleep(0.002)
        wall = self.clock()
        lat = np.array(self.latency) * 1000.0 if self.latency else np.zeros(1)
        lad = np.array(self.latency_admitted) * 1000.0 if self.latency_admitted else np.zeros(1)
        return {"wall_s": round(wall, 3), "outputs": self.outputs, "output_fps": round(self.outputs / wall, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "p50_ms_from_admission": round(float(np.percentile(lad, 50)), 2),
                "p99_ms_from_admission": round(float(np.percentile(lad, 99)), 2),
                "admitted": self.admitted, "dropped": self.dropped, "finished": self.finished,
                "still_queued": len(self.queue), "steps": self.steps}

    def close(self) -> None:
        for s in self.active:
            s.close(self.ctx)
        self.active = []
