"""Autotuner of the Analyse level-search schedule (svpf_analyse_set_tuning).

The level search reaches the same fixed point under every schedule setting; the
settings only decide how fast it gets there:

* ``slot_scale``   -- factor on the finest level's sweep-0 slot count.  More slots put
  more blocks at a segment start (evaluated before the segment above reaches them),
  fewer slots put more items of one slot in flight at once (evaluated before their
  predecessor is stored); both cost re-evaluations in the later sweeps.
* ``coarse_slots`` -- sweep-0 slots of the coarser levels, and ``coarse_run`` the
  positions of their visit order one item searches in order.
* ``sweep_grid``   -- workgroups of a later sweep (a persistent grid over the queue).
* ``fine_run``     -- consecutive finest-level positions one sweep-0 item searches in
  order (a block's reference rows stay cached for the next).

Their best values move with the batch (pairs, frame size), the pyramid and the kernels
themselves, so they are measured on the caller's own workload: coordinate descent, one
knob at a time over its candidate values, keeping the fastest, for a few rounds.  The
time of a trial is the search's own device time (HIP events around each level's
sweeps, svpf_profile_read "search"), the minimum over ``repeats`` runs.  After every
trial the payloads are hashed and compared with the first run's: a schedule that
changed a vector raises (it would be a bug, never a trade-off).
"""
from __future__ import annotations

import hashlib
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import svpf

# candidate values of each knob (the library's defaults are DEFAULT)
SPACE: Dict[str, Tuple] = {
    "slot_scale": (0.5, 0.7, 1.0, 1.4, 2.0, 2.8),
    "sweep_grid": (1024, 2048, 4096, 8192),
    "coarse_slots": (256, 512, 1024, 4096),
    "fine_run": (1, 2, 4, 8, 16, 32),
    "coarse_run": (4, 8, 16, 32),
}
DEFAULT: Dict[str, float] = {"slot_scale": 1.0, "sweep_grid": 2048, "coarse_slots": 256, "fine_run": 32, "coarse_run": 16}
ORDER = ("slot_scale", "fine_run", "sweep_grid", "coarse_slots", "coarse_run")


@dataclass
class TuneResult:
    best: Dict[str, float]
    best_ms: float
    default_ms: float
    trials: List[Tuple[Dict[str, float], float]] = field(default_factory=list)
    seconds: float = 0.0

    def as_dict(self) -> dict:
        return {"best": self.best, "best_search_ms": round(self.best_ms, 2),
                "default_search_ms": round(self.default_ms, 2), "trials": len(self.trials),
                "seconds": round(self.seconds, 1)}


def apply(ctx: svpf.Context, setting: Dict[str, float]) -> None:
    """Sets a schedule on the context (knobs absent from `setting` take their defaults)."""
    s = dict(DEFAULT)
    s.update(setting)
    ctx.set_tuning(slot_scale=float(s["slot_scale"]), coarse_slots=int(s["coarse_slots"]),
                   sweep_grid=int(s["sweep_grid"]), fine_run=int(s["fine_run"]), coarse_run=int(s["coarse_run"]),
                   graphs=int(s.get("graphs", 0)), fast_list_min=int(s.get("fast_list_min", 0)))


def descend(measure: Callable[[Dict[str, float]], float], space: Dict[str, Sequence] = SPACE,
            start: Optional[Dict[str, float]] = None, order: Sequence[str] = ORDER, rounds: int = 2,
            deadline: Optional[float] = None) -> Tuple[Dict[str, float], float, List[Tuple[Dict[str, float], float]]]:
    """Coordinate descent of measure(setting) (lower is better) over `space`.

    Each round tries every value of every knob in `order` with the other knobs at the
    best setting so far; a round that improves nothing ends the search, as does passing
    `deadline` (time.monotonic()).  Settings already measured are not measured again.
    Returns (best setting, its measure, every trial in order)."""
    best = dict(start if start is not None else DEFAULT)
    seen: Dict[Tuple, float] = {}
    trials: List[Tuple[Dict[str, float], float]] = []

    def cost(setting: Dict[str, float]) -> float:
        key = tuple(setting[k] for k in sorted(setting))
        if key not in seen:
            seen[key] = measure(dict(setting))
            trials.append((dict(setting), seen[key]))
        return seen[key]

    best_cost = cost(best)
    for _ in range(max(rounds, 1)):
        improved = False
        for knob in order:
            for v in space[knob]:
                if deadline is not None and time.monotonic() > deadline:
                    return best, best_cost, trials
                cand = dict(best)
                cand[knob] = v
                c = cost(cand)
                if c < best_cost:
                    best, best_cost, improved = cand, c, True
        if not improved:
            break
    return best, best_cost, trials


def _payload_hash(ctx: svpf.Context, payloads: Sequence[Tuple[int, int]]) -> str:
    h = hashlib.sha1()
    for ptr, n in payloads:
        h.update(ctx.download(ptr, n).tobytes())
    return h.hexdigest()


def tune(ctx: svpf.Context, run: Callable[[], None], payloads: Sequence[Tuple[int, int]], rounds: int = 2,
         repeats: int = 1, budget_s: Optional[float] = None, log: Optional[Callable[[str], None]] = None) -> TuneResult:
    """Tunes the schedule for `run` (one search call over the workload, e.g. a
    svpf.analyse_batch of a clip) and leaves the best setting applied on `ctx`.

    payloads: (device pointer, bytes) of the vector payloads `run` writes; their hash
    must not change between settings.  budget_s bounds the tuning's wall time (the
    trials measured by then decide)."""
    t_begin = time.monotonic()
    deadline = t_begin + budget_s if budget_s else None
    apply(ctx, DEFAULT)
    run()  # the workspace and descriptors exist before the first timed trial
    ctx.sync()
    reference = _payload_hash(ctx, payloads)

    def measure(setting: Dict[str, float]) -> float:
        apply(ctx, setting)
        best = float("inf")
        for _ in range(max(repeats, 1)):
            ctx.profile(True)
            run()
            ctx.sync()
            ms, _ = ctx.profile_read("search")
            ctx.profile(False)
            best = min(best, ms)
        got = _payload_hash(ctx, payloads)
        if got != reference:
            raise RuntimeError(f"autotune: schedule {setting} changed the vector payloads")
        if log:
            log(f"autotune {setting}: search {best:.2f} ms")
        return best

    best, best_ms, trials = descend(measure, rounds=rounds, deadline=deadline)
    apply(ctx, best)
    return TuneResult(best=best, best_ms=best_ms, default_ms=trials[0][1], trials=trials,
                      seconds=time.monotonic() - t_begin)
