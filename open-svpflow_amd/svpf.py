"""Python binding of libsvpf (include/svpf.h) over ctypes.

Host-side mirror of the reference's filter interfaces for the motion path
(svpflow1.Super / svpflow1.Analyse / svpflow2.SmoothFps): the same option strings
and the same data layouts (super-frame planes, vector payloads, output frames),
executed by the HIP kernels in libsvpf.so.  There is no CPU fallback: loading
fails loudly when the library is missing, and every call raises SvpfError when
the library reports an error.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsvpf.so")


class SvpfError(RuntimeError):
    pass


class AnalyseTuning(C.Structure):
    _fields_ = [("chase", C.c_int32), ("slot_scale", C.c_double), ("coarse_slots", C.c_int32),
                ("sweep_grid", C.c_int32), ("fine_run", C.c_int32), ("coarse_run", C.c_int32),
                ("graphs", C.c_int32), ("fast_list_min", C.c_int32)]


class Frame(C.Structure):
    _fields_ = [("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("stride_y", C.c_int32), ("stride_u", C.c_int32), ("stride_v", C.c_int32)]


class SuperParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pel", C.c_int32), ("gpu", C.c_int32),
                ("scale_down", C.c_int32), ("scale_up", C.c_int32), ("full", C.c_int32), ("rc", C.c_int32),
                ("levels", C.c_int32)]


_lib = None
_BINDERS = []  # per-section argtype setup, run once when the library loads


def lib() -> C.CDLL:
    """Loads libsvpf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SvpfError(f"{LIB_PATH} is missing: build it with open-svpflow_amd/build_lib.py")
        L = C.CDLL(LIB_PATH)
        L.svpf_last_error.restype = C.c_char_p
        L.svpf_stream.restype = C.c_void_p
        L.svpf_super_height.restype = C.c_int32
        L.svpf_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.svpf_destroy.argtypes = [C.c_void_p]
        for name in ("svpf_synchronize", "svpf_timer_start"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.svpf_timer_stop.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.svpf_device_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
        L.svpf_device_free.argtypes = [C.c_void_p, C.c_void_p]
        L.svpf_copy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        L.svpf_copy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        L.svpf_memset_d.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]
        L.svpf_copy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        L.svpf_super_params_from_opt.argtypes = [C.c_char_p, C.c_size_t, C.c_int32, C.c_int32,
                                                 C.POINTER(SuperParams), C.POINTER(C.c_int64), C.c_char_p,
                                                 C.c_size_t]
        L.svpf_super_height.argtypes = [C.POINTER(SuperParams)]
        L.svpf_super_build.argtypes = [C.c_void_p, C.POINTER(SuperParams), C.POINTER(Frame), C.POINTER(Frame)]
        L.svpf_super_build_batch.argtypes = [C.c_void_p, C.POINTER(SuperParams), C.c_int32, C.POINTER(Frame),
                                             C.POINTER(Frame)]
        L.svpf_profile_enable.argtypes = [C.c_void_p, C.c_int]
        L.svpf_profile_read.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.svpf_analyse_stats.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
        L.svpf_analyse_set_max_sweeps.argtypes = [C.c_void_p, C.c_int32]
        L.svpf_analyse_set_tuning.argtypes = [C.c_void_p, C.POINTER(AnalyseTuning)]
        for bind in _BINDERS:
            bind(L)
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise SvpfError(f"libsvpf error {rc}: {lib().svpf_last_error().decode(errors='replace')}")


def exported_symbols() -> list[str]:
    """Names declared in include/svpf.h (used by the symbol-export test)."""
    import re
    path = os.path.join(HERE, "..", "include", "svpf.h")
    text = open(path).read()
    return sorted(set(re.findall(r"\b(svpf_[a-z0-9_]+)\s*\(", text)))


_LIVE_CONTEXTS: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _close_live_contexts() -> None:
    """Ends every context still open when the interpreter exits (render sessions first),
    while the interpreter and the HIP runtime are whole: interpreter teardown would
    otherwise finalise sessions and contexts in any order."""
    for ctx in list(_LIVE_CONTEXTS):
        ctx.close()


class Context:
    """A libsvpf context: one device, one HIP stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().svpf_create(device, C.byref(h)))
        self.h = h
        self._allocs: list[int] = []
        self._sessions = weakref.WeakSet()  # render sessions to end before the context
        _LIVE_CONTEXTS.add(self)

    def close(self) -> None:
        if self.h:
            for s in list(self._sessions):
                s.close()
            for p in self._allocs:
                lib().svpf_device_free(self.h, p)
            self._allocs.clear()
            lib().svpf_destroy(self.h)
            self.h = None
        _LIVE_CONTEXTS.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- memory
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(lib().svpf_device_alloc(self.h, nbytes, C.byref(p)))
        self._allocs.append(p.value)
        return p.value

    def free(self, ptr: int) -> None:
        if ptr in self._allocs:
            self._allocs.remove(ptr)
        check(lib().svpf_device_free(self.h, ptr))

    def upload(self, dptr: int, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        check(lib().svpf_copy_h2d(self.h, dptr, a.ctypes.data, a.nbytes))

    def download(self, dptr: int, nbytes: int, dtype=np.uint8) -> np.ndarray:
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(lib().svpf_copy_d2h(self.h, out.ctypes.data, dptr, out.nbytes))
        return out

    def copy(self, dst: int, src: int, nbytes: int) -> None:
        """Device-to-device copy on the context's stream."""
        check(lib().svpf_copy_d2d(self.h, dst, src, nbytes))

    def memset(self, dptr: int, value: int, nbytes: int) -> None:
        check(lib().svpf_memset_d(self.h, dptr, value, nbytes))

    def sync(self) -> None:
        check(lib().svpf_synchronize(self.h))

    def timer_start(self) -> None:
        check(lib().svpf_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_float()
        check(lib().svpf_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def profile(self, on: bool) -> None:
        check(lib().svpf_profile_enable(self.h, 1 if on else 0))

    def set_max_sweeps(self, n: int) -> None:
        """Parallel sweeps per Analyse level before the sequential fallback (0 = default)."""
        check(lib().svpf_analyse_set_max_sweeps(self.h, n))

    def set_tuning(self, chase: int = 0, slot_scale: float = 0.0, coarse_slots: int = 0, sweep_grid: int = 0,
                   fine_run: int = 0, coarse_run: int = 0, graphs: int = 0, fast_list_min: int = 0):
        """svpf_analyse_set_tuning: zero keeps a default (chase 0: no chase; graphs -1: later
        sweeps launched one by one instead of replayed from a captured hipGraph;
        fast_list_min -1: no later sweep runs the tile-plan kernel)."""
        t = AnalyseTuning(chase, slot_scale, coarse_slots, sweep_grid, fine_run, coarse_run, graphs, fast_list_min)
        check(lib().svpf_analyse_set_tuning(self.h, C.byref(t)))

    def analyse_stats(self, level: int):
        """{runs, sweeps launched, evaluations skipped (inputs unchanged), evaluations run} of one level."""
        out = (C.c_int64 * 4)()
        check(lib().svpf_analyse_stats(self.h, level, out))
        return dict(zip(("runs", "sweeps", "skipped", "evaluations"), list(out)))

    def profile_read(self, kernel: str):
        """(total ms, launches) of one timed kernel since profiling was enabled."""
        ms, n = C.c_double(), C.c_int32()
        check(lib().svpf_profile_read(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


@dataclass
class DeviceFrame:
    """A YUV420P8 frame in device memory: three planes, each `rows` x `stride`."""
    ctx: Context
    y: int
    u: int
    v: int
    strides: tuple
    rows: tuple

    @classmethod
    def alloc(cls, ctx: Context, strides, rows) -> "DeviceFrame":
        ptrs = [ctx.alloc(s * r) for s, r in zip(strides, rows)]
        return cls(ctx, ptrs[0], ptrs[1], ptrs[2], tuple(strides), tuple(rows))

    def desc(self) -> Frame:
        return Frame(self.y, self.u, self.v, *self.strides)

    def upload(self, planes) -> None:
        for ptr, stride, rows, p in zip((self.y, self.u, self.v), self.strides, self.rows, planes):
            buf = np.zeros((rows, stride), np.uint8)
            buf[:p.shape[0], :p.shape[1]] = p
            self.ctx.upload(ptr, buf)

    def download(self):
        return [self.ctx.download(ptr, s * r).reshape(r, s)
                for ptr, s, r in zip((self.y, self.u, self.v), self.strides, self.rows)]


# ---------------------------------------------------------------- Super

def super_params(opt: str | bytes | None, width: int, height: int):
    """SuperOpts::from_opt (super_opts.rs:17-49) -> (SuperParams, sdata)."""
    p = SuperParams()
    sdata = C.c_int64()
    err = C.create_string_buffer(256)
    b = opt.encode() if isinstance(opt, str) else (opt or b"")
    rc = lib().svpf_super_params_from_opt(b, len(b), width, height, C.byref(p), C.byref(sdata), err, 256)
    if rc == -3:
        raise SvpfError(err.value.decode())
    check(rc)
    return p, sdata.value


def super_height(p: SuperParams) -> int:
    return lib().svpf_super_height(C.byref(p))


def super_build(ctx: Context, p: SuperParams, src: DeviceFrame, dst: DeviceFrame) -> None:
    """svpflow1.Super for one frame: d_src -> d_dst (all planes written)."""
    s, d = src.desc(), dst.desc()
    check(lib().svpf_super_build(ctx.h, C.byref(p), C.byref(s), C.byref(d)))


def super_build_batch(ctx: Context, p: SuperParams, srcs, dsts) -> None:
    n = len(srcs)
    S = (Frame * n)(*[f.desc() for f in srcs])
    D = (Frame * n)(*[f.desc() for f in dsts])
    check(lib().svpf_super_build_batch(ctx.h, C.byref(p), n, S, D))


# ---------------------------------------------------------------- Analyse

class AnalyseParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "vectors", "block_w", "block_h", "overlap_mode", "overlap_x", "overlap_y", "delta", "pel", "selector",
        "levels", "super_levels", "width", "height", "lambda_", "lsad")] + [("plevel", C.c_double)] + [
        (n, C.c_int32) for n in (
            "pnew", "pglobal", "pzero", "pnbour", "prev", "search_type", "coarse_search_type", "distance",
            "coarse_distance", "coarse_trymany", "coarse_bad_sad", "coarse_bad_range", "fine_satd", "coarse_satd",
            "has_refine", "refine_thsad", "refine_search_type", "refine_distance", "refine_satd", "sort")]


def _bind_analyse(L):
    L.svpf_analyse_params_from_opt.argtypes = [C.c_char_p, C.c_size_t, C.c_int64, C.c_int32, C.c_int32,
                                               C.POINTER(AnalyseParams), C.POINTER(C.c_int32),
                                               C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]
    L.svpf_analyse_pair.argtypes = [C.c_void_p, C.POINTER(AnalyseParams), C.POINTER(Frame), C.POINTER(Frame),
                                    C.c_void_p]
    L.svpf_analyse_batch.argtypes = [C.c_void_p, C.POINTER(AnalyseParams), C.c_int32, C.POINTER(Frame),
                                     C.POINTER(Frame), C.POINTER(C.c_void_p)]
    L.svpf_analyse_multi.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.POINTER(AnalyseParams)),
                                     C.POINTER(Frame), C.POINTER(Frame), C.POINTER(C.c_void_p)]


_BINDERS.append(_bind_analyse)


def analyse_params(opt, sdata: int, width: int = 0, height: int = 0):
    """AnalyseOpts::from_opt -> (AnalyseParams, vdata[16], payload_len)."""
    L = lib()
    p = AnalyseParams()
    vd = (C.c_int32 * 16)()
    plen = C.c_int32()
    err = C.create_string_buffer(256)
    b = opt.encode() if isinstance(opt, str) else (opt or b"")
    rc = L.svpf_analyse_params_from_opt(b, len(b), sdata, width, height, C.byref(p), vd, C.byref(plen), err, 256)
    if rc == -3:
        raise SvpfError(err.value.decode())
    check(rc)
    return p, list(vd), plen.value


def analyse_pair(ctx: Context, p: AnalyseParams, cur: DeviceFrame, ref: DeviceFrame, d_payload: int) -> None:
    """svpflow1.Analyse for one pair of super frames -> vector payload in device memory."""
    c, r = cur.desc(), ref.desc()
    check(lib().svpf_analyse_pair(ctx.h, C.byref(p), C.byref(c), C.byref(r), C.c_void_p(d_payload)))


def analyse_batch(ctx: Context, p: AnalyseParams, curs, refs, d_payloads) -> None:
    n = len(curs)
    Cu = (Frame * n)(*[f.desc() for f in curs])
    R = (Frame * n)(*[f.desc() for f in refs])
    P = (C.c_void_p * n)(*d_payloads)
    check(lib().svpf_analyse_batch(ctx.h, C.byref(p), n, Cu, R, P))


def analyse_multi(ctx: Context, params, curs, refs, d_payloads) -> None:
    n = len(curs)
    PP = (C.POINTER(AnalyseParams) * n)(*[C.pointer(p) for p in params])
    Cu = (Frame * n)(*[f.desc() for f in curs])
    R = (Frame * n)(*[f.desc() for f in refs])
    P = (C.c_void_p * n)(*d_payloads)
    check(lib().svpf_analyse_multi(ctx.h, n, PP, Cu, R, P))


# ---------------------------------------------------------------- SmoothFps

def _bind_smooth(L):
    L.svpf_smooth_create.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.c_int32, C.c_int32, C.c_int64,
                                     C.c_int64, C.c_int32, C.c_int64, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]
    L.svpf_smooth_destroy.argtypes = [C.c_void_p]
    L.svpf_smooth_destroy.restype = None
    L.svpf_smooth_set_sources.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Frame), C.POINTER(Frame),
                                          C.POINTER(C.c_void_p)]
    L.svpf_smooth_render.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Frame), C.POINTER(C.c_int32)]
    L.svpf_smooth_output_frames.argtypes = [C.c_void_p]
    L.svpf_smooth_output_frames.restype = C.c_int32
    L.svpf_smooth_output_size.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.svpf_smooth_release.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.svpf_smooth_set_float_render.argtypes = [C.c_void_p, C.c_int32]


_BINDERS.append(_bind_smooth)


class Smooth:
    """svpflow2.SmoothFps render session for one clip (YUV420P8).

    `opt` is the SmoothFps option string, `sdata` the Super clip's handle and
    `vdata` the 16-word vector header returned by analyse_params.
    """

    def __init__(self, ctx: Context, opt, width: int, height: int, fps_num: int, fps_den: int, nframes: int,
                 sdata: int, vdata):
        L = lib()
        self.ctx = ctx
        h = C.c_void_p()
        nout = C.c_int32()
        err = C.create_string_buffer(256)
        vd = (C.c_int32 * 16)(*vdata)
        b = opt.encode() if isinstance(opt, str) else (opt or b"")
        rc = L.svpf_smooth_create(ctx.h, b, len(b), width, height, fps_num, fps_den, nframes, sdata, vd,
                                  C.byref(h), C.byref(nout), err, 256)
        if rc == -3:
            raise SvpfError(err.value.decode())
        check(rc)
        self.h = h
        self.output_frames = nout.value
        ow, oh = C.c_int32(), C.c_int32()
        check(L.svpf_smooth_output_size(h, C.byref(ow), C.byref(oh)))
        self.output_size = (ow.value, oh.value)  # (width, height), light.* padding included
        ctx._sessions.add(self)

    def close(self) -> None:
        # a session is ended before its context (Context.close does so); a session met
        # after its context closed (interpreter teardown) is only dropped, since its
        # destroy call would use the context's stream
        if self.h and self.ctx.h:
            lib().svpf_smooth_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sources(self, first: int, srcs, supers, d_vectors) -> None:
        n = len(srcs)
        S = (Frame * n)(*[f.desc() for f in srcs])
        U = (Frame * n)(*[f.desc() for f in supers]) if supers is not None else None
        V = (C.c_void_p * n)(*d_vectors)
        check(lib().svpf_smooth_set_sources(self.h, first, n, S, U, V))

    def set_float_render(self, on: bool) -> None:
        """Rendered frames through the reference's float GPU renderer (True) or the
        bit-exact integer renderer (False, the default)."""
        check(lib().svpf_smooth_set_float_render(self.h, 1 if on else 0))

    def release(self, first: int, count: int) -> None:
        """Frees the device state derived from source frames first..first+count-1."""
        check(lib().svpf_smooth_release(self.h, first, count))

    def render(self, k0: int, outs, decisions: bool = False):
        n = len(outs)
        O = (Frame * n)(*[f.desc() for f in outs])
        D = (C.c_int32 * (2 * n))() if decisions else None
        check(lib().svpf_smooth_render(self.h, k0, n, O, D))
        if decisions:
            return [(D[2 * i], D[2 * i + 1]) for i in range(n)]
        return None
