"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the checker: tests compare libsvpf's HIP results with it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

if not os.path.exists(PATH):
    raise RuntimeError(f"{PATH} missing: run `make -C oracle`")
L = C.CDLL(PATH)
u8p = C.POINTER(C.c_uint8)
_i = C.c_int
L.oracle_opt_lookup.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, _i, C.POINTER(_i), C.POINTER(C.c_int64),
                                C.POINTER(C.c_double)]
L.oracle_super_opts.argtypes = [C.c_char_p, C.c_size_t, _i, _i, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]
L.oracle_super_build.argtypes = [C.c_char_p, C.c_size_t, _i, _i] + [u8p, _i] * 6


def _p(a: np.ndarray):
    return a.ctypes.data_as(u8p)


def opt_lookup(opt: str, path: str, kind: int):
    """kind 0 int / 1 float / 2 bool.  Returns (parsed, value or None)."""
    b = opt.encode()
    parsed = C.c_int()
    iv = C.c_int64()
    fv = C.c_double()
    hit = L.oracle_opt_lookup(b, len(b), path.encode(), kind, C.byref(parsed), C.byref(iv), C.byref(fv))
    if not hit:
        return bool(parsed.value), None
    return True, (fv.value if kind == 1 else (bool(iv.value) if kind == 2 else iv.value))


def super_opts(opt: str | None, w: int, h: int):
    b = (opt or "").encode()
    out = (C.c_int32 * 9)()
    sdata = C.c_int64()
    sh = C.c_int32()
    err = C.create_string_buffer(256)
    rc = L.oracle_super_opts(b, len(b), w, h, out, C.byref(sdata), C.byref(sh), err, 256)
    if rc:
        return None, err.value.decode()
    keys = ("width", "height", "pel", "gpu", "scale_down", "scale_up", "full", "rc", "levels")
    d = dict(zip(keys, list(out)))
    d["sdata"] = sdata.value
    d["super_h"] = sh.value
    return d, None


def super_build(opt: str | None, planes, strides=None, dst_strides=None):
    """get_frame_super on host planes -> list of 3 super planes (rows x stride)."""
    Y, U, V = planes
    h, w = Y.shape
    o, err = super_opts(opt, w, h)
    if o is None:
        raise ValueError(err)
    sh = o["super_h"]
    strides = strides or (w, w // 2, w // 2)
    dst_strides = dst_strides or strides
    src = []
    for p, s in zip(planes, strides):
        buf = np.zeros((p.shape[0], s), np.uint8)
        buf[:, :p.shape[1]] = p
        src.append(buf)
    dst = [np.zeros((sh if i == 0 else sh // 2, dst_strides[i]), np.uint8) for i in range(3)]
    b = (opt or "").encode()
    rc = L.oracle_super_build(b, len(b), w, h, _p(src[0]), strides[0], _p(src[1]), strides[1], _p(src[2]),
                              strides[2], _p(dst[0]), dst_strides[0], _p(dst[1]), dst_strides[1], _p(dst[2]),
                              dst_strides[2])
    assert rc == 0
    return dst

_sz = C.c_size_t
L.oracle_analyse_opts.argtypes = [C.c_char_p, _sz, C.c_int64, _i, _i, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.c_char_p, _sz]
L.oracle_analyse.argtypes = [C.c_char_p, _sz, C.c_int64, _i, _i] + [u8p, _i] * 6 + [
    u8p, _sz, C.POINTER(_sz), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _i]
L.oracle_candidate_cost.argtypes = [_i, _i, _i, _i] + [u8p, _i] * 6 + [_i] * 9 + [C.POINTER(C.c_uint32)] * 2
L.oracle_hadamard4.argtypes = [C.POINTER(C.c_int32)]
L.oracle_hadamard4.restype = C.c_uint32
L.oracle_level_order.argtypes = [_i, _i, _i, C.POINTER(C.c_int32), _i]
L.oracle_lambdas.argtypes = [_i, C.c_double, _i, _i, _i, _i, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
L.oracle_global.argtypes = [C.POINTER(C.c_int16), _i, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
L.oracle_rescale.argtypes = [C.POINTER(C.c_int16), C.POINTER(C.c_uint32), _i, _i, _i, _i]

ANALYSE_KEYS = ("vectors", "block_w", "block_h", "overlap_x", "overlap_y", "pel", "levels", "super_levels", "width",
                "height", "lambda", "lsad", "search_type", "coarse_search_type", "distance", "payload_len")


def analyse_opts(opt: str | None, sdata: int, width: int = 0, height: int = 0):
    b = (opt or "").encode()
    out = (C.c_int32 * 16)()
    vd = (C.c_int32 * 16)()
    err = C.create_string_buffer(256)
    rc = L.oracle_analyse_opts(b, len(b), sdata, width, height, out, vd, err, 256)
    if rc:
        return None, err.value.decode()
    d = dict(zip(ANALYSE_KEYS, list(out)))
    d["vdata"] = list(vd)
    return d, None


def analyse(opt: str | None, sdata: int, width: int, height: int, cur, ref, threads: bool = False,
            counters: bool = False):
    """get_frame_analyse on two super frames (lists of 3 planes rows x stride) -> payload bytes."""
    b = (opt or "").encode()
    o, err = analyse_opts(opt, sdata, width, height)
    if o is None:
        raise ValueError(err)
    out = np.zeros(o["payload_len"] + 64, np.uint8)
    n = _sz()
    ev = (C.c_uint64 * 16)()
    px = (C.c_uint64 * 16)()
    args = []
    for planes in (cur, ref):
        for p in planes:
            args += [_p(np.ascontiguousarray(p)), p.shape[1]]
    keep = [np.ascontiguousarray(p) for p in list(cur) + list(ref)]
    args = []
    for p in keep:
        args += [_p(p), p.shape[1]]
    rc = L.oracle_analyse(b, len(b), sdata, width, height, *args, _p(out), out.size, C.byref(n), ev, px,
                          1 if threads else 0)
    assert rc == 0, rc
    payload = out[:n.value].copy()
    if counters:
        return payload, list(ev), list(px)
    return payload


def candidate_cost(width, height, pel, super_levels, cur, ref, level, px, py, bw, bh, mvx, mvy, satd, chroma):
    keep = [np.ascontiguousarray(p) for p in list(cur) + list(ref)]
    args = []
    for p in keep:
        args += [_p(p), p.shape[1]]
    l = C.c_uint32()
    c = C.c_uint32()
    L.oracle_candidate_cost(width, height, pel, super_levels, *args, level, px, py, bw, bh, mvx, mvy, int(satd),
                            int(chroma), C.byref(l), C.byref(c))
    return l.value, c.value


def decode_payload(payload: np.ndarray):
    """Vector payload -> (header[16], [regions]) with each region an (N, 4) array dx, dy, score, luma."""
    hdr = payload[:64].view(np.int32).copy()
    regs = []
    off = 0x40
    n_reg = (1 if hdr[2] & 1 else 0) + (1 if hdr[2] & 2 else 0)
    count = int(hdr[12]) * int(hdr[13])
    for _ in range(n_reg):
        marker = payload[off:off + 4].view(np.int32)[0]
        assert marker == 2 * count + 1
        d = payload[off + 4:off + 4 + 8 * count].view(np.uint32).reshape(count, 2)
        dx = (d[:, 0] >> 16).astype(np.uint16).view(np.int16)
        dy = (d[:, 0] & 0xFFFF).astype(np.uint16).view(np.int16)
        regs.append(np.stack([dx.astype(np.int64), dy.astype(np.int64), (d[:, 1] & 0xFFFFFF).astype(np.int64),
                              (d[:, 1] >> 24).astype(np.int64)], 1))
        off += 4 + 8 * count
    return hdr, regs

_pp = C.POINTER(C.c_void_p)
L.oracle_smooth_opts.argtypes = [C.c_char_p, _sz, C.c_int64, C.c_int64, C.POINTER(C.c_int32), C.c_int32,
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p, _sz]
L.oracle_smooth_render.argtypes = [C.c_char_p, _sz, _i, _i, C.c_int64, C.c_int64, _i, C.c_int64,
                                   C.POINTER(C.c_int32)] + [_pp] * 6 + [_i, _pp, _sz, _i, u8p, u8p, u8p,
                                                                         C.POINTER(C.c_int32), _i]


def smooth_opts(opt, fps_num, fps_den, vdata, nframes):
    b = (opt or "").encode()
    vd = (C.c_int32 * 16)(*vdata)
    of, algo, sm = C.c_int32(), C.c_int32(), C.c_int32()
    err = C.create_string_buffer(256)
    rc = L.oracle_smooth_opts(b, len(b), fps_num, fps_den, vd, nframes, C.byref(of), C.byref(algo), C.byref(sm),
                              err, 256)
    if rc == -3:
        return None, err.value.decode()
    assert rc == 0
    return dict(out_frames=of.value, algo=algo.value, scene_mode=sm.value), None


class SmoothClip:
    """Host-side clip state for the oracle renderer: sources, super frames, vector payloads."""

    def __init__(self, opt, frames, supers, payloads, sdata, vdata, fps=(24, 1), float_render=False):
        self.float_render = 1 if float_render else 0
        self.opt = (opt or "").encode()
        self.h, self.w = frames[0][0].shape
        self.n = len(frames)
        self.fps = fps
        self.sdata = sdata
        self.vdata = (C.c_int32 * 16)(*vdata)
        self.keep = []

        def ptrs(arrs):
            arrs = [np.ascontiguousarray(a) for a in arrs]
            self.keep.extend(arrs)
            return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])

        self.src = [ptrs([f[p] for f in frames]) for p in range(3)]
        if supers is not None:
            self.sup = [ptrs([s[p][:, :(self.w if p == 0 else self.w // 2)] for s in supers]) for p in range(3)]
            self.super_rows = supers[0][0].shape[0]
        else:
            self.sup = [None, None, None]
            self.super_rows = 0
        self.vec = ptrs(payloads)
        self.vec_len = len(payloads[0])

    def padding(self):
        """Options::padding (options.rs:256-277): (pad_x, pad_y) on each side."""
        px, py = C.c_int32(), C.c_int32()
        rc = L.oracle_smooth_padding(self.opt, len(self.opt), self.w, self.h, self.fps[0], self.fps[1], self.vdata,
                                     C.byref(px), C.byref(py))
        assert rc == 0, rc
        return px.value, py.value

    def render(self, k):
        """Output frame k (padded by the light options' border, if any)."""
        px, py = self.padding()
        ow, oh = self.w + 2 * px, self.h + 2 * py
        Y = np.zeros((oh, ow), np.uint8)
        U = np.zeros((oh // 2, ow // 2), np.uint8)
        V = np.zeros((oh // 2, ow // 2), np.uint8)
        algo = C.c_int32()
        sup = [s if s is not None else C.cast(None, _pp) for s in self.sup]
        rc = L.oracle_smooth_render(self.opt, len(self.opt), self.w, self.h, self.fps[0], self.fps[1], self.n,
                                    self.sdata, self.vdata, self.src[0], self.src[1], self.src[2], sup[0], sup[1],
                                    sup[2], self.super_rows, self.vec, self.vec_len, k, _p(Y), _p(U), _p(V),
                                    C.byref(algo), self.float_render)
        assert rc >= 0, rc
        return (Y, U, V), rc, algo.value

L.oracle_smooth_padding.argtypes = [C.c_char_p, _sz, _i, _i, C.c_int64, C.c_int64, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
L.oracle_light_rng.argtypes = [C.c_uint32, _i, C.POINTER(C.c_int32)]
L.oracle_light_profile.argtypes = [C.POINTER(C.c_int64), _i, _i, _i, C.c_double, _i, C.POINTER(C.c_float)]
L.oracle_light_cubic.argtypes = [C.POINTER(C.c_int32), _i, _i, C.POINTER(C.c_float)]
L.oracle_light_corner_geometry.argtypes = [_i, _i, _i, C.POINTER(C.c_double), C.POINTER(C.c_double)]
L.oracle_light_apply.argtypes = [_i, _i, _i, _i, _i, _i, _i, C.c_double, _i, u8p, u8p, u8p]


def light_rng(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.int32)
    L.oracle_light_rng(seed & 0xFFFFFFFF, n, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def light_profile(accum, border: int, lights: int, cell: float, base: int) -> np.ndarray:
    a = np.ascontiguousarray(accum, np.int64)
    out = np.zeros(a.size, np.float32)
    L.oracle_light_profile(a.ctypes.data_as(C.POINTER(C.c_int64)), a.size, border, lights, cell, base,
                           out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def light_cubic(control, dim: int) -> np.ndarray:
    c = np.ascontiguousarray(control, np.int32)
    out = np.zeros(dim, np.float32)
    L.oracle_light_cubic(c.ctypes.data_as(C.POINTER(C.c_int32)), c.size, dim, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def light_corner_geometry(w: int, h: int, length: int):
    mix = np.zeros(w * h, np.float64)
    wt = np.zeros(w * h, np.float64)
    L.oracle_light_corner_geometry(w, h, length, mix.ctypes.data_as(C.POINTER(C.c_double)),
                                   wt.ctypes.data_as(C.POINTER(C.c_double)))
    return mix.reshape(h, w), wt.reshape(h, w)


def light_apply(planes, width: int, height: int, pad_x: int, pad_y: int, frame: int, border=12, lights=16,
                length=100, cell=1.0):
    """Draws the border of output frame `frame` into padded planes (modified in place)."""
    Y, U, V = planes
    L.oracle_light_apply(width, height, pad_x, pad_y, border, lights, length, cell, frame, _p(Y), _p(U), _p(V))
    return planes


# Per-block trace of the oracle's level searches (BlockTrace in oracle/analyse.h, 64 int32 words).
TRACE_FIELDS = ["level", "dir", "bx", "by", "px", "py", "bw", "bh", "edge", "smallest",
                "pred_dx", "pred_dy", "pred_score", "has_prev", "prev_dx", "prev_dy", "gx", "gy", "nseeds"] + \
               [f"seed{i}" for i in range(16)] + \
               ["lambda", "radius", "stype", "stage_mvx", "stage_mvy", "stage_sad", "mpred_x", "mpred_y",
                "found", "bad_before", "bad_limit", "expanded", "out_dx", "out_dy", "out_score", "out_luma"]
TRACE_IDX = {k: i for i, k in enumerate(TRACE_FIELDS)}
L.oracle_analyse_trace.argtypes = [C.c_char_p, _sz, C.c_int64, _i, _i] + [u8p, _i] * 6 + [
    u8p, _sz, C.POINTER(_sz), C.POINTER(C.c_int32), _sz, C.POINTER(_sz), C.POINTER(C.c_int32), _sz, C.POINTER(_sz)]


class Trace:
    """Rows of BlockTrace for one direction, in visit order.  `t[k]` is column k as
    int64 (scores are read back as unsigned 32-bit)."""

    def __init__(self, rows: np.ndarray):
        self.rows = rows.astype(np.int64)
        for k in ("pred_score", "stage_sad", "found", "out_score"):
            c = TRACE_IDX[k]
            self.rows[:, c] &= 0xFFFFFFFF

    def __len__(self):
        return len(self.rows)

    def __getitem__(self, k):
        return self.rows[:, TRACE_IDX[k]]

    def level(self, lv):
        return Trace.__new__(Trace)._set(self.rows[self.rows[:, 0] == lv])

    def _set(self, rows):
        self.rows = rows
        return self

    def rec(self, i):
        return {k: int(self.rows[i, j]) for j, k in enumerate(TRACE_FIELDS)}


def analyse_trace(opt: str | None, sdata: int, width: int, height: int, cur, ref):
    """get_frame_analyse with a trace: (payload, Trace dir 0 (cur->ref), Trace dir 1 (ref->cur))."""
    b = (opt or "").encode()
    o, err = analyse_opts(opt, sdata, width, height)
    if o is None:
        raise ValueError(err)
    out = np.zeros(o["payload_len"] + 64, np.uint8)
    n = _sz()
    keep = [np.ascontiguousarray(p) for p in list(cur) + list(ref)]
    args = []
    for p in keep:
        args += [_p(p), p.shape[1]]
    cap = 1
    while True:
        t0 = np.zeros((cap, 64), np.int32)
        t1 = np.zeros((cap, 64), np.int32)
        n0, n1 = _sz(), _sz()
        rc = L.oracle_analyse_trace(b, len(b), sdata, width, height, *args, _p(out), out.size, C.byref(n),
                                    t0.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n0),
                                    t1.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n1))
        if rc == 0:
            break
        assert rc == -1, rc
        cap = max(n0.value, n1.value, 1)
    return out[:n.value].copy(), Trace(t0[:n0.value]), Trace(t1[:n1.value])
