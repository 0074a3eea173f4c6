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
