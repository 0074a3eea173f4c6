"""A minimal VapourSynth API 3 core in Python, for testing the svp1 / svp2 plugin
shells without VapourSynth: it fills a VSAPI table at the byte offsets the plugins
read, keeps maps, nodes and frames as Python objects behind opaque handles, and runs
filters with VapourSynth's two-phase protocol (getFrame with arInitial collects the
requests, the requested frames are produced, then getFrame with arAllFramesReady).
Frames get 64-byte aligned strides, so padding bytes are exercised."""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np

AR_INITIAL, AR_ALL_FRAMES_READY = 0, 2
PE_UNSET, PE_INDEX = 1, 4


class Format(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("id", C.c_int32), ("color_family", C.c_int32),
                ("sample_type", C.c_int32), ("bits_per_sample", C.c_int32), ("bytes_per_sample", C.c_int32),
                ("sub_sampling_w", C.c_int32), ("sub_sampling_h", C.c_int32), ("num_planes", C.c_int32)]


class VideoInfo(C.Structure):
    _fields_ = [("format", C.c_void_p), ("fps_num", C.c_int64), ("fps_den", C.c_int64), ("width", C.c_int32),
                ("height", C.c_int32), ("num_frames", C.c_int32), ("flags", C.c_int32)]


# preset ids and layouts of the two formats the path uses
YUV420P8 = Format(b"YUV420P8", 3000010, 3000000, 0, 8, 1, 1, 1, 3)
GRAY8 = Format(b"Gray8", 1000010, 1000000, 0, 8, 1, 0, 0, 1)
YUV444P8 = Format(b"YUV444P8", 3000013, 3000000, 0, 8, 1, 0, 0, 3)
FORMATS = {f.id: f for f in (YUV420P8, GRAY8, YUV444P8)}

VP = C.c_void_p
PUBLIC = C.CFUNCTYPE(None, VP, VP, VP, VP, VP)
INIT = C.CFUNCTYPE(None, VP, VP, C.POINTER(VP), VP, VP, VP)
GETFRAME = C.CFUNCTYPE(VP, C.c_int, C.c_int, C.POINTER(VP), C.POINTER(VP), VP, VP, VP)
FREE = C.CFUNCTYPE(None, VP, VP, VP)
CONFIG = C.CFUNCTYPE(None, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, VP)
REGISTER = C.CFUNCTYPE(None, C.c_char_p, C.c_char_p, VP, VP, VP)


class VSError(RuntimeError):
    pass


class Frame:
    def __init__(self, fmt: Format, width: int, height: int, props=None):
        self.fmt, self.width, self.height = fmt, width, height
        self.planes = []
        for p in range(fmt.num_planes):
            w = width >> (fmt.sub_sampling_w if p else 0)
            h = height >> (fmt.sub_sampling_h if p else 0)
            stride = (w + 63) // 64 * 64
            self.planes.append(np.zeros((h, stride), np.uint8))
        self.props = {} if props is None else {k: list(v) for k, v in props.items()}

    def view(self, plane: int):
        """The plane without its row padding."""
        w = self.width >> (self.fmt.sub_sampling_w if plane else 0)
        return self.planes[plane][:, :w]


class Node:
    def __init__(self, vi: VideoInfo, frames=None):
        self.vi = vi
        self.frames = frames          # a source clip: list of Frame
        self.getframe = self.free = None
        self.instance = VP()
        self.name = "source"


class Core:
    """One fake core: load plugins, call their functions, fetch frames."""

    def __init__(self):
        self._ids = itertools.count(0x10000, 16)
        self.handles = {}             # handle -> object (Map dict, Node, Frame, frame context)
        self.plugins = {}             # namespace -> {"id", "name", "api", "functions": {name: (args, fn)}}
        self.cache = {}               # (node handle, n) -> Frame produced by a filter
        self.errors = []
        self.core_h = self._new({"core": True})
        self._keep = []
        self._table = (VP * 64)()
        def cb(ftype, off, fn):
            failed = None if ftype._restype_ is None else 0

            def guarded(*a):
                try:
                    return fn(*a)
                except Exception as e:  # surfaced after the plugin call returns
                    self.errors.append(e)
                    return failed
            f = ftype(guarded)
            self._keep.append(f)
            self._table[off // 8] = C.cast(f, VP)
        cb(C.CFUNCTYPE(None, VP), 48, self._free_frame)
        cb(C.CFUNCTYPE(None, VP), 56, lambda h: None)
        cb(C.CFUNCTYPE(VP, VP, C.c_int, C.c_int, VP, VP), 72, self._new_video_frame)
        cb(C.CFUNCTYPE(None, VP, VP, C.c_char_p, VP, VP, VP, C.c_int, C.c_int, VP, VP), 136, self._create_filter)
        cb(C.CFUNCTYPE(None, VP, C.c_char_p), 144, self._set_error)
        cb(C.CFUNCTYPE(VP, C.c_int, VP), 176, lambda i, c: C.addressof(FORMATS[i]) if i in FORMATS else None)
        cb(C.CFUNCTYPE(VP, C.c_int, VP, VP), 208, self._get_frame_filter)
        cb(C.CFUNCTYPE(None, C.c_int, VP, VP), 216, self._request_frame_filter)
        cb(C.CFUNCTYPE(C.c_int, VP, C.c_int), 240, lambda h, p: self.handles[h].planes[p].shape[1])
        cb(C.CFUNCTYPE(VP, VP, C.c_int), 248, lambda h, p: self.handles[h].planes[p].ctypes.data)
        cb(C.CFUNCTYPE(VP, VP, C.c_int), 256, lambda h, p: self.handles[h].planes[p].ctypes.data)
        cb(C.CFUNCTYPE(VP, VP), 304, lambda h: C.addressof(self.handles[h].vi))
        cb(C.CFUNCTYPE(None, VP, C.c_int, VP), 312, self._set_video_info)
        cb(C.CFUNCTYPE(VP, VP), 344, lambda h: self._props_handle(h))
        cb(C.CFUNCTYPE(VP, VP), 352, lambda h: self._props_handle(h))
        cb(C.CFUNCTYPE(C.c_int64, VP, C.c_char_p, C.c_int, C.POINTER(C.c_int)), 392, self._getter("int"))
        cb(C.CFUNCTYPE(C.c_double, VP, C.c_char_p, C.c_int, C.POINTER(C.c_int)), 400, self._getter("float"))
        cb(C.CFUNCTYPE(VP, VP, C.c_char_p, C.c_int, C.POINTER(C.c_int)), 408, self._getter("data"))
        cb(C.CFUNCTYPE(C.c_int, VP, C.c_char_p, C.c_int, C.POINTER(C.c_int)), 416, self._getter("size"))
        cb(C.CFUNCTYPE(VP, VP, C.c_char_p, C.c_int, C.POINTER(C.c_int)), 424, self._getter("node"))
        cb(C.CFUNCTYPE(C.c_int, VP, C.c_char_p, C.c_int64, C.c_int), 456, self._set_int)
        self.api = C.addressof(self._table)

    # ---------------------------------------------------------------- handles
    def _new(self, obj):
        h = next(self._ids)
        self.handles[h] = obj
        return h

    def _check(self):
        if self.errors:
            e = self.errors[0]
            self.errors = []
            raise e

    # ---------------------------------------------------------------- VSAPI entries
    def _free_frame(self, h):
        self.handles.pop(h, None)

    def _new_video_frame(self, fmt_p, w, h, prop_src, core):
        fmt = Format.from_address(fmt_p)
        props = self.handles[prop_src].props if prop_src else None
        return self._new(Frame(fmt, w, h, props))

    def _create_filter(self, in_h, out_h, name, init, getframe, free, mode, flags, instance, core):
        node = Node(VideoInfo())
        node.name = name.decode()
        node.getframe = C.cast(getframe, GETFRAME)
        node.free = C.cast(free, FREE)
        node.instance = VP(instance)
        nh = self._new(node)
        C.cast(init, INIT)(in_h, out_h, C.byref(node.instance), nh, self.core_h, self.api)
        self.handles[out_h].setdefault("clip", []).append(("node", nh))

    def _set_error(self, m, msg):
        self.handles[m]["_error"] = [("data", msg)]

    def _set_video_info(self, vi_p, n, node_h):
        C.memmove(C.addressof(self.handles[node_h].vi), vi_p, C.sizeof(VideoInfo))

    def _props_handle(self, frame_h):
        f = self.handles[frame_h]
        return self._new(f.props)

    def _getter(self, kind):
        def get(m, key, index, err):
            d = self.handles[m]
            vals = d.get(key.decode())
            if vals is None or index >= len(vals):
                if err:
                    err[0] = PE_UNSET if vals is None else PE_INDEX
                return 0 if kind in ("int", "float", "size") else None
            if err:
                err[0] = 0
            t, v = vals[index]
            if kind == "int":
                return int(v)
            if kind == "float":
                return float(v)
            if kind == "node":
                return v
            buf = v if isinstance(v, C.Array) else None
            if buf is None:
                buf = C.create_string_buffer(bytes(v), len(v))
                vals[index] = (t, buf)
            return C.addressof(buf) if kind == "data" else len(buf)
        return get

    def _set_int(self, m, key, v, append):
        d = self.handles[m]
        if append:
            d.setdefault(key.decode(), []).append(("int", v))
        else:
            d[key.decode()] = [("int", v)]
        return 0

    def _request_frame_filter(self, n, node_h, fc_h):
        self.handles[fc_h]["requests"].append((node_h, n))

    def _get_frame_filter(self, n, node_h, fc_h):
        node = self.handles[node_h]
        n = min(max(n, 0), node.vi.num_frames - 1)
        frame = self.handles[fc_h]["frames"].get((node_h, n))
        return self._new(frame) if frame is not None else None

    # ---------------------------------------------------------------- host side
    def load(self, path: str):
        lib = C.CDLL(path)
        lib.svpGetVersion.restype = C.c_int64
        info = {}

        def config(ident, ns, name, api, ro, plugin):
            info.update(id=ident.decode(), ns=ns.decode(), name=name.decode(), api=api, readonly=ro,
                        functions={})

        def register(name, args, fn, user, plugin):
            info["functions"][name.decode()] = (args.decode(), fn)

        cfg, reg = CONFIG(config), REGISTER(register)
        lib.VapourSynthPluginInit(cfg, reg, VP(0x1234))
        info["version"] = lib.svpGetVersion()
        info["lib"] = lib
        self.plugins[info["ns"]] = info
        return info

    def source(self, frames, fps=(24, 1), props=None):
        """A YUV420P8 clip from [(y, u, v)] arrays; props(i) -> dict of int frame props."""
        h, w = frames[0][0].shape
        vi = VideoInfo(C.addressof(YUV420P8), fps[0], fps[1], w, h, len(frames), 0)
        out = []
        for i, planes in enumerate(frames):
            f = Frame(YUV420P8, w, h)
            for p in range(3):
                f.view(p)[:] = planes[p]
            if props:
                f.props = {k: [("int", v)] for k, v in props(i).items()}
            out.append(f)
        return self._new(Node(vi, out))

    def invoke(self, ns: str, name: str, **args):
        """Calls a registered function; returns its output map (dict) or raises VSError."""
        fn = self.plugins[ns]["functions"][name][1]
        m = {}
        for k, v in args.items():
            if isinstance(v, bytes):
                m[k] = [("data", v)]
            elif isinstance(v, float):
                m[k] = [("float", v)]
            elif isinstance(v, int) and k in ("clip", "super", "vectors", "src"):
                m[k] = [("node", v)]
            else:
                m[k] = [("int", v)]
        in_h, out_h = self._new(m), self._new({})
        C.cast(fn, PUBLIC)(in_h, out_h, None, self.core_h, self.api)
        self._check()
        out = self.handles[out_h]
        if "_error" in out:
            raise VSError(out["_error"][0][1].decode())
        return out

    def video_info(self, node_h) -> VideoInfo:
        return self.handles[node_h].vi

    def frame(self, node_h, n) -> Frame:
        node = self.handles[node_h]
        n = min(max(n, 0), node.vi.num_frames - 1)
        if node.frames is not None:
            return node.frames[n]
        key = (node_h, n)
        if key in self.cache:
            return self.cache[key]
        ctx = {"requests": [], "frames": {}}
        fc = self._new(ctx)
        frame_data = VP()
        node.getframe(n, AR_INITIAL, C.byref(node.instance), C.byref(frame_data), fc, self.core_h, self.api)
        self._check()
        for rh, rn in ctx["requests"]:
            rn = min(max(rn, 0), self.handles[rh].vi.num_frames - 1)
            ctx["frames"][(rh, rn)] = self.frame(rh, rn)
        ret = node.getframe(n, AR_ALL_FRAMES_READY, C.byref(node.instance), C.byref(frame_data), fc, self.core_h,
                            self.api)
        self._check()
        if not ret:
            raise VSError(f"{node.name}: frame {n} failed")
        frame = self.handles.pop(ret)
        self.handles.pop(fc, None)
        self.cache[key] = frame
        return frame

    def close(self):
        """Frees every filter instance (VapourSynth's free callbacks)."""
        for h, obj in list(self.handles.items()):
            if isinstance(obj, Node) and obj.free:
                obj.free(obj.instance, self.core_h, self.api)
                obj.free = None
        self._check()
