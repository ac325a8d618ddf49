"""ctypes wrapper of the test VapourSynth host (fakehost.cpp): load the bm3d
plugin, build argument maps, invoke functions, read frames and their
properties.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "_build", "libvsfakehost.so")
PLUGIN = os.path.join(ROOT, "vapoursynth-bm3d_amd", "_build", "libbm3d_vs.so")

GRAY, RGB, YUV = 1, 2, 3
INTEGER, FLOAT = 0, 1

_lib = None


class HostError(Exception):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.fh_core_new.restype = vp
        L.fh_core_free.argtypes = [vp]
        L.fh_load_plugin.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int]
        L.fh_plugin_field.argtypes = [vp, C.c_int, C.c_int]
        L.fh_plugin_field.restype = C.c_char_p
        L.fh_plugin_int.argtypes = [vp, C.c_int, C.c_int]
        L.fh_function_count.argtypes = [vp, C.c_int]
        L.fh_function_field.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.fh_function_field.restype = C.c_char_p
        L.fh_map_new.restype = vp
        L.fh_map_free.argtypes = [vp]
        L.fh_map_append_int.argtypes = [vp, C.c_char_p, C.c_int64]
        L.fh_map_append_float.argtypes = [vp, C.c_char_p, C.c_double]
        L.fh_map_set_data.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.fh_map_append_node.argtypes = [vp, C.c_char_p, vp]
        L.fh_source_new.argtypes = [C.c_int] * 8
        L.fh_source_new.restype = vp
        L.fh_source_set_plane.argtypes = [vp, C.c_int, C.c_int, vp, C.c_ssize_t]
        L.fh_source_set_prop.argtypes = [vp, C.c_int, C.c_char_p, C.c_int64]
        L.fh_invoke.argtypes = [vp, C.c_char_p, C.c_char_p, vp, C.c_char_p, C.c_int]
        L.fh_invoke.restype = vp
        L.fh_node_info.argtypes = [vp, C.POINTER(C.c_int)]
        L.fh_get_frame.argtypes = [vp, C.c_int, C.c_char_p, C.c_int]
        L.fh_get_frame.restype = vp
        L.fh_get_frames_parallel.argtypes = [vp, C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.fh_frame_plane.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_ssize_t), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        L.fh_frame_prop.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(C.c_int64)]
        L.fh_frame_prop_count.argtypes = [vp, C.c_char_p]
        L.fh_frame_free.argtypes = [vp]
        L.fh_node_free.argtypes = [vp]
        L.fh_log_count.argtypes = [vp]
        L.fh_log_message.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        L.fh_log_message.restype = C.c_char_p
        _lib = L
    return _lib


def _dtype(st, bits):
    if st == FLOAT:
        return np.float32
    return np.uint8 if bits <= 8 else np.uint16


class Node:
    def __init__(self, host, ptr):
        self.host = host
        self.ptr = ptr

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fh_node_free(self.ptr)
            self.ptr = None

    @property
    def info(self):
        out = (C.c_int * 8)()
        lib().fh_node_info(self.ptr, out)
        keys = ("color_family", "sample_type", "bits", "ssw", "ssh", "width", "height", "num_frames")
        return dict(zip(keys, out))

    def get_frame(self, n):
        err = C.create_string_buffer(1024)
        f = lib().fh_get_frame(self.ptr, n, err, 1024)
        if not f:
            raise HostError(err.value.decode())
        try:
            return self._read(f)
        finally:
            lib().fh_frame_free(f)

    def get_frames_parallel(self, count, threads):
        err = C.create_string_buffer(1024)
        if lib().fh_get_frames_parallel(self.ptr, count, threads, err, 1024):
            raise HostError(err.value.decode())

    def _read(self, f):
        info = self.info
        dt = _dtype(info["sample_type"], info["bits"])
        planes = []
        for p in range(1 if info["color_family"] == GRAY else 3):
            data, stride, w, h = C.c_void_p(), C.c_ssize_t(), C.c_int(), C.c_int()
            lib().fh_frame_plane(f, p, C.byref(data), C.byref(stride), C.byref(w), C.byref(h))
            isz = np.dtype(dt).itemsize
            buf = (C.c_uint8 * (stride.value * h.value)).from_address(data.value)
            rows = np.frombuffer(buf, np.uint8).reshape(h.value, stride.value)[:, :w.value * isz]
            planes.append(rows.copy().view(dt))
        props = {}
        for key in ("BM3D_OPP", "_Range", "_Matrix", "BM3D_V_radius", "BM3D_V_process"):
            n = lib().fh_frame_prop_count(f, key.encode())
            if n > 0:
                vals = []
                for i in range(n):
                    v = C.c_int64()
                    lib().fh_frame_prop(f, key.encode(), i, C.byref(v))
                    vals.append(v.value)
                props[key] = vals if key == "BM3D_V_process" else vals[0]
        return planes, props


class Host:
    def __init__(self):
        self.core = lib().fh_core_new()
        self.plugin = None

    def __del__(self):
        if getattr(self, "core", None):
            lib().fh_core_free(self.core)
            self.core = None

    def load(self, path=PLUGIN):
        err = C.create_string_buffer(1024)
        idx = lib().fh_load_plugin(self.core, path.encode(), err, 1024)
        if idx < 0:
            raise HostError(err.value.decode())
        self.plugin = idx
        return idx

    def plugin_info(self, idx=None):
        idx = self.plugin if idx is None else idx
        L = lib()
        return dict(identifier=L.fh_plugin_field(self.core, idx, 0).decode(),
                    namespace=L.fh_plugin_field(self.core, idx, 1).decode(),
                    name=L.fh_plugin_field(self.core, idx, 2).decode(),
                    version=L.fh_plugin_int(self.core, idx, 0), api_version=L.fh_plugin_int(self.core, idx, 1),
                    flags=L.fh_plugin_int(self.core, idx, 2))

    def functions(self, idx=None):
        idx = self.plugin if idx is None else idx
        L = lib()
        return [(L.fh_function_field(self.core, idx, k, 0).decode(), L.fh_function_field(self.core, idx, k, 1).decode(),
                 L.fh_function_field(self.core, idx, k, 2).decode()) for k in range(L.fh_function_count(self.core, idx))]

    def source(self, cf, st, bits, frames, props=None, ssw=0, ssh=0):
        """frames: list of per-frame plane lists (2-D numpy arrays)."""
        h, w = frames[0][0].shape
        ptr = lib().fh_source_new(cf, st, bits, ssw, ssh, w, h, len(frames))
        if not ptr:
            raise HostError("bad source format")
        node = Node(self, ptr)
        for k, planes in enumerate(frames):
            for p, a in enumerate(planes):
                a = np.ascontiguousarray(a)
                lib().fh_source_set_plane(ptr, k, p, a.ctypes.data, a.strides[0])
            for key, v in (props or {}).items():
                lib().fh_source_set_prop(ptr, k, key.encode(), int(v))
        return node

    def invoke(self, fn, **kw):
        L = lib()
        m = L.fh_map_new()
        try:
            for k, v in kw.items():
                key = k.encode()
                vals = v if isinstance(v, (list, tuple)) else [v]
                for x in vals:
                    if isinstance(x, Node):
                        L.fh_map_append_node(m, key, x.ptr)
                    elif isinstance(x, str):
                        L.fh_map_set_data(m, key, x.encode())
                    elif isinstance(x, float):
                        L.fh_map_append_float(m, key, x)
                    else:
                        L.fh_map_append_int(m, key, int(x))
            err = C.create_string_buffer(1024)
            ptr = L.fh_invoke(self.core, b"bm3d", fn.encode(), m, err, 1024)
        finally:
            L.fh_map_free(m)
        if not ptr:
            raise HostError(err.value.decode())
        return Node(self, ptr)

    def log(self):
        out = []
        for i in range(lib().fh_log_count(self.core)):
            t = C.c_int()
            msg = lib().fh_log_message(self.core, i, C.byref(t)).decode()
            out.append((t.value, msg))
        return out
