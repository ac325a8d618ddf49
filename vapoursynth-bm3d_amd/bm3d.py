"""Host-side mirror of the VapourSynth ``bm3d`` namespace over the HIP C ABI.

The reference plugin exposes seven functions to VapourSynth scripts
(source/VSPlugin.cpp:430-517)::

    core.bm3d.RGB2OPP(clip input[, int sample=0])
    core.bm3d.OPP2RGB(clip input[, int sample=0])
    core.bm3d.Basic(clip input[, clip ref, string profile, float[] sigma, int block_size, ...])
    core.bm3d.Final(clip input, clip ref[, ...])
    core.bm3d.VBasic / VFinal(... radius, ps_num, ps_range, ps_step ...)
    core.bm3d.VAggregate(clip input[, int radius=1, int sample=0])

This module offers the same names, arguments and error messages for clips
held as numpy arrays (host) or torch tensors on the GPU (device).  A
:class:`Clip` is evaluated lazily frame by frame, like a VapourSynth node:
``clip.get_frame(n)`` asks the filter for the input frames it needs
(``bm3d_frame_range``) and runs ``bm3d_get_frame``.  Every filter runs on
the HIP library (libbm3d_amd.so); there is no CPU fallback -- if the library
or a GPU is missing, creating a filter raises :class:`Error`.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libbm3d_amd.so")

GRAY, RGB, YUV = 1, 2, 3
INTEGER, FLOAT = 0, 1


class Error(Exception):
    """Raised where VapourSynth would raise vs.Error."""


# --------------------------------------------------------------------------- C ABI

class ClipInfo(C.Structure):
    _fields_ = [("color_family", C.c_int), ("sample_type", C.c_int), ("bits_per_sample", C.c_int),
                ("subsampling_w", C.c_int), ("subsampling_h", C.c_int), ("width", C.c_int),
                ("height", C.c_int), ("num_frames", C.c_int)]


class Arg(C.Structure):
    _fields_ = [("key", C.c_char_p), ("type", C.c_int), ("count", C.c_int),
                ("ints", C.POINTER(C.c_int64)), ("floats", C.POINTER(C.c_double)), ("data", C.c_char_p)]


class Props(C.Structure):
    _fields_ = [("bm3d_opp", C.c_int64), ("range", C.c_int64), ("matrix", C.c_int64),
                ("v_radius", C.c_int64), ("v_process", C.c_int64 * 3)]


class FrameC(C.Structure):
    _fields_ = [("data", C.c_void_p * 3), ("stride", C.c_ssize_t * 3), ("width", C.c_int),
                ("height", C.c_int), ("mem", C.c_int), ("props", Props)]


class Stats(C.Structure):
    _fields_ = [("ref_blocks", C.c_int64), ("ssd_evals", C.c_int64), ("members", C.c_int64),
                ("members_sq", C.c_int64)]


class BatchShape(C.Structure):
    _fields_ = [("match_waves", C.c_int), ("match_strip", C.c_int), ("group_order", C.c_int)]


EXPORTS = ("bm3d_init", "bm3d_last_error", "bm3d_set_log", "bm3d_create", "bm3d_output_info",
           "bm3d_frame_range", "bm3d_get_frame", "bm3d_free", "bm3d_last_stats", "bm3d_profile",
           "bm3d_profile_read", "bm3d_debug_matches", "bm3d_debug_vmatches", "bm3d_get_frame_noalloc",
           "bm3d_noalloc_scratch_bytes", "bm3d_host_arith")

_lib = None
_initialised = False
_LOGFN = C.CFUNCTYPE(None, C.c_int, C.c_char_p, C.c_void_p)
_log_cb = None


def load_library(path: str = LIB_PATH):
    """Loads libbm3d_amd.so and declares the C signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise Error(f"bm3d: HIP library not built ({path}); run vapoursynth-bm3d_amd/build.py")
    # Device tensors handed over by torch must live in the HIP runtime the
    # library runs on: load torch (and the runtime it brings) first, so that
    # the library binds to that same runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    L.bm3d_init.argtypes = [C.c_int]
    L.bm3d_set_log.argtypes = [_LOGFN, C.c_void_p]
    L.bm3d_host_arith.restype = C.c_int
    L.bm3d_set_log.restype = None
    L.bm3d_last_error.restype = C.c_char_p
    L.bm3d_create.argtypes = [C.c_char_p, C.POINTER(ClipInfo), C.POINTER(ClipInfo), C.POINTER(Arg), C.c_int,
                              C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
    L.bm3d_output_info.argtypes = [C.c_void_p, C.POINTER(ClipInfo)]
    L.bm3d_frame_range.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.bm3d_get_frame.argtypes = [C.c_void_p, C.c_int, C.POINTER(FrameC), C.POINTER(FrameC), C.POINTER(FrameC)]
    L.bm3d_get_frame_noalloc.argtypes = [C.c_void_p, C.c_int, C.POINTER(FrameC), C.POINTER(FrameC),
                                         C.POINTER(FrameC)]
    L.bm3d_noalloc_scratch_bytes.restype = C.c_size_t
    L.bm3d_free.argtypes = [C.c_void_p]
    L.bm3d_free.restype = None
    L.bm3d_last_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.bm3d_profile.argtypes = [C.c_void_p, C.c_int]
    L.bm3d_profile_read.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]
    L.bm3d_debug_matches.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_int]
    L.bm3d_debug_vmatches.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_int]
    # learned block-matching proposal (bm3d_lbm.h)
    L.bm3d_lbm_param_count.restype = C.c_int
    L.bm3d_lbm_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_float)]
    L.bm3d_lbm_free.argtypes = [C.c_void_p]
    L.bm3d_lbm_free.restype = None
    L.bm3d_lbm_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.bm3d_lbm_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.bm3d_lbm_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float,
                                      C.POINTER(C.c_float)]
    L.bm3d_lbm_gradients.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                     C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.bm3d_lbm_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    L.bm3d_lbm_match_plane.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int),
                                       C.POINTER(C.c_float)]
    L.bm3d_filter_params.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.bm3d_set_matcher.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.bm3d_matcher_fallbacks.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.bm3d_match_lists.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                                   C.c_void_p]
    L.bm3d_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(Arg), C.c_int, C.c_char_p,
                                    C.c_size_t]
    L.bm3d_batch_free.argtypes = [C.c_void_p]
    L.bm3d_batch_free.restype = None
    L.bm3d_batch_run.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int,
                                 C.POINTER(C.c_float)]
    L.bm3d_batch_tune.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, C.c_char_p,
                                  C.c_size_t]
    L.bm3d_batch_get_shape.argtypes = [C.c_void_p, C.POINTER(BatchShape)]
    L.bm3d_batch_set_shape.argtypes = [C.c_void_p, C.POINTER(BatchShape)]
    L.bm3d_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    _lib = L
    return L


def host_arith() -> int:
    """bm3d_host_arith(): 0 = x86 SSE2 arithmetic, 1 = scalar (no GPU needed)."""
    return int(load_library().bm3d_host_arith())


def init(device: int = 0):
    global _initialised
    L = load_library()
    if not _initialised:
        if L.bm3d_init(device) != 0:
            raise Error("bm3d: " + L.bm3d_last_error().decode())
        _initialised = True
    return L


def set_log(fn: Optional[Callable[[int, str], None]]) -> None:
    """Routes the library's warnings -- vsapi->logMessage(mtWarning, ...) in
    the reference (include/BM3D_Base.h:202-207, VBM3D_Base.h:265-270,
    VAggregate.h:153-194) -- to fn(level, message); None restores stderr."""
    global _log_cb
    L = load_library()
    if fn is None:
        _log_cb = None
        L.bm3d_set_log(C.cast(None, _LOGFN), None)
        return
    _log_cb = _LOGFN(lambda level, msg, user: fn(level, msg.decode()))
    L.bm3d_set_log(_log_cb, None)


# --------------------------------------------------------------------------- clips

@dataclass
class VideoFormat:
    color_family: int
    sample_type: int
    bits_per_sample: int
    subsampling_w: int = 0
    subsampling_h: int = 0

    @property
    def num_planes(self):
        return 1 if self.color_family == GRAY else 3

    @property
    def bytes_per_sample(self):
        return 4 if self.sample_type == FLOAT else (2 if self.bits_per_sample > 8 else 1)

    @property
    def np_dtype(self):
        if self.sample_type == FLOAT:
            return np.float32
        return np.uint16 if self.bits_per_sample > 8 else np.uint8


GRAYS = VideoFormat(GRAY, FLOAT, 32)
GRAY16 = VideoFormat(GRAY, INTEGER, 16)
RGBS = VideoFormat(RGB, FLOAT, 32)
RGB48 = VideoFormat(RGB, INTEGER, 16)
YUV444PS = VideoFormat(YUV, FLOAT, 32)
YUV444P16 = VideoFormat(YUV, INTEGER, 16)


@dataclass
class Frame:
    """One frame: planes are 2-D numpy arrays (host) or torch tensors (GPU)."""
    planes: List
    props: Dict[str, object] = field(default_factory=dict)

    @property
    def on_device(self):
        return not isinstance(self.planes[0], np.ndarray)


class Clip:
    def __init__(self, fmt: VideoFormat, width: int, height: int, num_frames: int,
                 getter: Callable[[int], Frame], owner=None):
        self.format = fmt
        self.width = width
        self.height = height
        self.num_frames = num_frames
        self._get = getter
        self._owner = owner
        self._cache: Dict[int, Frame] = {}

    @property
    def filter(self):
        """The native filter behind this clip (None for source clips)."""
        return self._owner

    def clear_cache(self):
        self._cache.clear()

    def evict(self, n: int) -> None:
        """Drops cached frame n (recomputed if asked for again)."""
        self._cache.pop(n, None)

    def get_frame(self, n: int) -> Frame:
        if n < 0 or n >= self.num_frames:
            raise Error(f"frame {n} out of range")
        if n not in self._cache:
            self._cache[n] = self._get(n)
        return self._cache[n]

    def frames(self):
        return [self.get_frame(n) for n in range(self.num_frames)]

    def info(self) -> ClipInfo:
        f = self.format
        return ClipInfo(f.color_family, f.sample_type, f.bits_per_sample, f.subsampling_w, f.subsampling_h,
                        self.width, self.height, self.num_frames)

    @staticmethod
    def from_frames(fmt: VideoFormat, frames: Sequence[Frame]) -> "Clip":
        frames = list(frames)
        p0 = frames[0].planes[0]
        h, w = p0.shape[-2], p0.shape[-1]
        return Clip(fmt, w, h, len(frames), lambda n: frames[n])

    @staticmethod
    def from_arrays(fmt: VideoFormat, arrays, props=None) -> "Clip":
        """arrays: sequence of (P, H, W) arrays/tensors, one per frame."""
        frames = [Frame([a[p] for p in range(a.shape[0])], dict(props or {})) for a in arrays]
        return Clip.from_frames(fmt, frames)


# --------------------------------------------------------------------------- plumbing

def _props_to_c(props: Dict[str, object]) -> Props:
    p = Props(-1, -1, -1, -1, (C.c_int64 * 3)(-1, -1, -1))
    if "BM3D_OPP" in props:
        p.bm3d_opp = int(props["BM3D_OPP"])
    if "_Range" in props:
        p.range = int(props["_Range"])
    if "_Matrix" in props:
        p.matrix = int(props["_Matrix"])
    if "BM3D_V_radius" in props:
        p.v_radius = int(props["BM3D_V_radius"])
    if "BM3D_V_process" in props:
        v = list(props["BM3D_V_process"])
        for i in range(min(3, len(v))):
            p.v_process[i] = int(v[i])
    return p


def _props_from_c(p: Props, base: Dict[str, object]) -> Dict[str, object]:
    out = {k: v for k, v in base.items()
           if k not in ("BM3D_OPP", "_Range", "_Matrix", "BM3D_V_radius", "BM3D_V_process")}
    if p.bm3d_opp >= 0:
        out["BM3D_OPP"] = p.bm3d_opp
    if p.range >= 0:
        out["_Range"] = p.range
    if p.matrix >= 0:
        out["_Matrix"] = p.matrix
    if p.v_radius >= 0:
        out["BM3D_V_radius"] = p.v_radius
    if p.v_process[0] >= 0:
        out["BM3D_V_process"] = [p.v_process[i] for i in range(3)]
    return out


def _plane_ptr(a):
    if isinstance(a, np.ndarray):
        return a.ctypes.data, a.strides[0], 0
    return a.data_ptr(), a.stride(0) * a.element_size(), 1


def _sync_producer(device) -> None:
    """Waits for torch's current stream on device: the library's streams are
    ordered only after the legacy default stream, so device inputs produced on
    another torch stream must be complete before they are handed over."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def _frame_to_c(fr: Frame, keep: list) -> FrameC:
    c = FrameC()
    mem = None
    for p, a in enumerate(fr.planes[:3]):
        if isinstance(a, np.ndarray) and not a.flags.c_contiguous and a.strides[1] != a.itemsize:
            a = np.ascontiguousarray(a)
        keep.append(a)
        ptr, stride, m = _plane_ptr(a)
        c.data[p] = ptr
        c.stride[p] = stride
        mem = m
    c.width = fr.planes[0].shape[-1]
    c.height = fr.planes[0].shape[-2]
    c.mem = mem
    c.props = _props_to_c(fr.props)
    return c


def _alloc_like(fmt: VideoFormat, w: int, h: int, device):
    planes = []
    for p in range(fmt.num_planes):
        pw = w if p == 0 else w >> fmt.subsampling_w
        ph = h if p == 0 else h >> fmt.subsampling_h
        if device is None:
            planes.append(np.empty((ph, pw), fmt.np_dtype))
        else:
            import torch
            tdt = {np.float32: torch.float32, np.uint16: torch.uint16, np.uint8: torch.uint8}[fmt.np_dtype]
            planes.append(torch.empty((ph, pw), dtype=tdt, device=device))
    return planes


def _make_args(kwargs: Dict[str, object], keep: list):
    arr = (Arg * max(1, len(kwargs)))()
    i = 0
    for k, v in kwargs.items():
        if v is None:
            continue
        a = Arg()
        kb = k.encode()
        keep.append(kb)
        a.key = kb
        if isinstance(v, str):
            db = v.encode()
            keep.append(db)
            a.type, a.count, a.data = 2, 1, db
        else:
            vals = list(v) if isinstance(v, (list, tuple)) else [v]
            if any(isinstance(x, float) for x in vals) or k in ("sigma", "th_mse", "hard_thr"):
                buf = (C.c_double * len(vals))(*[float(x) for x in vals])
                a.type, a.count, a.floats = 1, len(vals), C.cast(buf, C.POINTER(C.c_double))
            else:
                buf = (C.c_int64 * len(vals))(*[int(x) for x in vals])
                a.type, a.count, a.ints = 0, len(vals), C.cast(buf, C.POINTER(C.c_int64))
            keep.append(buf)
        arr[i] = a
        i += 1
    return arr, i


class _Filter:
    def __init__(self, name: str, clip: Clip, ref: Optional[Clip], kwargs: Dict[str, object]):
        L = init()
        self.L = L
        self.name = name
        self.clip = clip
        self.ref = ref
        keep: list = []
        args, nargs = _make_args(kwargs, keep)
        handle = C.c_void_p()
        err = C.create_string_buffer(1024)
        ii = clip.info()
        ri = ref.info() if ref is not None else None
        rc = L.bm3d_create(name.encode(), C.byref(ii), C.byref(ri) if ri is not None else None, args, nargs,
                           C.byref(handle), err, 1024)
        if rc != 0:
            msg = err.value.decode() or L.bm3d_last_error().decode()
            raise Error(msg)
        self.h = handle
        info = ClipInfo()
        L.bm3d_output_info(self.h, C.byref(info))
        self.out_fmt = VideoFormat(info.color_family, info.sample_type, info.bits_per_sample, info.subsampling_w,
                                   info.subsampling_h)
        self.out_w, self.out_h, self.out_n = info.width, info.height, info.num_frames

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.bm3d_free(self.h)
        except Exception:
            pass

    def get_frame(self, n: int, noalloc: bool = False) -> Frame:
        """Output frame n.  noalloc: through bm3d_get_frame_noalloc (Basic and
        Final only; scratch from the library's fixed device area)."""
        first, last = C.c_int(), C.c_int()
        self.L.bm3d_frame_range(self.h, n, C.byref(first), C.byref(last))
        keep: list = []
        k = last.value - first.value + 1
        srcs = (FrameC * k)()
        refs = (FrameC * k)() if self.ref is not None else None
        for i, m in enumerate(range(first.value, last.value + 1)):
            srcs[i] = _frame_to_c(self.clip.get_frame(m), keep)
            if refs is not None:
                refs[i] = _frame_to_c(self.ref.get_frame(m), keep)
        src0 = self.clip.get_frame(n)
        device = None if not src0.on_device else src0.planes[0].device
        if device is not None:
            _sync_producer(device)
        planes = _alloc_like(self.out_fmt, self.out_w, self.out_h, device)
        out = Frame(planes, {})
        dst = _frame_to_c(out, keep)
        entry = self.L.bm3d_get_frame_noalloc if noalloc else self.L.bm3d_get_frame
        rc = entry(self.h, n, srcs, refs, C.byref(dst))
        if rc != 0:
            raise Error(f"bm3d.{self.name}: " + self.L.bm3d_last_error().decode())
        out.props = _props_from_c(dst.props, src0.props)
        return out

    def stats(self) -> Dict[str, int]:
        s = Stats()
        if self.L.bm3d_last_stats(self.h, C.byref(s)) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return dict(ref_blocks=s.ref_blocks, ssd_evals=s.ssd_evals, members=s.members, members_sq=s.members_sq)

    def profile(self, enable: bool = True):
        self.L.bm3d_profile(self.h, int(enable))

    def profile_read(self) -> Dict[str, tuple]:
        out = {}
        i = 0
        while True:
            name, ms, n = C.c_char_p(), C.c_double(), C.c_int64()
            if self.L.bm3d_profile_read(self.h, i, C.byref(name), C.byref(ms), C.byref(n)) != 0:
                break
            out[name.value.decode()] = (ms.value, n.value)
            i += 1
        return out

    def matches(self, b: int, cap: int = 256):
        cnt = C.c_int()
        ys = (C.c_int * cap)()
        xs = (C.c_int * cap)()
        ks = (C.c_float * cap)()
        rc = self.L.bm3d_debug_matches(self.h, b, C.byref(cnt), ys, xs, ks, cap)
        if rc != 0:
            raise Error(self.L.bm3d_last_error().decode())
        n = cnt.value
        return np.array(ks[:n], np.float32), np.array(ys[:n], np.int32), np.array(xs[:n], np.int32)

    def vmatches(self, b: int, cap: int = 512):
        """VBasic/VFinal: (keys, frames, ys, xs) of reference block b of the last frame."""
        cnt = C.c_int()
        zs = (C.c_int * cap)()
        ys = (C.c_int * cap)()
        xs = (C.c_int * cap)()
        ks = (C.c_float * cap)()
        rc = self.L.bm3d_debug_vmatches(self.h, b, C.byref(cnt), zs, ys, xs, ks, cap)
        if rc != 0:
            raise Error(self.L.bm3d_last_error().decode())
        n = cnt.value
        return (np.array(ks[:n], np.float32), np.array(zs[:n], np.int32), np.array(ys[:n], np.int32),
                np.array(xs[:n], np.int32))


    # -- learned block-matching proposal (no counterpart in the reference)
    def set_matcher(self, model, topk: int = 32):
        """Basic/Final: propose candidates with a trained lbm.Model (None:
        exhaustive search).  The cached frames of the clip are not cleared."""
        self._matcher = model  # keep the model alive while attached
        rc = self.L.bm3d_set_matcher(self.h, model.h if model is not None else None, topk)
        if rc != 0:
            raise Error(self.L.bm3d_last_error().decode())

    def params(self) -> Dict[str, int]:
        """Basic/Final matching parameters after profile defaults."""
        out = (C.c_int * 5)()
        if self.L.bm3d_filter_params(self.h, out) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return dict(block_size=out[0], block_step=out[1], group_size=out[2], bm_range=out[3], bm_step=out[4])

    def matcher_fallbacks(self) -> int:
        n = C.c_int()
        if self.L.bm3d_matcher_fallbacks(self.h, C.byref(n)) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return n.value

    def match_lists(self):
        """count[nb], pos[nb, G] as (y, x) pairs, key[nb, G] of the last frame."""
        nb, g = C.c_int64(), C.c_int()
        self.L.bm3d_match_lists(self.h, C.byref(nb), C.byref(g), None, None, None)
        cnt = np.zeros(nb.value, np.int32)
        pos = np.zeros((nb.value, g.value), np.int32)
        key = np.zeros((nb.value, g.value), np.float32)
        rc = self.L.bm3d_match_lists(self.h, C.byref(nb), C.byref(g), cnt.ctypes.data, pos.ctypes.data,
                                     key.ctypes.data)
        if rc != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return cnt, np.stack([pos >> 16, pos & 0xFFFF], axis=-1), key


def _node(name: str, clip: Clip, ref: Optional[Clip], kwargs) -> Clip:
    f = _Filter(name, clip, ref, kwargs)
    return Clip(f.out_fmt, f.out_w, f.out_h, f.out_n, f.get_frame, owner=f)


# --------------------------------------------------------------------------- the bm3d namespace

def RGB2OPP(input: Clip, sample: Optional[int] = None) -> Clip:
    return _node("RGB2OPP", input, None, {"sample": sample})


def OPP2RGB(input: Clip, sample: Optional[int] = None) -> Clip:
    return _node("OPP2RGB", input, None, {"sample": sample})


def Basic(input: Clip, ref: Optional[Clip] = None, profile: Optional[str] = None, sigma=None,
          block_size: Optional[int] = None, block_step: Optional[int] = None, group_size: Optional[int] = None,
          bm_range: Optional[int] = None, bm_step: Optional[int] = None, th_mse: Optional[float] = None,
          hard_thr: Optional[float] = None, matrix: Optional[int] = None, arith: Optional[int] = None) -> Clip:
    """arith (no counterpart in the reference): 0 = the x86 build's SSE2
    arithmetic, 1 = the scalar arithmetic of builds without SSE2; default:
    bm3d_host_arith()."""
    return _node("Basic", input, ref, dict(profile=profile, sigma=sigma, block_size=block_size,
                                           block_step=block_step, group_size=group_size, bm_range=bm_range,
                                           bm_step=bm_step, th_mse=th_mse, hard_thr=hard_thr, matrix=matrix,
                                           arith=arith))


def Final(input: Clip, ref: Clip, profile: Optional[str] = None, sigma=None, block_size: Optional[int] = None,
          block_step: Optional[int] = None, group_size: Optional[int] = None, bm_range: Optional[int] = None,
          bm_step: Optional[int] = None, th_mse: Optional[float] = None, matrix: Optional[int] = None,
          arith: Optional[int] = None) -> Clip:
    return _node("Final", input, ref, dict(profile=profile, sigma=sigma, block_size=block_size,
                                           block_step=block_step, group_size=group_size, bm_range=bm_range,
                                           bm_step=bm_step, th_mse=th_mse, matrix=matrix, arith=arith))


def VBasic(input: Clip, ref: Optional[Clip] = None, profile: Optional[str] = None, sigma=None,
           radius: Optional[int] = None, block_size: Optional[int] = None, block_step: Optional[int] = None,
           group_size: Optional[int] = None, bm_range: Optional[int] = None, bm_step: Optional[int] = None,
           ps_num: Optional[int] = None, ps_range: Optional[int] = None, ps_step: Optional[int] = None,
           th_mse: Optional[float] = None, hard_thr: Optional[float] = None, matrix: Optional[int] = None,
           arith: Optional[int] = None) -> Clip:
    return _node("VBasic", input, ref, dict(profile=profile, sigma=sigma, radius=radius, block_size=block_size,
                                            block_step=block_step, group_size=group_size, bm_range=bm_range,
                                            bm_step=bm_step, ps_num=ps_num, ps_range=ps_range, ps_step=ps_step,
                                            th_mse=th_mse, hard_thr=hard_thr, matrix=matrix, arith=arith))


def VFinal(input: Clip, ref: Clip, profile: Optional[str] = None, sigma=None, radius: Optional[int] = None,
           block_size: Optional[int] = None, block_step: Optional[int] = None, group_size: Optional[int] = None,
           bm_range: Optional[int] = None, bm_step: Optional[int] = None, ps_num: Optional[int] = None,
           ps_range: Optional[int] = None, ps_step: Optional[int] = None, th_mse: Optional[float] = None,
           matrix: Optional[int] = None, arith: Optional[int] = None) -> Clip:
    return _node("VFinal", input, ref, dict(profile=profile, sigma=sigma, radius=radius, block_size=block_size,
                                            block_step=block_step, group_size=group_size, bm_range=bm_range,
                                            bm_step=bm_step, ps_num=ps_num, ps_range=ps_range, ps_step=ps_step,
                                            th_mse=th_mse, matrix=matrix, arith=arith))


def VAggregate(input: Clip, radius: Optional[int] = None, sample: Optional[int] = None) -> Clip:
    return _node("VAggregate", input, None, {"radius": radius, "sample": sample})


# --------------------------------------------------------------------------- batched Basic -> Final

class Batch:
    """Basic then Final over many Gray float frames of uneven sizes, each
    with its own sigma, in one launch per stage (no counterpart in
    the reference, which filters frame by frame).  ``kwargs`` are the
    Basic/Final arguments shared by every frame (e.g. ``profile="np"``);
    ``pitches`` (elements) default to the widths.  Frames are float32 torch
    tensors on the GPU; every output equals Final(Basic(frame)) of the
    single-frame filters bit for bit."""

    def __init__(self, sizes: Sequence[tuple], sigmas: Sequence[float], pitches: Optional[Sequence[int]] = None,
                 **kwargs):
        L = init()
        self.L = L
        n = len(sizes)
        if n == 0 or len(sigmas) != n:
            raise Error("bm3d.Batch: one sigma per frame is required")
        self.sizes = [(int(h), int(w)) for h, w in sizes]
        self.pitches = [int(p) for p in pitches] if pitches is not None else [w for _, w in self.sizes]
        ws = (C.c_int * n)(*[w for _, w in self.sizes])
        hs = (C.c_int * n)(*[h for h, _ in self.sizes])
        ps = (C.c_int * n)(*self.pitches)
        sg = (C.c_double * n)(*[float(x) for x in sigmas])
        keep: list = []
        args, nargs = _make_args(kwargs, keep)
        handle = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = L.bm3d_batch_create(C.byref(handle), n, ws, hs, ps, sg, args, nargs, err, 1024)
        if rc != 0:
            raise Error(err.value.decode() or L.bm3d_last_error().decode())
        self.h = handle
        self.n = n

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.bm3d_batch_free(self.h)
                self.h = None
        except Exception:
            pass

    def _ptrs(self, src, dst):
        if len(src) != self.n or len(dst) != self.n:
            raise Error("bm3d.Batch: one source and one destination plane per frame")
        for k, (s, d) in enumerate(zip(src, dst)):
            h, w = self.sizes[k]
            for t in (s, d):
                if isinstance(t, np.ndarray) or not t.is_cuda or t.dtype.itemsize != 4 or t.dim() != 2:
                    raise Error("bm3d.Batch: planes must be 2-D float32 GPU tensors")
                if t.shape[0] != h or t.shape[1] != w or t.stride(1) != 1 or t.stride(0) != self.pitches[k]:
                    raise Error(f"bm3d.Batch: frame {k}: plane shape/pitch differs from the batch's")
        sp = (C.c_void_p * self.n)(*[s.data_ptr() for s in src])
        dp = (C.c_void_p * self.n)(*[d.data_ptr() for d in dst])
        _sync_producer(src[0].device)
        return sp, dp

    def run(self, src, dst, graph: bool = False) -> float:
        """Filters src[k] into dst[k]; returns the chain's GPU time in ms."""
        sp, dp = self._ptrs(src, dst)
        ms = C.c_float()
        if self.L.bm3d_batch_run(self.h, sp, dp, 1 if graph else 0, C.byref(ms)) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return ms.value

    def tune(self, src, dst, reps: int = 2) -> str:
        """Times the supported launch shapes and keeps the fastest; returns the report."""
        sp, dp = self._ptrs(src, dst)
        rep = C.create_string_buffer(8192)
        if self.L.bm3d_batch_tune(self.h, sp, dp, reps, rep, 8192) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return rep.value.decode()

    @property
    def shape(self) -> Dict[str, int]:
        s = BatchShape()
        if self.L.bm3d_batch_get_shape(self.h, C.byref(s)) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return dict(match_waves=s.match_waves, match_strip=s.match_strip, group_order=s.group_order)

    @shape.setter
    def shape(self, v: Dict[str, int]) -> None:
        s = BatchShape(int(v["match_waves"]), int(v["match_strip"]), int(v["group_order"]))
        if self.L.bm3d_batch_set_shape(self.h, C.byref(s)) != 0:
            raise Error(self.L.bm3d_last_error().decode())

    def info(self) -> Dict[str, int]:
        b, f, px = C.c_int64(), C.c_int64(), C.c_int64()
        if self.L.bm3d_batch_info(self.h, C.byref(b), C.byref(f), C.byref(px)) != 0:
            raise Error(self.L.bm3d_last_error().decode())
        return dict(basic_blocks=b.value, final_blocks=f.value, pixels=px.value)
