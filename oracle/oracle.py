"""ctypes wrapper around the CPU oracle (oracle/bm3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (the HIP library and its host
mirror) never imports this module.

Build with ``make -C oracle`` (done by ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbm3d_oracle.so")

MAX_GROUP = 256
ARITH_SSE2, ARITH_SCALAR = 0, 1
RCP_NATIVE, RCP_TABLE = 0, 1
TIES_LIBSTDCXX, TIES_INDEX = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("wiener", C.c_int),
        ("block_size", C.c_int), ("block_step", C.c_int), ("group_size", C.c_int),
        ("bm_range", C.c_int), ("bm_step", C.c_int),
        ("th_mse", C.c_double), ("lambda_", C.c_double),
        ("sigma", C.c_double * 3),
        ("process", C.c_int * 3),
        ("radius", C.c_int), ("ps_num", C.c_int), ("ps_range", C.c_int), ("ps_step", C.c_int),
    ]


class Filter(C.Structure):
    _fields_ = [
        ("p", Params),
        ("arith", C.c_int), ("rcp_mode", C.c_int), ("tie_mode", C.c_int),
        ("rcp_bits", C.c_int),
        ("rcp_table", C.POINTER(C.c_float)),
        ("final_amp", C.c_double * MAX_GROUP),
        ("thr", ((C.c_float * 4) * MAX_GROUP) * 3),
        ("wiener_sigma_sqr", (C.c_float * MAX_GROUP) * 3),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        L.or_profile_defaults.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(Params), C.c_double]
        L.or_matrix_norms.argtypes = [C.c_int, C.c_double * 3]
        L.or_default_matrix.argtypes = [C.c_int, C.c_int]
        L.or_filter_init.argtypes = [C.POINTER(Filter), C.POINTER(Params), C.c_int, C.c_int, fp, C.c_int, C.c_int]
        L.or_filter_init.restype = None
        L.or_ref_grid.argtypes = [C.c_int, C.c_int, C.c_int, ip]
        L.or_search_boundary.argtypes = [C.c_int] * 4
        L.or_ssd.argtypes = [fp, fp, C.c_ssize_t, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_ssd.restype = C.c_float
        L.or_block_match.argtypes = [C.POINTER(Filter), fp, C.c_int, C.c_int, C.c_ssize_t, C.c_int, C.c_int, fp, ip, ip]
        L.or_vblock_match.argtypes = [C.POINTER(Filter), C.POINTER(fp), C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_ssize_t, C.c_int, C.c_int, fp, ip, ip, ip]
        L.or_partial_sort_keys.argtypes = [fp, ip, C.c_int, C.c_int]
        L.or_partial_sort_keys.restype = None
        L.or_dct2_1d.argtypes = [fp, fp, C.c_int]
        L.or_dct3_1d.argtypes = [fp, fp, C.c_int]
        L.or_fdct3d.argtypes = [fp, C.c_int, C.c_int]
        L.or_idct3d.argtypes = [fp, C.c_int, C.c_int]
        for n in ("or_dct2_1d", "or_dct3_1d", "or_fdct3d", "or_idct3d"):
            getattr(L, n).restype = None
        L.or_rcp.argtypes = [C.POINTER(Filter), C.c_float]
        L.or_rcp.restype = C.c_float
        L.or_bm3d_kernel.argtypes = [C.POINTER(Filter), C.c_int, C.c_int, C.c_int, C.c_ssize_t,
                                     C.POINTER(fp), C.POINTER(fp), C.POINTER(fp)]
        L.or_bm3d_kernel_mt.argtypes = [C.POINTER(Filter), C.c_int, C.c_int, C.c_int, C.c_ssize_t,
                                     C.POINTER(fp), C.POINTER(fp), C.POINTER(fp), C.c_int]
        L.or_vbm3d_kernel.argtypes = [C.POINTER(Filter), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_ssize_t, C.POINTER(fp), C.POINTER(fp), C.POINTER(fp)]
        L.or_vaggregate_plane.argtypes = [C.c_int, C.c_int, C.c_ssize_t, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(fp), fp, C.c_ssize_t]
        L.or_vaggregate_plane.restype = None
        vp = C.c_void_p
        L.or_int2float.argtypes = [vp, C.c_int, C.c_ssize_t, fp, C.c_ssize_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_float2int.argtypes = [fp, C.c_ssize_t, vp, C.c_int, C.c_ssize_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_rgb2opp.argtypes = [C.POINTER(vp), C.c_int, C.c_ssize_t, C.POINTER(vp), C.c_int, C.c_ssize_t, C.c_int, C.c_int, C.c_int]
        L.or_rgb2opp_y.argtypes = [C.POINTER(vp), C.c_int, C.c_ssize_t, fp, C.c_ssize_t, C.c_int, C.c_int]
        L.or_opp2rgb.argtypes = [C.POINTER(vp), C.c_int, C.c_ssize_t, C.POINTER(vp), C.c_int, C.c_ssize_t, C.c_int, C.c_int, C.c_int]
        for n in ("or_int2float", "or_float2int", "or_rgb2opp", "or_rgb2opp_y", "or_opp2rgb"):
            getattr(L, n).restype = None
        L.or_chain_rgb48.argtypes = [C.POINTER(Filter), C.POINTER(Filter), C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(vp), C.POINTER(vp)]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


# --------------------------------------------------------------------------- params

def resolve(function, *, width, height, color_family, profile="fast", sigma=None, block_size=None,
            block_step=None, group_size=None, bm_range=None, bm_step=None, th_mse=None, hard_thr=None,
            matrix=None, radius=None, ps_num=None, ps_range=None, ps_step=None):
    """Resolve plugin arguments into oracle Params, restating
    BM3D_Data_Base::arguments_process / init_filter_data
    (source/BM3D_Base.cpp:32-283; VBM3D_Base.cpp:104-403; BM3D_Basic.cpp:32-65)
    for valid inputs.  color_family: 'gray' | 'yuv' | 'rgb'."""
    wiener = function in ("Final", "VFinal")
    temporal = function in ("VBasic", "VFinal")
    L = lib()
    sig = [10.0, 10.0, 10.0]
    if sigma is not None:
        s = list(sigma)[:3]
        sig = s + [s[-1]] * (3 - len(s))
    p = Params()
    if L.or_profile_defaults(profile.encode(), int(wiener), int(temporal), C.byref(p), sig[0]) != 0:
        raise ValueError("unknown profile")
    for name, val in (("block_size", block_size), ("block_step", block_step), ("group_size", group_size),
                      ("bm_range", bm_range), ("bm_step", bm_step), ("radius", radius), ("ps_num", ps_num),
                      ("ps_range", ps_range), ("ps_step", ps_step)):
        if val is not None:
            setattr(p, name, int(val))
    if th_mse is not None:
        p.th_mse = float(th_mse)
    if hard_thr is not None:
        p.lambda_ = float(hard_thr)
    if color_family == "rgb":
        mat = 100
    elif matrix is None or matrix == 2:
        mat = L.or_default_matrix(width, height)
    else:
        mat = int(matrix)
    norms = (C.c_double * 3)()
    L.or_matrix_norms(mat, norms)
    p.th_mse = p.th_mse * norms[0]
    for k in range(3):
        p.sigma[k] = sig[k] / 255.0 * norms[k]
        p.process[k] = 1 if (color_family == "rgb" or sig[k] != 0) else 0
    return p


def make_filter(params, arith=ARITH_SSE2, rcp_mode=RCP_NATIVE, rcp_table=None, rcp_bits=11,
                tie_mode=TIES_LIBSTDCXX):
    f = Filter()
    tbl = None
    if rcp_mode == RCP_TABLE:
        tbl = np.ascontiguousarray(rcp_table, dtype=np.float32)
        f._keep = tbl
    lib().or_filter_init(C.byref(f), C.byref(params), arith, rcp_mode,
                         _fp(tbl) if tbl is not None else None, rcp_bits, tie_mode)
    if tbl is not None:
        f._keep = tbl
    return f


# --------------------------------------------------------------------------- kernels

def bm3d_kernel(f, src, ref=None, nthreads=1):
    """src/ref: float32 arrays (P, H, W).  Returns float32 (P, H, W); planes
    that are not processed are returned as zeros.  nthreads > 1 spreads the
    matching and filtering of each band of block rows over threads (same
    result: the aggregation stays in scan order)."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    ref = src if ref is None else np.ascontiguousarray(ref, dtype=np.float32)
    P, H, W = src.shape
    dst = np.zeros_like(src)
    FP = C.POINTER(C.c_float)
    d = (FP * 3)(*[_fp(dst[k]) for k in range(P)] + [None] * (3 - P))
    s = (FP * 3)(*[_fp(src[k]) for k in range(P)] + [None] * (3 - P))
    r = (FP * 3)(*[_fp(ref[k]) for k in range(P)] + [None] * (3 - P))
    if nthreads > 1:
        lib().or_bm3d_kernel_mt(C.byref(f), P, W, H, W, d, s, r, int(nthreads))
    else:
        lib().or_bm3d_kernel(C.byref(f), P, W, H, W, d, s, r)
    return dst


def block_match(f, plane, j, i):
    plane = np.ascontiguousarray(plane, dtype=np.float32)
    H, W = plane.shape
    G = f.p.group_size
    keys = np.zeros(G, np.float32)
    ys = np.zeros(G, np.int32)
    xs = np.zeros(G, np.int32)
    n = lib().or_block_match(C.byref(f), _fp(plane), W, H, W, j, i, _fp(keys), _ip(ys), _ip(xs))
    return keys[:n].copy(), ys[:n].copy(), xs[:n].copy()


def vblock_match(f, planes, cur, j, i):
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    T, H, W = planes.shape
    G = f.p.group_size
    FP = C.POINTER(C.c_float)
    arr = (FP * T)(*[_fp(planes[k]) for k in range(T)])
    keys = np.zeros(G, np.float32)
    zs = np.zeros(G, np.int32)
    ys = np.zeros(G, np.int32)
    xs = np.zeros(G, np.int32)
    n = lib().or_vblock_match(C.byref(f), arr, T, cur, W, H, W, j, i, _fp(keys), _ip(zs), _ip(ys), _ip(xs))
    return keys[:n].copy(), zs[:n].copy(), ys[:n].copy(), xs[:n].copy()


def vbm3d_kernel(f, src, ref, cur):
    """src/ref: float32 (P, T, H, W) for the T = frames of the window around the
    output frame (index cur).  Returns the stacked (P, (2r+1)*2, H, W) buffer;
    slots of absent neighbours are left as NaN (the reference leaves them
    unwritten)."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    P, T, H, W = src.shape
    r = f.p.radius
    dst = np.full((P, (2 * r + 1) * 2, H, W), np.nan, np.float32)
    FP = C.POINTER(C.c_float)
    d = (FP * 3)(*[_fp(dst[k]) for k in range(P)] + [None] * (3 - P))
    s = (FP * (P * T))(*[_fp(src[p, t]) for p in range(P) for t in range(T)])
    rr = (FP * (P * T))(*[_fp(ref[p, t]) for p in range(P) for t in range(T)])
    lib().or_vbm3d_kernel(C.byref(f), P, T, cur, W, H, W, d, s, rr)
    return dst


def vaggregate_plane(stacks, radius, cur):
    """stacks: float32 (T, (2r+1)*2, H, W) of the T neighbours in frame order."""
    stacks = np.ascontiguousarray(stacks, dtype=np.float32)
    T, S, H, W = stacks.shape
    FP = C.POINTER(C.c_float)
    arr = (FP * T)(*[_fp(stacks[k]) for k in range(T)])
    dst = np.zeros((H, W), np.float32)
    lib().or_vaggregate_plane(W, H, W, radius, T, cur, arr, _fp(dst), W)
    return dst


# --------------------------------------------------------------------------- conversions

def _bits_of(a):
    if a.dtype == np.float32:
        return 0
    if a.dtype == np.uint8:
        return 8
    return 16


def int2float(plane, bits, chroma, full=True, clip=False):
    plane = np.ascontiguousarray(plane)
    H, W = plane.shape
    dst = np.zeros((H, W), np.float32)
    lib().or_int2float(plane.ctypes.data, bits, W, _fp(dst), W, W, H, int(chroma), int(full), int(clip))
    return dst


def float2int(plane, bits, chroma, full=True, clip=True):
    plane = np.ascontiguousarray(plane, dtype=np.float32)
    H, W = plane.shape
    dst = np.zeros((H, W), np.uint8 if bits <= 8 else np.uint16)
    lib().or_float2int(_fp(plane), W, dst.ctypes.data, bits, W, W, H, int(chroma), int(full), int(clip))
    return dst


def rgb2opp(rgb, src_bits, dst_bits, clip):
    rgb = np.ascontiguousarray(rgb)
    _, H, W = rgb.shape
    dt = np.float32 if dst_bits == 0 else (np.uint8 if dst_bits <= 8 else np.uint16)
    dst = np.zeros((3, H, W), dt)
    VP = C.c_void_p
    s = (VP * 3)(*[rgb[k].ctypes.data for k in range(3)])
    d = (VP * 3)(*[dst[k].ctypes.data for k in range(3)])
    lib().or_rgb2opp(s, src_bits, W, d, dst_bits, W, W, H, int(clip))
    return dst


def rgb2opp_y(rgb, src_bits):
    rgb = np.ascontiguousarray(rgb)
    _, H, W = rgb.shape
    dst = np.zeros((H, W), np.float32)
    VP = C.c_void_p
    s = (VP * 3)(*[rgb[k].ctypes.data for k in range(3)])
    lib().or_rgb2opp_y(s, src_bits, W, _fp(dst), W, W, H)
    return dst


def opp2rgb(yuv, src_bits, dst_bits, clip):
    yuv = np.ascontiguousarray(yuv)
    _, H, W = yuv.shape
    dt = np.float32 if dst_bits == 0 else (np.uint8 if dst_bits <= 8 else np.uint16)
    dst = np.zeros((3, H, W), dt)
    VP = C.c_void_p
    s = (VP * 3)(*[yuv[k].ctypes.data for k in range(3)])
    d = (VP * 3)(*[dst[k].ctypes.data for k in range(3)])
    lib().or_opp2rgb(s, src_bits, W, d, dst_bits, W, W, H, int(clip))
    return dst


def chain_rgb48(basic, final, frames, nthreads):
    """frames: uint16 (N, 3, H, W) RGB48.  Runs the CPU chain with nthreads."""
    frames = np.ascontiguousarray(frames, dtype=np.uint16)
    N, _, H, W = frames.shape
    out = np.zeros_like(frames)
    VP = C.c_void_p
    ins = (VP * (3 * N))(*[frames[n, p].ctypes.data for n in range(N) for p in range(3)])
    outs = (VP * (3 * N))(*[out[n, p].ctypes.data for n in range(N) for p in range(3)])
    lib().or_chain_rgb48(C.byref(basic), C.byref(final), N, nthreads, W, H, ins, outs)
    return out
