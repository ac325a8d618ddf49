"""Frame-level compositions of the CPU oracle that mirror the reference's
process_core_* dispatch (source/BM3D_Base.cpp:497-754) for the formats the
tests use.  Test infrastructure only."""
from __future__ import annotations

import numpy as np

import oracle as O


def family_name(fmt):
    return {1: "gray", 2: "rgb", 3: "yuv"}[fmt.color_family]


def make(function, fmt, width, height, tie_mode=O.TIES_INDEX, arith=O.ARITH_SSE2, **kw):
    p = O.resolve(function, width=width, height=height, color_family=family_name(fmt), **kw)
    return O.make_filter(p, arith=arith, tie_mode=tie_mode)


def bm3d_frame(f, fmt, src, ref=None, full=True):
    """One Basic/Final frame on the oracle.  src/ref: (P, H, W) arrays in the
    clip's sample format.  Returns the output frame in the same format
    (unprocessed planes copied from src)."""
    P = src.shape[0]
    cf = fmt.color_family
    bits = 0 if fmt.sample_type == 1 else fmt.bits_per_sample
    proc = [f.p.process[k] for k in range(3)]
    gray_path = cf == 1 or (cf == 3 and not proc[1] and not proc[2])
    if cf == 2:  # RGB: OPP float planes (process_core_rgb)
        s = O.rgb2opp(src, bits, 0, False).astype(np.float32)
        if ref is None:
            r = s
        elif f.p.wiener:
            r = O.rgb2opp(ref, bits, 0, False).astype(np.float32)
        else:
            r = np.zeros_like(s)
            r[0] = O.rgb2opp_y(ref, bits)
        d = O.bm3d_kernel(f, s, r)
        return O.opp2rgb(d, 0, bits, bits != 0)
    if gray_path:
        use = [0]
    else:
        use = [k for k in range(3) if proc[k]]
    H, W = src.shape[1:]
    s = np.zeros((3 if not gray_path else 1, H, W), np.float32)
    r = np.zeros_like(s)
    for k in range(s.shape[0]):
        s[k] = src[k] if bits == 0 else O.int2float(src[k], bits, k > 0, full, False)
        rr = src if ref is None else ref
        r[k] = rr[k] if bits == 0 else O.int2float(rr[k], bits, k > 0, full, False)
    if gray_path:
        f.p.process[1] = 0
        f.p.process[2] = 0
    d = O.bm3d_kernel(f, s, r)
    out = src.copy()
    for k in use:
        out[k] = d[k] if bits == 0 else O.float2int(d[k], bits, k > 0, full, True)
    return out
