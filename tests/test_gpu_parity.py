"""GPU parity: the HIP library against the CPU oracle on seeded inputs.

Bar: bit-exact.  The oracle restates the reference arithmetic; the kernels
evaluate the same operations in the same order (SSE2 lane-order SSD, the
oracle's DCT chains, the host's rcpps, the reference's aggregation order), so
every output sample and every match list must be identical.  Equal-key ties
are resolved by search index on both sides (TIES_INDEX)."""
import numpy as np
import pytest

import oracle as O
import synth
from oracle_chains import bm3d_frame, make

pytestmark = pytest.mark.gpu


def _gray_clip(bm3d, img, props=None):
    return bm3d.Clip.from_arrays(bm3d.GRAYS, [img[None]], props)


def _assert_equal(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if a.dtype == np.float32:
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    else:
        bad = a != b
    n = int(bad.sum())
    if n:
        idx = np.argwhere(bad)[:5]
        diff = np.abs(a.astype(np.float64) - b.astype(np.float64)).max()
        raise AssertionError(f"{what}: {n} samples differ, max |diff| {diff}, first at {idx.tolist()}")


@pytest.mark.parametrize("profile,h,w", [("fast", 64, 80), ("np", 72, 88)])
def test_basic_gray_float_matches_and_output(gpu, profile, h, w):
    noisy, _ = synth.gray_float(h, w, 10.0, seed=1001)
    clip = _gray_clip(gpu, noisy)
    node = gpu.Basic(clip, profile=profile)
    out = node.get_frame(0).planes[0]
    f = make("Basic", gpu.GRAYS, w, h, profile=profile)
    # match lists of every reference block
    flt = node._owner
    gy = np.zeros(h + 1, np.int32)
    gx = np.zeros(w + 1, np.int32)
    ny = O.lib().or_ref_grid(h, f.p.block_size, f.p.block_step, O._ip(gy))
    nx = O.lib().or_ref_grid(w, f.p.block_size, f.p.block_step, O._ip(gx))
    for b in range(ny * nx):
        j, i = int(gy[b // nx]), int(gx[b % nx])
        k_ref, y_ref, x_ref = O.block_match(f, noisy, j, i)
        k_gpu, y_gpu, x_gpu = flt.matches(b)
        assert len(k_gpu) == len(k_ref), (b, j, i)
        assert np.array_equal(y_gpu, y_ref) and np.array_equal(x_gpu, x_ref), (b, j, i, y_gpu, y_ref)
        assert np.array_equal(k_gpu.view(np.uint32), k_ref.view(np.uint32)), (b, j, i)
    ref = bm3d_frame(f, gpu.GRAYS, noisy[None])[0]
    _assert_equal(out, ref, "Basic gray float")


def test_final_gray_float(gpu):
    h, w = 64, 72
    noisy, _ = synth.gray_float(h, w, 10.0, seed=1002)
    fb = make("Basic", gpu.GRAYS, w, h, profile="fast")
    basic = bm3d_frame(fb, gpu.GRAYS, noisy[None])
    node = gpu.Final(_gray_clip(gpu, noisy), _gray_clip(gpu, basic[0]), profile="fast")
    out = node.get_frame(0).planes[0]
    ff = make("Final", gpu.GRAYS, w, h, profile="fast")
    ref = bm3d_frame(ff, gpu.GRAYS, noisy[None], basic)[0]
    _assert_equal(out, ref, "Final gray float")


def test_yuv16_basic_final_np(gpu):
    h, w = 56, 64
    rgb, _ = synth.rgb16(h, w, 10.0, seed=2001)
    opp = O.rgb2opp(rgb, 16, 16, True)
    fmt = gpu.YUV444P16
    props = {"BM3D_OPP": 1, "_Matrix": 2}
    clip = gpu.Clip.from_arrays(fmt, [opp], props)
    b = gpu.Basic(clip, profile="np", matrix=100).get_frame(0)
    fb = make("Basic", fmt, w, h, profile="np", matrix=100)
    ref_b = bm3d_frame(fb, fmt, opp)
    _assert_equal(np.stack(b.planes), ref_b, "Basic YUV444P16")
    bclip = gpu.Clip.from_arrays(fmt, [ref_b], props)
    fin = gpu.Final(clip, bclip, profile="np", matrix=100).get_frame(0)
    ff = make("Final", fmt, w, h, profile="np", matrix=100)
    ref_f = bm3d_frame(ff, fmt, opp, ref_b)
    _assert_equal(np.stack(fin.planes), ref_f, "Final YUV444P16")


def test_rgb48_basic_internal_opp(gpu):
    h, w = 48, 56
    rgb, _ = synth.rgb16(h, w, 10.0, seed=2002)
    fmt = gpu.RGB48
    out = gpu.Basic(gpu.Clip.from_arrays(fmt, [rgb]), profile="fast").get_frame(0)
    f = make("Basic", fmt, w, h, profile="fast")
    ref = bm3d_frame(f, fmt, rgb)
    _assert_equal(np.stack(out.planes), ref, "Basic RGB48")


def test_rgb2opp_opp2rgb(gpu):
    h, w = 40, 50
    rgb, _ = synth.rgb16(h, w, 10.0, seed=2003)
    fmt = gpu.RGB48
    for sample in (0, 1):
        o = gpu.RGB2OPP(gpu.Clip.from_arrays(fmt, [rgb]), sample=sample).get_frame(0)
        ref = O.rgb2opp(rgb, 16, 0 if sample else 16, sample == 0)
        _assert_equal(np.stack(o.planes), ref, f"RGB2OPP sample={sample}")
        assert o.props["BM3D_OPP"] == 1 and o.props["_Matrix"] == 2
        yfmt = gpu.YUV444PS if sample else gpu.YUV444P16
        back = gpu.OPP2RGB(gpu.Clip.from_arrays(yfmt, [ref], o.props), sample=0).get_frame(0)
        ref2 = O.opp2rgb(ref, 0 if sample else 16, 16, True)
        _assert_equal(np.stack(back.planes), ref2, f"OPP2RGB from sample={sample}")
        assert "BM3D_OPP" not in back.props and back.props["_Matrix"] == 0


def test_config1_640x480_basic_fast_full_frame(gpu):
    """The reference's CPU-runnable workload: one Gray float 640x480 frame,
    Basic profile 'fast', sigma 10: every match list and every output sample."""
    h, w = 480, 640
    noisy, _ = synth.gray_float(h, w, 10.0, seed=1000)
    node = gpu.Basic(_gray_clip(gpu, noisy), profile="fast")
    out = node.get_frame(0).planes[0]
    f = make("Basic", gpu.GRAYS, w, h, profile="fast")
    ref = bm3d_frame(f, gpu.GRAYS, noisy[None])[0]
    _assert_equal(out, ref, "Basic 640x480 fast")
    flt = node.filter
    gy = np.zeros(h + 1, np.int32)
    gx = np.zeros(w + 1, np.int32)
    ny = O.lib().or_ref_grid(h, 8, 8, O._ip(gy))
    nx = O.lib().or_ref_grid(w, 8, 8, O._ip(gx))
    assert ny * nx == 4800
    for b in range(0, ny * nx, 7):
        j, i = int(gy[b // nx]), int(gx[b % nx])
        k_ref, y_ref, x_ref = O.block_match(f, noisy, j, i)
        k_gpu, y_gpu, x_gpu = flt.matches(b)
        assert np.array_equal(y_gpu, y_ref) and np.array_equal(x_gpu, x_ref), (b, j, i)
        assert np.array_equal(k_gpu.view(np.uint32), k_ref.view(np.uint32)), (b, j, i)


@pytest.mark.parametrize("kw", [
    dict(block_size=4, block_step=3, group_size=8, bm_range=5),      # B < 8: scalar SSD, generic kernels
    dict(block_size=11, block_step=6, group_size=16, bm_range=7),    # SIMD part + residue columns
    dict(group_size=1),                                              # no matching
    dict(group_size=64, bm_range=6),                                 # G > 32: generic transform path
    dict(bm_step=2, bm_range=8),                                     # sparse search lattice
])
def test_basic_final_parameter_variants(gpu, kw):
    h, w = 40, 48
    noisy, clean = synth.gray_float(h, w, 15.0, seed=77)
    for fn in ("Basic", "Final"):
        if fn == "Basic":
            node = gpu.Basic(_gray_clip(gpu, noisy), profile="np", **kw)
            f = make("Basic", gpu.GRAYS, w, h, profile="np", **kw)
            ref = bm3d_frame(f, gpu.GRAYS, noisy[None])[0]
        else:
            node = gpu.Final(_gray_clip(gpu, noisy), _gray_clip(gpu, clean), profile="np", **kw)
            f = make("Final", gpu.GRAYS, w, h, profile="np", **kw)
            ref = bm3d_frame(f, gpu.GRAYS, noisy[None], clean[None])[0]
        _assert_equal(node.get_frame(0).planes[0], ref, f"{fn} {kw}")


def test_concurrent_get_frame_on_one_filter(gpu):
    """VapourSynth calls GetFrame of an fmParallel filter from several threads
    for different frames with one shared instance; the library serialises
    calls per filter.  Outputs must equal a sequential run."""
    from concurrent.futures import ThreadPoolExecutor
    frames = [synth.gray_float(48, 64, 10.0, seed=1200 + k)[0] for k in range(8)]
    clip = gpu.Clip.from_arrays(gpu.GRAYS, [f[None] for f in frames])
    seq = gpu.Basic(clip, profile="np")
    want = [np.asarray(seq.get_frame(n).planes[0]) for n in range(8)]
    par = gpu.Basic(clip, profile="np")
    fin = gpu.Final(clip, par, profile="np")
    with ThreadPoolExecutor(4) as pool:
        got = list(pool.map(lambda n: np.asarray(par.get_frame(n).planes[0]), range(8)))
        fin_got = list(pool.map(lambda n: np.asarray(fin.get_frame(n).planes[0]), range(8)))
    for n in range(8):
        _assert_equal(got[n], want[n], f"threaded Basic frame {n}")
    fin_seq = gpu.Final(clip, seq, profile="np")
    for n in range(8):
        _assert_equal(fin_got[n], np.asarray(fin_seq.get_frame(n).planes[0]), f"threaded Final frame {n}")
