"""GPU parity across the input formats, frame properties and parameter presets
the reference accepts, bit-exact against the oracle:

* 8- and 10-bit Gray and YUV444 input (Int2Float/Float2Int,
  include/Conversion.hpp:36-78);
* `_Range` (include/BM3D_Base.h:210-219): absent or 1 = full range, 0 =
  limited range (the reference's inversion of VapourSynth's convention);
* YUV420 with sigma=[s,0,0]: luma filtered, chroma passed through
  (source/BM3D_Base.cpp:230-250);
* the BM3D_OPP=1 warning for YUV input without matrix=100
  (include/BM3D_Base.h:202-207);
* Basic with a separate ref clip and explicit th_mse / hard_thr;
* profiles 'high' and 'vn' of Basic and Final (source/BM3D.cpp:33-130),
  and the V-BM3D profiles 'np', 'high' and 'vn' (source/VBM3D_Base.cpp:32-97)."""
import numpy as np
import pytest

import synth
from oracle_chains import bm3d_frame, make
from test_gpu_vbm3d import _check_stacks

pytestmark = pytest.mark.gpu


def _fmt(gpu, cf, bits, ssw=0, ssh=0):
    return gpu.VideoFormat(cf, gpu.INTEGER if bits else gpu.FLOAT, bits or 32, ssw, ssh)


def _quant(img, bits):
    m = (1 << bits) - 1
    q = np.clip(np.floor(img * m + 0.5), 0, m)
    return q.astype(np.uint8 if bits <= 8 else np.uint16)


def _planes(h, w, n, seed, sigma=10.0):
    return [synth.gray_float(h, w, sigma, seed=seed + k)[0] for k in range(n)]


def _same(got, want, what):
    got = np.stack([np.asarray(p) for p in got])
    want = np.asarray(want)
    assert got.shape == want.shape, what
    if got.dtype == np.float32:
        bad = got.view(np.uint32) != want.view(np.uint32)
    else:
        bad = got != want
    assert not bad.any(), f"{what}: {int(bad.sum())} samples differ, first at {np.argwhere(bad)[:3].tolist()}"


def _basic_final(gpu, fmt, src, props, kw_b, kw_f, full=True):
    H, W = src.shape[1:]
    clip = gpu.Clip.from_arrays(fmt, [src], props)
    basic = gpu.Basic(clip, **kw_b)
    b = basic.get_frame(0)
    fb = make("Basic", fmt, W, H, **kw_b)
    ref_b = bm3d_frame(fb, fmt, src, full=full)
    _same(b.planes, ref_b, f"Basic {kw_b}")
    final = gpu.Final(clip, basic, **kw_f)
    ff = make("Final", fmt, W, H, **kw_f)
    ref_f = bm3d_frame(ff, fmt, src, ref_b, full=full)
    _same(final.get_frame(0).planes, ref_f, f"Final {kw_f}")


@pytest.mark.parametrize("bits", [8, 10])
@pytest.mark.parametrize("cf", ["gray", "yuv"])
def test_integer_depths(gpu, bits, cf):
    h, w = 48, 64
    n = 1 if cf == "gray" else 3
    src = np.stack([_quant(p, bits) for p in _planes(h, w, n, 300 + bits)])
    fmt = _fmt(gpu, gpu.GRAY if cf == "gray" else gpu.YUV, bits)
    _basic_final(gpu, fmt, src, {}, dict(profile="np"), dict(profile="np"))


@pytest.mark.parametrize("rng_prop,full", [(0, False), (1, True)])
def test_range_property(gpu, rng_prop, full):
    h, w = 40, 56
    src = np.stack([_quant(p, 10) for p in _planes(h, w, 3, 320)])
    fmt = _fmt(gpu, gpu.YUV, 10)
    _basic_final(gpu, fmt, src, {"_Range": rng_prop}, dict(profile="fast"), dict(profile="fast"), full=full)


def test_yuv420_luma_only(gpu):
    h, w = 48, 64
    y = _quant(_planes(h, w, 1, 330)[0], 16)
    u = _quant(_planes(h // 2, w // 2, 1, 331)[0], 16)
    v = _quant(_planes(h // 2, w // 2, 1, 332)[0], 16)
    fmt = _fmt(gpu, gpu.YUV, 16, 1, 1)
    clip = gpu.Clip.from_frames(fmt, [gpu.Frame([y, u, v])])
    for fn in ("Basic", "Final"):
        if fn == "Basic":
            node = gpu.Basic(clip, profile="fast", sigma=[10.0, 0.0, 0.0])
            f = make("Basic", fmt, w, h, profile="fast", sigma=[10.0, 0.0, 0.0])
            want = bm3d_frame(f, fmt, y[None])[0]
        else:
            node = gpu.Final(clip, gpu.Basic(clip, profile="fast", sigma=[10.0, 0.0, 0.0]), profile="fast",
                             sigma=[10.0, 0.0, 0.0])
            fb = make("Basic", fmt, w, h, profile="fast", sigma=[10.0, 0.0, 0.0])
            ref = bm3d_frame(fb, fmt, y[None])
            f = make("Final", fmt, w, h, profile="fast", sigma=[10.0, 0.0, 0.0])
            want = bm3d_frame(f, fmt, y[None], ref)[0]
        out = node.get_frame(0)
        assert np.array_equal(np.asarray(out.planes[0]), want), fn
        assert np.array_equal(np.asarray(out.planes[1]), u) and np.array_equal(np.asarray(out.planes[2]), v), fn


def test_bm3d_opp_warning(gpu):
    h, w = 40, 48
    rgb, _ = synth.rgb16(h, w, 10.0, seed=340)
    import oracle as O
    opp = O.rgb2opp(rgb, 16, 16, True)
    fmt = gpu.YUV444P16
    props = {"BM3D_OPP": 1, "_Matrix": 2}
    logs = []
    gpu.set_log(lambda level, msg: logs.append((level, msg)))
    try:
        out = gpu.Basic(gpu.Clip.from_arrays(fmt, [opp], props), profile="fast").get_frame(0)
        assert any("BM3D_OPP=1" in m and "matrix=100" in m for _, m in logs), logs
        n_warn = len(logs)
        gpu.Basic(gpu.Clip.from_arrays(fmt, [opp], props), profile="fast", matrix=100).get_frame(0)
        assert len(logs) == n_warn, logs  # matrix=100 given: no warning
    finally:
        gpu.set_log(None)
    # only a warning: the frame is still filtered, full range (BM3D_OPP=1)
    f = make("Basic", fmt, w, h, profile="fast")
    _same(out.planes, bm3d_frame(f, fmt, opp, full=True), "Basic with BM3D_OPP=1")
    assert out.props["BM3D_OPP"] == 1


@pytest.mark.parametrize("cf", ["gray", "yuv16"])
def test_basic_separate_ref_and_explicit_thresholds(gpu, cf):
    h, w = 48, 56
    if cf == "gray":
        fmt = gpu.GRAYS
        src = np.stack(_planes(h, w, 1, 350, sigma=20.0))
        ref = np.stack(_planes(h, w, 1, 350, sigma=0.0))
    else:
        fmt = gpu.YUV444P16
        src = np.stack([_quant(p, 16) for p in _planes(h, w, 3, 360, sigma=20.0)])
        ref = np.stack([_quant(p, 16) for p in _planes(h, w, 3, 360, sigma=0.0)])
    kw = dict(profile="np", th_mse=1234.5, hard_thr=3.1, sigma=[20.0])
    node = gpu.Basic(gpu.Clip.from_arrays(fmt, [src]), gpu.Clip.from_arrays(fmt, [ref]), **kw)
    f = make("Basic", fmt, w, h, **kw)
    _same(node.get_frame(0).planes, bm3d_frame(f, fmt, src, ref), f"Basic ref {cf}")
    kwf = dict(profile="np", th_mse=321.0, sigma=[20.0])
    node = gpu.Final(gpu.Clip.from_arrays(fmt, [src]), gpu.Clip.from_arrays(fmt, [ref]), **kwf)
    f = make("Final", fmt, w, h, **kwf)
    _same(node.get_frame(0).planes, bm3d_frame(f, fmt, src, ref), f"Final th_mse {cf}")


@pytest.mark.parametrize("profile", ["high", "vn"])
@pytest.mark.parametrize("cf", ["gray", "yuvs"])
def test_profiles_high_vn(gpu, profile, cf):
    h, w = 52, 60
    fmt = gpu.GRAYS if cf == "gray" else gpu.YUV444PS
    n = 1 if cf == "gray" else 3
    src = np.stack(_planes(h, w, n, 370, sigma=15.0))
    kw = dict(profile=profile, sigma=[15.0])
    _basic_final(gpu, fmt, src, {}, kw, kw)


@pytest.mark.parametrize("profile", ["np", "high", "vn"])
def test_vprofiles_np_high_vn(gpu, profile):
    T, H, W = 5, 40, 52
    noisy, clean = synth.moving_video(T, H, W, 15.0, seed=380, planes=1)
    clip = gpu.Clip.from_arrays(gpu.GRAYS, list(noisy))
    vb = gpu.VBasic(clip, profile=profile)
    radius = {"np": 3, "high": 4, "vn": 4}[profile]
    f = make("VBasic", gpu.GRAYS, W, H, profile=profile)
    assert f.p.radius == radius
    _check_stacks(vb, f, noisy, noisy, radius, 1)
    guide = gpu.Clip.from_arrays(gpu.GRAYS, list(clean))
    vf = gpu.VFinal(clip, guide, profile=profile)
    f = make("VFinal", gpu.GRAYS, W, H, profile=profile)
    _check_stacks(vf, f, noisy, clean, radius, 1)
