"""bm3d_get_frame_noalloc (the allocation-free Basic/Final entry point) against
bm3d_get_frame, through the C ABI: bit-exact equality on two calls in a row,
the second with a smaller frame, plus host-memory frames and the size limit.
The fixed scratch area is shared and keeps the previous call's contents, so
the second, smaller call also checks that nothing stale is read."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def _dev(gpu, arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _gray_clip(gpu, img, device=True):
    import torch
    a = torch.from_numpy(img[None].copy()).cuda() if device else img[None].copy()
    return gpu.Clip.from_arrays(gpu.GRAYS, [a])


def _host(planes):
    return [p.cpu().numpy() if hasattr(p, "cpu") else np.asarray(p) for p in planes]


def _same(a, b, what):
    for k, (x, y) in enumerate(zip(_host(a), _host(b))):
        assert x.shape == y.shape, what
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{what}: plane {k} differs"


def _gray_chain(gpu, h, w, seed, device=True):
    noisy, _ = synth.gray_float(h, w, 10.0, seed=seed)
    clip = _gray_clip(gpu, noisy, device)
    basic = gpu.Basic(clip, profile="np")
    final = gpu.Final(clip, basic, profile="np")
    return basic, final


def test_noalloc_equals_get_frame_two_calls_second_smaller(gpu):
    L = gpu.load_library()
    assert L.bm3d_noalloc_scratch_bytes() == 64 << 20
    # first call: a larger Gray float np Basic -> Final on device frames
    print("noalloc: gray np 120x136, then 56x72", flush=True)
    b1, f1 = _gray_chain(gpu, 120, 136, seed=4101)
    ref_b1 = b1.get_frame(0).planes
    got_b1 = b1.filter.get_frame(0, noalloc=True).planes
    _same(got_b1, ref_b1, "Basic 120x136")
    got_f1 = f1.filter.get_frame(0, noalloc=True).planes
    ref_f1 = f1.filter.get_frame(0).planes
    _same(got_f1, ref_f1, "Final 120x136")
    # second call: a smaller frame (the area still holds the first call's data)
    b2, f2 = _gray_chain(gpu, 56, 72, seed=4102)
    got_f2 = f2.filter.get_frame(0, noalloc=True).planes
    ref_f2 = f2.filter.get_frame(0).planes
    _same(got_f2, ref_f2, "Final 56x72")
    got_b2 = b2.filter.get_frame(0, noalloc=True).planes
    _same(got_b2, b2.get_frame(0).planes, "Basic 56x72")


def test_noalloc_rgb48_and_host_frames(gpu):
    print("noalloc: RGB48 np 64x80 host frames", flush=True)
    rgb, _ = synth.rgb16(64, 80, 10.0, seed=4103)
    clip = gpu.Clip.from_arrays(gpu.RGB48, [rgb])
    basic = gpu.Basic(clip, profile="np")
    final = gpu.Final(clip, basic, profile="np")
    want_b = basic.get_frame(0).planes
    _same(basic.filter.get_frame(0, noalloc=True).planes, want_b, "RGB48 Basic (host frames)")
    want_f = final.filter.get_frame(0).planes
    _same(final.filter.get_frame(0, noalloc=True).planes, want_f, "RGB48 Final (host frames)")


def test_noalloc_rejects_frames_beyond_the_area(gpu):
    print("noalloc: 1024x1024 np Final must be rejected", flush=True)
    b, f = _gray_chain(gpu, 1024, 1024, seed=4104)
    with pytest.raises(gpu.Error, match="bytes of scratch"):
        f.filter.get_frame(0, noalloc=True)
    # the rejected call ran nothing; the filter still works through both entries
    b2, f2 = _gray_chain(gpu, 48, 64, seed=4105)
    _same(f2.filter.get_frame(0, noalloc=True).planes, f2.filter.get_frame(0).planes, "Final 48x64 after rejection")


def test_noalloc_rejects_other_filters(gpu):
    img, _ = synth.gray_float(32, 48, 10.0, seed=4106)
    v = gpu.VBasic(gpu.Clip.from_arrays(gpu.GRAYS, [img[None]] * 3))
    with pytest.raises(gpu.Error, match="only Basic and Final"):
        v.filter.get_frame(1, noalloc=True)
