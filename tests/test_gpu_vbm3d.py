"""GPU parity of V-BM3D (VBasic / VFinal stacks and VAggregate) against the
CPU oracle, bit-exact (ties resolved by search / concatenation index on
both sides)."""
import numpy as np
import pytest

import oracle as O
import synth
from oracle_chains import make

pytestmark = pytest.mark.gpu


def _oracle_stack(f, src, ref, n, radius):
    """src/ref: (T, P, H, W) whole clip; returns the stack of output frame n and
    the range of slots the reference writes."""
    T = src.shape[0]
    first, last = max(0, n - radius), min(T - 1, n + radius)
    win_s = np.ascontiguousarray(src[first:last + 1].transpose(1, 0, 2, 3))
    win_r = np.ascontiguousarray(ref[first:last + 1].transpose(1, 0, 2, 3))
    st = O.vbm3d_kernel(f, win_s, win_r, n - first)
    cur = n - first
    s0 = 2 * (radius - cur)
    return st, s0, s0 + 2 * (last - first + 1)


def _check_stacks(gpu_clip, f, src, ref, radius, planes):
    T = src.shape[0]
    H = src.shape[2]
    for n in range(T):
        fr = gpu_clip.get_frame(n)
        want, s0, s1 = _oracle_stack(f, src, ref, n, radius)
        for p in range(planes):
            got = np.asarray(fr.planes[p]).reshape((2 * radius + 1) * 2, H, -1)
            a, b = got[s0:s1], want[p, s0:s1]
            bad = a != b
            assert not bad.any(), (n, p, int(bad.sum()), float(np.abs(a - b).max()))
        assert fr.props["BM3D_V_radius"] == radius


@pytest.mark.parametrize("radius,profile", [(1, "fast"), (2, "lc")])
def test_vbasic_gray_stacks_and_vaggregate(gpu, radius, profile):
    T, H, W = 5, 48, 64
    noisy, _ = synth.moving_video(T, H, W, 15.0, seed=31, planes=1)
    clip = gpu.Clip.from_arrays(gpu.GRAYS, list(noisy))
    vb = gpu.VBasic(clip, profile=profile, radius=radius, ps_range=4)
    f = make("VBasic", gpu.GRAYS, W, H, profile=profile, radius=radius, ps_range=4)
    _check_stacks(vb, f, noisy, noisy, radius, 1)
    agg = gpu.VAggregate(vb, radius=radius, sample=1)
    for n in range(T):
        first, last = max(0, n - radius), min(T - 1, n + radius)
        stacks = np.stack([np.asarray(vb.get_frame(m).planes[0]).reshape(-1, H, W) for m in range(first, last + 1)])
        want = O.vaggregate_plane(stacks, radius, n - first)
        got = np.asarray(agg.get_frame(n).planes[0])
        assert np.array_equal(got, want), n
        assert "BM3D_V_radius" not in agg.get_frame(n).props


@pytest.mark.parametrize("shape", [
    dict(block_size=4, block_step=3, group_size=8, bm_range=5),     # B*B <= 64: per-frame member-list walk
    dict(block_size=11, block_step=6, group_size=8, bm_range=7),    # B*B > 64: the block walk
])
def test_vbasic_block_sizes(gpu, shape):
    T, H, W, radius = 4, 44, 60, 1
    noisy, _ = synth.moving_video(T, H, W, 15.0, seed=37, planes=1)
    clip = gpu.Clip.from_arrays(gpu.GRAYS, list(noisy))
    vb = gpu.VBasic(clip, radius=radius, ps_range=3, **shape)
    f = make("VBasic", gpu.GRAYS, W, H, radius=radius, ps_range=3, **shape)
    _check_stacks(vb, f, noisy, noisy, radius, 1)


def test_vfinal_yuv_float_and_vaggregate16(gpu):
    T, H, W, radius = 4, 40, 56, 1
    noisy, clean = synth.moving_video(T, H, W, 10.0, seed=32, planes=3)
    fmt = gpu.YUV444PS
    clip = gpu.Clip.from_arrays(fmt, list(noisy))
    guide = gpu.Clip.from_arrays(fmt, list(clean))
    vf = gpu.VFinal(clip, guide, profile="fast", radius=radius)
    f = make("VFinal", fmt, W, H, profile="fast", radius=radius)
    _check_stacks(vf, f, noisy, clean, radius, 3)
    agg = gpu.VAggregate(vf, radius=radius, sample=0)
    for n in range(T):
        first, last = max(0, n - radius), min(T - 1, n + radius)
        out = agg.get_frame(n)
        for p in range(3):
            stacks = np.stack([np.asarray(vf.get_frame(m).planes[p]).reshape(-1, H, W)
                               for m in range(first, last + 1)])
            fl = O.vaggregate_plane(stacks, radius, n - first)
            want = O.float2int(fl, 16, p > 0, True, True)
            assert np.array_equal(np.asarray(out.planes[p]), want), (n, p)


def test_vbasic_rgb_input_is_opp_stack(gpu):
    T, H, W, radius = 3, 32, 40, 1
    rgb = np.stack([synth.rgb16(H, W, 10.0, seed=40 + t)[0] for t in range(T)])
    clip = gpu.Clip.from_arrays(gpu.RGB48, list(rgb))
    vb = gpu.VBasic(clip, profile="fast", radius=radius)
    assert vb.format.color_family == gpu.YUV and vb.format.sample_type == gpu.FLOAT
    f = make("VBasic", gpu.RGB48, W, H, profile="fast", radius=radius)
    opp = np.stack([O.rgb2opp(rgb[t], 16, 0, False) for t in range(T)]).astype(np.float32)
    _check_stacks(vb, f, opp, opp, radius, 3)
    assert vb.get_frame(0).props["BM3D_OPP"] == 1
