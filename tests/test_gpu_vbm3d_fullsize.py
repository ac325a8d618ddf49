"""V-BM3D at its full working size: 1920x1080 YUV444 float, the chain
VBasic(radius=3, ps_num=2, ps_range=4) -> VAggregate(sample=1) ->
VFinal(same) -> VAggregate(sample=1) on a moving-object video, sigma 15.

The oracle finishes one full-size output frame in seconds, so the centre
frame of a 7-frame clip (whose window is the whole clip) is compared bit for
bit at every stage: the VBasic stack, VAggregate of the GPU's stacks, and the
VFinal stack computed against the GPU's aggregated basic estimate (each stage
is pinned on its own inputs).  Sampled reference blocks' match lists of the
full frame are compared with the oracle's block matcher as well."""
import numpy as np
import pytest

import oracle as O
import synth
from oracle_chains import make

pytestmark = pytest.mark.gpu

T, H, W, R = 7, 1080, 1920, 3
KW = dict(radius=R, ps_num=2, ps_range=4, sigma=[15.0])


@pytest.fixture(scope="module")
def video():
    noisy, _ = synth.moving_video(T, H, W, 15.0, seed=3000, planes=3)
    return noisy


def _stack(clip, n):
    fr = clip.get_frame(n)
    return np.stack([np.asarray(p).reshape(2 * (2 * R + 1), H, W) for p in fr.planes])


def _assert_bits(a, b, what):
    bad = a.view(np.uint32) != b.view(np.uint32)
    assert not bad.any(), f"{what}: {int(bad.sum())} samples differ, first at {np.argwhere(bad)[:3].tolist()}"


def test_fullsize_vbasic_vaggregate_vfinal(gpu, video):
    clip = gpu.Clip.from_arrays(gpu.YUV444PS, list(video))
    vb = gpu.VBasic(clip, **KW)
    agg = gpu.VAggregate(vb, radius=R, sample=1)
    vf = gpu.VFinal(clip, agg, **KW)
    cur = 3
    win = np.ascontiguousarray(video.transpose(1, 0, 2, 3))
    # VBasic stack of the centre frame
    fb = make("VBasic", gpu.YUV444PS, W, H, **KW)
    _assert_bits(_stack(vb, cur), O.vbm3d_kernel(fb, win, win, cur), "VBasic stack")
    # sampled match lists of the same frame (frame, y, x and key of every member)
    flt = vb._owner
    rng = np.random.RandomState(3)
    gy = np.zeros(H + 1, np.int32)
    gx = np.zeros(W + 1, np.int32)
    ny = O.lib().or_ref_grid(H, fb.p.block_size, fb.p.block_step, O._ip(gy))
    nx = O.lib().or_ref_grid(W, fb.p.block_size, fb.p.block_step, O._ip(gx))
    for b in rng.choice(ny * nx, 64, replace=False):
        j, i = int(gy[b // nx]), int(gx[b % nx])
        want = O.vblock_match(fb, win[0], cur, j, i)
        got = flt.vmatches(int(b))
        for w_, g_ in zip(want, got):
            assert np.array_equal(np.asarray(w_).view(np.uint32), np.asarray(g_).view(np.uint32)), (b, j, i)
    # VAggregate of the GPU's stacks
    stacks = [_stack(vb, m) for m in range(T)]
    out = agg.get_frame(cur)
    for p in range(3):
        want = O.vaggregate_plane(np.stack([s[p] for s in stacks]), R, cur)
        _assert_bits(np.asarray(out.planes[p]), want, f"VAggregate plane {p}")
    # VFinal stack against the GPU's aggregated basic estimate
    ref = np.stack([np.stack([np.asarray(q) for q in agg.get_frame(m).planes]) for m in range(T)])
    ff = make("VFinal", gpu.YUV444PS, W, H, **KW)
    want = O.vbm3d_kernel(ff, win, np.ascontiguousarray(ref.transpose(1, 0, 2, 3)), cur)
    _assert_bits(_stack(vf, cur), want, "VFinal stack")
