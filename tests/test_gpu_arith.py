"""GPU parity of the scalar arithmetic mode (arith=1): the reference built
without __SSE2__ (aarch64 and other non-x86 wheels) computes the SSD in plain
row-major order (include/Block.h:632-643) and the Wiener gain with an exact
division and one sequential L2 sum (source/BM3D_Final.cpp:106-112).  Every
output and match list must equal the oracle's OR_ARITH_SCALAR bit for bit, on
the 8x8 fast paths, the generic paths (block 4 and 11, group 64), V-BM3D, the
batched path and the learned matcher's exact re-rank.  A process whose host
probe reports no tabulable rcpps (BM3D_AMD_HOST_RCP=none) must use this mode
by default."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import synth
from oracle_chains import bm3d_frame, make
from test_gpu_vbm3d import _oracle_stack

pytestmark = pytest.mark.gpu

S = O.ARITH_SCALAR


def _same(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, what
    bad = got.view(np.uint32) != want.view(np.uint32) if got.dtype == np.float32 else got != want
    assert not bad.any(), f"{what}: {int(bad.sum())} samples differ, first at {np.argwhere(bad)[:3].tolist()}"


def _gray(gpu, img):
    return gpu.Clip.from_arrays(gpu.GRAYS, [img[None]])


@pytest.mark.parametrize("kw", [dict(profile="np"), dict(profile="fast"), dict(profile="np", block_size=4,
                                block_step=3, group_size=8, bm_range=5), dict(profile="vn"),
                                dict(profile="np", group_size=64, bm_range=6)])
def test_scalar_basic_final_gray(gpu, kw):
    h, w = 56, 72
    noisy, _ = synth.gray_float(h, w, 15.0, seed=4200)
    basic = gpu.Basic(_gray(gpu, noisy), arith=1, **kw)
    got_b = np.asarray(basic.get_frame(0).planes[0])
    fb = make("Basic", gpu.GRAYS, w, h, arith=S, **kw)
    ref_b = bm3d_frame(fb, gpu.GRAYS, noisy[None])
    _same(got_b, ref_b[0], f"Basic scalar {kw}")
    final = gpu.Final(_gray(gpu, noisy), _gray(gpu, ref_b[0]), arith=1, **kw)
    got_f = np.asarray(final.get_frame(0).planes[0])
    ff = make("Final", gpu.GRAYS, w, h, arith=S, **kw)
    _same(got_f, bm3d_frame(ff, gpu.GRAYS, noisy[None], ref_b)[0], f"Final scalar {kw}")
    for b in range(0, 40, 3):
        k_g, y_g, x_g = final.filter.matches(b)
        nx = len(range(0, w - ff.p.block_size, ff.p.block_step)) + 1
        j = min((b // nx) * ff.p.block_step, h - ff.p.block_size)
        i = min((b % nx) * ff.p.block_step, w - ff.p.block_size)
        k_o, y_o, x_o = O.block_match(ff, ref_b[0], j, i)
        assert np.array_equal(y_g, y_o) and np.array_equal(x_g, x_o), b
        assert np.array_equal(k_g.view(np.uint32), k_o.view(np.uint32)), b


def test_scalar_differs_from_sse2(gpu):
    """The two modes are different arithmetic: Final outputs differ somewhere."""
    h, w = 48, 64
    noisy, clean = synth.gray_float(h, w, 15.0, seed=4201)
    a = np.asarray(gpu.Final(_gray(gpu, noisy), _gray(gpu, clean), profile="np", arith=0).get_frame(0).planes[0])
    b = np.asarray(gpu.Final(_gray(gpu, noisy), _gray(gpu, clean), profile="np", arith=1).get_frame(0).planes[0])
    assert (a.view(np.uint32) != b.view(np.uint32)).any()


def test_scalar_yuv16_np_chain(gpu):
    h, w = 48, 56
    rgb, _ = synth.rgb16(h, w, 10.0, seed=4202)
    opp = O.rgb2opp(rgb, 16, 16, True)
    fmt = gpu.YUV444P16
    props = {"BM3D_OPP": 1}
    clip = gpu.Clip.from_arrays(fmt, [opp], props)
    b = gpu.Basic(clip, profile="np", matrix=100, arith=1)
    fb = make("Basic", fmt, w, h, profile="np", matrix=100, arith=S)
    ref_b = bm3d_frame(fb, fmt, opp)
    _same(np.stack(b.get_frame(0).planes), ref_b, "Basic YUV16 scalar")
    fin = gpu.Final(clip, b, profile="np", matrix=100, arith=1).get_frame(0)
    ff = make("Final", fmt, w, h, profile="np", matrix=100, arith=S)
    _same(np.stack(fin.planes), bm3d_frame(ff, fmt, opp, ref_b), "Final YUV16 scalar")


def test_scalar_vbm3d_stacks(gpu):
    T, H, W, radius = 4, 40, 56, 1
    noisy, clean = synth.moving_video(T, H, W, 15.0, seed=4203, planes=1)
    clip = gpu.Clip.from_arrays(gpu.GRAYS, list(noisy))
    guide = gpu.Clip.from_arrays(gpu.GRAYS, list(clean))
    for fn, node, ref in (("VBasic", gpu.VBasic(clip, radius=radius, arith=1), noisy),
                          ("VFinal", gpu.VFinal(clip, guide, radius=radius, arith=1), clean)):
        f = make(fn, gpu.GRAYS, W, H, radius=radius, arith=S)
        for n in range(T):
            want, s0, s1 = _oracle_stack(f, noisy, ref, n, radius)
            got = np.asarray(node.get_frame(n).planes[0]).reshape((2 * radius + 1) * 2, H, W)
            _same(got[s0:s1], want[0, s0:s1], f"{fn} scalar frame {n}")


def test_scalar_batch_equals_single_and_oracle(gpu):
    import torch
    frames = [(40, 56, 8.0), (33, 70, 20.0), (64, 48, 12.0)]
    hosts = [synth.gray_float(h, w, s, seed=4210 + k)[0] for k, (h, w, s) in enumerate(frames)]
    src = [torch.from_numpy(x).cuda() for x in hosts]
    dst = [torch.empty((h, w), dtype=torch.float32, device="cuda") for h, w, _ in frames]
    b = gpu.Batch([(h, w) for h, w, _ in frames], [s for *_, s in frames], profile="np", arith=1)
    b.run(src, dst)
    for k, (h, w, s) in enumerate(frames):
        basic = gpu.Basic(_gray(gpu, hosts[k]), profile="np", sigma=s, arith=1)
        single = np.asarray(gpu.Final(_gray(gpu, hosts[k]), basic, profile="np", sigma=s, arith=1)
                            .get_frame(0).planes[0])
        _same(dst[k].cpu().numpy(), single, f"batch frame {k}")
    h, w, s = frames[0]
    fb = make("Basic", gpu.GRAYS, w, h, profile="np", sigma=[s], arith=S)
    ff = make("Final", gpu.GRAYS, w, h, profile="np", sigma=[s], arith=S)
    rb = bm3d_frame(fb, gpu.GRAYS, hosts[0][None])
    _same(dst[0].cpu().numpy(), bm3d_frame(ff, gpu.GRAYS, hosts[0][None], rb)[0], "batch frame 0 vs oracle")


def test_scalar_learned_whole_window_equals_plain(gpu):
    import lbm
    h, w = 64, 96
    noisy, _ = synth.gray_float(h, w, 25.0, seed=4220)
    kw = dict(profile="fast", sigma=25.0, bm_range=3, group_size=16, arith=1)  # 48 candidates <= top-k 64
    plain = gpu.Basic(_gray(gpu, noisy), **kw)
    out_p = np.asarray(plain.get_frame(0).planes[0])
    learned = gpu.Basic(_gray(gpu, noisy), **kw)
    learned.filter.set_matcher(lbm.Model(seed=8), 64)
    out_l = np.asarray(learned.get_frame(0).planes[0])
    _same(out_l, out_p, "learned scalar vs plain scalar")
    f = make("Basic", gpu.GRAYS, w, h, profile="fast", sigma=[25.0], bm_range=3, group_size=16, arith=S)
    _same(out_p, bm3d_frame(f, gpu.GRAYS, noisy[None])[0], "plain scalar vs oracle")


def test_probe_failure_defaults_to_scalar(gpu):
    """A fresh process with BM3D_AMD_HOST_RCP=none: bm3d_init logs the fallback
    and a Final created without "arith" reproduces the scalar build."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = f"""
import sys
sys.path[:0] = [{os.path.join(root, 'vapoursynth-bm3d_amd')!r}, {os.path.join(root, 'oracle')!r},
                {os.path.join(root, 'tests')!r}]
import numpy as np, bm3d, synth, oracle as O
from oracle_chains import bm3d_frame, make
assert bm3d.host_arith() == 1
bm3d.init(0)
h, w = 40, 48
noisy, clean = synth.gray_float(h, w, 15.0, seed=4230)
c = lambda a: bm3d.Clip.from_arrays(bm3d.GRAYS, [a[None]])
got = np.asarray(bm3d.Final(c(noisy), c(clean), profile="np").get_frame(0).planes[0])
f = make("Final", bm3d.GRAYS, w, h, profile="np", arith=O.ARITH_SCALAR)
want = bm3d_frame(f, bm3d.GRAYS, noisy[None], clean[None])[0]
assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
print("scalar default ok")
"""
    env = dict(os.environ, BM3D_AMD_HOST_RCP="none")
    print("arith: subprocess with BM3D_AMD_HOST_RCP=none", flush=True)
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "scalar default ok" in out.stdout
    assert "scalar (non-SSE2) arithmetic" in out.stderr
