"""Full-size GPU parity for the benchmark workloads.

* 1920x1080 RGB48 through RGB2OPP -> Basic(np, matrix=100) -> Final(np,
  matrix=100) -> OPP2RGB: bit-exact against the oracle's chain on the same
  frame (every 16-bit sample), plus sampled match lists (positions and keys)
  of both stages.  At this size the matcher runs full 33x33 windows and the
  filter buckets ~229k groups over many chunks.
* The 64-frame batched workload (Gray float, 256x256 .. 2048x1152, sigma
  5..30 per frame, np): every batched frame equals the single-frame filters,
  the two extreme frames equal the oracle chain, and HIP-graph replays equal
  the eager run.
* 3840x2160 Gray float, Basic(fast, sigma 25, bm_range 24, group_size 16)
  with the trained learned matcher: sampled lists are exact re-ranks (keys
  are the SSE2-order SSD keys of their positions, sorted, self first), and the
  PSNR loss against the plain GPU output stays within 0.05 dB.

The oracle runs frame-parallel over bands of block rows (or_bm3d_kernel_mt,
identical results) so each case finishes in seconds on the host cores."""
import os

import numpy as np
import pytest

import oracle as O
import synth
from oracle_chains import make

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _grid(dim, B, s):
    last = dim - B
    out, j = [], 0
    while True:
        if j >= last + s:
            break
        out.append(min(j, last))
        j += s
    return out


def _sample_blocks(ny, nx, count, seed):
    rng = np.random.RandomState(seed)
    corners = [0, nx - 1, (ny - 1) * nx, ny * nx - 1, (ny // 2) * nx + nx // 2]
    return sorted(set(corners + rng.randint(0, ny * nx, count).tolist()))


def _check_lists(flt, f, plane, H, W, count, seed, what):
    B, s = f.p.block_size, f.p.block_step
    gy, gx = _grid(H, B, s), _grid(W, B, s)
    for b in _sample_blocks(len(gy), len(gx), count, seed):
        j, i = gy[b // len(gx)], gx[b % len(gx)]
        k_ref, y_ref, x_ref = O.block_match(f, plane, j, i)
        k_gpu, y_gpu, x_gpu = flt.matches(b)
        assert len(k_gpu) == len(k_ref), (what, b, j, i)
        assert np.array_equal(y_gpu, y_ref) and np.array_equal(x_gpu, x_ref), (what, b, j, i)
        assert np.array_equal(k_gpu.view(np.uint32), k_ref.view(np.uint32)), (what, b, j, i)


def _oracle_rgb48_chain(rgb, fb, ff):
    """The oracle's RGB48 np chain (or_chain_rgb48's steps, threaded per
    frame): returns the RGB48 output and the float luma planes both matching
    stages used."""
    opp = O.rgb2opp(rgb, 16, 16, True)
    s = np.stack([O.int2float(opp[p], 16, p > 0, True, False) for p in range(3)])
    d = O.bm3d_kernel(fb, s, s, nthreads=THREADS)
    bas = np.stack([O.float2int(d[p], 16, p > 0, True, True) for p in range(3)])
    r = np.stack([O.int2float(bas[p], 16, p > 0, True, False) for p in range(3)])
    d = O.bm3d_kernel(ff, s, r, nthreads=THREADS)
    fin = np.stack([O.float2int(d[p], 16, p > 0, True, True) for p in range(3)])
    return O.opp2rgb(fin, 16, 16, True), s[0], r[0]


def test_rgb48_np_chain_1080p_bit_exact(gpu):
    import torch
    H, W = 1080, 1920
    print(f"fullsize: RGB48 np chain {W}x{H}, oracle threads {THREADS}", flush=True)
    rgb, _ = synth.rgb16(H, W, 10.0, seed=2000)
    dev = [torch.from_numpy(rgb.view(np.int16)).cuda().view(torch.uint16)]
    src = gpu.Clip.from_arrays(gpu.RGB48, dev)
    opp = gpu.RGB2OPP(src)
    basic = gpu.Basic(opp, profile="np", sigma=[10.0] * 3, matrix=100)
    final = gpu.Final(opp, basic, profile="np", sigma=[10.0] * 3, matrix=100)
    out = gpu.OPP2RGB(final)
    got = np.stack([p.cpu().numpy() for p in out.get_frame(0).planes])
    fb = make("Basic", gpu.YUV444P16, W, H, profile="np", sigma=[10.0] * 3, matrix=100)
    ff = make("Final", gpu.YUV444P16, W, H, profile="np", sigma=[10.0] * 3, matrix=100)
    want, luma_b, luma_f = _oracle_rgb48_chain(rgb, fb, ff)
    bad = got != want
    assert not bad.any(), f"{int(bad.sum())} samples differ, first at {np.argwhere(bad)[:3].tolist()}"
    st_b, st_f = basic.filter.stats(), final.filter.stats()
    assert st_b["ref_blocks"] == 128851 and st_f["ref_blocks"] == 229401
    _check_lists(basic.filter, fb, luma_b, H, W, 300, 1, "Basic")
    _check_lists(final.filter, ff, luma_f, H, W, 300, 2, "Final")


def _gray_chain_oracle(fmt, img, sigma):
    H, W = img.shape
    fb = make("Basic", fmt, W, H, profile="np", sigma=[sigma])
    ff = make("Final", fmt, W, H, profile="np", sigma=[sigma])
    b = O.bm3d_kernel(fb, img[None], nthreads=THREADS)
    return O.bm3d_kernel(ff, img[None], b, nthreads=THREADS)[0]


def test_batched_64_uneven_frames_fullsize(gpu):
    import torch
    from bench import batch_workload
    sizes, sigmas = batch_workload()
    print(f"fullsize: batched {len(sizes)} frames", flush=True)
    hosts, src, dst = [], [], []
    for k, ((h, w), s) in enumerate(zip(sizes, sigmas)):
        img = synth.gray_float(h, w, s, seed=6400 + k)[0]
        hosts.append(img)
        src.append(torch.from_numpy(img).cuda())
        dst.append(torch.empty((h, w), dtype=torch.float32, device="cuda"))
    b = gpu.Batch(sizes, sigmas, profile="np")
    b.run(src, dst)
    eager = [d.clone() for d in dst]
    for _ in range(2):
        for d in dst:
            d.fill_(-1.0)
        b.run(src, dst, graph=True)
        torch.cuda.synchronize()
        for k in range(len(dst)):
            assert torch.equal(dst[k].view(torch.int32), eager[k].view(torch.int32)), ("graph replay", k)
    for k, ((h, w), s) in enumerate(zip(sizes, sigmas)):
        clip = gpu.Clip.from_arrays(gpu.GRAYS, [src[k][None]])
        basic = gpu.Basic(clip, profile="np", sigma=s)
        single = gpu.Final(clip, basic, profile="np", sigma=s).get_frame(0).planes[0]
        assert torch.equal(single.view(torch.int32), eager[k].view(torch.int32)), ("single frame", k, h, w, s)
    for k in (0, 1):  # 256x256 and 2048x1152
        want = _gray_chain_oracle(gpu.GRAYS, hosts[k], sigmas[k])
        got = eager[k].cpu().numpy()
        bad = got.view(np.uint32) != want.view(np.uint32)
        assert not bad.any(), (k, sizes[k], int(bad.sum()), np.argwhere(bad)[:3].tolist())


def test_learned_4k_lists_are_exact_reranks_and_psnr(gpu):
    import lbm
    H, W = 2160, 3840
    kw = dict(profile="fast", sigma=25.0, bm_range=24, group_size=16)
    weights = os.path.join(os.path.dirname(lbm.__file__), "lbm_weights.npz")
    print("fullsize: learned matching 4K", flush=True)
    noisy, clean = synth.gray_float(H, W, 25.0, seed=5999)
    clip = gpu.Clip.from_arrays(gpu.GRAYS, [noisy[None]])
    plain = gpu.Basic(clip, **kw)
    out_p = np.asarray(plain.get_frame(0).planes[0])
    learned = gpu.Basic(clip, **kw)
    learned.filter.set_matcher(lbm.Model.load(weights), 32)
    out_l = np.asarray(learned.get_frame(0).planes[0])
    cp, pp, _ = plain.filter.match_lists()
    cl, pl, kl = learned.filter.match_lists()
    recall = lbm.recall(cp, pp, cl, pl)
    loss = lbm.psnr(out_p, clean) - lbm.psnr(out_l, clean)
    print(f"recall {recall:.4f}, PSNR loss {loss:.4f} dB, fallbacks {learned.filter.matcher_fallbacks()}", flush=True)
    assert loss <= 0.05, loss
    dist_mul = 255.0 * 255.0 / 64.0
    rng = np.random.RandomState(3)
    img = np.ascontiguousarray(noisy)
    for b in rng.randint(0, len(cl), 400):
        n = int(cl[b])
        j, i = (int(v) for v in pl[b, 0])
        assert kl[b, 0] == 0.0
        assert np.all(np.diff(kl[b, 1:n]) >= 0), b
        rb = np.ascontiguousarray(img[j:j + 8, i:i + 8])
        for z in range(1, n):
            y, x = (int(v) for v in pl[b, z])
            assert (y, x) != (j, i) and abs(y - j) <= 24 and abs(x - i) <= 24, (b, z)
            d = O.lib().or_ssd(O._fp(rb), O._fp(img), W, y, x, 8, O.ARITH_SSE2)
            assert kl[b, z] == np.float32(np.float64(d) * dist_mul), (b, z)
