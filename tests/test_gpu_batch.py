"""GPU tests of the batched Basic -> Final path (many Gray float frames of
uneven sizes, one sigma per frame, one launch per stage over all frames).

Bar: bit-exact.  Every batched output must equal Final(Basic(frame)) of the
single-frame filters, and on small frames the oracle's chain; HIP-graph
replays (with unrelated eager GPU work and synchronisations between them)
must equal the eager run each time, also after the caller's planes move."""
import numpy as np
import pytest

import synth
from oracle_chains import bm3d_frame, make

pytestmark = pytest.mark.gpu

# (height, width, sigma): uneven sizes, some narrower than the search window
FRAMES = [(40, 56, 5.0), (64, 64, 12.5), (33, 90, 30.0), (72, 48, 8.0), (96, 128, 20.0), (24, 200, 15.0),
          (17, 23, 25.0)]


def _inputs(frames, seed0, pad=0):
    import torch
    hosts, src = [], []
    for k, (h, w, s) in enumerate(frames):
        noisy, _ = synth.gray_float(h, w, s, seed=seed0 + k)
        hosts.append(noisy)
        t = torch.zeros((h, w + pad), dtype=torch.float32, device="cuda")[:, :w]
        t.copy_(torch.from_numpy(noisy))
        src.append(t)
    return hosts, src


def _outputs(frames, pad=0):
    import torch
    return [torch.full((h, w + pad), -1.0, dtype=torch.float32, device="cuda")[:, :w] for h, w, _ in frames]


def _single(bm3d, noisy, sigma, **kw):
    clip = bm3d.Clip.from_arrays(bm3d.GRAYS, [noisy[None]])
    basic = bm3d.Basic(clip, sigma=sigma, **kw).get_frame(0).planes[0]
    bclip = bm3d.Clip.from_arrays(bm3d.GRAYS, [np.asarray(basic)[None]])
    return np.asarray(bm3d.Final(clip, bclip, sigma=sigma, **kw).get_frame(0).planes[0])


def _bits_equal(a, b, what):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape, what
    bad = a.view(np.uint32) != b.view(np.uint32)
    assert not bad.any(), f"{what}: {int(bad.sum())} samples differ, first at {np.argwhere(bad)[:3].tolist()}"


def _host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("profile", ["np", "fast"])
def test_batch_equals_single_frame_filters(gpu, profile):
    import torch
    hosts, src = _inputs(FRAMES, 7000, pad=13)
    pitches = [w + 13 for _, w, _ in FRAMES]
    b = gpu.Batch([(h, w) for h, w, _ in FRAMES], [s for *_, s in FRAMES], pitches=pitches, profile=profile)
    # destinations share the source pitch
    dst = [torch.full((h, w + 13), -1.0, dtype=torch.float32, device="cuda")[:, :w] for h, w, _ in FRAMES]
    b.run(src, dst)
    for k, (h, w, s) in enumerate(FRAMES):
        _bits_equal(_host(dst[k]), _single(gpu, hosts[k], s, profile=profile), f"frame {k} {h}x{w} sigma {s}")


def test_batch_matches_oracle_chain(gpu):
    frames = [(40, 56, 5.0), (33, 45, 22.0)]
    hosts, src = _inputs(frames, 7100)
    dst = _outputs(frames)
    b = gpu.Batch([(h, w) for h, w, _ in frames], [s for *_, s in frames], profile="np")
    b.run(src, dst)
    for k, (h, w, s) in enumerate(frames):
        fb = make("Basic", gpu.GRAYS, w, h, profile="np", sigma=[s])
        basic = bm3d_frame(fb, gpu.GRAYS, hosts[k][None])
        ff = make("Final", gpu.GRAYS, w, h, profile="np", sigma=[s])
        ref = bm3d_frame(ff, gpu.GRAYS, hosts[k][None], basic)[0]
        _bits_equal(_host(dst[k]), ref, f"frame {k} vs oracle")


def test_graph_replays_equal_eager(gpu):
    import torch
    hosts, src = _inputs(FRAMES, 7200)
    b = gpu.Batch([(h, w) for h, w, _ in FRAMES], [s for *_, s in FRAMES], profile="np")
    eager = _outputs(FRAMES)
    b.run(src, eager)
    want = [_host(t) for t in eager]
    dst = _outputs(FRAMES)
    scratch = torch.randn(1 << 20, device="cuda")
    for rep in range(3):
        for t in dst:
            t.fill_(float(rep + 7))
        scratch.mul_(1.5).add_(1.0)  # unrelated eager work between replays
        torch.cuda.synchronize()
        b.run(src, dst, graph=True)
        for k in range(len(FRAMES)):
            _bits_equal(_host(dst[k]), want[k], f"replay {rep} frame {k}")
    # the caller's planes move: the argument blob is rewritten, the graph kept
    dst2 = _outputs(FRAMES)
    src2 = [t.clone() for t in src]
    b.run(src2, dst2, graph=True)
    for k in range(len(FRAMES)):
        _bits_equal(_host(dst2[k]), want[k], f"moved planes frame {k}")


def test_autotune_keeps_results(gpu):
    hosts, src = _inputs(FRAMES, 7300)
    b = gpu.Batch([(h, w) for h, w, _ in FRAMES], [s for *_, s in FRAMES], profile="np")
    base = _outputs(FRAMES)
    b.run(src, base)
    want = [_host(t) for t in base]
    rep = b.tune(src, _outputs(FRAMES), reps=1)
    assert "best" in rep
    lines = [l for l in rep.splitlines() if l.startswith("match_waves")]
    assert len(lines) >= 6, rep
    for shape in ({"match_waves": 1, "match_strip": 8, "group_order": 0},
                  {"match_waves": 4, "match_strip": 11, "group_order": 1},
                  {"match_waves": 2, "match_strip": 11, "group_order": 0},
                  {"match_waves": 2, "match_strip": 8, "group_order": 1}):
        b.shape = shape
        assert b.shape == shape
        out = _outputs(FRAMES)
        b.run(src, out, graph=True)
        for k in range(len(FRAMES)):
            _bits_equal(_host(out[k]), want[k], f"shape {shape} frame {k}")


@pytest.mark.parametrize("order", [0, 1])
def test_large_frames_both_group_orders(gpu, order):
    # frames of more than one 4096-group chunk: the bucketing takes both its
    # single-frame-chunk and its frame-spanning-chunk paths
    frames = [(256, 256, 9.0), (40, 56, 21.0), (200, 330, 14.0), (24, 40, 6.0)]
    hosts, src = _inputs(frames, 7500)
    b = gpu.Batch([(h, w) for h, w, _ in frames], [s for *_, s in frames], profile="np")
    for nw in (1, 4):
        b.shape = {"match_waves": nw, "match_strip": 11, "group_order": order}
        dst = _outputs(frames)
        b.run(src, dst, graph=True)
        for k, (h, w, s) in enumerate(frames):
            _bits_equal(_host(dst[k]), _single(gpu, hosts[k], s, profile="np"), f"order {order} waves {nw} frame {k}")


def test_per_frame_fallback_shapes(gpu):
    # bm_step 2 and block_size 4 are outside the 8x8 batched kernels:
    # the batch runs one single-frame launch per frame and stage
    frames = FRAMES[:4]
    for kw in (dict(profile="fast", bm_step=2), dict(profile="lc", block_size=4, block_step=3)):
        hosts, src = _inputs(frames, 7400)
        b = gpu.Batch([(h, w) for h, w, _ in frames], [s for *_, s in frames], **kw)
        for graph in (False, True, True):
            dst = _outputs(frames)
            b.run(src, dst, graph=graph)
            for k, (h, w, s) in enumerate(frames):
                _bits_equal(_host(dst[k]), _single(gpu, hosts[k], s, **kw), f"{kw} frame {k} graph {graph}")


def test_batch_errors(gpu):
    with pytest.raises(gpu.Error, match="sigma"):
        gpu.Batch([(32, 32)], [10.0], sigma=5.0)
    with pytest.raises(gpu.Error, match="sigma must be positive"):
        gpu.Batch([(32, 32), (32, 40)], [10.0, 0.0])
    with pytest.raises(gpu.Error, match="pitch"):
        gpu.Batch([(32, 32)], [10.0], pitches=[16])
    with pytest.raises(gpu.Error, match="bm3d.Basic"):
        gpu.Batch([(32, 32)], [10.0], profile="nope")
    with pytest.raises(gpu.Error, match="smaller than block_size"):
        gpu.Batch([(32, 32), (6, 40)], [10.0, 10.0])
    b = gpu.Batch([(32, 32)], [10.0])
    with pytest.raises(gpu.Error, match="shape"):
        b.shape = {"match_waves": 3, "match_strip": 8, "group_order": 0}
    with pytest.raises(gpu.Error, match="shape"):
        b.shape = {"match_waves": 2, "match_strip": 8, "group_order": 2}
