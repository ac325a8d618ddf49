"""GPU tests of the learned block-matching proposal.

* the MFMA forward pass and the training gradients against the CPU oracle
  (lbm_oracle, which replays the kernels' bf16 rounding points; only the
  FP32 accumulation order differs, tolerance stated per check);
* training on exhaustive match lists lowers the loss and raises recall over
  the untrained network;
* every learned match list is an exact re-rank: its keys are the plain
  matcher's SSE2-order SSD keys of its positions, in (key, index) order, and
  with the whole window proposed the lists and the filtered frame equal the
  plain method's bit for bit."""
import numpy as np
import pytest

import lbm
import lbm_oracle as LO
import oracle as O
import synth

pytestmark = pytest.mark.gpu

DIST_MUL = 255.0 * 255.0 / 64.0


def _plane(h, w, sigma, seed):
    noisy, clean = synth.gray_float(h, w, sigma, seed=seed)
    return noisy, clean


def _rand_pairs(rng, h, w, n):
    return np.stack([rng.randint(0, h - 7, n), rng.randint(0, w - 7, n),
                     rng.randint(0, h - 7, n), rng.randint(0, w - 7, n)], axis=1).astype(np.int32)


def test_forward_matches_oracle(gpu):
    import torch
    rng = np.random.RandomState(11)
    noisy, _ = _plane(64, 80, 25.0, 900)
    pairs = _rand_pairs(rng, 64, 80, 200)
    model = lbm.Model(seed=5)
    e = model.forward(torch.from_numpy(noisy).cuda(), torch.from_numpy(pairs).cuda())
    _, eo = LO.forward(model.params(), LO.patches(noisy, pairs))
    err = np.abs(e - eo).max()
    assert err <= 1e-4 * max(1.0, np.abs(eo).max()), err


def test_gradients_match_oracle(gpu):
    import torch
    rng = np.random.RandomState(12)
    noisy, _ = _plane(72, 96, 25.0, 901)
    pairs = _rand_pairs(rng, 72, 96, 300)
    target = LO.pair_ssd(noisy, pairs).astype(np.float32)
    model = lbm.Model(seed=6)
    g, loss = model.gradients(torch.from_numpy(noisy).cuda(), torch.from_numpy(pairs).cuda(),
                              torch.from_numpy(target).cuda())
    lo, go = LO.loss_and_grad(model.params(), LO.patches(noisy, pairs), target)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    # per parameter block: |g - g_oracle| <= 2e-3 max|g_oracle| (FP32 sums of bf16 products)
    for a, b in ((LO.OFF_W1, LO.OFF_B1), (LO.OFF_B1, LO.OFF_W2), (LO.OFF_W2, LO.OFF_B2), (LO.OFF_B2, LO.N_PARAMS)):
        scale = np.abs(go[a:b]).max()
        assert np.abs(g[a:b] - go[a:b]).max() <= 2e-3 * scale, (a, b)


def _basic(bm3d, img, **kw):
    return bm3d.Basic(bm3d.Clip.from_arrays(bm3d.GRAYS, [img[None]]), **kw)


KW = dict(profile="fast", sigma=25.0, bm_range=12, group_size=16)


def _exact_key(img, j, i, y, x, mode=O.ARITH_SSE2):
    rb = np.ascontiguousarray(img[j:j + 8, i:i + 8])
    d = O.lib().or_ssd(O._fp(rb), O._fp(np.ascontiguousarray(img)), img.shape[1], y, x, 8, mode)
    return np.float32(np.float64(d) * DIST_MUL)


@pytest.mark.parametrize("arith,step", [(0, None), (1, None), (0, 3), (0, 5)])
def test_learned_lists_are_exact_reranks(gpu, arith, step):
    # block_step 8 (the profile's), 3 and 5: the proposal tiles and re-rank
    # runs cover blocks whose windows overlap by different amounts
    h, w = 96, 128
    noisy, _ = _plane(h, w, 25.0, 902)
    model = lbm.Model(seed=7)  # untrained: proposals are poor, the re-rank must still be exact
    kw = dict(KW, block_step=step) if step else KW
    node = _basic(gpu, noisy, arith=arith, **kw)
    node.filter.set_matcher(model, 32)
    node.get_frame(0)
    cnt, pos, key = node.filter.match_lists()
    mode = O.ARITH_SCALAR if arith else O.ARITH_SSE2
    for b in range(0, len(cnt), 5):
        n = int(cnt[b])
        j, i = pos[b, 0]
        assert key[b, 0] == 0.0
        ks = key[b, 1:n]
        assert np.all(np.diff(ks) >= 0), b
        assert len({(int(y), int(x)) for y, x in pos[b, :n]}) == n, ("duplicate position", b)
        for z in range(1, n):
            y, x = pos[b, z]
            assert (y, x) != (j, i)
            assert abs(int(y) - int(j)) <= 12 and abs(int(x) - int(i)) <= 12
            assert key[b, z] == _exact_key(noisy, j, i, y, x, mode), (b, z)


@pytest.mark.parametrize("bm_range", [3, 4])
def test_whole_window_proposals_equal_plain(gpu, bm_range):
    # windows of at most 9x9 positions hold at most 25 lattice points, all of
    # them proposed (top-k 64 -> 32 proposals): the 3x3 neighbourhoods then
    # cover every window position exactly once
    h, w = 64, 96
    noisy, _ = _plane(h, w, 25.0, 903)
    kw = dict(profile="fast", sigma=25.0, bm_range=bm_range, group_size=16)
    plain = _basic(gpu, noisy, **kw)
    out_p = np.asarray(plain.get_frame(0).planes[0])
    learned = _basic(gpu, noisy, **kw)
    learned.filter.set_matcher(lbm.Model(seed=8), 64)
    out_l = np.asarray(learned.get_frame(0).planes[0])
    cp, pp, kp = plain.filter.match_lists()
    cl, pl, kl = learned.filter.match_lists()
    assert np.array_equal(cp, cl)
    for b in range(len(cp)):
        n = cp[b]
        assert np.array_equal(pp[b, :n], pl[b, :n]) and np.array_equal(kp[b, :n].view(np.uint32), kl[b, :n].view(np.uint32))
    assert np.array_equal(out_p.view(np.uint32), out_l.view(np.uint32))


def test_training_lowers_loss_and_raises_recall(gpu):
    frames = [_plane(128, 128, 25.0, 910 + k)[0] for k in range(8)]
    model = lbm.Model(seed=9)
    untrained = lbm.Model(params=model.params())
    hist = lbm.train(model, frames, KW, epochs=6, lr=3e-3, batch_pairs=4096, seed=1)
    assert hist[-1] < 0.5 * hist[0], hist
    # held-out frame, wide window (bm_range 24): the 32 proposed lattice
    # positions are about 5% of each window's lattice, so recall measures
    # the network.  An untrained network is already a fair proxy (a random
    # ReLU projection roughly preserves patch distances); training must add
    # to it.
    wide = dict(KW, bm_range=24)
    held = _plane(160, 192, 25.0, 990)[0]
    plain = _basic(gpu, held, **wide)
    plain.get_frame(0)
    cp, pp, _ = plain.filter.match_lists()
    rec, fb = {}, {}
    for name, m in (("trained", model), ("untrained", untrained)):
        node = _basic(gpu, held, **wide)
        node.filter.set_matcher(m, 32)
        node.get_frame(0)
        cl, pl, _ = node.filter.match_lists()
        rec[name] = lbm.recall(cp, pp, cl, pl)
        fb[name] = node.filter.matcher_fallbacks()
    print("recall", rec, "fallbacks", fb, "blocks", len(cp))
    assert rec["trained"] > rec["untrained"] + 0.02, (rec, fb)
    assert fb["trained"] <= fb["untrained"], (rec, fb)
    assert rec["trained"] > 0.5, (rec, fb)


def test_fallback_blocks_get_the_exhaustive_lists(gpu):
    """A block with fewer than G-1 accepted re-ranked candidates is searched
    exhaustively again (the cooperative block-list kernel), so every learned
    list shorter than G must equal the plain matcher's list of that block.  A
    small th_mse leaves many blocks short of G-1 accepted candidates, and an
    untrained network makes most of them fall back."""
    h, w = 96, 128
    noisy, _ = _plane(h, w, 25.0, 904)
    kw = dict(profile="fast", sigma=25.0, bm_range=24, group_size=16, th_mse=600.0)
    plain = _basic(gpu, noisy, **kw)
    out_p = np.asarray(plain.get_frame(0).planes[0])
    learned = _basic(gpu, noisy, **kw)
    learned.filter.set_matcher(lbm.Model(seed=21), 32)
    learned.get_frame(0)
    cp, pp, kp = plain.filter.match_lists()
    cl, pl, kl = learned.filter.match_lists()
    fb = learned.filter.matcher_fallbacks()
    short = np.nonzero(cl < 16)[0]
    print("fallbacks", fb, "short lists", len(short), "of", len(cl), flush=True)
    assert fb > 0 and len(short) > 0
    for b in short:
        n = int(cl[b])
        assert n == int(cp[b]), b
        assert np.array_equal(pl[b, :n], pp[b, :n]), b
        assert np.array_equal(kl[b, :n].view(np.uint32), kp[b, :n].view(np.uint32)), b
    del out_p
