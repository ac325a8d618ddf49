"""Pins of the learned-matching oracle (CPU): the bfloat16 rounding against
hand-derived values and the network gradient against central finite
differences of the loss.  Plus the host-side helpers of lbm.py."""
import numpy as np

import lbm
import lbm_oracle as LO


def test_bf16_rounding_hand_values():
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    # 1 + 2^-8 + 2^-20 is above the halfway point -> 1 + 2^-7
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-3], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0029296875 + 0.0], np.float64)
    got = LO.bf16(x)
    assert got[:5].tolist() == want[:5].tolist()
    # 3e-3 = 1.536 * 2^-9; bf16 keeps 8 significant bits: 1.5390625 * 2^-9
    assert got[5] == 1.5390625 * 2 ** -9


def test_parameter_layout_matches_module():
    assert LO.N_PARAMS == lbm.N_PARAMS == 64 * 128 + 128 + 128 * 32 + 32
    assert (LO.OFF_B1, LO.OFF_W2, LO.OFF_B2) == (lbm.OFF_B1, lbm.OFF_W2, lbm.OFF_B2)


def test_gradient_matches_finite_differences():
    rng = np.random.RandomState(3)
    params = lbm.init_params(7).astype(np.float64)
    params[LO.OFF_B1:LO.OFF_W2] = rng.normal(0, 0.1, LO.K_HIDDEN)  # exercise the bias paths
    X = rng.rand(2 * 6, 64)
    target = rng.rand(6) * 5.0
    loss, grad = LO.loss_and_grad(params, X, target, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(LO.OFF_W1, LO.OFF_B1), 12, replace=False),
                          rng.choice(np.arange(LO.OFF_B1, LO.OFF_W2), 6, replace=False),
                          rng.choice(np.arange(LO.OFF_W2, LO.OFF_B2), 12, replace=False),
                          np.arange(LO.OFF_B2, LO.OFF_B2 + 4)])
    h = 1e-6
    for i in idx:
        p1 = params.copy()
        p1[i] += h
        p2 = params.copy()
        p2[i] -= h
        fd = (LO.loss_and_grad(p1, X, target, rounding=False)[0] - LO.loss_and_grad(p2, X, target, rounding=False)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_loss_is_zero_when_embedding_distance_equals_target():
    params = lbm.init_params(1).astype(np.float64)
    rng = np.random.RandomState(0)
    X = rng.rand(4, 64)
    _, E = LO.forward(params, X, rounding=False)
    d = ((E[0::2] - E[1::2]) ** 2).sum(axis=1)
    loss, grad = LO.loss_and_grad(params, X, d, rounding=False)
    assert loss < 1e-20 and np.abs(grad).max() < 1e-9


def test_sample_pairs_positives_and_window():
    rng = np.random.RandomState(0)
    count = np.array([3, 1], np.int32)
    pos = np.zeros((2, 4, 2), np.int32)
    pos[0, :3] = [(10, 12), (11, 15), (9, 9)]
    pos[1, :1] = [(0, 0)]
    pairs = lbm.sample_pairs(count, pos, W=40, H=30, R=4, rng=rng)
    assert pairs[:2].tolist() == [[10, 12, 11, 15], [10, 12, 9, 9]]
    negs = pairs[2:4]
    assert ((negs[:, 2] >= 6) & (negs[:, 2] <= 14) & (negs[:, 3] >= 8) & (negs[:, 3] <= 16)).all()
    assert len(pairs) == 4  # block 1 has no matches: no negatives either


def test_recall_counts_shared_matches():
    c = np.array([3, 2])
    a = np.array([[(0, 0), (1, 1), (2, 2)], [(5, 5), (6, 6), (0, 0)]])
    b = np.array([[(0, 0), (2, 2), (7, 7)], [(5, 5), (9, 9), (0, 0)]])
    assert lbm.recall(c, a, np.array([3, 2]), b) == 1 / 3
