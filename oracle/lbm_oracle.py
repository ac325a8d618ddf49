"""CPU oracle of the learned block-matching network -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module; the library never does.

The network has no counterpart in the reference (whose matching is the
exhaustive Block::BlockMatchingMulti, include/Block.h:683-729), so this file
restates the project's own definition, operation by operation, with the
rounding points of the HIP kernels:

  forward   X -> bf16;  Hpre = bf16(W1) bf16(X) + b1;  H = relu(Hpre)
            E = bf16(W2) bf16(H) + b2
  loss      d = |E_ref - E_cand|^2, w = 1 / (t + t0)^2,
            L = mean_p w (d - t)^2
  backward  dE_ref = 2 g (E_ref - E_cand), dE_cand = -dE_ref,
            g = 2 w (d - t) / npairs
            dHpre = (bf16(dE) bf16(W2)) * [H > 0]
            dW1 = bf16(dHpre)^T bf16(X), dW2 = bf16(dE)^T bf16(H)
            db1 = sum dHpre, db2 = sum dE

Products of bf16-rounded operands are exact in float64 and the sums are
accumulated in float64, so the only difference from the kernels is their
FP32 accumulation order.  Pins: the bf16 rounding against hand-derived
values, and the unrounded gradient (float64 throughout) against central
finite differences of the loss.
"""
from __future__ import annotations

import numpy as np

K_IN, K_HIDDEN, K_EMB = 64, 128, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + K_HIDDEN * K_IN
OFF_W2 = OFF_B1 + K_HIDDEN
OFF_B2 = OFF_W2 + K_EMB * K_HIDDEN
N_PARAMS = OFF_B2 + K_EMB
T0 = 1.0


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params):
    p = np.asarray(params, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(K_HIDDEN, K_IN)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(K_EMB, K_HIDDEN)
    b2 = p[OFF_B2:N_PARAMS]
    return W1, b1, W2, b2


def patches(plane: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    """[2 npairs, 64] patches: row 2p the reference of pair p, 2p+1 its candidate."""
    out = np.zeros((2 * len(pairs), K_IN), np.float32)
    for p, (ry, rx, qy, qx) in enumerate(np.asarray(pairs)):
        out[2 * p] = plane[ry:ry + 8, rx:rx + 8].reshape(-1)
        out[2 * p + 1] = plane[qy:qy + 8, qx:qx + 8].reshape(-1)
    return out


def pair_ssd(plane: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    X = patches(plane, pairs).astype(np.float64)
    return ((X[0::2] - X[1::2]) ** 2).sum(axis=1)


def forward(params, X, rounding=True):
    """Returns (H, E) for patches X [n, 64]."""
    W1, b1, W2, b2 = split(params)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    Hpre = r(X) @ r(W1).T + b1
    if rounding:
        Hpre = Hpre.astype(np.float32).astype(np.float64)
    H = np.maximum(Hpre, 0.0)
    E = r(H) @ r(W2).T + b2
    return H, E


def loss_and_grad(params, X, target, rounding=True, t0=T0):
    """Batch loss and gradient (parameter order) for patches X (pairs
    interleaved) and per-pair targets."""
    W1, b1, W2, b2 = split(params)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    H, E = forward(params, X, rounding)
    t = np.asarray(target, np.float64)
    n = len(t)
    diff = E[0::2] - E[1::2]
    d = (diff ** 2).sum(axis=1)
    w = 1.0 / (t + t0) ** 2
    res = d - t
    loss = float(np.mean(w * res * res))
    g = 2.0 * w * res / n
    dE = np.zeros_like(E)
    dE[0::2] = 2.0 * g[:, None] * diff
    dE[1::2] = -dE[0::2]
    if rounding:
        dE = dE.astype(np.float32).astype(np.float64)
    dH = (r(dE) @ r(W2)) * (H > 0)
    if rounding:
        dH = dH.astype(np.float32).astype(np.float64)
    gW1 = r(dH).T @ r(X)
    gW2 = r(dE).T @ r(H)
    grad = np.concatenate([gW1.reshape(-1), dH.sum(axis=0), gW2.reshape(-1), dE.sum(axis=0)])
    return loss, grad

