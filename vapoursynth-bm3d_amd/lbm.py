"""Learned block-matching proposal (bm3d_lbm.h) from Python.

A patch-embedding network (8x8 patch -> 128 ReLU -> 32, BF16 weights, FP32
accumulation) is trained on the GPU by the library's own forward, backward
and Adam kernels so that the squared distance of two embeddings follows the
SSD of the two patches.  Attached to a Basic/Final filter
(``clip.filter.set_matcher(model, topk)``) it scores the even-lattice positions
(even row and column) of every search window by embedding distance and
proposes the best min(topk, 32); their 3x3 neighbourhoods are re-ranked by
the exact FP32 SSD.  Reference blocks whose candidates do not give G-1
acceptable matches fall back to the exhaustive search, and windows of at most
topk candidates are searched exhaustively.  No counterpart in the reference
plugin.

Training data: the exhaustive match lists of the plain matcher on synthetic
noisy frames -- every reference block contributes its matches (positives)
and the same number of random candidates of its search window.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Sequence

import numpy as np

import bm3d

K_IN, K_HIDDEN, K_EMB = 64, 128, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + K_HIDDEN * K_IN
OFF_W2 = OFF_B1 + K_HIDDEN
OFF_B2 = OFF_W2 + K_EMB * K_HIDDEN
N_PARAMS = OFF_B2 + K_EMB


def init_params(seed: int) -> np.ndarray:
    """He-normal weights and zero biases (float32, parameter order)."""
    rng = np.random.RandomState(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0.0, np.sqrt(2.0 / K_IN), K_HIDDEN * K_IN)
    p[OFF_W2:OFF_B2] = rng.normal(0.0, np.sqrt(1.0 / K_HIDDEN), K_EMB * K_HIDDEN)
    return p


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Model:
    """One network on the GPU (FP32 master parameters + BF16 copies)."""

    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 0):
        self.L = bm3d.init()
        if self.L.bm3d_lbm_param_count() != N_PARAMS:
            raise bm3d.Error("bm3d_lbm: parameter count mismatch between library and module")
        p = np.ascontiguousarray(init_params(seed) if params is None else params, np.float32)
        if p.shape != (N_PARAMS,):
            raise bm3d.Error(f"bm3d_lbm: expected {N_PARAMS} parameters")
        h = C.c_void_p()
        if self.L.bm3d_lbm_create(C.byref(h), _fp(p)) != 0:
            raise bm3d.Error(self.L.bm3d_last_error().decode())
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.bm3d_lbm_free(self.h)
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise bm3d.Error(self.L.bm3d_last_error().decode())

    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, np.float32)
        self._check(self.L.bm3d_lbm_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.L.bm3d_lbm_set_params(self.h, _fp(p)))

    def save(self, path: str):
        np.savez(path, params=self.params())

    @classmethod
    def load(cls, path: str) -> "Model":
        with np.load(path, allow_pickle=False) as z:
            return cls(params=z["params"])

    # plane: torch float32 CUDA tensor (H, W), contiguous; pairs: torch int32 CUDA (n, 4)
    def train_step(self, plane, pairs, lr: float) -> float:
        bm3d._sync_producer(plane.device)
        loss = C.c_float()
        self._check(self.L.bm3d_lbm_train_step(self.h, plane.data_ptr(), plane.stride(0), pairs.data_ptr(),
                                               pairs.shape[0], lr, C.byref(loss)))
        return loss.value

    def gradients(self, plane, pairs, target=None):
        bm3d._sync_producer(plane.device)
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_float()
        self._check(self.L.bm3d_lbm_gradients(self.h, plane.data_ptr(), plane.stride(0), pairs.data_ptr(),
                                              target.data_ptr() if target is not None else None, pairs.shape[0],
                                              _fp(g), C.byref(loss)))
        return g, loss.value

    def forward(self, plane, pairs) -> np.ndarray:
        bm3d._sync_producer(plane.device)
        e = np.zeros((2 * pairs.shape[0], K_EMB), np.float32)
        self._check(self.L.bm3d_lbm_forward(self.h, plane.data_ptr(), plane.stride(0), pairs.data_ptr(),
                                            pairs.shape[0], _fp(e)))
        return e


def match_plane(flt, model: Optional[Model], plane, topk: int = 32):
    """Match lists of one device plane with the parameters of a Basic/Final
    filter (flt = clip.filter): model None is the exhaustive search.
    Returns (count, pos (y << 16 | x), key) as torch tensors, the number of
    fallbacks and the GPU time in ms."""
    import torch
    H, W = plane.shape
    prm = flt.params()
    G, B, S = prm["group_size"], prm["block_size"], prm["block_step"]
    nb = _ref_count(W, B, S) * _ref_count(H, B, S)
    dev = plane.device
    count = torch.zeros(nb, dtype=torch.int32, device=dev)
    pos = torch.zeros(nb * G, dtype=torch.int32, device=dev)
    key = torch.zeros(nb * G, dtype=torch.float32, device=dev)
    fb = C.c_int()
    ms = C.c_float()
    bm3d._sync_producer(dev)
    rc = flt.L.bm3d_lbm_match_plane(flt.h, model.h if model is not None else None, topk, plane.data_ptr(),
                                    plane.stride(0), W, H, count.data_ptr(), pos.data_ptr(), key.data_ptr(),
                                    C.byref(fb), C.byref(ms))
    if rc != 0:
        raise bm3d.Error(flt.L.bm3d_last_error().decode())
    return count, pos.view(nb, G), key.view(nb, G), fb.value, ms.value


def _ref_count(dim: int, B: int, s: int) -> int:
    last, n, j = dim - B, 0, 0
    while True:
        if j >= last + s:
            break
        if j > last:
            j = last
        n += 1
        j += s
    return n


def sample_pairs(count: np.ndarray, pos: np.ndarray, W: int, H: int, R: int, rng: np.random.RandomState,
                 negatives: Optional[int] = None) -> np.ndarray:
    """Training pairs from exhaustive match lists (count[nb], pos[nb, G, 2]):
    every match of a reference block (its first entry) is a positive, and as
    many random candidates of its search window are added.  Returns int32
    (n, 4) records (ref y, ref x, cand y, cand x)."""
    out = []
    for b in range(len(count)):
        n = int(count[b])
        j, i = int(pos[b, 0, 0]), int(pos[b, 0, 1])
        k = n - 1 if negatives is None else negatives
        for z in range(1, n):
            out.append((j, i, int(pos[b, z, 0]), int(pos[b, z, 1])))
        if k > 0:
            ys = rng.randint(max(0, j - R), min(H - 8, j + R) + 1, size=k)
            xs = rng.randint(max(0, i - R), min(W - 8, i + R) + 1, size=k)
            for y, x in zip(ys, xs):
                out.append((j, i, int(y), int(x)))
    return np.asarray(out, np.int32).reshape(-1, 4)


def recall(count_ref, pos_ref, count_new, pos_new) -> float:
    """Fraction of the reference lists' matches (self excluded) that the new
    lists also hold."""
    hit = tot = 0
    for b in range(len(count_ref)):
        n = int(count_ref[b])
        a = set(map(tuple, pos_ref[b, 1:n].tolist()))
        m = int(count_new[b])
        c = set(map(tuple, pos_new[b, 1:m].tolist()))
        hit += len(a & c)
        tot += len(a)
    return hit / max(tot, 1)


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def train(model: Model, frames: Sequence[np.ndarray], basic_kwargs: Dict[str, object], epochs: int = 2,
          lr: float = 2e-3, batch_pairs: int = 8192, seed: int = 0, log=None) -> List[float]:
    """Trains on the exhaustive match lists of Basic(**basic_kwargs) on each
    noisy Gray float frame (numpy (H, W)).  Returns the mean loss per epoch."""
    import torch
    rng = np.random.RandomState(seed)
    data = []
    for f in frames:
        H, W = f.shape
        clip = bm3d.Clip.from_arrays(bm3d.GRAYS, [f[None]])
        node = bm3d.Basic(clip, **basic_kwargs)
        node.get_frame(0)
        cnt, pos, _ = node.filter.match_lists()
        R = node.filter.params()["bm_range"]
        pairs = sample_pairs(cnt, pos, W, H, R, rng)
        data.append((torch.from_numpy(np.ascontiguousarray(f, np.float32)).cuda(),
                     torch.from_numpy(pairs).cuda()))
    hist = []
    steps_total = epochs * sum((p.shape[0] + batch_pairs - 1) // batch_pairs for _, p in data)
    step = 0
    for ep in range(epochs):
        tot, nb = 0.0, 0
        for fi in rng.permutation(len(data)):
            plane, pairs = data[fi]
            perm = torch.from_numpy(rng.permutation(pairs.shape[0])).cuda()
            pp = pairs[perm].contiguous()
            for s in range(0, pp.shape[0], batch_pairs):
                # cosine decay of the learning rate over all steps
                lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * step / max(steps_total, 1)))
                tot += model.train_step(plane, pp[s:s + batch_pairs].contiguous(), float(lr_t))
                nb += 1
                step += 1
        hist.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f}")
    return hist
