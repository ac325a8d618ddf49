"""Trains the learned block-matching proposal network on the GPU.

    python vapoursynth-bm3d_amd/lbm_train.py [--frames 200] [--epochs 4] [--out lbm_weights.npz]

Training data: the exhaustive match lists of Basic(profile='fast',
sigma=25, bm_range=24, group_size=16) -- the learned-matching workload -- on
200 seeded synthetic noisy Gray float frames (512x512; piecewise-smooth
images, AWGN sigma 25, seeds 5000 + k).  The network is trained by the
library's own forward, backward and Adam kernels (lbm.Model.train_step).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

LEARNED_KW = dict(profile="fast", sigma=25.0, bm_range=24, group_size=16)
WEIGHTS = os.path.join(HERE, "lbm_weights.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()

    import bm3d
    import lbm
    import synth

    bm3d.init(0)
    t0 = time.time()
    frames = [synth.gray_float(args.size, args.size, 25.0, seed=5000 + k)[0] for k in range(args.frames)]
    print(f"{len(frames)} frames {args.size}x{args.size} generated in {time.time() - t0:.1f} s", flush=True)
    model = lbm.Model(seed=args.seed)
    t0 = time.time()
    hist = lbm.train(model, frames, LEARNED_KW, epochs=args.epochs, lr=args.lr, batch_pairs=8192, seed=args.seed,
                     log=lambda m: print(m, flush=True))
    print(f"trained in {time.time() - t0:.1f} s; loss per epoch {[round(h, 5) for h in hist]}", flush=True)
    model.save(args.out)
    print(f"weights -> {args.out}", flush=True)


if __name__ == "__main__":
    main()
