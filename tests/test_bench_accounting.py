"""The benchmark's work accounting (the numerator of the roofline figures),
checked against operation counts taken from explicit implementations on tiny
inputs: a dense separable 3D DCT of one group and one exhaustive SSD, with
every multiply and add counted as it happens."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


class Counter:
    def __init__(self):
        self.mul = 0
        self.add = 0


def dense_dct_1d(x, ctr):
    """Dense DCT-II of a line: every output a full dot product (n multiplies,
    n - 1 adds), the form the algorithmic FLOP count assumes."""
    n = len(x)
    out = []
    for k in range(n):
        acc = None
        for j in range(n):
            t = np.cos(np.pi * k * (2 * j + 1) / (2 * n)) * x[j]
            ctr.mul += 1
            if acc is None:
                acc = t
            else:
                acc = acc + t
                ctr.add += 1
        out.append(acc)
    return out


def dense_dct_3d(g, ctr):
    """Separable 3D transform of a group g[z][y][x]: x lines, y lines, z lines."""
    n, B, _ = g.shape
    g = g.astype(object).copy()
    for z in range(n):
        for y in range(B):
            g[z, y, :] = dense_dct_1d(list(g[z, y, :]), ctr)
    for z in range(n):
        for x in range(B):
            g[z, :, x] = dense_dct_1d(list(g[z, :, x]), ctr)
    for y in range(B):
        for x in range(B):
            g[:, y, x] = dense_dct_1d(list(g[:, y, x]), ctr)
    return g


@pytest.mark.parametrize("n,B", [(1, 4), (3, 4), (5, 8)])
def test_transform_flops_match_an_explicit_count(n, B):
    ctr = Counter()
    rng = np.random.default_rng(n * 10 + B)
    dense_dct_3d(rng.standard_normal((n, B, B)), ctr)
    counted = ctr.mul + ctr.add
    # the formula counts a multiply-add pair per term (2 flops); the explicit
    # form has one fewer add per output, B*B*n outputs per axis pass
    per_transform_formula = 4 * B ** 3 * n + 2 * B * B * n * n
    assert counted + 3 * B * B * n == per_transform_formula
    # bench's per-stage figure: T transforms per group and plane plus the
    # elementwise work, for one group of n members
    stats = {"ssd_evals": 0, "members": n, "members_sq": n * n}
    ssd, dct = bench.flops_per_frame(stats, B, wiener=True, planes=1)
    assert ssd == 0
    assert dct == 3 * per_transform_formula + 8 * B * B * n
    ssd, dct = bench.flops_per_frame(stats, B, wiener=False, planes=3)
    assert dct == 3 * (2 * per_transform_formula + 4 * B * B * n)


def test_ssd_flops_match_an_explicit_count():
    B = 8
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((B, B)), rng.standard_normal((B, B))
    ctr = Counter()
    s = 0.0
    for y in range(B):
        for x in range(B):
            d = a[y, x] - b[y, x]
            ctr.add += 1
            s = s + d * d
            ctr.mul += 1
            ctr.add += 1
    assert s == pytest.approx(float(((a - b) ** 2).sum()))
    ssd, _ = bench.flops_per_frame({"ssd_evals": 7, "members": 0, "members_sq": 0}, B, wiener=False)
    assert ssd == 7 * (ctr.mul + ctr.add)
