"""Independent pins of the oracle's V-BM3D functions and conversion offsets.

The GPU kernels and the oracle were written together, so the oracle's
V-BM3D side (or_vblock_match, or_vbm3d_kernel's slot layout,
or_vaggregate_plane) is pinned here against references it does not share
code with:

* a float64 Python restatement of VBM3D_Process_Base::BlockMatching
  (source/VBM3D_Base.cpp:540-659) with Block::BlockMatchingMulti
  (include/Block.h:655-729), _SearchBoundary (:734-764) and
  AddSearchPos/MergeSearchPos/GenSearchPos (:771-827), written from the
  reference text: the member sets, frames, order and keys must agree on tiny
  clips away from near-ties;
* brute force: with ps_range covering the frame, the predictive windows are
  whole frames, so the list is self plus the G-1 smallest keys over every
  frame (self position allowed in the other frames);
* the stack layout (VBM3D_Base.cpp:697-698, include/VBM3D_Base.h:284-292):
  slots of absent neighbours stay unwritten; with group_size 1 the current
  slot holds exactly the single-frame Basic's numerator/denominator and the
  other present slots are zero;
* VAggregate (source/VAggregate.cpp:93-109, 194-202) on tagged stacks whose
  only finite, exact answer comes from reading slot r-(m-n) of neighbour m;
* conversion offsets (include/Conversion.hpp:59, 213, 249): hand-derived
  samples at the rounding boundary."""
import numpy as np
import pytest

import oracle as O

# --------------------------------------------------------------------------- restatement


def _search_boundary(pos, bound, rng, step):
    """Block::_SearchBoundary (include/Block.h:734-764)."""
    rng = rng // step * step
    if pos == bound:
        return bound
    if pos > bound:
        s = pos - rng
        while s < bound:
            s += step
        return s
    s = pos + rng
    while s > bound:
        s -= step
    return s


def _window(j, i, W, H, B, rng, step):
    """Row-major lattice of Block::AddSearchPos (include/Block.h:771-789)."""
    rng = rng // step * step
    l, r = _search_boundary(i, 0, rng, step), _search_boundary(i, W - B, rng, step)
    t, b = _search_boundary(j, 0, rng, step), _search_boundary(j, H - B, rng, step)
    return [(y, x) for y in range(t, b + 1, step) for x in range(l, r + 1, step)]


def _gen_search_pos(seeds, W, H, B, rng, step):
    """GenSearchPos (include/Block.h:811-827): std::merge of (y, x)-sorted
    windows one seed at a time, then std::unique."""
    acc = []
    for s in seeds:
        new = _window(s[0], s[1], W, H, B, rng, step)
        merged, a, c = [], 0, 0
        while a < len(acc) or c < len(new):  # std::merge: equal -> first range first
            if c >= len(new) or (a < len(acc) and not new[c] < acc[a]):
                merged.append(acc[a])
                a += 1
            else:
                merged.append(new[c])
                c += 1
        acc = merged
    out = []
    for p in acc:
        if not out or out[-1] != p:
            out.append(p)
    return out


class _Restated:
    """float64 restatement; records the keys every selection compared so that
    blocks whose outcome a float32 rounding could flip are skipped."""

    def __init__(self, planes, B, G, R, S, ps_num, ps_range, ps_step, th_sse, dist_mul):
        self.p = planes.astype(np.float64)
        self.B, self.G, self.R, self.S = B, G, R, S
        self.ps_num, self.ps_range, self.ps_step = ps_num, ps_range, ps_step
        self.th_sse, self.dist_mul = th_sse, dist_mul
        self.near = False

    def _core(self, ref, frame, cands):
        """BlockMatchingMulti core (include/Block.h:554-653): accept
        dist <= thSSE and dist != 0, key = dist * distMul."""
        B = self.B
        out = []
        for (y, x) in cands:
            d = float(((self.p[frame, y:y + B, x:x + B] - ref) ** 2).sum())
            if abs(d - self.th_sse) <= 1e-5 * self.th_sse:
                self.near = True
            if d <= self.th_sse and d != 0.0:
                out.append((d * self.dist_mul, (y, x)))
        return out

    def _cut(self, lst, first_fixed):
        """partial_sort to G when longer, else stable_sort (key only)."""
        G = self.G
        keys = sorted(k for k, _ in lst[1 if first_fixed else 0:])
        for a, b in zip(keys, keys[1:]):
            if b - a <= 1e-5 * b:
                self.near = True
        if len(lst) > G:
            head = lst[:1] if first_fixed else []
            rest = lst[1:] if first_fixed else lst
            rest = sorted(rest, key=lambda kp: kp[0])
            return head + rest[:G - len(head)]
        return sorted(lst, key=lambda kp: kp[0])

    def match(self, cur, j, i):
        B, T = self.B, self.p.shape[0]
        H, W = self.p.shape[1:]
        ref = self.p[cur, j:j + B, i:i + B]
        cands = [q for q in _window(j, i, W, H, B, self.R, self.S) if q != (j, i)]
        lst = self._cut([(0.0, (j, i))] + self._core(ref, cur, cands), True)
        code = [(k, cur, q) for k, q in lst]
        cur_seeds = [q for _, q in lst[:self.ps_num]]
        cur_pos = _gen_search_pos(cur_seeds, W, H, B, self.ps_range, self.ps_step)
        for frames in (range(cur - 1, -1, -1), range(cur + 1, T)):
            pos = cur_pos
            for f in frames:
                fm = self._cut(self._core(ref, f, pos), False)
                code += [(k, f, q) for k, q in fm]
                pos = _gen_search_pos([q for _, q in fm[:self.ps_num]], W, H, B, self.ps_range, self.ps_step)
        if len(code) > self.G:
            rest = sorted(code[1:], key=lambda c: c[0])
            ks = [c[0] for c in rest]
            g = self.G - 1
            if g < len(ks) and g > 0 and ks[g] - ks[g - 1] <= 1e-5 * ks[g]:
                self.near = True
            code = code[:1] + rest[:g]
        return code


def _clip(T, H, W, seed, shift=(1, 2), noise=0.03):
    rng = np.random.RandomState(seed)
    base = rng.rand(H + 8 * T, W + 8 * T)
    k = np.ones(3) / 3.0
    base = np.apply_along_axis(lambda r: np.convolve(r, k, "same"), 1, base)
    base = np.apply_along_axis(lambda c: np.convolve(c, k, "same"), 0, base)
    frames = []
    for t in range(T):
        dy, dx = shift[0] * t, shift[1] * t
        frames.append(base[dy:dy + H, dx:dx + W] + noise * rng.randn(H, W))
    return np.clip(np.stack(frames), 0.0, 1.0).astype(np.float32)


def _vfilter(H, W, **kw):
    p = O.resolve("VBasic", width=W, height=H, color_family="gray", **kw)
    return O.make_filter(p, tie_mode=O.TIES_INDEX)


def _grid(dim, B, s):
    last = dim - B
    out, j = [], 0
    while True:
        if j >= last + s:
            break
        out.append(min(j, last))
        j += s
    return out


@pytest.mark.parametrize("T,cur,kw", [
    (5, 2, dict(block_size=4, block_step=3, group_size=8, bm_range=5, bm_step=1, ps_num=2, ps_range=2, ps_step=1)),
    (4, 0, dict(block_size=4, block_step=2, group_size=6, bm_range=4, bm_step=2, ps_num=3, ps_range=3, ps_step=2)),
    (3, 2, dict(block_size=5, block_step=4, group_size=12, bm_range=6, bm_step=1, ps_num=1, ps_range=2, ps_step=1)),
])
def test_vblock_match_equals_restated_predictive_search(oracle_lib, T, cur, kw):
    H, W = 24, 32
    planes = _clip(T, H, W, seed=91 + T)
    f = _vfilter(H, W, th_mse=1500.0, radius=2, **kw)
    B, G = f.p.block_size, f.p.group_size
    mse2sse = float(B * B) / (255.0 * 255.0)
    th_sse = float(np.float32(f.p.th_mse * mse2sse))
    R = _Restated(planes, B, G, f.p.bm_range, f.p.bm_step, f.p.ps_num, f.p.ps_range, f.p.ps_step,
                  th_sse, 1.0 / mse2sse)
    checked = multi_frame = 0
    blocks = [(j, i) for j in _grid(H, B, f.p.block_step) for i in _grid(W, B, f.p.block_step)]
    for (j, i) in blocks:
        R.near = False
        want = R.match(cur, j, i)
        if R.near:
            continue
        keys, zs, ys, xs = O.vblock_match(f, planes, cur, j, i)
        got = list(zip(zs.tolist(), ys.tolist(), xs.tolist()))
        assert got == [(z, q[0], q[1]) for _, z, q in want], (j, i, got, want)
        np.testing.assert_allclose(keys, [k for k, _, _ in want], rtol=2e-5, atol=0)
        checked += 1
        multi_frame += len(set(zs.tolist())) > 1
    assert checked >= 0.8 * len(blocks), (checked, len(blocks))
    assert multi_frame > 0


def test_vblock_match_wide_predictive_window_is_global_best(oracle_lib):
    """ps_range covering the frame: every neighbour frame is searched whole,
    so the result is self + the G-1 smallest keys over all frames (the self
    position counts as a candidate in the other frames)."""
    T, H, W, cur = 4, 20, 24, 1
    planes = _clip(T, H, W, seed=7, shift=(0, 1), noise=0.05)
    f = _vfilter(H, W, block_size=4, block_step=4, group_size=10, bm_range=30, bm_step=1, ps_num=2,
                 ps_range=40, ps_step=1, th_mse=4000.0, radius=2)
    B, G = 4, 10
    mse2sse = float(B * B) / (255.0 * 255.0)
    th_sse = float(np.float32(f.p.th_mse * mse2sse))
    p64 = planes.astype(np.float64)
    checked = 0
    for j in _grid(H, B, 4):
        for i in _grid(W, B, 4):
            ref = p64[cur, j:j + B, i:i + B]
            allc = []
            for t in range(T):
                for y in range(H - B + 1):
                    for x in range(W - B + 1):
                        if t == cur and (y, x) == (j, i):
                            continue
                        d = float(((p64[t, y:y + B, x:x + B] - ref) ** 2).sum())
                        if d <= th_sse and d != 0.0:
                            allc.append((d / mse2sse, t, y, x))
            allc.sort()
            ks = [c[0] for c in allc]
            if any(b - a <= 1e-5 * b for a, b in zip(ks[:G], ks[1:G + 1])) or \
                    any(abs(c[0] * mse2sse - th_sse) <= 1e-5 * th_sse for c in allc):
                continue
            keys, zs, ys, xs = O.vblock_match(f, planes, cur, j, i)
            want = [(cur, j, i)] + [(t, y, x) for _, t, y, x in allc[:G - 1]]
            got = list(zip(zs.tolist(), ys.tolist(), xs.tolist()))
            if len(allc) + 1 > G:
                assert got == want, (j, i)
            else:  # no cut: concatenation order, compare as sets
                assert sorted(got) == sorted(want), (j, i)
            checked += 1
    assert checked >= 10


# --------------------------------------------------------------------------- stack layout


@pytest.mark.parametrize("cur,T", [(0, 3), (1, 4), (2, 5), (2, 3)])
def test_vbm3d_stack_absent_slots_untouched(oracle_lib, cur, T):
    """radius 2: frame n's stack has slots 2k/2k+1 for frame n-2+k; the window
    holds T frames with n at index cur, so slots k with cur-2+k outside
    [0, T) belong to absent neighbours and must stay unwritten."""
    H, W = 20, 24
    planes = _clip(T, H, W, seed=11)
    src = planes[None]
    f = _vfilter(H, W, block_size=4, block_step=3, group_size=6, bm_range=4, ps_num=2, ps_range=2, radius=2)
    st = O.vbm3d_kernel(f, src, src, cur)
    for k in range(5):
        present = 0 <= cur - 2 + k < T
        num, den = st[0, 2 * k], st[0, 2 * k + 1]
        if present:
            assert np.isfinite(num).all() and np.isfinite(den).all() and (den >= 0).all(), k
        else:
            assert np.isnan(num).all() and np.isnan(den).all(), k
    assert (st[0, 2 * 2 + 1] > 0).all()  # every pixel of the current frame is covered


def test_vbm3d_group_size_one_is_single_frame_basic(oracle_lib):
    """group_size 1: no matching (VBM3D_Base.cpp:543-546), so the current
    slot is the single-frame Basic's accumulation and neighbour slots are
    zero; num/den equals or_bm3d_kernel's Basic output."""
    H, W, T, cur = 24, 28, 5, 2
    planes = _clip(T, H, W, seed=12)
    kw = dict(block_size=4, block_step=3, group_size=1, bm_range=4, ps_num=1, ps_range=2)
    fv = _vfilter(H, W, radius=2, **kw)
    st = O.vbm3d_kernel(fv, planes[None], planes[None], cur)
    pb = O.resolve("Basic", width=W, height=H, color_family="gray", **kw)
    basic = O.bm3d_kernel(O.make_filter(pb, tie_mode=O.TIES_INDEX), planes[cur][None])[0]
    for k in range(5):
        if k == 2:
            continue
        assert (st[0, 2 * k] == 0).all() and (st[0, 2 * k + 1] == 0).all(), k
    out = (st[0, 4] / st[0, 5]).astype(np.float32)
    assert np.array_equal(out, basic)


# --------------------------------------------------------------------------- VAggregate


@pytest.mark.parametrize("radius,n,N", [(1, 0, 5), (2, 3, 8), (3, 1, 9), (2, 7, 8)])
def test_vaggregate_reads_slot_of_frame_n_from_each_neighbour(oracle_lib, radius, n, N):
    """Neighbour m = n+b..n+f holds frame n's sums in slot r-(m-n).  Tagged
    stacks: that slot carries num 2^t, den 1; every other slot num 2^20, den
    1000.  Sums of distinct powers of two are exact, so the only result is
    float32((2^T - 1) / T)."""
    H, W = 3, 5
    first, last = n - min(n, radius), n + min(N - 1 - n, radius)
    T, cur = last - first + 1, n - first
    S = (2 * radius + 1) * 2
    st = np.empty((T, S, H, W), np.float32)
    for t in range(T):
        m = first + t
        for k in range(S // 2):
            right = k == radius - (m - n)
            st[t, 2 * k] = 2.0 ** t if right else 2.0 ** 20
            st[t, 2 * k + 1] = 1.0 if right else 1000.0
    out = O.vaggregate_plane(st, radius, cur)
    want = np.float32(np.float32(2 ** T - 1) / np.float32(T))
    assert (out == want).all(), (out, want)


# --------------------------------------------------------------------------- conversion offsets


def test_pc_chroma_offset_0499999_rounds_to_half_in_float():
    """RangeConvert / MatrixConvert add FLType(0.499999) to a float offset of
    2^(b-1) (include/Conversion.hpp:59, 213, 249).  In binary32 the sum is
    exactly 2^(b-1) + 0.5 for every depth 8..16, so +0.499999 and +0.5 give
    the same outputs through the reference's arithmetic."""
    for b in range(8, 17):
        off = np.float32(np.float32(2 ** (b - 1)) + np.float32(0.499999))
        assert off == np.float32(2 ** (b - 1) + 0.5), b


@pytest.mark.parametrize("bits", [8, 10, 16])
def test_float2int_chroma_boundary_samples(oracle_lib, bits):
    """Hand-derived: PC-range chroma Float2Int of v = -0.5 is
    trunc(-0.5 * (2^b - 1) + (2^(b-1) + 0.5)) = trunc(1.0) = 1 (real-number
    arithmetic with +0.499999 would give 0); v = 0 gives 2^(b-1); v = 0.5 gives
    2^b - 1 after clipping; luma v = 1 gives 2^b - 1, v = 0.5/(2^b-1) gives 1."""
    m = 2 ** bits - 1
    chroma = np.array([[-0.5, 0.0, 0.5, 0.75]], np.float32)
    got = O.float2int(chroma, bits, True, True, True)
    assert got.tolist() == [[1, 2 ** (bits - 1), m, m]]
    luma = np.array([[1.0, 0.5 / m, 0.0, -0.25]], np.float32)
    got = O.float2int(luma, bits, False, True, True)
    assert got.tolist() == [[m, 1, 0, 0]]


def test_rgb2opp_16bit_boundary_samples(oracle_lib):
    """Hand-derived OPP quantisation at 16 bits (Conversion.hpp:205-236):
    gU = 0.5, gV = 0.25, offsetC = 32768.5 in binary32.  U(R=0,B=65535) =
    trunc(-32767.5 + 32768.5) = 1; V(G=65535) = trunc(-131070 * 0.25 +
    32768.5) = 1; U(R=65535,B=0) = trunc(65535.5) -> clipped 65535;
    Y(1,1,1) = trunc(3 * (1/3) + 0.5) = 1."""
    rgb = np.zeros((3, 1, 4), np.uint16)
    rgb[:, 0, 0] = (0, 0, 65535)
    rgb[:, 0, 1] = (0, 65535, 0)
    rgb[:, 0, 2] = (65535, 0, 0)
    rgb[:, 0, 3] = (1, 1, 1)
    yuv = O.rgb2opp(rgb, 16, 16, True)
    assert int(yuv[1, 0, 0]) == 1
    assert int(yuv[2, 0, 1]) == 1
    assert int(yuv[1, 0, 2]) == 65535
    assert int(yuv[0, 0, 3]) == 1
