"""Pins of the CPU oracle against what fixes the reference's arithmetic
independently of the oracle's own code: transform definitions (scipy's
unnormalised DCT-II/III are FFTW's REDFT10/REDFT01), hand-derived SSD
lane-order results, the real libstdc++ std::partial_sort / std::stable_sort,
the derived constants tabulated from the reference's formulas, round-trip
identities of the conversions, and float64 brute force on tiny frames.

The reference ships no tests, fixtures or golden vectors for this path, and
it cannot be compiled here (it needs VapourSynth and FFTW headers that the
image lacks), so these are the pins."""
import os
import subprocess
import textwrap

import numpy as np
import pytest
import scipy.fft as sf

import oracle as O
import synth


def _fp(a):
    return O._fp(a)


# ----------------------------------------------------------------- transforms

@pytest.mark.parametrize("N", list(range(1, 33)) + [40, 48, 64])
def test_dct_matches_redft_definition(oracle_lib, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N).astype(np.float32)
    X = np.zeros(N, np.float32)
    O.lib().or_dct2_1d(_fp(x), _fp(X), N)
    ref = sf.dct(x.astype(np.float64), type=2)   # 2 sum x_j cos(pi k (2j+1) / 2N)
    assert np.abs(X - ref).max() <= 4e-7 * N * np.abs(x).max() * 2
    y = np.zeros(N, np.float32)
    O.lib().or_dct3_1d(_fp(X), _fp(y), N)
    ref3 = sf.dct(X.astype(np.float64), type=3)  # X_0 + 2 sum_k>=1 X_k cos(...)
    assert np.abs(y - ref3).max() <= 4e-7 * N * np.abs(X).max() * 2
    # REDFT01 o REDFT10 = 2N * identity: the finalAMP factor per axis
    assert np.allclose(y / (2 * N), x, atol=1e-5 * N)


@pytest.mark.parametrize("n,B", [(1, 8), (5, 8), (16, 8), (32, 8), (3, 11), (7, 4)])
def test_3d_transform_and_final_amp(oracle_lib, n, B):
    rng = np.random.default_rng(100 * n + B)
    g = rng.random((n, B, B)).astype(np.float32)
    t = g.copy()
    O.lib().or_fdct3d(_fp(t), n, B)
    ref = sf.dctn(g.astype(np.float64), type=2)
    assert np.abs(t - ref).max() <= 1e-6 * np.abs(ref).max() * (n + 2 * B)
    O.lib().or_idct3d(_fp(t), n, B)
    amp = (2 * n) * (2 * B) * (2 * B)            # finalAMP (source/BM3D.cpp:154)
    assert np.allclose(t / amp, g, atol=2e-6 * (n + 2 * B))


# ----------------------------------------------------------------- SSD order

def _ssd(block_ref, cand, B, arith=O.ARITH_SSE2):
    r = np.ascontiguousarray(block_ref, np.float32)
    c = np.ascontiguousarray(cand, np.float32)
    return O.lib().or_ssd(_fp(r), _fp(c), c.shape[1], 0, 0, B, arith)


def test_ssd_sse2_lane_order_hand_derived(oracle_lib):
    """Block.h:582-613: lane k accumulates column k then k+4 of every row;
    lanes fold as ((s0+s1)+s2)+s3.  With d^2 = 2^24 at (0,0) and d^2 = 1 at
    (0,1) and (0,5): lane 1 holds 1+1 = 2, lane 0 holds 2^24, and the fold
    gives 2^24+2 exactly.  A row-major sum would round each +1 away (2^24),
    and a 'lane k = columns 2k,2k+1' reading would give 2^24 as well."""
    ref = np.zeros((8, 8), np.float32)
    cand = np.zeros((8, 8), np.float32)
    cand[0, 0] = 4096.0
    cand[0, 1] = 1.0
    cand[0, 5] = 1.0
    assert _ssd(ref, cand, 8) == np.float32(16777218.0)
    assert _ssd(ref, cand, 8, O.ARITH_SCALAR) == np.float32(16777216.0)
    # fold order: s0 = 2^24, s2 = 1, s3 = 1 -> ((2^24+0)+1)+1 = 2^24
    # (a pairwise fold (s0+s1)+(s2+s3) would give 2^24+2)
    cand[:] = 0
    cand[0, 0] = 4096.0
    cand[0, 2] = 1.0
    cand[0, 3] = 1.0
    assert _ssd(ref, cand, 8) == np.float32(16777216.0)
    # rows: lane 0 gets (0,0)=2^24 then (1,0)=1 and (1,4)=1 -> rounds away
    cand[:] = 0
    cand[0, 0] = 4096.0
    cand[1, 0] = 1.0
    cand[1, 4] = 1.0
    assert _ssd(ref, cand, 8) == np.float32(16777216.0)


def test_ssd_residue_columns_added_row_major_after_simd(oracle_lib):
    """Block.h:615-631: for B=11 the last 3 columns are added one by one,
    row-major, to the already folded SIMD sum.  2^24 in the SIMD part then
    1 at (0,8) and 1 at (0,9): each +1 rounds away -> 2^24."""
    ref = np.zeros((11, 11), np.float32)
    cand = np.zeros((11, 11), np.float32)
    cand[0, 0] = 4096.0
    cand[0, 8] = 1.0
    cand[0, 9] = 1.0
    assert _ssd(ref, cand, 11) == np.float32(16777216.0)
    # B < 8: no SIMD part at all, plain row-major
    ref4 = np.zeros((4, 4), np.float32)
    c4 = np.zeros((4, 4), np.float32)
    c4[0, 0] = 4096.0
    c4[0, 1] = 1.0
    c4[0, 2] = 1.0
    assert _ssd(ref4, c4, 4) == np.float32(16777216.0)


def test_ssd_matches_float64_on_random_blocks(oracle_lib):
    rng = np.random.default_rng(5)
    for B in (1, 4, 8, 11, 16):
        for _ in range(20):
            r = rng.random((B, B)).astype(np.float32)
            c = rng.random((B, B)).astype(np.float32)
            want = float(((r.astype(np.float64) - c) ** 2).sum())
            got = _ssd(r, c, B)
            assert abs(got - want) <= 1e-6 * max(want, 1e-3) * B


# ----------------------------------------------------------------- selection vs libstdc++

_PS_SRC = textwrap.dedent(r"""
    #include <algorithm>
    #include <cstdio>
    #include <vector>
    struct KP { float key; int pay; bool operator<(const KP &o) const { return key < o.key; } };
    int main() {
        int n, mid, mode;
        while (std::scanf("%d %d %d", &n, &mid, &mode) == 3) {
            std::vector<KP> v(n);
            for (int i = 0; i < n; ++i) { std::scanf("%f", &v[i].key); v[i].pay = i; }
            if (mode == 0) std::partial_sort(v.begin(), v.begin() + mid, v.end());
            else std::stable_sort(v.begin(), v.end());
            for (int i = 0; i < (mode == 0 ? mid : n); ++i) std::printf("%d ", v[i].pay);
            std::printf("\n");
        }
    }
""")


@pytest.fixture(scope="module")
def libstdcxx_sorter(tmp_path_factory):
    d = tmp_path_factory.mktemp("ps")
    src = d / "ps.cpp"
    src.write_text(_PS_SRC)
    exe = d / "ps"
    subprocess.run(["g++", "-O2", "-std=c++17", str(src), "-o", str(exe)], check=True)
    return str(exe)


def test_partial_sort_emulation_matches_libstdcxx(oracle_lib, libstdcxx_sorter):
    """The oracle's heap-select restatement against the real libstdc++ on key
    sequences with many ties (the only place the two tie orders differ)."""
    rng = np.random.default_rng(9)
    cases = []
    for t in range(300):
        n = int(rng.integers(2, 120))
        mid = int(rng.integers(1, n + 1))
        keys = rng.integers(0, max(2, n // 4), n).astype(np.float32) * 0.5
        cases.append((n, mid, keys))
    inp = "".join(f"{n} {mid} 0 " + " ".join(f"{k:.1f}" for k in keys) + "\n" for n, mid, keys in cases)
    out = subprocess.run([libstdcxx_sorter], input=inp, capture_output=True, text=True, check=True).stdout
    lines = out.strip().split("\n")
    for (n, mid, keys), line in zip(cases, lines):
        want = [int(v) for v in line.split()]
        k = keys.copy()
        pay = np.arange(n, dtype=np.int32)
        O.lib().or_partial_sort_keys(_fp(k), O._ip(pay), n, mid)
        assert list(pay[:mid]) == want, (n, mid)


# ----------------------------------------------------------------- search window / grid

@pytest.mark.parametrize("dim,B,step", [(64, 8, 8), (64, 8, 3), (67, 8, 4), (20, 11, 6), (9, 8, 1), (480, 8, 8)])
def test_ref_grid(oracle_lib, dim, B, step):
    out = np.zeros(dim + 1, np.int32)
    n = O.lib().or_ref_grid(dim, B, step, O._ip(out))
    L = dim - B
    want = sorted(set(list(range(0, L + 1, step)) + [L]))
    assert list(out[:n]) == want


@pytest.mark.parametrize("R,S", [(9, 1), (16, 1), (7, 3), (12, 5), (4, 4)])
def test_search_boundary_is_lattice_extent(oracle_lib, R, S):
    """_SearchBoundary walks pos -+ R' towards pos by S: the result is the
    extreme lattice point pos + k*S (|k*S| <= R' = R//S*S) inside [0, last]."""
    last = 50
    Rp = R // S * S
    for pos in range(0, last + 1):
        lat = [pos + k * S for k in range(-(Rp // S), Rp // S + 1) if 0 <= pos + k * S <= last]
        assert O.lib().or_search_boundary(pos, 0, R, S) == min(lat)
        assert O.lib().or_search_boundary(pos, last, R, S) == max(lat)


# ----------------------------------------------------------------- constants

@pytest.mark.parametrize("fn,profile,w,h,fam,matrix,sigma,want", [
    # (sigma'_Y, sigma'_U, sigma'_V, thMSE') tabulated from Specification.h / BM3D.cpp
    ("Basic", "fast", 640, 480, "gray", None, 10, (0.026218, None, None, 802.27)),
    ("Basic", "np", 1920, 1080, "yuv", 100, 10, (0.022641, 0.027730, 0.024015, 692.82)),
    ("Final", "np", 1920, 1080, "yuv", 100, 10, (0.022641, 0.027730, 0.024015, 173.21)),
    ("Basic", "np", 1920, 1080, "yuv", None, 10, (0.029397, 0.025162, 0.026550, 899.54)),
    ("VBasic", "fast", 1920, 1080, "yuv", None, 15, (0.044095, 0.037743, 0.039825, 1199.38)),
    ("VFinal", "fast", 1920, 1080, "yuv", None, 15, (0.044095, 0.037743, 0.039825, 262.37)),
    ("Basic", "fast", 3840, 2160, "gray", None, 25, (0.071522, None, None, 1750.87)),
])
def test_normalised_sigma_and_th_mse(oracle_lib, fn, profile, w, h, fam, matrix, sigma, want):
    p = O.resolve(fn, width=w, height=h, color_family=fam, profile=profile, sigma=[sigma] * 3, matrix=matrix)
    for k in range(3):
        if want[k] is not None:
            assert abs(p.sigma[k] - want[k]) < 5e-7
    assert abs(p.th_mse - want[3]) < 5e-3


@pytest.mark.parametrize("fn,profile,sigma,plane,G,want", [
    ("Basic", "fast", 10, 0, 8, 4.53044),     # thrBase(G) for 640x480 smpte170m
    ("Basic", "np", 10, 0, 16, 5.53297),      # OPP
])
def test_threshold_base(oracle_lib, fn, profile, sigma, plane, G, want):
    fam, w, h, m = ("gray", 640, 480, None) if profile == "fast" else ("yuv", 1920, 1080, 100)
    p = O.resolve(fn, width=w, height=h, color_family=fam, profile=profile, sigma=[sigma] * 3, matrix=m)
    f = O.make_filter(p)
    assert abs(f.thr[plane][G - 1][0] - want) < 5e-5
    # the other entries scale by sqrt(2)^#zero-frequency axes
    t = [f.thr[plane][G - 1][k] for k in range(4)]
    assert np.allclose(t, [want, want * 2 ** 0.5, want * 2, want * 8 ** 0.5], rtol=1e-5)
    assert f.final_amp[G - 1] == (2 * G) * 16 * 16


def test_wiener_sigma(oracle_lib):
    p = O.resolve("Final", width=1920, height=1080, color_family="yuv", profile="np", sigma=[10] * 3, matrix=100)
    f = O.make_filter(p)
    # (sigma' * sqrt(finalAMP))^2 for G=32 = 8.39882, 12.59823, 9.44867 (Y, U, V)
    for k, want in enumerate((8.39882, 12.59823, 9.44867)):
        assert abs(f.wiener_sigma_sqr[k][31] - want) < 5e-5


def test_profile_table(oracle_lib):
    """README.md:284-323 profile tables."""
    rows = {
        # profile: Basic(B, step, G, R), Final(...), VBasic, VFinal
        "fast": [(8, 8, 8, 9), (8, 7, 8, 9), (8, 8, 8, 7), (8, 7, 8, 7)],
        "lc": [(8, 6, 16, 9), (8, 5, 16, 9), (8, 6, 8, 9), (8, 5, 8, 9)],
        "np": [(8, 4, 16, 16), (8, 3, 32, 16), (8, 4, 8, 12), (8, 3, 8, 12)],
        "high": [(8, 3, 16, 16), (8, 2, 32, 16), (8, 3, 8, 16), (8, 2, 8, 16)],
        "vn": [(8, 4, 32, 16), (11, 6, 32, 16), (8, 4, 16, 12), (11, 6, 16, 12)],
    }
    vrows = {"fast": (1, 2, 4, 5), "lc": (2, 2, 4, 5), "np": (3, 2, 5, 6), "high": (4, 2, 7, 8), "vn": (4, 2, 5, 6)}
    for prof, r in rows.items():
        for fn, want in zip(("Basic", "Final", "VBasic", "VFinal"), r):
            p = O.resolve(fn, width=64, height=64, color_family="gray", profile=prof)
            assert (p.block_size, p.block_step, p.group_size, p.bm_range) == want, (prof, fn)
            assert p.bm_step == 1
        for fn, k in (("VBasic", 2), ("VFinal", 3)):
            p = O.resolve(fn, width=64, height=64, color_family="gray", profile=prof)
            rad, psn, psrb, psrf = vrows[prof]
            assert (p.radius, p.ps_num, p.ps_range) == (rad, psn, psrb if fn == "VBasic" else psrf)
    # th_mse defaults (thMSE_Default) and hard_thr
    for prof in ("fast", "np", "vn"):
        pb = O.resolve("Basic", width=64, height=64, color_family="gray", profile=prof, sigma=[7], matrix=0)
        pf = O.resolve("Final", width=64, height=64, color_family="gray", profile=prof, sigma=[7], matrix=0)
        vn = prof == "vn"
        assert pb.th_mse == (1000 + 7 * 150 if vn else 400 + 7 * 80)
        assert pf.th_mse == (400 + 7 * 40 if vn else 200 + 7 * 10)
        assert pb.lambda_ == (2.8 if vn else 2.7)


# ----------------------------------------------------------------- conversions

def test_int_float_round_trip_16bit(oracle_lib):
    v = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    for chroma in (False, True):
        f = O.int2float(v, 16, chroma)
        back = O.float2int(f, 16, chroma)
        lo, hi = (-0.5, 0.5) if chroma else (0.0, 1.0)
        assert f.min() >= lo and f.max() <= hi
        want = v.copy()
        if chroma:
            # PC-range chroma is always clipped to [-0.5, 0.5] on the way in
            # (Conversion.hpp:51-55): code 0 maps below -0.5 and comes back as 1
            want[0, 0] = 1
        assert np.array_equal(back, want)


def test_int_float_round_trip_8bit_limited(oracle_lib):
    v = np.arange(256, dtype=np.uint8).reshape(16, 16)
    for chroma in (False, True):
        f = O.int2float(v, 8, chroma, full=False)
        back = O.float2int(f, 8, chroma, full=False)
        # identity inside the limited range, clamped to it outside
        hi = 240 if chroma else 235
        assert np.array_equal(back, np.clip(v, 16, hi))


def test_opp_matrices(oracle_lib):
    """RGB->OPP (Y=(R+G+B)/3, U=(R-B)/2, V=(R-2G+B)/4) and its inverse
    (Specification.h:176-192, 253-269) within float rounding; integer round
    trip within 1 LSB."""
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 65536, (3, 32, 32)).astype(np.uint16)
    opp = O.rgb2opp(rgb, 16, 0, False)
    R, G, B = (rgb.astype(np.float64) / 65535.0)
    assert np.allclose(opp[0], (R + G + B) / 3, atol=1e-6)
    assert np.allclose(opp[1], (R - B) / 2, atol=1e-6)
    assert np.allclose(opp[2], (R - 2 * G + B) / 4, atol=1e-6)
    back = O.opp2rgb(opp, 0, 16, True)
    assert np.abs(back.astype(int) - rgb.astype(int)).max() <= 1
    opp16 = O.rgb2opp(rgb, 16, 16, True)
    back16 = O.opp2rgb(opp16, 16, 16, True)
    assert np.abs(back16.astype(int) - rgb.astype(int)).max() <= 2


# ----------------------------------------------------------------- brute force on tiny frames

def _grid(dim, B, s):
    L = dim - B
    return sorted(set(list(range(0, L + 1, s)) + [L]))


def _bm3d_basic_f64(img, f, matches):
    """float64 BM3D basic estimate from given match lists: gather, scipy
    DCT-II over (z, y, x), hard threshold with the reference's table, inverse,
    weights, aggregation, divide (BM3D_Basic.cpp:72-148, BM3D_Base.cpp:290-352)."""
    B = f.p.block_size
    H, W = img.shape
    num = np.zeros((H, W))
    den = np.zeros((H, W))
    kept_total = []
    for (j, i), (ys, xs) in matches.items():
        n = len(ys)
        g = np.stack([img[y:y + B, x:x + B] for y, x in zip(ys, xs)]).astype(np.float64)
        c = sf.dctn(g, type=2)
        thr = np.zeros_like(c)
        for z in range(n):
            for y in range(B):
                for x in range(B):
                    thr[z, y, x] = f.thr[0][n - 1][(x == 0) + (y == 0) + (z == 0)]
        keep = np.abs(c) > thr
        c[~keep] = 0
        r = int(keep.sum())
        kept_total.append(r)
        inv = sf.idctn(c, type=2) * (2 * n) * (2 * B) * (2 * B)  # unnormalised REDFT01
        inv = sf.dctn(c, type=3)
        dw = 1.0 if r < 1 else 1.0 / r
        nw = dw / f.final_amp[n - 1]
        for z, (y, x) in enumerate(zip(ys, xs)):
            num[y:y + B, x:x + B] += inv[z] * nw
            den[y:y + B, x:x + B] += dw
    return num / den


def test_basic_g1_is_sliding_dct_threshold(oracle_lib):
    """group_size=1: no matching, each reference block is its own group
    (BM3D_Base.cpp:478-481) -- a sliding 2D-DCT hard threshold computable
    independently with scipy."""
    noisy, _ = synth.gray_float(40, 48, 20.0, seed=11)
    p = O.resolve("Basic", width=48, height=40, color_family="gray", profile="fast", group_size=1, block_step=3)
    f = O.make_filter(p)
    got = O.bm3d_kernel(f, noisy[None])[0]
    B = p.block_size
    matches = {(j, i): ([j], [i]) for j in _grid(40, B, 3) for i in _grid(48, B, 3)}
    want = _bm3d_basic_f64(noisy, f, matches)
    assert np.abs(got - want).max() < 2e-5


def test_basic_matches_float64_brute_force(oracle_lib):
    """Full Basic on a tiny frame: the oracle's match lists are checked
    against float64 SSDs (same member sets away from ties/threshold), then
    filtering and aggregation against the float64 restatement."""
    H, W = 32, 40
    noisy, _ = synth.gray_float(H, W, 15.0, seed=12)
    p = O.resolve("Basic", width=W, height=H, color_family="gray", profile="np", bm_range=6, block_step=4)
    f = O.make_filter(p)
    B, G = p.block_size, p.group_size
    img64 = noisy.astype(np.float64)
    th = p.th_mse * B * B / 65025.0
    matches = {}
    for j in _grid(H, B, 4):
        for i in _grid(W, B, 4):
            k, ys, xs = O.block_match(f, noisy, j, i)
            assert ys[0] == j and xs[0] == i and k[0] == 0
            assert np.all(np.diff(k[1:]) >= 0)
            # float64 scores of every window position
            ref = img64[j:j + B, i:i + B]
            cand = []
            for y in range(max(0, j - 6), min(H - B, j + 6) + 1):
                for x in range(max(0, i - 6), min(W - B, i + 6) + 1):
                    if (y, x) == (j, i):
                        continue
                    d = ((img64[y:y + B, x:x + B] - ref) ** 2).sum()
                    if 0 < d <= th:
                        cand.append((d, y, x))
            cand.sort()
            want = cand[:G - 1]
            if len(want) == G - 1 and len(cand) > G - 1 and abs(cand[G - 1][0] - cand[G - 2][0]) < 1e-6:
                continue  # near-tie at the cut
            got = set(zip(ys[1:].tolist(), xs[1:].tolist()))
            assert got == set((y, x) for _, y, x in want), (j, i)
            matches[(j, i)] = (ys.tolist(), xs.tolist())
    assert len(matches) >= 0.9 * len(_grid(H, B, 4)) * len(_grid(W, B, 4))
    # filtering/aggregation with the oracle's lists for every block
    all_matches = {}
    for j in _grid(H, B, 4):
        for i in _grid(W, B, 4):
            _, ys, xs = O.block_match(f, noisy, j, i)
            all_matches[(j, i)] = (ys.tolist(), xs.tolist())
    got = O.bm3d_kernel(f, noisy[None])[0]
    want = _bm3d_basic_f64(noisy, f, all_matches)
    # float32 DCT rounding can flip a coefficient sitting on the threshold;
    # the bulk must agree to float precision
    err = np.abs(got - want)
    assert np.quantile(err, 0.99) < 1e-5
    assert err.max() < 5e-3


def test_final_matches_float64_with_host_rcp(oracle_lib):
    """Final on a tiny frame against float64 Wiener filtering that uses the
    same reciprocal (w = r^2 * rcp(r^2 + s^2)), BM3D_Final.cpp:50-127."""
    H, W = 32, 32
    noisy, clean = synth.gray_float(H, W, 10.0, seed=13)
    p = O.resolve("Final", width=W, height=H, color_family="gray", profile="fast")
    f = O.make_filter(p)
    basic = clean  # any reasonable guide image
    got = O.bm3d_kernel(f, noisy[None], basic[None])[0]
    B = p.block_size
    num = np.zeros((H, W))
    den = np.zeros((H, W))
    for j in _grid(H, B, p.block_step):
        for i in _grid(W, B, p.block_step):
            _, ys, xs = O.block_match(f, basic, j, i)
            n = len(ys)
            gs = sf.dctn(np.stack([noisy[y:y + B, x:x + B] for y, x in zip(ys, xs)]).astype(np.float64), type=2)
            gr = sf.dctn(np.stack([basic[y:y + B, x:x + B] for y, x in zip(ys, xs)]).astype(np.float64), type=2)
            s2 = f.wiener_sigma_sqr[0][n - 1]
            r2 = (gr.astype(np.float32) ** 2)
            rcp = np.array([O.lib().or_rcp(O.C.byref(f), float(v)) for v in (r2 + np.float32(s2)).ravel()])
            w = r2.ravel() * rcp
            l2 = max(float((w.astype(np.float64) ** 2).sum()), 1.1920929e-07)
            inv = sf.dctn((gs.ravel() * w).reshape(gs.shape), type=3)
            dw = 1.0 / l2
            nw = dw / f.final_amp[n - 1]
            for z, (y, x) in enumerate(zip(ys, xs)):
                num[y:y + B, x:x + B] += inv[z] * nw
                den[y:y + B, x:x + B] += dw
    want = num / den
    assert np.abs(got - want).max() < 2e-5


def test_rcp_native_equals_table_model(oracle_lib):
    """The reciprocal the Wiener gain uses is a function of the leading
    mantissa bits scaled by the exponent on this host (the GPU library relies
    on that model; it probes it again at init)."""
    p = O.resolve("Final", width=64, height=64, color_family="gray", profile="fast")
    fn = O.make_filter(p)
    xs = np.float32(1.0) + np.arange(0, 1 << 23, 997, dtype=np.float64) / (1 << 23)
    xs = xs.astype(np.float32)
    native = np.array([O.lib().or_rcp(O.C.byref(fn), float(v)) for v in xs], np.float32)
    rel = np.abs(native.astype(np.float64) * xs - 1)
    assert rel.max() < 1.5 * 2 ** -12 + 1e-9
    for e in (-20, -3, 5, 17):
        sc = np.array([O.lib().or_rcp(O.C.byref(fn), float(np.ldexp(v, e))) for v in xs[:200]], np.float32)
        assert np.array_equal(sc, np.ldexp(native[:200], -e).astype(np.float32))


def test_tie_orders_differ_only_at_equal_keys(oracle_lib):
    """The GPU orders equal keys by search index (TIES_INDEX); the reference's
    libstdc++ partial_sort may order / cut them differently.  On the
    640x480 Basic 'fast' workload every difference between the two match
    lists must be explained by equal keys."""
    H, W = 480, 640
    noisy, _ = synth.gray_float(H, W, 10.0, seed=1000)
    p = O.resolve("Basic", width=W, height=H, color_family="gray", profile="fast")
    f_ref = O.make_filter(p, tie_mode=O.TIES_LIBSTDCXX)
    f_idx = O.make_filter(p, tie_mode=O.TIES_INDEX)
    B, s = p.block_size, p.block_step
    differ = 0
    for j in _grid(H, B, s)[::3]:
        for i in _grid(W, B, s):
            k1, y1, x1 = O.block_match(f_ref, noisy, j, i)
            k2, y2, x2 = O.block_match(f_idx, noisy, j, i)
            assert np.array_equal(k1, k2)          # same key sequence always
            if not (np.array_equal(y1, y2) and np.array_equal(x1, x2)):
                differ += 1
                for q in np.nonzero((y1 != y2) | (x1 != x2))[0]:
                    assert np.sum(k1 == k1[q]) > 1  # position differs only inside a run of equal keys
    assert differ < 50


@pytest.mark.parametrize("fn,profile,planes", [("Basic", "np", 3), ("Final", "np", 3), ("Basic", "fast", 1),
                                               ("Final", "lc", 1)])
def test_threaded_kernel_equals_serial(oracle_lib, fn, profile, planes):
    """or_bm3d_kernel_mt (bands of block rows over threads, aggregation kept
    in scan order) gives the serial kernel's result bit for bit."""
    h, w = 88, 120
    rgb, _ = synth.rgb16(h, w, 10.0, seed=77)
    src = O.rgb2opp(rgb, 16, 0, False).astype(np.float32)[:planes]
    fam = "yuv" if planes == 3 else "gray"
    p = O.resolve(fn, width=w, height=h, color_family=fam, profile=profile, matrix=100 if planes == 3 else None)
    f = O.make_filter(p, tie_mode=O.TIES_INDEX)
    ref = np.clip(src * 0.95 + 0.02, -0.5, 1.0).astype(np.float32)
    a = O.bm3d_kernel(f, src, ref)
    for nt in (2, 5):
        b = O.bm3d_kernel(f, src, ref, nthreads=nt)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), nt


def test_final_scalar_matches_float64_exact_division(oracle_lib):
    """The scalar build's Final (BM3D_Final.cpp:106-112: w = r^2 / (r^2 + s^2),
    one sequential L2 sum) on a tiny frame against float64 Wiener filtering
    with exact division; row-major SSD keys (Block.h:632-643) drive the
    matching of both."""
    H, W = 32, 32
    noisy, clean = synth.gray_float(H, W, 10.0, seed=14)
    p = O.resolve("Final", width=W, height=H, color_family="gray", profile="fast")
    f = O.make_filter(p, arith=O.ARITH_SCALAR)
    got = O.bm3d_kernel(f, noisy[None], clean[None])[0]
    B = p.block_size
    num = np.zeros((H, W))
    den = np.zeros((H, W))
    for j in _grid(H, B, p.block_step):
        for i in _grid(W, B, p.block_step):
            _, ys, xs = O.block_match(f, clean, j, i)
            n = len(ys)
            gs = sf.dctn(np.stack([noisy[y:y + B, x:x + B] for y, x in zip(ys, xs)]).astype(np.float64), type=2)
            gr = sf.dctn(np.stack([clean[y:y + B, x:x + B] for y, x in zip(ys, xs)]).astype(np.float64), type=2)
            r2 = gr ** 2
            w = r2 / (r2 + f.wiener_sigma_sqr[0][n - 1])
            dw = 1.0 / max(float((w ** 2).sum()), 1.1920929e-07)
            inv = sf.dctn(gs * w, type=3)
            for z, (y, x) in enumerate(zip(ys, xs)):
                num[y:y + B, x:x + B] += inv[z] * dw / f.final_amp[n - 1]
                den[y:y + B, x:x + B] += dw
    assert np.abs(got - num / den).max() < 2e-5
    # and it is a different arithmetic from the SSE2 build's
    sse = O.bm3d_kernel(O.make_filter(p), noisy[None], clean[None])[0]
    assert (sse.view(np.uint32) != got.view(np.uint32)).any()
