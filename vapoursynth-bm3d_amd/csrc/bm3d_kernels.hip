// HIP kernels for the BM3D / V-BM3D filtering path on CDNA4 (gfx950).
//
// Every floating-point operation below is ordered to reproduce the reference
// plugin's arithmetic exactly (the host build compiles this file with
// -ffp-contract=off, so a*b+c is never fused unless __builtin_fmaf is
// written).  Kernels are wave64-native: one wavefront per reference block
// (matching), per group (transforms) or per 8x8 output tile (aggregation),
// with cross-lane ballots instead of shared-memory reductions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "bm3d_params.h"
#include "bm3d_device.h"
#include "bm3d_kernels.h"
#include "bm3d_dct_coefs.h"

namespace bm3d {

// ---------------------------------------------------------------------------
// K1+K2: full-search block matching, one wave per reference block
// ---------------------------------------------------------------------------
//
// Block::BlockMatchingMulti(range) (include/Block.h:683-729) with
// BM3D_Process_Base::BlockMatching (source/BM3D_Base.cpp:473-490): the
// lattice of the search window (row-major, self excluded) is scored, a
// candidate is accepted iff dist <= thSSE && dist != 0, key =
// float(double(dist) * distMul), and the G-1 best follow the self entry.
// Output per block b: count[b], pos[b*G + z] = (y << 16) | x, key[b*G + z].

constexpr int kMatchWaves = 4;

// Reference block b (the body of k_match and k_match_list); smem: the
// workgroup's LDS, this wave's part at wave * wave_words.
template <int BT>
__device__ __forceinline__ void match_block(const float* __restrict__ plane, const StageParams& p, int b,
                                            int max_cand, int wave_words, uint32_t* smem,
                                            int* __restrict__ out_count, int* __restrict__ out_pos,
                                            float* __restrict__ out_key) {
    const int wave = uniform(threadIdx.x >> 6);  // wave-uniform: scalar registers
    const int lane = lane_id();
    const int B = BT > 0 ? BT : p.block;
    const int G = p.group;
    // per-wave LDS: candidate keys, selection scratch (2G keys, 2G indices),
    // and the reference block when B is not a compile-time constant
    uint32_t* keys = smem + (size_t)wave * wave_words;
    uint32_t* sel_key = keys + max_cand;
    int* sel_idx = reinterpret_cast<int*>(sel_key + 2 * G);
    float* rbl = reinterpret_cast<float*>(sel_idx + 2 * G);

    const int bj = b / p.nx, bi = b - bj * p.nx;
    const int j = grid_pos(bj, p.block_step, p.height - B);
    const int i = grid_pos(bi, p.block_step, p.width - B);

    int* cnt_out = out_count + b;
    int* pos_out = out_pos + (size_t)b * G;
    float* key_out = out_key + (size_t)b * G;
    if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) {
        // BlockMatching short-cut: the group is the reference block alone
        if (lane == 0) {
            *cnt_out = 1;
            pos_out[0] = (j << 16) | i;
            key_out[0] = 0.0f;
        }
        return;
    }

    // reference block
    float rb_reg[BT > 0 ? BT * BT : 1];
    const float* rb = rbl;
    if constexpr (BT > 0) {
#pragma unroll
        for (int e = 0; e < BT * BT; ++e) rb_reg[e] = plane[(size_t)(j + e / BT) * p.pitch + i + e % BT];
        rb = rb_reg;
    } else {
        for (int e = lane; e < B * B; e += 64) rbl[e] = plane[(size_t)(j + e / B) * p.pitch + i + e % B];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    const int l = search_lo(i, p.range, p.step);
    const int r = search_hi(i, p.width - B, p.range, p.step);
    const int t = search_lo(j, p.range, p.step);
    const int bt = search_hi(j, p.height - B, p.range, p.step);
    const int ncols = (r - l) / p.step + 1;
    const int nrows = (bt - t) / p.step + 1;
    const int nc = ncols * nrows;

    int naccepted = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        int c = c0 + lane;
        uint32_t kv = 0xFFFFFFFFu;
        if (c < nc) {
            int cy = c / ncols;
            int y = t + cy * p.step;
            int x = l + (c - cy * ncols) * p.step;
            if (!(y == j && x == i)) {
                const float* q = plane + (size_t)y * p.pitch + x;
                float dist;
                if constexpr (BT > 0) dist = p.arith ? ssd_fixed<BT, true>(rb, q, p.pitch) : ssd_fixed<BT>(rb, q, p.pitch);
                else dist = ssd_generic(rb, q, p.pitch, B, p.arith != 0);
                if (dist <= p.th_sse && dist != 0.0f) {
                    float key = (float)((double)dist * p.dist_mul);
                    kv = __float_as_uint(key);
                }
            }
            keys[c] = kv;
        }
        naccepted += __popcll(__ballot(kv != 0xFFFFFFFFu));
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const int k = naccepted < G - 1 ? naccepted : G - 1;
    wave_select(keys, nc, naccepted, k, sel_key, sel_idx);
    if (lane == 0) {
        *cnt_out = k + 1;
        pos_out[0] = (j << 16) | i;
        key_out[0] = 0.0f;
    }
    for (int e = lane; e < k; e += 64) {
        int c = sel_idx[k + e];
        int cy = c / ncols;
        int y = t + cy * p.step;
        int x = l + (c - cy * ncols) * p.step;
        pos_out[1 + e] = (y << 16) | x;
        key_out[1 + e] = __uint_as_float(sel_key[k + e]);
    }
}

template <int BT>
__global__ __launch_bounds__(64 * kMatchWaves) void k_match(const float* __restrict__ plane, StageParams p,
                                                            int b0, int nb, int max_cand, int wave_words,
                                                            int* __restrict__ out_count,
                                                            int* __restrict__ out_pos,
                                                            float* __restrict__ out_key) {
    extern __shared__ uint32_t smem[];
    const int b = b0 + blockIdx.x * kMatchWaves + uniform(threadIdx.x >> 6);
    if (b >= nb) return;
    match_block<BT>(plane, p, b, max_cand, wave_words, smem, out_count, out_pos, out_key);
}

// Sparse re-run over the blocks blist[0 .. *nlist) (its length is known only
// on the device): the waves of a small grid stride over the list.
template <int BT>
__global__ __launch_bounds__(64 * kMatchWaves) void k_match_list(const float* __restrict__ plane, StageParams p,
                                                                 int nb, int max_cand, int wave_words,
                                                                 const int* __restrict__ blist,
                                                                 const int* __restrict__ nlist,
                                                                 int* __restrict__ out_count,
                                                                 int* __restrict__ out_pos,
                                                                 float* __restrict__ out_key) {
    extern __shared__ uint32_t smem[];
    const int n = uniform(*nlist);
    for (int k = blockIdx.x * kMatchWaves + uniform(threadIdx.x >> 6); k < n && k < nb;
         k += gridDim.x * kMatchWaves) {
        match_block<BT>(plane, p, uniform(blist[k]), max_cand, wave_words, smem, out_count, out_pos, out_key);
        wave_sync_lds();
    }
}

// Fast path for B = 8 and bm_step = 1: the search window is staged in LDS
// and each lane scores a horizontal strip of W neighbouring candidates, so a
// window row of W + 7 samples feeds 8W differences.  Lanes take consecutive
// rows of a strip and the window row stride is odd, so the lanes of a wave
// read distinct LDS banks; the reference block is read from LDS by broadcast.
// The per-candidate arithmetic is exactly ssd_fixed<8>'s.  Keys stay in
// registers.
//
// Selection of the k best (key, candidate index) pairs: the k-th smallest of
// the per-lane minimum keys, t0, bounds the k-th smallest key (k lane minima
// are <= t0), so every selected pair has key <= t0.  Those pairs (usually a
// few dozen) are compacted into LDS and the exact threshold is found by a
// radix descent over them alone; if they do not fit, the descent runs over all
// keys in registers.
constexpr int kSelCap = 256;

// Reference block b of one plane (the body of k_match8 and its batched
// form); wbase: the wave's wave_words LDS words.  Layout: the search window
// (rows of (ncols + 7) | 1 samples, then W - 1 padding samples that the last
// strips read past its last row) from word 0 and the reference block in the
// last 64 words; once every SSD is computed, the selection scratch
// (2 kSelCap + 4 G words) reuses the window's words.
template <int KMAX, int W>
__device__ __forceinline__ void match8_block(const float* __restrict__ plane, const StageParams& p, int b,
                                             uint32_t* wbase, int wave_words, int* __restrict__ out_count,
                                             int* __restrict__ out_pos, float* __restrict__ out_key) {
    const int lane = lane_id();
    constexpr int B = 8;
    constexpr int NK = KMAX * W;
    const int G = p.group;
    float* win = reinterpret_cast<float*>(wbase);
    const int bj = b / p.nx, bi = b - bj * p.nx;
    const int j = grid_pos(bj, p.block_step, p.height - B);
    const int i = grid_pos(bi, p.block_step, p.width - B);
    int* cnt_out = out_count + b;
    int* pos_out = out_pos + (size_t)b * G;
    float* key_out = out_key + (size_t)b * G;
    if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) {
        if (lane == 0) {
            *cnt_out = 1;
            pos_out[0] = (j << 16) | i;
            key_out[0] = 0.0f;
        }
        return;
    }
    const int l = search_lo(i, p.range, 1), r = search_hi(i, p.width - B, p.range, 1);
    const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
    const int ncols = r - l + 1, nrows = bt - t + 1;
    const int nstr = (ncols + W - 1) / W;
    const int WW = (ncols + 7) | 1;  // odd: lanes on consecutive rows read distinct banks
    const int WH = nrows + 7;
    {
        // element e = wy * WW + wx, advanced by 64 per step without a
        // division; eight steps' loads in flight before their LDS writes
        const int dq = 64 / WW, dr = 64 - dq * WW;
        const int n = WH * WW + W - 1;
        int wy = lane / WW, wx = lane - wy * WW;
        for (int e0 = 0; e0 < n; e0 += 8 * 64) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 64 * u + lane;
                const int gx = l + wx;
                v[u] = (e < n && wy < WH && gx < p.width) ? plane[(size_t)(t + wy) * p.pitch + gx] : 0.0f;
                wx += dr;
                wy += dq;
                if (wx >= WW) {
                    wx -= WW;
                    ++wy;
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 64 * u + lane;
                if (e < n) win[e] = v[u];
            }
        }
    }
    float* rbs = win + wave_words - 64;                             // [64] reference block
    uint32_t* s_key = wbase;                                        // [kSelCap], after the SSDs
    int* s_idx = reinterpret_cast<int*>(s_key + kSelCap);           // [kSelCap]
    uint32_t* sel_key = reinterpret_cast<uint32_t*>(s_idx + kSelCap);  // [2G]
    int* sel_idx = reinterpret_cast<int*>(sel_key + 2 * G);         // [2G]
    const int oy = j - t, ox = i - l;
    wave_sync_lds();
    rbs[lane] = win[(oy + (lane >> 3)) * WW + ox + (lane & 7)];
    wave_sync_lds();

    const int tasks = nrows * nstr;
    uint32_t keys[NK];
    int cbase[KMAX];  // candidate index of the round's first strip element: cy * ncols + sx
    int naccepted = 0;
#pragma unroll
    for (int tk = 0; tk < KMAX; ++tk) {
        const int task = lane + 64 * tk;
        const bool active = task < tasks;
        const int strip = active ? task / nrows : 0;
        const int cy = active ? task - strip * nrows : 0;
        const int sx = strip * W;
        cbase[tk] = cy * ncols + sx;
        float s[W][4];
#pragma unroll
        for (int k = 0; k < W; ++k)
#pragma unroll
            for (int c = 0; c < 4; ++c) s[k][c] = 0.0f;
        // rows one at a time: each row's loads stay with its arithmetic
#pragma unroll 1
        for (int rr = 0; rr < 8; ++rr) {
            const float4 ra = *reinterpret_cast<const float4*>(rbs + rr * 8);
            const float4 rc = *reinterpret_cast<const float4*>(rbs + rr * 8 + 4);
            const float rv[8] = {ra.x, ra.y, ra.z, ra.w, rc.x, rc.y, rc.z, rc.w};
            const float* wr = win + (cy + rr) * WW + sx;
            float row[W + 7];
#pragma unroll
            for (int q = 0; q < W + 7; ++q) row[q] = wr[q];
#pragma unroll
            for (int k = 0; k < W; ++k) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float d = rv[c] - row[k + c];
                    s[k][c] = s[k][c] + d * d;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float d = rv[4 + c] - row[k + 4 + c];
                    s[k][c] = s[k][c] + d * d;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            float dist = 0.0f + (((s[k][0] + s[k][1]) + s[k][2]) + s[k][3]);
            const int cx = sx + k;
            const bool valid = active && cx < ncols && !(cy == oy && cx == ox);
            uint32_t kv = 0xFFFFFFFFu;
            if (valid && dist <= p.th_sse && dist != 0.0f) kv = __float_as_uint((float)((double)dist * p.dist_mul));
            keys[tk * W + k] = kv;
            naccepted += __popcll(__ballot(kv != 0xFFFFFFFFu));
        }
    }
    // candidate index of register slot q = tk * W + k: cy * ncols + cx (row-major lattice order)
    auto cand_idx = [&](int q) { return cbase[q / W] + q % W; };
    // the selection scratch reuses the window's words
    wave_sync_lds();
    const int kk = naccepted < G - 1 ? naccepted : G - 1;
    const bool cut = naccepted > kk;
    int nsel = 0;  // entries written to sel_key/sel_idx[0..kk)
    if (!cut) {
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const bool sel = keys[q] != 0xFFFFFFFFu;
            const uint64_t m = __ballot(sel);
            const int at = nsel + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            if (sel) {
                sel_key[at] = keys[q];
                sel_idx[at] = cand_idx(q);
            }
            nsel += __popcll(m);
        }
    } else {
        // bound: t0 = the kk-th smallest lane minimum rounded up to a multiple
        // of 2^16 minus one (a 16-step descent over the high half): at least
        // kk lane minima are <= t0, so the kk smallest keys all are, and the
        // exact cut is found among the compacted keys below
        uint32_t lm = 0xFFFFFFFFu;
#pragma unroll
        for (int q = 0; q < NK; ++q) lm = keys[q] < lm ? keys[q] : lm;
        uint32_t t0 = 0;
        for (int bit = 31; bit >= 16; --bit) {
            const uint32_t trial = t0 | (1u << bit);
            if (__popcll(__ballot(lm < trial)) < kk) t0 = trial;
        }
        t0 |= 0xFFFFu;
        int M = 0;
        if (t0 != 0xFFFFFFFFu) {
#pragma unroll
            for (int q = 0; q < NK; ++q) M += __popcll(__ballot(keys[q] <= t0));
        }
        if (t0 != 0xFFFFFFFFu && M <= kSelCap) {
            int base = 0;
#pragma unroll
            for (int q = 0; q < NK; ++q) {
                const bool in = keys[q] <= t0;
                const uint64_t m = __ballot(in);
                const int at = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (in) {
                    s_key[at] = keys[q];
                    s_idx[at] = cand_idx(q);
                }
                base += __popcll(m);
            }
            wave_sync_lds();
            // the compacted pairs, SR per lane: as few registers as M needs
            auto pick = [&](auto src) {
                constexpr int SR = decltype(src)::value;
                uint32_t ck[SR];
                int ci[SR];
#pragma unroll
                for (int u = 0; u < SR; ++u) {
                    const int e = lane + 64 * u;
                    ck[u] = e < M ? s_key[e] : 0xFFFFFFFFu;
                    ci[u] = e < M ? s_idx[e] : 0x7FFFFFFF;
                }
                // exact k-th smallest key among the compacted pairs
                uint32_t tv = 0;
                for (int bit = 31; bit >= 0; --bit) {
                    const uint32_t trial = tv | (1u << bit);
                    int cnt = 0;
#pragma unroll
                    for (int u = 0; u < SR; ++u) cnt += __popcll(__ballot(ck[u] < trial));
                    if (cnt < kk) tv = trial;
                }
                int less = 0, eq = 0;
#pragma unroll
                for (int u = 0; u < SR; ++u) {
                    less += __popcll(__ballot(ck[u] < tv));
                    eq += __popcll(__ballot(ck[u] == tv));
                }
                const int need = kk - less;
                int ithr = 0x7FFFFFFF;
                if (eq > need) {
                    // equal keys at the cut: the earliest candidates win
                    int it = 0;
                    for (int bit = 20; bit >= 0; --bit) {
                        const int trial = it | (1 << bit);
                        int cnt = 0;
#pragma unroll
                        for (int u = 0; u < SR; ++u) cnt += __popcll(__ballot(ck[u] == tv && ci[u] < trial));
                        if (cnt < need) it = trial;
                    }
                    ithr = it + 1;
                }
#pragma unroll
                for (int u = 0; u < SR; ++u) {
                    const bool sel = ck[u] < tv || (ck[u] == tv && ci[u] < ithr);
                    const uint64_t m = __ballot(sel);
                    const int at = nsel + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                    if (sel) {
                        sel_key[at] = ck[u];
                        sel_idx[at] = ci[u];
                    }
                    nsel += __popcll(m);
                }
            };
            if (M <= 64) pick(std::integral_constant<int, 1>{});
            else if (M <= 128) pick(std::integral_constant<int, 2>{});
            else pick(std::integral_constant<int, kSelCap / 64>{});
        } else {
            // descent over every key in registers
            uint32_t tv = 0;
            for (int bit = 31; bit >= 0; --bit) {
                const uint32_t trial = tv | (1u << bit);
                int cnt = 0;
#pragma unroll
                for (int q = 0; q < NK; ++q) cnt += __popcll(__ballot(keys[q] < trial));
                if (cnt < kk) tv = trial;
            }
            int less = 0, eq = 0;
#pragma unroll
            for (int q = 0; q < NK; ++q) {
                less += __popcll(__ballot(keys[q] < tv));
                eq += __popcll(__ballot(keys[q] == tv));
            }
            const int need = kk - less;
            int ithr = 0x7FFFFFFF;
            if (eq > need) {
                int it = 0;
                for (int bit = 20; bit >= 0; --bit) {
                    const int trial = it | (1 << bit);
                    int cnt = 0;
#pragma unroll
                    for (int q = 0; q < NK; ++q) cnt += __popcll(__ballot(keys[q] == tv && cand_idx(q) < trial));
                    if (cnt < need) it = trial;
                }
                ithr = it + 1;
            }
#pragma unroll
            for (int q = 0; q < NK; ++q) {
                const uint32_t kv = keys[q];
                const int ci = cand_idx(q);
                const bool sel = kv < tv || (kv == tv && ci < ithr);
                const uint64_t m = __ballot(sel);
                const int at = nsel + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (sel) {
                    sel_key[at] = kv;
                    sel_idx[at] = ci;
                }
                nsel += __popcll(m);
            }
        }
    }
    wave_sync_lds();
    rank_sort_selected(sel_key, sel_idx, kk);
    if (lane == 0) {
        *cnt_out = kk + 1;
        pos_out[0] = (j << 16) | i;
        key_out[0] = 0.0f;
    }
    for (int e = lane; e < kk; e += 64) {
        const int c = sel_idx[kk + e];
        const int cy = c / ncols;
        pos_out[1 + e] = ((t + cy) << 16) | (l + c - cy * ncols);
        key_out[1 + e] = __uint_as_float(sel_key[kk + e]);
    }
}

// NW waves (reference blocks) per workgroup.  With blist, wave w of the grid
// takes reference block blist[w] for w < *nlist (a sparse re-run).
template <int KMAX, int W, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_match8(
    const float* __restrict__ plane, StageParams p, int nb, int wave_words, const int* __restrict__ blist,
    const int* __restrict__ nlist, int* __restrict__ out_count, int* __restrict__ out_pos,
    float* __restrict__ out_key) {
    extern __shared__ uint32_t smem[];
    const int wave = uniform(threadIdx.x >> 6);  // wave-uniform: scalar registers
    (void)blist;
    (void)nlist;
    const int b = xcd_item(nb, NW, wave);
    if (b < 0) return;
    match8_block<KMAX, W>(plane, p, b, smem + (size_t)wave * wave_words, wave_words, out_count, out_pos, out_key);
}

// Sparse re-run over the blocks blist[0 .. *nlist): a small grid whose waves
// stride over the list (its length is known only on the device).
template <int KMAX, int W, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_match8_list(
    const float* __restrict__ plane, StageParams p, int nb, int wave_words, const int* __restrict__ blist,
    const int* __restrict__ nlist, int* __restrict__ out_count, int* __restrict__ out_pos,
    float* __restrict__ out_key) {
    extern __shared__ uint32_t smem[];
    const int wave = uniform(threadIdx.x >> 6);
    const int n = uniform(*nlist);
    for (int k = blockIdx.x * NW + wave; k < n && k < nb; k += gridDim.x * NW) {
        match8_block<KMAX, W>(plane, p, uniform(blist[k]), smem + (size_t)wave * wave_words, wave_words,
                              out_count, out_pos, out_key);
        wave_sync_lds();
    }
}

// Sparse re-run over blist[0 .. *nlist) with the four waves of a workgroup
// on one reference block at a time (the blocks are few, so the latency of one
// block, not the throughput, sets the time): the workgroup stages the window,
// wave w scores strip tasks w, w + 4, ... (task = lane + 64 (w + 4 tk)) with
// match8_block's arithmetic, and each wave compacts its candidates at or below
// a bound of its kk-th smallest key (kk = min(accepted, G - 1) over the whole
// block), which holds every one of the block's kk best that the wave scored.
// The workgroup then ranks the compacted pairs by (key, candidate index).
// LDS: max(window + reference block, 4 waves x 64 KC W pairs) + 4 counts.
constexpr int kCoopW = 10;  // strip width of the cooperative re-run
__host__ __device__ constexpr size_t coop_words(int side, int KC, int W) {
    const size_t win = (size_t)(side + 7) * (size_t)((side + 7) | 1) + (size_t)W - 1 + 64;
    const size_t sel = (size_t)4 * 64 * KC * W * 2;
    return (win > sel ? win : sel) + 4;
}

template <int KC, int W>
__global__ __launch_bounds__(256) void k_match8_coop(const float* __restrict__ plane, StageParams p, int nb,
                                                     const int* __restrict__ blist, const int* __restrict__ nlist,
                                                     int* __restrict__ out_count, int* __restrict__ out_pos,
                                                     float* __restrict__ out_key, int words) {
    extern __shared__ uint32_t smem[];
    constexpr int B = 8;
    constexpr int NK = KC * W;
    const int lane = lane_id();
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int tid = (int)threadIdx.x;
    const int G = p.group;
    const int n = uniform(*nlist);
    int* wcnt = reinterpret_cast<int*>(smem + words - 4);  // per wave: accepted, then compacted
    float* win = reinterpret_cast<float*>(smem);
    for (int k = blockIdx.x; k < n && k < nb; k += gridDim.x) {
        const int b = uniform(blist[k]);
        const int bj = b / p.nx, bi = b - bj * p.nx;
        const int j = grid_pos(bj, p.block_step, p.height - B);
        const int i = grid_pos(bi, p.block_step, p.width - B);
        int* cnt_out = out_count + b;
        int* pos_out = out_pos + (size_t)b * G;
        float* key_out = out_key + (size_t)b * G;
        if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) {
            if (tid == 0) {
                *cnt_out = 1;
                pos_out[0] = (j << 16) | i;
                key_out[0] = 0.0f;
            }
            continue;
        }
        const int l = search_lo(i, p.range, 1), r = search_hi(i, p.width - B, p.range, 1);
        const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
        const int ncols = r - l + 1, nrows = bt - t + 1;
        const int nstr = (ncols + W - 1) / W;
        const int WW = (ncols + 7) | 1;
        const int WH = nrows + 7;
        const int nel = WH * WW + W - 1;
        // the window, eight loads in flight per thread
        for (int e0 = 0; e0 < nel; e0 += 8 * 256) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 256 * u + tid;
                const int wy = e / WW, wx = e - wy * WW;
                const int gx = l + wx;
                v[u] = (e < nel && wy < WH && gx < p.width) ? plane[(size_t)(t + wy) * p.pitch + gx] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 256 * u + tid;
                if (e < nel) win[e] = v[u];
            }
        }
        float* rbs = win + nel;  // [64] reference block
        const int oy = j - t, ox = i - l;
        __syncthreads();
        if (wave == 0) rbs[lane] = win[(oy + (lane >> 3)) * WW + ox + (lane & 7)];
        __syncthreads();
        const int tasks = nrows * nstr;
        uint32_t keys[NK];
        int cbase[KC];
        int nacc = 0;
#pragma unroll
        for (int tk = 0; tk < KC; ++tk) {
            const int task = lane + 64 * (wave + 4 * tk);
            const bool active = task < tasks;
            const int strip = active ? task / nrows : 0;
            const int cy = active ? task - strip * nrows : 0;
            const int sx = strip * W;
            cbase[tk] = cy * ncols + sx;
            float s[W][4];
#pragma unroll
            for (int kx = 0; kx < W; ++kx)
#pragma unroll
                for (int c = 0; c < 4; ++c) s[kx][c] = 0.0f;
#pragma unroll 1
            for (int rr = 0; rr < 8; ++rr) {
                const float4 ra = *reinterpret_cast<const float4*>(rbs + rr * 8);
                const float4 rc = *reinterpret_cast<const float4*>(rbs + rr * 8 + 4);
                const float rv[8] = {ra.x, ra.y, ra.z, ra.w, rc.x, rc.y, rc.z, rc.w};
                const float* wr = win + (cy + rr) * WW + sx;
                float row[W + 7];
#pragma unroll
                for (int q = 0; q < W + 7; ++q) row[q] = wr[q];
#pragma unroll
                for (int kx = 0; kx < W; ++kx) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        float d = rv[c] - row[kx + c];
                        s[kx][c] = s[kx][c] + d * d;
                    }
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        float d = rv[4 + c] - row[kx + 4 + c];
                        s[kx][c] = s[kx][c] + d * d;
                    }
                }
            }
#pragma unroll
            for (int kx = 0; kx < W; ++kx) {
                const float dist = 0.0f + (((s[kx][0] + s[kx][1]) + s[kx][2]) + s[kx][3]);
                const int cx = sx + kx;
                const bool valid = active && cx < ncols && !(cy == oy && cx == ox);
                uint32_t kv = 0xFFFFFFFFu;
                if (valid && dist <= p.th_sse && dist != 0.0f) kv = __float_as_uint((float)((double)dist * p.dist_mul));
                keys[tk * W + kx] = kv;
                nacc += __popcll(__ballot(kv != 0xFFFFFFFFu));
            }
        }
        if (lane == 0) wcnt[wave] = nacc;
        __syncthreads();  // every wave is done with the window; the counts are in
        const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        const int kk = total < G - 1 ? total : G - 1;
        // the wave's bound: its kk-th smallest key from the 16 leading bits
        uint32_t tp = 0;
        for (int bit = 30; bit >= 15; --bit) {
            const uint32_t trial = tp | (1u << bit);
            int cnt = 0;
#pragma unroll
            for (int q = 0; q < NK; ++q) cnt += __popcll(__ballot(keys[q] < trial));
            if (cnt < kk) tp = trial;
        }
        const uint32_t tb = tp | 0x7FFFu;
        uint32_t* ck = smem + (size_t)wave * (64 * NK * 2);
        int* ci = reinterpret_cast<int*>(ck + 64 * NK);
        int m = 0;
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const bool in = keys[q] <= tb && keys[q] != 0xFFFFFFFFu;
            const uint64_t bm = __ballot(in);
            if (in) {
                const int at = m + __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0));
                ck[at] = keys[q];
                ci[at] = cbase[q / W] + q % W;
            }
            m += __popcll(bm);
        }
        __syncthreads();  // the accepted counts have been read
        if (lane == 0) wcnt[wave] = m;
        __syncthreads();
        const int m0 = wcnt[0], m1 = wcnt[1], m2 = wcnt[2], m3 = wcnt[3];
        const int mt = m0 + m1 + m2 + m3;
        auto entry = [&](int e, uint32_t& kv, int& iv) {
            int w = 0;
            if (e >= m0) { e -= m0; w = 1;
                if (e >= m1) { e -= m1; w = 2;
                    if (e >= m2) { e -= m2; w = 3; } } }
            const uint32_t* kb = smem + (size_t)w * (64 * NK * 2);
            kv = kb[e];
            iv = reinterpret_cast<const int*>(kb + 64 * NK)[e];
        };
        for (int e1 = tid; e1 < mt; e1 += 256) {
            uint32_t k1;
            int i1;
            entry(e1, k1, i1);
            int rank = 0;
            for (int w = 0; w < 4; ++w) {
                const uint32_t* kb = smem + (size_t)w * (64 * NK * 2);
                const int* ib = reinterpret_cast<const int*>(kb + 64 * NK);
                const int mw = w == 0 ? m0 : w == 1 ? m1 : w == 2 ? m2 : m3;
                for (int e2 = 0; e2 < mw; ++e2) {
                    const uint32_t k2 = kb[e2];
                    rank += (k2 < k1) || (k2 == k1 && ib[e2] < i1);
                }
            }
            if (rank < kk) {
                const int cy = i1 / ncols;
                pos_out[1 + rank] = ((t + cy) << 16) | (l + i1 - cy * ncols);
                key_out[1 + rank] = __uint_as_float(k1);
            }
        }
        if (tid == 0) {
            *cnt_out = kk + 1;
            pos_out[0] = (j << 16) | i;
            key_out[0] = 0.0f;
        }
        __syncthreads();  // LDS is rewritten for the next block
    }
}

// The same search for a per-lane index.
__device__ __forceinline__ int lane_frame(const int* __restrict__ starts, int nframes, int idx) {
    int lo = 0, hi = nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= idx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Batched form: the reference blocks of every frame of a batch, numbered
// frame by frame, in one launch (one wave per block).
template <int KMAX, int W, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_match8_batch(
    const BatchFrame* __restrict__ frames, const int* __restrict__ starts, int nframes, int total, int wave_words,
    int G, int* __restrict__ count, int* __restrict__ pos, float* __restrict__ key) {
    extern __shared__ uint32_t smem[];
    const int wave = uniform(threadIdx.x >> 6);
    const int gb = xcd_item(total, NW, wave);
    if (gb < 0) return;
    const int f = batch_frame(starts, nframes, gb);
    const StageParams p = frames[f].p;
    const float* plane = frames[f].plane;
    const int s0 = uniform(starts[f]);
    match8_block<KMAX, W>(plane, p, gb - s0, smem + (size_t)wave * wave_words, wave_words, count + s0,
                          pos + (size_t)s0 * G, key + (size_t)s0 * G);
}

// ---------------------------------------------------------------------------
// K3-K7: gather, 3D DCT, shrinkage, inverse 3D DCT, weights
// ---------------------------------------------------------------------------
//
// Transform definition (identical chains in the CPU oracle): unnormalised
// DCT-II (FFTW REDFT10) forward and DCT-III (REDFT01) inverse along x, then
// y, then z (inverse: z, y, x), each evaluated with one even/odd split whose
// sums are fused multiply-add chains in increasing index order.
// coef(N, k, j) = float(2 cos(pi k (2j+1) / 2N)) lives at coef[off(N) + k*N + j].

__device__ __forceinline__ float cget(const float* __restrict__ coef, const int* __restrict__ coff, int N, int k,
                                      int j) {
    return coef[coff[N] + k * N + j];
}

// Forward 1D DCT-II of a line in LDS (element m at in[m*is]) into out[k*os].
__device__ void dct2_line(const float* in, int is, float* out, int os, int N, const float* __restrict__ coef,
                          const int* __restrict__ coff) {
    const int h = N >> 1;
    const bool odd = N & 1;
    for (int k = 0; k < N; ++k) {
        float acc;
        if ((k & 1) == 0) {
            if (h > 0) {
                acc = cget(coef, coff, N, k, 0) * (in[0] + in[(N - 1) * is]);
                for (int jj = 1; jj < h; ++jj)
                    acc = __builtin_fmaf(cget(coef, coff, N, k, jj), in[jj * is] + in[(N - 1 - jj) * is], acc);
                if (odd) acc = __builtin_fmaf(cget(coef, coff, N, k, h), in[h * is], acc);
            } else {
                acc = cget(coef, coff, N, k, 0) * in[0];
            }
        } else {
            acc = cget(coef, coff, N, k, 0) * (in[0] - in[(N - 1) * is]);
            for (int jj = 1; jj < h; ++jj)
                acc = __builtin_fmaf(cget(coef, coff, N, k, jj), in[jj * is] - in[(N - 1 - jj) * is], acc);
        }
        out[k * os] = acc;
    }
}

// Inverse 1D DCT-III of a line in LDS into out.
__device__ void dct3_line(const float* in, int is, float* out, int os, int N, const float* __restrict__ coef,
                          const int* __restrict__ coff) {
    const int h = N >> 1;
    const bool odd = N & 1;
    for (int jj = 0; jj < h + (odd ? 1 : 0); ++jj) {
        float e = in[0];
        for (int k = 2; k < N; k += 2) e = __builtin_fmaf(cget(coef, coff, N, k, jj), in[k * is], e);
        float o = 0.0f;
        if (N > 1 && jj < h) {
            o = cget(coef, coff, N, 1, jj) * in[is];
            for (int k = 3; k < N; k += 2) o = __builtin_fmaf(cget(coef, coff, N, k, jj), in[k * is], o);
        }
        if (jj < h) {
            out[jj * os] = e + o;
            out[(N - 1 - jj) * os] = e - o;
        } else {
            out[jj * os] = e;
        }
    }
}

// Forward (dir=+1) or inverse (dir=-1) separable 3D transform of the group
// g[z][y][x] (row stride B+1, slice stride B*(B+1)) using scratch t of the
// same shape.  The result ends up back in g.
__device__ void transform3d(float* g, float* t, int n, int B, bool forward, const float* __restrict__ coef,
                            const int* __restrict__ coff) {
    const int lane = lane_id();
    const int rs = B + 1, ss = B * (B + 1);
    if (forward) {
        for (int L = lane; L < n * B; L += 64) {  // x: lines (z, y)
            int z = L / B, y = L - z * B;
            dct2_line(g + z * ss + y * rs, 1, t + z * ss + y * rs, 1, B, coef, coff);
        }
        wave_sync_lds();
        for (int L = lane; L < n * B; L += 64) {  // y: lines (z, x)
            int z = L / B, x = L - z * B;
            dct2_line(t + z * ss + x, rs, g + z * ss + x, rs, B, coef, coff);
        }
        wave_sync_lds();
        for (int L = lane; L < B * B; L += 64) {  // z: lines (y, x)
            int y = L / B, x = L - y * B;
            dct2_line(g + y * rs + x, ss, t + y * rs + x, ss, n, coef, coff);
        }
        wave_sync_lds();
        for (int e = lane; e < n * ss; e += 64) g[e] = t[e];
        wave_sync_lds();
    } else {
        for (int L = lane; L < B * B; L += 64) {
            int y = L / B, x = L - y * B;
            dct3_line(g + y * rs + x, ss, t + y * rs + x, ss, n, coef, coff);
        }
        wave_sync_lds();
        for (int L = lane; L < n * B; L += 64) {
            int z = L / B, x = L - z * B;
            dct3_line(t + z * ss + x, rs, g + z * ss + x, rs, B, coef, coff);
        }
        wave_sync_lds();
        for (int L = lane; L < n * B; L += 64) {
            int z = L / B, y = L - z * B;
            dct3_line(g + z * ss + y * rs, 1, t + z * ss + y * rs, 1, B, coef, coff);
        }
        wave_sync_lds();
        for (int e = lane; e < n * ss; e += 64) g[e] = t[e];
        wave_sync_lds();
    }
}

// Collaborative filtering of one group on one plane (generic B and n):
// BM3D_Basic_Process::CollaborativeFilter (source/BM3D_Basic.cpp:72-148) and
// BM3D_Final_Process::CollaborativeFilter (source/BM3D_Final.cpp:50-127).
// Writes contrib[z][y][x] = float(c * numW) and denw.
__device__ void filter_group(const float* const* src_planes, const float* const* ref_planes, int pitch,
                             const int* zyx_pos, int n, int B, bool wiener, const PlaneTables& tb,
                             const float* __restrict__ rcp_tbl, int rcp_shift, const float* __restrict__ coef,
                             const int* __restrict__ coff, float* gs, float* gr, float* tmp,
                             float* __restrict__ contrib, float* __restrict__ denw_out, bool scalar) {
    const int lane = lane_id();
    const int rs = B + 1, ss = B * (B + 1);
    const int BB = B * B;
    // gather (BlockGroup::From, include/Block.h:1126-1168)
    for (int e = lane; e < n * BB; e += 64) {
        int z = e / BB, yx = e - z * BB, y = yx / B, x = yx - y * B;
        int pz = zyx_pos[3 * z], py = zyx_pos[3 * z + 1], px = zyx_pos[3 * z + 2];
        gs[z * ss + y * rs + x] = src_planes[pz][(size_t)(py + y) * pitch + px + x];
        if (wiener) gr[z * ss + y * rs + x] = ref_planes[pz][(size_t)(py + y) * pitch + px + x];
    }
    wave_sync_lds();
    transform3d(gs, tmp, n, B, true, coef, coff);
    float num_w, den_w;
    if (!wiener) {
        // hard threshold: keep iff |c| > thr (BM3D_Basic.cpp:94-134)
        int kept = 0;
        for (int e = lane; e < n * BB; e += 64) {
            int z = e / BB, yx = e - z * BB, y = yx / B, x = yx - y * B;
            int flag = (x == 0) + (y == 0) + (z == 0);
            float thr = tb.thr[n - 1][flag];
            float* c = gs + z * ss + y * rs + x;
            float s = *c;
            if (s > thr || s < -thr) ++kept;
            else *c = 0.0f;
        }
        for (int off = 32; off >= 1; off >>= 1) kept += __shfl_xor(kept, off);
        den_w = kept < 1 ? 1.0f : 1.0f / (float)kept;
    } else {
        transform3d(gr, tmp, n, B, true, coef, coff);
        // Wiener gain (BM3D_Final.cpp:74-112): flat index e of group[z][y][x];
        // the SSE2 path covers e < size - size%4 with w = r^2 * rcp(r^2 + s^2)
        // and 4 lane sums of w^2 (lane = e % 4); the tail divides exactly.
        // The scalar build divides exactly everywhere with one sequential sum.
        const int size = n * BB;
        const int sw = scalar ? 0 : size - size % 4;
        const float sig2 = tb.wiener_sigma_sqr[n - 1];
        for (int e = lane; e < size; e += 64) {
            int z = e / BB, yx = e - z * BB, y = yx / B, x = yx - y * B;
            int o = z * ss + y * rs + x;
            float r = gr[o];
            float r2 = r * r;
            float w = e < sw ? r2 * rcp_host(rcp_tbl, rcp_shift, r2 + sig2) : r2 / (r2 + sig2);
            gs[o] = gs[o] * w;
            tmp[e] = w * w;  // flat order for the ordered sums below
        }
        wave_sync_lds();
        float part = 0.0f;
        if (lane < 4) {
            for (int e = lane; e < sw; e += 4) part = part + tmp[e];
        }
        float p0 = __shfl(part, 0), p1 = __shfl(part, 1), p2 = __shfl(part, 2), p3 = __shfl(part, 3);
        float l2 = 0.0f;
        if (sw > 0) l2 = l2 + (((p0 + p1) + p2) + p3);
        for (int e = sw; e < size; ++e) l2 = l2 + tmp[e];
        wave_sync_lds();
        l2 = l2 > 1.1920928955078125e-07f ? l2 : 1.1920928955078125e-07f;  // max(FLT_EPSILON, L2)
        den_w = 1.0f / l2;
    }
    transform3d(gs, tmp, n, B, false, coef, coff);
    num_w = (float)((double)den_w / tb.final_amp[n - 1]);
    for (int e = lane; e < n * BB; e += 64) {
        int z = e / BB, yx = e - z * BB, y = yx / B, x = yx - y * B;
        contrib[e] = gs[z * ss + y * rs + x] * num_w;
    }
    if (lane == 0) *denw_out = den_w;
    wave_sync_lds();
}

// One wave per group; all processed planes of the group in turn.
__global__ __launch_bounds__(64) void k_filter(FilterArgs a) {
    extern __shared__ float fsm[];
    const int g = a.b0 + blockIdx.x;
    if (g >= a.nb) return;
    const int lane = lane_id();
    const int B = a.p.block, G = a.p.group;
    const int ss = B * (B + 1);
    const int n = uniform(a.count[g] < G ? a.count[g] : G);
    float* gs = fsm;
    float* gr = gs + G * ss;
    float* tmp = gr + G * ss;
    int* zyx = reinterpret_cast<int*>(tmp + G * ss);
    for (int z = lane; z < n; z += 64) {
        int pk = a.pos[(size_t)g * G + z];
        zyx[3 * z] = a.frame ? a.frame[(size_t)g * G + z] : 0;
        zyx[3 * z + 1] = pk >> 16;
        zyx[3 * z + 2] = pk & 0xffff;
    }
    wave_sync_lds();
    for (int pl = 0; pl < 3; ++pl) {
        if (!a.process[pl]) continue;
        filter_group(a.src[pl], a.ref[pl], a.p.pitch, zyx, n, B, a.wiener, a.tables[pl], a.rcp_table, a.rcp_shift, a.coef,
                     a.coef_off, gs, gr, tmp, a.contrib[pl] + (size_t)g * G * B * B, a.denw[pl] + g, a.p.arith != 0);
    }
}

// Fast path for B = 8, G <= GMAX: the group sits in LDS for the x and y
// passes (one lane per line, 8-point even/odd chains with coefficients from
// constant memory) while the z pass, the shrinkage and the inverse z pass run
// in registers with one lane per (y, x) column.  Every chain is the one
// dct2_line / dct3_line evaluates, so results equal the generic path's.
__constant__ float c_dct8[64];  // float(2 cos(pi k (2j+1) / 16)) at [k * 8 + j]
// Length-N coefficients for N = 1..32 at c_dctz[zoff(N) + k * N + j]; constant
// memory so that the wave-uniform reads of the z pass are scalar loads.
constexpr int kZMax = 32;
__constant__ float c_dctz[kZMax * (kZMax + 1) * (2 * kZMax + 1) / 6];
__host__ __device__ constexpr int zoff(int N) { return (N - 1) * N * (2 * N - 1) / 6; }

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

__device__ __forceinline__ void dct8_fwd(const float* in, float* out) {
    float u[4], v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        u[j] = in[j] + in[7 - j];
        v[j] = in[j] - in[7 - j];
    }
    static_for<0, 8>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const float* w = (k & 1) ? v : u;
        constexpr float c0 = DctC<8>::v[k * 8];
        float acc = c0 * w[0];
        static_for<1, 4>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            constexpr float c = DctC<8>::v[k * 8 + j];
            acc = __builtin_fmaf(c, w[j], acc);
        });
        out[k] = acc;
    });
}

__device__ __forceinline__ void dct8_inv(const float* in, float* out) {
    static_for<0, 4>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        float e = in[0];
        static_for<1, 4>([&](auto kc) {
            constexpr int k = 2 * decltype(kc)::value;
            constexpr float c = DctC<8>::v[k * 8 + j];
            e = __builtin_fmaf(c, in[k], e);
        });
        constexpr float c1 = DctC<8>::v[8 + j];
        float o = c1 * in[1];
        static_for<1, 4>([&](auto kc) {
            constexpr int k = 2 * decltype(kc)::value + 1;
            constexpr float c = DctC<8>::v[k * 8 + j];
            o = __builtin_fmaf(c, in[k], o);
        });
        out[j] = e + o;
        out[7 - j] = e - o;
    });
}

// Sixteen bytes of a row that starts at any float (dword alignment suffices
// for global vector loads on this target).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

template <int N>
struct RowLoads {
    static constexpr int kIters = (N * 8 + 63) / 64;
    f4u a[kIters], b[kIters];
};
template <int N>
__device__ __forceinline__ void rows_load(RowLoads<N>& r, const float* base0, const float* const* planes,
                                          const int* zyx, int pitch) {
    const int lane = lane_id();
    constexpr int NL = N * 8;
#pragma unroll
    for (int i = 0; i < RowLoads<N>::kIters; ++i) {
        const int L = i * 64 + lane;
        if (NL % 64 == 0 || L < NL) {
            const int z = L >> 3, y = L & 7;
            const float* base = base0 ? base0 : planes[zyx[3 * z]];
            const float* row = base + (size_t)(zyx[3 * z + 1] + y) * pitch + zyx[3 * z + 2];
            r.a[i] = *(const __attribute__((address_space(1))) f4u*)row;
            r.b[i] = *(const __attribute__((address_space(1))) f4u*)(row + 4);
        }
    }
}
template <int N>
__device__ __forceinline__ void rows_fwd(float* g, const RowLoads<N>& r) {
    const int lane = lane_id();
    constexpr int NL = N * 8;
#pragma unroll
    for (int i = 0; i < RowLoads<N>::kIters; ++i) {
        const int L = i * 64 + lane;
        if (NL % 64 == 0 || L < NL) {
            const float in[8] = {r.a[i].x, r.a[i].y, r.a[i].z, r.a[i].w, r.b[i].x, r.b[i].y, r.b[i].z, r.b[i].w};
            float out[8];
            dct8_fwd(in, out);
            float* dst = g + L * 9;
#pragma unroll
            for (int q = 0; q < 8; ++q) dst[q] = out[q];
        }
    }
    wave_sync_lds();
}

// y pass (forward or inverse) of the n slices in LDS: lines (z, x).
template <bool FWD>
__device__ __forceinline__ void pass_y(float* g, int n) {
    const int lane = lane_id();
    float in[8], out[8];
    for (int L = lane; L < n * 8; L += 64) {
        float* col = g + (L >> 3) * 72 + (L & 7);
#pragma unroll
        for (int q = 0; q < 8; ++q) in[q] = col[q * 9];
        if (FWD) dct8_fwd(in, out);
        else dct8_inv(in, out);
#pragma unroll
        for (int q = 0; q < 8; ++q) col[q * 9] = out[q];
    }
    wave_sync_lds();
}

// Inverse x pass fused with the contribution stores: line L = (z, y) of the
// group, inverse-transformed in registers, scaled by num_w and written as
// contrib[z * 64 + y * 8 + x] (two aligned 16-byte stores per line).
template <int N>
__device__ __forceinline__ void rows_inv_store(const float* g, float* __restrict__ contrib, float num_w) {
    const int lane = lane_id();
    constexpr int NL = N * 8;
#pragma unroll
    for (int L0 = 0; L0 < NL; L0 += 64) {
        const int L = L0 + lane;
        if (NL % 64 == 0 || L < NL) {
            const float* src = g + L * 9;
            float in[8], out[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) in[q] = src[q];
            dct8_inv(in, out);
            float4* dst = reinterpret_cast<float4*>(contrib + L * 8);
            dst[0] = float4{out[0] * num_w, out[1] * num_w, out[2] * num_w, out[3] * num_w};
            dst[1] = float4{out[4] * num_w, out[5] * num_w, out[6] * num_w, out[7] * num_w};
        }
    }
    wave_sync_lds();
}

// Forward z transform of the lane's column (y, x): g[z*72 + off] -> X[k].
// Forward z transform (length N, compile time) of the lane's column (y, x):
// g[z*72 + off] -> X[k], with the chains of dct2_line.
template <int N>
__device__ __forceinline__ void zfwdN(const float* g, int off, float* X) {
    constexpr int h = N / 2;
    float u[h > 0 ? h : 1], v[h > 0 ? h : 1];
#pragma unroll
    for (int j = 0; j < h; ++j) {
        const float a0 = g[j * 72 + off], b0 = g[(N - 1 - j) * 72 + off];
        u[j] = a0 + b0;
        v[j] = a0 - b0;
    }
    const float mid = (N & 1) ? g[h * 72 + off] : 0.0f;
    static_for<0, N>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        float acc;
        if constexpr ((k & 1) == 0) {
            if constexpr (h > 0) {
                constexpr float c0 = DctC<N>::v[k * N];
                acc = c0 * u[0];
                static_for<1, h>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    constexpr float c = DctC<N>::v[k * N + j];
                    acc = __builtin_fmaf(c, u[j], acc);
                });
                if constexpr (N & 1) {
                    constexpr float cm = DctC<N>::v[k * N + h];
                    acc = __builtin_fmaf(cm, mid, acc);
                }
            } else {
                constexpr float c0 = DctC<N>::v[0];
                acc = c0 * mid;
            }
        } else {
            constexpr float c0 = DctC<N>::v[k * N];
            acc = c0 * v[0];
            static_for<1, h>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                constexpr float c = DctC<N>::v[k * N + j];
                acc = __builtin_fmaf(c, v[j], acc);
            });
        }
        X[k] = acc;
    });
}

// Inverse z transform (chains of dct3_line) of X into the lane's column.
template <int N>
__device__ __forceinline__ void zinvN(float* g, int off, const float* X) {
    constexpr int h = N / 2;
    static_for<0, h + (N & 1)>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        float e = X[0];
        static_for<1, (N + 1) / 2>([&](auto kc) {
            constexpr int k = 2 * decltype(kc)::value;
            constexpr float c = DctC<N>::v[k * N + j];
            e = __builtin_fmaf(c, X[k], e);
        });
        if constexpr (j < h) {
            constexpr float c1 = DctC<N>::v[N + j];
            float o = c1 * X[1];
            static_for<1, N / 2>([&](auto kc) {
                constexpr int k = 2 * decltype(kc)::value + 1;
                constexpr float c = DctC<N>::v[k * N + j];
                o = __builtin_fmaf(c, X[k], o);
            });
            g[j * 72 + off] = e + o;
            g[(N - 1 - j) * 72 + off] = e - o;
        } else {
            g[j * 72 + off] = e;
        }
    });
}

// ONE: only plane 0 is processed (Gray, or YUV with the chroma planes
// passed through), which keeps the other planes' pointers out of registers.
template <int N, bool WIENER, bool ONE, bool SCALAR>
__device__ __forceinline__ void group8(const FilterArgs& a, int g_id, const int* zyx, float* gs) {
    const int lane = lane_id();
    const int G = a.p.group;
    const int y = lane >> 3, x = lane & 7;
    const int off = y * 9 + x;
    const bool single = a.frame == nullptr;
    // Gathers are prefetched: each one's loads are issued while the previous
    // gather is being transformed (hard threshold: the next plane's source
    // rows; Wiener: this plane's source rows during the reference transform,
    // the next plane's reference rows once the Wiener gains are looked up).
    // Global loads complete in order, so everything else the group reads
    // (plane bases, thresholds, Wiener sigma, amplitudes) is read up front: a
    // load issued later would wait for the gather in flight.
    const bool q0 = a.process[0] != 0, q1 = !ONE && a.process[1] != 0, q2 = !ONE && a.process[2] != 0;
    if (!q0 && !q1 && !q2) return;
    const float* sb0 = (single && q0) ? uniform_ptr(a.src[0][0]) : nullptr;
    const float* sb1 = (single && q1) ? uniform_ptr(a.src[1][0]) : nullptr;
    const float* sb2 = (single && q2) ? uniform_ptr(a.src[2][0]) : nullptr;
    const float* rb0 = (WIENER && single && q0) ? uniform_ptr(a.ref[0][0]) : nullptr;
    const float* rb1 = (WIENER && single && q1) ? uniform_ptr(a.ref[1][0]) : nullptr;
    const float* rb2 = (WIENER && single && q2) ? uniform_ptr(a.ref[2][0]) : nullptr;
    auto thr_of = [&](bool q, int pl, int k) { return (!WIENER && q) ? uniformf(a.tables[pl].thr[N - 1][k]) : 0.0f; };
    const float t00 = thr_of(q0, 0, 0), t01 = thr_of(q0, 0, 1), t02 = thr_of(q0, 0, 2), t03 = thr_of(q0, 0, 3);
    const float t10 = thr_of(q1, 1, 0), t11 = thr_of(q1, 1, 1), t12 = thr_of(q1, 1, 2), t13 = thr_of(q1, 1, 3);
    const float t20 = thr_of(q2, 2, 0), t21 = thr_of(q2, 2, 1), t22 = thr_of(q2, 2, 2), t23 = thr_of(q2, 2, 3);
    const float s20 = (WIENER && q0) ? uniformf(a.tables[0].wiener_sigma_sqr[N - 1]) : 0.0f;
    const float s21 = (WIENER && q1) ? uniformf(a.tables[1].wiener_sigma_sqr[N - 1]) : 0.0f;
    const float s22 = (WIENER && q2) ? uniformf(a.tables[2].wiener_sigma_sqr[N - 1]) : 0.0f;
    const double am0 = q0 ? uniformd(a.tables[0].final_amp[N - 1]) : 1.0;
    const double am1 = q1 ? uniformd(a.tables[1].final_amp[N - 1]) : 1.0;
    const double am2 = q2 ? uniformd(a.tables[2].final_amp[N - 1]) : 1.0;
    RowLoads<N> rl;
    // static plane indices throughout (a plane index known only at run time
    // would index the kernel-argument arrays through scratch)
#define BM3D_LOAD_SRC(P) rows_load<N>(rl, sb##P, a.src[P], zyx, a.p.pitch)
#define BM3D_LOAD_REF(P) rows_load<N>(rl, rb##P, a.ref[P], zyx, a.p.pitch)
    if (WIENER) {
        if (q0) BM3D_LOAD_REF(0); else if (q1) BM3D_LOAD_REF(1); else BM3D_LOAD_REF(2);
    } else {
        if (q0) BM3D_LOAD_SRC(0); else if (q1) BM3D_LOAD_SRC(1); else BM3D_LOAD_SRC(2);
    }
#pragma unroll
    for (int pl = 0; pl < (ONE ? 1 : 3); ++pl) {
        if (!(pl == 0 ? q0 : pl == 1 ? q1 : q2)) continue;
        float X[N];
        float den_w;
        if (!WIENER) {
            rows_fwd<N>(gs, rl);
            if (pl == 0) {
                if (q1) BM3D_LOAD_SRC(1); else if (q2) BM3D_LOAD_SRC(2);
            } else if (pl == 1) {
                if (q2) BM3D_LOAD_SRC(2);
            }
            float thr4[4];
            thr4[0] = pl == 0 ? t00 : pl == 1 ? t10 : t20;
            thr4[1] = pl == 0 ? t01 : pl == 1 ? t11 : t21;
            thr4[2] = pl == 0 ? t02 : pl == 1 ? t12 : t22;
            thr4[3] = pl == 0 ? t03 : pl == 1 ? t13 : t23;
            pass_y<true>(gs, N);
            zfwdN<N>(gs, off, X);
            int kept = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const int fl = (x == 0) + (y == 0) + (k == 0);
                const float thr = fl == 0 ? thr4[0] : fl == 1 ? thr4[1] : fl == 2 ? thr4[2] : thr4[3];
                const float s = X[k];
                if (s > thr || s < -thr) ++kept;
                else X[k] = 0.0f;
            }
            for (int o = 32; o >= 1; o >>= 1) kept += __shfl_xor(kept, o);
            den_w = kept < 1 ? 1.0f : 1.0f / (float)kept;
        } else {
            // one LDS buffer: the reference group's transform goes to registers
            // first, then the source group reuses the buffer
            rows_fwd<N>(gs, rl);
            if (pl == 0) BM3D_LOAD_SRC(0); else if (pl == 1) BM3D_LOAD_SRC(1); else BM3D_LOAD_SRC(2);
            pass_y<true>(gs, N);
            float R[N];
            zfwdN<N>(gs, off, R);
            wave_sync_lds();
            rows_fwd<N>(gs, rl);
            pass_y<true>(gs, N);
            zfwdN<N>(gs, off, X);
            wave_sync_lds();
            const float sig2 = pl == 0 ? s20 : pl == 1 ? s21 : s22;
            // w^2 stored chain-major: SSE2 lane c = x % 4 takes the flat
            // indices e = z*64 + y*8 + x in increasing order (BM3D_Final.cpp:80-104)
            float* w2 = gs;
            float l2 = 0.0f;
            if constexpr (SCALAR) {
                // builds without __SSE2__ (BM3D_Final.cpp:106-112): exact
                // division, w^2 summed in flat order e = z*64 + y*8 + x
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const float r2 = R[k] * R[k];
                    const float w = r2 / (r2 + sig2);
                    X[k] = X[k] * w;
                    w2[k * 64 + lane] = w * w;
                }
                wave_sync_lds();
                if (lane == 0)
                    for (int e = 0; e < N * 64; ++e) l2 = l2 + w2[e];
                l2 = __shfl(l2, 0);
            } else {
            const int c = x & 3;
            const int ibase = y * 2 + (x >> 2);
            // chain stride = 16 (mod 64): the four chains of one store start
            // in different banks (with stride 16N they would share banks for
            // N % 4 == 0: a four-way conflict on every store)
            constexpr int kChain = 16 * N + (80 - (16 * N) % 64) % 64;
            static_assert(3 * kChain + 16 * N <= (N <= 8 ? 8 : N <= 16 ? 16 : 32) * 72, "chains fit the group buffer");
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const float r2 = R[k] * R[k];
                const float w = r2 * rcp_host(a.rcp_table, a.rcp_shift, r2 + sig2);
                X[k] = X[k] * w;
                w2[c * kChain + k * 16 + ibase] = w * w;
            }
            wave_sync_lds();
            float part = 0.0f;
            if (lane < 4) {
                const float4* ch = reinterpret_cast<const float4*>(w2 + lane * kChain);
#pragma unroll
                for (int q0 = 0; q0 < 4 * N; q0 += 8) {
                    float4 v4[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (q0 + q < 4 * N) v4[q] = ch[q0 + q];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        if (q0 + q < 4 * N) {
                            part = part + v4[q].x;
                            part = part + v4[q].y;
                            part = part + v4[q].z;
                            part = part + v4[q].w;
                        }
                    }
                }
            }
            const float p0 = __shfl(part, 0), p1 = __shfl(part, 1), p2 = __shfl(part, 2), p3 = __shfl(part, 3);
            l2 = 0.0f + (((p0 + p1) + p2) + p3);
            }
            l2 = l2 > 1.1920928955078125e-07f ? l2 : 1.1920928955078125e-07f;
            den_w = 1.0f / l2;
            wave_sync_lds();
            if (pl == 0) {
                if (q1) BM3D_LOAD_REF(1); else if (q2) BM3D_LOAD_REF(2);
            } else if (pl == 1) {
                if (q2) BM3D_LOAD_REF(2);
            }
        }
        zinvN<N>(gs, off, X);
        wave_sync_lds();
        pass_y<false>(gs, N);
        const float num_w = (float)((double)den_w / (pl == 0 ? am0 : pl == 1 ? am1 : am2));
        rows_inv_store<N>(gs, pick3(a.contrib, pl) + (size_t)g_id * G * 64, num_w);
        if (lane == 0) pick3(a.denw, pl)[g_id] = den_w;
    }
}
#undef BM3D_LOAD_SRC
#undef BM3D_LOAD_REF

// One wave per group, taken from `order` (groups bucketed by size so that
// neighbouring waves run the same specialisation).
// Group g_id (the body of k_filter8 and its batched form).
template <int GMAX, bool WIENER, bool ONE, bool SCALAR>
__device__ __forceinline__ void filter8_group(const FilterArgs& a, int g_id, float* fsm) {
    const int lane = lane_id();
    const int G = a.p.group;
    const int n = uniform(a.count[g_id] < G ? a.count[g_id] : G);
    float* gs = fsm;                       // [GMAX][8][9], later the w^2 chains
    int* zyx = reinterpret_cast<int*>(gs + GMAX * 72);
    for (int z = lane; z < n; z += 64) {
        int pk = a.pos[(size_t)g_id * G + z];
        zyx[3 * z] = a.frame ? a.frame[(size_t)g_id * G + z] : 0;
        zyx[3 * z + 1] = pk >> 16;
        zyx[3 * z + 2] = pk & 0xffff;
    }
    wave_sync_lds();
    switch (n) {
#define BM3D_G8(NN)                                                         \
    case NN:                                                                \
        if constexpr (NN <= GMAX) group8<NN, WIENER, ONE, SCALAR>(a, g_id, zyx, gs); \
        break;
        BM3D_G8(1) BM3D_G8(2) BM3D_G8(3) BM3D_G8(4) BM3D_G8(5) BM3D_G8(6) BM3D_G8(7) BM3D_G8(8)
        BM3D_G8(9) BM3D_G8(10) BM3D_G8(11) BM3D_G8(12) BM3D_G8(13) BM3D_G8(14) BM3D_G8(15) BM3D_G8(16)
        BM3D_G8(17) BM3D_G8(18) BM3D_G8(19) BM3D_G8(20) BM3D_G8(21) BM3D_G8(22) BM3D_G8(23) BM3D_G8(24)
        BM3D_G8(25) BM3D_G8(26) BM3D_G8(27) BM3D_G8(28) BM3D_G8(29) BM3D_G8(30) BM3D_G8(31) BM3D_G8(32)
#undef BM3D_G8
    default:
        break;
    }
}

template <int GMAX, bool WIENER, bool ONE, bool SCALAR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_filter8(FilterArgs a) {
    extern __shared__ float fsm[];
    const int slot = xcd_item(a.nb, 1, 0);
    if (slot < 0) return;
    filter8_group<GMAX, WIENER, ONE, SCALAR>(a, a.order ? a.order[slot] : slot, fsm);
}

// Batched form: the groups of every frame (numbered frame by frame, taken in
// `order`) in one launch, one workgroup per group (XCD-contiguous slots).  a0
// holds what all frames share, with count, pos, contrib and denw indexed by
// the batch-wide group number; ff[f] the frame's pitch, plane pointer table
// (src p0..p2, ref p0..p2) and tables.  Plane 0 only (Gray frames).
template <int GMAX, bool WIENER, bool SCALAR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_filter8_batch(FilterArgs a0, const BatchFilterFrame* __restrict__ ff,
                                                      const int* __restrict__ starts, int nframes,
                                                      const int* __restrict__ order, int total) {
    extern __shared__ float fsm[];
    const int slot = xcd_item(total, 1, 0);
    if (slot < 0) return;
    const int gid = uniform(order[slot]);
    const int f = batch_frame(starts, nframes, gid);
    FilterArgs a = a0;
    a.p.pitch = uniform(ff[f].pitch);
    const float* const* pt = ff[f].planes;
    a.src[0] = pt;
    a.ref[0] = pt + 3;
    a.tables = ff[f].tables;
    filter8_group<GMAX, WIENER, true, SCALAR>(a, gid, fsm);
}

// Bucketing of groups by size, so that neighbouring waves run the same
// group-size specialisation.  Every group writes only its own contributions,
// so the processing order (and the order inside a bucket, which LDS atomics
// make arbitrary) does not affect any result.
__global__ void k_zero_ints(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

constexpr int kBucketThreads = 256;
constexpr int kBucketPerThread = 16;
constexpr int kBucketChunk = kBucketThreads * kBucketPerThread;

// Batched bucketing, frame-major: groups of frame 0 first, inside a frame the
// largest sizes first.  Bin of (frame f, size n): f * (G + 1) + G - n, so an
// exclusive scan in bin order gives every bin's first slot.  A chunk inside
// one frame histograms in LDS; a chunk spanning frames goes to the global
// bins lane by lane.
__global__ __launch_bounds__(kBucketThreads) void k_fsize_hist(const int* __restrict__ count,
                                                               const int* __restrict__ starts, int nframes, int nb,
                                                               int G, int* hist) {
    __shared__ int h[kMaxGroup + 1];
    const int lo = blockIdx.x * kBucketChunk;
    const int hi = min(nb, lo + kBucketChunk);
    const int f0 = batch_frame(starts, nframes, lo), f1 = batch_frame(starts, nframes, hi - 1);
    if (f0 != f1) {
        for (int b = lo + threadIdx.x; b < hi; b += kBucketThreads) {
            const int n = min(max(count[b], 1), G);
            atomicAdd(&hist[lane_frame(starts, nframes, b) * (G + 1) + G - n], 1);
        }
        return;
    }
    for (int i = threadIdx.x; i <= G; i += kBucketThreads) h[i] = 0;
    __syncthreads();
    for (int b = lo + threadIdx.x; b < hi; b += kBucketThreads) atomicAdd(&h[min(max(count[b], 1), G)], 1);
    __syncthreads();
    for (int i = threadIdx.x; i <= G; i += kBucketThreads)
        if (h[i]) atomicAdd(&hist[f0 * (G + 1) + G - i], h[i]);
}

// Exclusive prefix over nbins counters, in place (one workgroup).
__global__ __launch_bounds__(256) void k_bins_scan(int* hist, int nbins) {
    __shared__ int part[256];
    const int per = (nbins + 255) / 256;
    const int a = threadIdx.x * per, e = min(nbins, a + per);
    int sum = 0;
    for (int i = a; i < e; ++i) sum += hist[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < 256; ++i) {
            const int t = part[i];
            part[i] = acc;
            acc += t;
        }
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int i = a; i < e; ++i) {
        const int t = hist[i];
        hist[i] = acc;
        acc += t;
    }
}

__global__ __launch_bounds__(kBucketThreads) void k_fsize_scatter(const int* __restrict__ count,
                                                                  const int* __restrict__ starts, int nframes,
                                                                  int nb, int G, int* cursor, int* order) {
    __shared__ int h[kMaxGroup + 1];
    __shared__ int base[kMaxGroup + 1];
    const int lo = blockIdx.x * kBucketChunk;
    const int hi = min(nb, lo + kBucketChunk);
    const int f0 = batch_frame(starts, nframes, lo), f1 = batch_frame(starts, nframes, hi - 1);
    if (f0 != f1) {
        for (int b = lo + threadIdx.x; b < hi; b += kBucketThreads) {
            const int n = min(max(count[b], 1), G);
            order[atomicAdd(&cursor[lane_frame(starts, nframes, b) * (G + 1) + G - n], 1)] = b;
        }
        return;
    }
    for (int i = threadIdx.x; i <= G; i += kBucketThreads) h[i] = 0;
    __syncthreads();
    int rank[kBucketPerThread], size[kBucketPerThread];
#pragma unroll
    for (int k = 0; k < kBucketPerThread; ++k) {
        const int b = lo + threadIdx.x + k * kBucketThreads;
        size[k] = b < hi ? min(max(count[b], 1), G) : -1;
        rank[k] = size[k] >= 0 ? atomicAdd(&h[size[k]], 1) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= G; i += kBucketThreads)
        base[i] = h[i] ? atomicAdd(&cursor[f0 * (G + 1) + G - i], h[i]) : 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBucketPerThread; ++k)
        if (size[k] >= 1) order[base[size[k]] + rank[k]] = lo + threadIdx.x + k * kBucketThreads;
}

// Batched bucketing, chunk-local: every chunk of kBucketChunk consecutive
// groups is reordered in place by size (largest first), so the order stays
// spatially local.  No global counters.
__global__ __launch_bounds__(kBucketThreads) void k_chunk_bucket(const int* __restrict__ count, int nb, int G,
                                                                 int* order) {
    __shared__ int h[kMaxGroup + 1];
    const int lo = blockIdx.x * kBucketChunk;
    const int hi = min(nb, lo + kBucketChunk);
    for (int i = threadIdx.x; i <= G; i += kBucketThreads) h[i] = 0;
    __syncthreads();
    int rank[kBucketPerThread], size[kBucketPerThread];
#pragma unroll
    for (int k = 0; k < kBucketPerThread; ++k) {
        const int b = lo + threadIdx.x + k * kBucketThreads;
        size[k] = b < hi ? min(max(count[b], 1), G) : -1;
        rank[k] = size[k] >= 0 ? atomicAdd(&h[size[k]], 1) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int n = G; n >= 1; --n) {
            const int t = h[n];
            h[n] = acc;
            acc += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBucketPerThread; ++k)
        if (size[k] >= 1) order[lo + h[size[k]] + rank[k]] = lo + threadIdx.x + k * kBucketThreads;
}

hipError_t launch_bucket(const int* count, int nb, int G, int* order, hipStream_t st) {
    if (nb <= 0) return hipSuccess;
    const int chunks = (nb + kBucketChunk - 1) / kBucketChunk;
    hipLaunchKernelGGL(k_chunk_bucket, dim3(chunks), dim3(kBucketThreads), 0, st, count, nb, G, order);
    return hipGetLastError();
}

hipError_t launch_bucket_batch(const int* count, const int* starts, int nframes, int nb, int G, int mode, int* hist,
                               int* order, hipStream_t st) {
    if (nb <= 0) return hipSuccess;
    const int chunks = (nb + kBucketChunk - 1) / kBucketChunk;
    if (mode == kOrderChunk) {
        hipLaunchKernelGGL(k_chunk_bucket, dim3(chunks), dim3(kBucketThreads), 0, st, count, nb, G, order);
        return hipGetLastError();
    }
    const int nbins = nframes * (G + 1);
    hipLaunchKernelGGL(k_zero_ints, dim3((nbins + 255) / 256), dim3(256), 0, st, hist, nbins);
    hipLaunchKernelGGL(k_fsize_hist, dim3(chunks), dim3(kBucketThreads), 0, st, count, starts, nframes, nb, G, hist);
    hipLaunchKernelGGL(k_bins_scan, dim3(1), dim3(256), 0, st, hist, nbins);
    hipLaunchKernelGGL(k_fsize_scatter, dim3(chunks), dim3(kBucketThreads), 0, st, count, starts, nframes, nb, G, hist,
                       order);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// input conversions and standalone colour filters
// ---------------------------------------------------------------------------

// RangeConvert int/float -> float (Conversion.hpp:36-78) for one plane.
__global__ void k_to_float(const void* src, int bits, size_t spitch, float* dst, size_t dpitch, int W, int H,
                           float gain, float offset, int clip, float lo, float hi) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W || y >= H) return;
    float v = load_sample(src, bits, (size_t)y * spitch + x);
    if (bits != 0 || gain != 1.0f || offset != 0.0f || clip) {
        v = v * gain + offset;
        if (clip) v = clipf(v, lo, hi);
    }
    dst[(size_t)y * dpitch + x] = v;
}

// MatrixConvert_RGB2YUV / ConvertToY, OPP branch (Conversion.hpp:101-115,
// 205-236).  ny_only: write Y only.  dst_bits 0 = float output.
__global__ void k_rgb2opp(RgbArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= a.W || y >= a.H) return;
    size_t si = (size_t)y * a.spitch + x, di = (size_t)y * a.dpitch + x;
    float R = load_sample(a.src[0], a.src_bits, si);
    float G = load_sample(a.src[1], a.src_bits, si);
    float B = load_sample(a.src[2], a.src_bits, si);
    float t = ((R + G) + B) * a.gain[0] + a.offset[0];
    store_sample(a.dst[0], a.dst_bits, di, a.clip ? clipf(t, a.lo[0], a.hi[0]) : t);
    if (a.y_only) return;
    t = (R - B) * a.gain[1] + a.offset[1];
    store_sample(a.dst[1], a.dst_bits, di, a.clip ? clipf(t, a.lo[1], a.hi[1]) : t);
    t = ((R - G * 2.0f) + B) * a.gain[2] + a.offset[1];
    store_sample(a.dst[2], a.dst_bits, di, a.clip ? clipf(t, a.lo[1], a.hi[1]) : t);
}

// MatrixConvert_YUV2RGB, OPP branch (Conversion.hpp:369-392).
__global__ void k_opp2rgb(RgbArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= a.W || y >= a.H) return;
    size_t si = (size_t)y * a.spitch + x, di = (size_t)y * a.dpitch + x;
    float Y = load_sample(a.src[0], a.src_bits, si);
    float U = load_sample(a.src[1], a.src_bits, si);
    float V = load_sample(a.src[2], a.src_bits, si);
    float r = a.m[0] * Y + a.m[1] * U + a.m[2] * V + a.offset[0];
    float g = a.m[3] * Y + a.m[5] * V + a.offset[1];
    float b = a.m[6] * Y + a.m[7] * U + a.m[8] * V + a.offset[2];
    if (a.clip) {
        r = clipf(r, a.lo[0], a.hi[0]);
        g = clipf(g, a.lo[0], a.hi[0]);
        b = clipf(b, a.lo[0], a.hi[0]);
    }
    store_sample(a.dst[0], a.dst_bits, di, r);
    store_sample(a.dst[1], a.dst_bits, di, g);
    store_sample(a.dst[2], a.dst_bits, di, b);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------

size_t match_wave_words(int max_cand, int B, int G) {
    return (size_t)max_cand + 4 * (size_t)G + (B == 8 ? 0 : (size_t)B * B);
}

size_t match_lds_bytes(int max_cand, int B, int G) {
    return sizeof(uint32_t) * (size_t)kMatchWaves * match_wave_words(max_cand, B, G);
}

// Strip width and task passes of the B = 8 fast path for the largest
// (interior) window: the width with the fewest scored slots (ties: 8).
struct Match8Shape {
    int w, kmax;
    size_t words;
};

// (rounds, strip width) pairs the 8x8 matcher is built for: X(K, W, NW).
#define BM3D_M8_SHAPES(X, NW)                                                                  \
    X(1, 8, NW) X(2, 8, NW) X(3, 8, NW) X(4, 8, NW) X(1, 11, NW) X(2, 11, NW) X(3, 11, NW)     \
    X(4, 11, NW) X(2, 3, NW) X(4, 3, NW) X(6, 3, NW) X(3, 10, NW) X(4, 10, NW)
static const int kM8Shapes[][2] = {{1, 8},  {2, 8}, {3, 8}, {4, 8}, {1, 11}, {2, 11}, {3, 11},
                                   {4, 11}, {2, 3}, {4, 3}, {6, 3}, {3, 10}, {4, 10}};

// The fewest built rounds of strip width w that cover kmax rounds (0: none).
static int m8_rounds(int w, int kmax) {
    int best = 0;
    for (const auto& s : kM8Shapes)
        if (s[1] == w && s[0] >= kmax && (best == 0 || s[0] < best)) best = s[0];
    return best;
}

// Shape of the B = 8 fast path for strip width w: rounds of 64 strips and
// the wave's LDS words (match8_block's layout: the window with its padding,
// or the selection scratch that later reuses it, and the reference block).
static Match8Shape match8_shape_w(int max_cand, int G, int w) {
    const int side = (int)(0.5 + __builtin_sqrt((double)max_cand));
    const int nstr = (side + w - 1) / w;
    const int kmax = m8_rounds(w, (side * nstr + 63) / 64);
    const size_t ww = (size_t)((side + 7) | 1);
    const size_t win = (size_t)(side + 7) * ww + (size_t)w - 1;
    const size_t sel = 2 * (size_t)kSelCap + 4 * (size_t)G;
    return {w, kmax, (((win > sel ? win : sel) + 3) & ~(size_t)3) + 64};
}

// The built shape with the fewest strip slots (rounds x width: every round
// runs whole); among equals the wider strip (fewer window reads).
Match8Shape match8_shape(int max_cand, int G) {
    Match8Shape best{0, 0, 0};
    long best_slots = -1;
    for (int w : {11, 10, 8, 3}) {
        const Match8Shape sh = match8_shape_w(max_cand, G, w);
        if (sh.kmax == 0) continue;
        const long slots = (long)sh.kmax * w;
        if (best_slots < 0 || slots < best_slots) {
            best_slots = slots;
            best = sh;
        }
    }
    return best;
}

hipError_t launch_match(const float* plane, const StageParams& p, int nb, int max_cand, int* count, int* pos,
                        float* key, hipStream_t st, const int* blist, const int* nlist) {
    if (nb <= 0) return hipSuccess;
    // the 8x8 register-strip path computes the x86 (SSE2) lane order only;
    // the scalar flavour runs the generic matcher
    if (blist && p.block == 8 && p.step == 1 && p.arith == kArithSSE2) {
        // a block list: the cooperative form (four waves per block), windows up to 49 x 49
        const int side = (int)(0.5 + __builtin_sqrt((double)max_cand));
        const int nstr = (side + kCoopW - 1) / kCoopW;
        if (side * nstr <= 256) {
            const size_t words = coop_words(side, 1, kCoopW);
            // enough workgroups that a typical list (a few percent of the blocks) takes
            // one block per workgroup
            const int grid = nb < device_cus() * 8 ? nb : device_cus() * 8;
            hipLaunchKernelGGL((k_match8_coop<1, kCoopW>), dim3(grid), dim3(256), words * 4, st, plane, p, nb, blist,
                               nlist, count, pos, key, (int)words);
            return hipGetLastError();
        }
    }
    if (p.block == 8 && p.step == 1 && p.arith == kArithSSE2) {
        const Match8Shape sh = match8_shape(max_cand, p.group);
        const int km = sh.kmax;
        const size_t wbytes = sh.words * 4;
        const int nw = wbytes * 4 <= 65536 ? 4 : wbytes * 2 <= 65536 ? 2 : wbytes <= 65536 ? 1 : 0;
        if (km != 0 && nw != 0) {
            // xcd_item's grid; a block list: a grid of 4 workgroups per CU striding over it
            const int full = (((nb + nw - 1) / nw) + 7) / 8 * 8;
            const int blocks = blist ? (full < device_cus() * 4 ? full : device_cus() * 4) : full;
            const size_t lds = wbytes * nw;
#define BM3D_M8(K, WD, NW)                                                                                   \
    if (km == K && sh.w == WD && nw == NW) {                                                                 \
        if (blist)                                                                                           \
            hipLaunchKernelGGL((k_match8_list<K, WD, NW>), dim3(blocks), dim3(64 * NW), lds, st, plane, p,  \
                               nb, (int)sh.words, blist, nlist, count, pos, key);                            \
        else                                                                                                 \
            hipLaunchKernelGGL((k_match8<K, WD, NW>), dim3(blocks), dim3(64 * NW), lds, st, plane, p, nb,    \
                               (int)sh.words, blist, nlist, count, pos, key);                                \
        return hipGetLastError();                                                                            \
    }
            BM3D_M8_SHAPES(BM3D_M8, 4)
            BM3D_M8_SHAPES(BM3D_M8, 2)
            BM3D_M8_SHAPES(BM3D_M8, 1)
#undef BM3D_M8
        }
    }
    int blocks = (nb + kMatchWaves - 1) / kMatchWaves;
    size_t lds = match_lds_bytes(max_cand, p.block, p.group);
    int ww = (int)match_wave_words(max_cand, p.block, p.group);
    if (blist) {
        const int grid = blocks < device_cus() * 4 ? blocks : device_cus() * 4;
        if (p.block == 8)
            hipLaunchKernelGGL(k_match_list<8>, dim3(grid), dim3(64 * kMatchWaves), lds, st, plane, p, nb, max_cand,
                               ww, blist, nlist, count, pos, key);
        else
            hipLaunchKernelGGL(k_match_list<0>, dim3(grid), dim3(64 * kMatchWaves), lds, st, plane, p, nb, max_cand,
                               ww, blist, nlist, count, pos, key);
        return hipGetLastError();
    }
    if (p.block == 8)
        hipLaunchKernelGGL(k_match<8>, dim3(blocks), dim3(64 * kMatchWaves), lds, st, plane, p, 0, nb, max_cand,
                           ww, count, pos, key);
    else
        hipLaunchKernelGGL(k_match<0>, dim3(blocks), dim3(64 * kMatchWaves), lds, st, plane, p, 0, nb, max_cand,
                           ww, count, pos, key);
    return hipGetLastError();
}

size_t filter_lds_bytes(int B, int G) {
    return sizeof(float) * (size_t)3 * G * B * (B + 1) + sizeof(int) * 3 * (size_t)G;
}

static int filter8_gmax(int G) { return G <= 8 ? 8 : G <= 16 ? 16 : G <= 32 ? 32 : 0; }

size_t filter8_lds_bytes(int G, bool wiener) {
    int gm = filter8_gmax(G);
    return sizeof(float) * (size_t)gm * 72 + sizeof(int) * 3 * (size_t)gm;
}

// The kernels use the generated coefficient constants; they must equal what
// the host's std::cos gives (the tables of upload_coefs and the oracle).
template <int N>
static bool verify_dctc() {
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j) {
            float want = (float)(2.0 * std::cos(M_PI * (double)k * (double)(2 * j + 1) / (2.0 * (double)N)));
            if (__builtin_memcmp(&want, &DctC<N>::v[k * N + j], 4) != 0) return false;
        }
    if constexpr (N > 1) return verify_dctc<N - 1>();
    return true;
}

bool verify_dct_constants() { return verify_dctc<kDctCMax>(); }

bool filter_uses_buckets(int B, int G) { return B == 8 && filter8_gmax(G) > 0; }

bool filter_fits(int B, int G, bool wiener) {
    if (B == 8 && filter8_gmax(G) > 0) return filter8_lds_bytes(G, wiener) <= 65536;
    return filter_lds_bytes(B, G) <= 65536;
}

hipError_t upload_dct8(const float* c64) { return hipMemcpyToSymbol(HIP_SYMBOL(c_dct8), c64, 64 * sizeof(float)); }

hipError_t upload_dctz(const float* coef, int nmax) {
    // host table layout: [off(N) + k*N + j] with off(N) = sum_{M<N} M^2
    return hipMemcpyToSymbol(HIP_SYMBOL(c_dctz), coef, sizeof(float) * zoff(nmax + 1));
}

hipError_t launch_filter(const FilterArgs& a, hipStream_t st) {
    if (a.nb == 0) return hipSuccess;
    const int gm = filter8_gmax(a.p.group);
    if (a.p.block == 8 && gm > 0) {
        const size_t lds = filter8_lds_bytes(a.p.group, a.wiener);
        const bool one = !a.process[1] && !a.process[2];
        if (!one && a.wiener) {
            // Wiener on several planes: one launch per plane of the
            // single-plane kernel, whose waves hold one plane's state (G = 32:
            // 125 instead of 168 registers, four waves per SIMD instead of
            // three); each plane's results are those of the joint kernel.  The
            // hard threshold keeps the joint kernel (six waves per SIMD
            // already, and its cross-plane gather prefetch).
            for (int p = 0; p < 3; ++p) {
                if (!a.process[p]) continue;
                FilterArgs ap = a;
                ap.process[0] = 1;
                ap.process[1] = ap.process[2] = 0;
                ap.src[0] = a.src[p];
                ap.ref[0] = a.ref[p];
                ap.contrib[0] = a.contrib[p];
                ap.denw[0] = a.denw[p];
                ap.tables = a.tables + p;
                const hipError_t e = launch_filter(ap, st);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        }
        const int grid = (a.nb + 7) / 8 * 8;  // xcd_item: a multiple of 8
        // the hard threshold is the same in both arithmetic flavours; Wiener
        // reaches here with one plane only (split above)
        const bool sc = a.wiener && a.p.arith;
#define BM3D_F8(GM, W, O, S)                                                             \
    if (gm == GM && a.wiener == W && one == O && sc == S) {                              \
        hipLaunchKernelGGL((k_filter8<GM, W, O, S>), dim3(grid), dim3(64), lds, st, a);  \
        return hipGetLastError();                                                        \
    }
        BM3D_F8(8, 0, false, false)
        BM3D_F8(16, 0, false, false)
        BM3D_F8(32, 0, false, false)
        BM3D_F8(8, 0, true, false)
        BM3D_F8(16, 0, true, false)
        BM3D_F8(32, 0, true, false)
        BM3D_F8(8, 1, true, false)
        BM3D_F8(16, 1, true, false)
        BM3D_F8(32, 1, true, false)
        BM3D_F8(8, 1, true, true)
        BM3D_F8(16, 1, true, true)
        BM3D_F8(32, 1, true, true)
#undef BM3D_F8
    }
    size_t lds = filter_lds_bytes(a.p.block, a.p.group);
    hipLaunchKernelGGL(k_filter, dim3(a.nb), dim3(64), lds, st, a);
    return hipGetLastError();
}

int device_cus() {
    static int n = [] {
        int d = 0, v = 0;
        if (hipGetDevice(&d) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || v <= 0) return 256;
        return v;
    }();
    return n;
}

bool batch_match_supported(const StageParams& p, int max_cand, const BatchShape& sh) {
    if (p.block != 8 || p.step != 1 || p.arith != kArithSSE2) return false;
    if (sh.match_strip != 3 && sh.match_strip != 8 && sh.match_strip != 10 && sh.match_strip != 11) return false;
    if (sh.match_waves != 1 && sh.match_waves != 2 && sh.match_waves != 4) return false;
    const Match8Shape ms = match8_shape_w(max_cand, p.group, sh.match_strip);
    return ms.kmax >= 1 && ms.words * 4 * sh.match_waves <= 65536;
}

hipError_t launch_match_batch(const BatchFrame* frames, const int* starts, int nframes, int total,
                              const StageParams& p0, int max_cand, const BatchShape& sh, int* count, int* pos,
                              float* key, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (!batch_match_supported(p0, max_cand, sh)) return hipErrorInvalidValue;
    const Match8Shape ms = match8_shape_w(max_cand, p0.group, sh.match_strip);
    const int km = ms.kmax;
    const int nw = sh.match_waves;
    const size_t lds = ms.words * 4 * nw;
    const int groups = batch_groups(total, nw);
#define BM3D_M8B(K, WD, NW)                                                                                  \
    if (km == K && ms.w == WD && nw == NW) {                                                                 \
        hipLaunchKernelGGL((k_match8_batch<K, WD, NW>), dim3(groups), dim3(64 * NW), lds, st, frames, starts, \
                           nframes, total, (int)ms.words, p0.group, count, pos, key);                        \
        return hipGetLastError();                                                                            \
    }
    BM3D_M8_SHAPES(BM3D_M8B, 1)
    BM3D_M8_SHAPES(BM3D_M8B, 2)
    BM3D_M8_SHAPES(BM3D_M8B, 4)
#undef BM3D_M8B
    return hipErrorInvalidValue;
}

hipError_t launch_filter_batch(const FilterArgs& a0, const BatchFilterFrame* ff, const int* starts, int nframes,
                               const int* order, int total, const BatchShape& sh, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    const int G = a0.p.group;
    const int gm = filter8_gmax(G);
    if (a0.p.block != 8 || gm == 0) return hipErrorInvalidValue;
    const size_t lds = filter8_lds_bytes(G, a0.wiener);
    const int groups = batch_groups(total, 1);
    const bool sc = a0.wiener && a0.p.arith;
#define BM3D_F8B(GM, WI, S)                                                                                   \
    if (gm == GM && (a0.wiener != 0) == WI && sc == S) {                                                       \
        hipLaunchKernelGGL((k_filter8_batch<GM, WI, S>), dim3(groups), dim3(64), lds, st, a0, ff, starts, nframes, \
                           order, total);                                                                      \
        return hipGetLastError();                                                                              \
    }
    BM3D_F8B(8, false, false)
    BM3D_F8B(16, false, false)
    BM3D_F8B(32, false, false)
    BM3D_F8B(8, true, false)
    BM3D_F8B(16, true, false)
    BM3D_F8B(32, true, false)
    BM3D_F8B(8, true, true)
    BM3D_F8B(16, true, true)
    BM3D_F8B(32, true, true)
#undef BM3D_F8B
    return hipErrorInvalidValue;
}

hipError_t launch_to_float(const void* src, int bits, size_t spitch, float* dst, size_t dpitch, int W, int H,
                           float gain, float offset, int clip, float lo, float hi, hipStream_t st) {
    dim3 grid((W + 255) / 256, H);
    hipLaunchKernelGGL(k_to_float, grid, dim3(256), 0, st, src, bits, spitch, dst, dpitch, W, H, gain, offset, clip,
                       lo, hi);
    return hipGetLastError();
}

hipError_t launch_rgb2opp(const RgbArgs& a, hipStream_t st) {
    dim3 grid((a.W + 255) / 256, a.H);
    hipLaunchKernelGGL(k_rgb2opp, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_opp2rgb(const RgbArgs& a, hipStream_t st) {
    dim3 grid((a.W + 255) / 256, a.H);
    hipLaunchKernelGGL(k_opp2rgb, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace bm3d
