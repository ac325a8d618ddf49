// Device helpers shared by the HIP kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bm3d_params.h"

namespace bm3d {

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Position k of the reference-block grid along one axis: 0, s, 2s, ... with
// the last one clamped to dim - B (source/BM3D_Base.cpp:312-345).
__device__ __forceinline__ int grid_pos(int k, int step, int last) {
    int v = k * step;
    return v > last ? last : v;
}

// Block::_SearchBoundary (include/Block.h:734-764) in closed form: walk from
// pos -/+ range towards pos in steps of `step` until inside the plane.
__device__ __forceinline__ int search_lo(int pos, int range, int step) {
    int m = pos / step * step;  // largest multiple of step that keeps pos - m >= 0
    return pos - (m < range ? m : range);
}
__device__ __forceinline__ int search_hi(int pos, int last, int range, int step) {
    int m = (last - pos) / step * step;
    return pos + (m < range ? m : range);
}

// SSD of the reference block against the candidate at q, in the SSE2 lane
// order of Block::BlockMatchingMulti (include/Block.h:575-631): four lane
// accumulators, per row and 8-column chunk lane k takes column 8c+k and then
// 8c+4+k; the lanes are folded as ((s0+s1)+s2)+s3; columns beyond the last
// full chunk are then added one by one in row-major order.  Each operation
// rounds separately (no FMA).
template <int B>
__device__ __forceinline__ float ssd_fixed(const float* rb, const float* __restrict__ q, int pitch) {
    constexpr int SW = B - B % 8;
    float dist = 0.0f;
    if constexpr (SW > 0) {
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const float* qr = q + r * pitch;
#pragma unroll
            for (int c = 0; c < SW; c += 8) {
                float d0 = rb[r * B + c + 0] - qr[c + 0];
                float d1 = rb[r * B + c + 1] - qr[c + 1];
                float d2 = rb[r * B + c + 2] - qr[c + 2];
                float d3 = rb[r * B + c + 3] - qr[c + 3];
                s0 = s0 + d0 * d0;
                s1 = s1 + d1 * d1;
                s2 = s2 + d2 * d2;
                s3 = s3 + d3 * d3;
                float e0 = rb[r * B + c + 4] - qr[c + 4];
                float e1 = rb[r * B + c + 5] - qr[c + 5];
                float e2 = rb[r * B + c + 6] - qr[c + 6];
                float e3 = rb[r * B + c + 7] - qr[c + 7];
                s0 = s0 + e0 * e0;
                s1 = s1 + e1 * e1;
                s2 = s2 + e2 * e2;
                s3 = s3 + e3 * e3;
            }
        }
        dist = dist + (((s0 + s1) + s2) + s3);
    }
    if constexpr (B - SW > 0) {
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = SW; c < B; ++c) {
                float t = rb[r * B + c] - q[r * pitch + c];
                dist = dist + t * t;
            }
    }
    return dist;
}

// Runtime-B variant (reference block held in LDS).
__device__ inline float ssd_generic(const float* rb, const float* __restrict__ q, int pitch, int B) {
    const int SW = B - B % 8;
    float dist = 0.0f;
    if (SW > 0) {
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (int r = 0; r < B; ++r) {
            const float* qr = q + r * pitch;
            const float* rr = rb + r * B;
            for (int c = 0; c < SW; c += 8) {
                float d0 = rr[c + 0] - qr[c + 0];
                float d1 = rr[c + 1] - qr[c + 1];
                float d2 = rr[c + 2] - qr[c + 2];
                float d3 = rr[c + 3] - qr[c + 3];
                s0 = s0 + d0 * d0;
                s1 = s1 + d1 * d1;
                s2 = s2 + d2 * d2;
                s3 = s3 + d3 * d3;
                float e0 = rr[c + 4] - qr[c + 4];
                float e1 = rr[c + 5] - qr[c + 5];
                float e2 = rr[c + 6] - qr[c + 6];
                float e3 = rr[c + 7] - qr[c + 7];
                s0 = s0 + e0 * e0;
                s1 = s1 + e1 * e1;
                s2 = s2 + e2 * e2;
                s3 = s3 + e3 * e3;
            }
        }
        dist = dist + (((s0 + s1) + s2) + s3);
    }
    if (B - SW > 0) {
        for (int r = 0; r < B; ++r)
            for (int c = SW; c < B; ++c) {
                float t = rb[r * B + c] - q[r * pitch + c];
                dist = dist + t * t;
            }
    }
    return dist;
}

// Sorts the k (<= 256) entries sel_key/sel_idx[0..k) by (key, index) into
// sel_key/sel_idx[k..2k) by counting, for each entry, the entries before it.
__device__ inline void rank_sort_selected(uint32_t* sel_key, int* sel_idx, int k) {
    const int lane = lane_id();
    uint32_t rk[4];
    int ri[4], rr[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        int e = lane + 64 * s;
        rk[s] = e < k ? sel_key[e] : 0xFFFFFFFFu;
        ri[s] = e < k ? sel_idx[e] : 0x7FFFFFFF;
        rr[s] = 0;
    }
    for (int e2 = 0; e2 < k; ++e2) {
        uint32_t k2 = sel_key[e2];
        int i2 = sel_idx[e2];
#pragma unroll
        for (int s = 0; s < 4; ++s) rr[s] += (k2 < rk[s]) || (k2 == rk[s] && i2 < ri[s]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        int e = lane + 64 * s;
        if (e < k) {
            sel_key[k + rr[s]] = rk[s];
            sel_idx[k + rr[s]] = ri[s];
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// exact top-k selection on a wave
// ---------------------------------------------------------------------------
//
// keys[c] holds the IEEE bit pattern of candidate c's key (positive floats
// order like their bit patterns) or 0xFFFFFFFF when rejected.  Selects the
// k smallest by (key, c) -- the order std::stable_sort gives the accepted
// list, and the order std::partial_sort gives it except among equal keys --
// and writes them sorted to sel_key / sel_idx.  nc candidates, wave-uniform.
__device__ inline void wave_select(const uint32_t* keys, int nc, int naccepted, int k, uint32_t* sel_key,
                            int* sel_idx) {
    const int lane = lane_id();
    uint32_t thr = 0xFFFFFFFEu;  // all accepted keys qualify
    int need_eq = 0;             // how many keys equal to thr are admitted (when cut)
    const bool cut = naccepted > k;
    if (cut) {
        // radix descent for the k-th smallest key value T:
        // the largest T with count(key < T) < k
        uint32_t t = 0;
        for (int bit = 31; bit >= 0; --bit) {
            uint32_t trial = t | (1u << bit);
            int cnt = 0;
            for (int c0 = 0; c0 < nc; c0 += 64) {
                int c = c0 + lane;
                bool less = (c < nc) && (keys[c] < trial);
                cnt += __popcll(__ballot(less));
            }
            if (cnt < k) t = trial;
        }
        int less_cnt = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            int c = c0 + lane;
            less_cnt += __popcll(__ballot((c < nc) && (keys[c] < t)));
        }
        thr = t;
        need_eq = k - less_cnt;
    }
    // compaction in candidate order
    int base = 0, eq_seen = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        int c = c0 + lane;
        uint32_t kv = (c < nc) ? keys[c] : 0xFFFFFFFFu;
        bool sel;
        if (cut) {
            bool eq = kv == thr;
            uint64_t eqm = __ballot(eq);
            int eq_rank = eq_seen + __builtin_amdgcn_mbcnt_hi((uint32_t)(eqm >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)eqm, 0));
            sel = (kv < thr) || (eq && eq_rank < need_eq);
            eq_seen += __popcll(eqm);
        } else {
            sel = kv != 0xFFFFFFFFu;
        }
        uint64_t m = __ballot(sel);
        int pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        if (sel) {
            sel_key[pos] = kv;
            sel_idx[pos] = c;
        }
        base += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    rank_sort_selected(sel_key, sel_idx, k);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// arr[i] for a three-entry array of a kernel-argument struct, by selects: a
// dynamically indexed kernel-argument array would be copied to scratch.
template <typename T>
__device__ __forceinline__ T pick3(const T (&arr)[3], int i) {
    return i == 0 ? arr[0] : (i == 1 ? arr[1] : arr[2]);
}

// Load from global memory (the pointer may have lost its address space
// through a select; a flat load would also count against the LDS counter).
__device__ __forceinline__ float ld_global(const float* p) {
    return *(const __attribute__((address_space(1))) float*)p;
}

__device__ __forceinline__ float clipf(float v, float lo, float hi) { return v <= lo ? lo : (v >= hi ? hi : v); }

__device__ __forceinline__ void store_sample(void* base, int bits, size_t idx, float v) {
    if (bits == 0) static_cast<float*>(base)[idx] = v;
    else if (bits <= 8) static_cast<uint8_t*>(base)[idx] = (uint8_t)v;
    else static_cast<uint16_t*>(base)[idx] = (uint16_t)v;
}

__device__ __forceinline__ float load_sample(const void* base, int bits, size_t idx) {
    if (bits == 0) return static_cast<const float*>(base)[idx];
    if (bits <= 8) return (float)static_cast<const uint8_t*>(base)[idx];
    return (float)static_cast<const uint16_t*>(base)[idx];
}

// rcp as the host's _mm_rcp_ps computes it: a function of the top mantissa
// bits scaled by the exponent (table probed on the host at init).
__device__ __forceinline__ float rcp_host(const float* __restrict__ tbl, int shift, float x) {
    uint32_t u = __float_as_uint(x);
    int e = (int)((u >> 23) & 0xff) - 127;
    float r = __builtin_ldexpf(tbl[(u & 0x7fffff) >> shift], -e);
    return (u >> 31) ? -r : r;
}

// ---------------------------------------------------------------------------
// ordered aggregation into an LDS tile, shared by the spatial and temporal
// kernels
// ---------------------------------------------------------------------------
//
// One wave owns a kAggTH x kAggTW tile of numerator and denominator
// accumulators in LDS.  It walks the reference blocks that can reach the tile
// in scan order and, for each block, its members in group order; every member
// overlapping the tile is a record.  A record's B*B contributions are loaded
// contiguously (lane = element of the block) and added into the tile: within
// one record every lane owns a distinct pixel, and records are applied one
// after another, so each pixel receives its terms in (block, member) order.

constexpr int kAggTH = 16;               // tile rows
constexpr int kAggTW = 32;               // tile columns
constexpr int kAggStride = kAggTW + 8;   // LDS row stride (8 rows of a block hit distinct banks)
constexpr int kAggPlane = kAggTH * kAggStride;
constexpr int kAggLds = 6 * kAggPlane;   // num, den for three planes

__device__ __forceinline__ void reach_rect(const StageParams& p, int y0, int x0, int th, int tw, int reach,
                                           int& a_lo, int& a_hi, int& c_lo, int& c_hi) {
    const int B = p.block, s = p.block_step;
    const int lo_y = y0 - B + 1 - reach, hi_y = y0 + th - 1 + reach;
    const int lo_x = x0 - B + 1 - reach, hi_x = x0 + tw - 1 + reach;
    a_lo = lo_y > 0 ? lo_y / s - 1 : 0;
    a_hi = hi_y / s + 1;
    c_lo = lo_x > 0 ? lo_x / s - 1 : 0;
    c_hi = hi_x / s + 1;
    a_lo = a_lo < 0 ? 0 : a_lo;
    c_lo = c_lo < 0 ? 0 : c_lo;
    a_hi = a_hi > p.ny - 1 ? p.ny - 1 : a_hi;
    c_hi = c_hi > p.nx - 1 ? p.nx - 1 : c_hi;
}

// Accumulates, into the LDS tile acc at (y0, x0), every member (of frame fr
// when frame != null) of the reference blocks [a_lo, a_hi] x [c_lo, c_hi].
// NP processed planes with contributions c[k] and weights d[k];
// acc layout: [k][num | den][kAggTH][kAggStride].
//
// Loads are issued unconditionally from addresses that are always inside the
// record (lanes past B*B re-read element 0) and the results of uncovered
// lanes are simply not added, so a batch of records keeps all its loads in
// flight at once.
template <int NP>
__device__ __forceinline__ void agg_tile_lds(const StageParams& p, const int* __restrict__ count, const int* __restrict__ pos,
                                    const int* __restrict__ frame, int fr, int a_lo, int a_hi, int c_lo, int c_hi,
                                    int y0, int x0, const float* const* c, const float* const* d, float* acc) {
    const int lane = lane_id();
    for (int i = lane; i < NP * 2 * kAggPlane; i += 64) acc[i] = 0.0f;
    wave_sync_lds();
    const int B = p.block, G = p.group, BB = B * B;
    const int gp = G <= 1 ? 1 : G <= 2 ? 2 : G <= 4 ? 4 : G <= 8 ? 8 : G <= 16 ? 16 : G <= 32 ? 32 : 64;
    const int per = 64 / gp;  // reference blocks per walk step
    const int ncol = c_hi - c_lo + 1;
    const int chunks_per_row = (ncol + per - 1) / per;
    const int zchunks = (G + 63) / 64;
    const int total = (a_hi - a_lo + 1) * chunks_per_row * zchunks;
    const int y1 = y0 + kAggTH - 1, x1 = x0 + kAggTW - 1;
    const int lc = lane < BB ? lane : 0;  // element this lane carries (B*B <= 64 path)
    const int dy = lc / B, dx = lc - (lc / B) * B;
    const bool small = BB <= 64;

    // walk step it: lane -> (reference block b, member z); loads prefetched
    // one step ahead
    auto step_ids = [&](int it, int& b, int& z, bool& live) {
        const int zc = it % zchunks;
        const int rc = it / zchunks;
        const int ar = a_lo + rc / chunks_per_row;
        const int ac = c_lo + (rc % chunks_per_row) * per + lane / gp;
        z = zc * 64 + lane % gp;
        live = ac <= c_hi && z < G;
        b = ar * p.nx + (ac <= c_hi ? ac : c_hi);
        z = z < G ? z : G - 1;
    };
    int nb_n = 0, pk_n = 0, fz_n = fr;
    if (total > 0) {
        int b, z;
        bool live;
        step_ids(0, b, z, live);
        nb_n = count[b];
        pk_n = pos[b * G + z];
        if (frame) fz_n = frame[b * G + z];
    }
    for (int it = 0; it < total; ++it) {
        int b, z;
        bool live;
        step_ids(it, b, z, live);
        const int n = nb_n, pk = pk_n, fz = fz_n;
        if (it + 1 < total) {
            int b2, z2;
            bool l2;
            step_ids(it + 1, b2, z2, l2);
            nb_n = count[b2];
            pk_n = pos[b2 * G + z2];
            if (frame) fz_n = frame[b2 * G + z2];
        }
        const int e = b * G + z;
        const int qy = pk >> 16, qx = pk & 0xffff;
        const bool ov = live && z < n && fz == fr && qy <= y1 && qy + B > y0 && qx <= x1 && qx + B > x0;
        uint64_t m = __ballot(ov);
        if (small) {
            while (m) {
                // up to 8 records: all loads first, then the adds in order
                int rpk[8], re[8], rb[8];
                int cnt = 0;
                int l = __builtin_ctzll(m);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (m) {
                        l = __builtin_ctzll(m);
                        m &= m - 1;
                        cnt = j + 1;
                    }
                    rpk[j] = __builtin_amdgcn_readlane(pk, l);
                    re[j] = __builtin_amdgcn_readlane(e, l);
                    rb[j] = __builtin_amdgcn_readlane(b, l);
                }
                float v[NP][8], w[NP][8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const size_t off = (size_t)re[j] * BB + lc;
#pragma unroll
                    for (int k = 0; k < NP; ++k) {
                        v[k][j] = ld_global(c[k] + off);
                        w[k][j] = ld_global(d[k] + rb[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (j < cnt) {
                        const int ty = (rpk[j] >> 16) + dy - y0, tx = (rpk[j] & 0xffff) + dx - x0;
                        if (lane < BB && ty >= 0 && ty < kAggTH && tx >= 0 && tx < kAggTW) {
                            float* an = acc + ty * kAggStride + tx;
#pragma unroll
                            for (int k = 0; k < NP; ++k) {
                                an[k * 2 * kAggPlane] = an[k * 2 * kAggPlane] + v[k][j];
                                an[k * 2 * kAggPlane + kAggPlane] = an[k * 2 * kAggPlane + kAggPlane] + w[k][j];
                            }
                        }
                    }
                }
            }
        } else {
            // B*B > 64: one record at a time, 64 elements per pass
            while (m) {
                const int l = __builtin_ctzll(m);
                m &= m - 1;
                const int rpk = __builtin_amdgcn_readlane(pk, l);
                const int re = __builtin_amdgcn_readlane(e, l);
                const int rb = __builtin_amdgcn_readlane(b, l);
                const int qy2 = rpk >> 16, qx2 = rpk & 0xffff;
                for (int c0 = 0; c0 < BB; c0 += 64) {
                    const int el = c0 + lane;
                    const int ey = el / B, ex = el - (el / B) * B;
                    const int ty = qy2 + ey - y0, tx = qx2 + ex - x0;
                    if (el < BB && ty >= 0 && ty < kAggTH && tx >= 0 && tx < kAggTW) {
                        const size_t off = (size_t)re * BB + el;
                        float* an = acc + ty * kAggStride + tx;
#pragma unroll
                        for (int k = 0; k < NP; ++k) {
                            an[k * 2 * kAggPlane] = an[k * 2 * kAggPlane] + ld_global(c[k] + off);
                            an[k * 2 * kAggPlane + kAggPlane] = an[k * 2 * kAggPlane + kAggPlane] + ld_global(d[k] + rb);
                        }
                    }
                }
            }
        }
    }
    wave_sync_lds();
}

// Runs agg_tile_lds for the processed planes (k-th processed plane -> its
// contribution and weight arrays, selected without indexing so the pointers
// stay in registers); returns their number.
__device__ __forceinline__ int agg_tile_any(const StageParams& p, const int* process, const int* __restrict__ count,
                                            const int* __restrict__ pos, const int* __restrict__ frame, int fr,
                                            int a_lo, int a_hi, int c_lo, int c_hi, int y0, int x0,
                                            const float* const* contrib, const float* const* denw, float* acc) {
    const bool p0 = process[0] != 0, p1 = process[1] != 0, p2 = process[2] != 0;
    const int np = (int)p0 + (int)p1 + (int)p2;
    const int i0 = p0 ? 0 : p1 ? 1 : 2;
    const int i1 = (p0 && p1) ? 1 : 2;
    const float* c[3];
    const float* d[3];
    c[0] = i0 == 0 ? contrib[0] : i0 == 1 ? contrib[1] : contrib[2];
    d[0] = i0 == 0 ? denw[0] : i0 == 1 ? denw[1] : denw[2];
    c[1] = i1 == 1 ? contrib[1] : contrib[2];
    d[1] = i1 == 1 ? denw[1] : denw[2];
    c[2] = contrib[2];
    d[2] = denw[2];
    if (np == 3) agg_tile_lds<3>(p, count, pos, frame, fr, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc);
    else if (np == 2) agg_tile_lds<2>(p, count, pos, frame, fr, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc);
    else if (np == 1) agg_tile_lds<1>(p, count, pos, frame, fr, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc);
    return np;
}

}  // namespace bm3d
