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
// ordered gather aggregation shared by the spatial and temporal kernels
// ---------------------------------------------------------------------------

constexpr int kAggCap = 256;   // records buffered per wave (LDS)
constexpr int kAggWords = 5;   // per record: position, contribution base, den weight x 3

// Adds the buffered records [0, nrec) in order.  Loads of 8 records are
// issued before any of them is added, so their latencies overlap; the adds
// themselves keep the record order.
__device__ __forceinline__ void agg_consume(const int* rec, int nrec, int B, int px, int py, bool valid,
                                            const bool* proc, const float* __restrict__ c0,
                                            const float* __restrict__ c1, const float* __restrict__ c2, float* num,
                                            float* den) {
    const float* cp[3] = {c0, c1, c2};
    const int BB = B * B;
    for (int r0 = 0; r0 < nrec; r0 += 8) {
        float v[3][8], w[3][8];
        bool cov[8];
#pragma unroll
        for (int rr = 0; rr < 8; ++rr) {
            const int r = r0 + rr;
            cov[rr] = false;
            if (r < nrec) {
                const int* e = rec + r * kAggWords;
                const int q = e[0];
                const int dy = py - (q >> 16), dx = px - (q & 0xffff);
                cov[rr] = valid && dy >= 0 && dy < B && dx >= 0 && dx < B;
                const size_t off = (size_t)e[1] * BB + (size_t)(dy * B + dx);
#pragma unroll
                for (int pl = 0; pl < 3; ++pl) {
                    if (!proc[pl]) continue;
                    v[pl][rr] = cov[rr] ? cp[pl][off] : 0.0f;
                    w[pl][rr] = __int_as_float(e[2 + pl]);
                }
            }
        }
#pragma unroll
        for (int rr = 0; rr < 8; ++rr) {
            if (r0 + rr < nrec && cov[rr]) {
#pragma unroll
                for (int pl = 0; pl < 3; ++pl) {
                    if (!proc[pl]) continue;
                    num[pl] = num[pl] + v[pl][rr];
                    den[pl] = den[pl] + w[pl][rr];
                }
            }
        }
    }
}

// Walks the reference blocks of the rectangle [a_lo, a_hi] x [c_lo, c_hi] in
// scan order and accumulates every member (of frame fr, when frame != null)
// that overlaps the 8x8 tile at (y0, x0), in (block, member) order.
__device__ inline void agg_tile(const StageParams& p, const int* __restrict__ count, const int* __restrict__ pos,
                                const int* __restrict__ frame, int fr, int a_lo, int a_hi, int c_lo, int c_hi,
                                int y0, int x0, int px, int py, bool valid, const bool* proc,
                                const float* const* contrib, const float* const* denw, int* rec, float* num,
                                float* den) {
    const int lane = lane_id();
    const int B = p.block, G = p.group;
    const int gp = G <= 1 ? 1 : G <= 2 ? 2 : G <= 4 ? 4 : G <= 8 ? 8 : G <= 16 ? 16 : G <= 32 ? 32 : 64;
    const int per = 64 / gp;  // reference blocks per wave-load
    const int ncol = c_hi - c_lo + 1;
    const int chunks_per_row = (ncol + per - 1) / per;
    const int zchunks = (G + 63) / 64;
    const int total = (a_hi - a_lo + 1) * chunks_per_row * zchunks;
    int nrec = 0;
    for (int it = 0; it < total; ++it) {
        const int zc = it % zchunks;
        const int rc = it / zchunks;
        const int ar = a_lo + rc / chunks_per_row;
        const int ac = c_lo + (rc % chunks_per_row) * per + lane / gp;
        const int z = zc * 64 + lane % gp;
        bool ov = false;
        int pk = 0, b = 0;
        float dw[3] = {0.0f, 0.0f, 0.0f};
        if (ac <= c_hi && z < G) {
            b = ar * p.nx + ac;
            const int n = count[b];
            pk = pos[(size_t)b * G + z];
            const int fz = frame ? frame[(size_t)b * G + z] : fr;
#pragma unroll
            for (int pl = 0; pl < 3; ++pl)
                if (proc[pl]) dw[pl] = denw[pl][b];
            const int qy = pk >> 16, qx = pk & 0xffff;
            ov = z < n && fz == fr && qy <= y0 + 7 && qy + B > y0 && qx <= x0 + 7 && qx + B > x0;
        }
        const uint64_t m = __ballot(ov);
        const int cnt = __popcll(m);
        if (cnt == 0) continue;
        if (nrec + cnt > kAggCap) {
            wave_sync_lds();
            agg_consume(rec, nrec, B, px, py, valid, proc, contrib[0], contrib[1], contrib[2], num, den);
            wave_sync_lds();
            nrec = 0;
        }
        const int at = nrec + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        if (ov) {
            int* e = rec + at * kAggWords;
            e[0] = pk;
            e[1] = b * G + z;
            e[2] = __float_as_int(dw[0]);
            e[3] = __float_as_int(dw[1]);
            e[4] = __float_as_int(dw[2]);
        }
        nrec += cnt;
    }
    wave_sync_lds();
    agg_consume(rec, nrec, B, px, py, valid, proc, contrib[0], contrib[1], contrib[2], num, den);
}

// Conservative rectangle of reference blocks whose members can reach a tile
// when they lie at most `reach` from their reference block.
__device__ __forceinline__ void reach_rect(const StageParams& p, int y0, int x0, int reach, int& a_lo, int& a_hi,
                                           int& c_lo, int& c_hi) {
    const int B = p.block, s = p.block_step;
    const int lo_y = y0 - B + 1 - reach, hi_y = y0 + 7 + reach;
    const int lo_x = x0 - B + 1 - reach, hi_x = x0 + 7 + reach;
    a_lo = lo_y > 0 ? lo_y / s - 1 : 0;
    a_hi = hi_y / s + 1;
    c_lo = lo_x > 0 ? lo_x / s - 1 : 0;
    c_hi = hi_x / s + 1;
    a_lo = a_lo < 0 ? 0 : a_lo;
    c_lo = c_lo < 0 ? 0 : c_lo;
    a_hi = a_hi > p.ny - 1 ? p.ny - 1 : a_hi;
    c_hi = c_hi > p.nx - 1 ? p.nx - 1 : c_hi;
}

}  // namespace bm3d
