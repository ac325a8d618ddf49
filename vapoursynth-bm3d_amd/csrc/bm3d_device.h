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
__device__ __forceinline__ float uniformf(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniformd(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ const float* uniform_ptr(const float* p) {
    const unsigned long long b = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return (const float*)(((unsigned long long)hi << 32) | lo);
}

// Item of consumer `sub` (a wave) of this workgroup in a launch of
// 8 * ceil(ceil(total / per) / 8) workgroups with `per` consumers each, or -1.
// Workgroup g runs on XCD g % 8, so the items are split into eight contiguous
// ranges, one per XCD: neighbouring reference blocks and tiles, which read
// overlapping pixels and contributions, share an L2.
__device__ __forceinline__ int xcd_item(int total, int per, int sub) {
    const int span = (total + 7) >> 3;
    const int x = (int)(blockIdx.x & 7);
    const int i = x * span + (int)(blockIdx.x >> 3) * per + sub;
    return i < min(total, (x + 1) * span) ? i : -1;
}

// Frame of global item idx: the last f with starts[f] <= idx (starts[0] = 0,
// ascending, nframes entries; wave-uniform).
__device__ __forceinline__ int batch_frame(const int* __restrict__ starts, int nframes, int idx) {
    int lo = 0, hi = nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (uniform(starts[mid]) <= idx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

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
template <int B, bool SCALAR = false>
__device__ __forceinline__ float ssd_fixed(const float* rb, const float* __restrict__ q, int pitch) {
    constexpr int SW = B - B % 8;
    float dist = 0.0f;
    if constexpr (SCALAR) {
        // builds without __SSE2__ (Block.h:632-643): row-major, one accumulator
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                float t = rb[r * B + c] - q[r * pitch + c];
                dist = dist + t * t;
            }
        return dist;
    }
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
__device__ inline float ssd_generic(const float* rb, const float* __restrict__ q, int pitch, int B, bool scalar) {
    const int SW = scalar ? 0 : B - B % 8;  // scalar: every column row-major
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
    if (k <= 64) {
        // one entry per lane
        const uint32_t rk = lane < k ? sel_key[lane] : 0xFFFFFFFFu;
        const int ri = lane < k ? sel_idx[lane] : 0x7FFFFFFF;
        // the other entries come from the lanes' registers (no LDS round trip per step)
        int rr = 0;
        for (int e2 = 0; e2 < k; ++e2) {
            const uint32_t k2 = (uint32_t)__builtin_amdgcn_readlane((int)rk, e2);
            const int i2 = __builtin_amdgcn_readlane(ri, e2);
            rr += (k2 < rk) || (k2 == rk && i2 < ri);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < k) {
            sel_key[k + rr] = rk;
            sel_idx[k + rr] = ri;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        return;
    }
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

// wave_select for at most 64 * NR candidates, with the keys held in
// registers during the radix descent (one LDS read per key instead of one
// per bit).  Same result as wave_select.
template <int NR>
__device__ inline void wave_select_small(const uint32_t* keys, int nc, int naccepted, int k, uint32_t* sel_key,
                                         int* sel_idx) {
    const int lane = lane_id();
    uint32_t v[NR];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int c = lane + 64 * u;
        v[u] = c < nc ? keys[c] : 0xFFFFFFFFu;
    }
    uint32_t thr = 0xFFFFFFFEu;
    int need_eq = 0;
    const bool cut = naccepted > k;
    if (cut) {
        uint32_t t = 0;
        for (int bit = 31; bit >= 0; --bit) {
            const uint32_t trial = t | (1u << bit);
            int cnt = 0;
#pragma unroll
            for (int u = 0; u < NR; ++u) cnt += __popcll(__ballot(v[u] < trial));
            if (cnt < k) t = trial;
        }
        int less_cnt = 0;
#pragma unroll
        for (int u = 0; u < NR; ++u) less_cnt += __popcll(__ballot(v[u] < t));
        thr = t;
        need_eq = k - less_cnt;
    }
    int base = 0, eq_seen = 0;
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int c = lane + 64 * u;
        const uint32_t kv = v[u];
        bool sel;
        if (cut) {
            const bool eq = kv == thr;
            const uint64_t eqm = __ballot(eq);
            const int eq_rank = eq_seen + __builtin_amdgcn_mbcnt_hi((uint32_t)(eqm >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)eqm, 0));
            sel = (kv < thr) || (eq && eq_rank < need_eq);
            eq_seen += __popcll(eqm);
        } else {
            sel = kv != 0xFFFFFFFFu;
        }
        const uint64_t m = __ballot(sel);
        const int pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
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

constexpr int kAggRecCap = 128;    // records buffered per wave between walk and consume (small: occupancy)
constexpr int kAggRecWords = 3;    // position, member index b*G + z, reference block b
constexpr int kAggBatch = 8;       // records per load batch of the consume phase
constexpr int kAggWalk = 4;        // walk steps whose loads are in flight together

// One batch of kAggBatch records of the consume phase: lane j (< kAggBatch)
// holds record j's position and weights, every lane its element of each
// record's contribution block.
template <int NP>
struct AggBatch {
    int pk;                    // lane j: packed position of record j
    float w[NP];               // lane j: weight of record j
    float v[NP][kAggBatch];    // element lc of record j's block
};

// Issues the loads of records [r0, r0 + kAggBatch) (records past nrec repeat
// the last one and are never applied).  Every lane reads element lc of each
// block, which is always inside the block.
template <int NP>
__device__ __forceinline__ void agg_load_batch(const int* rec, int r0, int nrec, int BB, int lc,
                                               const float* const* c, const float* const* d, AggBatch<NP>& t) {
    const int lane = lane_id();
    const int rj = r0 + (lane < kAggBatch ? lane : 0);
    const int idx = (rj < nrec ? rj : nrec - 1) * kAggRecWords;
    t.pk = rec[idx];
    const int re_l = rec[idx + 1];
    const int rb_l = rec[idx + 2];
#pragma unroll
    for (int k = 0; k < NP; ++k) t.w[k] = ld_global(d[k] + rb_l);
    // the block's base is wave-uniform (scalar registers) and the lane's
    // element a 32-bit offset, so each load needs no vector address add
    const unsigned lcu = (unsigned)lc;
#pragma unroll
    for (int j = 0; j < kAggBatch; ++j) {
        const size_t blk = (size_t)(unsigned)__builtin_amdgcn_readlane(re_l, j) * (unsigned)BB;
#pragma unroll
        for (int k = 0; k < NP; ++k) t.v[k][j] = ld_global(c[k] + blk + lcu);
    }
}

// Adds, in order, the records [r0, r0 + kAggBatch) below nrec.  Branch-free:
// a lane whose element falls outside the tile (or lies past the block, or
// belongs to a record past nrec) adds into its own slot of the row padding
// (columns kAggTW.. of row lane / 8), which no output reads, so the records
// follow one another with no exec-mask changes and the compiler can compute
// the next record's address while this record's LDS round trip is in flight.
// The LDS operations of one wave complete in order, so a record's reads see
// the writes of the records before it.
static_assert(kAggStride - kAggTW >= 8 && kAggTH >= 8, "row padding holds one slot per lane");
template <int NP>
__device__ __forceinline__ void agg_apply_batch(const AggBatch<NP>& t, int r0, int nrec, int BB, int dy, int dx,
                                                int y0, int x0, float* acc) {
    const int lane = lane_id();
    const int spare = (lane >> 3) * kAggStride + kAggTW + (lane & 7);
    const int loff = dy * kAggStride + dx;  // the lane's element relative to the block origin
    const bool lane_ok = lane < BB;
#pragma unroll
    for (int j = 0; j < kAggBatch; ++j) {
        const int pk = __builtin_amdgcn_readlane(t.pk, j);
        const int sy = (pk >> 16) - y0, sx = (pk & 0xffff) - x0;  // block origin in the tile (uniform)
        const bool ok = r0 + j < nrec && lane_ok && (unsigned)(sy + dy) < (unsigned)kAggTH &&
                        (unsigned)(sx + dx) < (unsigned)kAggTW;
        float* an = acc + (ok ? sy * kAggStride + sx + loff : spare);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const float w = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, t.w[k]), j));
            an[k * 2 * kAggPlane] = an[k * 2 * kAggPlane] + t.v[k][j];
            an[k * 2 * kAggPlane + kAggPlane] = an[k * 2 * kAggPlane + kAggPlane] + w;
        }
    }
}

// Applies the whole list [0, nrec) in order, with the loads of the next two
// batches in flight while a batch is being added.
template <int NP>
__device__ __forceinline__ void agg_consume_list(const int* rec, int nrec, int B, int lc, int dy, int dx, int y0,
                                                 int x0, const float* const* c, const float* const* d,
                                                 float* acc) {
    if (nrec <= 0) return;
    const int BB = B * B;
    constexpr int S = kAggBatch;
    AggBatch<NP> ta, tb, tc;
    agg_load_batch<NP>(rec, 0, nrec, BB, lc, c, d, ta);
    if (S < nrec) agg_load_batch<NP>(rec, S, nrec, BB, lc, c, d, tb);
    for (int r0 = 0; r0 < nrec; r0 += 3 * S) {
        if (r0 + 2 * S < nrec) agg_load_batch<NP>(rec, r0 + 2 * S, nrec, BB, lc, c, d, tc);
        agg_apply_batch<NP>(ta, r0, nrec, BB, dy, dx, y0, x0, acc);
        if (r0 + S >= nrec) break;
        if (r0 + 3 * S < nrec) agg_load_batch<NP>(rec, r0 + 3 * S, nrec, BB, lc, c, d, ta);
        agg_apply_batch<NP>(tb, r0 + S, nrec, BB, dy, dx, y0, x0, acc);
        if (r0 + 2 * S >= nrec) break;
        if (r0 + 4 * S < nrec) agg_load_batch<NP>(rec, r0 + 4 * S, nrec, BB, lc, c, d, tb);
        agg_apply_batch<NP>(tc, r0 + 2 * S, nrec, BB, dy, dx, y0, x0, acc);
    }
}

// Accumulates, into the LDS tile acc at (y0, x0), every member (of frame fr
// when frame != null) of the reference blocks [a_lo, a_hi] x [c_lo, c_hi].
// NP processed planes with contributions c[k] and weights d[k];
// acc layout: [k][num | den][kAggTH][kAggStride]; rec: the wave's LDS record
// list (kAggRecCap x kAggRecWords ints).
//
// Walk: lanes take (reference block, member) pairs in scan order, four steps'
// loads at a time; members overlapping the tile are appended to the record
// list.  Consume: the list is applied in order (agg_consume_list).  B*B > 64
// (blocks larger than a wave) takes a direct path, one record at a time.
template <int NP>
__device__ __forceinline__ void agg_tile_lds(const StageParams& p, const int* __restrict__ count,
                                             const int* __restrict__ pos, const int* __restrict__ frame, int fr,
                                             int a_lo, int a_hi, int c_lo, int c_hi, int y0, int x0,
                                             const float* const* c, const float* const* d, float* acc, int* rec) {
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

    auto step_ids = [&](int it, int& b, int& z, bool& live) {
        const int zc = it % zchunks;
        const int rc = it / zchunks;
        const int ar = a_lo + rc / chunks_per_row;
        const int ac = c_lo + (rc % chunks_per_row) * per + lane / gp;
        z = zc * 64 + lane % gp;
        live = it < total && ac <= c_hi && z < G;
        b = ar * p.nx + (ac <= c_hi ? ac : c_hi);
        b = it < total ? b : a_lo * p.nx + c_lo;
        z = z < G ? z : G - 1;
    };
    int nrec = 0;
    for (int it = 0; it < total; it += kAggWalk) {
        int sb[kAggWalk], sz[kAggWalk], sn[kAggWalk], spk[kAggWalk], sfz[kAggWalk];
        bool slive[kAggWalk];
#pragma unroll
        for (int u = 0; u < kAggWalk; ++u) {
            step_ids(it + u, sb[u], sz[u], slive[u]);
            sn[u] = count[sb[u]];
            spk[u] = pos[sb[u] * G + sz[u]];
            sfz[u] = frame ? frame[sb[u] * G + sz[u]] : fr;
        }
#pragma unroll
        for (int u = 0; u < kAggWalk; ++u) {
            const int qy = spk[u] >> 16, qx = spk[u] & 0xffff;
            const bool ov = slive[u] && sz[u] < sn[u] && sfz[u] == fr && qy <= y1 && qy + B > y0 && qx <= x1 &&
                            qx + B > x0;
            uint64_t m = __ballot(ov);
            if (!small) {
                // B*B > 64: one record at a time, 64 elements per pass
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    const int rpk = __builtin_amdgcn_readlane(spk[u], l);
                    const int re = __builtin_amdgcn_readlane(sb[u] * G + sz[u], l);
                    const int rb = __builtin_amdgcn_readlane(sb[u], l);
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
                continue;
            }
            const int cnt = __popcll(m);
            if (cnt == 0) continue;
            if (nrec + cnt > kAggRecCap) {
                wave_sync_lds();
                agg_consume_list<NP>(rec, nrec, B, lc, dy, dx, y0, x0, c, d, acc);
                wave_sync_lds();
                nrec = 0;
            }
            if (ov) {
                const int at = nrec + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                int* e = rec + at * kAggRecWords;
                e[0] = spk[u];
                e[1] = sb[u] * G + sz[u];
                e[2] = sb[u];
            }
            nrec += cnt;
        }
    }
    wave_sync_lds();
    agg_consume_list<NP>(rec, nrec, B, lc, dy, dx, y0, x0, c, d, acc);
    wave_sync_lds();
}

// The same records in the same order from per-frame member lists: for each
// block row ar, ro[ar * (nx + 1) + c] counts the members of frame fr held by
// the row's blocks before column c (c = nx: the row's total), and
// rp[ar * nx * G + k] are those members' (packed position, member index
// b * G + z) in scan order.  A row's members in the reach rectangle are
// contiguous, so the walk reads only members of frame fr.  B*B <= 64 and at
// most 64 block rows in the rectangle.
template <int NP>
__device__ __forceinline__ void agg_tile_rows(const StageParams& p, const int* __restrict__ ro,
                                              const int2* __restrict__ rp, int a_lo, int a_hi, int c_lo, int c_hi,
                                              int y0, int x0, const float* const* c, const float* const* d,
                                              float* acc, int* rec) {
    const int lane = lane_id();
    for (int i = lane; i < NP * 2 * kAggPlane; i += 64) acc[i] = 0.0f;
    wave_sync_lds();
    const int B = p.block, G = p.group;
    const int y1 = y0 + kAggTH - 1, x1 = x0 + kAggTW - 1;
    const int BB = B * B;
    const int lc = lane < BB ? lane : 0;
    const int dy = lc / B, dx = lc - (lc / B) * B;
    const int rstride = p.nx + 1;
    int nrec = 0;
    auto append = [&](int2 pr, bool live) {
        const int qy = pr.x >> 16, qx = pr.x & 0xffff;
        const bool ov = live && qy <= y1 && qy + B > y0 && qx <= x1 && qx + B > x0;
        const uint64_t m = __ballot(ov);
        const int cnt = __popcll(m);
        if (cnt == 0) return;
        if (nrec + cnt > kAggRecCap) {
            wave_sync_lds();
            agg_consume_list<NP>(rec, nrec, B, lc, dy, dx, y0, x0, c, d, acc);
            wave_sync_lds();
            nrec = 0;
        }
        if (ov) {
            const int at = nrec + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            int* e = rec + at * kAggRecWords;
            e[0] = pr.x;
            e[1] = pr.y;
            e[2] = pr.y / G;
        }
        nrec += cnt;
    };
    // the rows' members form one list (row-major); lane r < nrows holds row
    // a_lo + r's start rs, length and exclusive prefix ex over the list; 64
    // consecutive list entries per step, four steps' loads in flight
    const int nrows = a_hi - a_lo + 1;  // <= 64 (checked by the caller)
    int rs = 0, rn = 0;
    if (lane < nrows) {
        rs = ro[(a_lo + lane) * rstride + c_lo];
        rn = ro[(a_lo + lane) * rstride + c_hi + 1] - rs;
    }
    int incl = rn;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    const int ex = incl - rn;
    for (int g0 = 0; g0 < total; g0 += 256) {
        int2 pr[4];
        bool live[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int g = g0 + 64 * u + lane;
            live[u] = g < total;
            int row = 0;  // the last row whose prefix is <= g
            for (int r = 1; r < nrows; ++r) row = __builtin_amdgcn_readlane(ex, r) <= g ? r : row;
            const int idx = __shfl(rs, row) + g - __shfl(ex, row);
            pr[u] = live[u] ? rp[(size_t)(a_lo + row) * p.nx * G + idx] : int2{0, 0};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) append(pr[u], live[u]);
    }
    wave_sync_lds();
    agg_consume_list<NP>(rec, nrec, B, lc, dy, dx, y0, x0, c, d, acc);
    wave_sync_lds();
}

// Runs agg_tile_lds for the processed planes (k-th processed plane -> its
// contribution and weight arrays, selected among values so that no
// kernel-argument array is indexed at run time); returns their number.
__device__ __forceinline__ int agg_tile_any(const StageParams& p, bool p0, bool p1, bool p2,
                                            const int* __restrict__ count, const int* __restrict__ pos,
                                            const int* __restrict__ frame, int fr, int a_lo, int a_hi, int c_lo,
                                            int c_hi, int y0, int x0, const float* c0, const float* c1,
                                            const float* c2, const float* d0, const float* d1, const float* d2,
                                            float* acc, int* rec) {
    const int np = (int)p0 + (int)p1 + (int)p2;
    const float* c[3];
    const float* d[3];
    c[0] = p0 ? c0 : (p1 ? c1 : c2);
    d[0] = p0 ? d0 : (p1 ? d1 : d2);
    c[1] = (p0 && p1) ? c1 : c2;
    d[1] = (p0 && p1) ? d1 : d2;
    c[2] = c2;
    d[2] = d2;
    if (np == 3) agg_tile_lds<3>(p, count, pos, frame, fr, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc, rec);
    else if (np == 2) agg_tile_lds<2>(p, count, pos, frame, fr, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc, rec);
    else if (np == 1) agg_tile_lds<1>(p, count, pos, frame, fr, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc, rec);
    return np;
}

// agg_tile_rows for the processed planes (selection as in agg_tile_any).
__device__ __forceinline__ void agg_rows_any(const StageParams& p, bool p0, bool p1, bool p2, const int* ro,
                                             const int2* rp, int a_lo, int a_hi, int c_lo, int c_hi, int y0, int x0,
                                             const float* c0, const float* c1, const float* c2, const float* d0,
                                             const float* d1, const float* d2, float* acc, int* rec) {
    const int np = (int)p0 + (int)p1 + (int)p2;
    const float* c[3];
    const float* d[3];
    c[0] = p0 ? c0 : (p1 ? c1 : c2);
    d[0] = p0 ? d0 : (p1 ? d1 : d2);
    c[1] = (p0 && p1) ? c1 : c2;
    d[1] = (p0 && p1) ? d1 : d2;
    c[2] = c2;
    d[2] = d2;
    if (np == 3) agg_tile_rows<3>(p, ro, rp, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc, rec);
    else if (np == 2) agg_tile_rows<2>(p, ro, rp, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc, rec);
    else if (np == 1) agg_tile_rows<1>(p, ro, rp, a_lo, a_hi, c_lo, c_hi, y0, x0, c, d, acc, rec);
}

}  // namespace bm3d
