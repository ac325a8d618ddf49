// Ordered gather aggregation of the group contributions, divide and output
// conversion (single-frame and batched launches).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bm3d_params.h"
#include "bm3d_device.h"
#include "bm3d_kernels.h"

namespace bm3d {

// ---------------------------------------------------------------------------
// K8-K10: ordered gather aggregation, divide and output conversion
// ---------------------------------------------------------------------------
//
// For each output pixel the reference adds float(c * numW) into num and
// denW into den in (reference-block scan order, group member order)
// (BlockGroup::AddTo / CountTo, include/Block.h:1239-1259, 1326-1345).  One
// wave owns a 32x16 tile of accumulators in LDS and walks the reference blocks
// that can reach it in scan order, so every pixel sees its contributions in
// exactly that order without atomics; then dst = num / den
// (BM3D_Base.cpp:455-469).

// Writes the filtered float triple/single to the destination format.
__device__ __forceinline__ void write_output(const OutputSpec& o, int x, int y, const float* v, const bool* proc) {
    if (o.mode == kOutFloat) {
#pragma unroll
        for (int pl = 0; pl < 3; ++pl)
            if (proc[pl]) static_cast<float*>(o.dst[pl])[(size_t)y * o.pitch[pl] + x] = v[pl];
    } else if (o.mode == kOutInt) {
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
            if (!proc[pl]) continue;
            float t = v[pl] * o.gain[pl] + o.offset[pl];
            t = clipf(t, o.lo[pl], o.hi[pl]);
            store_sample(o.dst[pl], o.bits, (size_t)y * o.pitch[pl] + x, t);
        }
    } else {  // kOutRGB: FloatYUV2RGB OPP (Conversion.hpp:369-392)
        const float Y = v[0], U = v[1], V = v[2];
        float r = o.m[0] * Y + o.m[1] * U + o.m[2] * V + o.offset[0];
        float g = o.m[3] * Y + o.m[5] * V + o.offset[1];
        float b = o.m[6] * Y + o.m[7] * U + o.m[8] * V + o.offset[2];
        if (o.clip) {
            r = clipf(r, o.lo[0], o.hi[0]);
            g = clipf(g, o.lo[0], o.hi[0]);
            b = clipf(b, o.lo[0], o.hi[0]);
        }
        size_t idx = (size_t)y * o.pitch[0] + x;
        store_sample(o.dst[0], o.bits, idx, r);
        store_sample(o.dst[1], o.bits, idx, g);
        store_sample(o.dst[2], o.bits, idx, b);
    }
}

// Tiles are dealt to the eight XCDs in contiguous runs (workgroup i runs on
// XCD i % 8), so neighbouring tiles, which read overlapping contributions,
// share an L2.
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per_xcd = (ntiles + 7) >> 3;
    return (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
}

// Output tile `tile` of one frame (the body of k_aggregate and its batched form).
__device__ __forceinline__ void aggregate_tile(const AggArgs& a, int tile, float* acc, int* rec) {
    const int lane = lane_id();
    const int tiles_x = (a.p.width + kAggTW - 1) / kAggTW;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int y0 = ty * kAggTH, x0 = tx * kAggTW;
    int a_lo, a_hi, c_lo, c_hi;
    reach_rect(a.p, y0, x0, kAggTH, kAggTW, a.reach, a_lo, a_hi, c_lo, c_hi);
    bool proc[3] = {a.process[0] != 0, a.process[1] != 0, a.process[2] != 0};
    agg_tile_any(a.p, proc[0], proc[1], proc[2], a.count, a.pos, nullptr, 0, a_lo, a_hi, c_lo, c_hi, y0, x0,
                 a.contrib[0], a.contrib[1], a.contrib[2], a.denw[0], a.denw[1], a.denw[2], acc, rec);
    for (int i = lane; i < kAggTH * kAggTW; i += 64) {
        const int ly = i / kAggTW, lx = i - ly * kAggTW;
        const int py = y0 + ly, px = x0 + lx;
        if (py >= a.p.height || px >= a.p.width) continue;
        float v[3];
        int k = 0;
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
            const float* an = acc + k * 2 * kAggPlane + ly * kAggStride + lx;
            v[pl] = proc[pl] ? an[0] / an[kAggPlane] : 0.0f;
            k += proc[pl] ? 1 : 0;
        }
        write_output(a.out, px, py, v, proc);
    }
}

// Processed planes of an aggregation (0..3).
__host__ __device__ __forceinline__ int agg_planes(const AggArgs& a) {
    return (a.process[0] != 0) + (a.process[1] != 0) + (a.process[2] != 0);
}

// PL: the planes the LDS accumulators are sized for (1 for Gray launches:
// a third of the three-plane LDS, so more waves fit on a CU; 3 otherwise).
// A frame with more processed planes returns without touching LDS (the host
// never launches one).
template <int PL>
__global__ __launch_bounds__(64) void k_aggregate(AggArgs a) {
    __shared__ float acc[PL * 2 * kAggPlane];
    __shared__ int rec[kAggRecCap * kAggRecWords];
    if (agg_planes(a) > PL) return;
    const int ntiles = ((a.p.width + kAggTW - 1) / kAggTW) * ((a.p.height + kAggTH - 1) / kAggTH);
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    aggregate_tile(a, tile, acc, rec);
}

// Batched form: the output tiles of every frame in one launch.
template <int PL>
__global__ __launch_bounds__(64) void k_aggregate_batch(const AggArgs* __restrict__ aa, const int* __restrict__ starts,
                                                        int nframes, int total) {
    __shared__ float acc[PL * 2 * kAggPlane];
    __shared__ int rec[kAggRecCap * kAggRecWords];
    const int t = xcd_item(total, 1, 0);
    if (t < 0) return;
    const int f = batch_frame(starts, nframes, t);
    const AggArgs a = aa[f];
    if (agg_planes(a) > PL) return;
    aggregate_tile(a, t - uniform(starts[f]), acc, rec);
}


// Three processed planes: one workgroup of three waves per tile, wave k
// accumulating processed plane k in its own LDS accumulators (a third of the
// one-wave kernel's LDS per wave, so three times the waves fit on a CU).  The
// walk is shared: per round each wave walks kSwWalk steps, the waves append
// the members overlapping the tile to one list in LDS in scan order (each
// wave's records after the earlier waves' counts), and every wave applies
// the whole list to its plane in that order (agg_consume_list).  Block sizes
// up to 8 (B * B <= 64 elements, one per lane).
constexpr int kSwWalk = 2;                        // walk steps per wave and round
constexpr int kSwCap = 3 * kSwWalk * 64;          // records per round, at most

__global__ __launch_bounds__(192) void k_aggregate3(AggArgs a) {
    __shared__ float acc[3 * 2 * kAggPlane];
    __shared__ int cnt[3];
    __shared__ int rec[kSwCap * kAggRecWords];
    const int ntiles = ((a.p.width + kAggTW - 1) / kAggTW) * ((a.p.height + kAggTH - 1) / kAggTH);
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const int lane = lane_id();
    const int w = (int)(threadIdx.x >> 6);
    const int tiles_x = (a.p.width + kAggTW - 1) / kAggTW;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int y0 = ty * kAggTH, x0 = tx * kAggTW;
    int a_lo, a_hi, c_lo, c_hi;
    reach_rect(a.p, y0, x0, kAggTH, kAggTW, a.reach, a_lo, a_hi, c_lo, c_hi);
    // wave w: processed plane w (all three planes are processed)
    const float* cw = w == 0 ? a.contrib[0] : (w == 1 ? a.contrib[1] : a.contrib[2]);
    const float* dw = w == 0 ? a.denw[0] : (w == 1 ? a.denw[1] : a.denw[2]);
    float* accw = acc + w * 2 * kAggPlane;
    for (int i = lane; i < 2 * kAggPlane; i += 64) accw[i] = 0.0f;
    const int B = a.p.block, G = a.p.group, BB = B * B;
    const int gp = G <= 1 ? 1 : G <= 2 ? 2 : G <= 4 ? 4 : G <= 8 ? 8 : G <= 16 ? 16 : G <= 32 ? 32 : 64;
    const int per = 64 / gp;
    const int ncol = c_hi - c_lo + 1;
    const int chunks_per_row = (ncol + per - 1) / per;
    const int zchunks = (G + 63) / 64;
    const int total = (a_hi - a_lo + 1) * chunks_per_row * zchunks;
    const int y1 = y0 + kAggTH - 1, x1 = x0 + kAggTW - 1;
    const int lc = lane < BB ? lane : 0;
    const int dy = lc / B, dx = lc - (lc / B) * B;
    for (int it0 = 0; it0 < total; it0 += 3 * kSwWalk) {
        int sb[kSwWalk], sz[kSwWalk], spk[kSwWalk], sn[kSwWalk];
        bool ov[kSwWalk];
        uint64_t msk[kSwWalk];
        int mine = 0;
#pragma unroll
        for (int u = 0; u < kSwWalk; ++u) {
            const int it = it0 + w * kSwWalk + u;
            const int zc = it % zchunks;
            const int rc = it / zchunks;
            const int ar = a_lo + rc / chunks_per_row;
            const int ac = c_lo + (rc % chunks_per_row) * per + lane / gp;
            const int z = zc * 64 + lane % gp;
            const bool live = it < total && ac <= c_hi && z < G;
            sb[u] = live ? ar * a.p.nx + ac : 0;
            sz[u] = live ? z : 0;
            ov[u] = live;
        }
#pragma unroll
        for (int u = 0; u < kSwWalk; ++u) {
            sn[u] = a.count[sb[u]];
            spk[u] = a.pos[(size_t)sb[u] * G + sz[u]];
        }
#pragma unroll
        for (int u = 0; u < kSwWalk; ++u) {
            const int qy = spk[u] >> 16, qx = spk[u] & 0xffff;
            ov[u] = ov[u] && sz[u] < sn[u] && qy <= y1 && qy + B > y0 && qx <= x1 && qx + B > x0;
            msk[u] = __ballot(ov[u]);
            mine += __popcll(msk[u]);
        }
        if (lane == 0) cnt[w] = mine;
        __syncthreads();
        const int c0 = cnt[0], c1 = cnt[1], c2 = cnt[2];
        int at = w == 0 ? 0 : (w == 1 ? c0 : c0 + c1);
        const int ntot = c0 + c1 + c2;
#pragma unroll
        for (int u = 0; u < kSwWalk; ++u) {
            if (ov[u]) {
                const int i = at + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(msk[u] >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)msk[u], 0));
                int* e = rec + i * kAggRecWords;
                e[0] = spk[u];
                e[1] = sb[u] * G + sz[u];
                e[2] = sb[u];
            }
            at += __popcll(msk[u]);
        }
        __syncthreads();
        agg_consume_list<1>(rec, ntot, B, lc, dy, dx, y0, x0, &cw, &dw, accw);
        __syncthreads();  // the list is rewritten by the next round
    }
    __syncthreads();  // every plane's sums are complete
    const bool proc[3] = {true, true, true};
    for (int i = (int)threadIdx.x; i < kAggTH * kAggTW; i += 192) {
        const int ly = i / kAggTW, lx = i - ly * kAggTW;
        const int py = y0 + ly, px = x0 + lx;
        if (py >= a.p.height || px >= a.p.width) continue;
        float v[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
            const float* an = acc + pl * 2 * kAggPlane + ly * kAggStride + lx;
            v[pl] = an[0] / an[kAggPlane];
        }
        write_output(a.out, px, py, v, proc);
    }
}

hipError_t launch_aggregate(const AggArgs& a, hipStream_t st) {
    const int tiles = ((a.p.width + kAggTW - 1) / kAggTW) * ((a.p.height + kAggTH - 1) / kAggTH);
    const int np = agg_planes(a);
    if (np <= 1)
        hipLaunchKernelGGL(k_aggregate<1>, dim3(8 * ((tiles + 7) / 8)), dim3(64), 0, st, a);
    else if (np == 3 && a.p.block * a.p.block <= 64)
        hipLaunchKernelGGL(k_aggregate3, dim3(8 * ((tiles + 7) / 8)), dim3(192), 0, st, a);
    else
        hipLaunchKernelGGL(k_aggregate<3>, dim3(8 * ((tiles + 7) / 8)), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_aggregate_batch(const AggArgs* aa, const int* starts, int nframes, int total, int planes,
                                  const BatchShape& sh, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (planes < 1 || planes > 3) return hipErrorInvalidValue;
    const int groups = batch_groups(total, 1);
    if (planes == 1)
        hipLaunchKernelGGL(k_aggregate_batch<1>, dim3(groups), dim3(64), 0, st, aa, starts, nframes, total);
    else
        hipLaunchKernelGGL(k_aggregate_batch<3>, dim3(groups), dim3(64), 0, st, aa, starts, nframes, total);
    return hipGetLastError();
}

int aggregate_tiles(int W, int H) { return ((W + kAggTW - 1) / kAggTW) * ((H + kAggTH - 1) / kAggTH); }

}  // namespace bm3d
