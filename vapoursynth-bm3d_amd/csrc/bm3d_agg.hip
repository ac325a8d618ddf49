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


hipError_t launch_aggregate(const AggArgs& a, hipStream_t st) {
    const int tiles = ((a.p.width + kAggTW - 1) / kAggTW) * ((a.p.height + kAggTH - 1) / kAggTH);
    if (agg_planes(a) <= 1)
        hipLaunchKernelGGL(k_aggregate<1>, dim3(8 * ((tiles + 7) / 8)), dim3(64), 0, st, a);
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
