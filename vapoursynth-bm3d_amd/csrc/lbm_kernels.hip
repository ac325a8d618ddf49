// Learned block-matching proposal (LBM) on CDNA4: a patch-embedding MLP
// 64 -> 128 (ReLU) -> 32 with BF16 weights and FP32 accumulation on MFMA
// (v_mfma_f32_16x16x16bf16_1k), its training kernels (forward, loss,
// backward, Adam), and the proposal path used by block matching:
// embeddings of the even-lattice patch positions and of the reference blocks,
// FP16 embedding-distance scores of the even lattice of each reference
// block's search window, the best lattice positions' 3x3 neighbourhoods as
// candidates, exact FP32 SSD re-rank with the plain matcher's acceptance and
// ordering rules, and a list of reference blocks that fall back to the
// exhaustive search.
//
// MFMA 16x16x16 operand layout (lane l, element i = 0..3):
//   A[m][k]: m = l % 16, k = 4 * (l / 16) + i
//   B[k][n]: n = l % 16, k = 4 * (l / 16) + i
//   D[m][n]: n = l % 16, m = 4 * (l / 16) + i
// The network is evaluated transposed, H^T = W1 X^T and E^T = W2 H^T, with
// the 16 patches of a wave along n: the D fragments of the first layer are
// then exactly the B fragments of the second.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bm3d_params.h"
#include "bm3d_device.h"
#include "lbm_kernels.h"

namespace bm3d {
namespace lbm {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Weight fragments of both layers, held in registers for the whole wave.
struct Frags {
    s4 w1[8][4];  // A of layer 1: tile t (hidden 16t..), k-step s (pixels 16s..)
    s4 w2[2][8];  // A of layer 2: tile u (embedding 16u..), k-step s (hidden 16s..)
    float b1[8][4];
    float b2[2][4];
};

__device__ __forceinline__ void load_frags(const Net& n, Frags& f) {
    const int lane = lane_id();
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const short* p = n.w1h + (16 * t + r) * kIn + 16 * s + 4 * q;
#pragma unroll
            for (int i = 0; i < 4; ++i) f.w1[t][s][i] = p[i];
        }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const short* p = n.w2h + (16 * u + r) * kHidden + 16 * s + 4 * q;
#pragma unroll
            for (int i = 0; i < 4; ++i) f.w2[u][s][i] = p[i];
        }
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) f.b1[t][i] = n.b1[16 * t + 4 * q + i];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) f.b2[u][i] = n.b2[16 * u + 4 * q + i];
}

// Patch pixels of the lane's patch (top-left (py, px)) for k-step s: the
// four pixels k = 16s + 4q .. +3, i.e. row 2s + q/2, columns 4(q&1) .. +3.
__device__ __forceinline__ s4 patch_frag(const float* __restrict__ plane, int pitch, int py, int px, int s,
                                         float* raw) {
    const int q = lane_id() >> 4;
    const float* p = plane + (size_t)(py + 2 * s + (q >> 1)) * pitch + px + 4 * (q & 1);
    float4 f;
    __builtin_memcpy(&f, p, sizeof(f));  // four consecutive samples (4-byte aligned)
    raw[0] = f.x;
    raw[1] = f.y;
    raw[2] = f.z;
    raw[3] = f.w;
    s4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = f2bf(raw[i]);
    return v;
}

// Evaluates the network on 16 patches (one per lane % 16).  x: the B
// fragments of the patches; h: post-ReLU hidden (D layout, fp32); e: the
// embeddings (D layout: lane l holds embedding 16u + 4q + i of patch l % 16).
__device__ __forceinline__ void forward16(const Frags& f, const s4 (&x)[4], float (&h)[8][4], float (&e)[2][4]) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {f.b1[t][0], f.b1[t][1], f.b1[t][2], f.b1[t][3]};  // bias as the initial sum
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(f.w1[t][s], x[s], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) h[t][i] = acc[i] > 0.0f ? acc[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        f4 acc = {f.b2[u][0], f.b2[u][1], f.b2[u][2], f.b2[u][3]};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            s4 hb;
#pragma unroll
            for (int i = 0; i < 4; ++i) hb[i] = f2bf(h[s][i]);
            acc = mfma(f.w2[u][s], hb, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) e[u][i] = acc[i];
    }
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------

// Forward pass of a training batch: patch 2p is pair p's reference block,
// 2p+1 its candidate.  Saves X (bf16), H (fp32) and E (fp32) per patch.
__global__ __launch_bounds__(256) void k_train_fwd(Net n, TrainBatch tb) {
    const int lane = lane_id();
    const int wave = uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int npatch = 2 * tb.npairs;
    const int p0 = wave * 16;
    if (p0 >= npatch) return;
    Frags f;
    load_frags(n, f);
    const int r = lane & 15, q = lane >> 4;
    const int patch = p0 + r < npatch ? p0 + r : npatch - 1;
    const int4 pr = tb.pairs[patch >> 1];
    const int py = (patch & 1) ? pr.z : pr.x;
    const int px = (patch & 1) ? pr.w : pr.y;
    s4 x[4];
    float raw[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        x[s] = patch_frag(tb.plane, tb.pitch, py, px, s, raw);
        if (p0 + r < npatch) {
            short* xs = tb.xs + (size_t)patch * kIn + 16 * s + 4 * q;
#pragma unroll
            for (int i = 0; i < 4; ++i) xs[i] = x[s][i];
        }
    }
    float h[8][4], e[2][4];
    forward16(f, x, h, e);
    if (p0 + r < npatch) {
#pragma unroll
        for (int t = 0; t < 8; ++t)
            *reinterpret_cast<float4*>(tb.hs + (size_t)patch * kHidden + 16 * t + 4 * q) =
                make_float4(h[t][0], h[t][1], h[t][2], h[t][3]);
#pragma unroll
        for (int u = 0; u < 2; ++u)
            *reinterpret_cast<float4*>(tb.es + (size_t)patch * kEmb + 16 * u + 4 * q) =
                make_float4(e[u][0], e[u][1], e[u][2], e[u][3]);
    }
}

// Loss per pair: w (d - t)^2 with d = |e_r - e_q|^2, w = 1 / (t + t0)^2;
// writes dL/dE for both patches (mean over pairs) and the per-block loss sum.
__global__ __launch_bounds__(256) void k_train_loss(TrainBatch tb, float t0, float* block_loss) {
    __shared__ float red[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float lp = 0.0f;
    if (p < tb.npairs) {
        const float* er = tb.es + (size_t)(2 * p) * kEmb;
        const float* eq = er + kEmb;
        float diff[kEmb];
        float d = 0.0f;
#pragma unroll
        for (int k = 0; k < kEmb; ++k) {
            diff[k] = er[k] - eq[k];
            d += diff[k] * diff[k];
        }
        const float t = tb.target[p];
        const float w = 1.0f / ((t + t0) * (t + t0));
        const float res = d - t;
        lp = w * res * res;
        const float g = 2.0f * w * res / (float)tb.npairs;
        float* gr = tb.des + (size_t)(2 * p) * kEmb;
        float* gq = gr + kEmb;
#pragma unroll
        for (int k = 0; k < kEmb; ++k) {
            gr[k] = 2.0f * g * diff[k];
            gq[k] = -2.0f * g * diff[k];
        }
    }
    red[threadIdx.x] = lp;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_loss[blockIdx.x] = red[0];
}

// dH_pre = (dE W2) * [H > 0], as dH^T = W2^T dE^T on MFMA, 16 patches per wave.
__global__ __launch_bounds__(256) void k_train_bwd_h(Net n, TrainBatch tb) {
    const int lane = lane_id();
    const int wave = uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int npatch = 2 * tb.npairs;
    const int p0 = wave * 16;
    if (p0 >= npatch) return;
    const int r = lane & 15, q = lane >> 4;
    const int patch = p0 + r < npatch ? p0 + r : npatch - 1;
    s4 b[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const float4 g = *reinterpret_cast<const float4*>(tb.des + (size_t)patch * kEmb + 16 * s + 4 * q);
        b[s][0] = f2bf(g.x);
        b[s][1] = f2bf(g.y);
        b[s][2] = f2bf(g.z);
        b[s][3] = f2bf(g.w);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            s4 a;  // W2^T[16t + r][16s + 4q + i] = W2[16s + 4q + i][16t + r]
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = n.w2h[(16 * s + 4 * q + i) * kHidden + 16 * t + r];
            acc = mfma(a, b[s], acc);
        }
        if (p0 + r < npatch) {
            const float4 hv = *reinterpret_cast<const float4*>(tb.hs + (size_t)patch * kHidden + 16 * t + 4 * q);
            const float4 o = make_float4(hv.x > 0.0f ? acc[0] : 0.0f, hv.y > 0.0f ? acc[1] : 0.0f,
                                         hv.z > 0.0f ? acc[2] : 0.0f, hv.w > 0.0f ? acc[3] : 0.0f);
            *reinterpret_cast<float4*>(tb.dhs + (size_t)patch * kHidden + 16 * t + 4 * q) = o;
        }
    }
}

// Weight gradients of a slice of kSlice patches per workgroup (partials):
// dW1 = dH_pre^T X, dW2 = dE^T H on MFMA with the patches as the k axis;
// db1, db2 = column sums.  Output: part[blockIdx.x][kParams] in parameter
// order [W1 | b1 | W2 | b2].
constexpr int kSlice = 256;
__global__ __launch_bounds__(256) void k_train_bwd_w(TrainBatch tb, float* part) {
    const int lane = lane_id();
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int npatch = 2 * tb.npairs;
    const int s0 = blockIdx.x * kSlice;
    const int r = lane & 15, q = lane >> 4;
    float* out = part + (size_t)blockIdx.x * kParams;
    // 48 output tiles: 32 of dW1 (8 hidden x 4 pixel tiles), 16 of dW2 (2 x 8)
    for (int tile = wave; tile < 48; tile += 4) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = s0; k0 < s0 + kSlice && k0 < npatch; k0 += 16) {
            s4 a, b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int pk = k0 + 4 * q + i;
                const bool ok = pk < npatch;
                if (tile < 32) {
                    const int mt = tile >> 2, nt = tile & 3;
                    a[i] = ok ? f2bf(tb.dhs[(size_t)pk * kHidden + 16 * mt + r]) : (short)0;
                    b[i] = ok ? tb.xs[(size_t)pk * kIn + 16 * nt + r] : (short)0;
                } else {
                    const int mt = (tile - 32) >> 3, nt = (tile - 32) & 7;
                    a[i] = ok ? f2bf(tb.des[(size_t)pk * kEmb + 16 * mt + r]) : (short)0;
                    b[i] = ok ? f2bf(tb.hs[(size_t)pk * kHidden + 16 * nt + r]) : (short)0;
                }
            }
            acc = mfma(a, b, acc);
        }
        // D[m][n]: m = 4q + i, n = r
        if (tile < 32) {
            const int mt = tile >> 2, nt = tile & 3;
#pragma unroll
            for (int i = 0; i < 4; ++i) out[kOffW1 + (16 * mt + 4 * q + i) * kIn + 16 * nt + r] = acc[i];
        } else {
            const int mt = (tile - 32) >> 3, nt = (tile - 32) & 7;
#pragma unroll
            for (int i = 0; i < 4; ++i) out[kOffW2 + (16 * mt + 4 * q + i) * kHidden + 16 * nt + r] = acc[i];
        }
    }
    // bias gradients: threads 0..127 -> b1, 128..159 -> b2
    const int t = threadIdx.x;
    if (t < kHidden + kEmb) {
        float sacc = 0.0f;
        for (int pk = s0; pk < s0 + kSlice && pk < npatch; ++pk)
            sacc += t < kHidden ? tb.dhs[(size_t)pk * kHidden + t] : tb.des[(size_t)pk * kEmb + (t - kHidden)];
        if (t < kHidden) out[kOffB1 + t] = sacc;
        else out[kOffB2 + t - kHidden] = sacc;
    }
}

// grad[i] = sum over slices of part[s][i], in slice order.
__global__ void k_train_reduce(const float* part, int nslices, float* grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float s = 0.0f;
    for (int k = 0; k < nslices; ++k) s += part[(size_t)k * kParams + i];
    grad[i] = s;
}

// Adam on the FP32 master parameters; refreshes the BF16 weight copies.
__global__ void k_train_adam(Net n, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                             float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    float w = n.master[i] - upd;
    n.master[i] = w;
    if (i < kOffB1) n.w1h[i] = f2bf(w);
    else if (i >= kOffW2 && i < kOffB2) n.w2h[i - kOffW2] = f2bf(w);
}

// BF16 copies of the master weights (after a host upload).
__global__ void k_sync_bf16(Net n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float w = n.master[i];
    if (i < kOffB1) n.w1h[i] = f2bf(w);
    else if (i >= kOffW2 && i < kOffB2) n.w2h[i - kOffW2] = f2bf(w);
}

// SSD (plain fp32, row-major) of each pair's two 8x8 patches: the training target.
__global__ void k_pair_ssd(const float* __restrict__ plane, int pitch, const int4* __restrict__ pairs, int npairs,
                           float* __restrict__ target) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const int4 pr = pairs[p];
    float s = 0.0f;
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            const float d = plane[(size_t)(pr.x + y) * pitch + pr.y + x] - plane[(size_t)(pr.z + y) * pitch + pr.w + x];
            s += d * d;
        }
    target[p] = s;
}

// ---------------------------------------------------------------------------
// inference: embeddings of the even lattice and of the reference blocks
// ---------------------------------------------------------------------------

// The proposal lattice: every kLat-th row and column.  lat_count(n): lattice
// points among positions 0..n-1; lat_first(v): the first lattice coordinate
// >= v; lat_span(f, b): lattice points from f (a lattice coordinate) to b.
__host__ __device__ constexpr int lat_count(int n) { return n > 0 ? (n + kLat - 1) / kLat : 0; }
__host__ __device__ constexpr int lat_first(int v) { return (v + kLat - 1) / kLat * kLat; }
__host__ __device__ constexpr int lat_span(int f, int b) { return f <= b ? (b - f) / kLat + 1 : 0; }

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));
typedef short sh2 __attribute__((ext_vector_type(2)));

// bf16(max(v, 0)) of two values: one packed conversion, then ReLU on the
// packed bf16 bits as signed 16-bit integers (negative values and -0 become
// +0), which equals rounding max(v, 0).
__device__ __forceinline__ sh2 relu_bf16x2(float a, float b) {
    const fl2 v = {a, b};
    const sh2 z = {0, 0};
    return __builtin_elementwise_max(__builtin_bit_cast(sh2, __builtin_convertvector(v, bf2)), z);
}

// Stores the embedding (fp16) and its squared norm (fp32 of the fp16 values)
// of the lane's patch at index idx.  Patches without an output (past the end
// of a run or a list) store to the spare slot at index `spare`, so the stores
// are unconditional and a run of patches stays one straight-line block.
__device__ __forceinline__ void store_emb(const float (&e)[2][4], int idx, int spare, _Float16* __restrict__ emb,
                                          float* __restrict__ norm) {
    const int q = lane_id() >> 4;
    float nrm = 0.0f;
    h4 eh[2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            eh[u][i] = (_Float16)e[u][i];
            const float ev = (float)eh[u][i];
            nrm += ev * ev;
        }
    // lanes r, r+16, r+32, r+48 hold the four quarters of one embedding; the
    // butterfly leaves the same sum in all four (fp addition commutes), and
    // all four store it
    nrm += __shfl_xor(nrm, 16);
    nrm += __shfl_xor(nrm, 32);
    const size_t at = (size_t)(idx >= 0 ? idx : spare);
#pragma unroll
    for (int u = 0; u < 2; ++u) __builtin_memcpy(emb + at * kEmb + 16 * u + 4 * q, &eh[u], sizeof(h4));
    norm[at] = nrm;
}

// The lattice embeddings run the network on v_mfma_f32_16x16x32_bf16 (half
// the instructions of the 16x16x16 form; same bf16 operands and roundings of
// the activations, FP32 accumulation in the instruction's own order).  A k-step
// pairs lane group q's eight values of A and B, so the contraction index can be
// assigned freely as long as both operands use the same assignment:
//   layer 1, k-step s: lane group q holds pixel row 2q + s (eight columns);
//   layer 2, k-step s: lane group q holds hidden units 16 (2s + j / 4) + 4q +
//   j % 4, j = 0..7 -- exactly the layer-1 outputs the lane already holds.
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

struct Frags32 {
    s8 w1[8][2];  // layer 1: tile t (hidden 16t..), k-step s
    s8 w2[2][4];  // layer 2: tile u (embedding 16u..), k-step s
    float b1[8][4];
    float b2[2][4];
};

__device__ __forceinline__ void load_frags32(const Net& n, Frags32& f) {
    const int lane = lane_id();
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int s = 0; s < 2; ++s) __builtin_memcpy(&f.w1[t][s], n.w1h + (16 * t + r) * kIn + (2 * q + s) * 8, 16);
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const short* row = n.w2h + (16 * u + r) * kHidden;
            s4 lo, hi;
            __builtin_memcpy(&lo, row + 16 * (2 * s) + 4 * q, 8);
            __builtin_memcpy(&hi, row + 16 * (2 * s + 1) + 4 * q, 8);
            f.w2[u][s] = s8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) f.b1[t][i] = n.b1[16 * t + 4 * q + i];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) f.b2[u][i] = n.b2[16 * u + 4 * q + i];
}

// The network on 16 patches: x[s] = the lane's pixel row 2q + s (bf16).
__device__ __forceinline__ void infer32(const Frags32& f, const s8 (&x)[2], float (&e)[2][4]) {
    s4 hb[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {f.b1[t][0], f.b1[t][1], f.b1[t][2], f.b1[t][3]};
        acc = mfma32(f.w1[t][0], x[0], acc);
        acc = mfma32(f.w1[t][1], x[1], acc);
        const sh2 lo = relu_bf16x2(acc[0], acc[1]), hi = relu_bf16x2(acc[2], acc[3]);
        hb[t] = s4{lo[0], lo[1], hi[0], hi[1]};
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        f4 acc = {f.b2[u][0], f.b2[u][1], f.b2[u][2], f.b2[u][3]};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const s4 a = hb[2 * s], c = hb[2 * s + 1];
            acc = mfma32(f.w2[u][s], s8{a[0], a[1], a[2], a[3], c[0], c[1], c[2], c[3]}, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) e[u][i] = acc[i];
    }
}

// four pixels -> bf16 (round to nearest even, packed conversion)
__device__ __forceinline__ s4 frag_cvt(const float4& v) {
    const fl2 a = {v.x, v.y}, b = {v.z, v.w};
    const sh2 lo = __builtin_bit_cast(sh2, __builtin_convertvector(a, bf2));
    const sh2 hi = __builtin_bit_cast(sh2, __builtin_convertvector(b, bf2));
    return s4{lo[0], lo[1], hi[0], hi[1]};
}

// eight pixels (two float4) -> bf16
__device__ __forceinline__ s8 row_cvt(const float4& a, const float4& b) {
    const s4 lo = frag_cvt(a), hi = frag_cvt(b);
    return s8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// Embeddings of the lattice: patch (kLat yy, kLat xx) -> index yy * ew2 + xx,
// ew2 = ceil((W - 7) / kLat), for yy < ceil((H - 7) / kLat).  One wave per
// column of 16 lattice positions and run of kEmbedRun lattice rows.  Lane
// group q holds pixel rows 2q and 2q + 1 of its patch, so one lattice row
// further down (kLat = 2 pixel rows) its rows are those group q + 1 held: a
// 16-lane shift, and group 3 loads the two new rows, issued kEmbedAhead
// lattice rows ahead.
constexpr int kEmbedRun = 32;
constexpr int kEmbedAhead = 4;
static_assert(kLat == 2, "row reuse between lattice rows assumes a stride of two rows");

__device__ __forceinline__ void embed_lattice_wg(int wg, const Net& n, const float* __restrict__ plane, int pitch,
                                                 int W, int H, _Float16* __restrict__ emb, float* __restrict__ norm) {
    const int ew2 = lat_count(W - 7), eh2 = lat_count(H - 7);
    const int gx = (ew2 + 15) / 16;
    const int runs = (eh2 + kEmbedRun - 1) / kEmbedRun;
    const int wave = uniform((int)(wg * 4 + (threadIdx.x >> 6)));
    if (wave >= gx * runs) return;
    const int ry = wave / gx;
    const int lane = lane_id();
    const int r = lane & 15, q = lane >> 4;
    const int xx = (wave - ry * gx) * 16 + r;
    const bool xok = xx < ew2;
    const float* col = plane + kLat * (xok ? xx : ew2 - 1);
    const int yy0 = ry * kEmbedRun;
    const int nrows = eh2 - yy0 < kEmbedRun ? eh2 - yy0 : kEmbedRun;
    const int y0 = kLat * yy0;
    // first pixel row group 3 loads for lattice row yy0 + k (k >= 1), clamped
    // for rows past the lattice (loaded, never used)
    auto new_row = [&](int k) {
        const int y = y0 + kLat * k + 6;
        return y < H - 1 ? y : H - 2;
    };
    Frags32 f;
    load_frags32(n, f);
    s8 x[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        float4 a, b;
        const float* src = col + (size_t)(y0 + 2 * q + s) * pitch;
        __builtin_memcpy(&a, src, sizeof(a));
        __builtin_memcpy(&b, src + 4, sizeof(b));
        x[s] = row_cvt(a, b);
    }
    // (every lane loads the new rows, so the loads stay outside branches and
    // in flight; group 3 uses them)
    float4 ahead[kEmbedAhead][4];
#pragma unroll
    for (int k = 0; k < kEmbedAhead; ++k) {
            const float* src = col + (size_t)new_row(k + 1) * pitch;
            __builtin_memcpy(&ahead[k][0], src, 16);
            __builtin_memcpy(&ahead[k][1], src + 4, 16);
            __builtin_memcpy(&ahead[k][2], src + pitch, 16);
            __builtin_memcpy(&ahead[k][3], src + pitch + 4, 16);
        }
    // every row of the run is evaluated (rows past the lattice store nothing):
    // no exit inside the unrolled body keeps the loads in flight across rows
#pragma unroll
    for (int k = 0; k < kEmbedRun; ++k) {
        float e[2][4];
        infer32(f, x, e);
        store_emb(e, xok && k < nrows ? (yy0 + k) * ew2 + xx : -1, ew2 * eh2, emb, norm);
        if (k + 1 == kEmbedRun) break;
        // next lattice row: group q takes group q + 1's rows, group 3 the new ones
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            s8 v;
#pragma unroll
            for (int i = 0; i < 8; i += 2) {
                const int w = __shfl_down(__builtin_bit_cast(int, sh2{x[s][i], x[s][i + 1]}), 16);
                const sh2 p2 = __builtin_bit_cast(sh2, w);
                v[i] = p2[0];
                v[i + 1] = p2[1];
            }
            const s8 nv = row_cvt(ahead[k % kEmbedAhead][2 * s], ahead[k % kEmbedAhead][2 * s + 1]);
            x[s] = q == 3 ? nv : v;
        }
        if (k + 1 + kEmbedAhead < kEmbedRun) {
            const float* src = col + (size_t)new_row(k + 1 + kEmbedAhead) * pitch;
            float4 (&d)[4] = ahead[k % kEmbedAhead];
            __builtin_memcpy(&d[0], src, 16);
            __builtin_memcpy(&d[1], src + 4, 16);
            __builtin_memcpy(&d[2], src + pitch, 16);
            __builtin_memcpy(&d[3], src + pitch + 4, 16);
        }
    }
}

// Embeddings of the reference blocks (scan order b = by * nx + bx), with the
// same network evaluation as the lattice: each wave embeds kRefGroups groups
// of 16 blocks, the next group's pixel rows loaded while a group runs (block
// indices past the end are clamped: computed, not stored).  The first lane of
// the grid also clears the fallback counter of the re-rank that follows.
constexpr int kRefGroups = 8;
__device__ __forceinline__ void embed_refs_wg(int wg, const Net& n, const float* __restrict__ plane,
                                              const StageParams& p, _Float16* __restrict__ emb,
                                              float* __restrict__ norm, int* __restrict__ nfallback) {
    if (wg == 0 && threadIdx.x == 0) *nfallback = 0;
    const int nb = p.nx * p.ny;
    const int wave = uniform((int)(wg * 4 + (threadIdx.x >> 6)));
    const int b0 = wave * 16 * kRefGroups;
    if (b0 >= nb) return;
    const int lane = lane_id();
    const int r = lane & 15, q = lane >> 4;
    auto load = [&](int g, float4 (&v)[4]) {
        const int b = b0 + 16 * g + r;
        const int bc = b < nb ? b : nb - 1;
        const int by = bc / p.nx;
        const int y = grid_pos(by, p.block_step, p.height - 8);
        const int x = grid_pos(bc - by * p.nx, p.block_step, p.width - 8);
        const float* src = plane + (size_t)(y + 2 * q) * p.pitch + x;
        __builtin_memcpy(&v[0], src, 16);
        __builtin_memcpy(&v[1], src + 4, 16);
        __builtin_memcpy(&v[2], src + p.pitch, 16);
        __builtin_memcpy(&v[3], src + p.pitch + 4, 16);
    };
    Frags32 f;
    load_frags32(n, f);
    float4 cur[4], nxt[4];
    load(0, cur);
#pragma unroll
    for (int g = 0; g < kRefGroups; ++g) {
        if (g + 1 < kRefGroups) load(g + 1, nxt);
        const s8 x[2] = {row_cvt(cur[0], cur[1]), row_cvt(cur[2], cur[3])};
        float e[2][4];
        infer32(f, x, e);
        const int b = b0 + 16 * g + r;
        store_emb(e, b < nb ? b : -1, nb, emb, norm);
        if (g + 1 < kRefGroups)
#pragma unroll
            for (int k = 0; k < 4; ++k) cur[k] = nxt[k];
    }
}

// Both embedding passes in one launch: workgroups [0, lat_wgs) embed the
// lattice, the rest the reference blocks (independent work, so the reference
// pass runs beside the lattice pass's last workgroups instead of after them).
__global__ __launch_bounds__(256) void k_embed(Net n, const float* __restrict__ plane, StageParams p, int lat_wgs,
                                               _Float16* __restrict__ emb, float* __restrict__ norm,
                                               _Float16* __restrict__ remb, float* __restrict__ rnorm,
                                               int* __restrict__ nfallback) {
    const int wg = (int)blockIdx.x;
    if (wg < lat_wgs) embed_lattice_wg(wg, n, plane, p.pitch, p.width, p.height, emb, norm);
    else embed_refs_wg(wg - lat_wgs, n, plane, p, remb, rnorm, nfallback);
}

// ---------------------------------------------------------------------------
// inference: proposals, exact re-rank and fallback list
// ---------------------------------------------------------------------------
//
// k_propose: one workgroup (four waves) per tile of kPropRows x kPropCols = 8
// x 2 reference blocks.  Embedding distance: score(r, c) = |e_r|^2 + |e_c|^2 -
// 2 e_r.e_c (fp32; fp16 in the per-block lists).  The dot products of the 16
// blocks with 16 lattice positions of one row of the union of their windows
// are one v_mfma_f32_16x16x32_f16 (a step); tall tiles keep the union small
// (at bm_range 24 and block_step 8: 53 x 29 lattice positions, 106 steps for
// 16 x 625 scores).  Each wave keeps kPropDepth steps' embedding loads in
// flight.  Tiles are dealt to the XCDs in contiguous runs (column-major, so
// vertically adjacent tiles, whose unions overlap most, share an L2).  Every
// even-lattice position (even row, even column) of each block's search window
// is scored, and each block proposes its M = min(K, 32) best lattice
// positions (ties to the earlier position) as a list of block-local lattice
// indices.
//
// k_rerank: the candidates of a block are the 3x3 neighbourhoods of its
// proposals inside the window, self excluded; a position covered by several
// proposals belongs to the first of them in lattice order.  Every candidate
// is re-ranked exactly like the plain matcher (SSE2 lane-order SSD, dist <=
// thSSE && dist != 0, key = float(double(dist) * distMul)) and the G-1 best
// by (key, window index) follow the self entry.  A block with fewer than G-1
// accepted candidates falls back to the exhaustive search, unless every
// lattice position of its window was proposed: the neighbourhoods then cover
// the whole window and the list is the exhaustive one.
constexpr int kPropRows = 8;
constexpr int kPropCols = 2;
constexpr int kPropDepth = 4;  // scoring steps in flight per wave
constexpr int kSubSide = (49 + kLat - 1) / kLat;  // lattice points per side of the largest window
constexpr int kSubCap = kSubSide * kSubSide;  // lattice scores of one block
static_assert(kPropRows * kPropCols == 16, "one MFMA row per block of a tile");

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4 mfma_h32(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Marks v[u] (16-bit values, 0xFFFF = absent; value index lane + 64 u) that
// are among the m smallest by (value, index).
template <int N>
__device__ __forceinline__ uint32_t wave_smallest16(const uint32_t (&v)[N], int m) {
    uint32_t tv = 0;  // the m-th smallest value: the largest tv with count(v < tv) < m
    for (int bit = 15; bit >= 0; --bit) {
        const uint32_t trial = tv | (1u << bit);
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < N; ++u) cnt += __popcll(__ballot(v[u] < trial));
        if (cnt < m) tv = trial;
    }
    int less = 0;
#pragma unroll
    for (int u = 0; u < N; ++u) less += __popcll(__ballot(v[u] < tv));
    const int need_eq = m - less;
    int eq_seen = 0;
    uint32_t sel = 0;
#pragma unroll
    for (int u = 0; u < N; ++u) {
        const bool eq = v[u] == tv && v[u] != 0xFFFFu;
        const uint64_t em = __ballot(eq);
        const int rank = eq_seen + __builtin_amdgcn_mbcnt_hi((uint32_t)(em >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)em, 0));
        if ((v[u] < tv && v[u] != 0xFFFFu) || (eq && rank < need_eq)) sel |= 1u << u;
        eq_seen += __popcll(em);
    }
    return sel;
}

__device__ __forceinline__ int mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Search window of block (by, bx) on the lattice: first lattice row/column
// (plane coordinates) and the number of lattice rows/columns.
struct LatWin {
    int y0, ny, x0, nx;
};
__device__ __forceinline__ LatWin lat_window(const StageParams& p, int by, int bx) {
    const int B = 8;
    const int j = grid_pos(by, p.block_step, p.height - B);
    const int i = grid_pos(bx, p.block_step, p.width - B);
    const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
    const int l = search_lo(i, p.range, 1), rr = search_hi(i, p.width - B, p.range, 1);
    LatWin w;
    w.y0 = lat_first(t);
    w.ny = lat_span(w.y0, bt);
    w.x0 = lat_first(l);
    w.nx = lat_span(w.x0, rr);
    return w;
}

__global__ __launch_bounds__(256) void k_propose(ProposeArgs a) {
    __shared__ unsigned short sub[16][kSubCap];  // lattice scores of the tile's blocks
    __shared__ uint32_t comp[4][128];            // per wave: compacted lattice scores and positions
    const StageParams& p = a.p;
    const int lane = lane_id();
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int r = lane & 15, q = lane >> 4;
    const int G = p.group;
    if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) return;  // k_rerank writes the self-only lists
    // tile of the workgroup (XCD-contiguous runs, column-major)
    const int row_tiles = (p.ny + kPropRows - 1) / kPropRows;
    const int col_tiles = (p.nx + kPropCols - 1) / kPropCols;
    const int ntiles = row_tiles * col_tiles;
    const int per_xcd = (ntiles + 7) >> 3;
    const int lin = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (lin >= ntiles) return;
    const int tcol = lin / row_tiles;
    const int by0 = (lin - tcol * row_tiles) * kPropRows;
    const int bx0 = tcol * kPropCols;
    const int by_last = by0 + kPropRows - 1 < p.ny - 1 ? by0 + kPropRows - 1 : p.ny - 1;
    const int bx_last = bx0 + kPropCols - 1 < p.nx - 1 ? bx0 + kPropCols - 1 : p.nx - 1;
    const int ew2 = lat_count(p.width - 7);
    const int K = a.topk;
    const int M = K < kPropSlots ? K : kPropSlots;
    // block m of the tile: row by0 + m / 2, column bx0 + m % 2 (clamped to the
    // plane's blocks; the copies past the edge score nothing)
    auto tile_block = [&](int m, int& by, int& bx) {
        by = by0 + (m >> 1);
        bx = bx0 + (m & 1);
        const bool ok = by <= by_last && bx <= bx_last;
        by = by <= by_last ? by : by_last;
        bx = bx <= bx_last ? bx : bx_last;
        return ok;
    };
    // union of the windows (lattice units from the first row/column)
    const LatWin wf = lat_window(p, by0, bx0), wl = lat_window(p, by_last, bx_last);
    const int uy0 = wf.y0, ux0 = wf.x0;
    const int nyu = (wl.y0 + kLat * (wl.ny - 1) - uy0) / kLat + 1;
    const int nxu = (wl.x0 + kLat * (wl.nx - 1) - ux0) / kLat + 1;
    const int ncb = nxu > 0 ? (nxu + 15) >> 4 : 0;  // 1 or 2 (nxu <= 29)
    const int ntile = nyu > 0 ? nyu * ncb : 0;
    // A fragments: lane r <-> block r, dims 8q .. 8q+7 (the same k assignment
    // as the B fragments)
    h8 af;
    {
        int by, bx;
        tile_block(r, by, bx);
        __builtin_memcpy(&af, a.remb + ((size_t)by * p.nx + bx) * kEmb + 8 * q, sizeof(h8));
    }
    // D rows 4q + ii <-> blocks 4q + ii: their windows in union lattice units
    int row_lo[4], nly[4], col_lo[4], nlx[4];
    float nref[4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        int by, bx;
        const bool ok = tile_block(4 * q + ii, by, bx);
        const LatWin w = lat_window(p, by, bx);
        row_lo[ii] = (w.y0 - uy0) / kLat;
        col_lo[ii] = (w.x0 - ux0) / kLat;
        nly[ii] = ok ? w.ny : 0;
        nlx[ii] = w.nx;
        nref[ii] = a.rnorm[(size_t)by * p.nx + bx];
    }
    for (int e = threadIdx.x; e < 16 * kSubCap; e += 256) (&sub[0][0])[e] = 0xFFFFu;
    __syncthreads();
    // step t: union lattice row t / ncb, columns 16 (t % ncb) .. +15 (ncb is 1
    // or 2: the union is at most 25 + 4 lattice columns wide); a wave scores
    // steps wave, wave + 4, ..., i.e. a fixed column chunk and every
    // (4 / ncb)-th row.  Loads run kPropDepth steps ahead, clamped to the last
    // union row past the end (never used); the unrolled body has no exit, so
    // the loads stay in flight across steps.
    const int lrow0 = uy0 / kLat, lcol0 = ux0 / kLat;
    const int dm = ncb > 0 ? 4 / ncb : 0;
    const int mr0 = ncb > 0 ? wave / ncb : 0;
    const int xi = (ncb > 0 ? wave % ncb : 0) * 16 + r;
    const int nsteps = ntile > wave ? (ntile - wave + 3) / 4 : 0;
    const size_t col_off = (size_t)lcol0 + (xi < nxu ? xi : nxu - 1);
    auto load_step = [&](int k, h8& b, float& nq) {
        const int mr = mr0 + dm * k;
        const size_t pq = (size_t)(lrow0 + (mr < nyu ? mr : nyu - 1)) * ew2 + col_off;
        __builtin_memcpy(&b, a.emb + pq * kEmb + 8 * q, sizeof(h8));
        nq = a.norm[pq];
    };
    int rx[4];
    bool colok[4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        rx[ii] = xi - col_lo[ii];
        colok[ii] = (unsigned)rx[ii] < (unsigned)nlx[ii];
    }
    h8 pb[kPropDepth];
    float pnq[kPropDepth];
    if (nsteps > 0) {
#pragma unroll
        for (int s = 0; s < kPropDepth; ++s) load_step(s, pb[s], pnq[s]);
    }
    for (int k0 = 0; k0 < nsteps; k0 += kPropDepth) {
#pragma unroll
        for (int s = 0; s < kPropDepth; ++s) {
            const int k = k0 + s;
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma_h32(af, pb[s], acc);
            const float nq = pnq[s];
            load_step(k + kPropDepth, pb[s], pnq[s]);
            if (k < nsteps) {
                const int mr = mr0 + dm * k;
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) {
                    const int ry = mr - row_lo[ii];
                    if (!colok[ii] || (unsigned)ry >= (unsigned)nly[ii]) continue;
                    float d = __builtin_fmaf(-2.0f, acc[ii], nref[ii] + nq);
                    d = d > 0.0f ? d : 0.0f;
                    sub[4 * q + ii][__umul24(ry, nlx[ii]) + rx[ii]] = __builtin_bit_cast(unsigned short, (_Float16)d);
                }
            }
        }
    }
    __syncthreads();
    // per block: the M best lattice positions, as a list in lattice order
    for (int k = wave; k < 16; k += 4) {
        int by, bx;
        if (!tile_block(k, by, bx)) continue;
        const LatWin w = lat_window(p, by, bx);
        const int nsub = w.ny * w.nx;
        constexpr int NS = (kSubCap + 63) / 64;
        uint32_t v[NS];
        uint32_t lmin = 0xFFFFu;
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = lane + 64 * u;
            v[u] = e < nsub ? (uint32_t)sub[k][e] : 0xFFFFu;
            lmin = v[u] < lmin ? v[u] : lmin;
        }
        uint16_t* out = a.props + ((size_t)by * p.nx + bx) * kPropSlots;
        // bound: the M-th smallest lane minimum tb (M lanes hold a value
        // <= tb, so the M best by (score, position) all score <= tb);
        // those values are compacted in position order, usually a few
        // dozen, and the exact selection runs over them alone
        uint32_t tb = 0;
        for (int bit = 15; bit >= 0; --bit) {
            const uint32_t trial = tb | (1u << bit);
            if (__popcll(__ballot(lmin < trial)) < M) tb = trial;
        }
        uint32_t* ck = comp[wave];
        int* ce = reinterpret_cast<int*>(ck + 64);
        int ncomp = 0;
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const bool in = v[u] <= tb && v[u] != 0xFFFFu;
            const uint64_t mk = __ballot(in);
            const int at = ncomp + mbcnt64(mk);
            if (in && at < 64) {
                ck[at] = v[u];
                ce[at] = lane + 64 * u;
            }
            ncomp += __popcll(mk);
        }
        wave_sync_lds();
        int nsel;
        if (ncomp <= 64) {
            const uint32_t cv = lane < ncomp ? ck[lane] : 0xFFFFu;
            uint32_t tv = 0;
            for (int bit = 15; bit >= 0; --bit) {
                const uint32_t trial = tv | (1u << bit);
                if (__popcll(__ballot(cv < trial)) < M) tv = trial;
            }
            const int less = __popcll(__ballot(cv < tv && cv != 0xFFFFu));
            const bool eq = cv == tv && cv != 0xFFFFu;
            const int rank = mbcnt64(__ballot(eq));
            const bool sel = (cv < tv && cv != 0xFFFFu) || (eq && rank < M - less);
            const uint64_t sm = __ballot(sel);
            if (sel) out[mbcnt64(sm)] = (uint16_t)ce[lane];
            nsel = __popcll(sm);
        } else {
            const uint32_t sel = wave_smallest16<NS>(v, M);
            nsel = 0;
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                const bool s_ = (sel >> u) & 1u;
                const uint64_t sm = __ballot(s_);
                if (s_) out[nsel + mbcnt64(sm)] = (uint16_t)(lane + 64 * u);
                nsel += __popcll(sm);
            }
        }
        if (lane >= nsel && lane < kPropSlots) out[lane] = 0xFFFFu;
        wave_sync_lds();  // comp is rewritten for this wave's next block
    }
}

// k_rerank: one workgroup (four waves) per run of kRerankBlocks consecutive
// reference blocks of one block row, two blocks per wave (lanes 0-31 and
// 32-63), one lane per proposal.  The union of the run's search windows
// (with the pixels their candidate blocks cover, one padding row above and
// below and two padding columns on the left; coordinates clamped to the
// plane) is staged in LDS with coalesced loads: at most 58 x 117 floats at
// bm_range 24.  Each lane reads the 10x10 pixels under its 3x3 neighbourhood
// of candidates (rows cy-1 .. cy+8, columns cx-1 .. cx+8, as 8-byte aligned
// pairs: cx is even) and evaluates the nine SSDs from registers against the
// reference block, in the SSE2 lane order of ssd_fixed<8> (scalar: the
// row-major order of builds without __SSE2__, Block.h:632-643): per candidate
// rows in increasing order, each row's columns 0-3 and then 4-7 into the four
// lane sums.  Selection per block: a radix descent over the leading bits of
// all keys bounds the kk-th smallest (kk = min(accepted, G-1)); the keys at or
// below the bound are compacted (into the window's LDS, free once every wave
// has its SSDs) and ranked by (key, window index).
constexpr int kRerankWaves = 4;
constexpr int kRerankBlocks = 2 * kRerankWaves;
constexpr int kWinRows = 49 + 7 + 2;  // search rows + block rows - 1 + padding above and below
constexpr int kWinStride = 120;       // >= 49 + 7 * 8 + 13 columns, a multiple of four
constexpr int kLatWords = (kSubCap + 31) / 32;  // lattice selection bitmap of one block
constexpr int kCompCap = 9 * kPropSlots;        // every candidate of a block
static_assert(kRerankBlocks * 2 * kCompCap <= kWinRows * kWinStride, "selection lists reuse the window");
struct RerankLds {
    alignas(16) float win[kWinRows * kWinStride];
    float ref[kRerankBlocks][64];
    uint32_t sel[kRerankBlocks][kLatWords];
};

template <bool SCALAR>
__device__ __forceinline__ void rerank_run(const ProposeArgs& a, RerankLds& s) {
    const StageParams& p = a.p;
    const int lane = lane_id();
    const int h = lane >> 5, sl = lane & 31;
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int B = 8;
    const int G = p.group;
    // the run of the workgroup (XCD-contiguous, row-major)
    const int runs_x = (p.nx + kRerankBlocks - 1) / kRerankBlocks;
    const int nruns = runs_x * p.ny;
    const int per_xcd = (nruns + 7) >> 3;
    const int lin = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (lin >= nruns) return;
    const int by = lin / runs_x;
    const int bx0 = (lin - by * runs_x) * kRerankBlocks;
    const int bxe = bx0 + kRerankBlocks - 1 < p.nx - 1 ? bx0 + kRerankBlocks - 1 : p.nx - 1;
    const int bx = bx0 + 2 * wave + h;
    const bool bok = bx <= bxe;
    const int bxc = bok ? bx : bxe;
    const int b = by * p.nx + bxc;
    const int j = grid_pos(by, p.block_step, p.height - B);
    const int i = grid_pos(bxc, p.block_step, p.width - B);
    int* cnt_out = a.count + b;
    int* pos_out = a.pos + (size_t)b * G;
    float* key_out = a.key + (size_t)b * G;
    if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) {
        if (sl == 0 && bok) {
            *cnt_out = 1;
            pos_out[0] = (j << 16) | i;
            key_out[0] = 0.0f;
        }
        return;
    }
    const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
    const int l = search_lo(i, p.range, 1), rr = search_hi(i, p.width - B, p.range, 1);
    // the staged window: plane rows t-1 .. bt+8, columns L0 .. L0 + wc - 1
    const int l_first = search_lo(grid_pos(bx0, p.block_step, p.width - B), p.range, 1);
    const int rr_last = search_hi(grid_pos(bxe, p.block_step, p.width - B), p.width - B, p.range, 1);
    const int L0 = (l_first - 2) & ~1;
    const int wc = rr_last + 10 - L0;
    const int wr = bt - t + 10;
    const bool inner = t >= 1 && bt + 8 <= p.height - 1 && L0 >= 0 && L0 + ((wc + 1) & ~1) <= p.width &&
                       (p.pitch & 1) == 0 && ((uintptr_t)a.plane & 7) == 0;
    if (inner) {
        // column pairs: thread -> pair tid % 60 of rows tid / 60, +4, +8, ...
        constexpr int kPairs = kWinStride / 2;
        constexpr int kRowsPer = 64 * kRerankWaves / kPairs;
        const int c2 = (int)threadIdx.x % kPairs, r0 = (int)threadIdx.x / kPairs;
        if (r0 < kRowsPer && 2 * c2 < wc) {
            const float* src = a.plane + (size_t)(t - 1 + r0) * p.pitch + L0 + 2 * c2;
            float* dst = s.win + r0 * kWinStride + 2 * c2;
#pragma unroll 4
            for (int r = r0; r < wr; r += kRowsPer) {
                float2 v;
                __builtin_memcpy(&v, src, sizeof(v));
                __builtin_memcpy(dst, &v, sizeof(v));
                src += (size_t)kRowsPer * p.pitch;
                dst += kRowsPer * kWinStride;
            }
        }
    } else {
        // windows at the plane's edges: every sample with clamped coordinates
        for (int e = (int)threadIdx.x; e < wr * kWinStride; e += 64 * kRerankWaves) {
            const int r = e / kWinStride, c = e - r * kWinStride;
            if (c >= wc) continue;
            int y = t - 1 + r, x = L0 + c;
            y = y < 0 ? 0 : (y > p.height - 1 ? p.height - 1 : y);
            x = x < 0 ? 0 : (x > p.width - 1 ? p.width - 1 : x);
            s.win[e] = a.plane[(size_t)y * p.pitch + x];
        }
    }
    const int slot = 2 * wave + h;
    float* rf = s.ref[slot];
    uint32_t* sb = s.sel[slot];
    rf[sl] = a.plane[(size_t)(j + (sl >> 3)) * p.pitch + i + (sl & 7)];
    rf[sl + 32] = a.plane[(size_t)(j + 4 + (sl >> 3)) * p.pitch + i + (sl & 7)];
    for (int w = sl; w < kLatWords; w += 32) sb[w] = 0u;
    const int ncols = rr - l + 1;
    const int ly0 = lat_first(t), lx0 = lat_first(l);
    const int nly = lat_span(ly0, bt), nlx = lat_span(lx0, rr);
    const int M = a.topk < kPropSlots ? a.topk : kPropSlots;
    const bool full = nly * nlx <= M;  // every lattice position proposed
    const int e = bok ? (int)a.props[(size_t)b * kPropSlots + sl] : 0xFFFF;
    const bool act = e < nly * nlx;  // 0xFFFF: no proposal in this slot
    const int my = act ? e / nlx : 0;
    const int mx = act ? e - my * nlx : 0;
    __syncthreads();
    if (act) __hip_atomic_fetch_or(&sb[e >> 5], 1u << (e & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    wave_sync_lds();
    auto proposed = [&](int yy, int xx) -> bool {
        if (yy < 0 || xx < 0 || xx >= nlx) return false;
        const int idx = yy * nlx + xx;
        return (sb[idx >> 5] >> (idx & 31)) & 1u;
    };
    // candidates this lane owns: inside the window, not self, and not in the
    // neighbourhood of an earlier proposal (rows above, or the left neighbour)
    const bool up = proposed(my - 1, mx), upl = proposed(my - 1, mx - 1), upr = proposed(my - 1, mx + 1);
    const bool lf = proposed(my, mx - 1);
    const int cy = ly0 + kLat * my, cx = lx0 + kLat * mx;
    uint32_t own = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const int y = cy + dy, x = cx + dx;
            bool ok = act && y >= t && y <= bt && x >= l && x <= rr && !(y == j && x == i);
            if (dy == -1 && (up || (dx == -1 && upl) || (dx == 1 && upr))) ok = false;
            if (dx == -1 && lf) ok = false;
            own |= (ok ? 1u : 0u) << ((dy + 1) * 3 + dx + 1);
        }
    // the nine SSDs; patch row pr = plane row cy - 1 + pr = window row cy - t + pr
    // (an 8-byte aligned pair index: cx - 2 - L0 and kWinStride are even)
    const float2* wbase = reinterpret_cast<const float2*>(s.win) + ((cy - t) * kWinStride + (cx - 2 - L0)) / 2;
    float acc[9][SCALAR ? 1 : 4];
#pragma unroll
    for (int u = 0; u < 9; ++u)
#pragma unroll
        for (int c = 0; c < (SCALAR ? 1 : 4); ++c) acc[u][c] = 0.0f;
    auto load_row = [&](int pr, float (&v)[10]) {
        float2 w[6];
        const float2* src = wbase + pr * (kWinStride / 2);
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = src[k];
        v[0] = w[0].y;
        v[1] = w[1].x;
        v[2] = w[1].y;
        v[3] = w[2].x;
        v[4] = w[2].y;
        v[5] = w[3].x;
        v[6] = w[3].y;
        v[7] = w[4].x;
        v[8] = w[4].y;
        v[9] = w[5].x;
    };
    // patch row pr is row pr - dy of the candidates in neighbourhood row dy
    auto add_row = [&](int pr, const float (&v)[10]) {
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const int rrow = pr - dy;
            if (rrow < 0 || rrow > 7) continue;
            float rb[8];
            __builtin_memcpy(&rb[0], rf + rrow * 8, 8 * sizeof(float));
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                float* sacc = acc[dy * 3 + dx];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const float d = rb[c] - v[dx + c];
                    if constexpr (SCALAR) sacc[0] = sacc[0] + d * d;
                    else sacc[c & 3] = sacc[c & 3] + d * d;
                }
            }
        }
    };
    float va[10], vb[10];
    load_row(0, va);
#pragma unroll 1
    for (int pr = 0; pr < 10; pr += 2) {
        load_row(pr + 1, vb);
        add_row(pr, va);
        if (pr + 2 < 10) load_row(pr + 2, va);
        add_row(pr + 1, vb);
    }
    // keys (0xFFFFFFFF: not a candidate or not accepted); the half-wave counts
    // are taken from the two halves of each ballot
    uint32_t kv[9];
    int nacc_lo = 0, nacc_hi = 0;
#pragma unroll
    for (int u = 0; u < 9; ++u) {
        float dist;
        if constexpr (SCALAR) dist = acc[u][0];
        else dist = 0.0f + (((acc[u][0] + acc[u][1]) + acc[u][2]) + acc[u][3]);
        const bool ok = ((own >> u) & 1u) && dist <= p.th_sse && dist != 0.0f;
        kv[u] = ok ? __float_as_uint((float)((double)dist * p.dist_mul)) : 0xFFFFFFFFu;
        const uint64_t bm = __ballot(ok);
        nacc_lo += __popc((uint32_t)bm);
        nacc_hi += __popc((uint32_t)(bm >> 32));
    }
    const int nacc = h ? nacc_hi : nacc_lo;
    const bool fb = nacc < G - 1 && !full;  // the exhaustive search decides
    if (fb && sl == 0 && bok) {
        const int at = atomicAdd(a.nfallback, 1);
        a.fallback[at] = b;
    }
    const int kk = nacc < G - 1 ? nacc : G - 1;
    // tb: an upper bound of the kk-th smallest key from its 16 leading bits
    // (keys are positive floats: bit 31 is clear), so that the compacted set
    // holds the kk best plus the few keys sharing their leading bits
    uint32_t tlo = 0, thi = 0;
    const int kk_lo = nacc_lo < G - 1 ? nacc_lo : G - 1, kk_hi = nacc_hi < G - 1 ? nacc_hi : G - 1;
    for (int bit = 30; bit >= 15; --bit) {
        const uint32_t trl = tlo | (1u << bit), trh = thi | (1u << bit);
        const uint32_t trial = h ? trh : trl;
        int clo = 0, chi = 0;
#pragma unroll
        for (int u = 0; u < 9; ++u) {
            const uint64_t bm = __ballot(kv[u] < trial);
            clo += __popc((uint32_t)bm);
            chi += __popc((uint32_t)(bm >> 32));
        }
        if (clo < kk_lo) tlo = trl;
        if (chi < kk_hi) thi = trh;
    }
    const uint32_t tb = (h ? thi : tlo) | 0x7FFFu;
    __syncthreads();  // every wave is done with the window
    uint32_t* ck = reinterpret_cast<uint32_t*>(s.win) + slot * 2 * kCompCap;
    int* ci = reinterpret_cast<int*>(ck + kCompCap);
    int nlo = 0, nhi = 0;
#pragma unroll
    for (int u = 0; u < 9; ++u) {
        const bool in = kv[u] <= tb && kv[u] != 0xFFFFFFFFu;
        const uint64_t m = __ballot(in);
        if (in) {
            const int dy = u / 3 - 1, dx = u % 3 - 1;
            const int at = (h ? nhi : nlo) + mbcnt64(m) - (h ? __popc((uint32_t)m) : 0);
            ck[at] = kv[u];
            ci[at] = (cy + dy - t) * ncols + (cx + dx - l);
        }
        nlo += __popc((uint32_t)m);
        nhi += __popc((uint32_t)(m >> 32));
    }
    const int ncomp = h ? nhi : nlo;
    wave_sync_lds();
    if (!fb && bok) {
        for (int e1 = sl; e1 < ncomp; e1 += 32) {
            const uint32_t k1 = ck[e1];
            const int w1 = ci[e1];
            // four entries' LDS reads in flight per step (one round trip per
            // entry otherwise)
            int rank = 0;
            int e2 = 0;
            for (; e2 + 4 <= ncomp; e2 += 4) {
                uint32_t k2[4];
                int w2[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    k2[q] = ck[e2 + q];
                    w2[q] = ci[e2 + q];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) rank += (k2[q] < k1) || (k2[q] == k1 && w2[q] < w1);
            }
            for (; e2 < ncomp; ++e2) {
                const uint32_t k2 = ck[e2];
                rank += (k2 < k1) || (k2 == k1 && ci[e2] < w1);
            }
            if (rank < kk) {
                const int wy = w1 / ncols;
                pos_out[1 + rank] = ((t + wy) << 16) | (l + w1 - wy * ncols);
                key_out[1 + rank] = __uint_as_float(k1);
            }
        }
        if (sl == 0) {
            *cnt_out = kk + 1;
            pos_out[0] = (j << 16) | i;
            key_out[0] = 0.0f;
        }
    }
}

// The arithmetic flavour is chosen once per launch (the LDS lives here so
// that both bodies share it).
__global__ __launch_bounds__(64 * kRerankWaves) void k_rerank(ProposeArgs a) {
    __shared__ RerankLds smem;
    if (a.p.arith) rerank_run<true>(a, smem);
    else rerank_run<false>(a, smem);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_train_fwd(const Net& n, const TrainBatch& tb, hipStream_t st) {
    const int waves = (2 * tb.npairs + 15) / 16;
    hipLaunchKernelGGL(k_train_fwd, dim3((waves + 3) / 4), dim3(256), 0, st, n, tb);
    return hipGetLastError();
}

hipError_t launch_train_loss(const TrainBatch& tb, float t0, float* block_loss, hipStream_t st) {
    hipLaunchKernelGGL(k_train_loss, dim3((tb.npairs + 255) / 256), dim3(256), 0, st, tb, t0, block_loss);
    return hipGetLastError();
}

hipError_t launch_train_bwd(const Net& n, const TrainBatch& tb, float* part, float* grad, hipStream_t st) {
    const int waves = (2 * tb.npairs + 15) / 16;
    hipLaunchKernelGGL(k_train_bwd_h, dim3((waves + 3) / 4), dim3(256), 0, st, n, tb);
    const int slices = (2 * tb.npairs + kSlice - 1) / kSlice;
    hipLaunchKernelGGL(k_train_bwd_w, dim3(slices), dim3(256), 0, st, tb, part);
    hipLaunchKernelGGL(k_train_reduce, dim3((kParams + 255) / 256), dim3(256), 0, st, part, slices, grad);
    return hipGetLastError();
}

int train_slices(int npairs) { return (2 * npairs + kSlice - 1) / kSlice; }

hipError_t launch_train_adam(const Net& n, const float* grad, float* m, float* v, float lr, float b1, float b2,
                             float eps, float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(k_train_adam, dim3((kParams + 255) / 256), dim3(256), 0, st, n, grad, m, v, lr, b1, b2, eps, c1,
                       c2);
    return hipGetLastError();
}

hipError_t launch_sync_bf16(const Net& n, hipStream_t st) {
    hipLaunchKernelGGL(k_sync_bf16, dim3((kParams + 255) / 256), dim3(256), 0, st, n);
    return hipGetLastError();
}

hipError_t launch_pair_ssd(const float* plane, int pitch, const int4* pairs, int npairs, float* target,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_pair_ssd, dim3((npairs + 255) / 256), dim3(256), 0, st, plane, pitch, pairs, npairs, target);
    return hipGetLastError();
}

hipError_t launch_embed(const Net& n, const float* plane, const StageParams& p, _Float16* emb, float* norm,
                        _Float16* remb, float* rnorm, int* nfallback, hipStream_t st) {
    const int ew2 = lat_count(p.width - 7), eh2 = lat_count(p.height - 7);
    const int lat_waves = ew2 > 0 && eh2 > 0 ? ((ew2 + 15) / 16) * ((eh2 + kEmbedRun - 1) / kEmbedRun) : 0;
    const int nb = p.nx * p.ny;
    const int ref_waves = nb > 0 ? (nb + 16 * kRefGroups - 1) / (16 * kRefGroups) : 1;  // >= 1: clears *nfallback
    const int lat_wgs = (lat_waves + 3) / 4, ref_wgs = (ref_waves + 3) / 4;
    hipLaunchKernelGGL(k_embed, dim3(lat_wgs + ref_wgs), dim3(256), 0, st, n, plane, p, lat_wgs, emb, norm, remb, rnorm,
                       nfallback);
    return hipGetLastError();
}

hipError_t launch_propose(const ProposeArgs& a, hipStream_t st) {
    const int tiles = ((a.p.ny + kPropRows - 1) / kPropRows) * ((a.p.nx + kPropCols - 1) / kPropCols);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_propose, dim3(8 * ((tiles + 7) / 8)), dim3(256), 0, st, a);
    const int runs = ((a.p.nx + kRerankBlocks - 1) / kRerankBlocks) * a.p.ny;
    hipLaunchKernelGGL(k_rerank, dim3(8 * ((runs + 7) / 8)), dim3(64 * kRerankWaves), 0, st, a);
    return hipGetLastError();
}

size_t propose_list_bytes(const StageParams& p) { return (size_t)p.nx * p.ny * kPropSlots * sizeof(uint16_t); }

}  // namespace lbm
}  // namespace bm3d
