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

// The network on 16 patches (inference form of forward16: same roundings).
__device__ __forceinline__ void infer16(const Frags& f, const s4 (&x)[4], float (&e)[2][4]) {
    s4 hb[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {f.b1[t][0], f.b1[t][1], f.b1[t][2], f.b1[t][3]};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(f.w1[t][s], x[s], acc);
        const sh2 lo = relu_bf16x2(acc[0], acc[1]), hi = relu_bf16x2(acc[2], acc[3]);
        hb[t] = s4{lo[0], lo[1], hi[0], hi[1]};
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        f4 acc = {f.b2[u][0], f.b2[u][1], f.b2[u][2], f.b2[u][3]};
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = mfma(f.w2[u][s], hb[s], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) e[u][i] = acc[i];
    }
}

// The lane's part of the patch at (py, px): rows 2 s + q / 2, columns
// 4 (q % 2) .. +3 for s = 0..3 (q = lane / 16), as loaded.
__device__ __forceinline__ void patch_load(const float* __restrict__ plane, int pitch, int py, int px,
                                           float4 (&v)[4]) {
    const int q = lane_id() >> 4;
#pragma unroll
    for (int s = 0; s < 4; ++s)
        __builtin_memcpy(&v[s], plane + (size_t)(py + 2 * s + (q >> 1)) * pitch + px + 4 * (q & 1), sizeof(float4));
}

// The MFMA fragments of a loaded patch part (packed bf16 conversion).
__device__ __forceinline__ void patch_cvt(const float4 (&v)[4], s4 (&x)[4]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const fl2 a = {v[s].x, v[s].y}, b = {v[s].z, v[s].w};
        const sh2 lo = __builtin_bit_cast(sh2, __builtin_convertvector(a, bf2));
        const sh2 hi = __builtin_bit_cast(sh2, __builtin_convertvector(b, bf2));
        x[s] = s4{lo[0], lo[1], hi[0], hi[1]};
    }
}

// Stores the embedding (fp16) and its squared norm (fp32 of the fp16 values)
// of the lane's patch at index idx (idx < 0: no patch).
__device__ __forceinline__ void store_emb(const float (&e)[2][4], int idx, _Float16* __restrict__ emb,
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
    // lanes r, r+16, r+32, r+48 hold the four quarters of one embedding
    nrm += __shfl_xor(nrm, 16);
    nrm += __shfl_xor(nrm, 32);
    if (idx >= 0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) __builtin_memcpy(emb + (size_t)idx * kEmb + 16 * u + 4 * q, &eh[u], sizeof(h4));
        if (q == 0) norm[idx] = nrm;
    }
}

// Each wave embeds kEmbedGroups groups of 16 patches; pos(g, r, y, x)
// returns the output index of lane r's patch in group g (and its top-left
// corner), or -1.  The patches of the next two groups are in flight while a
// group runs (loaded as floats, converted when their group starts).
constexpr int kEmbedGroups = 16;
template <typename Pos>
__device__ __forceinline__ void embed_groups(const Net& n, const float* __restrict__ plane, int pitch, Pos pos,
                                             _Float16* __restrict__ emb, float* __restrict__ norm) {
    const int r = lane_id() & 15;
    Frags f;
    load_frags(n, f);
    int y, x;
    int idq[2];
    float4 nv[2][4];
    idq[0] = pos(0, r, y, x);
    patch_load(plane, pitch, y, x, nv[0]);
    idq[1] = pos(1, r, y, x);
    patch_load(plane, pitch, y, x, nv[1]);
#pragma unroll
    for (int g = 0; g < kEmbedGroups; ++g) {
        if (__ballot(idq[g & 1] >= 0) == 0) break;
        s4 xf[4];
        patch_cvt(nv[g & 1], xf);
        const int cur = idq[g & 1];
        if (g + 2 < kEmbedGroups) {
            idq[g & 1] = pos(g + 2, r, y, x);
            patch_load(plane, pitch, y, x, nv[g & 1]);
        }
        float e[2][4];
        infer16(f, xf, e);
        store_emb(e, cur, emb, norm);
    }
}

// Embeddings of the lattice: patch (kLat yy, kLat xx) -> index yy * ew2 + xx,
// ew2 = ceil((W - 7) / kLat), for yy < ceil((H - 7) / kLat).
__global__ __launch_bounds__(256) void k_embed_lattice(Net n, const float* __restrict__ plane, int pitch, int W, int H,
                                                       _Float16* __restrict__ emb, float* __restrict__ norm) {
    const int ew2 = lat_count(W - 7), eh2 = lat_count(H - 7);
    const int per_row = (ew2 + 16 * kEmbedGroups - 1) / (16 * kEmbedGroups);
    const int wave = uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (wave >= per_row * eh2) return;
    const int yy = wave / per_row;
    const int xs = (wave - yy * per_row) * 16 * kEmbedGroups;
    embed_groups(
        n, plane, pitch,
        [&](int g, int r, int& y, int& x) {
            const int xx = xs + 16 * g + r;
            y = kLat * yy;
            x = kLat * (xx < ew2 ? xx : ew2 - 1);
            return xx < ew2 ? yy * ew2 + xx : -1;
        },
        emb, norm);
}

// Embeddings of the reference blocks (scan order b = by * nx + bx).
__global__ __launch_bounds__(256) void k_embed_refs(Net n, const float* __restrict__ plane, StageParams p,
                                                    _Float16* __restrict__ emb, float* __restrict__ norm) {
    const int nb = p.nx * p.ny;
    const int wave = uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int b0 = wave * 16 * kEmbedGroups;
    if (b0 >= nb) return;
    embed_groups(
        n, plane, p.pitch,
        [&](int g, int r, int& y, int& x) {
            const int b = b0 + 16 * g + r;
            const int bc = b < nb ? b : nb - 1;
            const int by = bc / p.nx;
            y = grid_pos(by, p.block_step, p.height - 8);
            x = grid_pos(bc - by * p.nx, p.block_step, p.width - 8);
            return b < nb ? b : -1;
        },
        emb, norm);
}

// ---------------------------------------------------------------------------
// inference: proposals, exact re-rank and fallback list
// ---------------------------------------------------------------------------
//
// One workgroup (four waves) per tile of kPropRows x kPropCols reference
// blocks.  Embedding distance: score(r, c) = |e_r|^2 + |e_c|^2 - 2 e_r.e_c
// (fp32; fp16 in the per-block lists).  The dot products of the 16 blocks of a
// tile row with 16 lattice positions of a union row are one 16x16x32 product
// (two v_mfma_f32_16x16x16_f16).  Tiles are dealt to the XCDs in column-major
// runs, so vertically adjacent tiles, whose unions overlap most, share an L2.
//
// Scoring (per tile row): every even-lattice position (even row, even column)
// of each block's search window is scored.  Each block proposes its M = K/2
// best lattice positions (ties to the earlier position); the candidates are
// the 3x3 neighbourhoods of the proposals inside the window, self excluded,
// marked in a per-block bitmap, so they come out in window scan order.
//
// Per block (one wave): every candidate is re-ranked exactly like the plain
// matcher (SSE2 lane-order SSD, dist <= thSSE && dist != 0, key =
// float(double(dist) * distMul)) and the G-1 best by (key, window index)
// follow the self entry.  A block with fewer than G-1 accepted candidates
// falls back to the exhaustive search; a window of at most K candidates is
// searched exhaustively here.
constexpr int kPropCols = 16;
constexpr int kPropRows = 4;
constexpr int kSubSide = (49 + kLat - 1) / kLat;  // lattice points per side of the largest window
constexpr int kSubCap = kSubSide * kSubSide;  // lattice scores per block of a tile row
constexpr int kListCap = 9 * (kMaxProposals / 2);
constexpr int kBitWords = (kMaxWindow + 31) / 32;

__device__ __forceinline__ f4 mfma_h(h4 a, h4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// SSD of the 8x8 block rb (row-major registers) against the block at q, in
// the SSE2 lane order of ssd_fixed<8> (scalar: the row-major order of builds
// without __SSE2__, Block.h:632-643).
template <bool SCALAR>
__device__ __forceinline__ float ssd8_rows(const float* rb, const float* __restrict__ q, int pitch) {
    if constexpr (SCALAR) {
        float dist = 0.0f;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float t = rb[r * 8 + c] - q[r * pitch + c];
                dist = dist + t * t;
            }
        return dist;
    }
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        float4 lo, hi;
        __builtin_memcpy(&lo, q + r * pitch, sizeof(lo));
        __builtin_memcpy(&hi, q + r * pitch + 4, sizeof(hi));
        const float d0 = rb[r * 8 + 0] - lo.x, d1 = rb[r * 8 + 1] - lo.y;
        const float d2 = rb[r * 8 + 2] - lo.z, d3 = rb[r * 8 + 3] - lo.w;
        s0 = s0 + d0 * d0;
        s1 = s1 + d1 * d1;
        s2 = s2 + d2 * d2;
        s3 = s3 + d3 * d3;
        const float e0 = rb[r * 8 + 4] - hi.x, e1 = rb[r * 8 + 5] - hi.y;
        const float e2 = rb[r * 8 + 6] - hi.z, e3 = rb[r * 8 + 7] - hi.w;
        s0 = s0 + e0 * e0;
        s1 = s1 + e1 * e1;
        s2 = s2 + e2 * e2;
        s3 = s3 + e3 * e3;
    }
    return 0.0f + (((s0 + s1) + s2) + s3);
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


__global__ __launch_bounds__(256) void k_propose(ProposeArgs a) {
    __shared__ unsigned short sub[kPropCols][kSubCap];  // lattice scores of one tile row
    __shared__ uint32_t bits[4][kBitWords];             // candidates of the block each wave marks
    __shared__ uint32_t comp[4][128];                   // per wave: compacted lattice scores and positions
    const StageParams& p = a.p;
    const int lane = lane_id();
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int r = lane & 15, q = lane >> 4;
    // tile of the workgroup (XCD-contiguous, column-major)
    const int row_tiles = (p.ny + kPropRows - 1) / kPropRows;
    const int col_tiles = (p.nx + kPropCols - 1) / kPropCols;
    const int ntiles = row_tiles * col_tiles;
    const int per_xcd = (ntiles + 7) >> 3;
    const int lin = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (lin >= ntiles) return;
    const int tcol = lin / row_tiles;
    const int bj0 = (lin - tcol * row_tiles) * kPropRows;
    const int bi0 = tcol * kPropCols;
    const int ew2 = lat_count(p.width - 7);
    const int B = 8;
    const int G = p.group;
    const int K = a.topk;
    const int M = K < kMaxProposals / 2 ? K : kMaxProposals / 2;
    const int last_bi = bi0 + kPropCols - 1 < p.nx - 1 ? bi0 + kPropCols - 1 : p.nx - 1;
    const bool shortcut = G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0);
    int jr[kPropRows], tr[kPropRows], br[kPropRows], bjr[kPropRows];
    bool rok[kPropRows];
#pragma unroll
    for (int m = 0; m < kPropRows; ++m) {
        rok[m] = bj0 + m < p.ny;
        bjr[m] = rok[m] ? bj0 + m : p.ny - 1;
        jr[m] = grid_pos(bjr[m], p.block_step, p.height - B);
        tr[m] = search_lo(jr[m], p.range, 1);
        br[m] = search_hi(jr[m], p.height - B, p.range, 1);
    }

    if (!shortcut) {
        // A fragments: lane r <-> block bi0 + r; D rows 4q + ii <-> block bi0 + 4q + ii
        const int bia = bi0 + r <= last_bi ? bi0 + r : last_bi;
        int lx0[4], nlx[4];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
            const int bi = bi0 + 4 * q + ii;
            const int ic = grid_pos(bi <= last_bi ? bi : last_bi, p.block_step, p.width - B);
            const int l = search_lo(ic, p.range, 1), rr = search_hi(ic, p.width - B, p.range, 1);
            lx0[ii] = lat_first(l);
            nlx[ii] = bi <= last_bi ? lat_span(lx0[ii], rr) : 0;
        }
        const int u_l = search_lo(grid_pos(bi0, p.block_step, p.width - B), p.range, 1);
        const int u_r = search_hi(grid_pos(last_bi, p.block_step, p.width - B), p.width - B, p.range, 1);
        const int ux0 = lat_first(u_l);
        const int nux = lat_span(ux0, u_r);
        const int ncb = (nux + 15) >> 4;
#pragma unroll
        for (int m = 0; m < kPropRows; ++m) {
            if (!rok[m]) break;
            h4 af0, af1;
            {
                const _Float16* er = a.remb + ((size_t)bjr[m] * p.nx + bia) * kEmb;
                __builtin_memcpy(&af0, er + 4 * q, sizeof(h4));
                __builtin_memcpy(&af1, er + 16 + 4 * q, sizeof(h4));
            }
            float nref[4];
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) {
                const int bi = bi0 + 4 * q + ii;
                nref[ii] = a.rnorm[(size_t)bjr[m] * p.nx + (bi <= last_bi ? bi : last_bi)];
            }
            for (int e = threadIdx.x; e < kPropCols * kSubCap; e += 256) (&sub[0][0])[e] = 0xFFFFu;
            __syncthreads();
            const int ly0 = lat_first(tr[m]);
            const int nly = lat_span(ly0, br[m]);
            const int ntile = nly * ncb;
            // one step = 16 lattice positions of one lattice row; the next step's
            // embeddings are loaded while this one is scored
            auto load_step = [&](int tl, h4& b0, h4& b1, float& nq) {
                const int mr = tl / ncb;
                const int xi = (tl - mr * ncb) * 16 + r;
                const int x = ux0 + kLat * (xi < nux ? xi : nux - 1);
                const size_t pq = (size_t)(ly0 / kLat + mr) * ew2 + x / kLat;
                __builtin_memcpy(&b0, a.emb + pq * kEmb + 4 * q, sizeof(h4));
                __builtin_memcpy(&b1, a.emb + pq * kEmb + 16 + 4 * q, sizeof(h4));
                nq = a.norm[pq];
            };
            h4 n0, n1;
            float nnq = 0.0f;
            if (wave < ntile) load_step(wave, n0, n1, nnq);
            for (int tl = wave; tl < ntile; tl += 4) {
                const int mr = tl / ncb;
                const int xi = (tl - mr * ncb) * 16 + r;
                const int x = ux0 + kLat * (xi < nux ? xi : nux - 1);
                const h4 b0 = n0, b1 = n1;
                const float nq = nnq;
                if (tl + 4 < ntile) load_step(tl + 4, n0, n1, nnq);
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma_h(af0, b0, acc);
                acc = mfma_h(af1, b1, acc);
                if (xi >= nux) continue;
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) {
                    const int mc = (x - lx0[ii]) / kLat;
                    if (x < lx0[ii] || mc >= nlx[ii]) continue;
                    float d = __builtin_fmaf(-2.0f, acc[ii], nref[ii] + nq);
                    d = d > 0.0f ? d : 0.0f;
                    sub[4 * q + ii][mr * nlx[ii] + mc] = __builtin_bit_cast(unsigned short, (_Float16)d);
                }
            }
            __syncthreads();
            // per block: the M best lattice positions -> 3x3 neighbourhoods
            for (int k = wave; k < kPropCols; k += 4) {
                const int bi = bi0 + k;
                if (bi > last_bi) break;
                const int i = grid_pos(bi, p.block_step, p.width - B);
                const int l = search_lo(i, p.range, 1), rr = search_hi(i, p.width - B, p.range, 1);
                const int ncols = rr - l + 1;
                const int lxk = lat_first(l);
                const int nlxk = lat_span(lxk, rr);
                const int nsub = nly * nlxk;
                const int j = jr[m], t = tr[m], bt = br[m];
                const int self = (j - t) * ncols + (i - l);
                constexpr int NS = (kSubCap + 63) / 64;
                uint32_t v[NS];
                uint32_t lmin = 0xFFFFu;
#pragma unroll
                for (int u = 0; u < NS; ++u) {
                    const int e = lane + 64 * u;
                    v[u] = e < nsub ? (uint32_t)sub[k][e] : 0xFFFFu;
                    lmin = v[u] < lmin ? v[u] : lmin;
                }
                uint32_t* bk = bits[wave];
                for (int w = lane; w < kBitWords; w += 64) bk[w] = 0u;
                // marks the 3x3 neighbourhood of lattice position e (self excluded)
                auto mark = [&](int e) {
                    const int my = e / nlxk;
                    const int y0 = ly0 + kLat * my, x0 = lxk + kLat * (e - my * nlxk);
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int yy = y0 + dy, xx = x0 + dx;
                            if (yy < t || yy > bt || xx < l || xx > rr) continue;
                            const int c = (yy - t) * ncols + (xx - l);
                            if (c == self) continue;
                            __hip_atomic_fetch_or(&bk[c >> 5], 1u << (c & 31), __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WAVEFRONT);
                        }
                };
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
                    const int at = ncomp + __builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0));
                    if (in && at < 64) {
                        ck[at] = v[u];
                        ce[at] = lane + 64 * u;
                    }
                    ncomp += __popcll(mk);
                }
                wave_sync_lds();
                if (ncomp <= 64) {
                    const uint32_t cv = lane < ncomp ? ck[lane] : 0xFFFFu;
                    uint32_t tv = 0;
                    for (int bit = 15; bit >= 0; --bit) {
                        const uint32_t trial = tv | (1u << bit);
                        if (__popcll(__ballot(cv < trial)) < M) tv = trial;
                    }
                    const int less = __popcll(__ballot(cv < tv && cv != 0xFFFFu));
                    const bool eq = cv == tv && cv != 0xFFFFu;
                    const uint64_t em = __ballot(eq);
                    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(em >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)em, 0));
                    if ((cv < tv && cv != 0xFFFFu) || (eq && rank < M - less)) mark(ce[lane]);
                } else {
                    const uint32_t sel = wave_smallest16<NS>(v, M);
#pragma unroll
                    for (int u = 0; u < NS; ++u)
                        if (sel & (1u << u)) mark(lane + 64 * u);
                }
                wave_sync_lds();
                // the block's candidate bitmap goes to k_rerank
                uint32_t* out = a.cand_bits + (size_t)(bjr[m] * p.nx + bi) * kBitWords;
                for (int w = lane; w < kBitWords; w += 64) out[w] = bk[w];
                wave_sync_lds();  // bk is cleared for this wave's next block
            }
            __syncthreads();  // sub is rewritten by the next tile row
        }
    }
}

// Per reference block (one wave): candidates from the bitmap of k_propose
// (window scan order), exact re-rank, output.  A separate kernel so that the
// latency-bound re-rank runs at the occupancy of its own small LDS footprint.
constexpr int kRerankWaves = 4;
constexpr int kRerankWords = 2 * kListCap + 4 * kMaxLearnedGroupSize + 64;
template <bool SCALAR>
__device__ __forceinline__ void rerank_body(const ProposeArgs& a, uint32_t (*smem)[kRerankWords]) {
    const StageParams& p = a.p;
    const int lane = lane_id();
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int b = (int)blockIdx.x * kRerankWaves + wave;
    if (b >= p.nx * p.ny) return;
    const int B = 8;
    const int G = p.group;
    const int K = a.topk;
    uint32_t* cand = smem[wave];
    uint32_t* keys = cand + kListCap;
    uint32_t* sel_key = keys + kListCap;
    int* sel_idx = reinterpret_cast<int*>(sel_key + 2 * kMaxLearnedGroupSize);
    float* rbs = reinterpret_cast<float*>(sel_idx + 2 * kMaxLearnedGroupSize);
    const int by = b / p.nx, bi = b - by * p.nx;
    const int j = grid_pos(by, p.block_step, p.height - B);
    const int i = grid_pos(bi, p.block_step, p.width - B);
    int* cnt_out = a.count + b;
    int* pos_out = a.pos + (size_t)b * G;
    float* key_out = a.key + (size_t)b * G;
    if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) {
        if (lane == 0) {
            *cnt_out = 1;
            pos_out[0] = (j << 16) | i;
            key_out[0] = 0.0f;
        }
        return;
    }
    const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
    const int l = search_lo(i, p.range, 1), rr = search_hi(i, p.width - B, p.range, 1);
    const int nrows = bt - t + 1;
    const int ncols = rr - l + 1, nc = ncols * nrows;
    const int self = (j - t) * ncols + (i - l);
    rbs[lane] = a.plane[(size_t)(j + (lane >> 3)) * p.pitch + i + (lane & 7)];
    int n = 0;
    if (nc - 1 <= K) {
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int c = c0 + lane;
            const bool in = c < nc && c != self;
            const uint64_t mk = __ballot(in);
            const int at = n + __builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0));
            if (in) cand[at] = (uint32_t)c;
            n += __popcll(mk);
        }
    } else {
        // the set bits of the block's bitmap, in window order
        const int nw = (nc + 31) >> 5;
        const uint32_t* bw = a.cand_bits + (size_t)b * kBitWords;
        const uint32_t w0 = lane < nw ? bw[lane] : 0u;
        const uint32_t w1 = lane + 64 < nw ? bw[lane + 64] : 0u;
        const int c0n = __popc(w0), c1n = __popc(w1);
        int incl = c0n, incl1 = c1n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v0 = __shfl_up(incl, o), v1 = __shfl_up(incl1, o);
            if (lane >= o) {
                incl += v0;
                incl1 += v1;
            }
        }
        const int tot0 = __shfl(incl, 63);
        n = tot0 + __shfl(incl1, 63);
        int at = incl - c0n;
        for (uint32_t w = w0; w; w &= w - 1) cand[at++] = (uint32_t)(lane * 32 + __builtin_ctz(w));
        at = tot0 + incl1 - c1n;
        for (uint32_t w = w1; w; w &= w - 1) cand[at++] = (uint32_t)((lane + 64) * 32 + __builtin_ctz(w));
    }
    wave_sync_lds();
    // exact SSD of every candidate, one lane per candidate (reference block
    // read from LDS by broadcast)
    int nacc = 0;
    for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        uint32_t kv = 0xFFFFFFFFu;
        if (e < n) {
            const int c = (int)cand[e];
            const int cy = c / ncols, cx = c - cy * ncols;
            const float dist = ssd8_rows<SCALAR>(rbs, a.plane + (size_t)(t + cy) * p.pitch + l + cx, p.pitch);
            if (dist <= p.th_sse && dist != 0.0f) kv = __float_as_uint((float)((double)dist * p.dist_mul));
            keys[e] = kv;
        }
        nacc += __popcll(__ballot(kv != 0xFFFFFFFFu));
    }
    wave_sync_lds();
    if (nacc < G - 1 && nc - 1 > K) {
        // not enough acceptable candidates: the exhaustive search decides
        if (lane == 0) {
            const int at = atomicAdd(a.nfallback, 1);
            a.fallback[at] = b;
        }
        return;
    }
    const int kk = nacc < G - 1 ? nacc : G - 1;
    // the kk best by (key, list position); list position order = window order
    wave_select_small<(kListCap + 63) / 64>(keys, n, nacc, kk, sel_key, sel_idx);
    for (int z = lane; z < kk; z += 64) {
        const int c = (int)cand[sel_idx[kk + z]];
        const int cy = c / ncols;
        pos_out[1 + z] = ((t + cy) << 16) | (l + c - cy * ncols);
        key_out[1 + z] = __uint_as_float(sel_key[kk + z]);
    }
    if (lane == 0) {
        *cnt_out = kk + 1;
        pos_out[0] = (j << 16) | i;
        key_out[0] = 0.0f;
    }
}

// One wave per reference block; the arithmetic flavour is chosen once per
// launch (the LDS lives here so that both bodies share it).
__global__ __launch_bounds__(64 * kRerankWaves) void k_rerank(ProposeArgs a) {
    __shared__ uint32_t smem[kRerankWaves][kRerankWords];
    if (a.p.arith) rerank_body<true>(a, smem);
    else rerank_body<false>(a, smem);
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

hipError_t launch_embed(const Net& n, const float* plane, int pitch, int W, int H, _Float16* emb, float* norm,
                        hipStream_t st) {
    const int ew2 = lat_count(W - 7), eh2 = lat_count(H - 7);
    if (ew2 <= 0 || eh2 <= 0) return hipSuccess;
    const int per_row = (ew2 + 16 * kEmbedGroups - 1) / (16 * kEmbedGroups);
    const int waves = per_row * eh2;
    hipLaunchKernelGGL(k_embed_lattice, dim3((waves + 3) / 4), dim3(256), 0, st, n, plane, pitch, W, H, emb, norm);
    return hipGetLastError();
}

hipError_t launch_embed_refs(const Net& n, const float* plane, const StageParams& p, _Float16* emb, float* norm,
                             hipStream_t st) {
    const int nb = p.nx * p.ny;
    if (nb <= 0) return hipSuccess;
    const int waves = (nb + 16 * kEmbedGroups - 1) / (16 * kEmbedGroups);
    hipLaunchKernelGGL(k_embed_refs, dim3((waves + 3) / 4), dim3(256), 0, st, n, plane, p, emb, norm);
    return hipGetLastError();
}

hipError_t launch_propose(const ProposeArgs& a, hipStream_t st) {
    const int tiles = ((a.p.ny + kPropRows - 1) / kPropRows) * ((a.p.nx + kPropCols - 1) / kPropCols);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_propose, dim3(8 * ((tiles + 7) / 8)), dim3(256), 0, st, a);
    const int nb = a.p.nx * a.p.ny;
    hipLaunchKernelGGL(k_rerank, dim3((nb + kRerankWaves - 1) / kRerankWaves), dim3(64 * kRerankWaves), 0, st, a);
    return hipGetLastError();
}

size_t propose_bits_words(const StageParams& p) { return (size_t)p.nx * p.ny * kBitWords; }

}  // namespace lbm
}  // namespace bm3d
