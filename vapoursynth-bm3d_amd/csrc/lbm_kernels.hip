// Learned block-matching proposal (LBM) on CDNA4: a patch-embedding MLP
// 64 -> 128 (ReLU) -> 32 with BF16 weights and FP32 accumulation on MFMA
// (v_mfma_f32_16x16x16bf16_1k), its training kernels (forward, loss,
// backward, Adam), and the proposal path used by block matching:
// embeddings of every patch position, FP16 embedding-distance scores of
// every candidate in each reference block's search window, top-K proposal
// selection, exact FP32 SSD re-rank with the plain matcher's acceptance and
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

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

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
    s4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        raw[i] = p[i];
        v[i] = f2bf(raw[i]);
    }
    return v;
}

// Evaluates the network on 16 patches (one per lane % 16).  x: the B
// fragments of the patches; h: post-ReLU hidden (D layout, fp32); e: the
// embeddings (D layout: lane l holds embedding 16u + 4q + i of patch l % 16).
__device__ __forceinline__ void forward16(const Frags& f, const s4 (&x)[4], float (&h)[8][4], float (&e)[2][4]) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(f.w1[t][s], x[s], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = acc[i] + f.b1[t][i];
            h[t][i] = v > 0.0f ? v : 0.0f;
        }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            s4 hb;
#pragma unroll
            for (int i = 0; i < 4; ++i) hb[i] = f2bf(h[s][i]);
            acc = mfma(f.w2[u][s], hb, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) e[u][i] = acc[i] + f.b2[u][i];
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
// inference: embeddings of every patch position
// ---------------------------------------------------------------------------

// emb[(y * ew + x) * 32 + k] (fp16) and norm[y * ew + x] = sum_k emb^2 (fp32 of
// the fp16 values) for y in [0, H-8], x in [0, W-8]; ew = W - 7.  Each wave
// runs kEmbedGroups groups of 16 consecutive positions of one row.
constexpr int kEmbedGroups = 8;
__global__ __launch_bounds__(256) void k_embed(Net n, const float* __restrict__ plane, int pitch, int W, int H,
                                               _Float16* __restrict__ emb, float* __restrict__ norm) {
    const int lane = lane_id();
    const int ew = W - 7, eh = H - 7;
    const int per_row = (ew + 16 * kEmbedGroups - 1) / (16 * kEmbedGroups);
    const int wave = uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (wave >= per_row * eh) return;
    const int y = wave / per_row;
    const int xs = (wave - y * per_row) * 16 * kEmbedGroups;
    Frags f;
    load_frags(n, f);
    const int r = lane & 15, q = lane >> 4;
    for (int g = 0; g < kEmbedGroups; ++g) {
        const int x0 = xs + 16 * g;
        if (x0 >= ew) break;
        const int x = x0 + r < ew ? x0 + r : ew - 1;
        s4 xf[4];
        float raw[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) xf[s] = patch_frag(plane, pitch, y, x, s, raw);
        float h[8][4], e[2][4];
        forward16(f, xf, h, e);
        float nrm = 0.0f;
        _Float16 eh16[2][4];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                eh16[u][i] = (_Float16)e[u][i];
                const float ev = (float)eh16[u][i];
                nrm += ev * ev;
            }
        // lanes r, r+16, r+32, r+48 hold the four quarters of one embedding
        nrm += __shfl_xor(nrm, 16);
        nrm += __shfl_xor(nrm, 32);
        if (x0 + r < ew) {
            const size_t base = ((size_t)y * ew + x0 + r) * kEmb;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                _Float16* o = emb + base + 16 * u + 4 * q;
#pragma unroll
                for (int i = 0; i < 4; ++i) o[i] = eh16[u][i];
            }
            if (q == 0) norm[(size_t)y * ew + x0 + r] = nrm;
        }
    }
}

// ---------------------------------------------------------------------------
// inference: scores of every candidate of every reference block
// ---------------------------------------------------------------------------
//
// score(r, c) = |e_r|^2 + |e_c|^2 - 2 e_r.e_c (clamped at 0), fp16, stored at
// scores[b * max_cand + cy * ncols + cx] in the window's row-major lattice
// order (the self position included; it is skipped by the selection).  A
// workgroup takes kTileRY x kTileRX reference blocks; the union of their
// windows is streamed through LDS kChunkRows rows at a time and every thread
// scores its union positions against every reference block of the tile
// whose window contains them.
constexpr int kTileRY = 4, kTileRX = 8, kTileRefs = kTileRY * kTileRX;
constexpr int kChunkRows = 8;

__device__ __forceinline__ float dot32(const uint32_t* a, const uint4 (&b)[4]) {
    float acc = 0.0f;
    const uint32_t bw[16] = {b[0].x, b[0].y, b[0].z, b[0].w, b[1].x, b[1].y, b[1].z, b[1].w,
                             b[2].x, b[2].y, b[2].z, b[2].w, b[3].x, b[3].y, b[3].z, b[3].w};
#pragma unroll
    for (int k = 0; k < 16; ++k)
        acc = __builtin_amdgcn_fdot2(__builtin_bit_cast(h2, a[k]), __builtin_bit_cast(h2, bw[k]), acc, false);
    return acc;
}

__global__ __launch_bounds__(256) void k_score(ScoreArgs a) {
    extern __shared__ uint4 lds4[];
    __shared__ uint32_t eref[kTileRefs][16];
    __shared__ float nref[kTileRefs];
    __shared__ int win[kTileRefs][5];  // t, l, nrows, ncols, valid
    const StageParams& p = a.p;
    const int tiles_x = (p.nx + kTileRX - 1) / kTileRX;
    const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
    const int ew = p.width - 7;
    const int B = 8;
    // tile geometry and the union of the windows
    int u_t = 1 << 30, u_b = -1, u_l = 1 << 30, u_r = -1;
    for (int k = 0; k < kTileRefs; ++k) {
        const int bj = ty * kTileRY + k / kTileRX, bi = tx * kTileRX + k % kTileRX;
        if (bj >= p.ny || bi >= p.nx) continue;
        const int j = grid_pos(bj, p.block_step, p.height - B), i = grid_pos(bi, p.block_step, p.width - B);
        const int l = search_lo(i, p.range, 1), rr = search_hi(i, p.width - B, p.range, 1);
        const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
        u_t = t < u_t ? t : u_t;
        u_b = bt > u_b ? bt : u_b;
        u_l = l < u_l ? l : u_l;
        u_r = rr > u_r ? rr : u_r;
    }
    const int tid = threadIdx.x;
    if (tid < kTileRefs) {
        const int k = tid;
        const int bj = ty * kTileRY + k / kTileRX, bi = tx * kTileRX + k % kTileRX;
        const bool ok = bj < p.ny && bi < p.nx;
        int j = 0, i = 0;
        if (ok) {
            j = grid_pos(bj, p.block_step, p.height - B);
            i = grid_pos(bi, p.block_step, p.width - B);
            const int l = search_lo(i, p.range, 1), rr = search_hi(i, p.width - B, p.range, 1);
            const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
            win[k][0] = t;
            win[k][1] = l;
            win[k][2] = bt - t + 1;
            win[k][3] = rr - l + 1;
        } else {
            win[k][0] = win[k][1] = win[k][2] = win[k][3] = 0;
        }
        win[k][4] = ok;
        const size_t e = (size_t)j * ew + i;
        const uint4* src = reinterpret_cast<const uint4*>(a.emb + e * kEmb);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint4 v = src[w];
            eref[k][4 * w + 0] = v.x;
            eref[k][4 * w + 1] = v.y;
            eref[k][4 * w + 2] = v.z;
            eref[k][4 * w + 3] = v.w;
        }
        nref[k] = a.norm[e];
    }
    const int ucols = u_r - u_l + 1;
    for (int c0 = u_t; c0 <= u_b; c0 += kChunkRows) {
        const int rows = (u_b - c0 + 1) < kChunkRows ? (u_b - c0 + 1) : kChunkRows;
        __syncthreads();
        for (int e = tid; e < rows * ucols * 4; e += 256) {
            const int pos = e >> 2, w = e & 3;
            const int yy = c0 + pos / ucols, xx = u_l + pos % ucols;
            lds4[e] = reinterpret_cast<const uint4*>(a.emb + ((size_t)yy * ew + xx) * kEmb)[w];
        }
        __syncthreads();
        for (int pos = tid; pos < rows * ucols; pos += 256) {
            const int yy = c0 + pos / ucols, xx = u_l + pos % ucols;
            uint4 eq[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) eq[w] = lds4[pos * 4 + w];
            const float nq = a.norm[(size_t)yy * ew + xx];
            for (int k = 0; k < kTileRefs; ++k) {
                const int t = win[k][0], l = win[k][1];
                const int cy = yy - t, cx = xx - l;
                if (!win[k][4] || cy < 0 || cy >= win[k][2] || cx < 0 || cx >= win[k][3]) continue;
                float d = nref[k] + nq - 2.0f * dot32(eref[k], eq);
                d = d > 0.0f ? d : 0.0f;
                const int bj = ty * kTileRY + k / kTileRX, bi = tx * kTileRX + k % kTileRX;
                const size_t b = (size_t)bj * p.nx + bi;
                a.scores[b * a.max_cand + cy * win[k][3] + cx] = (_Float16)d;
            }
        }
    }
}

size_t score_lds_bytes(const StageParams& p) {
    const int ucols = (kTileRX - 1) * p.block_step + 2 * p.range + 1;
    return (size_t)kChunkRows * ucols * 4 * sizeof(uint4);
}

// ---------------------------------------------------------------------------
// inference: proposal selection and exact re-rank
// ---------------------------------------------------------------------------
//
// One wave per reference block.  Keys (fp16 score bits << 16 | candidate
// index) are unique, so the K smallest are the K best scores with ties going
// to the earlier candidate.  The K proposals are then scored exactly like the
// plain matcher (ssd_fixed<8>, dist <= thSSE && dist != 0, key =
// float(double(dist) * distMul)) and the G-1 best by (key, candidate index)
// form the group after the self entry.  A block whose window holds no more
// than K candidates is searched exhaustively here; a block with fewer than
// G-1 accepted proposals is appended to the fallback list.
__global__ __launch_bounds__(256) void k_select(SelectArgs a) {
    __shared__ uint32_t sk[4][kMaxProposals];
    __shared__ uint32_t ek[4][kMaxProposals];
    const StageParams& p = a.p;
    const int lane = lane_id();
    const int wave = uniform((int)(threadIdx.x >> 6));
    const int b = uniform((int)(blockIdx.x * 4 + wave));
    if (b >= a.nb) return;
    constexpr int B = 8;
    const int G = p.group;
    const int K = a.topk;
    const int bj = b / p.nx, bi = b - bj * p.nx;
    const int j = grid_pos(bj, p.block_step, p.height - B), i = grid_pos(bi, p.block_step, p.width - B);
    const int l = search_lo(i, p.range, 1), r = search_hi(i, p.width - B, p.range, 1);
    const int t = search_lo(j, p.range, 1), bt = search_hi(j, p.height - B, p.range, 1);
    const int ncols = r - l + 1, nrows = bt - t + 1, nc = ncols * nrows;
    const int self = (j - t) * ncols + (i - l);
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
    // --- proposals: the K smallest keys among the nc - 1 candidates
    uint32_t* props = sk[wave];
    int nprop;
    if (nc - 1 <= K) {
        nprop = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int c = c0 + lane;
            const bool in = c < nc && c != self;
            const uint64_t m = __ballot(in);
            const int at = nprop + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            if (in) props[at] = (uint32_t)c;
            nprop += __popcll(m);
        }
    } else {
        const unsigned short* sc = reinterpret_cast<const unsigned short*>(a.scores) + (size_t)b * a.max_cand;
        constexpr int NS = (kMaxWindow + 63) / 64;
        uint32_t keys[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int c = lane + 64 * s;
            keys[s] = (c < nc && c != self) ? (((uint32_t)sc[c] << 16) | (uint32_t)c) : 0xFFFFFFFFu;
        }
        // radix descent for the K-th smallest key
        uint32_t tv = 0;
        for (int bit = 31; bit >= 0; --bit) {
            const uint32_t trial = tv | (1u << bit);
            int cnt = 0;
#pragma unroll
            for (int s = 0; s < NS; ++s) cnt += __popcll(__ballot(keys[s] < trial));
            if (cnt < K) tv = trial;
        }
        nprop = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const bool in = keys[s] <= tv;
            const uint64_t m = __ballot(in);
            const int at = nprop + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            if (in) props[at] = keys[s] & 0xffffu;
            nprop += __popcll(m);
        }
    }
    wave_sync_lds();
    // --- exact re-rank of the proposals (one lane per proposal)
    float rb[64];
#pragma unroll
    for (int e = 0; e < 64; ++e) rb[e] = a.plane[(size_t)(j + e / 8) * p.pitch + i + e % 8];
    uint32_t* ekey = ek[wave];
    int nacc = 0;
    for (int c0 = 0; c0 < nprop; c0 += 64) {
        const int s = c0 + lane;
        uint32_t kv = 0xFFFFFFFFu;
        int c = 0;
        if (s < nprop) {
            c = (int)props[s];
            const int cy = c / ncols, cx = c - cy * ncols;
            const float dist = ssd_fixed<8>(rb, a.plane + (size_t)(t + cy) * p.pitch + l + cx, p.pitch);
            if (dist <= p.th_sse && dist != 0.0f) kv = __float_as_uint((float)((double)dist * p.dist_mul));
        }
        const bool ok = kv != 0xFFFFFFFFu;
        nacc += __popcll(__ballot(ok));
        if (s < nprop) ekey[s] = kv;
    }
    wave_sync_lds();
    const int kk = nacc < G - 1 ? nacc : G - 1;
    if (nacc < G - 1 && nc - 1 > K) {
        // not enough acceptable proposals: the exhaustive search decides
        if (lane == 0) {
            const int at = atomicAdd(a.nfallback, 1);
            a.fallback[at] = b;
        }
        return;
    }
    // rank the accepted proposals by (key, candidate index); keep the kk best
    for (int s = lane; s < nprop; s += 64) {
        const uint32_t kv = ekey[s];
        if (kv == 0xFFFFFFFFu) continue;
        const int c = (int)props[s];
        int rank = 0;
        for (int o = 0; o < nprop; ++o) {
            const uint32_t k2 = ekey[o];
            const int c2 = (int)props[o];
            rank += (k2 < kv) || (k2 == kv && c2 < c);
        }
        if (rank < kk) {
            const int cy = c / ncols;
            pos_out[1 + rank] = ((t + cy) << 16) | (l + c - cy * ncols);
            key_out[1 + rank] = __uint_as_float(kv);
        }
    }
    if (lane == 0) {
        *cnt_out = kk + 1;
        pos_out[0] = (j << 16) | i;
        key_out[0] = 0.0f;
    }
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
    const int ew = W - 7, eh = H - 7;
    if (ew <= 0 || eh <= 0) return hipSuccess;
    const int per_row = (ew + 16 * kEmbedGroups - 1) / (16 * kEmbedGroups);
    const int waves = per_row * eh;
    hipLaunchKernelGGL(k_embed, dim3((waves + 3) / 4), dim3(256), 0, st, n, plane, pitch, W, H, emb, norm);
    return hipGetLastError();
}

hipError_t launch_score(const ScoreArgs& a, hipStream_t st) {
    const int tiles = ((a.p.ny + kTileRY - 1) / kTileRY) * ((a.p.nx + kTileRX - 1) / kTileRX);
    hipLaunchKernelGGL(k_score, dim3(tiles), dim3(256), score_lds_bytes(a.p), st, a);
    return hipGetLastError();
}

hipError_t launch_select(const SelectArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_select, dim3((a.nb + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace lbm
}  // namespace bm3d
