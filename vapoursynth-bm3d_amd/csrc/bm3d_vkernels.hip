// V-BM3D kernels: spatio-temporal block matching with predictive search,
// aggregation into the per-frame numerator/denominator stack, and the
// VAggregate division.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bm3d_device.h"
#include "bm3d_kernels.h"
#include "bm3d_params.h"

namespace bm3d {

// ---------------------------------------------------------------------------
// K11: V-BM3D block matching, one wave per reference block
// ---------------------------------------------------------------------------
//
// VBM3D_Process_Base::BlockMatching (source/VBM3D_Base.cpp:540-659):
// full search in the current frame (self first), then for the frames
// cur-1 ... 0 and cur+1 ... T-1 a predictive search over the (y, x)-sorted
// union of +-ps_range windows around the best ps_num matches of the
// previously searched frame in that direction (Block::GenSearchPos,
// include/Block.h:771-827); each frame keeps its G best, the lists are
// concatenated and, when longer than G, cut to the self entry plus the G-1
// smallest keys.  Ties are ordered by search / concatenation index.

constexpr int kVMatchWaves = 2;
constexpr int kVBitWords = 128;  // bitmap of the predictive search union (4096 positions)

// Per-wave LDS layout.  STAGE: the search windows are staged in LDS (the
// current frame's full-search window and one region per predictive seed) and
// the SSDs read them there; otherwise they read global memory.
struct VMatchLayout {
    int max_cand, ps_cap, cat_cap, G;
    int rbw;        // reference block copy (B*B) or 0
    int cw;         // current-frame window words (STAGE) or 0
    int rw, nreg;   // words per predictive region and regions (STAGE) or 0
    __host__ __device__ int words() const {
        // keys[max(max_cand, ps_cap)] + sel (4G) + ps raw/flag/sorted
        // (3 ps_cap) + concatenated key/pos/frame (3 cat_cap) + seeds (G)
        // + union bitmap + reference block + staged windows (the current
        // frame's window and the predictive regions share one buffer: the
        // regions are staged only after the current frame is searched)
        int m = max_cand > ps_cap ? max_cand : ps_cap;
        const int stage = cw > rw * nreg ? cw : rw * nreg;
        return m + 4 * G + 3 * ps_cap + 3 * cat_cap + G + kVBitWords + rbw + stage;
    }
};

// Window of the full search (range r) or of one predictive seed: (2r + 8)
// rows of (2r + 8) | 1 words (odd, so lanes reading consecutive candidates
// hit distinct banks).
__host__ __device__ inline int vwin_stride(int r) { return (2 * r + 8) | 1; }
__host__ __device__ inline int vwin_words(int r) { return (2 * r + 8) * vwin_stride(r); }

__host__ __device__ inline VMatchLayout vmatch_layout(const StageParams& p, int max_cand, int ps_cap, int cat_cap,
                                                      bool stage) {
    VMatchLayout L{max_cand, ps_cap, cat_cap, p.group, 0, 0, 0, 0};
    const int B = p.block;
    L.rbw = (stage || B != 8) ? B * B : 0;
    if (stage) {
        const int range = p.range / p.step * p.step;
        const int ps_range = p.ps_range / p.ps_step * p.ps_step;
        L.cw = vwin_words(range);
        L.rw = vwin_words(ps_range);
        L.nreg = p.ps_num < p.group ? p.ps_num : p.group;
        if (L.nreg < 1) L.nreg = 1;
    }
    return L;
}

// Sorted, de-duplicated union of the search windows around seeds[0..ns).
// raw/out hold packed (y << 16) | x; returns the count (wave-uniform).  When
// the union's bounding box fits the bitmap, positions are marked in it and
// read back in row-major order (ascending packed value); otherwise first
// occurrences are ranked among the distinct values.
__device__ int gen_search_pos(const int* seeds, int ns, const StageParams& p, int B, int* raw, int* flag,
                              int* out, uint32_t* bits) {
    const int lane = lane_id();
    const int range = p.ps_range / p.ps_step * p.ps_step;
    if (ns <= 0) return 0;
    int by0 = 0x7fffffff, by1 = -1, bx0 = 0x7fffffff, bx1 = -1;
    for (int s = 0; s < ns; ++s) {
        const int sp = uniform(seeds[s]);
        const int sy = sp >> 16, sx = sp & 0xffff;
        by0 = min(by0, search_lo(sy, range, p.ps_step));
        by1 = max(by1, search_hi(sy, p.height - B, range, p.ps_step));
        bx0 = min(bx0, search_lo(sx, range, p.ps_step));
        bx1 = max(bx1, search_hi(sx, p.width - B, range, p.ps_step));
    }
    const int bw = bx1 - bx0 + 1;
    const int nwords = ((by1 - by0 + 1) * bw + 31) >> 5;
    if (nwords <= kVBitWords) {
        for (int w = lane; w < nwords; w += 64) bits[w] = 0u;
        wave_sync_lds();
        for (int s = 0; s < ns; ++s) {
            const int sp = uniform(seeds[s]);
            const int sy = sp >> 16, sx = sp & 0xffff;
            const int l = search_lo(sx, range, p.ps_step), r = search_hi(sx, p.width - B, range, p.ps_step);
            const int t = search_lo(sy, range, p.ps_step), b = search_hi(sy, p.height - B, range, p.ps_step);
            const int nc = (r - l) / p.ps_step + 1, nr = (b - t) / p.ps_step + 1;
            // 64 / nc window rows per pass (one division per seed)
            const int rpp = nc <= 64 ? 64 / nc : 1;
            const int yo = nc <= 64 ? lane / nc : 0;
            const int xo = nc <= 64 ? lane - yo * nc : lane;
            if (yo < rpp)
                for (int ry = yo; ry < nr; ry += rpp)
                    for (int rx = xo; rx < nc; rx += (nc <= 64 ? nc : 64)) {
                        const int bit = (t + ry * p.ps_step - by0) * bw + (l + rx * p.ps_step - bx0);
                        __hip_atomic_fetch_or(&bits[bit >> 5], 1u << (bit & 31), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WAVEFRONT);
                    }
        }
        wave_sync_lds();
        // positions in bit order, 64 bits (two words) per pass: lane l tests bit
        // b0 + l; the row is tracked incrementally (no division)
        const int nbits = (by1 - by0 + 1) * bw;
        int total = 0;
        int ry = lane / bw, rx = lane - (lane / bw) * bw;  // row / column of bit `lane`
        const int step_y = 64 / bw, step_x = 64 - (64 / bw) * bw;
        for (int b0 = 0; b0 < nbits; b0 += 64) {
            const int bit = b0 + lane;
            const bool set = bit < nbits && ((bits[bit >> 5] >> (bit & 31)) & 1u);
            const uint64_t m = __ballot(set);
            if (set) {
                const int at = total + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                out[at] = ((by0 + ry) << 16) | (bx0 + rx);
            }
            total += __popcll(m);
            // advance this lane's bit by 64 = step_y rows and step_x columns
            ry += step_y;
            rx += step_x;
            if (rx >= bw) {
                rx -= bw;
                ++ry;
            }
        }
        wave_sync_lds();
        return total;
    }
    // window of each seed (Block::AddSearchPos): row-major lattice
    int nraw = 0;
    for (int s = 0; s < ns; ++s) {
        int sp = seeds[s];
        int sy = sp >> 16, sx = sp & 0xffff;
        int l = search_lo(sx, range, p.ps_step), r = search_hi(sx, p.width - B, range, p.ps_step);
        int t = search_lo(sy, range, p.ps_step), b = search_hi(sy, p.height - B, range, p.ps_step);
        int nc = (r - l) / p.ps_step + 1, nr = (b - t) / p.ps_step + 1;
        for (int e = lane; e < nc * nr; e += 64) {
            int ry = e / nc;
            raw[nraw + e] = ((t + ry * p.ps_step) << 16) | (l + (e - ry * nc) * p.ps_step);
        }
        nraw += nc * nr;
    }
    wave_sync_lds();
    // mark first occurrences, then place each at its rank among distinct values
    for (int e0 = 0; e0 < nraw; e0 += 64) {
        int e = e0 + lane;
        if (e < nraw) {
            int v = raw[e];
            int first = 1;
            for (int q = 0; q < e; ++q)
                if (raw[q] == v) {
                    first = 0;
                    break;
                }
            flag[e] = first;
        }
    }
    wave_sync_lds();
    int total = 0;
    for (int e0 = 0; e0 < nraw; e0 += 64) {
        int e = e0 + lane;
        bool first = e < nraw && flag[e];
        if (first) {
            int v = raw[e], rank = 0;
            for (int q = 0; q < nraw; ++q) rank += (flag[q] && raw[q] < v);
            out[rank] = v;
        }
        total += __popcll(__ballot(first));
    }
    wave_sync_lds();
    return total;
}

// Copies rows t..t+nr+6, columns l..l+nc+6 of a plane (the 8x8 blocks at
// nr x nc positions) into an LDS window of row stride ws.
__device__ __forceinline__ void stage_window(float* win, int ws, const float* __restrict__ plane, int pitch, int t,
                                             int l, int rows, int cols) {
    const int lane = lane_id();
    if (cols <= 64) {
        // 64 / cols rows per pass, one division per call
        const int rpp = 64 / cols;
        const int yo = lane / cols, x = lane - yo * cols;
        if (yo < rpp)
            for (int y = yo; y < rows; y += rpp) win[y * ws + x] = plane[(size_t)(t + y) * pitch + l + x];
    } else {
        for (int y = 0; y < rows; ++y)
            for (int x = lane; x < cols; x += 64) win[y * ws + x] = plane[(size_t)(t + y) * pitch + l + x];
    }
}

template <int BT, bool STAGE, bool SCALAR>
__device__ __forceinline__ void vmatch_body(const VMatchArgs& a) {
    extern __shared__ uint32_t vsm[];
    const int wave = uniform(threadIdx.x >> 6);  // wave-uniform: scalar registers
    const int lane = lane_id();
    const int b = blockIdx.x * kVMatchWaves + wave;
    const StageParams& p = a.p;
    if (b >= a.nb) return;
    const int B = BT > 0 ? BT : p.block;
    const int G = p.group;
    const VMatchLayout L = vmatch_layout(p, a.max_cand, a.ps_cap, a.cat_cap, STAGE);
    uint32_t* base = vsm + (size_t)wave * L.words();
    const int kcap = a.max_cand > a.ps_cap ? a.max_cand : a.ps_cap;
    uint32_t* keys = base;
    uint32_t* sel_key = keys + kcap;
    int* sel_idx = reinterpret_cast<int*>(sel_key + 2 * G);
    int* ps_raw = sel_idx + 2 * G;
    int* ps_flag = ps_raw + a.ps_cap;
    int* ps_pos = ps_flag + a.ps_cap;
    uint32_t* cat_key = reinterpret_cast<uint32_t*>(ps_pos + a.ps_cap);
    int* cat_pos = reinterpret_cast<int*>(cat_key + a.cat_cap);
    int* cat_frm = cat_pos + a.cat_cap;
    int* seeds = cat_frm + a.cat_cap;
    uint32_t* bits = reinterpret_cast<uint32_t*>(seeds + G);
    float* rbl = reinterpret_cast<float*>(bits + kVBitWords);
    float* cwin = rbl + L.rbw;
    float* regs = cwin;  // the current-frame window is dead once the predictive search starts

    const int bj = b / p.nx, bi = b - bj * p.nx;
    const int j = grid_pos(bj, p.block_step, p.height - B);
    const int i = grid_pos(bi, p.block_step, p.width - B);
    int* cnt_out = a.count + b;
    int* pos_out = a.pos + (size_t)b * G;
    int* frm_out = a.frame + (size_t)b * G;
    float* key_out = a.key + (size_t)b * G;
    if (G == 1 || !(p.th_sse > 0.0f && p.dist_mul > 0.0)) {
        if (lane == 0) {
            *cnt_out = 1;
            pos_out[0] = (j << 16) | i;
            frm_out[0] = a.cur;
            key_out[0] = 0.0f;
        }
        return;
    }
    const float* cur_plane = a.planes[a.cur];
    // --- current frame: full search with the self entry first
    const int l = search_lo(i, p.range, p.step);
    const int r = search_hi(i, p.width - B, p.range, p.step);
    const int t = search_lo(j, p.range, p.step);
    const int bt = search_hi(j, p.height - B, p.range, p.step);
    const int ncols = (r - l) / p.step + 1;
    const int nc = ncols * ((bt - t) / p.step + 1);
    const int range = p.range / p.step * p.step;
    const int cws = vwin_stride(range);
    float rb_reg[(BT > 0 && !STAGE) ? BT * BT : 1];
    const float* rb = rbl;
    if constexpr (STAGE) {
        stage_window(cwin, cws, cur_plane, p.pitch, t, l, bt - t + B, r - l + B);
        wave_sync_lds();
        for (int e = lane; e < B * B; e += 64) rbl[e] = cwin[(j - t + e / B) * cws + i - l + e % B];
        wave_sync_lds();
    } else if constexpr (BT > 0) {
#pragma unroll
        for (int e = 0; e < BT * BT; ++e) rb_reg[e] = cur_plane[(size_t)(j + e / BT) * p.pitch + i + e % BT];
        rb = rb_reg;
    } else {
        for (int e = lane; e < B * B; e += 64) rbl[e] = cur_plane[(size_t)(j + e / B) * p.pitch + i + e % B];
        wave_sync_lds();
    }
    auto ssd_at = [&](const float* q, int pitch) {
        if constexpr (BT > 0) return ssd_fixed<BT, SCALAR>(rb, q, pitch);
        else return ssd_generic(rb, q, pitch, B, SCALAR);
    };

    int nacc = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        int c = c0 + lane;
        uint32_t kv = 0xFFFFFFFFu;
        if (c < nc) {
            int cy = c / ncols;
            int y = t + cy * p.step, x = l + (c - cy * ncols) * p.step;
            if (!(y == j && x == i)) {
                float dist = STAGE ? ssd_at(cwin + (y - t) * cws + (x - l), cws)
                                   : ssd_at(cur_plane + (size_t)y * p.pitch + x, p.pitch);
                if (dist <= p.th_sse && dist != 0.0f) kv = __float_as_uint((float)((double)dist * p.dist_mul));
            }
            keys[c] = kv;
        }
        nacc += __popcll(__ballot(kv != 0xFFFFFFFFu));
    }
    wave_sync_lds();
    int k = nacc < G - 1 ? nacc : G - 1;
    if (nc <= 128) wave_select_small<2>(keys, nc, nacc, k, sel_key, sel_idx);
    else if (nc <= 256) wave_select_small<4>(keys, nc, nacc, k, sel_key, sel_idx);
    else wave_select(keys, nc, nacc, k, sel_key, sel_idx);
    int ncat = 1 + k;
    if (lane == 0) {
        cat_key[0] = 0u;
        cat_pos[0] = (j << 16) | i;
        cat_frm[0] = a.cur;
    }
    for (int e = lane; e < k; e += 64) {
        int c = sel_idx[k + e];
        int cy = c / ncols;
        cat_key[1 + e] = sel_key[k + e];
        cat_pos[1 + e] = ((t + cy * p.step) << 16) | (l + (c - cy * ncols) * p.step);
        cat_frm[1 + e] = a.cur;
    }
    wave_sync_lds();
    const int ncur_seeds = ncat < p.ps_num ? ncat : p.ps_num;
    const int ps_range = p.ps_range / p.ps_step * p.ps_step;
    const int rws = vwin_stride(ps_range);

    // --- predictive search, backward then forward
    for (int dir = -1; dir <= 1; dir += 2) {
        int ns = ncur_seeds;
        for (int e = lane; e < ns; e += 64) seeds[e] = cat_pos[e];
        wave_sync_lds();
        for (int f = a.cur + dir; f >= 0 && f < a.nframes; f += dir) {
            int np_ = gen_search_pos(seeds, ns, p, B, ps_raw, ps_flag, ps_pos, bits);
            const float* plane = a.planes[f];
            if constexpr (STAGE) {
                // one region per seed: the blocks of its window
                for (int s = 0; s < ns; ++s) {
                    const int sp = uniform(seeds[s]);
                    const int sy = sp >> 16, sx = sp & 0xffff;
                    const int sl = search_lo(sx, ps_range, p.ps_step);
                    const int sr = search_hi(sx, p.width - B, ps_range, p.ps_step);
                    const int st_ = search_lo(sy, ps_range, p.ps_step);
                    const int sb = search_hi(sy, p.height - B, ps_range, p.ps_step);
                    stage_window(regs + s * L.rw, rws, plane, p.pitch, st_, sl, sb - st_ + B, sr - sl + B);
                }
                wave_sync_lds();
            }
            int acc = 0;
            for (int c0 = 0; c0 < np_; c0 += 64) {
                int c = c0 + lane;
                uint32_t kv = 0xFFFFFFFFu;
                if (c < np_) {
                    int pk = ps_pos[c];
                    const int y = pk >> 16, x = pk & 0xffff;
                    float dist;
                    if constexpr (STAGE) {
                        // the first seed whose window holds (y, x)
                        int s = 0, ry = 0, rx = 0;
                        for (; s < ns; ++s) {
                            const int sp = seeds[s];
                            const int sy = sp >> 16, sx = sp & 0xffff;
                            ry = y - search_lo(sy, ps_range, p.ps_step);
                            rx = x - search_lo(sx, ps_range, p.ps_step);
                            if (ry >= 0 && rx >= 0 && y <= search_hi(sy, p.height - B, ps_range, p.ps_step) &&
                                x <= search_hi(sx, p.width - B, ps_range, p.ps_step))
                                break;
                        }
                        dist = ssd_at(regs + s * L.rw + ry * rws + rx, rws);
                    } else {
                        dist = ssd_at(plane + (size_t)y * p.pitch + x, p.pitch);
                    }
                    if (dist <= p.th_sse && dist != 0.0f) kv = __float_as_uint((float)((double)dist * p.dist_mul));
                    keys[c] = kv;
                }
                acc += __popcll(__ballot(kv != 0xFFFFFFFFu));
            }
            wave_sync_lds();
            int kf = acc < G ? acc : G;
            if (np_ <= 64) wave_select_small<1>(keys, np_, acc, kf, sel_key, sel_idx);
            else if (np_ <= 128) wave_select_small<2>(keys, np_, acc, kf, sel_key, sel_idx);
            else if (np_ <= 192) wave_select_small<3>(keys, np_, acc, kf, sel_key, sel_idx);
            else if (np_ <= 256) wave_select_small<4>(keys, np_, acc, kf, sel_key, sel_idx);
            else wave_select(keys, np_, acc, kf, sel_key, sel_idx);
            for (int e = lane; e < kf; e += 64) {
                cat_key[ncat + e] = sel_key[kf + e];
                cat_pos[ncat + e] = ps_pos[sel_idx[kf + e]];
                cat_frm[ncat + e] = f;
            }
            wave_sync_lds();
            // seeds of the next frame: the best ps_num matches of this one
            ns = kf < p.ps_num ? kf : p.ps_num;
            for (int e = lane; e < ns; e += 64) seeds[e] = cat_pos[ncat + e];
            ncat += kf;
            wave_sync_lds();
        }
    }

    // --- final cut (VBM3D_Base.cpp:652-656): keep order when within G
    if (ncat <= G) {
        for (int e = lane; e < ncat; e += 64) {
            pos_out[e] = cat_pos[e];
            frm_out[e] = cat_frm[e];
            key_out[e] = __uint_as_float(cat_key[e]);
        }
        if (lane == 0) *cnt_out = ncat;
        return;
    }
    // select G-1 of entries 1..ncat-1 by (key, concatenation index)
    int m = ncat - 1;
    if (m <= 64) wave_select_small<1>(cat_key + 1, m, m, G - 1, sel_key, sel_idx);
    else if (m <= 256) wave_select_small<4>(cat_key + 1, m, m, G - 1, sel_key, sel_idx);
    else wave_select(cat_key + 1, m, m, G - 1, sel_key, sel_idx);
    if (lane == 0) {
        *cnt_out = G;
        pos_out[0] = cat_pos[0];
        frm_out[0] = cat_frm[0];
        key_out[0] = __uint_as_float(cat_key[0]);
    }
    for (int e = lane; e < G - 1; e += 64) {
        int c = 1 + sel_idx[G - 1 + e];
        pos_out[1 + e] = cat_pos[c];
        frm_out[1 + e] = cat_frm[c];
        key_out[1 + e] = __uint_as_float(sel_key[G - 1 + e]);
    }
}

// ---------------------------------------------------------------------------
// V-BM3D aggregation into the stacked num/den slots
// ---------------------------------------------------------------------------
//
// VBM3D_Process_Base::Kernel writes each member into the numerator and
// denominator planes of its own frame (BlockGroup Pos3 AddTo / CountTo,
// include/Block.h:1283-1303, 1368-1387) in reference-block scan order, then
// member order.  One wave per (tile, frame): it walks the reference blocks
// that can reach the tile in that frame and accumulates the members of that
// frame in exactly that order.  Output slot 2*(radius + f - cur) (num) and
// +1 (den) of the output stack (VBM3D_Base.h:284-292).

template <int BT, bool STAGE>
__global__ __launch_bounds__(64 * kVMatchWaves) void k_vmatch(VMatchArgs a) {
    if (a.p.arith) vmatch_body<BT, STAGE, true>(a);
    else vmatch_body<BT, STAGE, false>(a);
}

// Per-frame member lists for the stack walk: one wave per (block row ar,
// frame f).  roff[(f * ny + ar) * (nx + 1) + c] = members of frame f held by
// the row's blocks before column c (c = nx: the row's total);
// rpair[(f * nb + ar * nx) * G + k] = (packed position, b * G + z) of those
// members in scan order (blocks by column, members by z).
__global__ __launch_bounds__(64) void k_vframe_rows(VStackArgs a) {
    const int ar = blockIdx.x, f = blockIdx.y;
    const int lane = lane_id();
    const int nx = a.p.nx, G = a.p.group;
    int* ro = a.roff + ((size_t)f * a.p.ny + ar) * (nx + 1);
    int2* rp = a.rpair + ((size_t)f * nx * a.p.ny + (size_t)ar * nx) * G;
    int carry = 0;
    for (int c0 = 0; c0 < nx; c0 += 64) {
        const int c = c0 + lane;
        const int b = ar * nx + c;
        const int n = c < nx ? a.count[b] : 0;
        int k = 0;
        for (int z = 0; z < n; ++z) k += a.frame[b * G + z] == f;
        int incl = k;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        int w = carry + incl - k;
        if (c < nx) ro[c] = w;
        for (int z = 0; z < n; ++z)
            if (a.frame[b * G + z] == f) rp[w++] = int2{a.pos[b * G + z], b * G + z};
        carry += __shfl(incl, 63);
    }
    if (lane == 0) ro[nx] = carry;
}

__global__ __launch_bounds__(64) void k_vstack(VStackArgs a) {
    __shared__ float acc[kAggLds];
    __shared__ int rec[kAggRecCap * kAggRecWords];
    const int lane = lane_id();
    const int tiles_x = (a.p.width + kAggTW - 1) / kAggTW, tiles_y = (a.p.height + kAggTH - 1) / kAggTH;
    const int ntiles = tiles_x * tiles_y;
    const int fr = blockIdx.y;
    const int per_xcd = (ntiles + 7) >> 3;
    const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int y0 = ty * kAggTH, x0 = tx * kAggTW;
    const int dist = fr > a.cur ? fr - a.cur : a.cur - fr;
    const int reach = a.p.range + dist * (a.p.ps_range / a.p.ps_step * a.p.ps_step);
    int a_lo, a_hi, c_lo, c_hi;
    reach_rect(a.p, y0, x0, kAggTH, kAggTW, reach, a_lo, a_hi, c_lo, c_hi);
    if (a.roff && a_hi - a_lo < 64)
        agg_rows_any(a.p, a.process[0] != 0, a.process[1] != 0, a.process[2] != 0,
                     a.roff + (size_t)fr * a.p.ny * (a.p.nx + 1), a.rpair + (size_t)fr * a.p.nx * a.p.ny * a.p.group,
                     a_lo, a_hi, c_lo, c_hi, y0, x0, a.contrib[0], a.contrib[1], a.contrib[2], a.denw[0], a.denw[1],
                     a.denw[2], acc, rec);
    else
        agg_tile_any(a.p, a.process[0] != 0, a.process[1] != 0, a.process[2] != 0, a.count, a.pos, a.frame, fr,
                     a_lo, a_hi, c_lo, c_hi, y0, x0, a.contrib[0], a.contrib[1], a.contrib[2], a.denw[0], a.denw[1],
                     a.denw[2], acc, rec);
    const int slot = 2 * (a.p.radius + fr - a.cur);
    for (int i = lane; i < kAggTH * kAggTW; i += 64) {
        const int ly = i / kAggTW, lx = i - ly * kAggTW;
        const int py = y0 + ly, px = x0 + lx;
        if (py >= a.p.height || px >= a.p.width) continue;
        int k = 0;
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
            if (!a.process[pl]) continue;
            const float* an = acc + k * 2 * kAggPlane + ly * kAggStride + lx;
            ++k;
            float* basep = a.dst[pl];
            basep[((size_t)slot * a.p.height + py) * a.dst_pitch[pl] + px] = an[0];
            basep[((size_t)(slot + 1) * a.p.height + py) * a.dst_pitch[pl] + px] = an[kAggPlane];
        }
    }
}

// ---------------------------------------------------------------------------
// K12: VAggregate (source/VAggregate.cpp:93-160)
// ---------------------------------------------------------------------------
//
// out = (sum_m num_m) / (sum_m den_m) over the neighbours m = n-b .. n+f in
// ascending order, slot radius - (m - n) of neighbour m's stack, then the
// Float2Int conversion.  Streaming, one thread per pixel.

__global__ void k_vagg(VAggArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int pl = blockIdx.z;
    if (x >= a.W || y >= a.H || !a.process[pl]) return;
    float num = 0.0f, den = 0.0f;
    for (int i = 0; i < a.nframes; ++i) {
        const int o = a.radius + a.cur - i;
        const float* st = a.stacks[pl * kMaxStackFrames + i];
        const size_t sp = a.stack_pitch[pl];
        num = num + st[((size_t)(2 * o) * a.H + y) * sp + x];
        den = den + st[((size_t)(2 * o + 1) * a.H + y) * sp + x];
    }
    float v = num / den;
    if (a.out_bits != 0) {
        v = v * a.gain[pl] + a.offset[pl];
        v = clipf(v, a.lo[pl], a.hi[pl]);
    }
    store_sample(a.dst[pl], a.out_bits, (size_t)y * a.dst_pitch[pl] + x, v);
}

// ---------------------------------------------------------------------------

// LDS of the global-read form (the staged form is used when it fits too).
size_t vmatch_lds_bytes(const StageParams& p, int max_cand, int ps_cap, int cat_cap) {
    return sizeof(uint32_t) * (size_t)kVMatchWaves * vmatch_layout(p, max_cand, ps_cap, cat_cap, false).words();
}

hipError_t launch_vmatch(const VMatchArgs& a, hipStream_t st) {
    const int blocks = (a.nb + kVMatchWaves - 1) / kVMatchWaves;
    if (blocks == 0) return hipSuccess;
    const size_t staged =
        sizeof(uint32_t) * (size_t)kVMatchWaves * vmatch_layout(a.p, a.max_cand, a.ps_cap, a.cat_cap, true).words();
    if (a.p.block == 8 && staged <= 65536) {
        hipLaunchKernelGGL((k_vmatch<8, true>), dim3(blocks), dim3(64 * kVMatchWaves), staged, st, a);
        return hipGetLastError();
    }
    const size_t lds = vmatch_lds_bytes(a.p, a.max_cand, a.ps_cap, a.cat_cap);
    if (a.p.block == 8)
        hipLaunchKernelGGL((k_vmatch<8, false>), dim3(blocks), dim3(64 * kVMatchWaves), lds, st, a);
    else
        hipLaunchKernelGGL((k_vmatch<0, false>), dim3(blocks), dim3(64 * kVMatchWaves), lds, st, a);
    return hipGetLastError();
}

size_t vframe_rows_words(const StageParams& p, int nframes) { return (size_t)nframes * p.ny * (p.nx + 1); }
size_t vframe_pairs(const StageParams& p, int nframes) { return (size_t)nframes * p.nx * p.ny * p.group; }

hipError_t launch_vstack(const VStackArgs& a, hipStream_t st) {
    if (a.roff) {
        if (a.p.block * a.p.block > 64) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_vframe_rows, dim3(a.p.ny, a.nframes), dim3(64), 0, st, a);
    }
    const int tiles = ((a.p.width + kAggTW - 1) / kAggTW) * ((a.p.height + kAggTH - 1) / kAggTH);
    hipLaunchKernelGGL(k_vstack, dim3(8 * ((tiles + 7) / 8), a.nframes), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_vagg(const VAggArgs& a, hipStream_t st) {
    dim3 grid((a.W + 255) / 256, a.H, 3);
    hipLaunchKernelGGL(k_vagg, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace bm3d
