/*
 * bm3d_oracle.c -- CPU restatement of the VapourSynth-BM3D filtering path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * bm3d_oracle.h for the list of functions, the reference file:line each one
 * follows and how each is pinned.  Build: oracle/Makefile (gcc, -O2,
 * -ffp-contract=off so that no multiply-add is fused behind our back: the
 * reference is built for baseline x86-64, which has no FMA).
 */
#include "bm3d_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__SSE2__)
#include <emmintrin.h>
#endif

/* ========================================================================== */
/* parameters                                                                 */
/* ========================================================================== */

int or_profile_defaults(const char *profile, int wiener, int temporal, or_params *o, double sigma0)
{
    memset(o, 0, sizeof(*o));
    o->wiener = wiener;
    /* BM3D_Para ctor, source/BM3D.cpp:33-107 */
    o->block_size = 8;
    o->bm_range = 16;
    o->bm_step = 1;
    if (!wiener) {
        o->block_step = 4;
        o->group_size = 16;
        o->lambda = 2.7;
    } else {
        o->block_step = 3;
        o->group_size = 32;
        o->lambda = 0.0; /* not used by the Wiener filter */
    }
    if (!strcmp(profile, "fast")) {
        o->bm_range = 9;
        o->group_size = 8;
        o->block_step = wiener ? 7 : 8;
    } else if (!strcmp(profile, "lc")) {
        o->bm_range = 9;
        if (!wiener) {
            o->block_step = 6;
        } else {
            o->block_step = 5;
            o->group_size = 16;
        }
    } else if (!strcmp(profile, "np")) {
    } else if (!strcmp(profile, "high")) {
        o->block_step = wiener ? 2 : 3;
    } else if (!strcmp(profile, "vn")) {
        if (!wiener) {
            o->block_step = 4;
            o->group_size = 32;
            o->lambda = 2.8;
        } else {
            o->block_size = 11;
            o->block_step = 6;
        }
    } else {
        return -1;
    }
    /* thMSE_Default, source/BM3D.cpp:110-130 */
    int vn = !strcmp(profile, "vn");
    if (!wiener)
        o->th_mse = vn ? 1000 + sigma0 * 150 : 400 + sigma0 * 80;
    else
        o->th_mse = vn ? 400 + sigma0 * 40 : 200 + sigma0 * 10;
    if (temporal) {
        /* VBM3D_Para ctor, source/VBM3D_Base.cpp:32-97 */
        o->radius = 3;
        o->group_size = 8;
        o->bm_range = 12;
        o->ps_num = 2;
        o->ps_step = 1;
        o->ps_range = wiener ? 6 : 5;
        if (!strcmp(profile, "fast")) {
            o->radius = 1;
            o->bm_range = 7;
            o->ps_range = wiener ? 5 : 4;
        } else if (!strcmp(profile, "lc")) {
            o->radius = 2;
            o->bm_range = 9;
            o->ps_range = wiener ? 5 : 4;
        } else if (!strcmp(profile, "high")) {
            o->radius = 4;
            o->bm_range = 16;
            o->ps_range = wiener ? 8 : 7;
        } else if (!strcmp(profile, "vn")) {
            o->radius = 4;
            o->group_size = 16;
        }
    }
    o->process[0] = o->process[1] = o->process[2] = 1;
    return 0;
}

/* ColorMatrix_Parameter (include/Specification.h:71-136), long double as the
 * non-special branch of ColorMatrix_RGB2YUV_Parameter uses. */
static void or_kr_kg_kb(int matrix, long double *kr, long double *kg, long double *kb)
{
    double r, g, b;
    switch (matrix) {
    case 0: r = 0; g = 1; b = 0; break;
    case 1: r = 0.2126; g = 0.7152; b = 0.0722; break;
    case 4: r = 0.30; g = 0.59; b = 0.11; break;
    case 5: r = 0.299; g = 0.587; b = 0.114; break;
    case 6: r = 0.299; g = 0.587; b = 0.114; break;
    case 7: r = 0.212; g = 0.701; b = 0.087; break;
    case 8: r = 0.25; g = 0.50; b = 0.25; break;
    case 9: r = 0.2627; g = 0.6780; b = 0.0593; break;
    case 10: r = 0.2627; g = 0.6780; b = 0.0593; break;
    case 100: *kr = 1.0 / 3.0; *kg = 1.0 / 3.0; *kb = 1.0 / 3.0; return;
    default: r = 0.2126; g = 0.7152; b = 0.0722; break;
    }
    *kr = r;
    *kg = g;
    *kb = b;
}

int or_matrix_norms(int matrix, double norms[3])
{
    /* ColorMatrix_RGB2YUV_Parameter<double> (Specification.h:139-213), then
     * the three row norms of init_filter_data (BM3D_Base.cpp:262-272). */
    double Yr, Yg, Yb, Ur, Ug, Ub, Vr, Vg, Vb;
    if (matrix == 0) {
        Yr = 0; Yg = 1; Yb = 0; Ur = 0; Ug = 0; Ub = 1; Vr = 1; Vg = 0; Vb = 0;
    } else if (matrix == 8) {
        Yr = (double)(1.0L / 4.0L); Yg = (double)(1.0L / 2.0L); Yb = (double)(1.0L / 4.0L);
        Ur = (double)(-1.0L / 4.0L); Ug = (double)(1.0L / 2.0L); Ub = (double)(-1.0L / 4.0L);
        Vr = (double)(1.0L / 2.0L); Vg = 0; Vb = (double)(-1.0L / 2.0L);
    } else if (matrix == 100) {
        Yr = (double)(1.0L / 3.0L); Yg = (double)(1.0L / 3.0L); Yb = (double)(1.0L / 3.0L);
        Ur = (double)(1.0L / 2.0L); Ug = 0; Ub = (double)(-1.0L / 2.0L);
        Vr = (double)(1.0L / 4.0L); Vg = (double)(-1.0L / 2.0L); Vb = (double)(1.0L / 4.0L);
    } else {
        long double kr, kg, kb;
        or_kr_kg_kb(matrix, &kr, &kg, &kb);
        Yr = (double)kr; Yg = (double)kg; Yb = (double)kb;
        Ur = (double)(-(long double)Yr * 0.5L / (1.0L - (long double)Yb));
        Ug = (double)(-(long double)Yg * 0.5L / (1.0L - (long double)Yb));
        Ub = 0.5;
        Vr = 0.5;
        Vg = (double)(-(long double)Yg * 0.5L / (1.0L - (long double)Yr));
        Vb = (double)(-(long double)Yb * 0.5L / (1.0L - (long double)Yr));
    }
    norms[0] = sqrt(Yr * Yr + Yg * Yg + Yb * Yb);
    norms[1] = sqrt(Ur * Ur + Ug * Ug + Ub * Ub);
    norms[2] = sqrt(Vr * Vr + Vg * Vg + Vb * Vb);
    return 0;
}

int or_default_matrix(int width, int height)
{
    /* ResLevel_Default / ColorMatrix_Default (Specification.h:295-313) */
    if (width > 2048 || height > 1536) return 9;
    if (width > 1024 || height > 576) return 1;
    return 6;
}

void or_filter_init(or_filter *f, const or_params *p, int arith, int rcp_mode,
                    const float *rcp_table, int rcp_bits, int tie_mode)
{
    memset(f, 0, sizeof(*f));
    f->p = *p;
    f->arith = arith;
    f->rcp_mode = rcp_mode;
    f->rcp_table = rcp_table;
    f->rcp_bits = rcp_bits;
    f->tie_mode = tie_mode;
    const int B = p->block_size;
    /* BM3D_FilterData ctor, source/BM3D.cpp:137-209 */
    for (int n = 1; n <= p->group_size && n <= OR_MAX_GROUP; ++n) {
        int amp_i = 2 * n * 2 * B * 2 * B;
        f->final_amp[n - 1] = (double)amp_i;
        double forward_amp = sqrt(f->final_amp[n - 1]);
        for (int pl = 0; pl < 3; ++pl) {
            double sigma = p->sigma[pl];
            if (p->wiener) {
                float v = (float)(sigma * forward_amp * sigma * forward_amp);
                f->wiener_sigma_sqr[pl][n - 1] = (v > FLT_MIN) ? v : FLT_MIN;
            } else {
                double thr_base = sigma * p->lambda * forward_amp;
                f->thr[pl][n - 1][0] = (float)thr_base;
                f->thr[pl][n - 1][1] = (float)(thr_base * sqrt(2.0));
                f->thr[pl][n - 1][2] = (float)(thr_base * 2.0);
                f->thr[pl][n - 1][3] = (float)(thr_base * sqrt(8.0));
            }
        }
    }
}

/* ========================================================================== */
/* scan grid and search window                                                */
/* ========================================================================== */

int or_ref_grid(int dim, int block, int step, int *out)
{
    /* BM3D_Base.cpp:312-345: 0, s, 2s, ... and the last position clamped. */
    const int last = dim - block;
    int n = 0;
    for (int j = 0;; j += step) {
        if (j >= last + step) break;
        if (j > last) j = last;
        if (out) out[n] = j;
        ++n;
    }
    return n;
}

int or_search_boundary(int pos, int plane_boundary, int range, int step)
{
    /* Block::_SearchBoundary, include/Block.h:734-764 */
    range = range / step * step;
    int sb;
    if (pos == plane_boundary) {
        sb = plane_boundary;
    } else if (pos > plane_boundary) {
        sb = pos - range;
        while (sb < plane_boundary) sb += step;
    } else {
        sb = pos + range;
        while (sb > plane_boundary) sb -= step;
    }
    return sb;
}

/* ========================================================================== */
/* SSD (Block.h:575-643)                                                      */
/* ========================================================================== */

float or_ssd(const float *ref_block, const float *plane, ptrdiff_t stride, int y, int x, int B,
             int arith)
{
    const float *q = plane + (ptrdiff_t)y * stride + x;
    float dist = 0.0f;
    if (arith == OR_ARITH_SSE2) {
        const int simd_width = B - B % 8;
        if (simd_width > 0) {
            /* four lane accumulators; per row and 8-column chunk, lane k
             * takes column 8c+k and then column 8c+4+k */
            float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int r = 0; r < B; ++r) {
                for (int c = 0; c < simd_width; c += 8) {
                    for (int k = 0; k < 4; ++k) {
                        float d1 = ref_block[r * B + c + k] - q[r * stride + c + k];
                        float d1s = d1 * d1;
                        s[k] = s[k] + d1s;
                    }
                    for (int k = 0; k < 4; ++k) {
                        float d2 = ref_block[r * B + c + 4 + k] - q[r * stride + c + 4 + k];
                        float d2s = d2 * d2;
                        s[k] = s[k] + d2s;
                    }
                }
            }
            float t = ((s[0] + s[1]) + s[2]) + s[3];
            dist = dist + t;
        }
        if (B - simd_width > 0) {
            for (int r = 0; r < B; ++r) {
                for (int c = simd_width; c < B; ++c) {
                    float t = ref_block[r * B + c] - q[r * stride + c];
                    float t2 = t * t;
                    dist = dist + t2;
                }
            }
        }
    } else {
        for (int r = 0; r < B; ++r) {
            for (int c = 0; c < B; ++c) {
                float t = ref_block[r * B + c] - q[r * stride + c];
                float t2 = t * t;
                dist = dist + t2;
            }
        }
    }
    return dist;
}

/* ========================================================================== */
/* selection: libstdc++ heap-based std::partial_sort and stable sort          */
/* ========================================================================== */

/* The heap helpers follow libstdc++'s published <bits/stl_heap.h>
 * (__adjust_heap / __push_heap / __pop_heap / __make_heap) and
 * <bits/stl_algo.h> (__heap_select, __sort_heap), comparing keys only, as
 * KeyPair::operator< does (include/Type.h:454-457). */
typedef struct { float key; int pay; } or_kp;

static void or_push_heap(or_kp *first, ptrdiff_t hole, ptrdiff_t top, or_kp value)
{
    ptrdiff_t parent = (hole - 1) / 2;
    while (hole > top && first[parent].key < value.key) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

static void or_adjust_heap(or_kp *first, ptrdiff_t hole, ptrdiff_t len, or_kp value)
{
    const ptrdiff_t top = hole;
    ptrdiff_t second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (first[second].key < first[second - 1].key) second--;
        first[hole] = first[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole = second - 1;
    }
    or_push_heap(first, hole, top, value);
}

static void or_make_heap(or_kp *first, ptrdiff_t len)
{
    if (len < 2) return;
    ptrdiff_t parent = (len - 2) / 2;
    for (;;) {
        or_kp value = first[parent];
        or_adjust_heap(first, parent, len, value);
        if (parent == 0) return;
        parent--;
    }
}

static void or_pop_heap(or_kp *first, ptrdiff_t last, ptrdiff_t result)
{
    or_kp value = first[result];
    first[result] = first[0];
    or_adjust_heap(first, 0, last, value);
}

static void or_partial_sort_kp(or_kp *a, ptrdiff_t n, ptrdiff_t middle)
{
    or_make_heap(a, middle);
    for (ptrdiff_t i = middle; i < n; ++i)
        if (a[i].key < a[0].key) or_pop_heap(a, middle, i);
    for (ptrdiff_t last = middle; last > 1;) {
        --last;
        or_pop_heap(a, last, last);
    }
}

void or_partial_sort_keys(float *keys, int *payload, int n, int middle)
{
    or_kp *a = (or_kp *)malloc(sizeof(or_kp) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        a[i].key = keys[i];
        a[i].pay = payload[i];
    }
    or_partial_sort_kp(a, n, middle);
    for (int i = 0; i < n; ++i) {
        keys[i] = a[i].key;
        payload[i] = a[i].pay;
    }
    free(a);
}

/* Stable insertion/merge sort by key only (std::stable_sort semantics). */
static void or_stable_sort_kp(or_kp *a, ptrdiff_t n, or_kp *tmp)
{
    if (n < 2) return;
    if (n <= 16) {
        for (ptrdiff_t i = 1; i < n; ++i) {
            or_kp v = a[i];
            ptrdiff_t j = i;
            while (j > 0 && v.key < a[j - 1].key) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = v;
        }
        return;
    }
    ptrdiff_t h = n / 2;
    or_stable_sort_kp(a, h, tmp);
    or_stable_sort_kp(a + h, n - h, tmp);
    ptrdiff_t i = 0, j = h, k = 0;
    while (i < h && j < n) {
        if (a[j].key < a[i].key) tmp[k++] = a[j++];
        else tmp[k++] = a[i++];
    }
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, sizeof(or_kp) * n);
}

/* Keep the `limit` smallest of a[from..n) after a[0..from) (from = 1 when a
 * self entry is pinned first).  Returns the new length.  Follows
 * BlockMatchingMulti's "partial_sort if over the limit, else stable_sort"
 * rule (Block.h:663-676, 714-728). */
static ptrdiff_t or_select(or_kp *a, ptrdiff_t n, ptrdiff_t limit, ptrdiff_t from, int tie_mode,
                           int sorted_when_small, or_kp *tmp)
{
    if (limit > 0 && n > limit) {
        if (tie_mode == OR_TIES_LIBSTDCXX) {
            or_partial_sort_kp(a + from, n - from, limit - from);
        } else {
            /* total order (key, original position): a stable sort of the
             * tail followed by truncation */
            or_stable_sort_kp(a + from, n - from, tmp);
        }
        return limit;
    }
    if (sorted_when_small) or_stable_sort_kp(a, n, tmp);
    return n;
}

/* ========================================================================== */
/* block matching                                                             */
/* ========================================================================== */

typedef struct {
    int y, x;
} or_pos;

static void or_load_block(const float *plane, ptrdiff_t stride, int y, int x, int B, float *blk)
{
    for (int r = 0; r < B; ++r)
        for (int c = 0; c < B; ++c) blk[r * B + c] = plane[(ptrdiff_t)(y + r) * stride + x + c];
}

/* BlockMatchingMulti core (Block.h:554-653): append accepted (key, index)
 * for each candidate position. */
static ptrdiff_t or_match_core(const or_filter *f, const float *blk, const float *plane,
                               ptrdiff_t stride, const or_pos *cand, ptrdiff_t ncand, or_kp *out,
                               ptrdiff_t nout, int pay_base)
{
    const int B = f->p.block_size;
    const double mse2sse = (double)(B * B) * 1.0f * 1.0f / (double)(255 * 255);
    const double dist_mul = 1.0 / mse2sse;
    const float th_sse = (float)(f->p.th_mse * mse2sse);
    for (ptrdiff_t c = 0; c < ncand; ++c) {
        float dist = or_ssd(blk, plane, stride, cand[c].y, cand[c].x, B, f->arith);
        if (dist <= th_sse && dist != 0.0f) {
            out[nout].key = (float)((double)dist * dist_mul);
            out[nout].pay = pay_base + (int)c;
            ++nout;
        }
    }
    return nout;
}

static ptrdiff_t or_window(int py, int px, int W, int H, int B, int range, int step,
                           int exclude_y, int exclude_x, or_pos *out)
{
    range = range / step * step;
    const int l = or_search_boundary(px, 0, range, step);
    const int r = or_search_boundary(px, W - B, range, step);
    const int t = or_search_boundary(py, 0, range, step);
    const int b = or_search_boundary(py, H - B, range, step);
    ptrdiff_t n = 0;
    for (int j = t; j <= b; j += step)
        for (int i = l; i <= r; i += step) {
            if (j == exclude_y && i == exclude_x) continue;
            out[n].y = j;
            out[n].x = i;
            ++n;
        }
    return n;
}

int or_block_match(const or_filter *f, const float *plane, int W, int H, ptrdiff_t stride, int j,
                   int i, float *keys, int *ys, int *xs)
{
    const int B = f->p.block_size, G = f->p.group_size;
    /* BM3D_Process_Base::BlockMatching (BM3D_Base.cpp:473-490) */
    if (G == 1 || f->p.th_mse <= 0) {
        keys[0] = 0.0f;
        ys[0] = j;
        xs[0] = i;
        return 1;
    }
    const int R = f->p.bm_range, S = f->p.bm_step;
    const int side = 2 * (R / S) + 1;
    or_pos *cand = (or_pos *)malloc(sizeof(or_pos) * side * side);
    or_kp *list = (or_kp *)malloc(sizeof(or_kp) * (side * side + 1));
    or_kp *tmp = (or_kp *)malloc(sizeof(or_kp) * (side * side + 1));
    float blk[64 * 64];
    or_load_block(plane, stride, j, i, B, blk);
    ptrdiff_t nc = or_window(j, i, W, H, B, R, S, j, i, cand);
    list[0].key = 0.0f;
    list[0].pay = -1;
    ptrdiff_t n = or_match_core(f, blk, plane, stride, cand, nc, list, 1, 0);
    n = or_select(list, n, G, 1, f->tie_mode, 1, tmp);
    for (ptrdiff_t k = 0; k < n; ++k) {
        keys[k] = list[k].key;
        if (list[k].pay < 0) {
            ys[k] = j;
            xs[k] = i;
        } else {
            ys[k] = cand[list[k].pay].y;
            xs[k] = cand[list[k].pay].x;
        }
    }
    free(cand);
    free(list);
    free(tmp);
    return (int)n;
}

static int or_pos_cmp(const void *a, const void *b)
{
    const or_pos *p = (const or_pos *)a, *q = (const or_pos *)b;
    if (p->y != q->y) return p->y < q->y ? -1 : 1;
    if (p->x != q->x) return p->x < q->x ? -1 : 1;
    return 0;
}

/* Block::GenSearchPos (Block.h:771-827): the (y, x)-sorted union of the
 * +-range windows around each seed, duplicates removed. */
static ptrdiff_t or_gen_search_pos(const or_pos *seeds, int nseeds, int W, int H, int B, int range,
                                   int step, or_pos *out)
{
    ptrdiff_t n = 0;
    for (int s = 0; s < nseeds; ++s)
        n += or_window(seeds[s].y, seeds[s].x, W, H, B, range, step, -1 << 30, -1 << 30, out + n);
    qsort(out, n, sizeof(or_pos), or_pos_cmp);
    ptrdiff_t m = 0;
    for (ptrdiff_t k = 0; k < n; ++k)
        if (m == 0 || or_pos_cmp(&out[m - 1], &out[k]) != 0) out[m++] = out[k];
    return m;
}

int or_vblock_match(const or_filter *f, const float *const *planes, int nframes, int cur, int W,
                    int H, ptrdiff_t stride, int j, int i, float *keys, int *zs, int *ys, int *xs)
{
    const int B = f->p.block_size, G = f->p.group_size;
    if (G == 1 || f->p.th_mse <= 0) {
        keys[0] = 0.0f;
        zs[0] = cur;
        ys[0] = j;
        xs[0] = i;
        return 1;
    }
    const int R = f->p.bm_range, S = f->p.bm_step;
    const int side = 2 * (R / S) + 1;
    const int ps_side = 2 * (f->p.ps_range / f->p.ps_step) + 1;
    const ptrdiff_t ps_cap = (ptrdiff_t)ps_side * ps_side * (f->p.ps_num > 0 ? f->p.ps_num : 1);
    const ptrdiff_t cap = side * side + 1 > ps_cap ? side * side + 1 : ps_cap;
    or_pos *cand = (or_pos *)malloc(sizeof(or_pos) * cap);
    or_kp *list = (or_kp *)malloc(sizeof(or_kp) * cap);
    or_kp *tmp = (or_kp *)malloc(sizeof(or_kp) * cap);
    /* the concatenated match code: key, frame, y, x */
    const ptrdiff_t total_cap = (ptrdiff_t)G * nframes + 1;
    or_kp *all = (or_kp *)malloc(sizeof(or_kp) * total_cap);
    or_kp *all_tmp = (or_kp *)malloc(sizeof(or_kp) * total_cap);
    int *az = (int *)malloc(sizeof(int) * total_cap);
    or_pos *ap = (or_pos *)malloc(sizeof(or_pos) * total_cap);
    or_pos cur_seeds[OR_MAX_GROUP], seeds[OR_MAX_GROUP];
    ptrdiff_t nall = 0;
    float blk[64 * 64];
    or_load_block(planes[cur], stride, j, i, B, blk);

    /* current frame: full search with self first (VBM3D_Base.cpp:556-570) */
    ptrdiff_t nc = or_window(j, i, W, H, B, R, S, j, i, cand);
    list[0].key = 0.0f;
    list[0].pay = -1;
    ptrdiff_t n = or_match_core(f, blk, planes[cur], stride, cand, nc, list, 1, 0);
    n = or_select(list, n, G, 1, f->tie_mode, 1, tmp);
    for (ptrdiff_t k = 0; k < n; ++k) {
        all[nall].key = list[k].key;
        all[nall].pay = (int)nall;
        az[nall] = cur;
        ap[nall] = list[k].pay < 0 ? (or_pos){j, i} : cand[list[k].pay];
        ++nall;
    }
    int ncur = (int)(n < f->p.ps_num ? n : f->p.ps_num);
    for (int k = 0; k < ncur; ++k) cur_seeds[k] = ap[k];

    /* predictive search: backward then forward (VBM3D_Base.cpp:584-649) */
    for (int dir = -1; dir <= 1; dir += 2) {
        int nseeds = ncur;
        memcpy(seeds, cur_seeds, sizeof(or_pos) * ncur);
        for (int fr = cur + dir; fr >= 0 && fr < nframes; fr += dir) {
            nc = or_gen_search_pos(seeds, nseeds, W, H, B, f->p.ps_range, f->p.ps_step, cand);
            n = or_match_core(f, blk, planes[fr], stride, cand, nc, list, 0, 0);
            n = or_select(list, n, G, 0, f->tie_mode, 1, tmp);
            for (ptrdiff_t k = 0; k < n; ++k) {
                all[nall].key = list[k].key;
                all[nall].pay = (int)nall;
                az[nall] = fr;
                ap[nall] = cand[list[k].pay];
                ++nall;
            }
            nseeds = (int)(n < f->p.ps_num ? n : f->p.ps_num);
            for (int k = 0; k < nseeds; ++k) seeds[k] = cand[list[k].pay];
        }
    }
    /* final cut (VBM3D_Base.cpp:652-656): no sort when within the limit */
    nall = or_select(all, nall, G, 1, f->tie_mode, 0, all_tmp);
    for (ptrdiff_t k = 0; k < nall; ++k) {
        keys[k] = all[k].key;
        zs[k] = az[all[k].pay];
        ys[k] = ap[all[k].pay].y;
        xs[k] = ap[all[k].pay].x;
    }
    free(cand);
    free(list);
    free(tmp);
    free(all);
    free(all_tmp);
    free(az);
    free(ap);
    return (int)nall;
}

/* ========================================================================== */
/* transforms                                                                 */
/* ========================================================================== */

/* Coefficient 2cos(pi k (2j+1) / 2N) rounded once to float.  The oracle
 * evaluates REDFT10 / REDFT01 (FFTW's unnormalised DCT-II / DCT-III) with one
 * even/odd split:
 *   forward  X[2m]   = sum_j c(2m, j)   * (x[j] + x[N-1-j])   (+ middle term)
 *            X[2m+1] = sum_j c(2m+1, j) * (x[j] - x[N-1-j])
 *   inverse  E[j] = X[0] + sum_m c(2m, j) X[2m];  O[j] = sum_m c(2m+1, j) X[2m+1]
 *            x[j] = E[j] + O[j];  x[N-1-j] = E[j] - O[j]
 * Each sum is a chain of fused multiply-adds in increasing index order whose
 * first term is a plain product (forward) or X[0] (inverse even part).  The
 * GPU kernels evaluate the identical chains. */
static float or_dct_coef_eval(int N, int k, int j)
{
    return (float)(2.0 * cos(M_PI * (double)k * (double)(2 * j + 1) / (2.0 * (double)N)));
}

/* Tabulated coefficients for N <= OR_COEF_TAB (same values, evaluated once). */
#define OR_COEF_TAB 64
static float *or_coef_tab[OR_COEF_TAB + 1];
static pthread_once_t or_coef_once = PTHREAD_ONCE_INIT;

static void or_coef_init(void)
{
    for (int N = 1; N <= OR_COEF_TAB; ++N) {
        or_coef_tab[N] = (float *)malloc(sizeof(float) * N * N);
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j) or_coef_tab[N][k * N + j] = or_dct_coef_eval(N, k, j);
    }
}

static inline float or_dct_coef(int N, int k, int j)
{
    if (N <= OR_COEF_TAB) return or_coef_tab[N][k * N + j];
    return or_dct_coef_eval(N, k, j);
}

void or_dct2_1d(const float *x, float *X, int N)
{
    pthread_once(&or_coef_once, or_coef_init);
    const int h = N / 2;
    const int odd = N & 1;
    float u[OR_MAX_GROUP], v[OR_MAX_GROUP];
    for (int j = 0; j < h; ++j) {
        u[j] = x[j] + x[N - 1 - j];
        v[j] = x[j] - x[N - 1 - j];
    }
    for (int k = 0; k < N; ++k) {
        float acc;
        if ((k & 1) == 0) {
            if (h > 0) {
                acc = or_dct_coef(N, k, 0) * u[0];
                for (int j = 1; j < h; ++j) acc = fmaf(or_dct_coef(N, k, j), u[j], acc);
                if (odd) acc = fmaf(or_dct_coef(N, k, h), x[h], acc);
            } else {
                acc = or_dct_coef(N, k, 0) * x[0];
            }
        } else {
            acc = or_dct_coef(N, k, 0) * v[0];
            for (int j = 1; j < h; ++j) acc = fmaf(or_dct_coef(N, k, j), v[j], acc);
        }
        X[k] = acc;
    }
}

void or_dct3_1d(const float *X, float *x, int N)
{
    pthread_once(&or_coef_once, or_coef_init);
    const int h = N / 2;
    const int odd = N & 1;
    for (int j = 0; j < h + odd; ++j) {
        float e = X[0];
        for (int k = 2; k < N; k += 2) e = fmaf(or_dct_coef(N, k, j), X[k], e);
        float o = 0.0f;
        if (N > 1 && j < h) {
            o = or_dct_coef(N, 1, j) * X[1];
            for (int k = 3; k < N; k += 2) o = fmaf(or_dct_coef(N, k, j), X[k], o);
        }
        if (j < h) {
            x[j] = e + o;
            x[N - 1 - j] = e - o;
        } else {
            x[j] = e;
        }
    }
}

/* 3D transforms on g[z][y][x]: forward along x, then y, then z; inverse
 * along z, then y, then x. */
void or_fdct3d(float *g, int n, int B)
{
    float in[OR_MAX_GROUP], out[OR_MAX_GROUP];
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < B; ++y) {
            float *row = g + ((ptrdiff_t)z * B + y) * B;
            memcpy(in, row, sizeof(float) * B);
            or_dct2_1d(in, row, B);
        }
    for (int z = 0; z < n; ++z)
        for (int x = 0; x < B; ++x) {
            for (int y = 0; y < B; ++y) in[y] = g[((ptrdiff_t)z * B + y) * B + x];
            or_dct2_1d(in, out, B);
            for (int y = 0; y < B; ++y) g[((ptrdiff_t)z * B + y) * B + x] = out[y];
        }
    for (int yx = 0; yx < B * B; ++yx) {
        for (int z = 0; z < n; ++z) in[z] = g[(ptrdiff_t)z * B * B + yx];
        or_dct2_1d(in, out, n);
        for (int z = 0; z < n; ++z) g[(ptrdiff_t)z * B * B + yx] = out[z];
    }
}

void or_idct3d(float *g, int n, int B)
{
    float in[OR_MAX_GROUP], out[OR_MAX_GROUP];
    for (int yx = 0; yx < B * B; ++yx) {
        for (int z = 0; z < n; ++z) in[z] = g[(ptrdiff_t)z * B * B + yx];
        or_dct3_1d(in, out, n);
        for (int z = 0; z < n; ++z) g[(ptrdiff_t)z * B * B + yx] = out[z];
    }
    for (int z = 0; z < n; ++z)
        for (int x = 0; x < B; ++x) {
            for (int y = 0; y < B; ++y) in[y] = g[((ptrdiff_t)z * B + y) * B + x];
            or_dct3_1d(in, out, B);
            for (int y = 0; y < B; ++y) g[((ptrdiff_t)z * B + y) * B + x] = out[y];
        }
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < B; ++y) {
            float *row = g + ((ptrdiff_t)z * B + y) * B;
            memcpy(in, row, sizeof(float) * B);
            or_dct3_1d(in, row, B);
        }
}

/* ========================================================================== */
/* reciprocal                                                                 */
/* ========================================================================== */

float or_rcp(const or_filter *f, float x)
{
    if (f->rcp_mode == OR_RCP_TABLE) {
        uint32_t u;
        memcpy(&u, &x, 4);
        int e = (int)((u >> 23) & 0xff) - 127;
        uint32_t idx = (u & 0x7fffff) >> (23 - f->rcp_bits);
        float r = ldexpf(f->rcp_table[idx], -e);
        return (u >> 31) ? -r : r;
    }
#if defined(__SSE2__)
    __m128 v = _mm_set1_ps(x);
    v = _mm_rcp_ps(v);
    return _mm_cvtss_f32(v);
#else
    return 1.0f / x;
#endif
}

/* ========================================================================== */
/* collaborative filters                                                      */
/* ========================================================================== */

/* BM3D_Basic_Process::CollaborativeFilter (source/BM3D_Basic.cpp:72-148):
 * g holds the gathered group; on return g holds the inverse transform and
 * the weights are set. */
static void or_basic_group(const or_filter *f, int plane, float *g, int n, float *num_w,
                           float *den_w)
{
    const int B = f->p.block_size;
    const ptrdiff_t size = (ptrdiff_t)n * B * B;
    or_fdct3d(g, n, B);
    const float *t = f->thr[plane][n - 1];
    int retained = 0;
    /* |c| > thr keeps the coefficient (SSE2 :98-122 and scalar :124-134 agree
     * element-wise; the count is an integer sum) */
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                ptrdiff_t idx = ((ptrdiff_t)z * B + y) * B + x;
                int flag = (x == 0) + (y == 0) + (z == 0);
                float thr = t[flag];
                float s = g[idx];
                if (s > thr || s < -thr) ++retained;
                else g[idx] = 0.0f;
            }
    (void)size;
    or_idct3d(g, n, B);
    float dw = retained < 1 ? 1.0f : 1.0f / (float)retained;
    *den_w = dw;
    *num_w = (float)((double)dw / f->final_amp[n - 1]);
}

/* BM3D_Final_Process::CollaborativeFilter (source/BM3D_Final.cpp:50-127). */
static void or_final_group(const or_filter *f, int plane, float *gs, float *gr, int n,
                           float *num_w, float *den_w)
{
    const int B = f->p.block_size;
    const ptrdiff_t size = (ptrdiff_t)n * B * B;
    or_fdct3d(gs, n, B);
    or_fdct3d(gr, n, B);
    const float sig2 = f->wiener_sigma_sqr[plane][n - 1];
    float l2 = 0.0f;
    ptrdiff_t k0 = 0;
    if (f->arith == OR_ARITH_SSE2) {
        const ptrdiff_t simd_width = size - size % 4;
        float lane[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (ptrdiff_t k = 0; k < simd_width; ++k) {
            float r = gr[k];
            float r2 = r * r;
            float w = r2 * or_rcp(f, r2 + sig2);
            gs[k] = gs[k] * w;
            float w2 = w * w;
            lane[k & 3] = lane[k & 3] + w2;
        }
        float t = ((lane[0] + lane[1]) + lane[2]) + lane[3];
        l2 = l2 + t;
        k0 = simd_width;
    }
    for (ptrdiff_t k = k0; k < size; ++k) {
        float r = gr[k];
        float r2 = r * r;
        float w = r2 / (r2 + sig2);
        gs[k] = gs[k] * w;
        float w2 = w * w;
        l2 = l2 + w2;
    }
    or_idct3d(gs, n, B);
    l2 = (l2 > FLT_EPSILON) ? l2 : FLT_EPSILON;
    float dw = 1.0f / l2;
    *den_w = dw;
    *num_w = (float)((double)dw / f->final_amp[n - 1]);
}

/* BlockGroup::AddTo(dst, stride, gain) / CountTo(dst, stride, value)
 * (include/Block.h:1239-1259, 1326-1345). */
static void or_aggregate(float *num, float *den, ptrdiff_t stride, const float *g, int n, int B,
                         const int *ys, const int *xs, float num_w, float den_w)
{
    for (int z = 0; z < n; ++z) {
        const float *blk = g + (ptrdiff_t)z * B * B;
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                ptrdiff_t o = (ptrdiff_t)(ys[z] + y) * stride + xs[z] + x;
                float c = blk[y * B + x] * num_w;
                num[o] = num[o] + c;
            }
    }
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                ptrdiff_t o = (ptrdiff_t)(ys[z] + y) * stride + xs[z] + x;
                den[o] = den[o] + den_w;
            }
}

/* ========================================================================== */
/* frame kernels                                                              */
/* ========================================================================== */

int or_bm3d_kernel(const or_filter *f, int nplanes, int W, int H, ptrdiff_t stride,
                   float *const *dst, const float *const *src, const float *const *ref)
{
    const int B = f->p.block_size, G = f->p.group_size;
    const ptrdiff_t count = (ptrdiff_t)H * stride;
    float *den[3] = {NULL, NULL, NULL};
    for (int p = 0; p < nplanes; ++p) {
        if (!f->p.process[p]) continue;
        den[p] = (float *)calloc(count, sizeof(float));
        memset(dst[p], 0, sizeof(float) * count);
    }
    int *gy = (int *)malloc(sizeof(int) * (H + 1));
    int *gx = (int *)malloc(sizeof(int) * (W + 1));
    int ny = or_ref_grid(H, B, f->p.block_step, gy);
    int nx = or_ref_grid(W, B, f->p.block_step, gx);
    float keys[OR_MAX_GROUP];
    int ys[OR_MAX_GROUP], xs[OR_MAX_GROUP];
    float *gs = (float *)malloc(sizeof(float) * (size_t)G * B * B);
    float *gr = (float *)malloc(sizeof(float) * (size_t)G * B * B);
    for (int a = 0; a < ny; ++a)
        for (int b = 0; b < nx; ++b) {
            int n = or_block_match(f, ref[0], W, H, stride, gy[a], gx[b], keys, ys, xs);
            if (n > G) n = G;
            for (int p = 0; p < nplanes; ++p) {
                if (!f->p.process[p]) continue;
                for (int z = 0; z < n; ++z) or_load_block(src[p], stride, ys[z], xs[z], B, gs + (ptrdiff_t)z * B * B);
                float nw, dw;
                if (f->p.wiener) {
                    for (int z = 0; z < n; ++z)
                        or_load_block(ref[p], stride, ys[z], xs[z], B, gr + (ptrdiff_t)z * B * B);
                    or_final_group(f, p, gs, gr, n, &nw, &dw);
                } else {
                    or_basic_group(f, p, gs, n, &nw, &dw);
                }
                or_aggregate(dst[p], den[p], stride, gs, n, B, ys, xs, nw, dw);
            }
        }
    for (int p = 0; p < nplanes; ++p) {
        if (!f->p.process[p]) continue;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                ptrdiff_t o = (ptrdiff_t)y * stride + x;
                dst[p][o] = dst[p][o] / den[p][o];
            }
        free(den[p]);
    }
    free(gy);
    free(gx);
    free(gs);
    free(gr);
    return 0;
}

/* or_bm3d_kernel with the matching and collaborative filtering of each band of
 * reference-block rows spread over nthreads threads; the aggregation stays
 * serial in the reference's scan order (BM3D_Base.cpp:309-351), so the
 * result is identical to or_bm3d_kernel's.  Test infrastructure: lets the
 * full-size parity tests check 1080p frames in seconds. */
typedef struct or_band_job {
    const or_filter *f;
    int nplanes, W, H;
    ptrdiff_t stride;
    const float *const *src, *const *ref;
    const int *gy, *gx;
    int nx, a0, a1;
    int next;            /* next row to take (atomic) */
    int *n;              /* [rows * nx] group sizes */
    int *ys, *xs;        /* [rows * nx * G] */
    float *grp;          /* [rows * nx][plane][G * B * B] filtered groups */
    float *nw, *dw;      /* [rows * nx][plane] */
} or_band_job;

static void *or_band_worker(void *arg)
{
    or_band_job *j = (or_band_job *)arg;
    const or_filter *f = j->f;
    const int B = f->p.block_size, G = f->p.group_size;
    const ptrdiff_t gsz = (ptrdiff_t)G * B * B;
    float keys[OR_MAX_GROUP];
    float *gr = (float *)malloc(sizeof(float) * gsz);
    for (;;) {
        const int a = __sync_fetch_and_add(&j->next, 1) + j->a0;
        if (a >= j->a1) break;
        for (int b = 0; b < j->nx; ++b) {
            const ptrdiff_t k = (ptrdiff_t)(a - j->a0) * j->nx + b;
            int *ys = j->ys + k * G, *xs = j->xs + k * G;
            int n = or_block_match(f, j->ref[0], j->W, j->H, j->stride, j->gy[a], j->gx[b], keys, ys, xs);
            if (n > G) n = G;
            j->n[k] = n;
            for (int p = 0; p < j->nplanes; ++p) {
                if (!f->p.process[p]) continue;
                float *gs = j->grp + (k * 3 + p) * gsz;
                for (int z = 0; z < n; ++z)
                    or_load_block(j->src[p], j->stride, ys[z], xs[z], B, gs + (ptrdiff_t)z * B * B);
                if (f->p.wiener) {
                    for (int z = 0; z < n; ++z)
                        or_load_block(j->ref[p], j->stride, ys[z], xs[z], B, gr + (ptrdiff_t)z * B * B);
                    or_final_group(f, p, gs, gr, n, &j->nw[k * 3 + p], &j->dw[k * 3 + p]);
                } else {
                    or_basic_group(f, p, gs, n, &j->nw[k * 3 + p], &j->dw[k * 3 + p]);
                }
            }
        }
    }
    free(gr);
    return NULL;
}

int or_bm3d_kernel_mt(const or_filter *f, int nplanes, int W, int H, ptrdiff_t stride,
                      float *const *dst, const float *const *src, const float *const *ref, int nthreads)
{
    const int B = f->p.block_size, G = f->p.group_size;
    const ptrdiff_t count = (ptrdiff_t)H * stride, gsz = (ptrdiff_t)G * B * B;
    if (nthreads < 1) nthreads = 1;
    float *den[3] = {NULL, NULL, NULL};
    for (int p = 0; p < nplanes; ++p) {
        if (!f->p.process[p]) continue;
        den[p] = (float *)calloc(count, sizeof(float));
        memset(dst[p], 0, sizeof(float) * count);
    }
    int *gy = (int *)malloc(sizeof(int) * (H + 1));
    int *gx = (int *)malloc(sizeof(int) * (W + 1));
    int ny = or_ref_grid(H, B, f->p.block_step, gy);
    int nx = or_ref_grid(W, B, f->p.block_step, gx);
    const int band = 2 * nthreads;
    or_band_job j;
    memset(&j, 0, sizeof(j));
    j.f = f;
    j.nplanes = nplanes;
    j.W = W;
    j.H = H;
    j.stride = stride;
    j.src = src;
    j.ref = ref;
    j.gy = gy;
    j.gx = gx;
    j.nx = nx;
    const ptrdiff_t nblk = (ptrdiff_t)band * nx;
    j.n = (int *)malloc(sizeof(int) * nblk);
    j.ys = (int *)malloc(sizeof(int) * nblk * G);
    j.xs = (int *)malloc(sizeof(int) * nblk * G);
    j.grp = (float *)malloc(sizeof(float) * nblk * 3 * gsz);
    j.nw = (float *)malloc(sizeof(float) * nblk * 3);
    j.dw = (float *)malloc(sizeof(float) * nblk * 3);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int a0 = 0; a0 < ny; a0 += band) {
        j.a0 = a0;
        j.a1 = a0 + band < ny ? a0 + band : ny;
        j.next = 0;
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, or_band_worker, &j);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
        for (ptrdiff_t k = 0; k < (ptrdiff_t)(j.a1 - a0) * nx; ++k)
            for (int p = 0; p < nplanes; ++p) {
                if (!f->p.process[p]) continue;
                or_aggregate(dst[p], den[p], stride, j.grp + (k * 3 + p) * gsz, j.n[k], B, j.ys + k * G,
                             j.xs + k * G, j.nw[k * 3 + p], j.dw[k * 3 + p]);
            }
    }
    for (int p = 0; p < nplanes; ++p) {
        if (!f->p.process[p]) continue;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                ptrdiff_t o = (ptrdiff_t)y * stride + x;
                dst[p][o] = dst[p][o] / den[p][o];
            }
        free(den[p]);
    }
    free(th);
    free(j.n);
    free(j.ys);
    free(j.xs);
    free(j.grp);
    free(j.nw);
    free(j.dw);
    free(gy);
    free(gx);
    return 0;
}

int or_vbm3d_kernel(const or_filter *f, int nplanes, int nframes, int cur, int W, int H,
                    ptrdiff_t stride, float *const *dst, const float *const *src,
                    const float *const *ref)
{
    const int B = f->p.block_size, G = f->p.group_size, r = f->p.radius;
    const ptrdiff_t slot = (ptrdiff_t)H * stride;
    const int base_slot = 2 * (r - cur); /* 2 * (radius + b_offset) */
    for (int p = 0; p < nplanes; ++p)
        if (f->p.process[p]) memset(dst[p] + base_slot * slot, 0, sizeof(float) * slot * 2 * nframes);
    int *gy = (int *)malloc(sizeof(int) * (H + 1));
    int *gx = (int *)malloc(sizeof(int) * (W + 1));
    int ny = or_ref_grid(H, B, f->p.block_step, gy);
    int nx = or_ref_grid(W, B, f->p.block_step, gx);
    float keys[OR_MAX_GROUP];
    int zs[OR_MAX_GROUP], ys[OR_MAX_GROUP], xs[OR_MAX_GROUP];
    float *gs = (float *)malloc(sizeof(float) * (size_t)G * B * B);
    float *gr = (float *)malloc(sizeof(float) * (size_t)G * B * B);
    const float **refy = (const float **)malloc(sizeof(float *) * nframes);
    for (int k = 0; k < nframes; ++k) refy[k] = ref[k];
    for (int a = 0; a < ny; ++a)
        for (int b = 0; b < nx; ++b) {
            int n = or_vblock_match(f, refy, nframes, cur, W, H, stride, gy[a], gx[b], keys, zs, ys, xs);
            for (int p = 0; p < nplanes; ++p) {
                if (!f->p.process[p]) continue;
                const float *const *sp = src + (ptrdiff_t)p * nframes;
                const float *const *rp = ref + (ptrdiff_t)p * nframes;
                for (int z = 0; z < n; ++z)
                    or_load_block(sp[zs[z]], stride, ys[z], xs[z], B, gs + (ptrdiff_t)z * B * B);
                float nw, dw;
                if (f->p.wiener) {
                    for (int z = 0; z < n; ++z)
                        or_load_block(rp[zs[z]], stride, ys[z], xs[z], B, gr + (ptrdiff_t)z * B * B);
                    or_final_group(f, p, gs, gr, n, &nw, &dw);
                } else {
                    or_basic_group(f, p, gs, n, &nw, &dw);
                }
                /* Pos3 AddTo / CountTo (Block.h:1283-1303, 1368-1387):
                 * member z goes to the num/den slots of its frame */
                for (int z = 0; z < n; ++z) {
                    float *num = dst[p] + (ptrdiff_t)(base_slot + 2 * zs[z]) * slot;
                    const float *blk = gs + (ptrdiff_t)z * B * B;
                    for (int y = 0; y < B; ++y)
                        for (int x = 0; x < B; ++x) {
                            ptrdiff_t o = (ptrdiff_t)(ys[z] + y) * stride + xs[z] + x;
                            float c = blk[y * B + x] * nw;
                            num[o] = num[o] + c;
                        }
                }
                for (int z = 0; z < n; ++z) {
                    float *den = dst[p] + (ptrdiff_t)(base_slot + 2 * zs[z] + 1) * slot;
                    for (int y = 0; y < B; ++y)
                        for (int x = 0; x < B; ++x) {
                            ptrdiff_t o = (ptrdiff_t)(ys[z] + y) * stride + xs[z] + x;
                            den[o] = den[o] + dw;
                        }
                }
            }
        }
    free(gy);
    free(gx);
    free(gs);
    free(gr);
    free(refy);
    return 0;
}

void or_vaggregate_plane(int W, int H, ptrdiff_t stride, int radius, int nframes, int cur,
                         const float *const *stacks, float *dst, ptrdiff_t dst_stride)
{
    const ptrdiff_t slot = (ptrdiff_t)H * stride;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float num = 0.0f, den = 0.0f;
            for (int i = 0; i < nframes; ++i) {
                int o = radius + cur - i; /* radius - (m - n) */
                num = num + stacks[i][(ptrdiff_t)(2 * o) * slot + (ptrdiff_t)y * stride + x];
                den = den + stacks[i][(ptrdiff_t)(2 * o + 1) * slot + (ptrdiff_t)y * stride + x];
            }
            dst[(ptrdiff_t)y * dst_stride + x] = num / den;
        }
}

/* ========================================================================== */
/* conversions (include/Conversion.hpp, include/Helper.h:285-377)             */
/* ========================================================================== */

static inline float or_clip(float v, float lo, float hi)
{
    return v <= lo ? lo : (v >= hi ? hi : v);
}

static inline float or_load_sample(const void *p, int bits, ptrdiff_t idx)
{
    if (bits == 0) return ((const float *)p)[idx];
    if (bits <= 8) return (float)((const uint8_t *)p)[idx];
    return (float)((const uint16_t *)p)[idx];
}

static inline void or_store_sample(void *p, int bits, ptrdiff_t idx, float v)
{
    if (bits == 0) ((float *)p)[idx] = v;
    else if (bits <= 8) ((uint8_t *)p)[idx] = (uint8_t)v;
    else ((uint16_t *)p)[idx] = (uint16_t)v;
}

/* GetQuanPara for an integer format (Helper.h:285-346). */
static void or_quan_int(int bits, int full, int chroma, int *floor_, int *neutral, int *ceil_)
{
    if (full) {
        *floor_ = 0;
        *neutral = chroma ? 1 << (bits - 1) : 0;
        *ceil_ = (1 << bits) - 1;
    } else {
        *floor_ = 16 << (bits - 8);
        *neutral = chroma ? 1 << (bits - 1) : *floor_;
        *ceil_ = chroma ? 240 << (bits - 8) : 235 << (bits - 8);
    }
}

void or_int2float(const void *src, int bits, ptrdiff_t src_stride, float *dst, ptrdiff_t dst_stride,
                  int W, int H, int chroma, int full, int clip)
{
    int sf, sn, sc;
    or_quan_int(bits, full, chroma, &sf, &sn, &sc);
    const float df = chroma ? -0.5f : 0.0f, dn = chroma ? 0.0f : df, dc = chroma ? 0.5f : 1.0f;
    const int srange = sc - sf;
    const float drange = dc - df;
    if (sf < sn && (sf + sc) % 2 == 1) clip = 1; /* PC-range chroma source */
    const float gain = drange / (float)srange;
    const float offset = -(float)sn * gain + dn;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float v = or_load_sample(src, bits, (ptrdiff_t)y * src_stride + x) * gain + offset;
            dst[(ptrdiff_t)y * dst_stride + x] = clip ? or_clip(v, df, dc) : v;
        }
}

void or_float2int(const float *src, ptrdiff_t src_stride, void *dst, int bits, ptrdiff_t dst_stride,
                  int W, int H, int chroma, int full, int clip)
{
    int df, dn, dc;
    or_quan_int(bits, full, chroma, &df, &dn, &dc);
    const float sf = chroma ? -0.5f : 0.0f, sn = chroma ? 0.0f : sf, sc = chroma ? 0.5f : 1.0f;
    const float srange = sc - sf;
    const int drange = dc - df;
    const int dst_pc = df < dn && (df + dc) % 2 == 1;
    const float gain = (float)drange / srange;
    float offset = -sn * gain + (float)dn;
    offset += dst_pc ? (float)0.499999 : 0.5f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float v = src[(ptrdiff_t)y * src_stride + x] * gain + offset;
            if (clip) v = or_clip(v, (float)df, (float)dc);
            or_store_sample(dst, bits, (ptrdiff_t)y * dst_stride + x, v);
        }
}

/* Source range (GetQuanPara 2-argument form, full range). */
static void or_src_range(int bits, float *sfloor, float *srange_f, int *is_int, int *srange_i)
{
    if (bits == 0) {
        *sfloor = 0.0f;
        *srange_f = 1.0f;
        *is_int = 0;
        *srange_i = 0;
    } else {
        *sfloor = 0.0f;
        *srange_i = (1 << bits) - 1;
        *srange_f = (float)*srange_i;
        *is_int = 1;
    }
}

void or_rgb2opp(const void *const *src, int src_bits, ptrdiff_t src_stride, void *const *dst,
                int dst_bits, ptrdiff_t dst_stride, int W, int H, int clip)
{
    float sf, sr;
    int si, sri;
    or_src_range(src_bits, &sf, &sr, &si, &sri);
    const int dst_float = dst_bits == 0;
    float dfy, dcy, dfc, dnc, dcc;
    int dst_pc = 0;
    if (dst_float) {
        dfy = 0.0f; dcy = 1.0f; dfc = -0.5f; dnc = 0.0f; dcc = 0.5f;
    } else {
        int a, b, c;
        or_quan_int(dst_bits, 1, 0, &a, &b, &c);
        dfy = (float)a; dcy = (float)c;
        or_quan_int(dst_bits, 1, 1, &a, &b, &c);
        dfc = (float)a; dnc = (float)b; dcc = (float)c;
        dst_pc = a < b && (a + c) % 2 == 1;
    }
    /* dRange computed in the destination type, then converted to float */
    const float drange_y = dcy - dfy, drange_c = dcc - dfc;
    const float gain_y = drange_y / (sr * 3.0f);
    float offset_y = -sf * 3.0f * gain_y + dfy;
    if (!dst_float) offset_y += 0.5f;
    const float gain_u = drange_c / (sr * 2.0f);
    const float gain_v = drange_c / (sr * 4.0f);
    float offset_c = dnc;
    if (!dst_float) offset_c += dst_pc ? (float)0.499999 : 0.5f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            ptrdiff_t si_ = (ptrdiff_t)y * src_stride + x, di = (ptrdiff_t)y * dst_stride + x;
            float R = or_load_sample(src[0], src_bits, si_);
            float G = or_load_sample(src[1], src_bits, si_);
            float B = or_load_sample(src[2], src_bits, si_);
            float t = ((R + G) + B) * gain_y + offset_y;
            or_store_sample(dst[0], dst_bits, di, clip ? or_clip(t, dfy, dcy) : t);
            t = (R - B) * gain_u + offset_c;
            or_store_sample(dst[1], dst_bits, di, clip ? or_clip(t, dfc, dcc) : t);
            t = ((R - G * 2.0f) + B) * gain_v + offset_c;
            or_store_sample(dst[2], dst_bits, di, clip ? or_clip(t, dfc, dcc) : t);
        }
}

void or_rgb2opp_y(const void *const *src, int src_bits, ptrdiff_t src_stride, float *dst,
                  ptrdiff_t dst_stride, int W, int H)
{
    float sf, sr;
    int si, sri;
    or_src_range(src_bits, &sf, &sr, &si, &sri);
    const float gain = 1.0f / (sr * 3.0f);
    const float offset = -sf * 3.0f * gain + 0.0f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            ptrdiff_t si_ = (ptrdiff_t)y * src_stride + x;
            float R = or_load_sample(src[0], src_bits, si_);
            float G = or_load_sample(src[1], src_bits, si_);
            float B = or_load_sample(src[2], src_bits, si_);
            dst[(ptrdiff_t)y * dst_stride + x] = ((R + G) + B) * gain + offset;
        }
}

void or_opp2rgb(const void *const *src, int src_bits, ptrdiff_t src_stride, void *const *dst,
                int dst_bits, ptrdiff_t dst_stride, int W, int H, int clip)
{
    const int dst_float = dst_bits == 0;
    float sfy, scy, sfc, snc, scc;
    if (src_bits == 0) {
        sfy = 0.0f; scy = 1.0f; sfc = -0.5f; snc = 0.0f; scc = 0.5f;
    } else {
        int a, b, c;
        or_quan_int(src_bits, 1, 0, &a, &b, &c);
        sfy = (float)a; scy = (float)c;
        or_quan_int(src_bits, 1, 1, &a, &b, &c);
        sfc = (float)a; snc = (float)b; scc = (float)c;
    }
    float dfl, dcl;
    if (dst_float) {
        dfl = 0.0f; dcl = 1.0f;
    } else {
        dfl = 0.0f; dcl = (float)((1 << dst_bits) - 1);
    }
    const float srange_y = scy - sfy, srange_c = scc - sfc;
    const float drange = dcl - dfl;
    const float gain_y = drange / srange_y, gain_c = drange / srange_c;
    /* ColorMatrix_YUV2RGB_Parameter OPP (Specification.h:253-269) */
    float Ry = 1.0f, Ru = 1.0f, Rv = (float)(2.0L / 3.0L);
    float Gy = 1.0f, Gu = 0.0f, Gv = (float)(-4.0L / 3.0L);
    float By = 1.0f, Bu = -1.0f, Bv = (float)(2.0L / 3.0L);
    Ry *= gain_y; Ru *= gain_c; Rv *= gain_c;
    Gy *= gain_y; Gu *= gain_c; Gv *= gain_c;
    By *= gain_y; Bu *= gain_c; Bv *= gain_c;
    float off_r = -sfy * Ry - snc * (Ru + Rv) + dfl;
    float off_g = -sfy * Gy - snc * (Gu + Gv) + dfl;
    float off_b = -sfy * By - snc * (Bu + Bv) + dfl;
    if (!dst_float) {
        off_r += 0.5f;
        off_g += 0.5f;
        off_b += 0.5f;
    }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            ptrdiff_t si_ = (ptrdiff_t)y * src_stride + x, di = (ptrdiff_t)y * dst_stride + x;
            float Y = or_load_sample(src[0], src_bits, si_);
            float U = or_load_sample(src[1], src_bits, si_);
            float V = or_load_sample(src[2], src_bits, si_);
            float t = Ry * Y + Ru * U + Rv * V + off_r;
            or_store_sample(dst[0], dst_bits, di, clip ? or_clip(t, dfl, dcl) : t);
            t = Gy * Y + Gv * V + off_g;
            or_store_sample(dst[1], dst_bits, di, clip ? or_clip(t, dfl, dcl) : t);
            t = By * Y + Bu * U + Bv * V + off_b;
            or_store_sample(dst[2], dst_bits, di, clip ? or_clip(t, dfl, dcl) : t);
        }
}

/* ========================================================================== */
/* CPU baseline runner                                                        */
/* ========================================================================== */

typedef struct {
    const or_filter *basic, *final_;
    int W, H;
    const uint16_t *const *in;
    uint16_t *const *out;
    int frame;
} or_job;

/* One frame of RGB2OPP(sample=0) -> Basic(matrix=100) -> Final(matrix=100)
 * -> OPP2RGB(sample=0), each stage exactly as the plugin's process_core_yuv
 * does for YUV444P16 (BM3D_Base.cpp:563-640). */
static void or_chain_one(const or_filter *basic, const or_filter *final_, int W, int H,
                         const uint16_t *const *rgb, uint16_t *const *out)
{
    const ptrdiff_t n = (ptrdiff_t)W * H;
    uint16_t *opp[3], *bas[3], *fin[3];
    float *s[3], *r[3], *d[3];
    for (int p = 0; p < 3; ++p) {
        opp[p] = (uint16_t *)malloc(n * 2);
        bas[p] = (uint16_t *)malloc(n * 2);
        fin[p] = (uint16_t *)malloc(n * 2);
        s[p] = (float *)malloc(n * 4);
        r[p] = (float *)malloc(n * 4);
        d[p] = (float *)malloc(n * 4);
    }
    or_rgb2opp((const void *const *)rgb, 16, W, (void *const *)opp, 16, W, W, H, 1);
    for (int p = 0; p < 3; ++p) or_int2float(opp[p], 16, W, s[p], W, W, H, p > 0, 1, 0);
    or_bm3d_kernel(basic, 3, W, H, W, d, (const float *const *)s, (const float *const *)s);
    for (int p = 0; p < 3; ++p) or_float2int(d[p], W, bas[p], 16, W, W, H, p > 0, 1, 1);
    for (int p = 0; p < 3; ++p) or_int2float(bas[p], 16, W, r[p], W, W, H, p > 0, 1, 0);
    or_bm3d_kernel(final_, 3, W, H, W, d, (const float *const *)s, (const float *const *)r);
    for (int p = 0; p < 3; ++p) or_float2int(d[p], W, fin[p], 16, W, W, H, p > 0, 1, 1);
    or_opp2rgb((const void *const *)fin, 16, W, (void *const *)out, 16, W, W, H, 1);
    for (int p = 0; p < 3; ++p) {
        free(opp[p]);
        free(bas[p]);
        free(fin[p]);
        free(s[p]);
        free(r[p]);
        free(d[p]);
    }
}

typedef struct {
    const or_filter *basic, *final_;
    int W, H, nframes, nthreads, tid;
    const uint16_t *const *in;
    uint16_t *const *out;
} or_worker;

static void *or_worker_main(void *arg)
{
    or_worker *w = (or_worker *)arg;
    for (int fr = w->tid; fr < w->nframes; fr += w->nthreads)
        or_chain_one(w->basic, w->final_, w->W, w->H, w->in + 3 * fr, w->out + 3 * fr);
    return NULL;
}

int or_chain_rgb48(const or_filter *basic, const or_filter *final_, int nframes, int nthreads,
                   int W, int H, const uint16_t *const *rgb_in, uint16_t *const *rgb_out)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    or_worker *ws = (or_worker *)malloc(sizeof(or_worker) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        ws[t] = (or_worker){basic, final_, W, H, nframes, nthreads, t, rgb_in, rgb_out};
        pthread_create(&th[t], NULL, or_worker_main, &ws[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(ws);
    return 0;
}
