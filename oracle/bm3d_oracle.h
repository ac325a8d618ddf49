/*
 * bm3d_oracle.h -- CPU restatement of the VapourSynth-BM3D filtering path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never calls into it.
 *
 * Every function restates one piece of the reference plugin's arithmetic
 * (file:line citations below refer to the reference repository).  Where the
 * reference delegates to FFTW (r2r REDFT10/REDFT01), the oracle evaluates the
 * published transform definition with a fixed, documented operation order;
 * FFTW's own float rounding is not reproducible (timing-chosen plans) and is
 * therefore "parity unpinned" at that boundary.
 *
 * Pins (see tests/): closed forms (DCT definition vs float64, REDFT01 o
 * REDFT10 = 2N, threshold/Wiener constants vs hand-computed values), hand-
 * derived SSD lane-order cases, brute-force float64 BM3D on tiny frames
 * (scipy DCT, python sort), libstdc++ heap-select semantics on crafted key
 * sequences, and the invariants listed with each function.  The V-BM3D
 * functions are pinned in tests/test_oracle_vbm3d_pins.py: or_vblock_match
 * against a float64 restatement of VBM3D_Base.cpp:540-659 / Block.h:655-827
 * and against brute force with whole-frame predictive windows;
 * or_vbm3d_kernel's stack layout (absent-neighbour slots untouched, group
 * size 1 = the single-frame Basic); or_vaggregate_plane on tagged stacks
 * with one exact answer; the conversions' rounding offsets at hand-derived
 * boundary samples.
 */
#ifndef BM3D_ORACLE_H
#define BM3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_GROUP 256

/* Arithmetic flavour of the reference build (include/Helper.h:45-59). */
enum { OR_ARITH_SSE2 = 0, OR_ARITH_SCALAR = 1 };
/* Reciprocal used by the Wiener gain in the SSE2 path (BM3D_Final.cpp:94). */
enum { OR_RCP_NATIVE = 0, OR_RCP_TABLE = 1 };
/* How equal keys are ordered by the selection step.
 * LIBSTDCXX: exact libstdc++ std::partial_sort / std::stable_sort behaviour
 *            (Block.h:663-676, 714-728; VBM3D_Base.cpp:652-656).
 * INDEX:     total order on (key, search index) -- the order the GPU uses. */
enum { OR_TIES_LIBSTDCXX = 0, OR_TIES_INDEX = 1 };

/* Resolved filter parameters (what BM3D_Data_Base::arguments_process and
 * init_filter_data leave behind, source/BM3D_Base.cpp:32-283). */
typedef struct or_params {
    int wiener;                 /* 0 = Basic (hard threshold), 1 = Final (Wiener) */
    int block_size, block_step, group_size, bm_range, bm_step;
    double th_mse;              /* already multiplied by normY (BM3D_Base.cpp:274) */
    double lambda;              /* hard_thr */
    double sigma[3];            /* sigma/255 * norm{Y,U,V} (BM3D_Base.cpp:277-282) */
    int process[3];
    int radius, ps_num, ps_range, ps_step;   /* V-BM3D (VBM3D_Base.h:35-47) */
} or_params;

/* Per-plane precomputed data (source/BM3D.cpp:137-209). */
typedef struct or_filter {
    or_params p;
    int arith, rcp_mode, tie_mode;
    int rcp_bits;                   /* table mode: mantissa bits used as index */
    const float *rcp_table;         /* table mode: rcp of 1.m for each index */
    double final_amp[OR_MAX_GROUP]; /* finalAMP[n-1] = 2n * 2B * 2B */
    float thr[3][OR_MAX_GROUP][4];  /* hard threshold by n and #zero-frequency axes */
    float wiener_sigma_sqr[3][OR_MAX_GROUP];
} or_filter;

/* ---- parameters ---------------------------------------------------------- */
/* Profile defaults, BM3D_Para ctor + thMSE_Default (source/BM3D.cpp:33-130)
 * and VBM3D_Para (source/VBM3D_Base.cpp:32-97).  profile: fast lc np high vn.
 * Returns 0 or -1 for an unknown profile. */
int or_profile_defaults(const char *profile, int wiener, int temporal, or_params *out,
                        double sigma0);
/* Norms of the RGB->YUV rows, ColorMatrix_RGB2YUV_Parameter
 * (include/Specification.h:139-213) evaluated as init_filter_data does. */
int or_matrix_norms(int matrix, double norms[3]);
/* ColorMatrix_Default (include/Specification.h:305-313). */
int or_default_matrix(int width, int height);
void or_filter_init(or_filter *f, const or_params *p, int arith, int rcp_mode,
                    const float *rcp_table, int rcp_bits, int tie_mode);

/* ---- primitives ---------------------------------------------------------- */
/* Reference-block grid along one axis (BM3D_Base.cpp:312-345). */
int or_ref_grid(int dim, int block, int step, int *out);
/* Block::_SearchBoundary (include/Block.h:734-764). */
int or_search_boundary(int pos, int plane_boundary, int range, int step);
/* SSD of a BxB block, SSE2 lane order or scalar order (Block.h:575-643). */
float or_ssd(const float *ref_block, const float *plane, ptrdiff_t stride, int y, int x, int B,
             int arith);
/* Full-search block matching, Block::BlockMatchingMulti(range) (Block.h:683-729).
 * Writes up to G (key, y, x) entries, self first.  Returns the count. */
int or_block_match(const or_filter *f, const float *plane, int W, int H, ptrdiff_t stride, int j,
                   int i, float *keys, int *ys, int *xs);
/* V-BM3D block matching (VBM3D_Base.cpp:540-659).  frames: [nframes] luma
 * planes, cur index.  Writes up to G (key, z, y, x).  Returns the count. */
int or_vblock_match(const or_filter *f, const float *const *planes, int nframes, int cur, int W,
                    int H, ptrdiff_t stride, int j, int i, float *keys, int *zs, int *ys, int *xs);
/* libstdc++ std::partial_sort on keys (permuting payload indices). */
void or_partial_sort_keys(float *keys, int *payload, int n, int middle);

/* Unnormalised DCT-II / DCT-III of length N with the oracle's fixed order. */
void or_dct2_1d(const float *x, float *X, int N);
void or_dct3_1d(const float *X, float *x, int N);
/* In-place 3D transforms on group[z][y][x] (n, B, B). */
void or_fdct3d(float *g, int n, int B);
void or_idct3d(float *g, int n, int B);
/* rcp as used by the Wiener gain (native _mm_rcp_ps or table). */
float or_rcp(const or_filter *f, float x);

/* ---- frame kernels ------------------------------------------------------- */
/* BM3D_Process_Base::Kernel (BM3D_Base.cpp:290-470) on float planes.
 * nplanes 1 (gray) or 3; matching on ref[0].  dst planes are written where
 * process[p] is set.  All planes W x H with a common stride (elements). */
int or_bm3d_kernel(const or_filter *f, int nplanes, int W, int H, ptrdiff_t stride,
                   float *const *dst, const float *const *src, const float *const *ref);
/* The same result with the matching and filtering of each band of block
 * rows spread over nthreads threads (aggregation stays serial, in order). */
int or_bm3d_kernel_mt(const or_filter *f, int nplanes, int W, int H, ptrdiff_t stride,
                      float *const *dst, const float *const *src, const float *const *ref, int nthreads);
/* VBM3D_Process_Base::Kernel (VBM3D_Base.cpp:410-537).  src/ref:
 * [plane * nframes + frame].  dst: [plane] -> base of the (2r+1)*2 slot stack
 * of this output frame (slot s at dst + s*H*stride).  cur = index of the
 * output frame among nframes; b_offset = -cur. */
int or_vbm3d_kernel(const or_filter *f, int nplanes, int nframes, int cur, int W, int H,
                    ptrdiff_t stride, float *const *dst, const float *const *src,
                    const float *const *ref);
/* VAggregate_Process::Kernel (source/VAggregate.cpp:93-160) for one plane.
 * stacks: [frames] pointers to each neighbour's stack base (ascending frame
 * order), slot index radius - (m - n). */
void or_vaggregate_plane(int W, int H, ptrdiff_t stride, int radius, int nframes, int cur,
                         const float *const *stacks, float *dst, ptrdiff_t dst_stride);

/* ---- conversions (include/Conversion.hpp) -------------------------------- */
/* Int2Float (Conversion.hpp:424-437): bits 8..16 integer plane -> float. */
void or_int2float(const void *src, int bits, ptrdiff_t src_stride, float *dst, ptrdiff_t dst_stride,
                  int W, int H, int chroma, int full, int clip);
/* Float2Int (Conversion.hpp:440-452). */
void or_float2int(const float *src, ptrdiff_t src_stride, void *dst, int bits, ptrdiff_t dst_stride,
                  int W, int H, int chroma, int full, int clip);
/* MatrixConvert_RGB2YUV OPP branch (Conversion.hpp:205-236) for arbitrary
 * src/dst sample formats: bits 0 = float32. */
void or_rgb2opp(const void *const *src, int src_bits, ptrdiff_t src_stride, void *const *dst,
                int dst_bits, ptrdiff_t dst_stride, int W, int H, int clip);
/* ConvertToY OPP branch (Conversion.hpp:101-115): float Y only. */
void or_rgb2opp_y(const void *const *src, int src_bits, ptrdiff_t src_stride, float *dst,
                  ptrdiff_t dst_stride, int W, int H);
/* MatrixConvert_YUV2RGB OPP branch (Conversion.hpp:369-392). */
void or_opp2rgb(const void *const *src, int src_bits, ptrdiff_t src_stride, void *const *dst,
                int dst_bits, ptrdiff_t dst_stride, int W, int H, int clip);

/* ---- CPU baseline runner -------------------------------------------------- */
/* RGB48 -> RGB2OPP -> Basic -> Final -> OPP2RGB on nframes frames with
 * nthreads worker threads, one frame per thread (VapourSynth fmParallel).
 * rgb_in / rgb_out: [frame][plane] uint16 planes of W x H, stride W. */
int or_chain_rgb48(const or_filter *basic, const or_filter *final_, int nframes, int nthreads,
                   int W, int H, const uint16_t *const *rgb_in, uint16_t *const *rgb_out);

#ifdef __cplusplus
}
#endif
#endif
