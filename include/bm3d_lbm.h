/*
 * bm3d_lbm.h -- learned block-matching proposal for the MI355X BM3D library.
 *
 * No counterpart in the reference: its block matching is the exhaustive
 * search of Block::BlockMatchingMulti (include/Block.h:683-729).  This
 * extension scores the even-lattice positions (even row and column) of each
 * search window with a small patch-embedding network (64 -> 128 ReLU -> 32,
 * BF16 weights, FP32 accumulation on MFMA), proposes the best min(topk, 32)
 * of them, re-ranks the 3x3 neighbourhoods of the proposals by the exact FP32
 * SSD with the reference's acceptance rule (dist <= thSSE && dist != 0) and
 * key (float(dist * distMul)), and falls back to the exhaustive search for a
 * reference block whose candidates do not yield G-1 acceptable matches.
 * Windows of at most topk candidates are searched exhaustively.  Filtering
 * and aggregation are unchanged.  Block size 8, bm_step 1, bm_range <= 24,
 * group_size <= 64.
 *
 * All plane, pair and match-list pointers are device pointers unless stated.
 * Planes are float32 with a row pitch in elements.  Calls are serialised on
 * the model's (or the filter's) own HIP stream and return when done.  That
 * stream is ordered after the legacy default stream only: device inputs
 * written on another stream must be complete before the call.
 */
#ifndef BM3D_LBM_H
#define BM3D_LBM_H

#include <stdint.h>

#include "bm3d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bm3d_lbm bm3d_lbm;

/* Number of network parameters, laid out [W1 (128x64) | b1 (128) | W2 (32x128) | b2 (32)], row-major. */
int bm3d_lbm_param_count(void);

/* Creates a model from host parameters (bm3d_lbm_param_count() floats). */
int bm3d_lbm_create(bm3d_lbm **out, const float *params);
void bm3d_lbm_free(bm3d_lbm *m);
int bm3d_lbm_get_params(const bm3d_lbm *m, float *params);
/* Replaces the parameters and resets the optimiser state. */
int bm3d_lbm_set_params(bm3d_lbm *m, const float *params);

/* Training.  pairs: npairs records (ref y, ref x, cand y, cand x) of 8x8
 * patches of plane.  target: per-pair SSD, or NULL to use the pairs' own
 * SSD.  Loss = mean over pairs of (d - t)^2 / (t + 1)^2, d the squared
 * embedding distance.  bm3d_lbm_train_step applies one Adam step (lr) and
 * returns the batch loss before the step; bm3d_lbm_gradients only computes
 * the gradient (host array of bm3d_lbm_param_count() floats). */
int bm3d_lbm_train_step(bm3d_lbm *m, const float *plane, int pitch, const int *pairs, int npairs, float lr,
                        float *loss);
int bm3d_lbm_gradients(bm3d_lbm *m, const float *plane, int pitch, const int *pairs, const float *target,
                       int npairs, float *grad, float *loss);
/* Embeddings of the pairs' patches (host out, [2 npairs][32]: reference then candidate). */
int bm3d_lbm_forward(bm3d_lbm *m, const float *plane, int pitch, const int *pairs, int npairs, float *emb);

/* Block matching of one plane with the parameters of a Basic/Final filter
 * (block_size 8, bm_step 1): m = NULL runs the exhaustive search, otherwise
 * the proposal path of m with topk (1..64) as above.  count/pos/key receive
 * the match lists in the layout of bm3d_match_lists.  *fallback (host, may be
 * NULL) receives the number of reference blocks that fell back to the
 * exhaustive search; *ms (host, may be NULL) the GPU time of the matching. */
int bm3d_lbm_match_plane(const bm3d_filter *f, const bm3d_lbm *m, int topk, const float *plane, int pitch, int width,
                         int height, int *count, int *pos, float *key, int *fallback, float *ms);

/* Matching parameters of a Basic/Final filter after profile defaults:
 * out[0..4] = block_size, block_step, group_size, bm_range, bm_step. */
int bm3d_filter_params(const bm3d_filter *f, int *out);

/* Makes a Basic/Final filter use the proposal path of m (NULL: exhaustive). */
int bm3d_set_matcher(bm3d_filter *f, bm3d_lbm *m, int topk);
/* Reference blocks that fell back to the exhaustive search in the last frame. */
int bm3d_matcher_fallbacks(const bm3d_filter *f, int *fallback);

/* Host copy of all match lists of the last Basic/Final frame: count[nb],
 * pos[nb * G] packed (y << 16) | x, key[nb * G]; *nb and *group receive the
 * sizes (call with NULL arrays to query them). */
int bm3d_match_lists(const bm3d_filter *f, int64_t *nb, int *group, int *count, int *pos, float *key);

#ifdef __cplusplus
}
#endif

#endif
