/*
 * bm3d_batch.h -- batched Basic+Final of many Gray float frames of uneven
 * sizes on one MI355X.
 *
 * No counterpart in the reference, whose filters run frame by frame (one
 * VapourSynth frame request per call, source/VSPlugin.cpp:180).  A batch
 * holds nframes frames, each with its own width, height and sigma (and so
 * its own default matrix and thresholds, exactly as a separate Basic/Final
 * call would derive them, source/BM3D_Base.cpp:106-236); the remaining
 * arguments (profile, block_size, ...) are shared.  Every stage -- block
 * matching, collaborative filtering, aggregation -- runs as one
 * launch over the work of all frames; the Basic -> Final chain can be
 * captured once into a HIP graph and replayed.  Each output equals
 * Final(Basic(frame)) of the single-frame filters bit for bit.
 *
 * Frames are device planes of float32 with a row pitch in elements.  Calls
 * run on the batch's stream (ordered after the legacy default stream) and
 * return when the outputs are written; calls on one batch are serialised.
 */
#ifndef BM3D_BATCH_H
#define BM3D_BATCH_H

#include <stdint.h>

#include "bm3d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bm3d_batch bm3d_batch;

/* Launch shape of the batched stages (chosen by bm3d_batch_tune). */
typedef struct bm3d_batch_shape {
    int match_waves;  /* reference blocks (waves) per matching workgroup: 1, 2 or 4 */
    int match_strip;  /* candidates per lane strip of the LDS search tile: 3, 8, 10 or 11 */
    int group_order;  /* 0: groups bucketed by size per frame; 1: per chunk of 4096 groups */
} bm3d_batch_shape;

/* widths, heights, pitches (elements), sigmas: nframes entries each.  args:
 * Basic/Final arguments shared by all frames (e.g. profile), "sigma"
 * excluded.  Errors are reported as bm3d_create's, prefixed by the frame. */
int bm3d_batch_create(bm3d_batch **out, int nframes, const int *widths, const int *heights, const int *pitches,
                      const double *sigmas, const bm3d_arg *args, int nargs, char *err, size_t errlen);
void bm3d_batch_free(bm3d_batch *b);

/* Runs Basic then Final on every frame: src[f] -> dst[f] (device planes with
 * the pitches given at creation).  graph != 0: the chain is captured into a
 * HIP graph on the first such call and replayed afterwards.  *ms (may be
 * NULL) receives the GPU time of the chain. */
int bm3d_batch_run(bm3d_batch *b, const float *const *src, float *const *dst, int graph, float *ms);

/* Autotuner: times every supported launch shape on the chain (reps runs
 * each, outputs into dst) and keeps the fastest.  report (may be NULL)
 * receives one line per shape. */
int bm3d_batch_tune(bm3d_batch *b, const float *const *src, float *const *dst, int reps, char *report,
                    size_t report_len);
int bm3d_batch_get_shape(const bm3d_batch *b, bm3d_batch_shape *shape);
int bm3d_batch_set_shape(bm3d_batch *b, const bm3d_batch_shape *shape);

/* Totals of the batch: reference blocks of Basic and Final, pixels. */
int bm3d_batch_info(const bm3d_batch *b, int64_t *basic_blocks, int64_t *final_blocks, int64_t *pixels);

#ifdef __cplusplus
}
#endif

#endif
