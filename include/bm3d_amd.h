/*
 * bm3d_amd.h -- C ABI of the MI355X BM3D / V-BM3D library.
 *
 * The reference is a VapourSynth API-v4 plugin (namespace "bm3d") whose
 * filters are created from a VSMap of arguments and then asked for frames
 * (source/VSPlugin.cpp:424-518).  This header is the thin boundary a
 * VapourSynth plugin shim (or any other host) binds: plain structs,
 * pointers and sizes, integer error codes, no exceptions.  Each entry point
 * names the reference interface it replaces.
 *
 * Threading: calls on different filters may run concurrently; calls on one
 * filter are serialised internally (the reference's GetFrame is fmParallel,
 * source/VSPlugin.cpp:180).
 */
#ifndef BM3D_AMD_H
#define BM3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM3D_AMD_VERSION_MAJOR 10
#define BM3D_AMD_VERSION_MINOR 1

/* Error codes. */
enum {
    BM3D_OK = 0,
    BM3D_E_ARG = 1,      /* argument validation failed (message as mapSetError) */
    BM3D_E_HIP = 2,      /* HIP runtime error */
    BM3D_E_STATE = 3,    /* library not initialised / bad handle */
    BM3D_E_UNSUPPORTED = 4
};

/* VapourSynth colour families and sample types (VSColorFamily / VSSampleType). */
enum { BM3D_CF_GRAY = 1, BM3D_CF_RGB = 2, BM3D_CF_YUV = 3 };
enum { BM3D_ST_INTEGER = 0, BM3D_ST_FLOAT = 1 };

/* Argument value types (the VS type strings of source/VSPlugin.cpp:430-517). */
enum { BM3D_ARG_INT = 0, BM3D_ARG_FLOAT = 1, BM3D_ARG_DATA = 2 };

/* A clip's constant format (VSVideoInfo).  width/height 0 = variable. */
typedef struct bm3d_clip_info {
    int color_family;
    int sample_type;
    int bits_per_sample;
    int subsampling_w, subsampling_h;
    int width, height;
    int num_frames;
} bm3d_clip_info;

/* One named argument (an entry of the VSMap passed to the create callback). */
typedef struct bm3d_arg {
    const char *key;
    int type;
    int count;               /* number of elements */
    const int64_t *ints;     /* BM3D_ARG_INT */
    const double *floats;    /* BM3D_ARG_FLOAT */
    const char *data;        /* BM3D_ARG_DATA */
} bm3d_arg;

/* Frame properties the filters read or write.  -1 = absent. */
typedef struct bm3d_props {
    int64_t bm3d_opp;          /* "BM3D_OPP" */
    int64_t range;             /* "_Range" */
    int64_t matrix;            /* "_Matrix" */
    int64_t v_radius;          /* "BM3D_V_radius" */
    int64_t v_process[3];      /* "BM3D_V_process" (-1 = absent) */
} bm3d_props;

/* One video frame: up to three planes.  stride in bytes.  mem: 0 = host
 * memory, 1 = device memory of the library's HIP device. */
typedef struct bm3d_frame {
    void *data[3];
    ptrdiff_t stride[3];
    int width, height;
    int mem;
    bm3d_props props;
} bm3d_frame;

typedef struct bm3d_filter bm3d_filter;

/* Log sink for warnings (vsapi->logMessage(mtWarning, ...) in the
 * reference: BM3D_Base.h:202-207, VBM3D_Base.h:265-270, VAggregate.h:153-194). */
typedef void (*bm3d_log_fn)(int level, const char *msg, void *user);

/* Selects the HIP device, probes the host reciprocal used by the Wiener
 * stage and uploads the constant tables.  Replaces plugin load
 * (VapourSynthPluginInit2, source/VSPlugin.cpp:424-428). */
int bm3d_init(int device);

/* Arithmetic the filters reproduce by default on this host: 0 = the x86
 * build's SSE2 paths (lane-order SSD, include/Block.h:582-631; rcpps Wiener
 * gain, source/BM3D_Final.cpp:80-104), chosen when the host's _mm_rcp_ps can
 * be tabulated; 1 = the scalar paths of builds without __SSE2__ (row-major
 * SSD, Block.h:632-643; exact Wiener division, BM3D_Final.cpp:106-112).
 * Runs the host probe once (no GPU needed); the environment variable
 * BM3D_AMD_HOST_RCP=none makes the probe report failure.  Basic, Final,
 * VBasic and VFinal also accept an int argument "arith" (0 or 1) that
 * overrides the default; it has no counterpart in the reference's argument
 * list (arith=0 is rejected where the probe failed). */
int bm3d_host_arith(void);

/* Message of the last error on this thread ("" if none). */
const char *bm3d_last_error(void);

void bm3d_set_log(bm3d_log_fn fn, void *user);

/* Creates one of the seven filters: "RGB2OPP", "OPP2RGB", "Basic", "Final",
 * "VBasic", "VFinal", "VAggregate".  Replaces the *_Create callbacks
 * (source/VSPlugin.cpp:56-181, 199-231, 265-294, 328-357, 389-417) and the
 * arguments_process they call (source/BM3D_Base.cpp:32-259,
 * source/BM3D_Basic.cpp:32-65, source/BM3D_Final.cpp:32-43,
 * source/VBM3D_Base.cpp:104-379, source/VAggregate.cpp:30-90,
 * include/RGB2OPP.h:54-94, include/OPP2RGB.h:54-94).  ref may be NULL.
 * On BM3D_E_ARG, err receives "bm3d.<Function>: <message>" exactly as
 * VSData::setError formats it (include/Helper.h:519-523). */
int bm3d_create(const char *function, const bm3d_clip_info *input, const bm3d_clip_info *ref,
                const bm3d_arg *args, int nargs, bm3d_filter **out, char *err, size_t errlen);

/* Output clip format (the VSVideoInfo given to createVideoFilter,
 * source/VSPlugin.cpp:75, 125, 180, 230, 284-286, 347-349, 409-411). */
int bm3d_output_info(const bm3d_filter *f, bm3d_clip_info *out);

/* Frames [first, last] of input (and ref) that frame n needs -- the
 * arInitial requests of GetFrame (source/VSPlugin.cpp:144-145, 244-253,
 * 307-316, 370-378). */
int bm3d_frame_range(const bm3d_filter *f, int n, int *first, int *last);

/* Computes output frame n.  src/ref: arrays holding frames first..last of
 * bm3d_frame_range (ref NULL when the filter has no ref clip).  dst planes
 * are provided by the caller (host or device memory); dst->props receives
 * the output frame's properties.  Replaces the *_GetFrame arAllFramesReady
 * path (VSProcess::process, include/Helper.h:695-725, and the
 * process_core / Kernel functions behind it).  Runs on the filter's stream,
 * which is ordered after the legacy default stream: device planes written on
 * another stream must be complete before the call.  Returns when dst is
 * written. */
int bm3d_get_frame(bm3d_filter *f, int n, const bm3d_frame *src, const bm3d_frame *ref, bm3d_frame *dst);

/* Allocation-free form of bm3d_get_frame for Basic and Final filters.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  The call makes no hipMalloc/hipFree or other
 * allocating call: every scratch buffer (converted input planes, match
 * lists, group contributions and weights, output staging) comes from one
 * fixed device area of bm3d_noalloc_scratch_bytes() bytes per GPU, held by
 * the library until the process ends.  A frame whose scratch exceeds it is
 * rejected with BM3D_E_UNSUPPORTED before anything runs (e.g. 64 MiB covers
 * np Final on 256x256 Gray float, or np Basic+Final on 192x192 RGB48).
 * Results equal bm3d_get_frame's.  All calls on one GPU share its area:
 * calls that overlap in time on one GPU (on any filters) race, so callers
 * serialise them.  The learned matcher (bm3d_set_matcher) is not supported
 * here.  The match lists read by bm3d_debug_matches after such a call live in
 * the shared area until the next call on that GPU. */
int bm3d_get_frame_noalloc(bm3d_filter *f, int n, const bm3d_frame *src, const bm3d_frame *ref, bm3d_frame *dst);

/* Size in bytes of the fixed scratch area of bm3d_get_frame_noalloc. */
size_t bm3d_noalloc_scratch_bytes(void);

/* Releases the filter (the *_Free callbacks, source/VSPlugin.cpp:80-84 etc.). */
void bm3d_free(bm3d_filter *f);

/* Work done by the last Basic/Final frame: reference blocks, SSD
 * evaluations, group members (sum of n) and sum of n^2 over groups -- the
 * inputs of the algorithmic FLOP count.  No counterpart in the reference. */
typedef struct bm3d_stats {
    int64_t ref_blocks;
    int64_t ssd_evals;
    int64_t members;
    int64_t members_sq;
} bm3d_stats;
int bm3d_last_stats(const bm3d_filter *f, bm3d_stats *out);

/* Per-kernel timing with HIP events recorded on the filter's own stream
 * around each launch.  bm3d_profile(f, 1) resets and starts accumulating;
 * bm3d_profile_read returns entry idx (BM3D_E_ARG past the last one).  No
 * counterpart in the reference. */
int bm3d_profile(bm3d_filter *f, int enable);
int bm3d_profile_read(const bm3d_filter *f, int idx, const char **name, double *total_ms, int64_t *launches);

/* Copies the match list of reference block b from the last Basic/Final
 * frame: positions (y, x), keys and count.  For parity tests of block
 * matching against Block::BlockMatchingMulti (include/Block.h:683-729);
 * no counterpart in the reference's API. */
int bm3d_debug_matches(const bm3d_filter *f, int64_t b, int *count, int *ys, int *xs, float *keys, int cap);

/* The same for the last VBasic/VFinal frame, with each member's frame
 * (index inside the frame's temporal window).  For parity tests against
 * VBM3D_Process_Base::BlockMatching (source/VBM3D_Base.cpp:540-660); no
 * counterpart in the reference's API. */
int bm3d_debug_vmatches(const bm3d_filter *f, int64_t b, int *count, int *frames, int *ys, int *xs, float *keys,
                        int cap);

#ifdef __cplusplus
}
#endif

#endif
