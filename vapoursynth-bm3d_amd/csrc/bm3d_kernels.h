// Kernel argument blocks and launchers (implemented in bm3d_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "bm3d_params.h"

namespace bm3d {

enum OutMode { kOutFloat = 0, kOutInt = 1, kOutRGB = 2 };

// How the aggregation kernel stores num/den: the float planes as is, a
// Float2Int per plane (Conversion.hpp:440-452), or FloatYUV2RGB OPP.
struct OutputSpec {
    int mode;
    int bits;          // integer output bits (0 = float)
    int clip;
    void* dst[3];
    size_t pitch[3];   // elements
    float gain[3], offset[3], lo[3], hi[3];
    float m[9];        // YUV->RGB coefficients (kOutRGB)
};

struct FilterArgs {
    StageParams p;
    int b0, nb;
    int wiener;
    int process[3];
    const int* count;          // [nb]
    const int* pos;            // [nb * G] packed (y << 16) | x
    const int* frame;          // [nb * G] frame index (V-BM3D) or null
    const int* order;          // processing order of the groups (bucketed by size) or null
    const float* const* src[3];  // per plane: device array of frame plane pointers
    const float* const* ref[3];
    const PlaneTables* tables;   // device array, one per plane
    const float* rcp_table;    // rcp of 1.m indexed by the top mantissa bits
    int rcp_shift;             // 23 - (table index bits)
    const float* coef;         // DCT coefficients
    const int* coef_off;       // offsets by length N
    float* contrib[3];         // [nb * G * B * B]
    float* denw[3];            // [nb]
};

struct AggArgs {
    StageParams p;
    int process[3];
    int reach;                 // max |member - reference| displacement
    const int* count;
    const int* pos;
    const float* contrib[3];
    const float* denw[3];
    OutputSpec out;
};

struct RgbArgs {
    const void* src[3];
    void* dst[3];
    int src_bits, dst_bits;
    size_t spitch, dpitch;
    int W, H;
    int clip, y_only;
    float gain[3], offset[3], lo[2], hi[2];
    float m[9];
};

// One frame of a batched matching stage: its parameters and matching plane.
struct BatchFrame {
    StageParams p;
    const float* plane;
};

// Per-frame part of a batched filter stage.
struct BatchFilterFrame {
    int pitch;                   // elements
    const float* const* planes;  // device table: src p0..p2, ref p0..p2
    const PlaneTables* tables;   // device array, one per plane
};

// Group orders of a batched filter stage.
enum { kOrderFrame = 0, kOrderChunk = 1 };

// Tunable launch shapes of the batched stages.
struct BatchShape {
    int match_waves;   // reference blocks (waves) per matching workgroup: 1, 2 or 4
    int match_strip;   // candidates per lane strip in the LDS search tile: 3, 8, 10 or 11
    int group_order;   // kOrderFrame: size buckets per frame; kOrderChunk: per chunk of groups
};

constexpr int kMaxStackFrames = 33;  // 2 * radius + 1 with radius <= 16

struct VMatchArgs {
    StageParams p;
    int nb, nframes, cur;
    int max_cand, ps_cap, cat_cap;
    const float* planes[kMaxStackFrames];  // reference luma of the window's frames
    int* count;
    int* pos;
    int* frame;
    float* key;
};

struct VStackArgs {
    StageParams p;
    int nframes, cur;
    int process[3];
    const int* count;
    const int* pos;
    const int* frame;
    const float* contrib[3];
    const float* denw[3];
    float* dst[3];          // base of the (2r+1)*2 slot stack per plane
    size_t dst_pitch[3];    // elements
    int* roff;              // per-frame member lists (k_vframe_rows), or null: frame filter walk
    int2* rpair;
};

struct VAggArgs {
    int W, H, nframes, cur, radius;
    int process[3];
    const float* stacks[3 * kMaxStackFrames];  // [plane * kMaxStackFrames + neighbour]
    size_t stack_pitch[3];
    void* dst[3];
    size_t dst_pitch[3];
    int out_bits;
    float gain[3], offset[3], lo[3], hi[3];
};

size_t vmatch_lds_bytes(const StageParams& p, int max_cand, int ps_cap, int cat_cap);
hipError_t launch_vmatch(const VMatchArgs& a, hipStream_t st);
hipError_t launch_vstack(const VStackArgs& a, hipStream_t st);
size_t vframe_rows_words(const StageParams& p, int nframes);  // roff ints
size_t vframe_pairs(const StageParams& p, int nframes);       // rpair entries
hipError_t launch_vagg(const VAggArgs& a, hipStream_t st);

size_t match_wave_words(int max_cand, int B, int G);
size_t match_lds_bytes(int max_cand, int B, int G);
size_t filter_lds_bytes(int B, int G);
hipError_t upload_dct8(const float* c64);
hipError_t upload_dctz(const float* coef, int nmax);
bool filter_fits(int B, int G, bool wiener);
hipError_t launch_bucket(const int* count, int nb, int G, int* order, hipStream_t st);
bool filter_uses_buckets(int B, int G);
bool verify_dct_constants();
hipError_t launch_match(const float* plane, const StageParams& p, int nb, int max_cand, int* count, int* pos,
                        float* key, hipStream_t st, const int* blist = nullptr, const int* nlist = nullptr);
hipError_t launch_filter(const FilterArgs& a, hipStream_t st);
// Batched stages over frames numbered by starts[] (device, nframes entries).
bool batch_match_supported(const StageParams& p, int max_cand, const BatchShape& sh);
hipError_t launch_match_batch(const BatchFrame* frames, const int* starts, int nframes, int total,
                              const StageParams& p0, int max_cand, const BatchShape& sh, int* count, int* pos,
                              float* key, hipStream_t st);
hipError_t launch_filter_batch(const FilterArgs& a0, const BatchFilterFrame* ff, const int* starts, int nframes,
                               const int* order, int total, const BatchShape& sh, hipStream_t st);
hipError_t launch_bucket_batch(const int* count, const int* starts, int nframes, int nb, int G, int mode, int* hist,
                               int* order, hipStream_t st);
// planes: processed planes of every frame in the batch (LDS is sized for it).
hipError_t launch_aggregate_batch(const AggArgs* aa, const int* starts, int nframes, int total, int planes,
                                  const BatchShape& sh, hipStream_t st);
int device_cus();
int aggregate_tiles(int W, int H);

// Workgroups of a batched launch (xcd_item): one per `per` items, rounded up
// to a multiple of 8 so that every XCD gets the same number.
inline int batch_groups(int items, int per) {
    const long need = ((long)items + per - 1) / per;
    return (int)((need + 7) / 8 * 8);
}
hipError_t launch_aggregate(const AggArgs& a, hipStream_t st);
hipError_t launch_to_float(const void* src, int bits, size_t spitch, float* dst, size_t dpitch, int W, int H,
                           float gain, float offset, int clip, float lo, float hi, hipStream_t st);
hipError_t launch_rgb2opp(const RgbArgs& a, hipStream_t st);
hipError_t launch_opp2rgb(const RgbArgs& a, hipStream_t st);

}  // namespace bm3d
