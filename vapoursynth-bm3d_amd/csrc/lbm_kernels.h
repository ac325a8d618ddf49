// Learned block-matching proposal: network layout, argument blocks and
// launchers (lbm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "bm3d_params.h"

namespace bm3d {
namespace lbm {

constexpr int kIn = 64;       // 8x8 patch
constexpr int kHidden = 128;
constexpr int kEmb = 32;
// parameter vector: [W1 (kHidden x kIn) | b1 | W2 (kEmb x kHidden) | b2], row-major
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kIn;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kEmb * kHidden;
constexpr int kParams = kOffB2 + kEmb;
constexpr int kMaxProposals = 64;   // largest top-K
constexpr int kLat = 2;             // proposal lattice stride (rows and columns)
constexpr int kMaxWindow = 49 * 49; // largest search window (bm_range 24)
constexpr int kPropSlots = kMaxProposals / 2;  // lattice points proposed per reference block (at most)

// Device view of one network: FP32 master parameters and the BF16 copies of
// the two weight matrices the MFMA kernels read.
struct Net {
    float* master;       // [kParams]
    short* w1h;          // [kHidden][kIn] bf16 bits
    short* w2h;          // [kEmb][kHidden] bf16 bits
    const float* b1;     // master + kOffB1
    const float* b2;     // master + kOffB2
};

// A training batch of npairs (reference patch, candidate patch) pairs of one
// plane; pairs[p] = (ref y, ref x, cand y, cand x).  Patch 2p is the
// reference, 2p+1 the candidate.  Work arrays are per patch.
struct TrainBatch {
    const float* plane;
    int pitch;
    const int4* pairs;
    int npairs;
    const float* target;  // [npairs] SSD of the pair
    short* xs;            // [2 npairs][kIn] bf16 patches
    float* hs;            // [2 npairs][kHidden] post-ReLU hidden
    float* es;            // [2 npairs][kEmb] embeddings
    float* des;           // [2 npairs][kEmb] dL/dE
    float* dhs;           // [2 npairs][kHidden] dL/dH before the ReLU
};

struct ProposeArgs {
    StageParams p;
    const float* plane;
    const _Float16* emb;   // even lattice: [ceil((H-7)/2) * ceil((W-7)/2) + 1][kEmb] (last: spare slot)
    const float* norm;     // [ceil((H-7)/2) * ceil((W-7)/2) + 1]
    const _Float16* remb;  // reference blocks: [nx * ny + 1][kEmb]
    const float* rnorm;    // [nx * ny + 1]
    int max_cand;         // (2 range + 1)^2
    int topk;
    int* count;
    int* pos;
    float* key;
    int* fallback;   // reference blocks left to the exhaustive search
    int* nfallback;  // their number (device counter, cleared by the caller)
    uint16_t* props;      // [nx * ny][kPropSlots] proposed lattice points (block-local index, 0xFFFF = none)
};

hipError_t launch_train_fwd(const Net& n, const TrainBatch& tb, hipStream_t st);
hipError_t launch_train_loss(const TrainBatch& tb, float t0, float* block_loss, hipStream_t st);
hipError_t launch_train_bwd(const Net& n, const TrainBatch& tb, float* part, float* grad, hipStream_t st);
int train_slices(int npairs);
hipError_t launch_train_adam(const Net& n, const float* grad, float* m, float* v, float lr, float b1, float b2,
                             float eps, float c1, float c2, hipStream_t st);
hipError_t launch_sync_bf16(const Net& n, hipStream_t st);
hipError_t launch_pair_ssd(const float* plane, int pitch, const int4* pairs, int npairs, float* target,
                           hipStream_t st);
// The lattice embeddings (emb/norm: ceil((W-7)/2) * ceil((H-7)/2) + 1 entries, the last a spare slot) and
// the reference blocks' (remb/rnorm: nx * ny + 1 entries) in one launch; also clears *nfallback for the
// re-rank that follows.
hipError_t launch_embed(const Net& n, const float* plane, const StageParams& p, _Float16* emb, float* norm,
                        _Float16* remb, float* rnorm, int* nfallback, hipStream_t st);
constexpr int kMaxLearnedGroupSize = 64;
hipError_t launch_propose(const ProposeArgs& a, hipStream_t st);
size_t propose_list_bytes(const StageParams& p);

}  // namespace lbm
}  // namespace bm3d
