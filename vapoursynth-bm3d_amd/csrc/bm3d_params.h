// Parameter blocks shared by the host engine and the HIP kernels.
//
// The values mirror what the reference keeps in BM3D_Para / VBM3D_Para and
// BM3D_FilterData after argument processing (source/BM3D.cpp:33-209,
// source/BM3D_Base.cpp:262-283, source/VBM3D_Base.cpp:32-97, 382-403).
#pragma once

#include <cstdint>

namespace bm3d {

constexpr int kMaxGroup = 256;   // group_size range [1, 256] (BM3D_Base.cpp:160)
constexpr int kMaxBlock = 64;    // block_size range [1, 64]  (BM3D_Base.cpp:136)
constexpr int kRcpBits = 11;     // mantissa bits indexing the Wiener reciprocal table

// Arithmetic flavour of the reference build being reproduced
// (include/Helper.h:45-59): the x86 build's SSE2 paths (lane-order SSD,
// Block.h:582-631; rcpps Wiener gain, BM3D_Final.cpp:80-104) or the scalar
// paths of builds without __SSE2__ (row-major SSD, Block.h:632-643; exact
// Wiener division and a sequential L2 sum, BM3D_Final.cpp:106-112).
enum { kArithSSE2 = 0, kArithScalar = 1 };

// Geometry and matching parameters of one filter stage.
struct StageParams {
    int width, height, pitch;     // plane size and row pitch (floats)
    int block, block_step;        // B, reference-block step
    int group;                    // G
    int range, step;              // search range (rounded down to a multiple of step) and step
    int nx, ny;                   // reference grid columns / rows
    float th_sse;                 // float(thMSE' * B^2 / 65025)
    double dist_mul;              // 65025 / B^2
    int radius, ps_num, ps_range, ps_step;   // V-BM3D only
    int arith;                    // kArithSSE2 or kArithScalar
};

// Per-plane shrinkage constants indexed by actual group size n (1..G).
struct PlaneTables {
    float thr[kMaxGroup][4];          // hard threshold by #zero-frequency axes
    float wiener_sigma_sqr[kMaxGroup];
    double final_amp[kMaxGroup];      // 2n * 2B * 2B
};

}  // namespace bm3d
