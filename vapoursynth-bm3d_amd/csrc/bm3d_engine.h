// Internal declarations shared by the host engine translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/bm3d_amd.h"
#include "bm3d_params.h"

namespace bm3d {

int set_error(int code, const std::string& msg);
// Fails unless bm3d_init ran; makes the library's device current.
int ensure_init();
void log_warning(const std::string& msg);
int rcp_bits();
const float* rcp_dev();
// Arithmetic the reference build on this host would use: kArithSSE2 when the
// host's _mm_rcp_ps can be tabulated, else kArithScalar.  Needs no GPU.
int host_default_arith();

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::bm3d::set_error(BM3D_E_HIP, std::string("HIP error: ") +                  \
                                                     hipGetErrorString(e_) + " at " #expr);    \
    } while (0)

enum Fn { kRGB2OPP, kOPP2RGB, kBasic, kFinal, kVBasic, kVFinal, kVAggregate };

// Device allocation that only grows (reused across frames of one filter).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want) {
        if (want <= bytes) return BM3D_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (want == 0) return BM3D_OK;
        HIP_TRY(hipMalloc(&p, want));
        bytes = want;
        return BM3D_OK;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Work buffers of the learned proposal path of one filter.
struct LbmWork {
    DevBuf emb, norm, remb, rnorm, fallback, nfallback, props;
};

// Where one frame call takes its device scratch from.  A null Scratch*
// means the filter's own DevBufs (grown with hipMalloc as needed).
// Otherwise consecutive 256-byte aligned pieces of a fixed area of cap
// bytes at base; with dry set nothing is launched, copied or recorded and
// only `used` advances (the sizing pass of a call that must not allocate).
struct Scratch {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    bool dry = false;
};

// `bytes` of device scratch for buffer `buf` of the call (see Scratch).
int take(Scratch* sc, DevBuf& buf, size_t bytes, void** out);
template <typename T>
int take_as(Scratch* sc, DevBuf& buf, size_t bytes, T** out) {
    void* p = nullptr;
    int rc = take(sc, buf, bytes, &p);
    *out = static_cast<T*>(p);
    return rc;
}

}  // namespace bm3d

struct bm3d_lbm;

struct bm3d_filter {
    bm3d::Fn fn;
    bm3d_clip_info vi{}, rvi{}, ovi{};
    bool rdef = false;
    bool wiener = false, temporal = false;
    // BM3D_Para / VBM3D_Para after argument processing
    std::string profile = "fast";
    double sigma[3] = {10, 10, 10};
    int B = 8, step = 8, G = 8, R = 9, S = 1;
    double th_mse = 0, lambda = 0;   // th_mse already multiplied by normY
    int radius = 1, ps_num = 2, ps_range = 4, ps_step = 1;
    int matrix = 2;
    int arith = 0;                   // kArithSSE2 / kArithScalar (non-reference argument "arith")
    int process[3] = {1, 1, 1};
    int sample = 0;                  // RGB2OPP / OPP2RGB / VAggregate
    // constant tables
    bm3d::PlaneTables tables[3]{};
    bm3d::DevBuf tables_dev, coef_dev, coef_off_dev;
    // per-frame workspace (reused)
    bm3d::DevBuf in_raw[2][3 * 33];  // [src/ref][plane * frames] staged host input
    bm3d::DevBuf fplane[2][3 * 33];  // converted float planes
    bm3d::DevBuf out_stage[3];
    bm3d::DevBuf m_count, m_pos, m_key, m_frame, m_order, m_vroff, m_vpair, contrib[3], denw[3], ptrs;
    // learned block-matching proposal (null: exhaustive search)
    bm3d_lbm* lbm = nullptr;
    int lbm_topk = 32;
    bm3d::LbmWork lbm_work;
    const float* ptrs_host[6] = {};  // source of the device plane-pointer table
    hipStream_t stream = nullptr;
    mutable std::mutex mu;
    bm3d_stats stats{};
    int last_nb = 0;
    // match lists of the last frame (the filter's buffers, or the fixed
    // scratch area after an allocation-free call)
    const int* last_count = nullptr;
    const int* last_pos = nullptr;
    const float* last_key = nullptr;
    const int* last_frame = nullptr;
    int last_W = 0, last_H = 0;
    // per-kernel event timing
    static constexpr int kMaxTimers = 8;
    bool prof = false;
    const char* prof_name[kMaxTimers] = {};
    double prof_ms[kMaxTimers] = {};
    int64_t prof_n[kMaxTimers] = {};
    hipEvent_t prof_ev[2 * kMaxTimers] = {};
    bool prof_pending[kMaxTimers] = {};
    ~bm3d_filter() {
        for (auto& e : prof_ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace bm3d {
int run_frame(bm3d_filter* f, int n, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame* dst);
// Basic/Final frame with every scratch buffer taken from sc (non-null).
int run_bm3d_scratch(bm3d_filter* f, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame* dst, Scratch* sc);
// Registers the fixed scratch area of the allocation-free entry point for
// the current device (called by bm3d_init with that device current).
int noalloc_setup(int device);
// Timer slots used by the launch sequence.
enum Timer { kTMatch = 0, kTFilter = 1, kTAggregate = 2, kTConvert = 3, kTVAggregate = 4 };
void prof_begin(bm3d_filter* f, int t, const char* name);
void prof_end(bm3d_filter* f, int t);
void prof_collect(bm3d_filter* f);
StageParams stage_params(const bm3d_filter* f, int W, int H, size_t pitch);
// Block matching of one plane (exhaustive, or the proposal path of m).
int lbm_match(const bm3d_lbm* m, int topk, LbmWork& w, const float* plane, const StageParams& sp, int* count,
              int* pos, float* key, hipStream_t st);
}
