// Allocation-free Basic/Final frame entry point (bm3d_get_frame_noalloc).
//
// Callers that may not call hipMalloc (code under HIP graph capture, tight
// serving loops) pass device frames in and out as with bm3d_get_frame; every
// scratch buffer of the call (staged or converted input planes, match lists,
// the size-bucket order, the plane-pointer table, group contributions and
// weights, output staging) is carved from one fixed device array of
// kNoAllocScratchBytes per GPU, which the library holds until the process
// ends.  No state carries over between calls: each call writes every scratch
// word it later reads, so nothing has to start at zero.  All calls on one GPU
// share that GPU's array, so calls that overlap in time on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/bm3d_amd.h"
#include "bm3d_engine.h"

namespace bm3d {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

// The fixed scratch area.  Host code never uses this name as an address: the
// device address of each GPU's copy is looked up once per device at init.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices] = {};

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(BM3D_E_STATE, "bm3d_init: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return BM3D_OK;
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(BM3D_E_HIP, "bm3d_init: scratch area has no device address");
    g_scratch_ptr[device] = static_cast<char*>(p);
    return BM3D_OK;
}

static char* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace bm3d

using namespace bm3d;

extern "C" {

size_t bm3d_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int bm3d_get_frame_noalloc(bm3d_filter* f, int n, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame* dst) {
    (void)n;
    if (!f || !src || !dst) return set_error(BM3D_E_STATE, "bm3d_get_frame_noalloc: null argument");
    if (f->fn != kBasic && f->fn != kFinal)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d_get_frame_noalloc: only Basic and Final filters");
    if (f->rdef && !ref) return set_error(BM3D_E_STATE, "bm3d_get_frame_noalloc: the filter needs a ref frame");
    int rc = ensure_init();
    if (rc) return rc;
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    char* base = scratch_of(dev);
    if (!base) return set_error(BM3D_E_STATE, "bm3d_get_frame_noalloc: bm3d_init has not been called on this device");
    std::lock_guard<std::mutex> lk(f->mu);
    // sizing pass: launches nothing, touches no memory
    Scratch dry;
    dry.dry = true;
    rc = run_bm3d_scratch(f, src, ref, dst, &dry);
    if (rc) return rc;
    if (dry.used > kNoAllocScratchBytes)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d_get_frame_noalloc: the frame needs " + std::to_string(dry.used) +
                                                 " bytes of scratch; the fixed area holds " +
                                                 std::to_string(kNoAllocScratchBytes));
    Scratch sc;
    sc.base = base;
    sc.cap = kNoAllocScratchBytes;
    return run_bm3d_scratch(f, src, ref, dst, &sc);
}

}  // extern "C"
