// Host engine behind the C ABI (include/bm3d_amd.h): argument processing of
// the seven bm3d functions, constant tables, device workspaces and the
// per-frame launch sequence.  Numerics of every stage live in the kernels;
// this file restates the reference's parameter logic
// (source/BM3D.cpp:33-209, source/BM3D_Base.cpp:32-283,
// source/VBM3D_Base.cpp:32-403, source/VAggregate.cpp:30-90,
// include/RGB2OPP.h:54-94, include/OPP2RGB.h:54-94,
// include/Specification.h:139-313) and its per-frame dispatch
// (include/Helper.h:695-725, source/BM3D_Base.cpp:497-754).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#if defined(__SSE2__)
#include <emmintrin.h>
#endif

#include "../../include/bm3d_amd.h"
#include "bm3d_engine.h"
#include "bm3d_kernels.h"
#include "bm3d_params.h"

namespace bm3d {

// ---------------------------------------------------------------------------
// errors, logging, device state
// ---------------------------------------------------------------------------

thread_local std::string g_last_error;
static bm3d_log_fn g_log_fn = nullptr;
static void* g_log_user = nullptr;

void log_warning(const std::string& msg) {
    if (g_log_fn) g_log_fn(1, msg.c_str(), g_log_user);
    else std::fprintf(stderr, "%s\n", msg.c_str());
}

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

struct DeviceState {
    bool ready = false;
    int device = 0;
    int rcp_bits = 0;          // 0: exact division fallback
    std::vector<float> rcp_host;
    float* rcp_dev = nullptr;
    std::mutex mu;
};
static DeviceState g_dev;

// Probes _mm_rcp_ps on this host: the Wiener gain of the reference's SSE2
// path (source/BM3D_Final.cpp:94) depends on it.  Finds the smallest number
// of leading mantissa bits that determines the result for inputs in [1, 2)
// and checks that other exponents only rescale it.
static int probe_rcp(std::vector<float>& table, int& bits) {
#if defined(__SSE2__)
    const uint32_t N = 1u << 23;
    std::vector<uint32_t> r(N);
    for (uint32_t m = 0; m < N; m += 4) {
        __m128i v = _mm_setr_epi32(0x3f800000 | m, 0x3f800000 | (m + 1), 0x3f800000 | (m + 2), 0x3f800000 | (m + 3));
        __m128 o = _mm_rcp_ps(_mm_castsi128_ps(v));
        _mm_storeu_si128(reinterpret_cast<__m128i*>(r.data() + m), _mm_castps_si128(o));
    }
    bits = 23;
    for (int k = 1; k <= 23; ++k) {
        const uint32_t sh = 23 - k;
        bool ok = true;
        for (uint32_t m = 0; m < N && ok; ++m) ok = r[m] == r[(m >> sh) << sh];
        if (ok) {
            bits = k;
            break;
        }
    }
    table.resize(size_t(1) << bits);
    for (uint32_t i = 0; i < (1u << bits); ++i) {
        uint32_t u = r[i << (23 - bits)];
        std::memcpy(&table[i], &u, 4);
    }
    // exponent independence over the normal range that keeps results normal
    uint32_t seed = 12345;
    for (int t = 0; t < 200000; ++t) {
        seed = seed * 1664525u + 1013904223u;
        uint32_t m = (seed >> 5) & (N - 1);
        seed = seed * 1664525u + 1013904223u;
        int e = (int)(seed % 250u) - 124;
        float x = std::ldexp(1.0f + (float)m / (float)N, e);
        float got = _mm_cvtss_f32(_mm_rcp_ps(_mm_set1_ps(x)));
        float want = std::ldexp(table[m >> (23 - bits)], -e);
        if (std::memcmp(&got, &want, 4) != 0) return -1;
    }
    return 0;
#else
    (void)table;
    bits = 0;
    return -1;
#endif
}

// The host probe, run once per process.  BM3D_AMD_HOST_RCP=none makes it
// report failure, which selects the scalar arithmetic (a host without a
// tabulable rcpps behaves like the reference's non-SSE2 builds).
struct HostRcp {
    int status = -1;   // 0: tabulated
    int bits = 0;
    std::vector<float> table;
};
static const HostRcp& host_rcp() {
    static HostRcp h;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* env = std::getenv("BM3D_AMD_HOST_RCP");
        if (env && std::strcmp(env, "none") == 0) {
            h.status = -1;
            return;
        }
        h.status = probe_rcp(h.table, h.bits);
    });
    return h;
}

int host_default_arith() { return host_rcp().status == 0 ? kArithSSE2 : kArithScalar; }

int ensure_init() {
    if (!g_dev.ready) return set_error(BM3D_E_STATE, "bm3d_init has not been called");
    HIP_TRY(hipSetDevice(g_dev.device));
    return BM3D_OK;
}

// ---------------------------------------------------------------------------
// argument access (VSMap semantics: a missing key is "error", not a failure)
// ---------------------------------------------------------------------------

struct Args {
    const bm3d_arg* a;
    int n;
    const bm3d_arg* find(const char* key) const {
        for (int i = 0; i < n; ++i)
            if (a[i].key && std::strcmp(a[i].key, key) == 0) return &a[i];
        return nullptr;
    }
    int num(const char* key) const {
        const bm3d_arg* x = find(key);
        return x ? x->count : -1;
    }
    // mapGetIntSaturated
    bool get_int(const char* key, int& v, int idx = 0) const {
        const bm3d_arg* x = find(key);
        if (!x || idx >= x->count) return false;
        int64_t t = x->type == BM3D_ARG_INT ? x->ints[idx] : (int64_t)x->floats[idx];
        v = t > INT32_MAX ? INT32_MAX : (t < INT32_MIN ? INT32_MIN : (int)t);
        return true;
    }
    bool get_int64(const char* key, int64_t& v) const {
        const bm3d_arg* x = find(key);
        if (!x || x->count < 1) return false;
        v = x->type == BM3D_ARG_INT ? x->ints[0] : (int64_t)x->floats[0];
        return true;
    }
    bool get_float(const char* key, double& v, int idx = 0) const {
        const bm3d_arg* x = find(key);
        if (!x || idx >= x->count) return false;
        v = x->type == BM3D_ARG_FLOAT ? x->floats[idx] : (double)x->ints[idx];
        return true;
    }
    bool get_data(const char* key, std::string& v) const {
        const bm3d_arg* x = find(key);
        if (!x || x->type != BM3D_ARG_DATA || !x->data) return false;
        v = x->data;
        return true;
    }
};

// ---------------------------------------------------------------------------
// colour constants (include/Specification.h)
// ---------------------------------------------------------------------------

static void matrix_norms(int matrix, double norms[3]) {
    // ColorMatrix_RGB2YUV_Parameter<double> then init_filter_data's row norms
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
        // ColorMatrix_Parameter (Specification.h:71-136): double literals
        double r, g, b;
        switch (matrix) {
        case 4: r = 0.30; g = 0.59; b = 0.11; break;
        case 5: case 6: r = 0.299; g = 0.587; b = 0.114; break;
        case 7: r = 0.212; g = 0.701; b = 0.087; break;
        case 9: case 10: r = 0.2627; g = 0.6780; b = 0.0593; break;
        default: r = 0.2126; g = 0.7152; b = 0.0722; break;
        }
        long double kr = r, kg = g, kb = b;
        Yr = (double)kr; Yg = (double)kg; Yb = (double)kb;
        Ur = (double)(-(long double)Yr * 0.5L / (1.0L - (long double)Yb));
        Ug = (double)(-(long double)Yg * 0.5L / (1.0L - (long double)Yb));
        Ub = 0.5;
        Vr = 0.5;
        Vg = (double)(-(long double)Yg * 0.5L / (1.0L - (long double)Yr));
        Vb = (double)(-(long double)Yb * 0.5L / (1.0L - (long double)Yr));
    }
    norms[0] = std::sqrt(Yr * Yr + Yg * Yg + Yb * Yb);
    norms[1] = std::sqrt(Ur * Ur + Ug * Ug + Ub * Ub);
    norms[2] = std::sqrt(Vr * Vr + Vg * Vg + Vb * Vb);
}

static int default_matrix(int w, int h) {
    if (w > 2048 || h > 1536) return 9;   // UHD -> bt2020nc
    if (w > 1024 || h > 576) return 1;    // HD  -> bt709
    return 6;                             // SD  -> smpte170m
}

static bool is_valid_matrix(int m) {
    return m == 0 || m == 1 || m == 4 || m == 5 || m == 6 || m == 7 || m == 8 || m == 9 || m == 10 || m == 100;
}

// ---------------------------------------------------------------------------
// filter object
// ---------------------------------------------------------------------------

static const char* fn_name(Fn f) {
    static const char* n[] = {"RGB2OPP", "OPP2RGB", "Basic", "Final", "VBasic", "VFinal", "VAggregate"};
    return n[f];
}

}  // namespace bm3d

namespace bm3d {

// Fills PlaneTables exactly as BM3D_FilterData (source/BM3D.cpp:137-209).
static void build_tables(bm3d_filter* f, const double sig[3]) {
    for (int n = 1; n <= f->G; ++n) {
        int amp_i = 2 * n * 2 * f->B * 2 * f->B;
        double amp = (double)amp_i;
        double forward_amp = std::sqrt(amp);
        for (int p = 0; p < 3; ++p) {
            PlaneTables& t = f->tables[p];
            t.final_amp[n - 1] = amp;
            double sigma = sig[p];
            float v = (float)(sigma * forward_amp * sigma * forward_amp);
            t.wiener_sigma_sqr[n - 1] = v > FLT_MIN ? v : FLT_MIN;
            double base = sigma * f->lambda * forward_amp;
            t.thr[n - 1][0] = (float)base;
            t.thr[n - 1][1] = (float)(base * std::sqrt(2.0));
            t.thr[n - 1][2] = (float)(base * 2.0);
            t.thr[n - 1][3] = (float)(base * std::sqrt(8.0));
        }
    }
}

// DCT coefficient table for lengths 1..Nmax: coef[off(N) + k*N + j].
static int upload_coefs(bm3d_filter* f) {
    const int nmax = std::max(f->B, f->G);
    std::vector<int> off(nmax + 2, 0);
    for (int N = 1; N <= nmax; ++N) off[N + 1] = off[N] + N * N;
    std::vector<float> c(off[nmax + 1]);
    for (int N = 1; N <= nmax; ++N)
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                c[off[N] + k * N + j] =
                    (float)(2.0 * std::cos(M_PI * (double)k * (double)(2 * j + 1) / (2.0 * (double)N)));
    int rc;
    if ((rc = f->coef_dev.ensure(c.size() * sizeof(float)))) return rc;
    if ((rc = f->coef_off_dev.ensure(off.size() * sizeof(int)))) return rc;
    HIP_TRY(hipMemcpy(f->coef_dev.p, c.data(), c.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(f->coef_off_dev.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    if ((rc = f->tables_dev.ensure(sizeof(f->tables)))) return rc;
    HIP_TRY(hipMemcpy(f->tables_dev.p, f->tables, sizeof(f->tables), hipMemcpyHostToDevice));
    return BM3D_OK;
}

// BM3D_Data_Base / VBM3D_Data_Base::arguments_process (+ Basic's hard_thr)
static std::string parse_bm3d(bm3d_filter* f, const Args& args) {
    const bm3d_clip_info& vi = f->vi;
    if (vi.width <= 0 || vi.height <= 0 || vi.color_family == 0)
        return "Invalid input clip, only constant format input supported";
    if ((vi.sample_type == BM3D_ST_INTEGER && vi.bits_per_sample > 16) ||
        (vi.sample_type == BM3D_ST_FLOAT && vi.bits_per_sample != 32))
        return "Invalid input clip, only 8-16 bit integer or 32 bit float formats supported";
    if (vi.width > 65535 || vi.height > 32767)
        return "frames wider than 65535 or taller than 32767 pixels are not supported";
    if (f->rdef) {
        const bm3d_clip_info& r = f->rvi;
        if (r.width <= 0 || r.height <= 0 || r.color_family == 0)
            return "Invalid clip \"ref\", only constant format input supported";
        if (r.color_family != vi.color_family || r.sample_type != vi.sample_type ||
            r.bits_per_sample != vi.bits_per_sample || r.subsampling_w != vi.subsampling_w ||
            r.subsampling_h != vi.subsampling_h)
            return "input clip and clip \"ref\" must be of the same format";
        if (r.width != vi.width || r.height != vi.height)
            return "input clip and clip \"ref\" must be of the same width and height";
        if (r.num_frames != vi.num_frames) return "input clip and clip \"ref\" must have the same number of frames";
    } else {
        f->rvi = vi;
    }
    std::string profile;
    f->profile = args.get_data("profile", profile) ? profile : "fast";
    if (f->profile != "fast" && f->profile != "lc" && f->profile != "np" && f->profile != "high" && f->profile != "vn")
        return "Unrecognized \"profile\" specified, should be \"fast\", \"lc\", \"np\", \"high\" or \"vn\"";
    // profile defaults: BM3D_Para (source/BM3D.cpp:33-107), VBM3D_Para (VBM3D_Base.cpp:32-97)
    const bool w = f->wiener;
    const std::string& pr = f->profile;
    int dB = 8, dR = 16, dS = 1, dstep = w ? 3 : 4, dG = w ? 32 : 16;
    double dlambda = w ? 0.0 : 2.7;
    if (pr == "fast") {
        dR = 9; dG = 8; dstep = w ? 7 : 8;
    } else if (pr == "lc") {
        dR = 9;
        if (!w) dstep = 6;
        else { dstep = 5; dG = 16; }
    } else if (pr == "high") {
        dstep = w ? 2 : 3;
    } else if (pr == "vn") {
        if (!w) { dstep = 4; dG = 32; dlambda = 2.8; }
        else { dB = 11; dstep = 6; }
    }
    int dradius = 0, dps_num = 0, dps_range = 0, dps_step = 0;
    if (f->temporal) {
        dradius = 3; dG = 8; dR = 12; dps_num = 2; dps_step = 1; dps_range = w ? 6 : 5;
        if (pr == "fast") { dradius = 1; dR = 7; dps_range = w ? 5 : 4; }
        else if (pr == "lc") { dradius = 2; dR = 9; dps_range = w ? 5 : 4; }
        else if (pr == "high") { dradius = 4; dR = 16; dps_range = w ? 8 : 7; }
        else if (pr == "vn") { dradius = 4; dG = 16; }
    }
    // sigma
    int m = args.num("sigma");
    if (m > 0) {
        if (m > 3) m = 3;
        int i;
        for (i = 0; i < m; ++i) {
            args.get_float("sigma", f->sigma[i], i);
            if (f->sigma[i] < 0) return "Invalid \"sigma\" assigned, must be a non-negative floating point number";
        }
        for (; i < 3; ++i) f->sigma[i] = f->sigma[i - 1];
    } else {
        f->sigma[0] = f->sigma[1] = f->sigma[2] = 10.0;
    }
    if (f->temporal) {
        if (!args.get_int("radius", f->radius)) f->radius = dradius;
        else if (f->radius < 1 || f->radius > 16) return "Invalid \"radius\" assigned, must be an integer in [1, 16]";
    }
    if (!args.get_int("block_size", f->B)) f->B = dB;
    else if (f->B < 1 || f->B > 64) return "Invalid \"block_size\" assigned, must be an integer in [1, 64]";
    else if (f->B > vi.width || f->B > vi.height)
        return "Invalid \"block_size\" assigned, must not exceed width or height of the frame";
    if (!args.get_int("block_step", f->step)) f->step = dstep;
    else if (f->step < 1 || f->step > f->B) return "Invalid \"block_step\" assigned, must be an integer in [1, block_size]";
    if (!args.get_int("group_size", f->G)) f->G = dG;
    else if (f->G < 1 || f->G > 256) return "Invalid \"group_size\" assigned, must be an integer in [1, 256]";
    if (!args.get_int("bm_range", f->R)) f->R = dR;
    else if (f->R < 1) return "Invalid \"bm_range\" assigned, must be a positive integer";
    if (!args.get_int("bm_step", f->S)) f->S = dS;
    else if (f->S < 1 || f->S > f->R) return "Invalid \"bm_step\" assigned, must be an integer in [1, bm_range]";
    if (f->temporal) {
        if (!args.get_int("ps_num", f->ps_num)) f->ps_num = dps_num;
        else if (f->ps_num < 1 || f->ps_num > f->G) return "Invalid \"ps_num\" assigned, must be an integer in [1, group_size]";
        if (!args.get_int("ps_range", f->ps_range)) f->ps_range = dps_range;
        else if (f->ps_range < 1) return "Invalid \"ps_range\" assigned, must be a positive integer";
        if (!args.get_int("ps_step", f->ps_step)) f->ps_step = dps_step;
        else if (f->ps_step < 1 || f->ps_step > f->ps_range)
            return "Invalid \"ps_step\" assigned, must be an integer in [1, ps_range]";
    }
    // th_mse (thMSE_Default, source/BM3D.cpp:110-130)
    if (!args.get_float("th_mse", f->th_mse)) {
        const bool vn = pr == "vn";
        if (!w) f->th_mse = vn ? 1000 + f->sigma[0] * 150 : 400 + f->sigma[0] * 80;
        else f->th_mse = vn ? 400 + f->sigma[0] * 40 : 200 + f->sigma[0] * 10;
    } else if (f->th_mse <= 0) {
        return "Invalid \"th_mse\" assigned, must be a positive floating point number";
    }
    // matrix
    int64_t mat = 2;
    bool has_matrix = args.get_int64("matrix", mat);
    if (vi.color_family == BM3D_CF_RGB) f->matrix = 100;
    else if (!has_matrix || mat == 2) f->matrix = default_matrix(vi.width, vi.height);
    else if (!is_valid_matrix((int)mat) || mat != (int)mat) return "Unsupported \"matrix\" specified";
    else f->matrix = (int)mat;
    for (int i = 0; i < 3; ++i) f->process[i] = (vi.color_family != BM3D_CF_RGB && f->sigma[i] == 0) ? 0 : 1;
    if (f->process[1] || f->process[2]) {
        if (vi.subsampling_h || vi.subsampling_w)
            return "input clip: sub-sampled format is not supported when chroma is processed, convert it to YUV444 or RGB first. "
                   "For the best quality, RGB colorspace is recommended as input.";
        if (f->rvi.subsampling_h || f->rvi.subsampling_w)
            return "clip \"ref\": sub-sampled format is not supported when chroma is processed, convert it to YUV444 or RGB first. "
                   "For the best quality, RGB colorspace is recommended as input.";
    }
    if (!w) {
        // hard_thr (source/BM3D_Basic.cpp:44-53)
        if (!args.get_float("hard_thr", f->lambda)) f->lambda = dlambda;
        else if (f->lambda <= 0) return "Invalid \"hard_thr\" assigned, must be a positive floating point number";
    } else {
        f->lambda = dlambda;
    }
    // "arith" has no counterpart in the reference's argument list: it selects
    // which reference build's arithmetic to reproduce (0: x86 SSE2 paths,
    // 1: the scalar paths of builds without __SSE2__); default: what the
    // reference would use on this host.
    if (!args.get_int("arith", f->arith)) f->arith = host_default_arith();
    else if (f->arith != kArithSSE2 && f->arith != kArithScalar)
        return "Invalid \"arith\" assigned, must be 0 (x86 SSE2 arithmetic) or 1 (scalar arithmetic)";
    else if (f->arith == kArithSSE2 && host_default_arith() != kArithSSE2)
        return "\"arith=0\" needs the host's SSE2 rcpps, which could not be tabulated on this host";
    return "";
}

static std::string parse_color(bm3d_filter* f, const Args& args, bool to_opp) {
    const bm3d_clip_info& vi = f->vi;
    if (vi.width <= 0 || vi.height <= 0 || vi.color_family == 0)
        return "Invalid input clip, only constant format input supported";
    if ((vi.sample_type == BM3D_ST_INTEGER && vi.bits_per_sample > 16) ||
        (vi.sample_type == BM3D_ST_FLOAT && vi.bits_per_sample != 32))
        return "Invalid input clip, only 8-16 bit int or 32 bit float formats supported";
    if (to_opp && vi.color_family != BM3D_CF_RGB) return "Invalid input clip, must be of RGB color family";
    if (!to_opp && vi.color_family != BM3D_CF_YUV) return "Invalid input clip, must be of YUV color family";
    int64_t s = 0;
    if (!args.get_int64("sample", s)) s = 0;
    else if (s != 0 && s != 1)
        return "Invalid 'sample' assigned, must be 0 (integer sample type) or 1 (float sample type)";
    f->sample = (int)s;
    return "";
}

static std::string parse_vaggregate(bm3d_filter* f, const Args& args) {
    const bm3d_clip_info& vi = f->vi;
    if (vi.width <= 0 || vi.height <= 0 || vi.color_family == 0)
        return "Invalid input clip, only constant format input supported";
    if (vi.sample_type != BM3D_ST_FLOAT || vi.bits_per_sample != 32)
        return "Invalid input clip, only accept 32 bit float format clip from bm3d.VBasic or bm3d.VFinal";
    if (vi.color_family == BM3D_CF_RGB) return "Invalid input clip, must be of Gray or YUV color family";
    if (!args.get_int("radius", f->radius)) f->radius = 1;
    else if (f->radius < 1 || f->radius > 16) return "Invalid \"radius\" assigned, must be an integer in [1, 16]";
    int64_t s = 0;
    if (!args.get_int64("sample", s)) s = 0;
    else if (s != 0 && s != 1)
        return "Invalid 'sample' assigned, must be 0 (integer sample type) or 1 (float sample type)";
    f->sample = (int)s;
    return "";
}

}  // namespace bm3d

using namespace bm3d;

extern "C" {

int bm3d_init(int device) {
    std::lock_guard<std::mutex> lk(g_dev.mu);
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(BM3D_E_STATE, "bm3d_init: no such HIP device");
    HIP_TRY(hipSetDevice(device));
    if (g_dev.ready && g_dev.device == device) return BM3D_OK;
    g_dev.device = device;
    const HostRcp& hr = host_rcp();
    std::vector<float> tbl = hr.table;
    int bits = hr.bits;
    if (hr.status != 0) {
        // No tabulable rcpps on this host: filters default to the reference's
        // scalar (non-SSE2) arithmetic -- row-major SSD, exact Wiener
        // division -- which needs no table (a one-entry placeholder is kept so
        // the device pointer is valid).
        log_warning("bm3d: host _mm_rcp_ps could not be tabulated; filters use the scalar (non-SSE2) arithmetic");
        bits = 0;
        tbl.assign(1, 1.0f);
    }
    g_dev.rcp_bits = bits;
    g_dev.rcp_host = tbl;
    if (g_dev.rcp_dev) (void)hipFree(g_dev.rcp_dev);
    HIP_TRY(hipMalloc(&g_dev.rcp_dev, tbl.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(g_dev.rcp_dev, tbl.data(), tbl.size() * sizeof(float), hipMemcpyHostToDevice));
    // 8-point DCT coefficients of the B = 8 kernels (same expression as the
    // length-N tables of upload_coefs)
    float c8[64];
    for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 8; ++j)
            c8[k * 8 + j] = (float)(2.0 * std::cos(M_PI * (double)k * (double)(2 * j + 1) / (2.0 * 8.0)));
    HIP_TRY(upload_dct8(c8));
    // z-axis coefficients for N = 1..32 (same expression and layout)
    std::vector<float> cz;
    for (int N = 1; N <= 32; ++N)
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                cz.push_back((float)(2.0 * std::cos(M_PI * (double)k * (double)(2 * j + 1) / (2.0 * (double)N))));
    HIP_TRY(upload_dctz(cz.data(), 32));
    if (!verify_dct_constants())
        return set_error(BM3D_E_STATE, "bm3d_init: compiled DCT constants differ from this host's cos(); rebuild");
    int rc = noalloc_setup(device);
    if (rc) return rc;
    g_dev.ready = true;
    return BM3D_OK;
}

const char* bm3d_last_error(void) { return g_last_error.c_str(); }

int bm3d_host_arith(void) { return host_default_arith(); }

void bm3d_set_log(bm3d_log_fn fn, void* user) {
    g_log_fn = fn;
    g_log_user = user;
}

int bm3d_create(const char* function, const bm3d_clip_info* input, const bm3d_clip_info* ref, const bm3d_arg* args,
                int nargs, bm3d_filter** out, char* err, size_t errlen) {
    if (out) *out = nullptr;
    int rc;
    if (!function || !input || !out) return set_error(BM3D_E_ARG, "bm3d_create: null argument");
    static const char* names[] = {"RGB2OPP", "OPP2RGB", "Basic", "Final", "VBasic", "VFinal", "VAggregate"};
    int fi = -1;
    for (int i = 0; i < 7; ++i)
        if (std::strcmp(function, names[i]) == 0) fi = i;
    if (fi < 0) return set_error(BM3D_E_ARG, std::string("bm3d_create: unknown function ") + function);
    std::unique_ptr<bm3d_filter> f(new bm3d_filter());
    f->fn = (Fn)fi;
    f->vi = *input;
    f->rdef = ref != nullptr;
    if (ref) f->rvi = *ref;
    f->wiener = f->fn == kFinal || f->fn == kVFinal;
    f->temporal = f->fn == kVBasic || f->fn == kVFinal;
    Args a{args, nargs};
    std::string msg;
    if ((f->fn == kFinal || f->fn == kVFinal) && !ref) msg = "Argument ref is required";
    else if (f->fn == kRGB2OPP || f->fn == kOPP2RGB) msg = parse_color(f.get(), a, f->fn == kRGB2OPP);
    else if (f->fn == kVAggregate) msg = parse_vaggregate(f.get(), a);
    else msg = parse_bm3d(f.get(), a);
    if (!msg.empty()) {
        std::string full = std::string("bm3d.") + fn_name(f->fn) + ": " + msg;
        if (err && errlen) {
            std::strncpy(err, full.c_str(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return set_error(BM3D_E_ARG, full);
    }
    // output format (source/VSPlugin.cpp: createVideoFilter video infos)
    f->ovi = f->vi;
    switch (f->fn) {
    case kRGB2OPP:
        f->ovi.color_family = BM3D_CF_YUV;
        f->ovi.sample_type = f->sample ? BM3D_ST_FLOAT : BM3D_ST_INTEGER;
        f->ovi.bits_per_sample = f->sample ? 32 : 16;
        f->ovi.subsampling_w = f->ovi.subsampling_h = 0;
        break;
    case kOPP2RGB:
        f->ovi.color_family = BM3D_CF_RGB;
        f->ovi.sample_type = f->sample ? BM3D_ST_FLOAT : BM3D_ST_INTEGER;
        f->ovi.bits_per_sample = f->sample ? 32 : 16;
        f->ovi.subsampling_w = f->ovi.subsampling_h = 0;
        break;
    case kVBasic:
    case kVFinal:
        f->ovi.color_family = f->vi.color_family == BM3D_CF_RGB ? BM3D_CF_YUV : f->vi.color_family;
        f->ovi.sample_type = BM3D_ST_FLOAT;
        f->ovi.bits_per_sample = 32;
        f->ovi.height = f->vi.height * (f->radius * 2 + 1) * 2;
        break;
    case kVAggregate:
        f->ovi.sample_type = f->sample ? BM3D_ST_FLOAT : BM3D_ST_INTEGER;
        f->ovi.bits_per_sample = f->sample ? 32 : 16;
        f->ovi.height = f->vi.height / (f->radius * 2 + 1) / 2;
        break;
    default:
        break;
    }
    // argument validation needs no device; everything below does
    if ((rc = ensure_init())) return rc;
    if (f->fn == kBasic || f->fn == kFinal || f->fn == kVBasic || f->fn == kVFinal) {
        // init_filter_data (source/BM3D_Base.cpp:262-283)
        double norms[3];
        matrix_norms(f->matrix, norms);
        f->th_mse *= norms[0];
        double sig[3];
        for (int p = 0; p < 3; ++p) sig[p] = f->sigma[p] / double(255) * norms[p];
        build_tables(f.get(), sig);
        if ((rc = upload_coefs(f.get()))) return rc;
    }
    // blocking stream: ordered after work on the legacy default stream (where
    // callers such as torch produce device inputs) and before later work on it
    HIP_TRY(hipStreamCreate(&f->stream));
    *out = f.release();
    return BM3D_OK;
}

int bm3d_output_info(const bm3d_filter* f, bm3d_clip_info* out) {
    if (!f || !out) return set_error(BM3D_E_STATE, "bm3d_output_info: null handle");
    *out = f->ovi;
    return BM3D_OK;
}

int bm3d_frame_range(const bm3d_filter* f, int n, int* first, int* last) {
    if (!f || !first || !last) return set_error(BM3D_E_STATE, "bm3d_frame_range: null handle");
    int r = (f->fn == kVBasic || f->fn == kVFinal || f->fn == kVAggregate) ? f->radius : 0;
    int N = f->vi.num_frames;
    *first = n - std::min(n, r);
    *last = N > 0 ? n + std::min(N - 1 - n, r) : n + r;
    return BM3D_OK;
}

void bm3d_free(bm3d_filter* f) { delete f; }

int bm3d_last_stats(const bm3d_filter* f, bm3d_stats* out) {
    if (!f || !out) return set_error(BM3D_E_STATE, "bm3d_last_stats: null handle");
    std::lock_guard<std::mutex> lk(f->mu);
    bm3d_stats s = f->stats;
    if (f->last_nb > 0 && f->last_count && (f->fn == kBasic || f->fn == kFinal)) {
        std::vector<int> c(f->last_nb);
        HIP_TRY(hipMemcpy(c.data(), f->last_count, sizeof(int) * f->last_nb, hipMemcpyDeviceToHost));
        s.members = s.members_sq = 0;
        for (int v : c) {
            int n = std::min(v, f->G);
            s.members += n;
            s.members_sq += (int64_t)n * n;
        }
    }
    *out = s;
    return BM3D_OK;
}

int bm3d_profile(bm3d_filter* f, int enable) {
    if (!f) return set_error(BM3D_E_STATE, "bm3d_profile: null handle");
    std::lock_guard<std::mutex> lk(f->mu);
    for (int i = 0; i < bm3d_filter::kMaxTimers; ++i) {
        f->prof_ms[i] = 0;
        f->prof_n[i] = 0;
        f->prof_pending[i] = false;
    }
    if (enable)
        for (auto& e : f->prof_ev)
            if (!e) HIP_TRY(hipEventCreate(&e));
    f->prof = enable != 0;
    return BM3D_OK;
}

int bm3d_profile_read(const bm3d_filter* f, int idx, const char** name, double* total_ms, int64_t* launches) {
    if (!f) return set_error(BM3D_E_STATE, "bm3d_profile_read: null handle");
    std::lock_guard<std::mutex> lk(f->mu);
    int k = 0;
    for (int i = 0; i < bm3d_filter::kMaxTimers; ++i) {
        if (!f->prof_name[i]) continue;
        if (k == idx) {
            if (name) *name = f->prof_name[i];
            if (total_ms) *total_ms = f->prof_ms[i];
            if (launches) *launches = f->prof_n[i];
            return BM3D_OK;
        }
        ++k;
    }
    return BM3D_E_ARG;
}

// Match list of block b of the last frame; the caller holds f->mu.
static int debug_matches_locked(const bm3d_filter* f, int64_t b, int* count, int* ys, int* xs, float* keys, int cap) {
    if (b < 0 || b >= f->last_nb || !f->last_count) return set_error(BM3D_E_ARG, "bm3d_debug_matches: no such block");
    int n = 0;
    HIP_TRY(hipMemcpy(&n, f->last_count + b, sizeof(int), hipMemcpyDeviceToHost));
    n = std::min(n, f->G);
    std::vector<int> pos(f->G);
    std::vector<float> k(f->G);
    HIP_TRY(hipMemcpy(pos.data(), f->last_pos + b * f->G, sizeof(int) * f->G, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(k.data(), f->last_key + b * f->G, sizeof(float) * f->G, hipMemcpyDeviceToHost));
    *count = n;
    for (int z = 0; z < n && z < cap; ++z) {
        ys[z] = (int)((uint32_t)pos[z] >> 16);
        xs[z] = pos[z] & 0xffff;
        keys[z] = k[z];
    }
    return BM3D_OK;
}

int bm3d_debug_matches(const bm3d_filter* f, int64_t b, int* count, int* ys, int* xs, float* keys, int cap) {
    if (!f) return set_error(BM3D_E_STATE, "bm3d_debug_matches: null handle");
    std::lock_guard<std::mutex> lk(f->mu);
    return debug_matches_locked(f, b, count, ys, xs, keys, cap);
}

int bm3d_debug_vmatches(const bm3d_filter* f, int64_t b, int* count, int* frames, int* ys, int* xs, float* keys,
                        int cap) {
    if (!f) return set_error(BM3D_E_STATE, "bm3d_debug_vmatches: null handle");
    std::lock_guard<std::mutex> lk(f->mu);
    if (!f->temporal || !f->last_frame) return set_error(BM3D_E_ARG, "bm3d_debug_vmatches: not a VBasic/VFinal filter");
    int rc = debug_matches_locked(f, b, count, ys, xs, keys, cap);
    if (rc) return rc;
    std::vector<int> fr(f->G);
    HIP_TRY(hipMemcpy(fr.data(), f->last_frame + b * f->G, sizeof(int) * f->G, hipMemcpyDeviceToHost));
    for (int z = 0; z < *count && z < cap; ++z) frames[z] = fr[z];
    return BM3D_OK;
}

int bm3d_get_frame(bm3d_filter* f, int n, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame* dst) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!f || !src || !dst) return set_error(BM3D_E_STATE, "bm3d_get_frame: null argument");
    std::lock_guard<std::mutex> lk(f->mu);
    return run_frame(f, n, src, ref, dst);
}

}  // extern "C"

namespace bm3d {

void prof_begin(bm3d_filter* f, int t, const char* name) {
    if (!f->prof) return;
    f->prof_name[t] = name;
    (void)hipEventRecord(f->prof_ev[2 * t], f->stream);
}

void prof_end(bm3d_filter* f, int t) {
    if (!f->prof) return;
    (void)hipEventRecord(f->prof_ev[2 * t + 1], f->stream);
    f->prof_pending[t] = true;
}

// Call after the stream has been synchronised.
void prof_collect(bm3d_filter* f) {
    if (!f->prof) return;
    for (int t = 0; t < bm3d_filter::kMaxTimers; ++t) {
        if (!f->prof_pending[t]) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, f->prof_ev[2 * t], f->prof_ev[2 * t + 1]) == hipSuccess) {
            f->prof_ms[t] += ms;
            f->prof_n[t] += 1;
        }
        f->prof_pending[t] = false;
    }
}

int rcp_bits() { return g_dev.rcp_bits; }
const float* rcp_dev() { return g_dev.rcp_dev; }
}  // namespace bm3d
