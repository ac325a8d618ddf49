// Per-frame dispatch: what VSProcess::process and the process_core_* /
// Kernel functions of the reference do for one output frame
// (include/Helper.h:695-725, source/BM3D_Base.cpp:290-754,
// include/RGB2OPP.h:188-209, include/OPP2RGB.h:188-209), as a sequence of
// kernel launches on the filter's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "bm3d_engine.h"
#include "bm3d_kernels.h"

namespace bm3d {

int take(Scratch* sc, DevBuf& buf, size_t bytes, void** out) {
    if (!sc) {
        int rc = buf.ensure(bytes);
        if (rc) return rc;
        *out = buf.p;
        return BM3D_OK;
    }
    const size_t at = (sc->used + 255) & ~(size_t)255;
    if (!sc->dry && at + bytes > sc->cap)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: scratch area exhausted (" + std::to_string(at + bytes) +
                                                 " bytes needed, " + std::to_string(sc->cap) + " available)");
    // the sizing pass hands out distinct non-null addresses that are never used
    *out = sc->dry ? reinterpret_cast<void*>((uintptr_t)256 + at) : sc->base + at;
    sc->used = at + bytes;
    return BM3D_OK;
}

// Builds the stage parameters of a Basic/Final filter for a W x H plane.
StageParams stage_params(const bm3d_filter* f, int W, int H, size_t pitch) {
    StageParams p{};
    p.width = W;
    p.height = H;
    p.pitch = (int)pitch;
    p.block = f->B;
    p.block_step = f->step;
    p.group = f->G;
    p.step = f->S;
    p.range = f->R / f->S * f->S;
    auto count = [](int dim, int B, int s) {
        int last = dim - B, n = 0;
        for (int j = 0;; j += s) {
            if (j >= last + s) break;
            if (j > last) j = last;
            ++n;
        }
        return n;
    };
    p.nx = count(W, f->B, f->step);
    p.ny = count(H, f->B, f->step);
    // Block::BlockMatchingMulti (include/Block.h:558-560)
    const double mse2sse = (double)(f->B * f->B) * 1.0f * 1.0f / (double)(255 * 255);
    p.dist_mul = 1.0 / mse2sse;
    p.th_sse = (float)(f->th_mse * mse2sse);
    if (f->G == 1 || f->th_mse <= 0) {
        p.th_sse = 0.0f;  // kernel takes the short-cut
        p.dist_mul = 0.0;
    }
    p.radius = f->radius;
    p.ps_num = f->ps_num;
    p.ps_range = f->ps_range;
    p.ps_step = f->ps_step;
    p.arith = f->arith;
    return p;
}

namespace {

constexpr int kPitchAlign = 64;  // floats: 256-byte aligned rows

int bytes_of(const bm3d_clip_info& vi) { return vi.sample_type == BM3D_ST_FLOAT ? 4 : (vi.bits_per_sample > 8 ? 2 : 1); }
int bits_code(const bm3d_clip_info& vi) { return vi.sample_type == BM3D_ST_FLOAT ? 0 : vi.bits_per_sample; }
int nplanes(const bm3d_clip_info& vi) { return vi.color_family == BM3D_CF_GRAY ? 1 : 3; }
int plane_w(const bm3d_clip_info& vi, int p) { return p == 0 ? vi.width : vi.width >> vi.subsampling_w; }
int plane_h(const bm3d_clip_info& vi, int p) { return p == 0 ? vi.height : vi.height >> vi.subsampling_h; }

// GetQuanPara for integer formats (include/Helper.h:285-346).
void quan_int(int bits, bool full, bool chroma, int& fl, int& ne, int& ce) {
    if (full) {
        fl = 0;
        ne = chroma ? 1 << (bits - 1) : 0;
        ce = (1 << bits) - 1;
    } else {
        fl = 16 << (bits - 8);
        ne = chroma ? 1 << (bits - 1) : fl;
        ce = chroma ? 240 << (bits - 8) : 235 << (bits - 8);
    }
}

// Int2Float coefficients: RangeConvert<float, int> (include/Conversion.hpp:36-78).
struct ToFloat {
    float gain = 1, offset = 0, lo = 0, hi = 1;
    int clip = 0;
};
ToFloat int2float_coefs(int bits, bool full, bool chroma) {
    ToFloat t;
    int sf, sn, sc;
    quan_int(bits, full, chroma, sf, sn, sc);
    const float df = chroma ? -0.5f : 0.0f, dn = chroma ? 0.0f : df, dc = chroma ? 0.5f : 1.0f;
    const int srange = sc - sf;
    const float drange = dc - df;
    t.clip = (sf < sn && (sf + sc) % 2 == 1) ? 1 : 0;  // PC-range chroma source always clips
    t.gain = drange / (float)srange;
    t.offset = -(float)sn * t.gain + dn;
    t.lo = df;
    t.hi = dc;
    return t;
}

// Float2Int coefficients: RangeConvert<int, float> with clip.
void float2int_coefs(int bits, bool full, bool chroma, float& gain, float& offset, float& lo, float& hi) {
    int df, dn, dc;
    quan_int(bits, full, chroma, df, dn, dc);
    const float sf = chroma ? -0.5f : 0.0f, sn = chroma ? 0.0f : sf, sc = chroma ? 0.5f : 1.0f;
    const float srange = sc - sf;
    const int drange = dc - df;
    const bool pc = df < dn && (df + dc) % 2 == 1;
    gain = (float)drange / srange;
    offset = -sn * gain + (float)dn;
    offset += pc ? (float)0.499999 : 0.5f;
    lo = (float)df;
    hi = (float)dc;
}

// MatrixConvert_RGB2YUV, OPP branch coefficients (Conversion.hpp:205-236).
// src_bits 0 = float; dst_bits 0 = float.
void rgb2opp_coefs(int src_bits, int dst_bits, RgbArgs& a) {
    float sf = 0.0f, sr;
    if (src_bits == 0) sr = 1.0f;
    else sr = (float)((1 << src_bits) - 1);
    const bool dst_float = dst_bits == 0;
    float dfy, dcy, dfc, dnc, dcc;
    bool pc = false;
    if (dst_float) {
        dfy = 0.0f; dcy = 1.0f; dfc = -0.5f; dnc = 0.0f; dcc = 0.5f;
    } else {
        int a0, b0, c0;
        quan_int(dst_bits, true, false, a0, b0, c0);
        dfy = (float)a0; dcy = (float)c0;
        quan_int(dst_bits, true, true, a0, b0, c0);
        dfc = (float)a0; dnc = (float)b0; dcc = (float)c0;
        pc = a0 < b0 && (a0 + c0) % 2 == 1;
    }
    const float dry = dcy - dfy, drc = dcc - dfc;
    a.gain[0] = dry / (sr * 3.0f);
    a.offset[0] = -sf * 3.0f * a.gain[0] + dfy;
    if (!dst_float) a.offset[0] += 0.5f;
    a.gain[1] = drc / (sr * 2.0f);
    a.gain[2] = drc / (sr * 4.0f);
    a.offset[1] = dnc;
    if (!dst_float) a.offset[1] += pc ? (float)0.499999 : 0.5f;
    a.lo[0] = dfy; a.hi[0] = dcy;
    a.lo[1] = dfc; a.hi[1] = dcc;
}

// MatrixConvert_YUV2RGB, OPP branch coefficients (Conversion.hpp:290-417,
// Specification.h:253-269).  m = Ry Ru Rv Gy Gu Gv By Bu Bv.
void opp2rgb_coefs(int src_bits, int dst_bits, float m[9], float off[3], float& lo, float& hi) {
    float sfy, scy, sfc, snc, scc;
    if (src_bits == 0) {
        sfy = 0.0f; scy = 1.0f; sfc = -0.5f; snc = 0.0f; scc = 0.5f;
    } else {
        int a0, b0, c0;
        quan_int(src_bits, true, false, a0, b0, c0);
        sfy = (float)a0; scy = (float)c0;
        quan_int(src_bits, true, true, a0, b0, c0);
        sfc = (float)a0; snc = (float)b0; scc = (float)c0;
    }
    const bool dst_float = dst_bits == 0;
    const float dfl = 0.0f, dcl = dst_float ? 1.0f : (float)((1 << dst_bits) - 1);
    const float sry = scy - sfy, src_ = scc - sfc, dr = dcl - dfl;
    const float gy = dr / sry, gc = dr / src_;
    float Ry = 1.0f, Ru = 1.0f, Rv = (float)(2.0L / 3.0L);
    float Gy = 1.0f, Gu = 0.0f, Gv = (float)(-4.0L / 3.0L);
    float By = 1.0f, Bu = -1.0f, Bv = (float)(2.0L / 3.0L);
    Ry *= gy; Ru *= gc; Rv *= gc;
    Gy *= gy; Gu *= gc; Gv *= gc;
    By *= gy; Bu *= gc; Bv *= gc;
    off[0] = -sfy * Ry - snc * (Ru + Rv) + dfl;
    off[1] = -sfy * Gy - snc * (Gu + Gv) + dfl;
    off[2] = -sfy * By - snc * (Bu + Bv) + dfl;
    if (!dst_float) {
        off[0] += 0.5f;
        off[1] += 0.5f;
        off[2] += 0.5f;
    }
    m[0] = Ry; m[1] = Ru; m[2] = Rv; m[3] = Gy; m[4] = Gu; m[5] = Gv; m[6] = By; m[7] = Bu; m[8] = Bv;
    lo = dfl;
    hi = dcl;
}

size_t align_pitch(int w) { return (size_t)((w + kPitchAlign - 1) / kPitchAlign) * kPitchAlign; }

// Float planes already in device memory are read in place when every plane a
// call needs shares one row stride: returns that pitch (elements), else 0.
size_t direct_float_pitch(const bm3d_clip_info& vi, const bm3d_frame* const* frames, int nframes, const bool need[3]) {
    if (vi.sample_type != BM3D_ST_FLOAT || vi.color_family == BM3D_CF_RGB) return 0;
    ptrdiff_t stride = 0;
    for (int t = 0; t < nframes; ++t)
        for (int p = 0; p < 3; ++p) {
            if (!need[p]) continue;
            const bm3d_frame& fr = *frames[t];
            if (fr.mem != 1 || fr.stride[p] <= 0 || fr.stride[p] % (ptrdiff_t)sizeof(float) != 0) return 0;
            if (stride == 0) stride = fr.stride[p];
            if (fr.stride[p] != stride) return 0;
        }
    if (stride == 0 || (size_t)stride / sizeof(float) < (size_t)vi.width) return 0;
    return (size_t)stride / sizeof(float);
}

// Returns a device pointer + element pitch for plane p of frame fr, staging
// host memory into scratch (buffer buf of the call) when needed.
int device_plane(const bm3d_frame& fr, int p, int w, int h, int bps, DevBuf& buf, hipStream_t st,
                 const void*& ptr, size_t& pitch_elems, Scratch* sc = nullptr) {
    if (fr.mem == 1) {
        ptr = fr.data[p];
        pitch_elems = (size_t)fr.stride[p] / bps;
        return BM3D_OK;
    }
    void* d = nullptr;
    int rc = take(sc, buf, (size_t)w * h * bps, &d);
    if (rc) return rc;
    if (!(sc && sc->dry))
        HIP_TRY(hipMemcpy2DAsync(d, (size_t)w * bps, fr.data[p], fr.stride[p], (size_t)w * bps, h,
                                 hipMemcpyHostToDevice, st));
    ptr = d;
    pitch_elems = w;
    return BM3D_OK;
}

bool full_range(const bm3d_props& pr) {
    // NewFrame (include/BM3D_Base.h:189-223): BM3D_OPP=1 or no _Range => full;
    // otherwise full = (_Range != 0), the inverse of VapourSynth's convention.
    if (pr.bm3d_opp == 1 || pr.range < 0) return true;
    return pr.range != 0;
}

// Copies an unprocessed plane through (the reference leaves it uninitialised).
int copy_plane(const bm3d_frame& s, bm3d_frame& d, int p, int w, int h, int bps, hipStream_t st) {
    HIP_TRY(hipMemcpy2DAsync(d.data[p], d.stride[p], s.data[p], s.stride[p], (size_t)w * bps, h, hipMemcpyDefault, st));
    return BM3D_OK;
}

int run_color(bm3d_filter* f, const bm3d_frame& s, bm3d_frame& d) {
    const bm3d_clip_info& vi = f->vi;
    const int W = vi.width, H = vi.height;
    const int sb = bytes_of(vi);
    const int dbits = f->sample ? 0 : 16, db = f->sample ? 4 : 2;
    RgbArgs a{};
    int rc;
    for (int p = 0; p < 3; ++p) {
        size_t pe;
        if ((rc = device_plane(s, p, W, H, sb, f->in_raw[0][p], f->stream, a.src[p], pe))) return rc;
        a.spitch = pe;
    }
    const bool dev_out = d.mem == 1;
    for (int p = 0; p < 3; ++p) {
        if (dev_out) {
            a.dst[p] = d.data[p];
            a.dpitch = (size_t)d.stride[p] / db;
        } else {
            if ((rc = f->out_stage[p].ensure((size_t)W * H * db))) return rc;
            a.dst[p] = f->out_stage[p].p;
            a.dpitch = W;
        }
    }
    a.src_bits = bits_code(vi);
    a.dst_bits = dbits;
    a.W = W;
    a.H = H;
    a.clip = dbits != 0;
    if (f->fn == kRGB2OPP) {
        rgb2opp_coefs(a.src_bits, dbits, a);
        prof_begin(f, kTConvert, "k_rgb2opp");
        HIP_TRY(launch_rgb2opp(a, f->stream));
        prof_end(f, kTConvert);
        d.props = s.props;
        d.props.matrix = 2;
        d.props.bm3d_opp = 1;
    } else {
        float lo, hi;
        opp2rgb_coefs(a.src_bits, dbits, a.m, a.offset, lo, hi);
        a.lo[0] = lo;
        a.hi[0] = hi;
        prof_begin(f, kTConvert, "k_opp2rgb");
        HIP_TRY(launch_opp2rgb(a, f->stream));
        prof_end(f, kTConvert);
        d.props = s.props;
        d.props.matrix = 0;
        d.props.bm3d_opp = -1;
    }
    if (!dev_out)
        for (int p = 0; p < 3; ++p)
            HIP_TRY(hipMemcpy2DAsync(d.data[p], d.stride[p], a.dst[p], (size_t)W * db, (size_t)W * db, H,
                                     hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    prof_collect(f);
    return BM3D_OK;
}


// Exact number of SSD evaluations of a frame: every lattice position of
// every reference block's search window except the block itself
// (include/Block.h:683-707).
int64_t ssd_count(const StageParams& p) {
    auto axis = [&](int pos, int last) {
        int lo = pos - std::min(pos / p.step * p.step, p.range);
        int hi = pos + std::min((last - pos) / p.step * p.step, p.range);
        return (int64_t)((hi - lo) / p.step + 1);
    };
    std::vector<int64_t> cy(p.ny), cx(p.nx);
    for (int a = 0; a < p.ny; ++a) cy[a] = axis(std::min(a * p.block_step, p.height - p.block), p.height - p.block);
    for (int c = 0; c < p.nx; ++c) cx[c] = axis(std::min(c * p.block_step, p.width - p.block), p.width - p.block);
    int64_t sy = 0, sx = 0;
    for (auto v : cy) sy += v;
    for (auto v : cx) sx += v;
    if (p.th_sse <= 0.0f) return 0;
    return sy * sx - (int64_t)p.ny * p.nx;
}

// sc: where the call's scratch comes from (null: the filter's own buffers).
int run_bm3d(bm3d_filter* f, const bm3d_frame& s, const bm3d_frame* r, bm3d_frame& d, Scratch* sc) {
    const bool dry = sc && sc->dry;
    const bm3d_clip_info& vi = f->vi;
    const int W = vi.width, H = vi.height;
    const int bps = bytes_of(vi), bits = bits_code(vi);
    const bool rgb = vi.color_family == BM3D_CF_RGB;
    const bool isfloat = vi.sample_type == BM3D_ST_FLOAT;
    const int P = nplanes(vi);
    const bool gray_path = vi.color_family == BM3D_CF_GRAY || (vi.color_family == BM3D_CF_YUV && !f->process[1] && !f->process[2]);
    const bool full = rgb ? true : full_range(s.props);
    if (!dry && s.props.bm3d_opp == 1 && !rgb && f->matrix != 100)
        log_warning(std::string(f->wiener ? "bm3d.Basic/bm3d.Final" : "bm3d.Basic/bm3d.Final") +
                    " - warning: There's a frame property \"BM3D_OPP=1\" indicating opponent color space input. "
                    "You should specify \"matrix=100\" in the filter's argument.");
    size_t P_ = align_pitch(W);
    const hipStream_t st = f->stream;
    int rc;
    int proc[3] = {f->process[0], gray_path ? 0 : f->process[1], gray_path ? 0 : f->process[2]};
    if (P == 1) proc[1] = proc[2] = 0;
    if (!rgb && !proc[0] && !proc[1] && !proc[2]) {
        // VSProcess::process returns the source frame untouched when no plane
        // is processed (include/Helper.h:697-700)
        if (dry) return BM3D_OK;
        for (int p = 0; p < P; ++p)
            if ((rc = copy_plane(s, d, p, plane_w(vi, p), plane_h(vi, p), bps, st))) return rc;
        d.props = s.props;
        HIP_TRY(hipStreamSynchronize(st));
        return BM3D_OK;
    }

    // --- inputs as float planes (process_core_{gray,yuv,rgb})
    const float* S[3] = {nullptr, nullptr, nullptr};
    const float* Rf[3] = {nullptr, nullptr, nullptr};
    const bm3d_frame& rf = f->rdef ? *r : s;
    if (rgb) {
        for (int which = 0; which < (f->rdef ? 2 : 1); ++which) {
            const bm3d_frame& fr = which ? rf : s;
            RgbArgs a{};
            for (int p = 0; p < 3; ++p) {
                size_t pe;
                if ((rc = device_plane(fr, p, W, H, bps, f->in_raw[which][p], st, a.src[p], pe, sc))) return rc;
                a.spitch = pe;
                if ((rc = take(sc, f->fplane[which][p], P_ * H * sizeof(float), &a.dst[p]))) return rc;
            }
            a.dpitch = P_;
            a.src_bits = bits;
            a.dst_bits = 0;
            a.W = W;
            a.H = H;
            a.clip = 0;
            a.y_only = (which == 1 && !f->wiener) ? 1 : 0;  // RGB2FloatY for a Basic ref
            rgb2opp_coefs(bits, 0, a);
            if (!dry) HIP_TRY(launch_rgb2opp(a, st));
            for (int p = 0; p < 3; ++p) (which ? Rf : S)[p] = static_cast<const float*>(a.dst[p]);
        }
        if (!f->rdef)
            for (int p = 0; p < 3; ++p) Rf[p] = S[p];
    } else {
        bool need_w[2][3];
        for (int p = 0; p < 3; ++p) {
            need_w[0][p] = p < P && (proc[p] || (p == 0 && !f->rdef));
            need_w[1][p] = p < P && (p == 0 || (f->wiener && proc[p]));
        }
        // float planes already in device memory with one common stride are read in place
        const bm3d_frame* fs = &s;
        const bm3d_frame* fr0 = &rf;
        const size_t ps = direct_float_pitch(vi, &fs, 1, need_w[0]);
        const size_t pr = f->rdef ? direct_float_pitch(vi, &fr0, 1, need_w[1]) : ps;
        const bool direct = ps != 0 && ps == pr;
        if (direct) {
            P_ = ps;
            for (int p = 0; p < 3; ++p) {
                if (need_w[0][p]) S[p] = static_cast<const float*>(s.data[p]);
                if (f->rdef && need_w[1][p]) Rf[p] = static_cast<const float*>(rf.data[p]);
            }
        }
        for (int which = 0; which < (f->rdef ? 2 : 1) && !direct; ++which) {
            const bm3d_frame& fr = which ? rf : s;
            for (int p = 0; p < P; ++p) {
                const bool need = need_w[which][p];
                if (!need) continue;
                const void* ptr;
                size_t pe;
                if ((rc = device_plane(fr, p, W, H, bps, f->in_raw[which][p], st, ptr, pe, sc))) return rc;
                float* dstp = nullptr;
                if ((rc = take_as(sc, f->fplane[which][p], P_ * H * sizeof(float), &dstp))) return rc;
                if (dry) {
                } else if (isfloat) {
                    HIP_TRY(launch_to_float(ptr, 0, pe, dstp, P_, W, H, 1.0f, 0.0f, 0, 0.0f, 1.0f, st));
                } else {
                    ToFloat t = int2float_coefs(bits, full, p > 0);
                    HIP_TRY(launch_to_float(ptr, bits, pe, dstp, P_, W, H, t.gain, t.offset, t.clip, t.lo, t.hi, st));
                }
                (which ? Rf : S)[p] = dstp;
            }
        }
        if (!f->rdef)
            for (int p = 0; p < 3; ++p) Rf[p] = S[p];
    }

    // --- block matching on the reference luma
    StageParams sp = stage_params(f, W, H, P_);
    const int nb = sp.nx * sp.ny;
    const int G = f->G, B = f->B;
    const int side = 2 * (sp.range / sp.step) + 1;
    const int max_cand = side * side;
    int *m_count = nullptr, *m_pos = nullptr;
    float* m_key = nullptr;
    if ((rc = take_as(sc, f->m_count, sizeof(int) * nb, &m_count))) return rc;
    if ((rc = take_as(sc, f->m_pos, sizeof(int) * (size_t)nb * G, &m_pos))) return rc;
    if ((rc = take_as(sc, f->m_key, sizeof(float) * (size_t)nb * G, &m_key))) return rc;
    if (match_lds_bytes(max_cand, B, G) > 65536)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: search window too large for the matching kernel");
    if (sc && f->lbm)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: the learned matcher has no fixed-scratch form");
    if (!dry) {
        prof_begin(f, kTMatch, f->wiener ? "k_match(final)" : "k_match(basic)");
        if (f->lbm) {
            if ((rc = lbm_match(f->lbm, f->lbm_topk, f->lbm_work, Rf[0], sp, m_count, m_pos, m_key, st))) return rc;
        } else {
            HIP_TRY(launch_match(Rf[0], sp, nb, max_cand, m_count, m_pos, m_key, st));
        }
        prof_end(f, kTMatch);
        f->last_nb = nb;
        f->last_count = m_count;
        f->last_pos = m_pos;
        f->last_key = m_key;
        f->last_frame = nullptr;
    }

    // --- collaborative filtering
    FilterArgs fa{};
    fa.p = sp;
    fa.b0 = 0;
    fa.nb = nb;
    fa.wiener = f->wiener;
    for (int p = 0; p < 3; ++p) fa.process[p] = proc[p];
    fa.count = m_count;
    fa.pos = m_pos;
    fa.frame = nullptr;
    // device pointer table: [src p0..p2][ref p0..p2]
    const float** dp = nullptr;
    if ((rc = take_as(sc, f->ptrs, sizeof(float*) * 6, &dp))) return rc;
    // the table's source must outlive the asynchronous copy
    f->ptrs_host[0] = S[0]; f->ptrs_host[1] = S[1]; f->ptrs_host[2] = S[2];
    f->ptrs_host[3] = Rf[0]; f->ptrs_host[4] = Rf[1]; f->ptrs_host[5] = Rf[2];
    if (!dry) HIP_TRY(hipMemcpyAsync(dp, f->ptrs_host, sizeof(float*) * 6, hipMemcpyHostToDevice, st));
    for (int p = 0; p < 3; ++p) {
        fa.src[p] = dp + p;
        fa.ref[p] = dp + 3 + p;
        if (!proc[p]) continue;
        if ((size_t)nb * G >= ((size_t)1 << 31))
            return set_error(BM3D_E_UNSUPPORTED, "bm3d: reference blocks x group_size must stay below 2^31 per frame");
        if ((rc = take_as(sc, f->contrib[p], sizeof(float) * (size_t)nb * G * B * B, &fa.contrib[p]))) return rc;
        if ((rc = take_as(sc, f->denw[p], sizeof(float) * nb, &fa.denw[p]))) return rc;
    }
    fa.tables = f->tables_dev.as<PlaneTables>();
    fa.rcp_table = rcp_dev();
    fa.rcp_shift = 23 - rcp_bits();
    fa.coef = f->coef_dev.as<float>();
    fa.coef_off = f->coef_off_dev.as<int>();
    if (!filter_fits(B, G, f->wiener))
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: block_size/group_size too large for the on-chip group buffer");
    int* order = nullptr;
    if (filter_uses_buckets(B, G))
        if ((rc = take_as(sc, f->m_order, sizeof(int) * (size_t)nb, &order))) return rc;
    if (!dry) {
        prof_begin(f, kTFilter, f->wiener ? "k_filter(wiener)" : "k_filter(hard)");
        if (order) {
            HIP_TRY(launch_bucket(fa.count, nb, G, order, st));
            fa.order = order;
        }
        HIP_TRY(launch_filter(fa, st));
        prof_end(f, kTFilter);
    }

    // --- aggregation + divide + output conversion
    AggArgs ag{};
    ag.p = sp;
    for (int p = 0; p < 3; ++p) {
        ag.process[p] = proc[p];
        ag.contrib[p] = fa.contrib[p];
        ag.denw[p] = fa.denw[p];
    }
    ag.reach = sp.range;
    ag.count = fa.count;
    ag.pos = fa.pos;
    OutputSpec& o = ag.out;
    const bool dev_out = d.mem == 1;
    const int obits = bits;
    const int ob = bps;
    for (int p = 0; p < 3; ++p) {
        bool used = rgb || proc[p];
        if (!used) continue;
        if (dev_out) {
            o.dst[p] = d.data[p];
            o.pitch[p] = (size_t)d.stride[p] / ob;
        } else {
            if ((rc = take(sc, f->out_stage[p], (size_t)W * H * ob, &o.dst[p]))) return rc;
            o.pitch[p] = W;
        }
    }
    if (dry) return BM3D_OK;
    o.bits = obits;
    if (rgb) {
        o.mode = kOutRGB;
        float lo, hi;
        opp2rgb_coefs(0, obits, o.m, o.offset, lo, hi);
        o.lo[0] = lo;
        o.hi[0] = hi;
        o.clip = obits != 0;
    } else if (isfloat) {
        o.mode = kOutFloat;
    } else {
        o.mode = kOutInt;
        o.clip = 1;
        for (int p = 0; p < 3; ++p) float2int_coefs(obits, full, p > 0, o.gain[p], o.offset[p], o.lo[p], o.hi[p]);
    }
    prof_begin(f, kTAggregate, f->wiener ? "k_aggregate(final)" : "k_aggregate(basic)");
    HIP_TRY(launch_aggregate(ag, st));
    prof_end(f, kTAggregate);

    // --- outputs
    for (int p = 0; p < P; ++p) {
        const int pw = plane_w(vi, p), ph = plane_h(vi, p);
        bool written = rgb || proc[p];
        if (written && !dev_out) {
            HIP_TRY(hipMemcpy2DAsync(d.data[p], d.stride[p], o.dst[p], (size_t)W * ob, (size_t)W * ob, H,
                                     hipMemcpyDeviceToHost, st));
        } else if (!written) {
            if ((rc = copy_plane(s, d, p, pw, ph, bps, st))) return rc;
        }
    }
    d.props = s.props;
    HIP_TRY(hipStreamSynchronize(st));
    prof_collect(f);
    f->stats.ref_blocks = nb;
    f->stats.ssd_evals = ssd_count(sp);
    f->last_W = W;
    f->last_H = H;
    return BM3D_OK;
}

// Converts the planes of one input frame to float the way process_core_*
// does (source/BM3D_Base.cpp:497-754, source/VBM3D_Base.cpp:666-1039).
// which: 0 = src, 1 = ref.  slot: frame index inside the window.  need[p]
// tells which planes are required.  Results go to out[p].
int convert_frame(bm3d_filter* f, const bm3d_frame& fr, int which, int slot, const bool need[3], bool full,
                  size_t P_, const float* out[3]) {
    const bm3d_clip_info& vi = f->vi;
    const int W = vi.width, H = vi.height;
    const int bps = bytes_of(vi), bits = bits_code(vi);
    const hipStream_t st = f->stream;
    int rc;
    if (vi.color_family == BM3D_CF_RGB) {
        RgbArgs a{};
        for (int p = 0; p < 3; ++p) {
            size_t pe;
            if ((rc = device_plane(fr, p, W, H, bps, f->in_raw[which][3 * slot + p], st, a.src[p], pe))) return rc;
            a.spitch = pe;
            if ((rc = f->fplane[which][3 * slot + p].ensure(P_ * H * sizeof(float)))) return rc;
            a.dst[p] = f->fplane[which][3 * slot + p].p;
            out[p] = f->fplane[which][3 * slot + p].as<float>();
        }
        a.dpitch = P_;
        a.src_bits = bits;
        a.dst_bits = 0;
        a.W = W;
        a.H = H;
        a.clip = 0;
        a.y_only = (need[0] && !need[1] && !need[2]) ? 1 : 0;
        rgb2opp_coefs(bits, 0, a);
        HIP_TRY(launch_rgb2opp(a, st));
        return BM3D_OK;
    }
    for (int p = 0; p < nplanes(vi); ++p) {
        if (!need[p]) continue;
        const void* ptr;
        size_t pe;
        if ((rc = device_plane(fr, p, W, H, bps, f->in_raw[which][3 * slot + p], st, ptr, pe))) return rc;
        if ((rc = f->fplane[which][3 * slot + p].ensure(P_ * H * sizeof(float)))) return rc;
        float* dstp = f->fplane[which][3 * slot + p].as<float>();
        if (bits == 0) {
            HIP_TRY(launch_to_float(ptr, 0, pe, dstp, P_, W, H, 1.0f, 0.0f, 0, 0.0f, 1.0f, st));
        } else {
            ToFloat t = int2float_coefs(bits, full, p > 0);
            HIP_TRY(launch_to_float(ptr, bits, pe, dstp, P_, W, H, t.gain, t.offset, t.clip, t.lo, t.hi, st));
        }
        out[p] = dstp;
    }
    return BM3D_OK;
}

// VBM3D_Process_Base::process_core_* + Kernel (source/VBM3D_Base.cpp:410-1039):
// the output frame is the float num/den stack of the window's frames.
int run_vbm3d(bm3d_filter* f, int n, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame& d) {
    const bm3d_clip_info& vi = f->vi;
    const int W = vi.width, H = vi.height;
    const bool rgb = vi.color_family == BM3D_CF_RGB;
    const int P = nplanes(vi);
    int first, last;
    bm3d_frame_range(f, n, &first, &last);
    const int T = last - first + 1, cur = n - first;
    if (T > kMaxStackFrames) return set_error(BM3D_E_UNSUPPORTED, "bm3d: temporal window too large");
    const bm3d_frame& s = src[cur];
    const bool gray_path = vi.color_family == BM3D_CF_GRAY || (vi.color_family == BM3D_CF_YUV && !f->process[1] && !f->process[2]);
    const bool full = rgb ? true : full_range(s.props);
    if (s.props.bm3d_opp == 1 && !rgb && f->matrix != 100)
        log_warning("bm3d.VBasic/bm3d.VFinal - warning: There's a frame property \"BM3D_OPP=1\" indicating opponent "
                    "color space input. You should specify \"matrix=100\" in the filter's argument.");
    int proc[3] = {f->process[0], gray_path ? 0 : f->process[1], gray_path ? 0 : f->process[2]};
    if (P == 1) proc[1] = proc[2] = 0;
    if (rgb) proc[0] = proc[1] = proc[2] = 1;
    size_t P_ = align_pitch(W);
    const hipStream_t st = f->stream;
    int rc;
    // --- float planes of every frame of the window
    std::vector<const float*> S(3 * T, nullptr), Rf(3 * T, nullptr);
    bool need_s0[3], need_r0[3];
    for (int p = 0; p < 3; ++p) {
        need_s0[p] = p < P && (proc[p] || (p == 0 && !f->rdef));
        need_r0[p] = p < P && (p == 0 || (f->wiener && proc[p]));
    }
    size_t direct = 0;  // float device planes read in place
    {
        std::vector<const bm3d_frame*> fs(T), fr(T);
        for (int t = 0; t < T; ++t) {
            fs[t] = &src[t];
            fr[t] = f->rdef ? &ref[t] : &src[t];
        }
        const size_t ps = direct_float_pitch(vi, fs.data(), T, need_s0);
        const size_t pr = f->rdef ? direct_float_pitch(vi, fr.data(), T, need_r0) : ps;
        if (ps != 0 && ps == pr) direct = ps;
    }
    if (direct) {
        P_ = direct;
        for (int t = 0; t < T; ++t)
            for (int p = 0; p < 3; ++p) {
                if (need_s0[p]) S[3 * t + p] = static_cast<const float*>(src[t].data[p]);
                if (f->rdef) {
                    if (need_r0[p]) Rf[3 * t + p] = static_cast<const float*>(ref[t].data[p]);
                } else {
                    Rf[3 * t + p] = S[3 * t + p];
                }
            }
    }
    for (int t = 0; t < T && !direct; ++t) {
        bool need_s[3], need_r[3];
        for (int p = 0; p < 3; ++p) {
            need_s[p] = need_s0[p];
            need_r[p] = need_r0[p];
        }
        if (rgb) need_s[0] = need_s[1] = need_s[2] = true;
        const float* so[3] = {nullptr, nullptr, nullptr};
        if ((rc = convert_frame(f, src[t], 0, t, need_s, full, P_, so))) return rc;
        for (int p = 0; p < 3; ++p) S[3 * t + p] = so[p];
        if (f->rdef) {
            const float* ro[3] = {nullptr, nullptr, nullptr};
            if (rgb && !f->wiener) need_r[1] = need_r[2] = false;
            if ((rc = convert_frame(f, ref[t], 1, t, need_r, full, P_, ro))) return rc;
            for (int p = 0; p < 3; ++p) Rf[3 * t + p] = ro[p];
        } else {
            for (int p = 0; p < 3; ++p) Rf[3 * t + p] = so[p];
        }
    }
    // --- matching
    StageParams sp = stage_params(f, W, H, P_);
    const int nb = sp.nx * sp.ny, G = f->G, B = f->B;
    const int side = 2 * (sp.range / sp.step) + 1;
    const int ps_range = f->ps_range / f->ps_step * f->ps_step;
    const int ps_side = 2 * (ps_range / f->ps_step) + 1;
    VMatchArgs va{};
    va.p = sp;
    va.nb = nb;
    va.nframes = T;
    va.cur = cur;
    va.max_cand = side * side;
    va.ps_cap = ps_side * ps_side * std::max(1, std::min(f->ps_num, G));
    va.cat_cap = G * T;
    if (vmatch_lds_bytes(sp, va.max_cand, va.ps_cap, va.cat_cap) > 65536)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: search window too large for the V-BM3D matching kernel");
    for (int t = 0; t < T; ++t) va.planes[t] = Rf[3 * t];
    if ((rc = f->m_count.ensure(sizeof(int) * nb))) return rc;
    if ((rc = f->m_pos.ensure(sizeof(int) * (size_t)nb * G))) return rc;
    if ((rc = f->m_key.ensure(sizeof(float) * (size_t)nb * G))) return rc;
    if ((rc = f->m_frame.ensure(sizeof(int) * (size_t)nb * G))) return rc;
    va.count = f->m_count.as<int>();
    va.pos = f->m_pos.as<int>();
    va.frame = f->m_frame.as<int>();
    va.key = f->m_key.as<float>();
    prof_begin(f, kTMatch, f->wiener ? "k_vmatch(final)" : "k_vmatch(basic)");
    HIP_TRY(launch_vmatch(va, st));
    prof_end(f, kTMatch);
    f->last_nb = nb;
    f->last_count = va.count;
    f->last_pos = va.pos;
    f->last_key = va.key;
    f->last_frame = va.frame;
    // --- collaborative filtering over frames
    FilterArgs fa{};
    fa.p = sp;
    fa.nb = nb;
    fa.wiener = f->wiener;
    for (int p = 0; p < 3; ++p) fa.process[p] = proc[p];
    fa.count = va.count;
    fa.pos = va.pos;
    fa.frame = va.frame;
    if ((rc = f->ptrs.ensure(sizeof(float*) * 6 * T))) return rc;
    std::vector<const float*> hp(6 * T);
    for (int p = 0; p < 3; ++p)
        for (int t = 0; t < T; ++t) {
            hp[p * T + t] = S[3 * t + p];
            hp[3 * T + p * T + t] = Rf[3 * t + p];
        }
    HIP_TRY(hipMemcpyAsync(f->ptrs.p, hp.data(), sizeof(float*) * 6 * T, hipMemcpyHostToDevice, st));
    const float* const* dp = f->ptrs.as<const float* const>();
    for (int p = 0; p < 3; ++p) {
        fa.src[p] = dp + p * T;
        fa.ref[p] = dp + 3 * T + p * T;
        if (!proc[p]) continue;
        if ((size_t)nb * G >= ((size_t)1 << 31))
            return set_error(BM3D_E_UNSUPPORTED, "bm3d: reference blocks x group_size must stay below 2^31 per frame");
        if ((rc = f->contrib[p].ensure(sizeof(float) * (size_t)nb * G * B * B))) return rc;
        if ((rc = f->denw[p].ensure(sizeof(float) * nb))) return rc;
        fa.contrib[p] = f->contrib[p].as<float>();
        fa.denw[p] = f->denw[p].as<float>();
    }
    fa.tables = f->tables_dev.as<PlaneTables>();
    fa.rcp_table = rcp_dev();
    fa.rcp_shift = 23 - rcp_bits();
    fa.coef = f->coef_dev.as<float>();
    fa.coef_off = f->coef_off_dev.as<int>();
    if (!filter_fits(B, G, f->wiener))
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: block_size/group_size too large for the on-chip group buffer");
    prof_begin(f, kTFilter, f->wiener ? "k_filter(wiener)" : "k_filter(hard)");
    if (filter_uses_buckets(B, G)) {
        if ((rc = f->m_order.ensure(sizeof(int) * (size_t)nb))) return rc;
        HIP_TRY(launch_bucket(fa.count, nb, G, f->m_order.as<int>(), st));
        fa.order = f->m_order.as<int>();
    }
    HIP_TRY(launch_filter(fa, st));
    prof_end(f, kTFilter);
    // --- aggregation into the stack
    const int slots = (2 * f->radius + 1) * 2;
    const bool dev_out = d.mem == 1;
    VStackArgs vs{};
    vs.p = sp;
    vs.nframes = T;
    vs.cur = cur;
    vs.count = va.count;
    vs.pos = va.pos;
    vs.frame = va.frame;
    for (int p = 0; p < 3; ++p) {
        vs.process[p] = proc[p];
        vs.contrib[p] = fa.contrib[p];
        vs.denw[p] = fa.denw[p];
        if (!proc[p]) continue;
        if (dev_out) {
            vs.dst[p] = static_cast<float*>(d.data[p]);
            vs.dst_pitch[p] = (size_t)d.stride[p] / sizeof(float);
        } else {
            if ((rc = f->out_stage[p].ensure((size_t)W * H * slots * sizeof(float)))) return rc;
            vs.dst[p] = f->out_stage[p].as<float>();
            vs.dst_pitch[p] = W;
        }
    }
    if (B * B <= 64) {
        // per-frame member lists: each frame slot's walk reads only its members
        if ((rc = f->m_vroff.ensure(sizeof(int) * vframe_rows_words(sp, T)))) return rc;
        if ((rc = f->m_vpair.ensure(sizeof(int2) * vframe_pairs(sp, T)))) return rc;
        vs.roff = f->m_vroff.as<int>();
        vs.rpair = f->m_vpair.as<int2>();
    }
    prof_begin(f, kTAggregate, f->wiener ? "k_vstack(final)" : "k_vstack(basic)");
    HIP_TRY(launch_vstack(vs, st));
    prof_end(f, kTAggregate);
    if (!dev_out) {
        const int s0 = 2 * (f->radius - cur);
        for (int p = 0; p < 3; ++p) {
            if (!proc[p]) continue;
            char* hd = static_cast<char*>(d.data[p]) + (size_t)s0 * H * d.stride[p];
            const float* sd = vs.dst[p] + (size_t)s0 * H * W;
            HIP_TRY(hipMemcpy2DAsync(hd, d.stride[p], sd, (size_t)W * sizeof(float), (size_t)W * sizeof(float),
                                     (size_t)2 * T * H, hipMemcpyDeviceToHost, st));
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    prof_collect(f);
    f->stats.ref_blocks = nb;
    d.props = s.props;
    if (rgb) d.props.bm3d_opp = 1;
    d.props.v_radius = f->radius;
    for (int p = 0; p < 3; ++p) d.props.v_process[p] = f->process[p];
    return BM3D_OK;
}

// VAggregate_Process (source/VAggregate.cpp:93-334, include/VAggregate.h:98-231).
int run_vaggregate(bm3d_filter* f, int n, const bm3d_frame* src, bm3d_frame& d) {
    const bm3d_clip_info& vi = f->vi;
    const int W = vi.width;
    const int r = f->radius;
    const int H = vi.height / (r * 2 + 1) / 2;
    const int P = nplanes(vi);
    int first, last;
    bm3d_frame_range(f, n, &first, &last);
    const int T = last - first + 1, cur = n - first;
    if (T > kMaxStackFrames) return set_error(BM3D_E_UNSUPPORTED, "bm3d: temporal window too large");
    const bm3d_frame& s = src[cur];
    if (s.props.v_radius < 0)
        log_warning("bm3d.VAggregate - warning: No frame property \"BM3D_V_radius\" exists in the input frame. "
                    "Make sure you call bm3d.VAggregate next to bm3d.VBasic or bm3d.VFinal.");
    else if (s.props.v_radius != r)
        log_warning("bm3d.VAggregate - warning: Mismatch between argument \"radius=" + std::to_string(r) +
                    "\" and the input frame property \"BM3D_V_radius=" + std::to_string(s.props.v_radius) +
                    "\" which indicates the radius used in previous filter (bm3d.VBasic or bm3d.VFinal).");
    int proc[3] = {1, 1, 1};
    if (s.props.v_process[0] < 0 || s.props.v_process[1] < 0 || s.props.v_process[2] < 0) {
        log_warning("bm3d.VAggregate - warning: No frame property \"BM3D_V_process\" exists in the input frame. "
                    "Make sure you call bm3d.VAggregate next to bm3d.VBasic or bm3d.VFinal.");
    } else {
        for (int p = 0; p < 3; ++p) proc[p] = (int)s.props.v_process[p];
    }
    const bool gray_path = vi.color_family == BM3D_CF_GRAY || (!proc[1] && !proc[2]);
    if (P == 1 || gray_path) proc[1] = proc[2] = 0;
    const bool full = s.props.bm3d_opp == 1 || s.props.range < 0 ? true : s.props.range != 0;
    const hipStream_t st = f->stream;
    int rc;
    VAggArgs a{};
    a.W = W;
    a.H = H;
    a.nframes = T;
    a.cur = cur;
    a.radius = r;
    const int out_bits = f->sample ? 0 : 16;
    const int ob = f->sample ? 4 : 2;
    a.out_bits = out_bits;
    const bool dev_out = d.mem == 1;
    for (int p = 0; p < 3; ++p) {
        a.process[p] = proc[p];
        if (!proc[p]) continue;
        for (int t = 0; t < T; ++t) {
            const void* ptr;
            size_t pe;
            if ((rc = device_plane(src[t], p, W, vi.height, 4, f->in_raw[0][3 * t + p], st, ptr, pe))) return rc;
            a.stacks[p * kMaxStackFrames + t] = static_cast<const float*>(ptr);
            a.stack_pitch[p] = pe;
        }
        if (out_bits) float2int_coefs(16, full, p > 0, a.gain[p], a.offset[p], a.lo[p], a.hi[p]);
        if (dev_out) {
            a.dst[p] = d.data[p];
            a.dst_pitch[p] = (size_t)d.stride[p] / ob;
        } else {
            if ((rc = f->out_stage[p].ensure((size_t)W * H * ob))) return rc;
            a.dst[p] = f->out_stage[p].p;
            a.dst_pitch[p] = W;
        }
    }
    prof_begin(f, kTVAggregate, "k_vagg");
    HIP_TRY(launch_vagg(a, st));
    prof_end(f, kTVAggregate);
    if (!dev_out)
        for (int p = 0; p < 3; ++p)
            if (proc[p])
                HIP_TRY(hipMemcpy2DAsync(d.data[p], d.stride[p], a.dst[p], (size_t)W * ob, (size_t)W * ob, H,
                                         hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    prof_collect(f);
    d.props = s.props;
    d.props.v_radius = -1;
    d.props.v_process[0] = d.props.v_process[1] = d.props.v_process[2] = -1;
    return BM3D_OK;
}

}  // namespace

int run_bm3d_scratch(bm3d_filter* f, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame* dst, Scratch* sc) {
    return run_bm3d(f, *src, ref, *dst, sc);
}

int run_frame(bm3d_filter* f, int n, const bm3d_frame* src, const bm3d_frame* ref, bm3d_frame* dst) {
    (void)n;
    switch (f->fn) {
    case kRGB2OPP:
    case kOPP2RGB:
        return run_color(f, src[0], *dst);
    case kBasic:
    case kFinal:
        return run_bm3d(f, src[0], ref, *dst, nullptr);
    case kVBasic:
    case kVFinal:
        return run_vbm3d(f, n, src, ref, *dst);
    case kVAggregate:
        return run_vaggregate(f, n, src, *dst);
    default:
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: function not implemented yet");
    }
}

}  // namespace bm3d
