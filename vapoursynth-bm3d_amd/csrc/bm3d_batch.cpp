// Batched Basic -> Final over many Gray float frames of uneven sizes.
//
// Each frame keeps its own Basic and Final filter (created by bm3d_create
// with the frame's sigma, so its thresholds, matrix tables and shrinkage
// tables are exactly those of a single-frame call, source/BM3D_Base.cpp:
// 106-236).  The work of all frames is numbered frame by frame -- reference
// blocks for matching and filtering, 32x16 output tiles for aggregation --
// and each stage is one launch over all frames whose waves look up the frame of
// the item they take (starts[] arrays in device memory).  The per-frame
// argument blocks (plane pointers, stage parameters, output planes) live in
// one device blob that is rewritten before a run only when the caller's
// planes change, so a captured HIP graph of the chain stays valid.
//
// Shapes the 8x8 specialised kernels do not cover (block_size != 8,
// bm_step != 1, group_size > 32) run the same chain with one single-frame
// launch per frame and stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bm3d_batch.h"
#include "bm3d_engine.h"
#include "bm3d_kernels.h"

using namespace bm3d;

namespace {

struct FilterDel {
    void operator()(bm3d_filter* f) const { bm3d_free(f); }
};
using FilterPtr = std::unique_ptr<bm3d_filter, FilterDel>;

// One stage (Basic or Final) of the batch.
struct Stage {
    std::vector<FilterPtr> filt;        // per frame
    std::vector<StageParams> sp;        // per frame
    std::vector<int> start;             // first global block of each frame (nframes + 1)
    int total = 0;                      // reference blocks of the batch
    int G = 0, B = 0, max_cand = 0;
    bool wiener = false;
    bool batched = false;               // the 8x8 batched kernels apply
    // offsets (bytes) inside the device blob
    size_t off_start = 0, off_frames = 0, off_ff = 0, off_ag = 0, off_ptrs = 0;
    std::vector<BatchFrame> frames;     // host images of the blob sections
    std::vector<BatchFilterFrame> ff;
    std::vector<AggArgs> ag;
    std::vector<FilterArgs> fa;         // per-frame filter arguments (per-frame launches)
    FilterArgs a0{};                    // shared filter arguments (batched launch)
    std::vector<const float*> ptrs;     // 6 per frame: src p0..p2, ref p0..p2
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int stage_max_cand(const StageParams& p) {
    const int side = 2 * (p.range / p.step) + 1;
    return side * side;
}

}  // namespace

struct bm3d_batch {
    int n = 0;
    std::vector<int> W, H, pitch;
    std::vector<size_t> est_off;        // basic-estimate plane offsets (floats)
    std::vector<int> tstart;            // first global output tile of each frame (n + 1)
    int ttotal = 0;
    Stage st[2];                        // Basic, Final
    BatchShape shape{};
    DevBuf blob, est, count, pos, key, order, hist, contrib, denw;
    size_t off_tstart = 0, blob_bytes = 0;
    std::vector<char> host_blob;
    std::vector<const float*> last_src;
    std::vector<float*> last_dst;
    bool blob_valid = false;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int64_t pixels = 0;
    std::mutex mu;  // calls on one batch are serialised (one blob, one workspace)
    ~bm3d_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

void copy_err(char* err, size_t errlen, const std::string& msg) {
    if (err && errlen) {
        std::strncpy(err, msg.c_str(), errlen - 1);
        err[errlen - 1] = 0;
    }
}

// Fills the argument blocks of stage s that depend on the caller's planes.
void fill_stage(bm3d_batch* b, int which, const float* const* src, float* const* dst) {
    Stage& s = b->st[which];
    float* est = b->est.as<float>();
    for (int f = 0; f < b->n; ++f) {
        const int s0 = s.start[f];
        const float* srcp = src[f];
        const float* refp = which == 0 ? src[f] : est + b->est_off[f];
        BatchFrame& bf = s.frames[f];
        bf.p = s.sp[f];
        bf.plane = refp;
        const float** pt = &s.ptrs[6 * f];
        pt[0] = srcp;
        pt[1] = pt[2] = nullptr;
        pt[3] = refp;
        pt[4] = pt[5] = nullptr;
        const float* const* dpt =
            reinterpret_cast<const float* const*>(b->blob.as<char>() + s.off_ptrs) + 6 * f;
        const bm3d_filter* flt = s.filt[f].get();
        s.ff[f] = BatchFilterFrame{b->pitch[f], dpt, flt->tables_dev.as<PlaneTables>()};
        FilterArgs& fa = s.fa[f];
        fa = FilterArgs{};
        fa.p = s.sp[f];
        fa.b0 = 0;
        fa.nb = s.start[f + 1] - s0;
        fa.wiener = s.wiener;
        fa.process[0] = 1;
        fa.count = b->count.as<int>() + s0;
        fa.pos = b->pos.as<int>() + (size_t)s0 * s.G;
        fa.frame = nullptr;
        fa.order = s.batched ? nullptr : b->order.as<int>() + s0;
        for (int p = 0; p < 3; ++p) {
            fa.src[p] = dpt + p;
            fa.ref[p] = dpt + 3 + p;
        }
        fa.tables = flt->tables_dev.as<PlaneTables>();
        fa.rcp_table = rcp_dev();
        fa.rcp_shift = 23 - rcp_bits();
        fa.coef = flt->coef_dev.as<float>();
        fa.coef_off = flt->coef_off_dev.as<int>();
        fa.contrib[0] = b->contrib.as<float>() + (size_t)s0 * s.G * s.B * s.B;
        fa.denw[0] = b->denw.as<float>() + s0;
        AggArgs& ag = s.ag[f];
        ag = AggArgs{};
        ag.p = s.sp[f];
        ag.process[0] = 1;
        ag.reach = s.sp[f].range;
        ag.count = fa.count;
        ag.pos = fa.pos;
        ag.contrib[0] = fa.contrib[0];
        ag.denw[0] = fa.denw[0];
        ag.out.mode = kOutFloat;
        ag.out.bits = 0;
        ag.out.dst[0] = which == 0 ? static_cast<void*>(est + b->est_off[f]) : static_cast<void*>(dst[f]);
        ag.out.pitch[0] = (size_t)b->pitch[f];
    }
    // batched launch: the stage-wide arrays indexed by the batch-wide group number
    FilterArgs& a0 = s.a0;
    a0 = s.fa[0];
    a0.nb = s.total;
    a0.count = b->count.as<int>();
    a0.pos = b->pos.as<int>();
    a0.order = nullptr;
    a0.contrib[0] = b->contrib.as<float>();
    a0.denw[0] = b->denw.as<float>();
}

// Writes the blob image on the host and uploads it (ordered on the stream).
int upload_blob(bm3d_batch* b, const float* const* src, float* const* dst) {
    fill_stage(b, 0, src, dst);
    fill_stage(b, 1, src, dst);
    char* h = b->host_blob.data();
    for (int which = 0; which < 2; ++which) {
        Stage& s = b->st[which];
        std::memcpy(h + s.off_start, s.start.data(), sizeof(int) * b->n);
        std::memcpy(h + s.off_frames, s.frames.data(), sizeof(BatchFrame) * b->n);
        std::memcpy(h + s.off_ff, s.ff.data(), sizeof(BatchFilterFrame) * b->n);
        std::memcpy(h + s.off_ag, s.ag.data(), sizeof(AggArgs) * b->n);
        std::memcpy(h + s.off_ptrs, s.ptrs.data(), sizeof(float*) * 6 * b->n);
    }
    std::memcpy(h + b->off_tstart, b->tstart.data(), sizeof(int) * b->n);
    HIP_TRY(hipMemcpyAsync(b->blob.p, h, b->blob_bytes, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));  // the host image may be rewritten by the next call
    b->last_src.assign(src, src + b->n);
    b->last_dst.assign(dst, dst + b->n);
    b->blob_valid = true;
    return BM3D_OK;
}

template <typename T>
const T* blob_at(const bm3d_batch* b, size_t off) {
    return reinterpret_cast<const T*>(b->blob.as<const char>() + off);
}

// Enqueues one stage (matching, bucketing, filtering, aggregation).
int enqueue_stage(bm3d_batch* b, int which) {
    Stage& s = b->st[which];
    const hipStream_t st = b->stream;
    int* cnt = b->count.as<int>();
    int* pos = b->pos.as<int>();
    float* key = b->key.as<float>();
    if (s.batched) {
        const int* starts = blob_at<int>(b, s.off_start);
        HIP_TRY(launch_match_batch(blob_at<BatchFrame>(b, s.off_frames), starts, b->n, s.total, s.sp[0], s.max_cand,
                                   b->shape, cnt, pos, key, st));
        HIP_TRY(launch_bucket_batch(cnt, starts, b->n, s.total, s.G, b->shape.group_order, b->hist.as<int>(),
                                    b->order.as<int>(), st));
        HIP_TRY(launch_filter_batch(s.a0, blob_at<BatchFilterFrame>(b, s.off_ff), starts, b->n, b->order.as<int>(),
                                    s.total, b->shape, st));
        // Gray frames: one processed plane (process[0] only)
        HIP_TRY(launch_aggregate_batch(blob_at<AggArgs>(b, s.off_ag), blob_at<int>(b, b->off_tstart), b->n, b->ttotal,
                                       1, b->shape, st));
        return BM3D_OK;
    }
    for (int f = 0; f < b->n; ++f) {
        const int s0 = s.start[f], nb = s.start[f + 1] - s0;
        HIP_TRY(launch_match(s.frames[f].plane, s.sp[f], nb, s.max_cand, cnt + s0, pos + (size_t)s0 * s.G,
                             key + (size_t)s0 * s.G, st));
        if (filter_uses_buckets(s.B, s.G))
            HIP_TRY(launch_bucket(cnt + s0, nb, s.G, b->order.as<int>() + s0, st));
        FilterArgs fa = s.fa[f];
        if (!filter_uses_buckets(s.B, s.G)) fa.order = nullptr;
        HIP_TRY(launch_filter(fa, st));
        HIP_TRY(launch_aggregate(s.ag[f], st));
    }
    return BM3D_OK;
}

int enqueue_chain(bm3d_batch* b) {
    int rc;
    if ((rc = enqueue_stage(b, 0))) return rc;
    return enqueue_stage(b, 1);
}

bool same_planes(const bm3d_batch* b, const float* const* src, float* const* dst) {
    if (!b->blob_valid) return false;
    for (int f = 0; f < b->n; ++f)
        if (b->last_src[f] != src[f] || b->last_dst[f] != dst[f]) return false;
    return true;
}

int run_chain(bm3d_batch* b, const float* const* src, float* const* dst, bool graph, float* ms) {
    int rc;
    if (!same_planes(b, src, dst)) {
        if ((rc = upload_blob(b, src, dst))) return rc;
        // the per-frame launches carry plane pointers in their arguments
        if (!(b->st[0].batched && b->st[1].batched)) b->drop_graph();
    }
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    if (graph) {
        if (!b->exec) {
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_chain(b);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(e);
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        if ((rc = enqueue_chain(b))) return rc;
    }
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    if (ms) HIP_TRY(hipEventElapsedTime(ms, b->ev[0], b->ev[1]));
    return BM3D_OK;
}

std::string shape_name(const BatchShape& s) {
    char buf[96];
    std::snprintf(buf, sizeof(buf), "match_waves=%d match_strip=%d group_order=%d", s.match_waves, s.match_strip,
                  s.group_order);
    return buf;
}

bool shape_ok(const bm3d_batch* b, const BatchShape& sh) {
    if (sh.group_order != kOrderFrame && sh.group_order != kOrderChunk) return false;
    for (const Stage& s : b->st)
        if (s.batched && !batch_match_supported(s.sp[0], s.max_cand, sh)) return false;
    return true;
}

}  // namespace

extern "C" {

int bm3d_batch_create(bm3d_batch** out, int nframes, const int* widths, const int* heights, const int* pitches,
                      const double* sigmas, const bm3d_arg* args, int nargs, char* err, size_t errlen) {
    if (out) *out = nullptr;
    if (!out || nframes <= 0 || !widths || !heights || !pitches || !sigmas || (nargs > 0 && !args))
        return set_error(BM3D_E_ARG, "bm3d_batch_create: null argument or no frames");
    for (int i = 0; i < nargs; ++i)
        if (args[i].key && std::strcmp(args[i].key, "sigma") == 0)
            return set_error(BM3D_E_ARG, "bm3d_batch_create: sigma is given per frame, not in args");
    std::unique_ptr<bm3d_batch> b(new bm3d_batch());
    b->n = nframes;
    b->W.assign(widths, widths + nframes);
    b->H.assign(heights, heights + nframes);
    b->pitch.assign(pitches, pitches + nframes);
    std::vector<bm3d_arg> a(args, args + nargs);
    a.push_back(bm3d_arg{});
    bm3d_arg& sig = a.back();
    sig.key = "sigma";
    sig.type = BM3D_ARG_FLOAT;
    sig.count = 1;
    int rc;
    for (int which = 0; which < 2; ++which) {
        Stage& s = b->st[which];
        s.wiener = which == 1;
        s.start.assign(nframes + 1, 0);
        int64_t total = 0;
        for (int f = 0; f < nframes; ++f) {
            const int W = widths[f], H = heights[f];
            if (W <= 0 || H <= 0 || pitches[f] < W || W > 65535 || H > 32767)
                return set_error(BM3D_E_ARG, "bm3d_batch_create: frame " + std::to_string(f) +
                                                 ": invalid width, height or pitch");
            if (!(sigmas[f] > 0.0))
                return set_error(BM3D_E_ARG, "bm3d_batch_create: frame " + std::to_string(f) +
                                                 ": sigma must be positive");
            sig.floats = &sigmas[f];
            bm3d_clip_info vi{BM3D_CF_GRAY, BM3D_ST_FLOAT, 32, 0, 0, W, H, 1};
            bm3d_filter* flt = nullptr;
            char e[512] = {0};
            rc = bm3d_create(which ? "Final" : "Basic", &vi, which ? &vi : nullptr, a.data(), (int)a.size(), &flt, e,
                             sizeof(e));
            if (rc) {
                std::string msg = "bm3d_batch_create: frame " + std::to_string(f) + ": " +
                                  (e[0] ? std::string(e) : std::string(bm3d_last_error()));
                copy_err(err, errlen, msg);
                return set_error(rc, msg);
            }
            s.filt.emplace_back(flt);
            if (W < flt->B || H < flt->B)
                return set_error(BM3D_E_ARG, "bm3d_batch_create: frame " + std::to_string(f) +
                                                 ": smaller than block_size");
            s.sp.push_back(stage_params(flt, W, H, (size_t)pitches[f]));
            s.start[f] = (int)total;
            total += (int64_t)s.sp.back().nx * s.sp.back().ny;
        }
        s.start[nframes] = (int)total;
        const bm3d_filter* f0 = s.filt[0].get();
        s.G = f0->G;
        s.B = f0->B;
        s.max_cand = stage_max_cand(s.sp[0]);
        if (total * s.G >= ((int64_t)1 << 31))
            return set_error(BM3D_E_UNSUPPORTED, "bm3d_batch_create: reference blocks x group_size of the batch must "
                                                 "stay below 2^31");
        s.total = (int)total;
        if (match_lds_bytes(s.max_cand, s.B, s.G) > 65536)
            return set_error(BM3D_E_UNSUPPORTED, "bm3d_batch_create: search window too large for the matching kernel");
        if (!filter_fits(s.B, s.G, s.wiener))
            return set_error(BM3D_E_UNSUPPORTED,
                             "bm3d_batch_create: block_size/group_size too large for the on-chip group buffer");
        s.frames.resize(nframes);
        s.ff.resize(nframes);
        s.fa.resize(nframes);
        s.ag.resize(nframes);
        s.ptrs.assign(6 * (size_t)nframes, nullptr);
    }
    // default launch shape: 2 waves per matching workgroup, the first strip
    // width of 3, 11, 8 that fits, size buckets per chunk of groups
    b->shape = BatchShape{2, 11, kOrderChunk};
    for (int w : {3, 11, 8}) {
        BatchShape t{2, w, kOrderChunk};
        bool ok = true;
        for (const Stage& s : b->st)
            ok = ok && s.B == 8 && s.sp[0].step == 1 && batch_match_supported(s.sp[0], s.max_cand, t);
        if (ok) {
            b->shape = t;
            break;
        }
    }
    for (Stage& s : b->st)
        s.batched = s.B == 8 && s.sp[0].step == 1 && s.G <= 32 && filter_uses_buckets(s.B, s.G) &&
                    batch_match_supported(s.sp[0], s.max_cand, b->shape);
    // tiles and basic-estimate planes
    b->tstart.assign(nframes + 1, 0);
    b->est_off.assign(nframes, 0);
    int64_t tiles = 0, est_floats = 0;
    for (int f = 0; f < nframes; ++f) {
        b->tstart[f] = (int)tiles;
        tiles += aggregate_tiles(widths[f], heights[f]);
        b->est_off[f] = (size_t)est_floats;
        est_floats += (int64_t)align_up((size_t)pitches[f] * heights[f], 64);
        b->pixels += (int64_t)widths[f] * heights[f];
    }
    b->tstart[nframes] = (int)tiles;
    b->ttotal = (int)tiles;
    // device blob layout
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    for (Stage& s : b->st) {
        s.off_start = take(sizeof(int) * nframes);
        s.off_frames = take(sizeof(BatchFrame) * nframes);
        s.off_ff = take(sizeof(BatchFilterFrame) * nframes);
        s.off_ag = take(sizeof(AggArgs) * nframes);
        s.off_ptrs = take(sizeof(float*) * 6 * nframes);
    }
    b->off_tstart = take(sizeof(int) * nframes);
    b->blob_bytes = off;
    b->host_blob.assign(off, 0);
    size_t maxb = 0, maxbg = 0;
    for (const Stage& s : b->st) {
        maxb = std::max(maxb, (size_t)s.total);
        maxbg = std::max(maxbg, (size_t)s.total * s.G);
    }
    size_t maxc = 0;
    for (const Stage& s : b->st) maxc = std::max(maxc, (size_t)s.total * s.G * s.B * s.B);
    int gmax = std::max(b->st[0].G, b->st[1].G);
    if ((rc = b->blob.ensure(b->blob_bytes))) return rc;
    if ((rc = b->est.ensure(sizeof(float) * (size_t)est_floats))) return rc;
    if ((rc = b->count.ensure(sizeof(int) * maxb))) return rc;
    if ((rc = b->pos.ensure(sizeof(int) * maxbg))) return rc;
    if ((rc = b->key.ensure(sizeof(float) * maxbg))) return rc;
    if ((rc = b->order.ensure(sizeof(int) * maxb))) return rc;
    if ((rc = b->hist.ensure(sizeof(int) * (size_t)nframes * (gmax + 1)))) return rc;
    if ((rc = b->contrib.ensure(sizeof(float) * maxc))) return rc;
    if ((rc = b->denw.ensure(sizeof(float) * maxb))) return rc;
    HIP_TRY(hipStreamCreate(&b->stream));  // ordered with the legacy default stream
    HIP_TRY(hipEventCreate(&b->ev[0]));
    HIP_TRY(hipEventCreate(&b->ev[1]));
    *out = b.release();
    return BM3D_OK;
}

void bm3d_batch_free(bm3d_batch* b) { delete b; }

int bm3d_batch_run(bm3d_batch* b, const float* const* src, float* const* dst, int graph, float* ms) {
    if (!b || !src || !dst) return set_error(BM3D_E_STATE, "bm3d_batch_run: null argument");
    for (int f = 0; f < b->n; ++f)
        if (!src[f] || !dst[f]) return set_error(BM3D_E_ARG, "bm3d_batch_run: null plane");
    std::lock_guard<std::mutex> lk(b->mu);
    return run_chain(b, src, dst, graph != 0, ms);
}

int bm3d_batch_get_shape(const bm3d_batch* b, bm3d_batch_shape* shape) {
    if (!b || !shape) return set_error(BM3D_E_STATE, "bm3d_batch_get_shape: null argument");
    shape->match_waves = b->shape.match_waves;
    shape->match_strip = b->shape.match_strip;
    shape->group_order = b->shape.group_order;
    return BM3D_OK;
}

int bm3d_batch_set_shape(bm3d_batch* b, const bm3d_batch_shape* shape) {
    if (!b || !shape) return set_error(BM3D_E_STATE, "bm3d_batch_set_shape: null argument");
    BatchShape t{shape->match_waves, shape->match_strip, shape->group_order};
    if (!shape_ok(b, t)) return set_error(BM3D_E_ARG, "bm3d_batch_set_shape: unsupported shape " + shape_name(t));
    std::lock_guard<std::mutex> lk(b->mu);
    b->shape = t;
    b->drop_graph();
    return BM3D_OK;
}

int bm3d_batch_tune(bm3d_batch* b, const float* const* src, float* const* dst, int reps, char* report,
                    size_t report_len) {
    if (!b || !src || !dst) return set_error(BM3D_E_STATE, "bm3d_batch_tune: null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    reps = std::max(1, reps);
    std::string rep;
    auto time_shape = [&](const BatchShape& t, float& best) -> int {
        b->shape = t;
        b->drop_graph();
        float sum = 0.0f;
        int rc;
        if ((rc = run_chain(b, src, dst, false, nullptr))) return rc;  // warm-up
        for (int r = 0; r < reps; ++r) {
            float ms = 0.0f;
            if ((rc = run_chain(b, src, dst, false, &ms))) return rc;
            sum += ms;
        }
        best = sum / reps;
        rep += shape_name(t) + " ms=" + std::to_string(best) + "\n";
        return BM3D_OK;
    };
    const bool any_batched = b->st[0].batched || b->st[1].batched;
    BatchShape best = b->shape;
    float best_ms = -1.0f;
    int rc;
    if (any_batched) {
        // matching shape first (at the default group order), then the order
        for (int nw : {1, 2, 4})
            for (int w : {3, 8, 10, 11}) {
                BatchShape t{nw, w, best.group_order};
                if (!shape_ok(b, t)) continue;
                float ms;
                if ((rc = time_shape(t, ms))) return rc;
                if (best_ms < 0 || ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
        BatchShape t = best;
        t.group_order = best.group_order == kOrderFrame ? kOrderChunk : kOrderFrame;
        float ms;
        if ((rc = time_shape(t, ms))) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    b->shape = best;
    b->drop_graph();
    rep += "best " + shape_name(best) + "\n";
    copy_err(report, report_len, rep);
    return BM3D_OK;
}

int bm3d_batch_info(const bm3d_batch* b, int64_t* basic_blocks, int64_t* final_blocks, int64_t* pixels) {
    if (!b) return set_error(BM3D_E_STATE, "bm3d_batch_info: null handle");
    if (basic_blocks) *basic_blocks = b->st[0].total;
    if (final_blocks) *final_blocks = b->st[1].total;
    if (pixels) *pixels = b->pixels;
    return BM3D_OK;
}

}  // extern "C"
