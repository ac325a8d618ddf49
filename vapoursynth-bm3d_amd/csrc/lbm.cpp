// Host side of the learned block-matching proposal: model state, training
// steps, the proposal matching pipeline and the C ABI of bm3d_lbm.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bm3d_lbm.h"
#include "bm3d_engine.h"
#include "bm3d_kernels.h"
#include "lbm_kernels.h"

using bm3d::DevBuf;
namespace L = bm3d::lbm;

struct bm3d_lbm {
    DevBuf master, w1h, w2h, adam_m, adam_v, grad, part, block_loss;
    DevBuf xs, hs, es, des, dhs, target;
    bm3d::LbmWork work;  // standalone bm3d_lbm_match_plane calls
    int64_t steps = 0;
    hipStream_t stream = nullptr;
    L::Net net() const {
        L::Net n{};
        n.master = master.as<float>();
        n.w1h = w1h.as<short>();
        n.w2h = w2h.as<short>();
        n.b1 = n.master + L::kOffB1;
        n.b2 = n.master + L::kOffB2;
        return n;
    }
    ~bm3d_lbm() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace bm3d {

namespace {

constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;
constexpr float kLossT0 = 1.0f;

int ensure_batch(bm3d_lbm* m, int npairs) {
    const size_t np = 2 * (size_t)npairs;
    int rc;
    if ((rc = m->xs.ensure(np * L::kIn * sizeof(short)))) return rc;
    if ((rc = m->hs.ensure(np * L::kHidden * sizeof(float)))) return rc;
    if ((rc = m->es.ensure(np * L::kEmb * sizeof(float)))) return rc;
    if ((rc = m->des.ensure(np * L::kEmb * sizeof(float)))) return rc;
    if ((rc = m->dhs.ensure(np * L::kHidden * sizeof(float)))) return rc;
    if ((rc = m->target.ensure((size_t)npairs * sizeof(float)))) return rc;
    if ((rc = m->part.ensure((size_t)L::train_slices(npairs) * L::kParams * sizeof(float)))) return rc;
    if ((rc = m->block_loss.ensure((size_t)((npairs + 255) / 256) * sizeof(float)))) return rc;
    return BM3D_OK;
}

L::TrainBatch make_batch(bm3d_lbm* m, const float* plane, int pitch, const int* pairs, int npairs,
                         const float* target) {
    L::TrainBatch tb{};
    tb.plane = plane;
    tb.pitch = pitch;
    tb.pairs = reinterpret_cast<const int4*>(pairs);
    tb.npairs = npairs;
    tb.target = target ? target : m->target.as<float>();
    tb.xs = m->xs.as<short>();
    tb.hs = m->hs.as<float>();
    tb.es = m->es.as<float>();
    tb.des = m->des.as<float>();
    tb.dhs = m->dhs.as<float>();
    return tb;
}

// Forward, loss and gradient of one batch on m's stream; *loss = batch loss.
int batch_gradient(bm3d_lbm* m, const float* plane, int pitch, const int* pairs, const float* target, int npairs,
                   float* loss) {
    if (!plane || !pairs || npairs <= 0 || pitch <= 0) return set_error(BM3D_E_ARG, "bm3d_lbm: bad batch");
    int rc;
    if ((rc = ensure_batch(m, npairs))) return rc;
    hipStream_t st = m->stream;
    if (!target) HIP_TRY(L::launch_pair_ssd(plane, pitch, reinterpret_cast<const int4*>(pairs), npairs,
                                            m->target.as<float>(), st));
    const L::TrainBatch tb = make_batch(m, plane, pitch, pairs, npairs, target);
    const L::Net n = m->net();
    HIP_TRY(L::launch_train_fwd(n, tb, st));
    HIP_TRY(L::launch_train_loss(tb, kLossT0, m->block_loss.as<float>(), st));
    HIP_TRY(L::launch_train_bwd(n, tb, m->part.as<float>(), m->grad.as<float>(), st));
    if (loss) {
        const int nbk = (npairs + 255) / 256;
        std::vector<float> bl(nbk);
        HIP_TRY(hipMemcpyAsync(bl.data(), m->block_loss.p, nbk * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        double s = 0.0;
        for (float v : bl) s += v;
        *loss = (float)(s / npairs);
    }
    return BM3D_OK;
}

}  // namespace

int lbm_match(const bm3d_lbm* m, int topk, LbmWork& w, const float* plane, const StageParams& sp, int* count,
              int* pos, float* key, hipStream_t st) {
    const int nb = sp.nx * sp.ny;
    const int side = 2 * (sp.range / sp.step) + 1;
    const int max_cand = side * side;
    if (!m) {
        HIP_TRY(launch_match(plane, sp, nb, max_cand, count, pos, key, st));
        return BM3D_OK;
    }
    if (sp.block != 8 || sp.step != 1)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: the learned matcher needs block_size=8 and bm_step=1");
    if (max_cand > L::kMaxWindow)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: the learned matcher supports bm_range <= 24");
    if (topk < 1 || topk > L::kMaxProposals)
        return set_error(BM3D_E_ARG, "bm3d: learned matcher top-k must be in [1, 64]");
    if (sp.group > L::kMaxLearnedGroupSize)
        return set_error(BM3D_E_UNSUPPORTED, "bm3d: the learned matcher supports group_size <= 64");
    const int W = sp.width, H = sp.height;
    const size_t npos = (size_t)((W - 7 + L::kLat - 1) / L::kLat) * ((H - 7 + L::kLat - 1) / L::kLat);  // lattice
    int rc;
    // (one spare slot each: the embedding kernels' stores of patches without an output)
    if ((rc = w.emb.ensure((npos + 1) * L::kEmb * sizeof(_Float16)))) return rc;
    if ((rc = w.norm.ensure((npos + 1) * sizeof(float)))) return rc;
    if ((rc = w.remb.ensure(((size_t)nb + 1) * L::kEmb * sizeof(_Float16)))) return rc;
    if ((rc = w.rnorm.ensure(((size_t)nb + 1) * sizeof(float)))) return rc;
    if ((rc = w.fallback.ensure((size_t)nb * sizeof(int)))) return rc;
    if ((rc = w.nfallback.ensure(sizeof(int)))) return rc;
    if ((rc = w.props.ensure(L::propose_list_bytes(sp)))) return rc;
    const L::Net n = m->net();
    HIP_TRY(L::launch_embed(n, plane, sp, w.emb.as<_Float16>(), w.norm.as<float>(), w.remb.as<_Float16>(),
                            w.rnorm.as<float>(), w.nfallback.as<int>(), st));
    L::ProposeArgs pa{};
    pa.p = sp;
    pa.plane = plane;
    pa.emb = w.emb.as<_Float16>();
    pa.norm = w.norm.as<float>();
    pa.remb = w.remb.as<_Float16>();
    pa.rnorm = w.rnorm.as<float>();
    pa.max_cand = max_cand;
    pa.topk = topk;
    pa.count = count;
    pa.pos = pos;
    pa.key = key;
    pa.fallback = w.fallback.as<int>();
    pa.nfallback = w.nfallback.as<int>();
    pa.props = w.props.as<uint16_t>();
    HIP_TRY(L::launch_propose(pa, st));
    HIP_TRY(launch_match(plane, sp, nb, max_cand, count, pos, key, st, w.fallback.as<int>(), w.nfallback.as<int>()));
    return BM3D_OK;
}

}  // namespace bm3d

using bm3d::set_error;

extern "C" {

int bm3d_lbm_param_count(void) { return L::kParams; }

int bm3d_lbm_set_params(bm3d_lbm* m, const float* params) {
    if (!m || !params) return set_error(BM3D_E_ARG, "bm3d_lbm_set_params: null argument");
    HIP_TRY(hipMemcpy(m->master.p, params, L::kParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->adam_m.p, 0, L::kParams * sizeof(float)));
    HIP_TRY(hipMemset(m->adam_v.p, 0, L::kParams * sizeof(float)));
    m->steps = 0;
    HIP_TRY(L::launch_sync_bf16(m->net(), m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return BM3D_OK;
}

int bm3d_lbm_create(bm3d_lbm** out, const float* params) {
    if (!out || !params) return set_error(BM3D_E_ARG, "bm3d_lbm_create: null argument");
    *out = nullptr;
    bm3d_lbm* m = new bm3d_lbm();
    int rc = BM3D_OK;
    auto fail = [&](int r) {
        delete m;
        return r;
    };
    if (hipStreamCreate(&m->stream) != hipSuccess)  // ordered with the legacy default stream
        return fail(set_error(BM3D_E_HIP, "bm3d_lbm_create: stream creation failed"));
    if ((rc = m->master.ensure(L::kParams * sizeof(float)))) return fail(rc);
    if ((rc = m->w1h.ensure(L::kHidden * L::kIn * sizeof(short)))) return fail(rc);
    if ((rc = m->w2h.ensure(L::kEmb * L::kHidden * sizeof(short)))) return fail(rc);
    if ((rc = m->adam_m.ensure(L::kParams * sizeof(float)))) return fail(rc);
    if ((rc = m->adam_v.ensure(L::kParams * sizeof(float)))) return fail(rc);
    if ((rc = m->grad.ensure(L::kParams * sizeof(float)))) return fail(rc);
    if ((rc = bm3d_lbm_set_params(m, params))) return fail(rc);
    *out = m;
    return BM3D_OK;
}

void bm3d_lbm_free(bm3d_lbm* m) { delete m; }

int bm3d_lbm_get_params(const bm3d_lbm* m, float* params) {
    if (!m || !params) return set_error(BM3D_E_ARG, "bm3d_lbm_get_params: null argument");
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(params, m->master.p, L::kParams * sizeof(float), hipMemcpyDeviceToHost));
    return BM3D_OK;
}

int bm3d_lbm_gradients(bm3d_lbm* m, const float* plane, int pitch, const int* pairs, const float* target,
                       int npairs, float* grad, float* loss) {
    if (!m || !grad) return set_error(BM3D_E_ARG, "bm3d_lbm_gradients: null argument");
    int rc;
    if ((rc = bm3d::batch_gradient(m, plane, pitch, pairs, target, npairs, loss))) return rc;
    HIP_TRY(hipMemcpyAsync(grad, m->grad.p, L::kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return BM3D_OK;
}

int bm3d_lbm_train_step(bm3d_lbm* m, const float* plane, int pitch, const int* pairs, int npairs, float lr,
                        float* loss) {
    if (!m) return set_error(BM3D_E_ARG, "bm3d_lbm_train_step: null model");
    int rc;
    if ((rc = bm3d::batch_gradient(m, plane, pitch, pairs, nullptr, npairs, loss))) return rc;
    m->steps += 1;
    const float c1 = 1.0f - std::pow(bm3d::kAdamB1, (float)m->steps);
    const float c2 = 1.0f - std::pow(bm3d::kAdamB2, (float)m->steps);
    HIP_TRY(L::launch_train_adam(m->net(), m->grad.as<float>(), m->adam_m.as<float>(), m->adam_v.as<float>(), lr,
                                 bm3d::kAdamB1, bm3d::kAdamB2, bm3d::kAdamEps, c1, c2, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return BM3D_OK;
}

int bm3d_lbm_forward(bm3d_lbm* m, const float* plane, int pitch, const int* pairs, int npairs, float* emb) {
    if (!m || !emb || !plane || !pairs || npairs <= 0) return set_error(BM3D_E_ARG, "bm3d_lbm_forward: bad argument");
    int rc;
    if ((rc = bm3d::ensure_batch(m, npairs))) return rc;
    const bm3d::lbm::TrainBatch tb = bm3d::make_batch(m, plane, pitch, pairs, npairs, nullptr);
    HIP_TRY(L::launch_train_fwd(m->net(), tb, m->stream));
    HIP_TRY(hipMemcpyAsync(emb, m->es.p, 2 * (size_t)npairs * L::kEmb * sizeof(float), hipMemcpyDeviceToHost,
                           m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return BM3D_OK;
}

int bm3d_lbm_match_plane(const bm3d_filter* f, const bm3d_lbm* m, int topk, const float* plane, int pitch, int width,
                         int height, int* count, int* pos, float* key, int* fallback, float* ms) {
    if (!f || !plane || !count || !pos || !key) return set_error(BM3D_E_ARG, "bm3d_lbm_match_plane: null argument");
    if (f->fn != bm3d::kBasic && f->fn != bm3d::kFinal)
        return set_error(BM3D_E_ARG, "bm3d_lbm_match_plane: needs a Basic or Final filter");
    if (width < f->B || height < f->B || pitch < width)
        return set_error(BM3D_E_ARG, "bm3d_lbm_match_plane: bad plane size");
    const bm3d::StageParams sp = bm3d::stage_params(f, width, height, (size_t)pitch);
    bm3d_lbm* mm = const_cast<bm3d_lbm*>(m);
    bm3d::LbmWork tmp;
    bm3d::LbmWork& w = mm ? mm->work : tmp;
    hipStream_t st = f->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, st));
    }
    int rc = bm3d::lbm_match(m, topk, w, plane, sp, count, pos, key, st);
    if (rc) return rc;
    if (ms) {
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (fallback) {
        *fallback = 0;
        if (m) HIP_TRY(hipMemcpy(fallback, w.nfallback.p, sizeof(int), hipMemcpyDeviceToHost));
    }
    return BM3D_OK;
}

int bm3d_filter_params(const bm3d_filter* f, int* out) {
    if (!f || !out) return set_error(BM3D_E_ARG, "bm3d_filter_params: null argument");
    out[0] = f->B;
    out[1] = f->step;
    out[2] = f->G;
    out[3] = f->R;
    out[4] = f->S;
    return BM3D_OK;
}

int bm3d_set_matcher(bm3d_filter* f, bm3d_lbm* m, int topk) {
    if (!f) return set_error(BM3D_E_ARG, "bm3d_set_matcher: null filter");
    if (f->fn != bm3d::kBasic && f->fn != bm3d::kFinal)
        return set_error(BM3D_E_ARG, "bm3d_set_matcher: only Basic and Final use a matcher");
    if (m && (topk < 1 || topk > L::kMaxProposals))
        return set_error(BM3D_E_ARG, "bm3d_set_matcher: top-k must be in [1, 64]");
    std::lock_guard<std::mutex> lk(f->mu);
    f->lbm = m;
    f->lbm_topk = topk;
    return BM3D_OK;
}

int bm3d_matcher_fallbacks(const bm3d_filter* f, int* fallback) {
    if (!f || !fallback) return set_error(BM3D_E_ARG, "bm3d_matcher_fallbacks: null argument");
    *fallback = 0;
    if (!f->lbm || !f->lbm_work.nfallback.p) return BM3D_OK;
    HIP_TRY(hipStreamSynchronize(f->stream));
    HIP_TRY(hipMemcpy(fallback, f->lbm_work.nfallback.p, sizeof(int), hipMemcpyDeviceToHost));
    return BM3D_OK;
}

int bm3d_match_lists(const bm3d_filter* f, int64_t* nb, int* group, int* count, int* pos, float* key) {
    if (!f) return set_error(BM3D_E_ARG, "bm3d_match_lists: null filter");
    std::lock_guard<std::mutex> lk(f->mu);
    if (nb) *nb = f->last_nb;
    if (group) *group = f->G;
    if (!count && !pos && !key) return BM3D_OK;
    if (f->last_nb <= 0 || !f->last_count) return set_error(BM3D_E_STATE, "bm3d_match_lists: no frame processed yet");
    HIP_TRY(hipStreamSynchronize(f->stream));
    const size_t nbz = (size_t)f->last_nb;
    if (count) HIP_TRY(hipMemcpy(count, f->last_count, nbz * sizeof(int), hipMemcpyDeviceToHost));
    if (pos) HIP_TRY(hipMemcpy(pos, f->last_pos, nbz * f->G * sizeof(int), hipMemcpyDeviceToHost));
    if (key) HIP_TRY(hipMemcpy(key, f->last_key, nbz * f->G * sizeof(float), hipMemcpyDeviceToHost));
    return BM3D_OK;
}

}  // extern "C"
