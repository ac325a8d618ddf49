// VapourSynth API v4 plugin "bm3d" over the MI355X library's C ABI.
//
// The plugin surface is the reference's (source/VSPlugin.cpp:424-518): the
// same identifier, namespace, description, version and the seven functions
// with their argument strings, return type, filter modes and dependency
// request patterns.  Each callback maps onto one C-ABI call:
//   *_Create            -> bm3d_create (argument processing and errors) and
//                          bm3d_output_info (the output VSVideoInfo)
//   GetFrame/arInitial  -> bm3d_frame_range (the frames to request)
//   GetFrame/arAllFramesReady -> bm3d_get_frame on the frames' host planes,
//                          then the output frame properties
//   *_Free              -> bm3d_free
// Warnings go to vsapi->logMessage(mtWarning, ...) as in the reference.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bm3d_amd.h"
#include "vs4_api.h"

namespace {

enum ArgKind { kInt, kFloat, kData };

struct ArgSpec {
    const char* key;
    ArgKind kind;
};

// The optional arguments of each function (the registered strings below), in
// the C ABI's value types.
const ArgSpec kBM3DArgs[] = {
    {"profile", kData},  {"sigma", kFloat},    {"radius", kInt},   {"block_size", kInt}, {"block_step", kInt},
    {"group_size", kInt}, {"bm_range", kInt},  {"bm_step", kInt},  {"ps_num", kInt},     {"ps_range", kInt},
    {"ps_step", kInt},   {"th_mse", kFloat},   {"hard_thr", kFloat}, {"matrix", kInt},
};
const ArgSpec kSampleArgs[] = {{"sample", kInt}, {"radius", kInt}};

struct Function {
    const char* name;
    const char* args;
    bool temporal;  // frames requested around n: rpGeneral dependencies
};

const Function kFunctions[] = {
    {"RGB2OPP", "input:vnode;sample:int:opt;", false},
    {"OPP2RGB", "input:vnode;sample:int:opt;", false},
    {"Basic",
     "input:vnode;ref:vnode:opt;profile:data:opt;sigma:float[]:opt;block_size:int:opt;block_step:int:opt;"
     "group_size:int:opt;bm_range:int:opt;bm_step:int:opt;th_mse:float:opt;hard_thr:float:opt;matrix:int:opt;",
     false},
    {"Final",
     "input:vnode;ref:vnode;profile:data:opt;sigma:float[]:opt;block_size:int:opt;block_step:int:opt;"
     "group_size:int:opt;bm_range:int:opt;bm_step:int:opt;th_mse:float:opt;matrix:int:opt;",
     false},
    {"VBasic",
     "input:vnode;ref:vnode:opt;profile:data:opt;sigma:float[]:opt;radius:int:opt;block_size:int:opt;"
     "block_step:int:opt;group_size:int:opt;bm_range:int:opt;bm_step:int:opt;ps_num:int:opt;ps_range:int:opt;"
     "ps_step:int:opt;th_mse:float:opt;hard_thr:float:opt;matrix:int:opt;",
     true},
    {"VFinal",
     "input:vnode;ref:vnode;profile:data:opt;sigma:float[]:opt;radius:int:opt;block_size:int:opt;"
     "block_step:int:opt;group_size:int:opt;bm_range:int:opt;bm_step:int:opt;ps_num:int:opt;ps_range:int:opt;"
     "ps_step:int:opt;th_mse:float:opt;matrix:int:opt;",
     true},
    {"VAggregate", "input:vnode;radius:int:opt;sample:int:opt;", true},
};

// ---------------------------------------------------------------- library

std::once_flag g_init_once;
int g_init_rc = BM3D_E_STATE;
std::string g_init_msg;

std::mutex g_log_mu;
const VSAPI* g_log_api = nullptr;
VSCore* g_log_core = nullptr;

void log_to_core(int level, const char* msg, void*) {
    (void)level;
    std::lock_guard<std::mutex> lk(g_log_mu);
    if (g_log_api && g_log_core) g_log_api->logMessage(mtWarning, msg, g_log_core);
    else std::fprintf(stderr, "%s\n", msg);
}

// bm3d_init on first use (BM3D_AMD_DEVICE selects the HIP device, default 0).
int init_library(std::string& msg) {
    std::call_once(g_init_once, [] {
        const char* d = std::getenv("BM3D_AMD_DEVICE");
        g_init_rc = bm3d_init(d ? std::atoi(d) : 0);
        if (g_init_rc != BM3D_OK) g_init_msg = bm3d_last_error();
    });
    msg = g_init_msg;
    return g_init_rc;
}

// ---------------------------------------------------------------- filters

struct Instance {
    const Function* fn = nullptr;
    bm3d_filter* f = nullptr;
    VSNode* node = nullptr;
    VSNode* rnode = nullptr;
    VSVideoInfo vo{};
};

bm3d_clip_info clip_info(const VSVideoInfo* vi) {
    bm3d_clip_info c{};
    c.color_family = vi->format.colorFamily;
    c.sample_type = vi->format.sampleType;
    c.bits_per_sample = vi->format.bitsPerSample;
    c.subsampling_w = vi->format.subSamplingW;
    c.subsampling_h = vi->format.subSamplingH;
    c.width = vi->width;
    c.height = vi->height;
    c.num_frames = vi->numFrames;
    return c;
}

int64_t prop_int(const VSAPI* vsapi, const VSMap* m, const char* key, int idx = 0) {
    int err = 0;
    const int64_t v = vsapi->mapGetInt(m, key, idx, &err);
    return err ? -1 : v;
}

bm3d_props read_props(const VSAPI* vsapi, const VSFrame* fr) {
    const VSMap* m = vsapi->getFramePropertiesRO(fr);
    bm3d_props p{};
    p.bm3d_opp = prop_int(vsapi, m, "BM3D_OPP");
    p.range = prop_int(vsapi, m, "_Range");
    p.matrix = prop_int(vsapi, m, "_Matrix");
    p.v_radius = prop_int(vsapi, m, "BM3D_V_radius");
    for (int k = 0; k < 3; ++k) p.v_process[k] = prop_int(vsapi, m, "BM3D_V_process", k);
    return p;
}

void write_props(const VSAPI* vsapi, VSMap* m, const bm3d_props& p) {
    auto set = [&](const char* key, int64_t v) {
        if (v >= 0) vsapi->mapSetInt(m, key, v, maReplace);
        else vsapi->mapDeleteKey(m, key);
    };
    set("BM3D_OPP", p.bm3d_opp);
    set("_Range", p.range);
    set("_Matrix", p.matrix);
    set("BM3D_V_radius", p.v_radius);
    if (p.v_process[0] >= 0) {
        const int64_t v[3] = {p.v_process[0], p.v_process[1], p.v_process[2]};
        vsapi->mapSetIntArray(m, "BM3D_V_process", v, 3);
    } else {
        vsapi->mapDeleteKey(m, "BM3D_V_process");
    }
}

bm3d_frame host_frame(const VSAPI* vsapi, const VSFrame* fr, int planes) {
    bm3d_frame h{};
    for (int p = 0; p < planes; ++p) {
        h.data[p] = const_cast<uint8_t*>(vsapi->getReadPtr(fr, p));
        h.stride[p] = vsapi->getStride(fr, p);
    }
    h.width = vsapi->getFrameWidth(fr, 0);
    h.height = vsapi->getFrameHeight(fr, 0);
    h.mem = 0;
    h.props = read_props(vsapi, fr);
    return h;
}

const VSFrame* VS_CC get_frame(int n, int reason, void* instance, void** frame_data, VSFrameContext* ctx,
                               VSCore* core, const VSAPI* vsapi) {
    (void)frame_data;
    Instance* d = static_cast<Instance*>(instance);
    int first = n, last = n;
    bm3d_frame_range(d->f, n, &first, &last);
    if (reason == arInitial) {
        for (int m = first; m <= last; ++m) {
            vsapi->requestFrameFilter(m, d->node, ctx);
            if (d->rnode) vsapi->requestFrameFilter(m, d->rnode, ctx);
        }
        return nullptr;
    }
    if (reason != arAllFramesReady) return nullptr;
    const int T = last - first + 1;
    std::vector<const VSFrame*> src(T, nullptr), ref(d->rnode ? T : 0, nullptr);
    std::vector<bm3d_frame> hs(T), hr(d->rnode ? T : 0);
    const int in_planes = vsapi->getVideoInfo(d->node)->format.numPlanes;
    for (int t = 0; t < T; ++t) {
        src[t] = vsapi->getFrameFilter(first + t, d->node, ctx);
        hs[t] = host_frame(vsapi, src[t], in_planes);
        if (d->rnode) {
            ref[t] = vsapi->getFrameFilter(first + t, d->rnode, ctx);
            hr[t] = host_frame(vsapi, ref[t], in_planes);
        }
    }
    const VSFrame* cur = src[n - first];
    VSFrame* dst = vsapi->newVideoFrame(&d->vo.format, d->vo.width, d->vo.height, cur, core);
    bm3d_frame hd{};
    for (int p = 0; p < d->vo.format.numPlanes; ++p) {
        hd.data[p] = vsapi->getWritePtr(dst, p);
        hd.stride[p] = vsapi->getStride(dst, p);
    }
    hd.width = d->vo.width;
    hd.height = d->vo.height;
    hd.mem = 0;
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        g_log_api = vsapi;
        g_log_core = core;
    }
    const int rc = bm3d_get_frame(d->f, n, hs.data(), d->rnode ? hr.data() : nullptr, &hd);
    for (const VSFrame* fr : src) vsapi->freeFrame(fr);
    for (const VSFrame* fr : ref) vsapi->freeFrame(fr);
    if (rc != BM3D_OK) {
        const std::string msg = std::string("bm3d.") + d->fn->name + ": " + bm3d_last_error();
        vsapi->setFilterError(msg.c_str(), ctx);
        vsapi->freeFrame(dst);
        return nullptr;
    }
    write_props(vsapi, vsapi->getFramePropertiesRW(dst), hd.props);
    return dst;
}

void VS_CC free_instance(void* instance, VSCore* core, const VSAPI* vsapi) {
    (void)core;
    Instance* d = static_cast<Instance*>(instance);
    bm3d_free(d->f);
    vsapi->freeNode(d->node);
    if (d->rnode) vsapi->freeNode(d->rnode);
    delete d;
}

// The values of one VSMap entry in the C ABI's representation.
struct ArgValues {
    std::string key;
    ArgKind kind;
    std::vector<int64_t> ints;
    std::vector<double> floats;
    std::string data;
};

void VS_CC create(const VSMap* in, VSMap* out, void* user_data, VSCore* core, const VSAPI* vsapi) {
    const Function* fn = static_cast<const Function*>(user_data);
    int err = 0;
    VSNode* node = vsapi->mapGetNode(in, "input", 0, &err);
    VSNode* rnode = vsapi->mapGetNode(in, "ref", 0, &err);
    if (err) rnode = nullptr;
    auto fail = [&](const std::string& msg) {
        vsapi->mapSetError(out, msg.c_str());
        if (node) vsapi->freeNode(node);
        if (rnode) vsapi->freeNode(rnode);
    };
    if (!node) return fail(std::string("bm3d.") + fn->name + ": Invalid input clip");
    // the argument values, in the registered types
    const bool colour = !std::strcmp(fn->name, "RGB2OPP") || !std::strcmp(fn->name, "OPP2RGB") ||
                        !std::strcmp(fn->name, "VAggregate");
    const ArgSpec* specs = colour ? kSampleArgs : kBM3DArgs;
    const size_t nspec = colour ? sizeof(kSampleArgs) / sizeof(kSampleArgs[0]) : sizeof(kBM3DArgs) / sizeof(kBM3DArgs[0]);
    std::vector<ArgValues> vals;
    for (size_t k = 0; k < nspec; ++k) {
        const int ne = vsapi->mapNumElements(in, specs[k].key);
        if (ne <= 0) continue;
        ArgValues v{specs[k].key, specs[k].kind, {}, {}, {}};
        int e = 0;
        if (specs[k].kind == kInt) {
            const int64_t* a = vsapi->mapGetIntArray(in, specs[k].key, &e);
            if (!e && a) v.ints.assign(a, a + ne);
        } else if (specs[k].kind == kFloat) {
            const double* a = vsapi->mapGetFloatArray(in, specs[k].key, &e);
            if (!e && a) v.floats.assign(a, a + ne);
        } else {
            const char* s = vsapi->mapGetData(in, specs[k].key, 0, &e);
            if (!e && s) v.data = s;
        }
        if (e) continue;
        vals.push_back(std::move(v));
    }
    std::vector<bm3d_arg> args;
    for (const ArgValues& v : vals) {
        bm3d_arg a{};
        a.key = v.key.c_str();
        if (v.kind == kInt) {
            a.type = BM3D_ARG_INT;
            a.count = (int)v.ints.size();
            a.ints = v.ints.data();
        } else if (v.kind == kFloat) {
            a.type = BM3D_ARG_FLOAT;
            a.count = (int)v.floats.size();
            a.floats = v.floats.data();
        } else {
            a.type = BM3D_ARG_DATA;
            a.count = 1;
            a.data = v.data.c_str();
        }
        args.push_back(a);
    }
    const bm3d_clip_info ci = clip_info(vsapi->getVideoInfo(node));
    bm3d_clip_info ri{};
    if (rnode) ri = clip_info(vsapi->getVideoInfo(rnode));
    // argument errors are reported before the device is touched
    bm3d_filter* f = nullptr;
    char msg[1024] = {0};
    int rc = bm3d_create(fn->name, &ci, rnode ? &ri : nullptr, args.data(), (int)args.size(), &f, msg, sizeof msg);
    if (rc == BM3D_E_STATE && !f) {
        // arguments were valid; the library needs its device first
        std::string imsg;
        if (init_library(imsg) != BM3D_OK) return fail(std::string("bm3d.") + fn->name + ": " + imsg);
        rc = bm3d_create(fn->name, &ci, rnode ? &ri : nullptr, args.data(), (int)args.size(), &f, msg, sizeof msg);
    }
    if (rc != BM3D_OK) return fail(msg[0] ? std::string(msg) : std::string("bm3d.") + fn->name + ": " + bm3d_last_error());
    // the library's warnings go to this core's log (also when another user
    // of the library in this process initialised it)
    bm3d_set_log(log_to_core, nullptr);
    bm3d_clip_info oi{};
    bm3d_output_info(f, &oi);
    Instance* d = new Instance();
    d->fn = fn;
    d->f = f;
    d->node = node;
    d->rnode = rnode;
    d->vo = *vsapi->getVideoInfo(node);
    vsapi->queryVideoFormat(&d->vo.format, oi.color_family, oi.sample_type, oi.bits_per_sample, oi.subsampling_w,
                            oi.subsampling_h, core);
    d->vo.width = oi.width;
    d->vo.height = oi.height;
    const int pattern = fn->temporal ? rpGeneral : rpStrictSpatial;
    VSFilterDependency deps[2] = {{node, pattern}, {rnode, pattern}};
    vsapi->createVideoFilter(out, fn->name, &d->vo, get_frame, free_instance, fmParallel, deps, rnode ? 2 : 1, d, core);
}

}  // namespace

VS_EXTERNAL_API(void) VapourSynthPluginInit2(VSPlugin* plugin, const VSPLUGINAPI* vspapi) {
    vspapi->configPlugin("com.vapoursynth.bm3d", "bm3d", "Implementation of BM3D denoising filter for VapourSynth.",
                         VS_MAKE_VERSION(10, 1), VAPOURSYNTH_API_VERSION, 0, plugin);
    for (const Function& fn : kFunctions)
        vspapi->registerFunction(fn.name, fn.args, "clip:vnode;", create, const_cast<Function*>(&fn), plugin);
}
