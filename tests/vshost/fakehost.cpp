// A minimal VapourSynth API v4 host for tests: loads a plugin with dlopen,
// calls its VapourSynthPluginInit2, validates invocation arguments against the
// registered argument strings the way the VapourSynth core does (unknown
// names, missing required arguments, types, arrays), and serves frames with a
// synchronous scheduler: arInitial collects the requests, the requested
// frames are produced (recursively, cached per node), then arAllFramesReady.
// Frames live in host memory with 64-byte aligned rows.  Test
// infrastructure only; exposed to Python through the fh_* C functions.
#include <dlfcn.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../vapoursynth-bm3d_amd/vs/vs4_api.h"

// ------------------------------------------------------------------ objects

struct VSMap {
    struct Entry {
        std::string key;
        int type = ptUnset;
        std::vector<int64_t> i;
        std::vector<double> f;
        std::vector<std::string> d;
        std::vector<VSNode*> n;
        std::vector<const VSFrame*> fr;
    };
    std::vector<Entry> e;
    std::string error;
    Entry* find(const char* key) {
        for (auto& x : e)
            if (x.key == key) return &x;
        return nullptr;
    }
    const Entry* find(const char* key) const {
        for (auto& x : e)
            if (x.key == key) return &x;
        return nullptr;
    }
    void clear();
    ~VSMap() { clear(); }
};

struct VSFrame {
    std::atomic<int> refs{1};
    VSVideoFormat fmt{};
    int w = 0, h = 0;
    std::vector<uint8_t> buf[3];
    uint8_t* ptr[3] = {nullptr, nullptr, nullptr};
    ptrdiff_t stride[3] = {0, 0, 0};
    VSMap props;
};

struct VSNode {
    std::atomic<int> refs{1};
    VSVideoInfo vi{};
    std::string name;
    // source
    std::vector<VSFrame*> frames;
    // filter
    VSFilterGetFrame getframe = nullptr;
    VSFilterFree freefn = nullptr;
    void* instance = nullptr;
    VSCore* core = nullptr;
    std::mutex mu;
    std::map<int, const VSFrame*> cache;
};

struct VSFrameContext {
    std::vector<std::pair<VSNode*, int>> requests;
    std::vector<std::pair<std::pair<VSNode*, int>, const VSFrame*>> ready;
    std::string error;
};

struct RegFunction {
    std::string name, args, ret;
    VSPublicFunction fn = nullptr;
    void* data = nullptr;
};

struct VSPlugin {
    std::string id, ns, name;
    int version = 0, api = 0, flags = 0;
    std::vector<RegFunction> funcs;
    void* dl = nullptr;
};

struct VSCore {
    std::vector<std::unique_ptr<VSPlugin>> plugins;
    std::mutex log_mu;
    std::vector<std::pair<int, std::string>> log;
};

static const VSAPI* api();
static void frame_release(const VSFrame* f);
static void node_release(VSNode* n);

void VSMap::clear() {
    for (auto& x : e) {
        for (VSNode* n : x.n) node_release(n);
        for (const VSFrame* f : x.fr) frame_release(f);
    }
    e.clear();
    error.clear();
}

static void frame_release(const VSFrame* f) {
    if (!f) return;
    VSFrame* m = const_cast<VSFrame*>(f);
    if (--m->refs == 0) delete m;
}

static void node_release(VSNode* n) {
    if (!n) return;
    if (--n->refs != 0) return;
    if (n->freefn) n->freefn(n->instance, n->core, api());
    for (VSFrame* f : n->frames) frame_release(f);
    for (auto& kv : n->cache) frame_release(kv.second);
    delete n;
}

static int bytes_of(int st, int bits) { return st == stFloat ? (bits == 16 ? 2 : 4) : (bits > 8 ? 2 : 1); }

static VSFrame* frame_new(const VSVideoFormat& fmt, int w, int h) {
    VSFrame* f = new VSFrame();
    f->fmt = fmt;
    f->w = w;
    f->h = h;
    for (int p = 0; p < fmt.numPlanes; ++p) {
        const int pw = p ? w >> fmt.subSamplingW : w, ph = p ? h >> fmt.subSamplingH : h;
        const ptrdiff_t stride = ((ptrdiff_t)pw * fmt.bytesPerSample + 63) & ~(ptrdiff_t)63;
        f->buf[p].assign((size_t)stride * ph + 64, 0);
        uintptr_t a = (reinterpret_cast<uintptr_t>(f->buf[p].data()) + 63) & ~(uintptr_t)63;
        f->ptr[p] = reinterpret_cast<uint8_t*>(a);
        f->stride[p] = stride;
    }
    return f;
}

static void map_copy(const VSMap& s, VSMap& d) {
    d.clear();
    for (const auto& x : s.e) {
        d.e.push_back(x);
        for (VSNode* n : x.n) ++n->refs;
        for (const VSFrame* f : x.fr) ++const_cast<VSFrame*>(f)->refs;
    }
}

// ------------------------------------------------------------------ VSAPI

static int VS_CC query_format(VSVideoFormat* f, int cf, int st, int bits, int ssw, int ssh, VSCore*) noexcept {
    if (cf < cfGray || cf > cfYUV || (st != stInteger && st != stFloat) || bits < 8 || bits > 32) {
        std::memset(f, 0, sizeof(*f));
        return 0;
    }
    f->colorFamily = cf;
    f->sampleType = st;
    f->bitsPerSample = bits;
    f->bytesPerSample = bytes_of(st, bits);
    f->subSamplingW = cf == cfYUV ? ssw : 0;
    f->subSamplingH = cf == cfYUV ? ssh : 0;
    f->numPlanes = cf == cfGray ? 1 : 3;
    return 1;
}

static void VS_CC create_video_filter(VSMap* out, const char* name, const VSVideoInfo* vi, VSFilterGetFrame gf,
                                      VSFilterFree fr, int mode, const VSFilterDependency* deps, int ndeps,
                                      void* inst, VSCore* core) noexcept {
    (void)mode;
    (void)deps;
    (void)ndeps;
    VSNode* n = new VSNode();
    n->vi = *vi;
    n->name = name;
    n->getframe = gf;
    n->freefn = fr;
    n->instance = inst;
    n->core = core;
    VSMap::Entry e;
    e.key = "clip";
    e.type = ptVideoNode;
    e.n.push_back(n);
    out->e.push_back(std::move(e));
}

static void VS_CC free_node(VSNode* n) noexcept { node_release(n); }
static VSNode* VS_CC add_node_ref(VSNode* n) noexcept {
    ++n->refs;
    return n;
}
static const VSVideoInfo* VS_CC get_video_info(VSNode* n) noexcept { return &n->vi; }

static VSFrame* VS_CC new_video_frame(const VSVideoFormat* fmt, int w, int h, const VSFrame* prop_src,
                                      VSCore*) noexcept {
    VSFrame* f = frame_new(*fmt, w, h);
    if (prop_src) map_copy(prop_src->props, f->props);
    return f;
}
static void VS_CC free_frame(const VSFrame* f) noexcept { frame_release(f); }
static const VSFrame* VS_CC add_frame_ref(const VSFrame* f) noexcept {
    ++const_cast<VSFrame*>(f)->refs;
    return f;
}
static const VSMap* VS_CC props_ro(const VSFrame* f) noexcept { return &f->props; }
static VSMap* VS_CC props_rw(VSFrame* f) noexcept { return &f->props; }
static ptrdiff_t VS_CC get_stride(const VSFrame* f, int p) noexcept { return f->stride[p]; }
static const uint8_t* VS_CC get_read_ptr(const VSFrame* f, int p) noexcept { return f->ptr[p]; }
static uint8_t* VS_CC get_write_ptr(VSFrame* f, int p) noexcept { return f->ptr[p]; }
static const VSVideoFormat* VS_CC get_frame_format(const VSFrame* f) noexcept { return &f->fmt; }
static int VS_CC frame_width(const VSFrame* f, int p) noexcept { return p ? f->w >> f->fmt.subSamplingW : f->w; }
static int VS_CC frame_height(const VSFrame* f, int p) noexcept { return p ? f->h >> f->fmt.subSamplingH : f->h; }

static const VSFrame* produce(VSNode* node, int n, std::string& err);

static const VSFrame* VS_CC get_frame_filter(int n, VSNode* node, VSFrameContext* ctx) noexcept {
    for (auto& r : ctx->ready)
        if (r.first.first == node && r.first.second == n) return add_frame_ref(r.second);
    return nullptr;
}
static void VS_CC request_frame_filter(int n, VSNode* node, VSFrameContext* ctx) noexcept {
    for (auto& r : ctx->requests)
        if (r.first == node && r.second == n) return;
    ctx->requests.emplace_back(node, n);
}
static void VS_CC set_filter_error(const char* msg, VSFrameContext* ctx) noexcept { ctx->error = msg; }

static VSMap* VS_CC create_map() noexcept { return new VSMap(); }
static void VS_CC free_map(VSMap* m) noexcept { delete m; }
static void VS_CC map_set_error(VSMap* m, const char* msg) noexcept {
    m->clear();
    m->error = msg ? msg : "";
}
static const char* VS_CC map_get_error(const VSMap* m) noexcept { return m->error.empty() ? nullptr : m->error.c_str(); }
static int VS_CC map_delete_key(VSMap* m, const char* key) noexcept {
    for (size_t k = 0; k < m->e.size(); ++k)
        if (m->e[k].key == key) {
            for (VSNode* n : m->e[k].n) node_release(n);
            for (const VSFrame* f : m->e[k].fr) frame_release(f);
            m->e.erase(m->e.begin() + k);
            return 1;
        }
    return 0;
}
static int entry_size(const VSMap::Entry& x) {
    switch (x.type) {
    case ptInt: return (int)x.i.size();
    case ptFloat: return (int)x.f.size();
    case ptData: return (int)x.d.size();
    case ptVideoNode: return (int)x.n.size();
    case ptVideoFrame: return (int)x.fr.size();
    default: return 0;
    }
}
static int VS_CC map_num_elements(const VSMap* m, const char* key) noexcept {
    const VSMap::Entry* x = m->find(key);
    return x ? entry_size(*x) : -1;
}
static int VS_CC map_get_type(const VSMap* m, const char* key) noexcept {
    const VSMap::Entry* x = m->find(key);
    return x ? x->type : ptUnset;
}
template <typename T>
static const VSMap::Entry* typed(const VSMap* m, const char* key, int type, int index, int* err) {
    const VSMap::Entry* x = m->find(key);
    int e = peSuccess;
    if (!x) e = peUnset;
    else if (x->type != type) e = peType;
    else if (index < 0 || index >= entry_size(*x)) e = peIndex;
    if (err) *err = e;
    return e == peSuccess ? x : nullptr;
}
static int64_t VS_CC map_get_int(const VSMap* m, const char* key, int index, int* err) noexcept {
    const VSMap::Entry* x = typed<int64_t>(m, key, ptInt, index, err);
    return x ? x->i[index] : 0;
}
static int VS_CC map_get_int_sat(const VSMap* m, const char* key, int index, int* err) noexcept {
    int64_t v = map_get_int(m, key, index, err);
    return v > INT32_MAX ? INT32_MAX : v < INT32_MIN ? INT32_MIN : (int)v;
}
static const int64_t* VS_CC map_get_int_array(const VSMap* m, const char* key, int* err) noexcept {
    const VSMap::Entry* x = typed<int64_t>(m, key, ptInt, 0, err);
    return x ? x->i.data() : nullptr;
}
static VSMap::Entry& slot(VSMap* m, const char* key, int type, int append) {
    VSMap::Entry* x = m->find(key);
    if (x && (!append || x->type != type)) {
        map_delete_key(m, key);
        x = nullptr;
    }
    if (!x) {
        VSMap::Entry e;
        e.key = key;
        e.type = type;
        m->e.push_back(std::move(e));
        x = &m->e.back();
    }
    return *x;
}
static int VS_CC map_set_int(VSMap* m, const char* key, int64_t v, int append) noexcept {
    slot(m, key, ptInt, append).i.push_back(v);
    return 0;
}
static int VS_CC map_set_int_array(VSMap* m, const char* key, const int64_t* v, int size) noexcept {
    VSMap::Entry& x = slot(m, key, ptInt, 0);
    x.i.assign(v, v + size);
    return 0;
}
static double VS_CC map_get_float(const VSMap* m, const char* key, int index, int* err) noexcept {
    const VSMap::Entry* x = typed<double>(m, key, ptFloat, index, err);
    return x ? x->f[index] : 0.0;
}
static const double* VS_CC map_get_float_array(const VSMap* m, const char* key, int* err) noexcept {
    const VSMap::Entry* x = typed<double>(m, key, ptFloat, 0, err);
    return x ? x->f.data() : nullptr;
}
static int VS_CC map_set_float(VSMap* m, const char* key, double v, int append) noexcept {
    slot(m, key, ptFloat, append).f.push_back(v);
    return 0;
}
static const char* VS_CC map_get_data(const VSMap* m, const char* key, int index, int* err) noexcept {
    const VSMap::Entry* x = typed<std::string>(m, key, ptData, index, err);
    return x ? x->d[index].c_str() : nullptr;
}
static int VS_CC map_set_data(VSMap* m, const char* key, const char* data, int size, int type, int append) noexcept {
    (void)type;
    slot(m, key, ptData, append).d.push_back(size < 0 ? std::string(data) : std::string(data, size));
    return 0;
}
static VSNode* VS_CC map_get_node(const VSMap* m, const char* key, int index, int* err) noexcept {
    const VSMap::Entry* x = typed<VSNode*>(m, key, ptVideoNode, index, err);
    return x ? add_node_ref(x->n[index]) : nullptr;
}
static int VS_CC map_set_node(VSMap* m, const char* key, VSNode* n, int append) noexcept {
    slot(m, key, ptVideoNode, append).n.push_back(add_node_ref(n));
    return 0;
}

static void VS_CC log_message(int type, const char* msg, VSCore* core) noexcept {
    std::lock_guard<std::mutex> lk(core->log_mu);
    core->log.emplace_back(type, msg);
}

static int VS_CC get_api_version() noexcept { return VAPOURSYNTH_API_VERSION; }

static const VSAPI* api() {
    static VSAPI a = [] {
        VSAPI v;
        std::memset(&v, 0, sizeof(v));
        v.createVideoFilter = create_video_filter;
        v.freeNode = free_node;
        v.addNodeRef = add_node_ref;
        v.getVideoInfo = get_video_info;
        v.newVideoFrame = new_video_frame;
        v.freeFrame = free_frame;
        v.addFrameRef = add_frame_ref;
        v.getFramePropertiesRO = props_ro;
        v.getFramePropertiesRW = props_rw;
        v.getStride = get_stride;
        v.getReadPtr = get_read_ptr;
        v.getWritePtr = get_write_ptr;
        v.getVideoFrameFormat = get_frame_format;
        v.getFrameWidth = frame_width;
        v.getFrameHeight = frame_height;
        v.queryVideoFormat = query_format;
        v.getFrameFilter = get_frame_filter;
        v.requestFrameFilter = request_frame_filter;
        v.setFilterError = set_filter_error;
        v.createMap = create_map;
        v.freeMap = free_map;
        v.mapSetError = map_set_error;
        v.mapGetError = map_get_error;
        v.mapDeleteKey = map_delete_key;
        v.mapNumElements = map_num_elements;
        v.mapGetType = map_get_type;
        v.mapGetInt = map_get_int;
        v.mapGetIntSaturated = map_get_int_sat;
        v.mapGetIntArray = map_get_int_array;
        v.mapSetInt = map_set_int;
        v.mapSetIntArray = map_set_int_array;
        v.mapGetFloat = map_get_float;
        v.mapGetFloatArray = map_get_float_array;
        v.mapSetFloat = map_set_float;
        v.mapGetData = map_get_data;
        v.mapSetData = map_set_data;
        v.mapGetNode = map_get_node;
        v.mapSetNode = map_set_node;
        v.getAPIVersion = get_api_version;
        v.logMessage = log_message;
        return v;
    }();
    return &a;
}

// ------------------------------------------------------------------ plugin API

static int VS_CC config_plugin(const char* id, const char* ns, const char* name, int version, int api_version,
                               int flags, VSPlugin* p) noexcept {
    p->id = id;
    p->ns = ns;
    p->name = name;
    p->version = version;
    p->api = api_version;
    p->flags = flags;
    return 1;
}
static int VS_CC register_function(const char* name, const char* args, const char* ret, VSPublicFunction fn,
                                   void* data, VSPlugin* p) noexcept {
    p->funcs.push_back(RegFunction{name, args, ret, fn, data});
    return 1;
}
static const VSPLUGINAPI* plugin_api() {
    static VSPLUGINAPI a{get_api_version, config_plugin, register_function};
    return &a;
}

// ------------------------------------------------------------------ scheduling

static const VSFrame* produce(VSNode* node, int n, std::string& err) {
    if (n < 0 || (node->vi.numFrames > 0 && n >= node->vi.numFrames)) {
        err = "frame " + std::to_string(n) + " out of range";
        return nullptr;
    }
    if (!node->getframe) {
        if (n >= (int)node->frames.size() || !node->frames[n]) {
            err = "source frame " + std::to_string(n) + " not set";
            return nullptr;
        }
        return add_frame_ref(node->frames[n]);
    }
    {
        std::lock_guard<std::mutex> lk(node->mu);
        auto it = node->cache.find(n);
        if (it != node->cache.end()) return add_frame_ref(it->second);
    }
    VSFrameContext ctx;
    void* frame_data = nullptr;
    node->getframe(n, arInitial, node->instance, &frame_data, &ctx, node->core, api());
    for (auto& r : ctx.requests) {
        const VSFrame* f = produce(r.first, r.second, err);
        if (!f) {
            node->getframe(n, arError, node->instance, &frame_data, &ctx, node->core, api());
            for (auto& x : ctx.ready) frame_release(x.second);
            return nullptr;
        }
        ctx.ready.push_back({r, f});
    }
    const VSFrame* out = node->getframe(n, arAllFramesReady, node->instance, &frame_data, &ctx, node->core, api());
    for (auto& x : ctx.ready) frame_release(x.second);
    if (!ctx.error.empty() || !out) {
        err = ctx.error.empty() ? "filter returned no frame" : ctx.error;
        frame_release(out);
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(node->mu);
        auto it = node->cache.find(n);
        if (it == node->cache.end()) node->cache[n] = add_frame_ref(out);
    }
    return out;
}

// ------------------------------------------------------------------ argument checks

struct ArgDecl {
    std::string name, type;
    bool array = false, opt = false;
};

static std::vector<ArgDecl> parse_args(const std::string& s) {
    std::vector<ArgDecl> out;
    size_t a = 0;
    while (a < s.size()) {
        size_t b = s.find(';', a);
        if (b == std::string::npos) b = s.size();
        std::string item = s.substr(a, b - a);
        a = b + 1;
        if (item.empty()) continue;
        std::vector<std::string> parts;
        size_t c = 0;
        while (true) {
            size_t d = item.find(':', c);
            parts.push_back(item.substr(c, d == std::string::npos ? std::string::npos : d - c));
            if (d == std::string::npos) break;
            c = d + 1;
        }
        ArgDecl dcl;
        dcl.name = parts[0];
        dcl.type = parts.size() > 1 ? parts[1] : "";
        if (dcl.type.size() > 2 && dcl.type.compare(dcl.type.size() - 2, 2, "[]") == 0) {
            dcl.array = true;
            dcl.type.resize(dcl.type.size() - 2);
        }
        for (size_t k = 2; k < parts.size(); ++k)
            if (parts[k] == "opt") dcl.opt = true;
        out.push_back(dcl);
    }
    return out;
}

// The checks the core makes before calling a plugin function; the messages
// follow the core's wording.
static std::string check_args(const std::string& fname, const std::string& decl, VSMap& in) {
    const std::vector<ArgDecl> ds = parse_args(decl);
    for (const auto& x : in.e) {
        bool known = false;
        for (const auto& d : ds) known = known || d.name == x.key;
        if (!known) return fname + ": Function does not take argument(s) named " + x.key;
    }
    for (const auto& d : ds) {
        VSMap::Entry* x = in.find(d.name.c_str());
        if (!x) {
            if (!d.opt) return fname + ": argument " + d.name + " is required";
            continue;
        }
        const int n = entry_size(*x);
        if (!d.array && n != 1) return fname + ": argument " + d.name + " is not of array type but more than one value was supplied";
        int want = d.type == "int" ? ptInt : d.type == "float" ? ptFloat : d.type == "data" ? ptData
                 : d.type == "vnode" ? ptVideoNode : ptUnset;
        if (want == ptFloat && x->type == ptInt) {  // the core converts int to float arguments
            x->f.assign(x->i.begin(), x->i.end());
            x->i.clear();
            x->type = ptFloat;
        }
        if (x->type != want) return fname + ": argument " + d.name + " is not of the correct type";
    }
    return "";
}

// ------------------------------------------------------------------ C API for the tests

extern "C" {

typedef VSCore fh_core;

fh_core* fh_core_new(void) { return new VSCore(); }

void fh_core_free(fh_core* c) { delete c; }

int fh_load_plugin(fh_core* c, const char* path, char* err, int errlen) {
    void* dl = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!dl) {
        std::snprintf(err, errlen, "%s", dlerror());
        return -1;
    }
    auto init = reinterpret_cast<void(VS_CC*)(VSPlugin*, const VSPLUGINAPI*)>(dlsym(dl, "VapourSynthPluginInit2"));
    if (!init) {
        std::snprintf(err, errlen, "no VapourSynthPluginInit2 in %s", path);
        dlclose(dl);
        return -2;
    }
    std::unique_ptr<VSPlugin> p(new VSPlugin());
    p->dl = dl;
    init(p.get(), plugin_api());
    c->plugins.push_back(std::move(p));
    return (int)c->plugins.size() - 1;
}

const char* fh_plugin_field(fh_core* c, int idx, int field) {
    const VSPlugin& p = *c->plugins.at(idx);
    return field == 0 ? p.id.c_str() : field == 1 ? p.ns.c_str() : p.name.c_str();
}

int fh_plugin_int(fh_core* c, int idx, int field) {
    const VSPlugin& p = *c->plugins.at(idx);
    return field == 0 ? p.version : field == 1 ? p.api : p.flags;
}

int fh_function_count(fh_core* c, int idx) { return (int)c->plugins.at(idx)->funcs.size(); }

const char* fh_function_field(fh_core* c, int idx, int fn, int field) {
    const RegFunction& f = c->plugins.at(idx)->funcs.at(fn);
    return field == 0 ? f.name.c_str() : field == 1 ? f.args.c_str() : f.ret.c_str();
}

VSMap* fh_map_new(void) { return new VSMap(); }
void fh_map_free(VSMap* m) { delete m; }
void fh_map_append_int(VSMap* m, const char* key, int64_t v) { map_set_int(m, key, v, maAppend); }
void fh_map_append_float(VSMap* m, const char* key, double v) { map_set_float(m, key, v, maAppend); }
void fh_map_set_data(VSMap* m, const char* key, const char* s) { map_set_data(m, key, s, -1, dtUtf8, maReplace); }
void fh_map_append_node(VSMap* m, const char* key, VSNode* n) { map_set_node(m, key, n, maAppend); }

VSNode* fh_source_new(int cf, int st, int bits, int ssw, int ssh, int w, int h, int nframes) {
    VSNode* n = new VSNode();
    if (!query_format(&n->vi.format, cf, st, bits, ssw, ssh, nullptr)) {
        delete n;
        return nullptr;
    }
    n->vi.fpsNum = 25;
    n->vi.fpsDen = 1;
    n->vi.width = w;
    n->vi.height = h;
    n->vi.numFrames = nframes;
    n->name = "Source";
    n->frames.assign(nframes, nullptr);
    for (int k = 0; k < nframes; ++k) n->frames[k] = frame_new(n->vi.format, w, h);
    return n;
}

int fh_source_set_plane(VSNode* n, int k, int p, const void* data, ptrdiff_t stride) {
    if (k < 0 || k >= (int)n->frames.size() || p < 0 || p >= n->vi.format.numPlanes) return -1;
    VSFrame* f = n->frames[k];
    const int pw = frame_width(f, p), ph = frame_height(f, p);
    for (int y = 0; y < ph; ++y)
        std::memcpy(f->ptr[p] + y * f->stride[p], static_cast<const uint8_t*>(data) + y * stride,
                    (size_t)pw * f->fmt.bytesPerSample);
    return 0;
}

int fh_source_set_prop(VSNode* n, int k, const char* key, int64_t v) {
    if (k < 0 || k >= (int)n->frames.size()) return -1;
    map_set_int(&n->frames[k]->props, key, v, maReplace);
    return 0;
}

// Invokes ns.fn(args): the node it returns, or null with err set.
VSNode* fh_invoke(fh_core* c, const char* ns, const char* fn, VSMap* args, char* err, int errlen) {
    for (auto& p : c->plugins) {
        if (p->ns != ns) continue;
        for (auto& f : p->funcs) {
            if (f.name != fn) continue;
            const std::string fname = std::string(ns) + "." + fn;
            const std::string bad = check_args(fname, f.args, *args);
            if (!bad.empty()) {
                std::snprintf(err, errlen, "%s", bad.c_str());
                return nullptr;
            }
            VSMap out;
            f.fn(args, &out, f.data, c, api());
            if (!out.error.empty()) {
                std::snprintf(err, errlen, "%s", out.error.c_str());
                return nullptr;
            }
            VSMap::Entry* e = out.find("clip");
            if (!e || e->n.empty()) {
                std::snprintf(err, errlen, "%s: no clip returned", fname.c_str());
                return nullptr;
            }
            return add_node_ref(e->n[0]);
        }
        std::snprintf(err, errlen, "no function %s in namespace %s", fn, ns);
        return nullptr;
    }
    std::snprintf(err, errlen, "no plugin with namespace %s", ns);
    return nullptr;
}

int fh_node_info(VSNode* n, int* out) {
    const VSVideoInfo& v = n->vi;
    const int vals[8] = {v.format.colorFamily, v.format.sampleType, v.format.bitsPerSample, v.format.subSamplingW,
                         v.format.subSamplingH, v.width, v.height, v.numFrames};
    std::memcpy(out, vals, sizeof(vals));
    return 0;
}

const VSFrame* fh_get_frame(VSNode* n, int k, char* err, int errlen) {
    std::string e;
    const VSFrame* f = produce(n, k, e);
    if (!f) std::snprintf(err, errlen, "%s", e.c_str());
    return f;
}

// Produces frames 0..count-1 of n on nthreads threads (frame k on thread
// k % nthreads), as a frame-parallel VapourSynth core would; the results stay
// in the node's cache.  Returns the number of failures.
int fh_get_frames_parallel(VSNode* n, int count, int nthreads, char* err, int errlen) {
    std::atomic<int> failures{0};
    std::mutex emu;
    std::string first;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t] {
            for (int k = t; k < count; k += nthreads) {
                std::string e;
                const VSFrame* f = produce(n, k, e);
                if (!f) {
                    ++failures;
                    std::lock_guard<std::mutex> lk(emu);
                    if (first.empty()) first = e;
                }
                frame_release(f);
            }
        });
    for (auto& x : th) x.join();
    if (failures) std::snprintf(err, errlen, "%s", first.c_str());
    return failures;
}

int fh_frame_plane(const VSFrame* f, int p, const void** data, ptrdiff_t* stride, int* w, int* h) {
    if (p < 0 || p >= f->fmt.numPlanes) return -1;
    *data = f->ptr[p];
    *stride = f->stride[p];
    *w = frame_width(f, p);
    *h = frame_height(f, p);
    return 0;
}

int fh_frame_prop(const VSFrame* f, const char* key, int idx, int64_t* v) {
    int e = 0;
    *v = map_get_int(&f->props, key, idx, &e);
    return e;
}

int fh_frame_prop_count(const VSFrame* f, const char* key) { return map_num_elements(&f->props, key); }

void fh_frame_free(const VSFrame* f) { frame_release(f); }

void fh_node_free(VSNode* n) { node_release(n); }

int fh_log_count(fh_core* c) {
    std::lock_guard<std::mutex> lk(c->log_mu);
    return (int)c->log.size();
}

const char* fh_log_message(fh_core* c, int i, int* type) {
    std::lock_guard<std::mutex> lk(c->log_mu);
    *type = c->log.at(i).first;
    return c->log.at(i).second.c_str();
}

}  // extern "C"
