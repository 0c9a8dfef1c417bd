// capi.cpp -- C ABI of libswegl_hip.so (include/swegl_hip.h): contexts,
// resident scenes and surfaces, and the launch sequence of one render.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swegl_hip.h"
#include "hostmath.h"
#include "jobs.h"
#include "launch.h"

using namespace sw;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int rc, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return rc;
}

// shared with the asset loader (asset.cpp)
int sw_fail_msg(int rc, const std::string &msg)
{
    g_err = msg;
    return rc;
}

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess)                                                       \
            return fail(SWEGL_EIO, "%s: %s", #x, hipGetErrorString(e_));            \
    } while (0)

// grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int ensure(size_t bytes)
    {
        if (bytes <= n && p)
            return SWEGL_OK;
        release();
        size_t want = std::max<size_t>(bytes, 256);
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return fail(SWEGL_ENOMEM, "hipMalloc(%zu) failed", want);
        }
        n = want;
        return SWEGL_OK;
    }
    template <class T> T *as() const { return (T *)p; }
};

struct swegl_target {
    int kind;
    int32_t w, h;
    DevBuf buf;
};

// Per-viewport scratch of the context lane.  The front end's buffers come in
// kFrontSets sets (the eager path's pipeline, see swegl_ctx), so that the
// front ends of the next renders can run while earlier renders' rasters
// still read their own.
static const int kFrontSets = 3;
struct FrontScratch {
    DevBuf ndc, yes, rect, big;
};
struct ViewScratch {
    FrontScratch front[kFrontSets];
    DevBuf dof, tflat;
};

struct swegl_scene {
    DevBuf px, py, pz, nx, ny, nz, tu, tv, tri, prims, nodes, texels, texs, lights, world, nworld;
    uint32_t n_vertices = 0, n_tris = 0, n_prims = 0, n_lights = 0;
    bool no_material = false;   // some primitive has material_id -1
    std::vector<PrimDev> h_prims;
    std::vector<swegl_node_desc> h_nodes;
    std::vector<int32_t> roots, children;
    std::vector<uint32_t> prim_node;   // node of each primitive
    float ambient = 0, sun_I = 0;
    f3 sun = {0, 0, 0};
    std::vector<std::unique_ptr<ViewScratch>> views;
    uint64_t resident = 0;
};

struct ProfRec {
    const char *name;
    hipEvent_t a, b;
};

struct StageSlot {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool used = false;
};

// Device buffers one launch sequence uses.  The eager path keeps one set per
// context; every captured graph owns a set, since their addresses are baked
// into the graph's kernel arguments.
struct Resources {
    DevBuf jobs, tile_count, tile_offset, bins, ctr;
    size_t bytes() const { return jobs.n + tile_count.n + tile_offset.n + bins.n + ctr.n; }
};

// The addresses a launch sequence runs on: a Resources set, or regions of
// the allocation-free entry point's fixed scratch array.
struct ResPtr {
    char *jobs = nullptr;
    uint32_t *tile_count = nullptr, *tile_offset = nullptr, *bins = nullptr, *ctr = nullptr;
};
static ResPtr res_ptr(Resources &R)
{
    ResPtr p;
    p.jobs = R.jobs.as<char>();
    p.tile_count = R.tile_count.as<uint32_t>();
    p.tile_offset = R.tile_offset.as<uint32_t>();
    p.bins = R.bins.as<uint32_t>();
    p.ctr = R.ctr.as<uint32_t>();
    return p;
}

// Bump allocation over a fixed device region (the allocation-free path):
// take() hands out 256-byte aligned pieces and fails once the region is full.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0, want = 0;   // want: bytes asked for, met or not
    void *take(size_t bytes)
    {
        const size_t off = (used + 255) & ~(size_t)255;
        want = off + bytes;
        if (!base || off + bytes > cap)
            return nullptr;
        used = off + bytes;
        return base + off;
    }
};

// A captured launch sequence.  Only the resource addresses and grid sizes
// are baked in; every pointer the kernels follow lives in the job table,
// which is rewritten into one of the graph's two pinned buffers and copied
// to the device ahead of every replay on the same stream.
struct GraphEntry {
    Resources res;
    void *pinned[2] = {nullptr, nullptr};
    size_t job_bytes = 0;
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int next = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;   // swegl_ctx::graph_tick of the last replay (LRU)
    size_t bytes() const { return res.bytes() + 2 * job_bytes; }
    ~GraphEntry()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        for (int i = 0; i < 2; i++) {
            if (pinned[i])
                (void)hipHostFree(pinned[i]);
            if (done[i])
                (void)hipEventDestroy(done[i]);
        }
    }
};

// An execution lane: a stream with its own launch resources and upload
// staging.  The context's lane keeps per-viewport scratch with each scene;
// a server lane (own_views) carves all of a batch's scratch -- per-viewport
// buffers and launch resources -- from one grow-only region of its own, so
// that batches on different lanes run concurrently and a lane holds exactly
// one region whatever scenes pass through it.
struct Lane {
    hipStream_t stream = nullptr;
    Resources res[kFrontSets];   // the context lane uses one set per pipeline slot; others res[0]
    DevBuf region;
    StageSlot stage[4];
    int stage_next = 0;
    bool own_views = false;
    ViewScratch *scratch(swegl_scene *s, int use);
    int init(bool own);
    ~Lane();
};

struct swegl_ctx {
    int device = 0;
    int n_cus = 256;
    hipStream_t stream = nullptr;
    int tw = 32, th = 32;
    uint32_t big_tiles = 16; // bin limit (swegl_ctx_set_bin_limit)
    uint32_t groups_per_cu = 4;   // persistent raster workgroups per CU
    bool tile_auto = true;        // smaller tiles for waves with few of them (swegl_ctx_set_tile_auto)
    unsigned long long *h_stats = nullptr;   // pinned
    void *h_stage = nullptr;                 // pinned staging for read-backs (grown on demand)
    size_t h_stage_n = 0;
    bool prof = false;
    std::vector<ProfRec> pending;
    std::map<std::string, std::pair<double, uint64_t>> prof_acc;
    std::vector<hipEvent_t> event_pool;
    Lane lane0;   // the context stream
    std::map<std::string, std::unique_ptr<GraphEntry>> graphs;
    uint64_t graph_tick = 0;
    size_t graph_bytes = 0;                // device + pinned bytes of the cached graphs
    std::vector<Lane *> server_lanes;      // lanes of this context's servers (scene scratch)
    char *scratch = nullptr;               // this device's fixed scratch array (allocation-free path)
    // The eager path's two-stream pipeline.  A render with one launch wave
    // runs its front end (job table upload, k_vertex, k_front, k_bin_fill)
    // on fstream and its raster and depth of field on the context stream,
    // which waits for that front end.  Renders take the kFrontSets sets of
    // front-end buffers in turn (slot), so a render's front end only waits
    // for the render that last used its set, kFrontSets renders back: the
    // front ends run ahead, in the gaps the rasters and depth-of-field
    // passes of the renders before leave, and are done before their own
    // raster is due.  Any other work on the context stream (uploads, pose
    // updates, target writes, other render paths) bumps op_seq; the next
    // pipelined render then starts its front end after everything on the
    // context stream.
    hipStream_t fstream = nullptr;
    hipEvent_t f_done[kFrontSets] = {}, r_done[kFrontSets] = {}, tail = nullptr;
    bool r_recorded[kFrontSets] = {}, f_recorded[kFrontSets] = {};
    // pinned host copy of each slot's job table, reused once that slot's
    // last front end (which copied it) is done: f_done, no event of its own
    void *job_pinned[kFrontSets] = {};
    size_t job_pinned_n[kFrontSets] = {};
    int front_slot = 0;   // the next render's set of front-end buffers
    uint64_t op_seq = 1, pipe_seq = 0;   // equal: only pipelined renders since the last one
    std::vector<char> noalloc_tables;      // host job tables of the allocation-free path
    hipEvent_t noalloc_done = nullptr;     // their last upload
    bool noalloc_used = false;
};

// The graph cache keeps at most this many captured launch sequences and
// this many bytes of their resources; the least recently replayed go first.
// The limits hold a working set of a thousand launch shapes (a 256-scene
// set replayed one scene at a time needs 256 entries and under 1 GiB).
static const size_t kGraphCacheEntries = 1024;
static const size_t kGraphCacheBytes = (size_t)16 << 30;

// --------------------------------------------------------------- helpers
static hipEvent_t take_event(swegl_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

static void drain_profile(swegl_ctx *c)
{
    for (ProfRec &r : c->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &acc = c->prof_acc[r.name];
            acc.first += ms;
            acc.second += 1;
        }
        c->event_pool.push_back(r.a);
        c->event_pool.push_back(r.b);
    }
    c->pending.clear();
}

// Runs `launch` on the context stream, bracketed by events when profiling.
template <class F>
static hipError_t timed(swegl_ctx *c, hipStream_t stream, const char *name, bool capturing, F launch)
{
    if (!c->prof || capturing)
        return launch();
    ProfRec r{name, take_event(c), take_event(c)};
    (void)hipEventRecord(r.a, stream);
    hipError_t e = launch();
    (void)hipEventRecord(r.b, stream);
    c->pending.push_back(r);
    return e;
}

// A pinned staging slot for host->device uploads issued on a lane's stream;
// a slot is reused only after the copy that read it has completed.
static int stage_acquire(Lane &L, size_t bytes, StageSlot **out)
{
    StageSlot &s = L.stage[L.stage_next];
    L.stage_next = (L.stage_next + 1) % 4;
    if (s.used)
        HIPCHK(hipEventSynchronize(s.done));
    if (s.cap < bytes) {
        if (s.host)
            (void)hipHostFree(s.host);
        s.host = nullptr;
        size_t want = std::max<size_t>(bytes, 4096);
        HIPCHK(hipHostMalloc(&s.host, want, hipHostMallocDefault));
        s.cap = want;
    }
    *out = &s;
    return SWEGL_OK;
}

static int upload(Lane &L, void *dst, const void *src, size_t bytes, hipStream_t st = nullptr)
{
    if (!bytes)
        return SWEGL_OK;
    if (!st)
        st = L.stream;
    StageSlot *s;
    int rc = stage_acquire(L, bytes, &s);
    if (rc)
        return rc;
    memcpy(s->host, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, s->host, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(s->done, st));
    s->used = true;
    return SWEGL_OK;
}

// The context lane's scratch for the use-th viewport of scene s in a wave.
ViewScratch *Lane::scratch(swegl_scene *s, int use)
{
    while ((int)s->views.size() <= use)
        s->views.emplace_back(new ViewScratch);
    return s->views[use].get();
}

int Lane::init(bool own)
{
    own_views = own;
    if (own)
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (StageSlot &st : stage)
        // only tells the host that the copy out of the slot has ended: no fence
        HIPCHK(hipEventCreateWithFlags(&st.done, hipEventDisableTiming | hipEventDisableSystemFence));
    return SWEGL_OK;
}

Lane::~Lane()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    for (StageSlot &st : stage) {
        if (st.host)
            (void)hipHostFree(st.host);
        if (st.done)
            (void)hipEventDestroy(st.done);
    }
    if (own_views && stream)
        (void)hipStreamDestroy(stream);
}

// ------------------------------------------------- allocation-free scratch
// swegl_render_noalloc runs on this array instead of buffers the library
// allocates: SWEGL_NOALLOC_SCRATCH_BYTES on every device the library runs
// on, held from the code object's load until the process ends.  In host
// code the array's name is not a device address: swegl_ctx_create looks up
// each device's address once (hipGetSymbolAddress, with that device current)
// and every use goes through that address.
static __device__ __align__(256) char scratch[SWEGL_NOALLOC_SCRATCH_BYTES];
static const int kMaxDevices = 64;
static std::mutex g_device_mu;
static char *g_scratch_ptr[kMaxDevices];   // per device: the address of `scratch`
static bool g_kernels_resolved[kMaxDevices];

// Once per device, with the device current: the scratch array's address,
// and every kernel resolved (code object loaded, symbols looked up) before
// any stream capture, so that no kernel's first use falls inside a capture.
static int device_setup(int device, char **scratch_out)
{
    if (device >= kMaxDevices)
        return fail(SWEGL_ENODEV, "device %d above the library's limit of %d", device, kMaxDevices);
    std::lock_guard<std::mutex> g(g_device_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SWEGL_EIO, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = (char *)p;
    }
    if (!g_kernels_resolved[device]) {
        hipError_t e = resolve_kernels();
        if (e != hipSuccess)
            return fail(SWEGL_EIO, "resolving the kernels on device %d: %s", device, hipGetErrorString(e));
        g_kernels_resolved[device] = true;
    }
    *scratch_out = g_scratch_ptr[device];
    return SWEGL_OK;
}

// ------------------------------------------------------------------ ctx
extern "C" const char *swegl_last_error(void) { return g_err.c_str(); }

extern "C" int swegl_ctx_create(int device, swegl_ctx **out)
{
    if (!out)
        return fail(SWEGL_EINVAL, "null out");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return fail(SWEGL_ENODEV, "no HIP device %d (count %d)", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SWEGL_ENODEV, "device %d is %s, this build targets gfx950", device,
                    prop.gcnArchName);
    std::unique_ptr<swegl_ctx> c(new swegl_ctx);
    c->device = device;
    c->n_cus = prop.multiProcessorCount;
    int src = device_setup(device, &c->scratch);
    if (src)
        return src;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->fstream, hipStreamNonBlocking));
    // These events order work between the device's own streams: each kernel
    // ends with its own device-scope release, so they skip the system-scope
    // fence (an L2 write-back and invalidate on every XCD, ~10 us) that the
    // default event record adds for the host's and other devices' sake.
    const unsigned dev_ev = hipEventDisableTiming | hipEventDisableSystemFence;
    for (int k = 0; k < kFrontSets; k++) {
        HIPCHK(hipEventCreateWithFlags(&c->f_done[k], dev_ev));
        HIPCHK(hipEventCreateWithFlags(&c->r_done[k], dev_ev));
    }
    HIPCHK(hipEventCreateWithFlags(&c->tail, dev_ev));
    HIPCHK(hipHostMalloc((void **)&c->h_stats, SW_CTR_CONTROL * sizeof(uint32_t), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&c->noalloc_done, hipEventDisableTiming));
    c->lane0.stream = c->stream;
    int lrc = c->lane0.init(false);
    if (lrc)
        return lrc;
    *out = c.release();
    return SWEGL_OK;
}

extern "C" int swegl_ctx_destroy(swegl_ctx *c)
{
    if (!c)
        return SWEGL_OK;
    (void)hipSetDevice(c->device);
    if (c->fstream)
        (void)hipStreamSynchronize(c->fstream);
    (void)hipStreamSynchronize(c->stream);
    drain_profile(c);
    for (hipEvent_t e : c->event_pool)
        (void)hipEventDestroy(e);
    if (c->noalloc_done)
        (void)hipEventDestroy(c->noalloc_done);
    if (c->h_stats)
        (void)hipHostFree(c->h_stats);
    if (c->h_stage)
        (void)hipHostFree(c->h_stage);
    c->graphs.clear();
    for (int k = 0; k < kFrontSets; k++)
        for (hipEvent_t e : {c->f_done[k], c->r_done[k]})
            if (e)
                (void)hipEventDestroy(e);
    if (c->tail)
        (void)hipEventDestroy(c->tail);
    for (void *h : c->job_pinned)
        if (h)
            (void)hipHostFree(h);
    if (c->fstream)
        (void)hipStreamDestroy(c->fstream);
    (void)hipStreamDestroy(c->stream);
    c->lane0.stream = nullptr;   // destroyed above: the lane must not wait on it
    delete c;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_synchronize(swegl_ctx *c)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    HIPCHK(hipStreamSynchronize(c->fstream));
    HIPCHK(hipStreamSynchronize(c->stream));
    drain_profile(c);
    return SWEGL_OK;
}

extern "C" int swegl_device_info(swegl_ctx *c, int *n_cus, char *arch, int arch_cap, uint64_t *hbm)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device));
    if (n_cus)
        *n_cus = prop.multiProcessorCount;
    if (arch && arch_cap > 0)
        snprintf(arch, arch_cap, "%s", prop.gcnArchName);
    if (hbm)
        *hbm = prop.totalGlobalMem;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_set_tile(swegl_ctx *c, int tw, int th)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    static const int ok[][2] = {{8, 8}, {16, 16}, {32, 16}, {32, 32}, {64, 32}, {64, 64}};
    for (auto &p : ok)
        if (p[0] == tw && p[1] == th) {
            c->tw = tw;
            c->th = th;
            return SWEGL_OK;
        }
    return fail(SWEGL_EINVAL, "unsupported tile %dx%d", tw, th);
}

extern "C" int swegl_ctx_set_tile_auto(swegl_ctx *c, int on)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    c->tile_auto = on != 0;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_get_tile(swegl_ctx *c, int *tw, int *th)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    if (tw)
        *tw = c->tw;
    if (th)
        *th = c->th;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_set_bin_limit(swegl_ctx *c, uint32_t tiles)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    if (tiles > 4096)
        return fail(SWEGL_EINVAL, "bin limit %u above 4096", tiles);
    c->big_tiles = tiles;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_set_raster_groups(swegl_ctx *c, uint32_t per_cu)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    if (per_cu < 1 || per_cu > 16)
        return fail(SWEGL_EINVAL, "raster groups per CU %u outside [1,16]", per_cu);
    c->groups_per_cu = per_cu;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_set_profiling(swegl_ctx *c, int on)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    HIPCHK(hipStreamSynchronize(c->stream));
    drain_profile(c);
    c->prof_acc.clear();
    c->prof = on != 0;
    return SWEGL_OK;
}

extern "C" int swegl_ctx_kernel_times(swegl_ctx *c, char *names, int name_cap, double *ms,
                                      uint64_t *counts, int cap, int *n)
{
    if (!c)
        return fail(SWEGL_EINVAL, "null ctx");
    HIPCHK(hipStreamSynchronize(c->stream));
    drain_profile(c);
    int i = 0, pos = 0;
    for (auto &kv : c->prof_acc) {
        if (i >= cap)
            break;
        int len = (int)kv.first.size() + 1;
        if (pos + len > name_cap)
            break;
        memcpy(names + pos, kv.first.c_str(), len);
        pos += len;
        ms[i] = kv.second.second ? kv.second.first / (double)kv.second.second : 0.0;
        counts[i] = kv.second.second;
        i++;
    }
    if (n)
        *n = i;
    return SWEGL_OK;
}

// ---------------------------------------------------------------- targets
extern "C" int swegl_target_create(swegl_ctx *c, int kind, int32_t w, int32_t h, swegl_target **out)
{
    if (!c || !out)
        return fail(SWEGL_EINVAL, "null argument");
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768 || kind < 0 || kind > 2)
        return fail(SWEGL_EINVAL, "bad target %dx%d kind %d", w, h, kind);
    std::unique_ptr<swegl_target> t(new swegl_target);
    t->kind = kind;
    t->w = w;
    t->h = h;
    int rc = t->buf.ensure((size_t)w * h * 4);
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(t->buf.p, 0, (size_t)w * h * 4, c->stream));
    *out = t.release();
    return SWEGL_OK;
}

extern "C" int swegl_target_destroy(swegl_ctx *c, swegl_target *t)
{
    if (!t)
        return SWEGL_OK;
    if (c)
        (void)hipStreamSynchronize(c->stream);
    delete t;
    return SWEGL_OK;
}

// Device rows to host memory through the context's pinned staging buffer:
// one DMA into pinned memory, then a host copy, which is much faster than a
// device copy straight into pageable memory (a frame read back every frame
// is the presentation path of an interactive loop).
static int read_rows(swegl_ctx *c, const char *src, size_t src_pitch, size_t row, int32_t h, void *dst,
                     size_t pitch)
{
    const size_t n = row * (size_t)h;
    if (n > c->h_stage_n) {
        if (c->h_stage)
            HIPCHK(hipHostFree(c->h_stage));
        c->h_stage = nullptr;
        c->h_stage_n = 0;
        HIPCHK(hipHostMalloc(&c->h_stage, n, hipHostMallocDefault));
        c->h_stage_n = n;
    }
    if (src_pitch == row)
        HIPCHK(hipMemcpyAsync(c->h_stage, src, n, hipMemcpyDeviceToHost, c->stream));
    else
        HIPCHK(hipMemcpy2DAsync(c->h_stage, row, src, src_pitch, row, h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (pitch == row) {
        memcpy(dst, c->h_stage, n);
    } else {
        for (int32_t r = 0; r < h; r++)
            memcpy((char *)dst + (size_t)r * pitch, (const char *)c->h_stage + (size_t)r * row, row);
    }
    drain_profile(c);
    return SWEGL_OK;
}

extern "C" int swegl_target_read(swegl_ctx *c, swegl_target *t, void *dst, int32_t pitch)
{
    if (!c || !t || !dst || pitch < t->w * 4)
        return fail(SWEGL_EINVAL, "bad read arguments");
    return read_rows(c, (const char *)t->buf.p, (size_t)t->w * 4, (size_t)t->w * 4, t->h, dst, (size_t)pitch);
}

extern "C" int swegl_target_read_rect(swegl_ctx *c, swegl_target *t, int32_t x, int32_t y, int32_t w,
                                      int32_t h, void *dst, int32_t pitch)
{
    if (!c || !t || !dst || w <= 0 || h <= 0 || x < 0 || y < 0 || x + w > t->w || y + h > t->h ||
        pitch < w * 4)
        return fail(SWEGL_EINVAL, "bad read_rect arguments");
    const char *src = (const char *)t->buf.p + ((size_t)y * t->w + x) * 4;
    return read_rows(c, src, (size_t)t->w * 4, (size_t)w * 4, h, dst, (size_t)pitch);
}

extern "C" int swegl_target_copy(swegl_ctx *c, swegl_target *dst, swegl_target *src)
{
    if (!c || !dst || !src || dst->w != src->w || dst->h != src->h || dst->kind != src->kind)
        return fail(SWEGL_EINVAL, "targets differ");
    HIPCHK(hipMemcpyAsync(dst->buf.p, src->buf.p, (size_t)src->w * src->h * 4, hipMemcpyDeviceToDevice,
                          c->stream));
    return SWEGL_OK;
}

extern "C" int swegl_target_write(swegl_ctx *c, swegl_target *t, const void *src, int32_t pitch)
{
    if (!c || !t || !src || pitch < t->w * 4)
        return fail(SWEGL_EINVAL, "bad write arguments");
    HIPCHK(hipMemcpy2DAsync(t->buf.p, (size_t)t->w * 4, src, pitch, (size_t)t->w * 4, t->h,
                            hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SWEGL_OK;
}

namespace sw {
hipError_t launch_fill32(uint32_t *p, uint64_t n, uint32_t word, hipStream_t s);
}

extern "C" int swegl_target_fill(swegl_ctx *c, swegl_target *t, uint32_t word)
{
    if (!c || !t)
        return fail(SWEGL_EINVAL, "null argument");
    HIPCHK(launch_fill32(t->buf.as<uint32_t>(), (uint64_t)t->w * t->h, word, c->stream));
    return SWEGL_OK;
}

// ----------------------------------------------------------------- scenes
// Node matrices: original_to_world = parent * get_local_world_matrix()
// (vertex_shaders.hpp:16-28, model.hpp:56-61) and the normal matrix
// scale(rotation, scale) (vertex_shaders.hpp:63).
static int node_matrices(swegl_scene *s, std::vector<NodeDev> &out)
{
    const size_t nn = s->h_nodes.size();
    out.assign(nn, NodeDev{});
    std::vector<float> o2w(nn * 16);
    std::vector<char> seen(nn, 0);
    struct Item { int node; std::vector<float> parent; };
    std::vector<Item> stack;
    for (auto it = s->roots.rbegin(); it != s->roots.rend(); ++it)
        stack.push_back({*it, std::vector<float>(hm::kIdentity, hm::kIdentity + 16)});
    size_t guard = 0;
    while (!stack.empty()) {
        Item it = std::move(stack.back());
        stack.pop_back();
        if (++guard > 64 * (nn + 1))
            return fail(SWEGL_EINVAL, "node hierarchy has a cycle");
        const swegl_node_desc &d = s->h_nodes[it.node];
        float local[16], rs[16], m[16];
        hm::local_matrix(d.rotation, d.scale, d.translation, local, rs);
        hm::matmul(it.parent.data(), local, m);
        memcpy(&o2w[(size_t)it.node * 16], m, sizeof m);
        seen[it.node] = 1;
        NodeDev &nd = out[it.node];
        memcpy(nd.M.m, m, 12 * sizeof(float));
        memcpy(nd.RS.m, rs, 12 * sizeof(float));
        for (int k = (int)d.n_children - 1; k >= 0; k--)
            stack.push_back({s->children[d.first_child + k], std::vector<float>(m, m + 16)});
    }
    for (size_t i = 0; i < nn; i++)
        if (!seen[i])
            return fail(SWEGL_EINVAL, "node %zu is not reachable from a root node", i);
    return SWEGL_OK;
}

static int scene_pose(swegl_ctx *c, swegl_scene *s, const swegl_scene_desc *d)
{
    c->op_seq++;   // world, normals, nodes and lights change on the context stream
    s->h_nodes.assign(d->nodes, d->nodes + d->n_nodes);
    s->roots.assign(d->root_nodes, d->root_nodes + d->n_root_nodes);
    s->children.assign(d->child_indices, d->child_indices + d->n_child_indices);
    for (int32_t r : s->roots)
        if (r < 0 || (uint32_t)r >= d->n_nodes)
            return fail(SWEGL_EINVAL, "root node %d out of range", r);
    for (const swegl_node_desc &nd : s->h_nodes) {
        if ((uint64_t)nd.first_child + nd.n_children > d->n_child_indices)
            return fail(SWEGL_EINVAL, "child range out of range");
        for (uint32_t k = 0; k < nd.n_children; k++) {
            int32_t ci = d->child_indices[nd.first_child + k];
            if (ci < 0 || (uint32_t)ci >= d->n_nodes)
                return fail(SWEGL_EINVAL, "child index %d out of range", ci);
        }
    }
    std::vector<NodeDev> nodes;
    int rc = node_matrices(s, nodes);
    if (rc)
        return rc;
    std::vector<Light4> lights(std::max<uint32_t>(1, d->n_point_lights));
    for (uint32_t i = 0; i < d->n_point_lights; i++)
        lights[i] = Light4{d->point_lights[i].position[0], d->point_lights[i].position[1],
                           d->point_lights[i].position[2], d->point_lights[i].intensity};
    s->n_lights = d->n_point_lights;
    s->ambient = d->ambient_light_intensity;
    s->sun = f3{d->sun_direction[0], d->sun_direction[1], d->sun_direction[2]};
    s->sun_I = d->sun_intensity;
    if ((rc = s->nodes.ensure(nodes.size() * sizeof(NodeDev))) ||
        (rc = s->lights.ensure(lights.size() * sizeof(Light4))))
        return rc;
    if ((rc = upload(c->lane0, s->nodes.p, nodes.data(), nodes.size() * sizeof(NodeDev))) ||
        (rc = upload(c->lane0, s->lights.p, lights.data(), lights.size() * sizeof(Light4))))
        return rc;
    // v_world and normal_world for these poses
    WorldArgs W;
    W.px = s->px.as<float>();
    W.py = s->py.as<float>();
    W.pz = s->pz.as<float>();
    W.nx = s->nx.as<float>();
    W.ny = s->ny.as<float>();
    W.nz = s->nz.as<float>();
    W.tu = s->tu.as<float>();
    W.tv = s->tv.as<float>();
    W.prims = s->prims.as<PrimDev>();
    W.nodes = s->nodes.as<NodeDev>();
    W.n_prims = s->n_prims;
    W.n_vertices = s->n_vertices;
    W.world = s->world.as<float4>();
    W.nworld = s->nworld.as<float4>();
    HIPCHK(launch_world(W, c->stream));
    return SWEGL_OK;
}

extern "C" int swegl_scene_upload(swegl_ctx *c, const swegl_scene_desc *d, swegl_scene **out)
{
    if (!c || !d || !out)
        return fail(SWEGL_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    std::unique_ptr<swegl_scene> s(new swegl_scene);
    // ---- validate and flatten primitives in draw order (nodes, then primitives)
    std::vector<float> px, py, pz, nx, ny, nz, tu, tv;
    std::vector<uint32_t> tri;
    for (uint32_t ni = 0; ni < d->n_nodes; ni++) {
        const swegl_node_desc &nd = d->nodes[ni];
        if ((uint64_t)nd.first_primitive + nd.n_primitives > d->n_primitives)
            return fail(SWEGL_EINVAL, "node %u primitive range out of range", ni);
        for (uint32_t k = 0; k < nd.n_primitives; k++) {
            const swegl_primitive_desc &p = d->primitives[nd.first_primitive + k];
            // points and lines are kept but never drawn: the reference's marking
            // and drawing loops handle the triangle modes only (renderer.cpp:83-185, 197-229)
            if (p.mode < SWEGL_POINTS || p.mode > SWEGL_TRIANGLE_FAN)
                return fail(SWEGL_EINVAL, "primitive mode %d is not a primitive_t::index_mode_t", p.mode);
            // material_id -1: pixel_shader_t takes default_material (pixel_shaders.cpp:14-31);
            // the texture shaders would index materials[-1], so renders with them refuse
            // such a scene (swegl_scene::no_material)
            if (p.material_id < -1 || p.material_id >= (int)d->n_materials)
                return fail(SWEGL_EINVAL, "material_id %d out of range", p.material_id);
            const bool nomat = p.material_id == -1;
            const swegl_material_desc &m = nomat ? d->default_material : d->materials[p.material_id];
            if (!nomat && (m.texture_idx < -1 || m.texture_idx >= (int)d->n_textures))
                return fail(SWEGL_EINVAL, "texture_idx %d out of range", m.texture_idx);
            s->no_material = s->no_material || nomat;
            if (p.n_vertices && (!p.positions || !p.normals || !p.tex_coords))
                return fail(SWEGL_EINVAL, "null vertex arrays");
            for (uint32_t i = 0; i < p.n_indices; i++)
                if (p.indices[i] >= p.n_vertices)
                    return fail(SWEGL_EINVAL, "index %u out of range (%u vertices)", p.indices[i],
                                p.n_vertices);
            PrimDev pd;
            pd.first_vertex = (uint32_t)px.size();
            pd.n_vertices = p.n_vertices;
            pd.first_tri = (uint32_t)(tri.size() / 3);
            pd.node = ni;
            pd.tex = nomat ? -1 : m.texture_idx;
            pd.color = m.color;
            pd.double_sided = (!nomat && m.double_sided) ? 1 : 0;   // renderer.cpp:89-90
            const uint32_t base = pd.first_vertex;
            for (uint32_t v = 0; v < p.n_vertices; v++) {
                px.push_back(p.positions[3 * v]);
                py.push_back(p.positions[3 * v + 1]);
                pz.push_back(p.positions[3 * v + 2]);
                nx.push_back(p.normals[3 * v]);
                ny.push_back(p.normals[3 * v + 1]);
                nz.push_back(p.normals[3 * v + 2]);
                tu.push_back(p.tex_coords[2 * v]);
                tv.push_back(p.tex_coords[2 * v + 1]);
            }
            const uint32_t *I = p.indices;
            const size_t n = p.n_indices;
            // draw order and winding of renderer.cpp:197-229
            if (p.mode == SWEGL_TRIANGLE_STRIP)
                for (size_t i = 2; i < n; i++) {
                    tri.push_back(base + I[i - 2]);
                    tri.push_back(base + I[i - 1 + (i & 1)]);
                    tri.push_back(base + I[i - (i & 1)]);
                }
            if (p.mode == SWEGL_TRIANGLE_FAN)
                for (size_t i = 2; i < n; i++) {
                    tri.push_back(base + I[0]);
                    tri.push_back(base + I[i - 1]);
                    tri.push_back(base + I[i]);
                }
            if (p.mode == SWEGL_TRIANGLES)
                for (size_t i = 2; i < n; i += 3) {
                    tri.push_back(base + I[i - 2]);
                    tri.push_back(base + I[i - 1]);
                    tri.push_back(base + I[i]);
                }
            pd.n_tris = (uint32_t)(tri.size() / 3) - pd.first_tri;
            s->h_prims.push_back(pd);
        }
    }
    if (px.size() > 0x7FFFFFFFu || tri.size() / 3 > 0x7FFFFFFFu)
        return fail(SWEGL_EINVAL, "scene too large");
    s->n_vertices = (uint32_t)px.size();
    s->n_tris = (uint32_t)(tri.size() / 3);
    s->n_prims = (uint32_t)s->h_prims.size();
    // ---- textures
    std::vector<TexDev> texs(std::max<uint32_t>(1, d->n_textures));
    uint64_t texels = 0;
    for (uint32_t i = 0; i < d->n_textures; i++) {
        const swegl_texture_desc &t = d->textures[i];
        if (!t.bitmap || t.width == 0 || t.height == 0 || t.width > 65536 || t.height > 65536)
            return fail(SWEGL_EINVAL, "texture %u is empty or too large", i);
        texs[i] = TexDev{texels, (int32_t)t.width, (int32_t)t.height};
        texels += (uint64_t)t.width * t.height;
    }
    int rc;
    const size_t V = std::max<size_t>(1, px.size());
    if ((rc = s->px.ensure(V * 4)) || (rc = s->py.ensure(V * 4)) || (rc = s->pz.ensure(V * 4)) ||
        (rc = s->nx.ensure(V * 4)) || (rc = s->ny.ensure(V * 4)) || (rc = s->nz.ensure(V * 4)) ||
        (rc = s->tu.ensure(V * 4)) || (rc = s->tv.ensure(V * 4)) ||
        (rc = s->world.ensure(V * 16)) || (rc = s->nworld.ensure(V * 16)) ||
        (rc = s->tri.ensure(std::max<size_t>(3, tri.size()) * 4)) ||
        (rc = s->prims.ensure(std::max<size_t>(1, s->h_prims.size()) * sizeof(PrimDev))) ||
        (rc = s->texs.ensure(texs.size() * sizeof(TexDev))) ||
        (rc = s->texels.ensure(std::max<uint64_t>(1, texels) * 4)))
        return rc;
    auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
        if (!bytes)
            return SWEGL_OK;
        hipError_t e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess)
            return fail(SWEGL_EIO, "upload: %s", hipGetErrorString(e));
        return SWEGL_OK;
    };
    if ((rc = up(s->px, px.data(), px.size() * 4)) || (rc = up(s->py, py.data(), py.size() * 4)) ||
        (rc = up(s->pz, pz.data(), pz.size() * 4)) || (rc = up(s->nx, nx.data(), nx.size() * 4)) ||
        (rc = up(s->ny, ny.data(), ny.size() * 4)) || (rc = up(s->nz, nz.data(), nz.size() * 4)) ||
        (rc = up(s->tu, tu.data(), tu.size() * 4)) || (rc = up(s->tv, tv.data(), tv.size() * 4)) ||
        (rc = up(s->tri, tri.data(), tri.size() * 4)) ||
        (rc = up(s->prims, s->h_prims.data(), s->h_prims.size() * sizeof(PrimDev))) ||
        (rc = up(s->texs, texs.data(), texs.size() * sizeof(TexDev))))
        return rc;
    for (uint32_t i = 0; i < d->n_textures; i++) {
        const swegl_texture_desc &t = d->textures[i];
        HIPCHK(hipMemcpyAsync(s->texels.as<uint32_t>() + texs[i].offset, t.bitmap,
                              (size_t)t.width * t.height * 4, hipMemcpyHostToDevice, c->stream));
    }
    if ((rc = scene_pose(c, s.get(), d)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));   // host vectors go out of scope
    s->resident = V * (8 * 4 + 32) + tri.size() * 4 + texels * 4 +
                  s->h_prims.size() * sizeof(PrimDev) + d->n_nodes * sizeof(NodeDev);
    *out = s.release();
    return SWEGL_OK;
}

extern "C" int swegl_scene_update_pose(swegl_ctx *c, swegl_scene *s, const swegl_scene_desc *d)
{
    if (!c || !s || !d)
        return fail(SWEGL_EINVAL, "null argument");
    if (d->n_nodes != s->h_nodes.size())
        return fail(SWEGL_EINVAL, "pose update changes the node count");
    for (uint32_t i = 0; i < d->n_nodes; i++)
        if (d->nodes[i].first_primitive != s->h_nodes[i].first_primitive ||
            d->nodes[i].n_primitives != s->h_nodes[i].n_primitives)
            return fail(SWEGL_EINVAL, "pose update changes primitives");
    // renders on any lane read the scene's node and world buffers
    HIPCHK(hipDeviceSynchronize());
    return scene_pose(c, s, d);
}

extern "C" int swegl_scene_destroy(swegl_ctx *c, swegl_scene *s)
{
    if (!s)
        return SWEGL_OK;
    if (c)
        (void)hipStreamSynchronize(c->stream);
    delete s;
    return SWEGL_OK;
}

extern "C" int swegl_scene_bytes(swegl_ctx *c, swegl_scene *s, uint64_t *bytes)
{
    if (!c || !s || !bytes)
        return fail(SWEGL_EINVAL, "null argument");
    uint64_t b = s->resident;
    for (auto &v : s->views) {
        b += v->dof.n + v->tflat.n;
        for (const FrontScratch &f : v->front)
            b += f.ndc.n + f.yes.n + f.rect.n + f.big.n;
    }
    *bytes = b;
    return SWEGL_OK;
}

// ----------------------------------------------------------------- render
struct JobIn {
    swegl_scene *scene;
    const swegl_viewport_desc *vp;
};

static int check_viewport(const swegl_viewport_desc &v)
{
    if (!v.color || v.color->kind != SWEGL_TARGET_COLOR)
        return fail(SWEGL_EINVAL, "viewport colour target missing or not a colour surface");
    if (v.w <= 0 || v.h <= 0 || v.x < 0 || v.y < 0 || v.x + v.w > v.color->w || v.y + v.h > v.color->h)
        return fail(SWEGL_EINVAL, "viewport rect (%d,%d,%d,%d) outside the %dx%d surface", v.x, v.y, v.w,
                    v.h, v.color->w, v.color->h);
    if (v.w > 65535 * 8 || v.h > 65535 * 8)
        return fail(SWEGL_EINVAL, "viewport too large");
    if (!v.depth || v.depth->kind != SWEGL_TARGET_FLOAT || v.depth->w * v.depth->h != v.w * v.h)
        return fail(SWEGL_EINVAL, "viewport depth target must hold w*h floats");
    if (v.triangle_id && (v.triangle_id->kind != SWEGL_TARGET_UINT ||
                          v.triangle_id->w * v.triangle_id->h != v.w * v.h))
        return fail(SWEGL_EINVAL, "triangle id target must hold w*h uints");
    static_assert(SW_MAX_LAYERS == SWEGL_MAX_TRANSPARENCY_LAYERS, "layer limit");
    if (v.transparency_layer_count < 0 || v.transparency_layer_count > SWEGL_MAX_TRANSPARENCY_LAYERS)
        return fail(SWEGL_ENOTSUP, "transparency_layer_count %d outside [0,%d]", v.transparency_layer_count,
                    SWEGL_MAX_TRANSPARENCY_LAYERS);
    if (v.transparency_layer_count > 0 && v.post_shader != SWEGL_POST_NULL)
        return fail(SWEGL_ENOTSUP, "transparency layers with a depth-of-field post shader");
    if (v.pixel_shader < SWEGL_SHADER_PHONG_BILINEAR || v.pixel_shader > SWEGL_SHADER_LIGHTS_PHONG)
        return fail(SWEGL_ENOTSUP, "pixel shader %d is not a swegl shader type", v.pixel_shader);
    if (v.post_shader < SWEGL_POST_NULL || v.post_shader > SWEGL_POST_DEPTH_TAP9)
        return fail(SWEGL_EINVAL, "post shader %d (the reference dereferences an unset post shader)",
                    v.post_shader);
    if (v.post_shader != SWEGL_POST_NULL &&
        (v.dof_max_radius < 0 || v.dof_max_radius > SW_MAX_DOF_RADIUS))
        return fail(SWEGL_EINVAL, "dof_max_radius %d outside [0,%d]", v.dof_max_radius, SW_MAX_DOF_RADIUS);
    if (v.post_shader != SWEGL_POST_NULL &&
        (!std::isfinite(v.focal_distance) || !std::isfinite(v.focal_depth)))
        return fail(SWEGL_EINVAL, "focal distance and depth must be finite");
    return SWEGL_OK;
}

static bool rects_overlap(const swegl_viewport_desc &a, const swegl_viewport_desc &b)
{
    return a.x < b.x + b.w && b.x < a.x + a.w && a.y < b.y + b.h && b.y < a.y + a.h;
}
static bool conflicts(const swegl_viewport_desc &a, const swegl_viewport_desc &b)
{
    if (a.color == b.color && rects_overlap(a, b))
        return true;
    if (a.depth == b.depth || (a.triangle_id && a.triangle_id == b.triangle_id))
        return true;
    return false;
}

static void m34_from(const float *m16, m34 &o) { memcpy(o.m, m16, 12 * sizeof(float)); }


// One wave of independent jobs, prepared (buffers allocated, job table
// built) before anything is launched.
struct Wave {
    std::vector<JobDev> jd;
    uint32_t vblk = 0, tblk = 0, dblk = 0, tiles = 0, n_ctr = 0;
    int rmax = 0;             // largest DoF radius (sizes the DoF LDS)
    int tw = 32, th = 32;     // raster tile of this wave
    bool generic = false;     // some job uses a shader other than Phong + bilinear
    bool layers = false;      // the wave's (single) job has transparency layers
    uint64_t cap = 0;
    bool fixed = false;       // k_front's fixed layout (BatchDev::fixed; per job JobDev::tile_cap)
    uint64_t fixed_total = 0; // entries of the fixed layout, ahead of the counted layout in the bins
    uint32_t big_tiles = 0;   // bin limit of this wave (BatchDev::big_tiles)
    size_t job_offset = 0;   // byte offset of this wave's table in Resources::jobs
};

// Tile size of a wave: the context's, or with tile_auto, the next smaller
// size of 64x64 > 64x32 > 32x32 > 32x16 > 16x16 while the wave's viewports
// give fewer tiles than there are persistent raster workgroups (a small
// viewport otherwise leaves most CUs idle).
static void wave_tile(const swegl_ctx *c, const std::vector<JobIn> &jobs, int &tw, int &th)
{
    tw = c->tw;
    th = c->th;
    for (const JobIn &j : jobs)
        if (j.vp->transparency_layer_count > 0) {
            tw = SW_LAYER_TW;   // the in-order layer raster's tile
            th = SW_LAYER_TH;
            return;
        }
    if (!c->tile_auto)
        return;
    // A wave of few triangles (at most 4,096: a cube, a few quads) is all
    // shading, evenly spread: half as many tiles as workgroups is one round
    // of larger tiles, where a full count leaves a thin second round (the
    // 640x480 cube: 32x16 tiles 21.1k frames/s, 16x16 18.9k).
    uint64_t tris = 0;
    for (const JobIn &j : jobs)
        tris += j.scene->n_tris;
    const uint64_t want = (uint64_t)std::max(1, c->n_cus) * c->groups_per_cu / (tris <= 4096 ? 2 : 1);
    for (;;) {
        uint64_t tiles = 0;
        for (const JobIn &j : jobs)
            tiles += (uint64_t)((j.vp->w + tw - 1) / tw) * ((j.vp->h + th - 1) / th);
        if (tiles >= want || (tw <= 16 && th <= 16))
            return;
        if (tw > th)
            tw /= 2;    // 64x32 -> 32x32, 32x16 -> 16x16
        else
            th /= 2;    // 64x64 -> 64x32, 32x32 -> 32x16
    }
}

// Per-viewport scratch of one job: frustum coordinates, marks, tile
// rectangles, the big list, and with DoF the pre-DoF colour and code bytes,
// with layers the flattened layer colour.
struct JobScratch {
    float4 *ndc = nullptr;
    uint8_t *yes = nullptr;
    uint2 *rect = nullptr;
    uint32_t *big = nullptr;
    uint32_t *dof = nullptr;
    uint32_t *tflat = nullptr;
};
static size_t job_scratch_sizes(const swegl_scene *s, const swegl_viewport_desc &v, size_t sz[6])
{
    const size_t V = std::max<uint32_t>(1, s->n_vertices), T = std::max<uint32_t>(1, s->n_tris);
    const size_t P = (size_t)v.w * v.h;
    sz[0] = V * 16;
    sz[1] = V;
    sz[2] = T * 8;
    sz[3] = T * 4;
    sz[4] = v.post_shader != SWEGL_POST_NULL ? P * 5 : 0;   // colour, then one code byte per pixel
    sz[5] = v.transparency_layer_count > 0 ? P * 4 : 0;
    return sz[0] + sz[1] + sz[2] + sz[3] + sz[4] + sz[5];
}
static int job_scratch(Lane &L, Arena *arena, swegl_scene *s, int use, const swegl_viewport_desc &v,
                       JobScratch &out, int slot)
{
    size_t sz[6];
    job_scratch_sizes(s, v, sz);
    void *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (arena) {
        for (int k = 0; k < 6; k++)
            if (sz[k] && !(p[k] = arena->take(sz[k])))
                return fail(SWEGL_ERANGE, "allocation-free render: the viewports' scratch needs more than "
                                          "the %zu bytes of the fixed scratch array", arena->cap);
    } else {
        ViewScratch &vs = *L.scratch(s, use);
        FrontScratch &f = vs.front[slot];
        DevBuf *b[6] = {&f.ndc, &f.yes, &f.rect, &f.big, &vs.dof, &vs.tflat};
        for (int k = 0; k < 6; k++)
            if (sz[k]) {
                int rc = b[k]->ensure(sz[k]);
                if (rc)
                    return rc;
                p[k] = b[k]->p;
            }
    }
    out.ndc = (float4 *)p[0];
    out.yes = (uint8_t *)p[1];
    out.rect = (uint2 *)p[2];
    out.big = (uint32_t *)p[3];
    out.dof = (uint32_t *)p[4];
    out.tflat = (uint32_t *)p[5];
    return SWEGL_OK;
}

static int build_wave(swegl_ctx *c, Lane &L, const std::vector<JobIn> &jobs, Wave &W, Arena *arena, int slot)
{
    const uint32_t nj = (uint32_t)jobs.size();
    W.jd.assign(nj, JobDev{});
    wave_tile(c, jobs, W.tw, W.th);
    std::map<swegl_scene *, int> scene_uses;
    uint64_t tris_total = 0;
    for (uint32_t j = 0; j < nj; j++) {
        swegl_scene *s = jobs[j].scene;
        const swegl_viewport_desc &v = *jobs[j].vp;
        JobScratch vs;
        int rc = job_scratch(L, arena, s, scene_uses[s]++, v, vs, slot);
        if (rc)
            return rc;
        const bool dof = v.post_shader != SWEGL_POST_NULL;
        JobDev &J = W.jd[j];
        memset(&J, 0, sizeof J);
        J.px = s->px.as<float>();
        J.py = s->py.as<float>();
        J.pz = s->pz.as<float>();
        J.nx = s->nx.as<float>();
        J.ny = s->ny.as<float>();
        J.nz = s->nz.as<float>();
        J.tu = s->tu.as<float>();
        J.tv = s->tv.as<float>();
        J.tri = s->tri.as<uint32_t>();
        J.prims = s->prims.as<PrimDev>();
        J.nodes = s->nodes.as<NodeDev>();
        J.texels = s->texels.as<uint32_t>();
        J.texs = s->texs.as<TexDev>();
        J.lights = s->lights.as<Light4>();
        J.n_prims = s->n_prims;
        J.n_lights = s->n_lights;
        J.n_vertices = s->n_vertices;
        J.n_tris = s->n_tris;
        J.ambient = s->ambient;
        J.sun = s->sun;
        J.sun_I = s->sun_I;
        J.world = s->world.as<float4>();
        J.nworld = s->nworld.as<float4>();
        J.ndc = vs.ndc;
        J.yes = vs.yes;
        J.rect = vs.rect;
        J.big = vs.big;
        m34_from(v.view, J.V.view);
        m34_from(v.projection, J.V.proj);
        hm::viewport_matrix(v.x, v.y, v.w, v.h, &J.V.m00, &J.V.m03, &J.V.m11, &J.V.m13);
        J.V.vx = v.x;
        J.V.vy = v.y;
        J.V.vw = v.w;
        J.V.vh = v.h;
        J.V.cam = f3{v.camera_position[0], v.camera_position[1], v.camera_position[2]};
        J.surface = v.color->buf.as<uint32_t>();
        J.surface_pitch = v.color->w;
        J.dofc = nullptr;
        if (dof) {
            J.out_color = vs.dof;
            J.dofc = (uint8_t *)vs.dof + (size_t)v.w * v.h * 4;
            J.out_pitch = v.w;
            J.out_ox = v.x;
            J.out_oy = v.y;
        } else {
            J.out_color = J.surface;
            J.out_pitch = J.surface_pitch;
            J.out_ox = 0;
            J.out_oy = 0;
        }
        J.zbuf = v.depth->buf.as<float>();
        J.tid = v.triangle_id ? v.triangle_id->buf.as<uint32_t>() : nullptr;
        J.shader = v.pixel_shader;
        J.layers = v.transparency_layer_count;
        J.tflat = v.transparency_layer_count > 0 ? vs.tflat : nullptr;
        W.layers = W.layers || v.transparency_layer_count > 0;
        W.generic = W.generic || v.pixel_shader != SWEGL_SHADER_PHONG_BILINEAR;
        J.post = v.post_shader;
        J.fdist = v.focal_distance;
        J.fdepth = v.focal_depth;
        J.rmax = v.dof_max_radius;
        J.tw = W.tw;
        J.th = W.th;
        J.tiles_x = (v.w + W.tw - 1) / W.tw;
        J.tiles_y = (v.h + W.th - 1) / W.th;
        J.tile_base = W.tiles;
        W.tiles += (uint32_t)(J.tiles_x * J.tiles_y);
        J.vblock0 = W.vblk;
        W.vblk += (s->n_vertices + 255) / 256;
        J.tblock0 = W.tblk;
        W.tblk += (s->n_tris + 255) / 256;
        J.dblock0 = W.dblk;
        if (dof) {
            W.dblk += dof_blocks(v.w, v.h);
            W.rmax = std::max(W.rmax, v.dof_max_radius);
        }
        tris_total += s->n_tris;
    }
    W.n_ctr = SW_CTR_BIG0 + 2 * nj;   // big-list lengths, then fixed-layout overflow words
    W.big_tiles = std::max<uint32_t>(1, W.layers ? std::max<uint32_t>(c->big_tiles, SW_LAYER_BIG_TILES)
                                                 : c->big_tiles);
    // every binned triangle contributes at most big_tiles entries
    W.cap = std::max<uint64_t>(4096, tris_total * W.big_tiles);
    if (W.cap > 0xFFFFFFF0ull)
        return fail(SWEGL_ERANGE, "too many triangles in one batch");
    // k_front's fixed layout (the context lane's buffers and server lanes'
    // regions, whose request bound counts it; not the allocation-free path's
    // fixed array; no layers): each job's tiles get room for sixteen times
    // the job's per-tile share of 1.25 entries per triangle, at least 512 (a
    // batch mixes scenes of very different densities, and a crowded tile --
    // a horizon, a stack of meshes -- holds 8-19 times its job's share); a
    // job whose busiest tile needs more takes the filled layout, the others
    // keep theirs
    W.fixed = false;
    W.fixed_total = 0;
    if ((!arena || L.own_views) && !W.layers && W.tiles && tris_total < 0x80000000ull) {   // (SW_BIN_CHECK: top bit)
        uint64_t total = 0;
        for (JobDev &J : W.jd) {
            const uint64_t tiles = (uint64_t)J.tiles_x * J.tiles_y;
            const uint64_t per = tiles ? ((uint64_t)J.n_tris * 20 + tiles - 1) / tiles : 0;   // 16 x 1.25
            const uint64_t tc = (std::max<uint64_t>(512, per) + 63) & ~(uint64_t)63;
            J.tile_cap = (uint32_t)std::min<uint64_t>(tc, 0xFFFFFFC0ull);
            J.bin_base = total;
            total += tiles * J.tile_cap;
        }
        if (total <= 0xFFFFFFF0ull) {   // entry offsets are 32-bit
            W.fixed = true;
            W.fixed_total = total;
        }
    }
    if (!W.fixed)
        for (JobDev &J : W.jd) {
            J.tile_cap = 0;
            J.bin_base = 0;
        }
    return SWEGL_OK;
}

// Place each wave's job table in Resources::jobs; returns the bytes used.
// Every call that uploads or copies job tables places them this way, so a
// replayed graph finds each wave's table where it was captured.
static size_t place_job_tables(std::vector<Wave> &waves)
{
    size_t job_bytes = 0;
    for (Wave &W : waves) {
        W.job_offset = job_bytes;
        job_bytes += (W.jd.size() * sizeof(JobDev) + 255) & ~(size_t)255;
    }
    return job_bytes;
}

// Bytes of each launch resource for a list of waves: job tables, tile
// counters (one padded line per tile), tile offsets, bins, counter words.
// The counter block always holds the statistics and control words that
// read_stats copies.
static void resource_sizes(std::vector<Wave> &waves, size_t sz[5])
{
    const size_t job_bytes = place_job_tables(waves);
    uint32_t tiles = 0, n_ctr = SW_CTR_BIG0;
    uint64_t cap = 0;
    for (Wave &W : waves) {
        tiles = std::max(tiles, W.tiles);
        n_ctr = std::max(n_ctr, W.n_ctr);
        cap = std::max(cap, W.fixed_total + W.cap);   // the fixed layout, then the counted one
    }
    sz[0] = job_bytes;
    sz[1] = (size_t)tiles * SW_TILE_PAD * 4 + 8;
    sz[2] = (size_t)tiles * 8 + 8;   // offsets [0, tiles], then the claim order
    sz[3] = cap * 4;
    sz[4] = (size_t)n_ctr * 4;
}

// Size a resource set for a list of waves (grow-only).
static int size_resources(Resources &R, std::vector<Wave> &waves)
{
    size_t sz[5];
    resource_sizes(waves, sz);
    int rc;
    if ((rc = R.jobs.ensure(sz[0])) || (rc = R.tile_count.ensure(sz[1])) || (rc = R.tile_offset.ensure(sz[2])) ||
        (rc = R.bins.ensure(sz[3])) || (rc = R.ctr.ensure(sz[4])))
        return rc;
    return SWEGL_OK;
}

// The same resources carved from an arena (the allocation-free path).
static int carve_resources(Arena &A, std::vector<Wave> &waves, ResPtr &out)
{
    size_t sz[5];
    resource_sizes(waves, sz);
    void *p[5];
    for (int k = 0; k < 5; k++)
        if (!(p[k] = A.take(sz[k])))
            return fail(SWEGL_ERANGE, "allocation-free render needs %zu bytes of scratch or more, the fixed "
                                      "scratch array holds %zu", A.want, A.cap);
    out.jobs = (char *)p[0];
    out.tile_count = (uint32_t *)p[1];
    out.tile_offset = (uint32_t *)p[2];
    out.bins = (uint32_t *)p[3];
    out.ctr = (uint32_t *)p[4];
    return SWEGL_OK;
}

// Enqueue the kernels of one wave (no allocation: safe under graph capture).
// `untimed`: no profiling events (capture, and the allocation-free path).
// The front-end stages go on `fs` (the lane's stream when null); with
// another stream, `fdone` is recorded there after them and the lane's
// stream waits for it before the raster.
static int launch_wave(swegl_ctx *c, Lane &L, const ResPtr &R, const Wave &W, bool first_wave, bool untimed,
                       bool want_stats, hipStream_t fs = nullptr, hipEvent_t fdone = nullptr)
{
    const bool capturing = untimed;
    BatchDev B;
    B.jobs = (const JobDev *)(R.jobs + W.job_offset);
    B.n_jobs = (uint32_t)W.jd.size();
    B.n_tiles = W.tiles;
    B.tile_count = R.tile_count;
    B.tile_offset = R.tile_offset;
    B.tile_order = R.tile_offset + W.tiles + 1;
    B.tile_cursor = R.tile_count + 1;
    B.bins_fixed = R.bins;
    B.bins = R.bins + W.fixed_total;
    B.bin_capacity = (uint32_t)W.cap;
    B.fixed = W.fixed ? 1u : 0u;
    B.ctr = R.ctr;
    B.big_tiles = W.big_tiles;
    B.stats = want_stats ? 1u : 0u;
    // statistics are zeroed once per render by the first wave; control words
    // and big-list lengths by every wave
    const uint32_t zero0 = first_wave ? 0 : SW_CTR_CONTROL;
    const uint32_t groups = (uint32_t)std::max(1, c->n_cus) * c->groups_per_cu;
    hipStream_t st = L.stream;
#define STAGE(name, strm, call)                                                        \
    do {                                                                               \
        hipError_t e_ = timed(c, strm, name, capturing, [&] { return call; });          \
        if (e_ != hipSuccess)                                                          \
            return fail(SWEGL_EIO, "%s launch: %s", name, hipGetErrorString(e_));       \
    } while (0)
    if (!fs)
        fs = st;
    STAGE("vertex", fs, launch_vertex(B, W.vblk, zero0, W.n_ctr - zero0, fs));
    STAGE("front", fs, launch_front(B, W.tblk, fs));
    // with two streams, the fill's own completion records fdone (no
    // separate marker on fstream, which would hold up the next front end)
    STAGE("bin_fill", fs, launch_bin_fill(B, W.tblk, fs, fs != st ? fdone : nullptr));
    if (fs != st)
        HIPCHK(hipStreamWaitEvent(st, fdone, 0));
    if (W.layers) {
        const uint32_t pixels = (uint32_t)(W.jd[0].V.vw * W.jd[0].V.vh);
        STAGE("raster", st, launch_layers(B, W.generic, pixels, st));
    } else {
        STAGE("raster", st, launch_raster(B, W.tw, W.th, groups, W.generic, st));
        STAGE("dof", st, launch_dof(B, W.dblk, W.rmax, st));
    }
#undef STAGE
    return SWEGL_OK;
}

static int prepare(swegl_ctx *c, Lane &L, const std::vector<JobIn> &all, std::vector<Wave> &waves,
                   Arena *arena = nullptr, int slot = 0)
{
    HIPCHK(hipSetDevice(c->device));
    for (const JobIn &j : all) {
        if (!j.scene)
            return fail(SWEGL_EINVAL, "null scene");
        int rc = check_viewport(*j.vp);
        if (rc)
            return rc;
        const int k = j.vp->pixel_shader;
        const bool textured = k <= SWEGL_SHADER_FLAT_NEAREST || k == SWEGL_SHADER_TEXTURE_NEAREST ||
                              k == SWEGL_SHADER_TEXTURE_BILINEAR;
        if (textured && j.scene->no_material)
            return fail(SWEGL_EINVAL, "a primitive has material_id -1, which the texture shaders index "
                                      "(materials[-1], pixel_shaders.cpp:213-215, 288-290)");
    }
    // waves: a job conflicting with an earlier job of the current wave (same
    // surface region or planes) starts a new wave, which keeps the
    // reference's viewport order for overlapping viewports.
    std::vector<std::vector<JobIn>> groups(1);
    for (const JobIn &j : all) {
        bool clash = false;
        for (const JobIn &k : groups.back())
            if (conflicts(*j.vp, *k.vp) || j.vp->transparency_layer_count > 0 ||
                k.vp->transparency_layer_count > 0)
                clash = true;   // a job with layers reads and writes outside its viewport: alone
        if (clash)
            groups.emplace_back();
        groups.back().push_back(j);
    }
    waves.clear();
    for (size_t w = 0; w < groups.size(); w++) {
        if (groups[w].empty())
            continue;
        waves.emplace_back();
        int rc = build_wave(c, L, groups[w], waves.back(), arena, slot);
        if (rc)
            return rc;
    }
    return SWEGL_OK;
}

static int read_stats(swegl_ctx *c, Lane &L, const uint32_t *ctr, swegl_frame_stats *stats)
{
    HIPCHK(hipMemcpyAsync(c->h_stats, ctr, SW_CTR_CONTROL * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          L.stream));
    HIPCHK(hipStreamSynchronize(L.stream));
    drain_profile(c);
    const unsigned long long *sl = (const unsigned long long *)((const uint32_t *)c->h_stats + SW_CTR_STATS);
    uint64_t sum[8] = {0};
    for (int k = 0; k < SW_STAT_SLOTS; k++)
        for (int i = 0; i < 8; i++)
            sum[i] += sl[k * 8 + i];
    stats->fragments = sum[SW_ST_FRAGMENTS];
    stats->nan_fragments = sum[SW_ST_NAN];
    stats->drawn_subtriangles = sum[SW_ST_DRAWN];
    stats->bin_entries = sum[SW_ST_BINS];
    stats->texel_oob = sum[SW_ST_OOB];
    stats->big_triangles = sum[SW_ST_BIG];
    return SWEGL_OK;
}

// No viewport, no wave: the statistics are zero (nothing ran to count them).
static int no_waves(swegl_frame_stats *stats)
{
    if (stats)
        memset(stats, 0, sizeof *stats);
    return SWEGL_OK;
}

// A server lane's batch: the scratch is measured by a pass over a region of
// unbounded size at a dummy base, the lane's region grown to that size if
// needed, then the batch prepared again on the region.
static int prepare_in_region(swegl_ctx *c, Lane &L, const std::vector<JobIn> &all, std::vector<Wave> &waves,
                             ResPtr &R)
{
    Arena M;
    M.base = (char *)(uintptr_t)256;   // never dereferenced: sizes only
    M.cap = (size_t)1 << 62;
    int rc = prepare(c, L, all, waves, &M);
    if (rc)
        return rc;
    if (waves.empty())
        return SWEGL_OK;
    if ((rc = carve_resources(M, waves, R)))
        return rc;
    if ((rc = L.region.ensure(M.used)))
        return rc;
    Arena A;
    A.base = L.region.as<char>();
    A.cap = L.region.n;
    if ((rc = prepare(c, L, all, waves, &A)))
        return rc;
    return carve_resources(A, waves, R);
}

static int render_jobs(swegl_ctx *c, Lane &L, const std::vector<JobIn> &all, swegl_frame_stats *stats)
{
    std::vector<Wave> waves;
    ResPtr R;
    int rc;
    if (L.own_views) {
        if ((rc = prepare_in_region(c, L, all, waves, R)))
            return rc;
        if (waves.empty())
            return no_waves(stats);
    } else {
        // the context lane: the front-end buffers of this render's slot
        const int p = c->front_slot;
        c->front_slot = (p + 1) % kFrontSets;
        if ((rc = prepare(c, L, all, waves, nullptr, p)))
            return rc;
        if (waves.empty())
            return no_waves(stats);
        Resources &RS = L.res[p];
        if ((rc = size_resources(RS, waves)))
            return rc;
        R = res_ptr(RS);
        if (waves.size() == 1 && !waves[0].layers && !c->prof) {
            // Pipelined (not while kernels are timed, so that each stage's
            // events time that stage alone): the front end on fstream, after
            // the raster that last used this slot's buffers, or after
            // everything on the context stream when other work was queued
            // there since the last pipelined render
            hipStream_t fs = c->fstream;
            if (c->pipe_seq != c->op_seq) {
                HIPCHK(hipEventRecord(c->tail, c->stream));
                HIPCHK(hipStreamWaitEvent(fs, c->tail, 0));
            } else if (c->r_recorded[p]) {
                HIPCHK(hipStreamWaitEvent(fs, c->r_done[p], 0));
            }
            c->r_recorded[p] = false;
            c->pipe_seq = c->op_seq - 1;   // until this render is fully queued
            const size_t jb = waves[0].jd.size() * sizeof(JobDev);
            if (c->f_recorded[p])
                HIPCHK(hipEventSynchronize(c->f_done[p]));   // the slot's last copy has ended
            if (c->job_pinned_n[p] < jb) {
                if (c->job_pinned[p])
                    (void)hipHostFree(c->job_pinned[p]);
                c->job_pinned[p] = nullptr;
                c->job_pinned_n[p] = 0;
                HIPCHK(hipHostMalloc(&c->job_pinned[p], std::max<size_t>(jb, 4096), hipHostMallocDefault));
                c->job_pinned_n[p] = std::max<size_t>(jb, 4096);
            }
            memcpy(c->job_pinned[p], waves[0].jd.data(), jb);
            c->f_recorded[p] = true;
            hipError_t ce = hipMemcpyAsync(R.jobs + waves[0].job_offset, c->job_pinned[p], jb,
                                           hipMemcpyHostToDevice, fs);
            if ((rc = ce != hipSuccess ? fail(SWEGL_EIO, "job table copy: %s", hipGetErrorString(ce)) : 0) ||
                (rc = launch_wave(c, L, R, waves[0], true, false, stats != nullptr, fs, c->f_done[p]))) {
                // whatever reached fstream is waited for by the context stream
                (void)hipEventRecord(c->f_done[p], fs);
                (void)hipStreamWaitEvent(c->stream, c->f_done[p], 0);
                return rc;
            }
            HIPCHK(hipEventRecord(c->r_done[p], c->stream));
            c->r_recorded[p] = true;
            c->pipe_seq = c->op_seq;
            return stats ? read_stats(c, L, R.ctr, stats) : SWEGL_OK;
        }
        c->op_seq++;   // serial: the next pipelined render waits for it
    }
    for (size_t w = 0; w < waves.size(); w++) {
        if ((rc = upload(L, R.jobs + waves[w].job_offset, waves[w].jd.data(),
                         waves[w].jd.size() * sizeof(JobDev))) ||
            (rc = launch_wave(c, L, R, waves[w], w == 0, false, stats != nullptr)))
            return rc;
    }
    return stats ? read_stats(c, L, R.ctr, stats) : SWEGL_OK;
}

// The allocation-free variant of render_jobs: every piece of device scratch
// comes from the fixed array (per-viewport scratch first, then the launch
// resources), the job tables go in by a copy from host memory the call owns,
// and nothing is allocated on the device or pinned on the host.  Sizes that
// do not fit are refused before anything is enqueued.
static int render_jobs_noalloc(swegl_ctx *c, const std::vector<JobIn> &all, swegl_frame_stats *stats)
{
    c->op_seq++;
    if (!c->scratch)
        return fail(SWEGL_EIO, "no scratch array address for device %d", c->device);
    Arena A;
    A.base = c->scratch;
    A.cap = SWEGL_NOALLOC_SCRATCH_BYTES;
    std::vector<Wave> waves;
    int rc = prepare(c, c->lane0, all, waves, &A);
    if (rc)
        return rc;
    if (waves.empty())
        return no_waves(stats);
    ResPtr R;
    if ((rc = carve_resources(A, waves, R)))
        return rc;
    size_t job_bytes = 0;
    for (const Wave &W : waves)
        job_bytes = std::max(job_bytes, W.job_offset + W.jd.size() * sizeof(JobDev));
    // the host copy of the tables is rewritten only once the previous call's
    // upload from it has completed
    if (c->noalloc_used)
        HIPCHK(hipEventSynchronize(c->noalloc_done));
    std::vector<char> &tables = c->noalloc_tables;
    if (tables.size() < job_bytes)
        tables.resize(job_bytes);
    for (const Wave &W : waves)
        memcpy(tables.data() + W.job_offset, W.jd.data(), W.jd.size() * sizeof(JobDev));
    HIPCHK(hipMemcpyAsync(R.jobs, tables.data(), job_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->noalloc_done, c->stream));
    c->noalloc_used = true;
    for (size_t w = 0; w < waves.size(); w++)
        if ((rc = launch_wave(c, c->lane0, R, waves[w], w == 0, true, stats != nullptr)))
            return rc;
    return stats ? read_stats(c, c->lane0, R.ctr, stats) : SWEGL_OK;
}

// ----------------------------------------------------------- graph path

// A captured graph holds every launch argument by value: grids, tile shapes,
// buffer capacities, LDS sizes and kernel variants.  The key names all of
// them, so a replay runs only where the capture's launches fit.
static std::string graph_key(swegl_ctx *c, const std::vector<Wave> &waves)
{
    std::string k = std::to_string(c->big_tiles) + "/" + std::to_string(c->groups_per_cu);
    for (const Wave &W : waves) {
        k += "|" + std::to_string(W.tw) + "x" + std::to_string(W.th) + ":" + std::to_string(W.jd.size()) + ":" +
             std::to_string(W.vblk) + "," + std::to_string(W.tblk) +
             "," + std::to_string(W.dblk) + "," + std::to_string(W.tiles) + "," + std::to_string(W.cap) + "," +
             std::to_string(W.rmax) + "," + std::to_string(W.big_tiles) + "," + std::to_string(W.fixed_total) +
             (W.generic ? "g" : "") +
             (W.layers ? "L" : "");
        // the layer waves' flatten and copy grids are sized by the viewport's pixels
        if (W.layers)
            k += std::to_string((uint64_t)W.jd[0].V.vw * W.jd[0].V.vh);
    }
    return k;
}

// Before a new capture: drop the least recently replayed graphs while the
// cache is at its entry or byte limit.  A dropped graph may still be running
// or be about to read its pinned tables, so the stream is drained first.
static int evict_graphs(swegl_ctx *c)
{
    if (c->graphs.size() < kGraphCacheEntries && c->graph_bytes < kGraphCacheBytes)
        return SWEGL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    while (!c->graphs.empty() &&
           (c->graphs.size() >= kGraphCacheEntries || c->graph_bytes >= kGraphCacheBytes)) {
        auto oldest = c->graphs.begin();
        for (auto it = c->graphs.begin(); it != c->graphs.end(); ++it)
            if (it->second->last_use < oldest->second->last_use)
                oldest = it;
        c->graph_bytes -= std::min(c->graph_bytes, oldest->second->bytes());
        c->graphs.erase(oldest);
    }
    return SWEGL_OK;
}

static int render_jobs_graph(swegl_ctx *c, const std::vector<JobIn> &all, swegl_frame_stats *stats)
{
    c->op_seq++;   // the per-viewport scratch of slot 0, on the context stream
    std::vector<Wave> waves;
    int rc = prepare(c, c->lane0, all, waves);
    if (rc)
        return rc;
    if (waves.empty())
        return no_waves(stats);
    // A wave with depth of field and a shader other than Phong + bilinear
    // runs eagerly: one such graph replay ended in a GPU memory fault whose
    // cause is not known (the same jobs pass eagerly and in other graphs).
    for (const Wave &W : waves)
        if (W.generic && W.dblk)
            return render_jobs(c, c->lane0, all, stats);
    const std::string key = graph_key(c, waves) + (stats ? "/stats" : "");
    auto it = c->graphs.find(key);
    GraphEntry *G;
    if (it == c->graphs.end()) {
        if ((rc = evict_graphs(c)))
            return rc;
        std::unique_ptr<GraphEntry> g(new GraphEntry);
        if ((rc = size_resources(g->res, waves)))
            return rc;
        for (const Wave &W : waves)
            g->job_bytes = std::max(g->job_bytes, W.job_offset + W.jd.size() * sizeof(JobDev));
        for (int i = 0; i < 2; i++) {
            HIPCHK(hipHostMalloc(&g->pinned[i], g->job_bytes, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&g->done[i], hipEventDisableTiming));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int lrc = SWEGL_OK;
        const ResPtr GR = res_ptr(g->res);
        for (size_t w = 0; lrc == SWEGL_OK && w < waves.size(); w++)
            lrc = launch_wave(c, c->lane0, GR, waves[w], w == 0, true, stats != nullptr);
        hipGraph_t graph = nullptr;
        hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        g->graph = graph;   // destroyed with the entry if anything below fails
        if (lrc)
            return lrc;
        if (ee != hipSuccess)
            return fail(SWEGL_EIO, "hipStreamEndCapture: %s", hipGetErrorString(ee));
        HIPCHK(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
        G = g.get();
        c->graph_bytes += G->bytes();
        c->graphs[key] = std::move(g);
    } else {
        G = it->second.get();
        // the same wave layout as the capture (the key says so): the same places
        place_job_tables(waves);
        for (const Wave &W : waves)
            if (W.job_offset + W.jd.size() * sizeof(JobDev) > G->job_bytes)
                return fail(SWEGL_EIO, "graph replay: job tables larger than at capture");
    }
    // write this frame's job tables into the next pinned buffer, once the
    // replay that last read it has finished
    const int slot = G->next;
    G->next ^= 1;
    if (G->used[slot])
        HIPCHK(hipEventSynchronize(G->done[slot]));
    for (const Wave &W : waves)
        memcpy((char *)G->pinned[slot] + W.job_offset, W.jd.data(), W.jd.size() * sizeof(JobDev));
    // the job table goes in by a plain copy ahead of the replay on the same
    // stream: cheaper on the host than rewriting a copy node of the graph
    HIPCHK(hipMemcpyAsync(G->res.jobs.p, G->pinned[slot], G->job_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipGraphLaunch(G->exec, c->stream));
    HIPCHK(hipEventRecord(G->done[slot], c->stream));
    G->used[slot] = true;
    G->last_use = ++c->graph_tick;
    return stats ? read_stats(c, c->lane0, G->res.ctr.as<uint32_t>(), stats) : SWEGL_OK;
}

extern "C" int swegl_render(swegl_ctx *c, swegl_scene *s, const swegl_viewport_desc *vps, uint32_t n,
                            swegl_frame_stats *stats)
{
    if (!c || !s || (n && !vps))
        return fail(SWEGL_EINVAL, "null argument");
    std::vector<JobIn> jobs;
    for (uint32_t i = 0; i < n; i++)
        jobs.push_back({s, &vps[i]});
    return render_jobs(c, c->lane0, jobs, stats);
}

// Allocation-free variant of swegl_render (no counterpart in the reference):
// the same kernels on the device's fixed scratch array.
extern "C" int swegl_render_noalloc(swegl_ctx *c, swegl_scene *s, const swegl_viewport_desc *vps, uint32_t n,
                                    swegl_frame_stats *stats)
{
    if (!c || !s || (n && !vps))
        return fail(SWEGL_EINVAL, "null argument");
    std::vector<JobIn> jobs;
    for (uint32_t i = 0; i < n; i++)
        jobs.push_back({s, &vps[i]});
    return render_jobs_noalloc(c, jobs, stats);
}

// The scene fields the reference's render writes (model.hpp:20-29), for
// parity tests: render one viewport, then read back the resident v_world and
// normal_world and the viewport's frustum coordinates and marks, with the
// viewport transform applied to the marked vertices as frustum_to_viewport
// does (vertex_shaders.hpp:72-84).
extern "C" int swegl_debug_vertex_state(swegl_ctx *c, swegl_scene *s, const swegl_viewport_desc *vp,
                                        float *v_world, float *normal_world, float *v_viewport, uint8_t *yes)
{
    if (!c || !s || !vp || (s->n_vertices && (!v_world || !normal_world || !v_viewport || !yes)))
        return fail(SWEGL_EINVAL, "null argument");
    const int slot = c->front_slot;   // the set of front-end buffers the render below uses
    int rc = swegl_render(c, s, vp, 1, nullptr);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t V = s->n_vertices;
    if (!V)
        return SWEGL_OK;
    if (s->views.empty())
        return fail(SWEGL_EINVAL, "no vertex state for this scene");
    const ViewScratch &vs = *s->views[0];
    std::vector<float4> w(V), n(V), f(V);
    std::vector<uint8_t> y(V);
    HIPCHK(hipMemcpy(w.data(), s->world.p, V * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(n.data(), s->nworld.p, V * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(f.data(), vs.front[slot].ndc.p, V * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(y.data(), vs.front[slot].yes.p, V, hipMemcpyDeviceToHost));
    float m00, m03, m11, m13;
    hm::viewport_matrix(vp->x, vp->y, vp->w, vp->h, &m00, &m03, &m11, &m13);
    for (size_t i = 0; i < V; i++) {
        v_world[3 * i] = w[i].x;
        v_world[3 * i + 1] = w[i].y;
        v_world[3 * i + 2] = w[i].z;
        normal_world[3 * i] = n[i].x;
        normal_world[3 * i + 1] = n[i].y;
        normal_world[3 * i + 2] = n[i].z;
        yes[i] = y[i] ? 1 : 0;
        v_viewport[3 * i] = y[i] ? m00 * f[i].x + m03 : f[i].x;
        v_viewport[3 * i + 1] = y[i] ? m11 * f[i].y + m13 : f[i].y;
        v_viewport[3 * i + 2] = f[i].z;
    }
    return SWEGL_OK;
}

extern "C" int swegl_render_graph(swegl_ctx *c, swegl_scene *s, const swegl_viewport_desc *vps,
                                  uint32_t n, swegl_frame_stats *stats)
{
    if (!c || !s || (n && !vps))
        return fail(SWEGL_EINVAL, "null argument");
    std::vector<JobIn> jobs;
    for (uint32_t i = 0; i < n; i++)
        jobs.push_back({s, &vps[i]});
    return render_jobs_graph(c, jobs, stats);
}

static int collect(const swegl_job *jobs, uint32_t n, std::vector<JobIn> &all)
{
    for (uint32_t i = 0; i < n; i++) {
        if (jobs[i].n_viewports && !jobs[i].viewports)
            return fail(SWEGL_EINVAL, "job %u has no viewports", i);
        for (uint32_t k = 0; k < jobs[i].n_viewports; k++)
            all.push_back({jobs[i].scene, &jobs[i].viewports[k]});
    }
    return SWEGL_OK;
}

extern "C" int swegl_render_batch(swegl_ctx *c, const swegl_job *jobs, uint32_t n, swegl_frame_stats *stats)
{
    if (!c || (n && !jobs))
        return fail(SWEGL_EINVAL, "null argument");
    std::vector<JobIn> all;
    int rc = collect(jobs, n, all);
    return rc ? rc : render_jobs(c, c->lane0, all, stats);
}

extern "C" int swegl_render_batch_graph(swegl_ctx *c, const swegl_job *jobs, uint32_t n,
                                        swegl_frame_stats *stats)
{
    if (!c || (n && !jobs))
        return fail(SWEGL_EINVAL, "null argument");
    std::vector<JobIn> all;
    int rc = collect(jobs, n, all);
    return rc ? rc : render_jobs_graph(c, all, stats);
}

// ----------------------------------------------------------------- server
struct SrvReq {
    uint64_t ticket;
    swegl_scene *scene;
    std::vector<swegl_viewport_desc> vps;
    uint64_t scratch, targets;   // request_bytes
};

struct swegl_server {
    swegl_ctx *c = nullptr;
    swegl_server_config cfg{};
    struct LaneState {
        std::unique_ptr<Lane> lane;
        hipEvent_t done = nullptr;
        bool busy = false;
        std::vector<SrvReq> reqs;
        uint64_t scratch = 0, targets = 0;   // of the batch in flight
        // device bytes the lane holds: its region (kept between batches),
        // or while busy at least its batch's scratch bound
        uint64_t held() const { return std::max<uint64_t>(lane->region.n, busy ? scratch : 0); }
    };
    std::vector<LaneState> lanes;
    std::deque<SrvReq> pending;
    std::deque<uint64_t> finished;
    uint64_t next_ticket = 1, in_flight = 0, peak = 0, submitted = 0, completed = 0, batches = 0;
    uint64_t targets_in_flight = 0;
    uint32_t largest = 0;
    ~swegl_server()
    {
        for (LaneState &l : lanes) {
            if (l.lane && l.lane->stream)
                (void)hipStreamSynchronize(l.lane->stream);
            if (l.done)
                (void)hipEventDestroy(l.done);
        }
    }
};

// Device bytes of a request, split into the scratch its batch carves from
// the lane's region and the target planes it writes.  The scratch is an
// upper bound of the request's share of a batch's region, by the formulas
// build_wave and resource_sizes use: per viewport the six scratch pieces
// (each 256-byte aligned), its tiles' counters and offsets at the smallest
// tile a wave can take (16x16, 8x8 with layers), its triangles' bins at
// the wave's bin limit, its job table entry, plus a fixed share covering a
// batch's minimum bin array, counter block and alignment.
struct ReqBytes {
    uint64_t scratch = 0, targets = 0;
};
static ReqBytes request_bytes(const swegl_ctx *c, const swegl_scene *s, const std::vector<swegl_viewport_desc> &vps)
{
    ReqBytes b;
    b.scratch = 16384 + (uint64_t)SW_CTR_BIG0 * 4 + 8 * 256;
    for (const swegl_viewport_desc &v : vps) {
        size_t sz[6];
        job_scratch_sizes(s, v, sz);
        for (size_t k : sz)
            b.scratch += (k + 255) & ~(uint64_t)255;
        const bool layers = v.transparency_layer_count > 0;
        const int t = layers ? SW_LAYER_TW : 16;
        const uint64_t tiles = (uint64_t)((v.w + t - 1) / t) * ((v.h + t - 1) / t);
        const uint32_t bt = std::max<uint32_t>(1, layers ? std::max<uint32_t>(c->big_tiles, SW_LAYER_BIG_TILES)
                                                          : c->big_tiles);
        b.scratch += tiles * (SW_TILE_PAD * 4 + 8) + 16 + (uint64_t)std::max<uint32_t>(1, s->n_tris) * bt * 4 +
                     sizeof(JobDev) + 8;   // the job's table entry and its two counter words
        // its share of the fixed bin layout (build_wave): tiles x cap entries
        // with cap <= max(512, 20 n_tris / tiles + 1) + 63, largest at the
        // smallest tile
        if (!layers)
            b.scratch += (tiles * 640 + (uint64_t)s->n_tris * 20) * 4;
        const uint64_t P = (uint64_t)v.w * v.h;
        b.targets += (uint64_t)v.color->w * v.color->h * 4 + P * 4 + (v.triangle_id ? P * 4 : 0);
    }
    return b;
}

extern "C" int swegl_server_create(swegl_ctx *c, const swegl_server_config *cfg, swegl_server **out)
{
    if (!c || !cfg || !out)
        return fail(SWEGL_EINVAL, "null argument");
    std::unique_ptr<swegl_server> sv(new swegl_server);
    sv->c = c;
    sv->cfg = *cfg;
    if (!sv->cfg.max_batch)
        sv->cfg.max_batch = 256;
    if (!sv->cfg.lanes)
        sv->cfg.lanes = 2;
    if (sv->cfg.lanes > 8)
        return fail(SWEGL_EINVAL, "at most 8 lanes");
    HIPCHK(hipSetDevice(c->device));
    sv->lanes.resize(sv->cfg.lanes);
    for (auto &l : sv->lanes) {
        l.lane.reset(new Lane);
        int rc = l.lane->init(true);
        if (rc)
            return rc;
        HIPCHK(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    }
    *out = sv.release();
    return SWEGL_OK;
}

extern "C" int swegl_server_destroy(swegl_server *sv)
{
    delete sv;
    return SWEGL_OK;
}

extern "C" int swegl_server_submit(swegl_server *sv, const swegl_job *job, uint64_t *ticket)
{
    if (!sv || !job || !job->scene || (job->n_viewports && !job->viewports))
        return fail(SWEGL_EINVAL, "null argument");
    SrvReq r;
    r.scene = job->scene;
    r.vps.assign(job->viewports, job->viewports + job->n_viewports);
    for (const swegl_viewport_desc &v : r.vps) {
        int rc = check_viewport(v);
        if (rc)
            return rc;
    }
    const ReqBytes rb = request_bytes(sv->c, r.scene, r.vps);
    r.scratch = rb.scratch;
    r.targets = rb.targets;
    if (r.scratch + r.targets > sv->cfg.hbm_budget_bytes)
        return fail(SWEGL_EAGAIN, "request needs %llu bytes, budget is %llu",
                    (unsigned long long)(r.scratch + r.targets), (unsigned long long)sv->cfg.hbm_budget_bytes);
    r.ticket = sv->next_ticket++;
    if (ticket)
        *ticket = r.ticket;
    sv->pending.push_back(std::move(r));
    sv->submitted++;
    return SWEGL_OK;
}

// Device bytes the server holds with lane `l` taking a batch of `scratch`
// and `targets` bytes: every lane's region (a busy lane at least its
// batch's bound), lane l's at least the new batch's, and the target planes
// of every request in flight.
static uint64_t projected_bytes(const swegl_server *sv, size_t l, uint64_t scratch, uint64_t targets)
{
    uint64_t b = sv->targets_in_flight + targets;
    for (size_t m = 0; m < sv->lanes.size(); m++)
        b += m == l ? std::max<uint64_t>(sv->lanes[m].lane->region.n, scratch) : sv->lanes[m].held();
    return b;
}

static int server_pump(swegl_server *sv)
{
    for (auto &l : sv->lanes) {
        if (!l.busy)
            continue;
        hipError_t q = hipEventQuery(l.done);
        if (q == hipErrorNotReady)
            continue;
        if (q != hipSuccess)
            return fail(SWEGL_EIO, "batch failed: %s", hipGetErrorString(q));
        for (const SrvReq &r : l.reqs)
            sv->finished.push_back(r.ticket);
        sv->completed += l.reqs.size();
        sv->in_flight -= l.scratch + l.targets;
        sv->targets_in_flight -= l.targets;
        l.reqs.clear();
        l.scratch = l.targets = 0;
        l.busy = false;
    }
    const uint64_t budget = sv->cfg.hbm_budget_bytes;
    for (size_t li = 0; li < sv->lanes.size(); li++) {
        auto &l = sv->lanes[li];
        if (l.busy || sv->pending.empty())
            continue;
        uint64_t scratch = 0, targets = 0;
        for (int pass = 0; pass < 2 && l.reqs.empty(); pass++) {
            while (!sv->pending.empty() && l.reqs.size() < sv->cfg.max_batch &&
                   projected_bytes(sv, li, scratch + sv->pending.front().scratch,
                                   targets + sv->pending.front().targets) <= budget) {
                scratch += sv->pending.front().scratch;
                targets += sv->pending.front().targets;
                l.reqs.push_back(std::move(sv->pending.front()));
                sv->pending.pop_front();
            }
            if (!l.reqs.empty() || pass)
                break;
            // the head does not fit: give back the regions idle lanes keep
            bool freed = false;
            for (auto &m : sv->lanes)
                if (!m.busy && m.lane->region.n) {
                    HIPCHK(hipStreamSynchronize(m.lane->stream));
                    m.lane->region.release();
                    freed = true;
                }
            if (!freed)
                break;
        }
        if (l.reqs.empty())
            break;   // head of the queue waits for memory
        std::vector<JobIn> all;
        for (const SrvReq &r : l.reqs)
            for (const swegl_viewport_desc &v : r.vps)
                all.push_back({r.scene, &v});
        int rc = render_jobs(sv->c, *l.lane, all, nullptr);
        if (rc)
            return rc;
        HIPCHK(hipEventRecord(l.done, l.lane->stream));
        l.busy = true;
        l.scratch = scratch;
        l.targets = targets;
        sv->in_flight += scratch + targets;
        sv->targets_in_flight += targets;
        sv->peak = std::max(sv->peak, projected_bytes(sv, sv->lanes.size(), 0, 0));
        sv->batches++;
        sv->largest = std::max<uint32_t>(sv->largest, (uint32_t)l.reqs.size());
    }
    return SWEGL_OK;
}

extern "C" int swegl_server_poll(swegl_server *sv, uint64_t *done, uint32_t cap, uint32_t *n_done)
{
    if (!sv)
        return fail(SWEGL_EINVAL, "null server");
    int rc = server_pump(sv);
    uint32_t n = 0;
    while (n < cap && !sv->finished.empty()) {
        if (done)
            done[n] = sv->finished.front();
        sv->finished.pop_front();
        n++;
    }
    if (n_done)
        *n_done = n;
    return rc;
}

extern "C" int swegl_server_drain(swegl_server *sv)
{
    if (!sv)
        return fail(SWEGL_EINVAL, "null server");
    while (sv->completed < sv->submitted) {
        int rc = server_pump(sv);
        if (rc)
            return rc;
    }
    return SWEGL_OK;
}

extern "C" int swegl_server_get_stats(swegl_server *sv, swegl_server_stats *st)
{
    if (!sv || !st)
        return fail(SWEGL_EINVAL, "null argument");
    st->submitted = sv->submitted;
    st->completed = sv->completed;
    st->batches = sv->batches;
    st->bytes_in_flight = sv->in_flight;
    st->peak_bytes_in_flight = sv->peak;
    st->largest_batch = sv->largest;
    return SWEGL_OK;
}

extern "C" int swegl_server_run_trace(swegl_server *sv, const swegl_job *jobs, const double *arrival_s,
                                      uint32_t n, double *latency_s)
{
    if (!sv || (n && (!jobs || !arrival_s || !latency_s)))
        return fail(SWEGL_EINVAL, "null argument");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
    std::unordered_map<uint64_t, uint32_t> index;        // ticket -> request
    std::unordered_map<const void *, uint64_t> writer;   // target -> ticket in flight
    std::unordered_map<uint64_t, std::vector<const void *>> writes;
    uint32_t next = 0, done_n = 0;
    std::vector<uint64_t> fin(256);
    while (done_n < n) {
        const double t = now();
        while (next < n && arrival_s[next] <= t) {
            const swegl_job &j = jobs[next];
            bool busy = false;
            for (uint32_t k = 0; k < j.n_viewports && !busy; k++) {
                const void *tg[3] = {j.viewports[k].color, j.viewports[k].depth, j.viewports[k].triangle_id};
                for (const void *p : tg)
                    if (p && writer.count(p))
                        busy = true;
            }
            if (busy)
                break;   // its surfaces are still being written: wait in order
            uint64_t tk;
            int rc = swegl_server_submit(sv, &j, &tk);
            if (rc)
                return rc;
            index[tk] = next;
            for (uint32_t k = 0; k < j.n_viewports; k++) {
                const void *tg[3] = {j.viewports[k].color, j.viewports[k].depth, j.viewports[k].triangle_id};
                for (const void *p : tg)
                    if (p) {
                        writer[p] = tk;
                        writes[tk].push_back(p);
                    }
            }
            next++;
        }
        uint32_t got = 0;
        int rc = swegl_server_poll(sv, fin.data(), (uint32_t)fin.size(), &got);
        if (rc)
            return rc;
        if (got) {
            const double tc = now();
            for (uint32_t q = 0; q < got; q++) {
                const uint32_t i = index[fin[q]];
                latency_s[i] = tc - arrival_s[i];
                for (const void *p : writes[fin[q]]) {
                    auto w = writer.find(p);
                    if (w != writer.end() && w->second == fin[q])
                        writer.erase(w);
                }
                writes.erase(fin[q]);
                index.erase(fin[q]);
                done_n++;
            }
        }
    }
    return SWEGL_OK;
}

// ----------------------------------------------------- host math exports
extern "C" void swegl_matrix_rotate_x(float m[16], float a) { hm::rotate_x(m, a); }
extern "C" void swegl_matrix_rotate_y(float m[16], float a) { hm::rotate_y(m, a); }
extern "C" void swegl_matrix_rotate_z(float m[16], float a) { hm::rotate_z(m, a); }
extern "C" void swegl_matrix_translate(float m[16], float x, float y, float z) { hm::translate(m, x, y, z); }
extern "C" void swegl_camera_init(float aspect, float view[16], float proj[16], float center[3])
{
    hm::camera_init(aspect, view, proj, center);
}
extern "C" void swegl_camera_translate(float view[16], float center[3], float x, float y, float z)
{
    hm::camera_translate(view, center, x, y, z);
}
extern "C" void swegl_camera_rotate(float view[16], int axis, float a)
{
    if (axis == 0)
        hm::rotate_x(view, -a);
    else if (axis == 1)
        hm::rotate_y(view, -a);
    else
        hm::rotate_z(view, -a);
}
