// C ABI of the arc-length engine: context + hipMalloc arena, table objects,
// per-kernel event timing, and the orchestration of the gfx950 kernels.
//
// Host orchestration replaces ArclengthCalculateCuda
// (gcurval/detail/cuda/arclength_calculate.h:93-507): no per-level host
// round trips, no free+malloc growth of task buffers, no device syncs inside a
// build; every allocation happens when a table is created.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gcurval_c.h"
#include "launch.h"

using namespace gcurval;
using namespace gcurval::rt;

// ===========================================================================
// Context
// ===========================================================================
struct ProfRec {
    std::string name;
    hipEvent_t s, e;
    hipStream_t stream;
};

// Auxiliary streams of a context: the learned build's per-bin subdivision
// launches are spread over them so one bin's tail overlaps the next bin.
constexpr int kAuxStreams = 3;

// Scratch of gcv_sample_noalloc: the block -> curve and block -> interval
// maps of up to kNoAllocBlocks sample blocks (kNoAllocBlocks * kSampleBlock
// samples), 8 MiB held by every device the library runs on until the process
// ends. In host code `scratch` only names the symbol; kernels get the device
// address that hipGetSymbolAddress returns for the current device.
constexpr int64_t kNoAllocBlocks = int64_t(1) << 20;
static __device__ __align__(256) int32_t scratch[2 * kNoAllocBlocks];

// Device address of `scratch` per device, resolved once per device (with
// that device current) by the first context created on it.
static std::mutex g_scratch_mu;
static std::map<int, int32_t*> g_scratch_ptr;

static hipError_t resolve_scratch(int device, int32_t** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return e != hipSuccess ? e : hipErrorInvalidSymbol;
    g_scratch_ptr[device] = (int32_t*)ptr;
    *out = (int32_t*)ptr;
    return hipSuccess;
}

struct gcv_ctx {
    int device = 0;
    int32_t* scratch_ptr = nullptr;  // this device's `scratch` (gcv_sample_noalloc)
    int n_cu = 0;
    char dev_name[256] = {0};
    hipStream_t stream = nullptr;
    hipStream_t aux[kAuxStreams] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kAuxStreams] = {};
    std::string err;
    // arena: hipMalloc chunks, first-fit with coalescing inside a chunk
    // (256-byte aligned); one fixed chunk, or a pool that adds chunks
    bool pooled = false, grow = false;
    std::map<uintptr_t, size_t> chunks;             // chunk base -> size
    size_t cap = 0, used = 0;
    std::map<uintptr_t, size_t> free_blocks;        // address -> size
    std::unordered_map<void*, size_t> live;         // pointer -> size
    // timing
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool profiling = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int64_t>> prof;
    ProfRec pending;
    bool pending_on = false;
    bool capturing = false;  // the stream is being captured into a graph
};

static int set_err(gcv_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

// Calls that synchronise the stream are refused while it is being captured.
static int capture_guard(gcv_ctx* c, const char* what) {
    if (c && c->capturing)
        return set_err(c, GCV_ERR_STATE, "%s synchronises the stream: not allowed during graph capture", what);
    return GCV_OK;
}


// Every entry point runs on its context's device, whatever device the calling
// thread has current (hipSetDevice is per host thread).
static void use_device(const gcv_ctx* c) {
    int d = -1;
    if (c && hipGetDevice(&d) == hipSuccess && d != c->device) hipSetDevice(c->device);
}

#define HIPCHK(ctx, expr)                                                                            \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess)                                                                        \
            return set_err((ctx), GCV_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                           __FILE__, __LINE__);                                                      \
    } while (0)

static hipEvent_t take_event(gcv_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

namespace gcurval {
namespace rt {
void prof_begin(void* owner, const char* name, hipStream_t stream) {
    gcv_ctx* c = (gcv_ctx*)owner;
    if (!c || !c->profiling || c->capturing) return;
    c->pending.name = name;
    c->pending.stream = stream ? stream : c->stream;
    c->pending.s = take_event(c);
    c->pending.e = take_event(c);
    if (c->pending.s) hipEventRecord(c->pending.s, c->pending.stream);
    c->pending_on = true;
}
void prof_end(void* owner, hipStream_t) {
    gcv_ctx* c = (gcv_ctx*)owner;
    if (!c || !c->pending_on) return;
    if (c->pending.e) hipEventRecord(c->pending.e, c->pending.stream);
    c->recs.push_back(c->pending);
    c->pending_on = false;
}
}  // namespace rt
}  // namespace gcurval

static void harvest(gcv_ctx* c) {
    if (c->recs.empty()) return;
    hipStreamSynchronize(c->stream);  // aux-stream work always joins the context stream
    for (ProfRec& r : c->recs) {
        float ms = 0.f;
        if (r.s && r.e && hipEventElapsedTime(&ms, r.s, r.e) == hipSuccess) {
            auto& slot = c->prof[r.name];
            slot.first += ms;
            slot.second += 1;
        }
        if (r.s) c->event_pool.push_back(r.s);
        if (r.e) c->event_pool.push_back(r.e);
    }
    c->recs.clear();
}

static LaunchCfg cfg_of(gcv_ctx* c) { return LaunchCfg{c->stream, c->n_cu, c, 0, 100}; }

// Streams and events of a context (each may be unset), then the object.
static void ctx_release(gcv_ctx* c) {
    for (int i = 0; i < kAuxStreams; i++) {
        if (c->join_ev[i]) hipEventDestroy(c->join_ev[i]);
        if (c->aux[i]) hipStreamDestroy(c->aux[i]);
    }
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    if (c->t0) hipEventDestroy(c->t0);
    if (c->t1) hipEventDestroy(c->t1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// Fork: the auxiliary streams wait for the work queued so far on the context
// stream; join: the context stream waits for everything queued on them. Both
// are stream-ordered (no host synchronisation) and are recorded into a graph
// being captured.
static hipError_t aux_fork(gcv_ctx* c) {
    hipError_t e = hipEventRecord(c->fork_ev, c->stream);
    for (int i = 0; e == hipSuccess && i < kAuxStreams; i++) e = hipStreamWaitEvent(c->aux[i], c->fork_ev, 0);
    return e;
}
static hipError_t aux_join(gcv_ctx* c) {
    hipError_t e = hipSuccess;
    for (int i = 0; e == hipSuccess && i < kAuxStreams; i++) {
        e = hipEventRecord(c->join_ev[i], c->aux[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->join_ev[i], 0);
    }
    return e;
}

extern "C" {

const char* gcv_version(void) { return "gcurval-amd 0.1 (gfx950)"; }

void gcv_opts_default(gcv_opts* o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->precision = GCV_FP32;
    o->gl_order = 3;
    o->n_subdivided = 64;
    o->max_depth = 0;
    o->tolerance = 0.0;
    o->normalize = 0;
    o->newton_iters = 0;
}

// One more arena chunk (hipMalloc, 256-byte multiple) on the free list.
constexpr size_t kPoolChunk = (size_t)256 << 20;  // growth step of a pool context
static int add_chunk(gcv_ctx* c, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return set_err(c, GCV_ERR_OOM, "hipMalloc of a %zu-byte chunk failed", bytes);
    c->chunks[(uintptr_t)p] = bytes;
    c->free_blocks[(uintptr_t)p] = bytes;
    c->cap += bytes;
    return GCV_OK;
}

int gcv_ctx_create(int device, size_t arena_bytes, gcv_ctx** out) {
    if (!out) return GCV_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GCV_ERR_HIP;
    if (device < 0 || device >= n) return GCV_ERR_INVALID;
    gcv_ctx* c = new gcv_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return GCV_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->n_cu = prop.multiProcessorCount;
        snprintf(c->dev_name, sizeof c->dev_name, "%s", prop.gcnArchName);
    }
    if (c->n_cu <= 0) c->n_cu = 256;
    // every stream at the default priority
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&c->t0) == hipSuccess && hipEventCreate(&c->t1) == hipSuccess &&
              hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < kAuxStreams; i++)
        ok = hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming) == hipSuccess;
    if (!ok || resolve_scratch(device, &c->scratch_ptr) != hipSuccess) {
        ctx_release(c);
        return GCV_ERR_HIP;
    }
    if (arena_bytes == GCV_ARENA_POOL) {
        c->pooled = c->grow = true;  // chunks are added on demand
    } else if (arena_bytes > 0) {
        c->pooled = true;
        if (add_chunk(c, arena_bytes) != GCV_OK) {
            ctx_release(c);
            return GCV_ERR_OOM;
        }
    }
    *out = c;
    return GCV_OK;
}

int gcv_ctx_stream(gcv_ctx* c, void** stream) {
    use_device(c);
    if (!c || !stream) return GCV_ERR_INVALID;
    *stream = (void*)c->stream;
    return GCV_OK;
}

int gcv_ctx_destroy(gcv_ctx* c) {
    use_device(c);
    if (!c) return GCV_OK;
    if (c->capturing) {  // an unfinished capture: end it and drop the graph
        hipGraph_t g = nullptr;
        hipStreamEndCapture(c->stream, &g);
        if (g) hipGraphDestroy(g);
        c->capturing = false;
    }
    hipStreamSynchronize(c->stream);
    harvest(c);
    for (hipEvent_t e : c->event_pool) hipEventDestroy(e);
    c->event_pool.clear();
    if (c->pooled)
        for (auto& kv : c->chunks) hipFree((void*)kv.first);
    else
        for (auto& kv : c->live) hipFree(kv.first);
    c->chunks.clear();
    ctx_release(c);
    return GCV_OK;
}

const char* gcv_last_error(const gcv_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gcv_alloc(gcv_ctx* c, size_t bytes, void** dptr) {
    use_device(c);
    if (!c || !dptr) return GCV_ERR_INVALID;
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    bytes = (bytes + 255) & ~(size_t)255;
    if (!c->pooled) {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess)
            return set_err(c, GCV_ERR_OOM, "hipMalloc of %zu bytes failed", bytes);
        c->live[p] = bytes;
        c->used += bytes;
        *dptr = p;
        return GCV_OK;
    }
    for (int attempt = 0; attempt < 2; attempt++) {
        for (auto it = c->free_blocks.begin(); it != c->free_blocks.end(); ++it) {
            if (it->second >= bytes) {
                const uintptr_t at = it->first;
                const size_t sz = it->second;
                c->free_blocks.erase(it);
                if (sz > bytes) c->free_blocks[at + bytes] = sz - bytes;
                void* p = (void*)at;
                c->live[p] = bytes;
                c->used += bytes;
                *dptr = p;
                return GCV_OK;
            }
        }
        if (!c->grow || attempt > 0) break;
        // a growing pool adds a chunk of at least kPoolChunk bytes (a
        // hipMalloc: refused while the stream is captured)
        if (int rc = capture_guard(c, "gcv_alloc growing the pool")) return rc;
        if (int rc = add_chunk(c, std::max(bytes, kPoolChunk))) return rc;
    }
    return set_err(c, GCV_ERR_OOM, "arena exhausted: %zu bytes requested, %zu of %zu in use", bytes, c->used,
                   c->cap);
}

int gcv_free(gcv_ctx* c, void* p) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    if (!p) return GCV_OK;
    auto it = c->live.find(p);
    if (it == c->live.end()) return set_err(c, GCV_ERR_INVALID, "gcv_free of unknown pointer %p", p);
    size_t sz = it->second;
    c->live.erase(it);
    c->used -= sz;
    if (!c->pooled) {
        if (c->capturing) {  // keep the bookkeeping consistent: refuse before forgetting the block
            c->live[p] = sz;
            c->used += sz;
            return capture_guard(c, "gcv_free without an arena");
        }
        hipStreamSynchronize(c->stream);
        HIPCHK(c, hipFree(p));
        return GCV_OK;
    }
    // stream-ordered reuse: later work on the same stream may take the
    // block; neighbours merge only inside the block's chunk (separate
    // hipMalloc chunks may happen to be adjacent)
    const uintptr_t off = (uintptr_t)p;
    auto ch = std::prev(c->chunks.upper_bound(off));
    const uintptr_t lo = ch->first, hi = ch->first + ch->second;
    auto nx = c->free_blocks.lower_bound(off);
    if (nx != c->free_blocks.end() && nx->first == off + sz && nx->first < hi) {
        sz += nx->second;
        nx = c->free_blocks.erase(nx);
    }
    if (nx != c->free_blocks.begin()) {
        auto pv = std::prev(nx);
        if (pv->first >= lo && pv->first + pv->second == off) {
            pv->second += sz;
            return GCV_OK;
        }
    }
    c->free_blocks[off] = sz;
    return GCV_OK;
}

int gcv_arena_stats(const gcv_ctx* c, size_t* used, size_t* capacity) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    if (used) *used = c->used;
    if (capacity) *capacity = c->cap;
    return GCV_OK;
}

int gcv_copy_h2d(gcv_ctx* c, void* dst, const void* src, size_t bytes) {
    use_device(c);
    if (!c || (bytes && (!dst || !src))) return GCV_ERR_INVALID;
    if (int rc = capture_guard(c, "gcv_copy_h2d")) return rc;
    if (!bytes) return GCV_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GCV_OK;
}

int gcv_copy_d2h(gcv_ctx* c, void* dst, const void* src, size_t bytes) {
    use_device(c);
    if (!c || (bytes && (!dst || !src))) return GCV_ERR_INVALID;
    if (int rc = capture_guard(c, "gcv_copy_d2h")) return rc;
    if (!bytes) return GCV_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GCV_OK;
}

int gcv_copy_d2d(gcv_ctx* c, void* dst, const void* src, size_t bytes) {
    use_device(c);
    if (!c || (bytes && (!dst || !src))) return GCV_ERR_INVALID;
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return set_err(c, GCV_ERR_INVALID, "gcv_copy_d2d needs 16-byte aligned pointers and sizes");
    const int64_t n16 = (int64_t)(bytes / 16);
    if (!n16) return GCV_OK;
    const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, (int64_t)c->n_cu * 32);
    prof_begin(c, "k_copy16");
    hipLaunchKernelGGL(dev::k_copy16, dim3((unsigned)blocks), dim3(256), 0, c->stream, (uint4*)dst,
                       (const uint4*)src, n16);
    hipError_t e = hipGetLastError();
    prof_end(c);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "copy launch failed: %s", hipGetErrorString(e));
    return GCV_OK;
}

int gcv_valu_calibration(gcv_ctx* c, int kind, double* dbuf, int64_t n_blocks, int iters) {
    use_device(c);
    if (!c || !dbuf || n_blocks <= 0 || n_blocks > INT32_MAX || iters < 0 || kind < 0 || kind > 2)
        return c ? set_err(c, GCV_ERR_INVALID, "gcv_valu_calibration: kind 0..2, n_blocks > 0, a buffer of "
                           "n_blocks * 256 doubles") : GCV_ERR_INVALID;
    static const char* names[3] = {"k_valu_fma32", "k_valu_fma64", "k_valu_sqrt32"};
    prof_begin(c, names[kind]);
    if (kind == 0)
        hipLaunchKernelGGL(dev::k_valu_chain<0>, dim3((unsigned)n_blocks), dim3(256), 0, c->stream, dbuf, iters);
    else if (kind == 1)
        hipLaunchKernelGGL(dev::k_valu_chain<1>, dim3((unsigned)n_blocks), dim3(256), 0, c->stream, dbuf, iters);
    else
        hipLaunchKernelGGL(dev::k_valu_chain<2>, dim3((unsigned)n_blocks), dim3(256), 0, c->stream, dbuf, iters);
    hipError_t e = hipGetLastError();
    prof_end(c);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "calibration launch failed: %s", hipGetErrorString(e));
    return GCV_OK;
}

int gcv_sync(gcv_ctx* c) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    if (int rc = capture_guard(c, "gcv_sync")) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GCV_OK;
}

int gcv_device_info(gcv_ctx* c, int* n_cu, char* name, int name_len) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    if (n_cu) *n_cu = c->n_cu;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", c->dev_name);
    return GCV_OK;
}

int gcv_timer_start(gcv_ctx* c) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    HIPCHK(c, hipEventRecord(c->t0, c->stream));
    return GCV_OK;
}

int gcv_timer_stop(gcv_ctx* c, double* ms) {
    use_device(c);
    if (!c || !ms) return GCV_ERR_INVALID;
    if (int rc = capture_guard(c, "gcv_timer_stop")) return rc;
    HIPCHK(c, hipEventRecord(c->t1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->t1));
    float f = 0.f;
    HIPCHK(c, hipEventElapsedTime(&f, c->t0, c->t1));
    *ms = f;
    return GCV_OK;
}

int gcv_profile_enable(gcv_ctx* c, int on) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    c->profiling = on != 0;
    return GCV_OK;
}

int gcv_profile_reset(gcv_ctx* c) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    harvest(c);
    c->prof.clear();
    return GCV_OK;
}

int gcv_profile_query(gcv_ctx* c, const char* prefix, double* total_ms, int64_t* count) {
    use_device(c);
    if (!c || !prefix) return GCV_ERR_INVALID;
    harvest(c);
    double t = 0;
    int64_t n = 0;
    const size_t lp = strlen(prefix);
    for (auto& kv : c->prof)
        if (kv.first.compare(0, lp, prefix) == 0) {
            t += kv.second.first;
            n += kv.second.second;
        }
    if (total_ms) *total_ms = t;
    if (count) *count = n;
    return GCV_OK;
}

}  // extern "C"

// ===========================================================================
// Tables
// ===========================================================================
enum { KIND_ELLIPSE = 0, KIND_BSPLINE = 1 };

struct ClassRange {
    int p, D;
    int64_t offset, count;  // into class_list
};

struct gcv_table {
    gcv_ctx* ctx = nullptr;
    int kind = KIND_ELLIPSE;
    gcv_opts o{};
    int n_curve = 0;
    int64_t n_int = 0;
    int log2_nsub = 6;
    double eps0 = 0;  // per-root level-0 tolerance, exact value of the build's Dt
    int stack_cap = 0;
    bool built = false;
    // curve data (caller-owned, device)
    const double* a = nullptr;
    const double* b = nullptr;
    gcv_bspline_batch bat{};
    std::vector<ClassRange> classes;
    int uniform_class = 0;
    // owned device arrays
    int64_t* ioff = nullptr;
    double* int_cum = nullptr;
    double* root_prefix = nullptr;
    uint64_t* deep = nullptr;
    double* curve_len = nullptr;
    double* curve_scale = nullptr;
    double* int_scale = nullptr;
    int32_t* int_curve = nullptr;
    // cumulative scan tiles (B-spline tables): status words, aggregates, prefixes
    unsigned long long* scan_status = nullptr;
    double* scan_agg = nullptr;
    double* scan_prefix = nullptr;
    unsigned scan_epoch = 0;
    int64_t max_ipc = 0;                  // most intervals in one curve
    int scan_mode = 0;                    // 0 by curve length, 1 per-curve waves, 2 look-back
    double* coef64 = nullptr;             // FP64 builds: [n_int * cstride]
    int cstride = 1;                      // coefficient values per interval (max p * D over the batch)
    float* coef32 = nullptr;
    double* u0 = nullptr;
    double* h = nullptr;
    int32_t* class_list = nullptr;
    int32_t* class_of_curve = nullptr;
    // seeded builds: per-interval k, root offsets, (class, k) buckets
    int8_t* seed_k = nullptr;             // [n_int] k used by uniform builds
    int8_t* seed_floor = nullptr;         // [n_int] safety floor of seeded builds
    int tune_wpb = 0;                     // subdivision launch shape picked by gcv_table_autotune
    int tune_grid_pct = 100;
    int tune_streams = 1;                 // streams of the seeded build's per-bin launches (default_streams)
    int8_t* cls_idx = nullptr;            // [n_curve] class index (mixed batches)
    int8_t* root_k = nullptr;             // [n_int] log2 root count of a seeded build
    int64_t* root_off = nullptr;          // [n_int] first root of each interval (seeded builds: dense roots)
    int64_t* bucket_root_chunk = nullptr; // [kBucketMaxWaves] roots of each bucketing chunk, then their offsets
    int32_t* bucket_list = nullptr;       // [n_int] intervals sorted by (class, k)
    int32_t* bucket_hist = nullptr;       // [kMaxBins * kBucketMaxWaves] per-wave bin counts
    int8_t* bucket_bin = nullptr;         // [n_int] bin of each interval
    int32_t* bucket_totals = nullptr;     // [kMaxBins] intervals per bin
    int64_t* bucket_ranges = nullptr;     // [2 kMaxBins] {first slot, count}
    int32_t* rerun_list = nullptr;        // [n_int] intervals of seeded tiles that outgrew the small stack
    int seed_stack_cap = 0;               // task-stack entries per wave of the seeded bins kernel
    bool seeded = false;                  // last build used per-interval root counts
    int8_t* start_ok = nullptr;           // [n_int] learned Newton-start decision (output 3 of the predictor)
    bool has_start = false;               // start_ok holds the predictions of the last (learned) build
    int newton_start = GCV_START_PLAIN;   // sampler Newton start (gcv_table_set_newton_start)
    int8_t* iters_out = nullptr;          // optional per-sample Newton step counts (gcv_table_set_iters_out)
    int built_ls = 6;                     // log2 roots per interval of the last uniform build
    unsigned long long* stats = nullptr;
    std::vector<void*> owned;
    // sampling scratch: curve of the first sample of each sample block
    int32_t* blk_curve = nullptr;
    int32_t* blk_int = nullptr;           // interval of each sample block's first sample (uniform positions)
    int64_t blk_cap = 0;
};

// Default spread of a build's per-class passes (coefficients, then the
// subdivision launches): over two auxiliary streams for batches of several
// evaluator classes, so one class's tail overlaps the next (six classes of
// 11.9M intervals: learned build 2.45 ms with three streams, 2.36 ms with
// two, where two classes' launches fill the GPU and a third only competes),
// on the context stream for one class (where the fork and join cost more
// than the overlap gains).
static int default_streams(const gcv_table* t) { return t->classes.size() > 1 ? 2 : 1; }

template <typename T>
static int talloc(gcv_table* t, T** p, size_t count) {
    void* v = nullptr;
    int rc = gcv_alloc(t->ctx, sizeof(T) * (count ? count : 1), &v);
    if (rc) return rc;
    t->owned.push_back(v);
    *p = (T*)v;
    return GCV_OK;
}

static void table_free(gcv_table* t) {
    if (!t) return;
    if (t->ctx) {
        hipStreamSynchronize(t->ctx->stream);
        for (void* p : t->owned) gcv_free(t->ctx, p);
        if (t->blk_curve) gcv_free(t->ctx, t->blk_curve);
        if (t->blk_int) gcv_free(t->ctx, t->blk_int);
    }
    delete t;
}

static int resolve_opts(gcv_ctx* c, const gcv_opts* in, gcv_opts* o) {
    gcv_opts_default(o);
    if (in) *o = *in;
    if (o->precision != GCV_FP32 && o->precision != GCV_FP64)
        return set_err(c, GCV_ERR_INVALID, "precision must be GCV_FP32 or GCV_FP64");
    if (o->gl_order != 3 && o->gl_order != 5)
        return set_err(c, GCV_ERR_UNSUPPORTED, "gl_order %d not built (3 or 5)", o->gl_order);
    if (o->n_subdivided < 1 || o->n_subdivided > 64 || (o->n_subdivided & (o->n_subdivided - 1)))
        return set_err(c, GCV_ERR_INVALID, "n_subdivided must be a power of two in [1, 64]");
    if (o->max_depth == 0) o->max_depth = o->precision == GCV_FP32 ? 14 : 23;
    const int maxd = o->precision == GCV_FP32 ? 18 : 31;
    if (o->max_depth < 0 || o->max_depth > maxd)
        return set_err(c, GCV_ERR_INVALID, "max_depth must be in [0, %d]", maxd);
    if (o->tolerance == 0.0) o->tolerance = o->precision == GCV_FP32 ? 1e-6 : 1e-14;
    if (!(o->tolerance > 0.0)) return set_err(c, GCV_ERR_INVALID, "tolerance must be positive");
    // FP32 with the FP64 step: one FP32 Newton step brings the lane into the
    // quadratic regime the FP64 step finishes; lanes not there yet keep going
    if (o->newton_iters == 0) o->newton_iters = o->precision == GCV_FP32 ? (o->skip_fp64_step ? 2 : 1) : 3;
    if (o->newton_iters < 0 || o->newton_iters > 32) return set_err(c, GCV_ERR_INVALID, "newton_iters out of range");
    return GCV_OK;
}

static int log2i(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

// Allocation shared by both families.
static int table_alloc_common(gcv_table* t) {
    int rc;
    // at most 64 roots per interval (the plain build; seeded builds store fewer, densely)
    const int64_t nroots = t->n_int << 6;
    if ((rc = talloc(t, &t->ioff, t->n_curve + 1))) return rc;
    if ((rc = talloc(t, &t->int_cum, t->n_int))) return rc;
    if ((rc = talloc(t, &t->root_prefix, nroots))) return rc;
    if ((rc = talloc(t, &t->deep, t->n_int))) return rc;
    if ((rc = talloc(t, &t->curve_len, t->n_curve))) return rc;
    if ((rc = talloc(t, &t->stats, dev::kStatBufWords))) return rc;
    // seeded-build arrays (allocated once, so builds never allocate)
    if ((rc = talloc(t, &t->seed_k, t->n_int)) || (rc = talloc(t, &t->seed_floor, t->n_int)) ||
        (rc = talloc(t, &t->start_ok, t->n_int)) ||
        (rc = talloc(t, &t->root_k, t->n_int)) || (rc = talloc(t, &t->root_off, t->n_int)) ||
        (rc = talloc(t, &t->bucket_root_chunk, dev::kBucketMaxWaves)) || (rc = talloc(t, &t->bucket_list, t->n_int)) ||
        (rc = talloc(t, &t->bucket_hist, dev::kMaxBins * dev::kBucketMaxWaves)) ||
        (rc = talloc(t, &t->bucket_bin, t->n_int)) || (rc = talloc(t, &t->bucket_totals, dev::kMaxBins)) ||
        (rc = talloc(t, &t->bucket_ranges, 2 * dev::kMaxBins)) || (rc = talloc(t, &t->rerun_list, t->n_int)))
        return rc;
    // depth bound of the wave-local stack: 64 (max_depth + 1), plus one block of slack
    t->stack_cap = 64 * (t->o.max_depth + 2);
    // the seeded bins kernel keeps a small stack (tiles that outgrow it are re-run)
    t->seed_stack_cap = 256;
    const int nsub = 1 << t->log2_nsub;
    if (t->o.precision == GCV_FP32) t->eps0 = (double)((float)t->o.tolerance / (float)nsub);
    else t->eps0 = t->o.tolerance / (double)nsub;
    return GCV_OK;
}

// Runs one input-check kernel (k_check_*) and reports its finding: the first
// bad curve and what is wrong with it, or GCV_OK.
template <class Launch>
static int check_input(gcv_ctx* c, const char* family, Launch launch) {
    void* dbad = nullptr;
    int rc = gcv_alloc(c, sizeof(unsigned long long), &dbad);
    if (rc) return rc;
    unsigned long long bad = ~0ull;
    if ((rc = gcv_copy_h2d(c, dbad, &bad, sizeof(bad)))) { gcv_free(c, dbad); return rc; }
    launch((unsigned long long*)dbad);
    if (hipGetLastError() != hipSuccess) { gcv_free(c, dbad); return set_err(c, GCV_ERR_HIP, "input check launch failed"); }
    rc = gcv_copy_d2h(c, &bad, dbad, sizeof(bad));
    gcv_free(c, dbad);
    if (rc) return rc;
    if (bad == ~0ull) return GCV_OK;
    static const char* bspline_why[4] = {"", "a knot is not finite", "its knots decrease",
                                         "a control point is not finite"};
    const int why = (int)(bad & 3);
    return set_err(c, GCV_ERR_DATA, "%s curve %llu: %s", family, bad >> 2,
                   family[0] == 'e' ? "a semi-axis is not finite" : bspline_why[why]);
}

extern "C" int gcv_table_create_ellipse(gcv_ctx* c, int n_curve, const double* d_a, const double* d_b,
                                        const gcv_opts* opts, gcv_table** out) {
    use_device(c);
    if (!c || !out) return GCV_ERR_INVALID;
    *out = nullptr;
    if (n_curve < 0 || (n_curve > 0 && (!d_a || !d_b))) return set_err(c, GCV_ERR_INVALID, "bad ellipse batch");
    gcv_table* t = new gcv_table;
    t->ctx = c;
    t->kind = KIND_ELLIPSE;
    int rc = resolve_opts(c, opts, &t->o);
    if (rc) { delete t; return rc; }
    if (n_curve > 0 && (rc = check_input(c, "ellipse", [&](unsigned long long* bad) {
            hipLaunchKernelGGL(dev::k_check_ellipse, dim3((unsigned)((n_curve + 255) / 256)), dim3(256), 0,
                               c->stream, n_curve, d_a, d_b, bad);
        }))) {
        delete t;
        return rc;
    }
    t->n_curve = n_curve;
    t->n_int = n_curve;
    t->log2_nsub = log2i(t->o.n_subdivided);
    t->a = d_a;
    t->b = d_b;
    if ((rc = table_alloc_common(t))) { table_free(t); return rc; }
    std::vector<int64_t> ioff(n_curve + 1);
    for (int i = 0; i <= n_curve; i++) ioff[i] = i;
    if ((rc = gcv_copy_h2d(c, t->ioff, ioff.data(), sizeof(int64_t) * (n_curve + 1)))) { table_free(t); return rc; }
    *out = t;
    return GCV_OK;
}

extern "C" int gcv_table_create_bspline(gcv_ctx* c, const gcv_bspline_batch* B, const gcv_opts* opts,
                                        gcv_table** out) {
    use_device(c);
    if (!c || !B || !out) return GCV_ERR_INVALID;
    *out = nullptr;
    const int n = B->n_curve;
    if (n < 0) return set_err(c, GCV_ERR_INVALID, "n_curve < 0");
    if (n > 0 && (!B->degree || !B->dim || !B->knot_off || !B->ctrl_off || !B->knots || !B->ctrl))
        return set_err(c, GCV_ERR_INVALID, "null array in B-spline batch");
    if (B->ctrl_stride != 2 && B->ctrl_stride != 3) return set_err(c, GCV_ERR_INVALID, "ctrl_stride must be 2 or 3");
    gcv_table* t = new gcv_table;
    t->ctx = c;
    t->kind = KIND_BSPLINE;
    int rc = resolve_opts(c, opts, &t->o);
    if (rc) { delete t; return rc; }
    t->bat = *B;
    t->n_curve = n;
    t->log2_nsub = log2i(t->o.n_subdivided);
    // per-curve metadata to host for validation and planning
    std::vector<int32_t> deg(n), dim(n);
    std::vector<int64_t> ko(n + 1), co(n + 1);
    if (n > 0) {
        if ((rc = gcv_copy_d2h(c, deg.data(), B->degree, sizeof(int32_t) * n)) ||
            (rc = gcv_copy_d2h(c, dim.data(), B->dim, sizeof(int32_t) * n)) ||
            (rc = gcv_copy_d2h(c, ko.data(), B->knot_off, sizeof(int64_t) * (n + 1))) ||
            (rc = gcv_copy_d2h(c, co.data(), B->ctrl_off, sizeof(int64_t) * (n + 1)))) {
            delete t;
            return rc;
        }
    } else {
        ko[0] = co[0] = 0;
    }
    std::vector<int64_t> ioff(n + 1, 0);
    std::map<int, std::vector<int>> by_class;
    for (int i = 0; i < n; i++) {
        const int p = deg[i], D = dim[i];
        const int64_t nk = ko[i + 1] - ko[i], np = co[i + 1] - co[i];
        if (p < 1 || p > 5) { delete t; return set_err(c, GCV_ERR_UNSUPPORTED, "curve %d: degree %d not in [1, 5]", i, p); }
        if (D < 2 || D > 3 || D > B->ctrl_stride) { delete t; return set_err(c, GCV_ERR_INVALID, "curve %d: bad dim %d", i, D); }
        if (nk < 2 * p + 2) { delete t; return set_err(c, GCV_ERR_INVALID, "curve %d: %lld knots < 2p+2", i, (long long)nk); }
        if (np != nk - p - 1) { delete t; return set_err(c, GCV_ERR_INVALID, "curve %d: %lld control points != n_knots-p-1", i, (long long)np); }
        ioff[i + 1] = ioff[i] + (nk - 2 * p - 1);
        by_class[bspline_class(p, D)].push_back(i);
    }
    t->n_int = ioff[n];
    for (int i = 0; i < n; i++) t->max_ipc = std::max(t->max_ipc, ioff[i + 1] - ioff[i]);
    if (t->n_int > INT32_MAX) { delete t; return set_err(c, GCV_ERR_UNSUPPORTED, "more than 2^31-1 intervals"); }
    if (n > 0) {
        const dev::BSplineDev BD{n, B->ctrl_stride, B->degree, B->dim, B->knot_off, B->ctrl_off, B->knots, B->ctrl};
        rc = check_input(c, "B-spline", [&](unsigned long long* bad) {
            const unsigned blocks = (unsigned)std::min<int64_t>((n + 3) / 4, 65536);
            hipLaunchKernelGGL(dev::k_check_bspline, dim3(blocks), dim3(256), 0, c->stream, BD, bad);
        });
        if (rc) { delete t; return rc; }
    }
    if ((rc = table_alloc_common(t))) { table_free(t); return rc; }
    if ((rc = gcv_copy_h2d(c, t->ioff, ioff.data(), sizeof(int64_t) * (n + 1)))) { table_free(t); return rc; }
    const int64_t n_tiles = (t->n_int + dev::kScanTile - 1) / dev::kScanTile;
    if ((rc = talloc(t, &t->curve_scale, n)) || (rc = talloc(t, &t->int_scale, t->n_int)) ||
        (rc = talloc(t, &t->int_curve, t->n_int)) || (rc = talloc(t, &t->u0, t->n_int)) ||
        (rc = talloc(t, &t->h, t->n_int)) || (rc = talloc(t, &t->scan_status, std::max<int64_t>(n_tiles, 1))) ||
        (rc = talloc(t, &t->scan_agg, std::max<int64_t>(n_tiles, 1))) ||
        (rc = talloc(t, &t->scan_prefix, std::max<int64_t>(n_tiles, 1)))) {
        table_free(t);
        return rc;
    }
    // status words start at epoch 0, which no build uses
    hipLaunchKernelGGL(dev::k_fill_u64, dim3((unsigned)((std::max<int64_t>(n_tiles, 1) + 255) / 256)), dim3(256), 0,
                       c->stream, t->scan_status, (int)std::max<int64_t>(n_tiles, 1), 0ull);
    if (hipGetLastError() != hipSuccess) {
        table_free(t);
        return set_err(c, GCV_ERR_HIP, "scan status fill failed");
    }
    for (auto& kv : by_class) t->cstride = std::max(t->cstride, (kv.first / 10) * (kv.first % 10));
    // FP32 builds keep FP32 coefficients only (the sampler's FP64 step evaluates them in FP64)
    rc = t->o.precision == GCV_FP32 ? talloc(t, &t->coef32, t->n_int * t->cstride)
                                    : talloc(t, &t->coef64, t->n_int * t->cstride);
    if (rc) {
        table_free(t);
        return rc;
    }
    // class plan: a single class needs no interval list
    if (by_class.size() <= 1) {
        if (!by_class.empty()) {
            const int cls = by_class.begin()->first;
            t->classes.push_back(ClassRange{cls / 10, cls % 10, 0, t->n_int});
            t->uniform_class = cls;
        }
    } else {
        std::vector<int32_t> list;
        list.reserve(t->n_int);
        std::vector<int32_t> coc(n);
        for (auto& kv : by_class) {
            ClassRange r{kv.first / 10, kv.first % 10, (int64_t)list.size(), 0};
            for (int cv : kv.second) {
                coc[cv] = kv.first;
                for (int64_t i = ioff[cv]; i < ioff[cv + 1]; i++) list.push_back((int32_t)i);
            }
            r.count = (int64_t)list.size() - r.offset;
            t->classes.push_back(r);
        }
        std::vector<int8_t> cidx(n);
        for (int cv = 0; cv < n; cv++)
            for (size_t r = 0; r < t->classes.size(); r++)
                if (bspline_class(t->classes[r].p, t->classes[r].D) == coc[cv]) cidx[cv] = (int8_t)r;
        if ((rc = talloc(t, &t->class_list, list.size())) ||
            (rc = gcv_copy_h2d(c, t->class_list, list.data(), sizeof(int32_t) * list.size())) ||
            (rc = talloc(t, &t->class_of_curve, n)) ||
            (rc = gcv_copy_h2d(c, t->class_of_curve, coc.data(), sizeof(int32_t) * n)) ||
            (rc = talloc(t, &t->cls_idx, n)) || (rc = gcv_copy_h2d(c, t->cls_idx, cidx.data(), n))) {
            table_free(t);
            return rc;
        }
    }
    t->tune_streams = default_streams(t);
    *out = t;
    return GCV_OK;
}

static LaunchCfg cfg_of_table(const gcv_table* t) {
    LaunchCfg L = cfg_of(t->ctx);
    L.sub_wpb = t->tune_wpb;
    L.sub_grid_pct = t->tune_grid_pct;
    return L;
}

template <typename Dt>
static dev::SubdivArgs<Dt> subdiv_args(gcv_table* t) {
    dev::SubdivArgs<Dt> A{};
    A.n_slots = t->n_int;
    A.int_list = nullptr;
    A.log2_nsub = t->log2_nsub;
    A.eps0 = (Dt)t->eps0;
    A.max_depth = t->o.max_depth;
    A.root_prefix = t->root_prefix;
    A.int_len = t->int_cum;
    A.curve_len = t->kind == KIND_ELLIPSE ? t->curve_len : nullptr;
    A.deep = t->deep;
    A.stack_cap = t->stack_cap;
    A.stats = t->stats;
    return A;
}

static BSplineTableDev bspline_dev(gcv_table* t) {
    return BSplineTableDev{t->coef64, t->coef32, t->cstride, t->int_scale, t->u0, t->h};
}

#define LAUNCH_OR_RETURN_STATUS(name, ...)                                                              \
    do {                                                                                               \
        prof_begin(c, name);                                                                           \
        hipLaunchKernelGGL(__VA_ARGS__);                                                               \
        hipError_t _le = hipGetLastError();                                                            \
        prof_end(c);                                                                                   \
        if (_le != hipSuccess) return set_err(c, GCV_ERR_HIP, "%s launch failed: %s", name,            \
                                              hipGetErrorString(_le));                                 \
    } while (0)

#define LAUNCH_OR_RETURN(name, ...)            \
    do {                                       \
        prof_begin(c, name);                   \
        hipLaunchKernelGGL(__VA_ARGS__);       \
        hipError_t _le = hipGetLastError();    \
        prof_end(c);                           \
        if (_le != hipSuccess) return _le;     \
    } while (0)

// The learned seed predicted inside the coefficient pass (B-spline tables).
struct SeedPredict {
    dev::MlpParams P;
    dev::SeedOut O;
};

// Build phases. Setup: statistics reset, and for B-splines the per-curve
// scale / interval map and the per-interval derivative polynomials (with
// `sp`, also the learned seed's k and safety floor of every interval).
template <typename Dt, int NGL>
static hipError_t build_setup(gcv_table* t, bool reset_stats = true, const SeedPredict* sp = nullptr) {
    gcv_ctx* c = t->ctx;
    if (reset_stats)  // statistics rows and the seeded build's queue words
        LAUNCH_OR_RETURN("k_fill_stats", dev::k_fill_u64, dim3((dev::kStatBufWords + 255) / 256), dim3(256), 0,
                         c->stream, t->stats, dev::kStatBufWords, 0ull);
    if (t->kind == KIND_ELLIPSE) return hipSuccess;
    dev::BSplineDev B{t->n_curve, t->bat.ctrl_stride, t->bat.degree, t->bat.dim, t->bat.knot_off,
                      t->bat.ctrl_off, t->bat.knots, t->bat.ctrl};
    if (t->n_curve > 0) {
        // a wave per curve: the per-curve loops are latency chains, so every
        // curve gets its own wave (up to 64 Ki blocks, then grid-stride)
        const int64_t blocks = std::min<int64_t>((t->n_curve + 3) / 4, 65536);
        LAUNCH_OR_RETURN("k_bspline_curve_setup", dev::k_bspline_curve_setup, dim3((unsigned)blocks), dim3(256), 0,
                         c->stream, B, t->ioff, t->o.normalize, t->curve_scale, t->int_scale, t->int_curve);
    }
    // the classes' coefficient passes are independent and latency bound:
    // with several classes they run on the auxiliary streams, so one pass's
    // tail overlaps the next
    const int n_streams = std::min(std::max(t->tune_streams, 1), kAuxStreams);
    const bool multi = n_streams > 1 && t->classes.size() > 1;
    hipError_t e = multi ? aux_fork(c) : hipSuccess;
    int ci = 0;
    for (const ClassRange& r : t->classes) {
        if (e != hipSuccess) break;
        if (r.count == 0) continue;
        const hipStream_t st = multi ? c->aux[ci++ % n_streams] : c->stream;
        const int32_t* list = t->class_list ? t->class_list + r.offset : nullptr;
        const unsigned blocks = (unsigned)((r.count + 255) / 256);
        // fused pass: 4 waves of 64 intervals per block, grid-stride
        const unsigned sblocks = (unsigned)std::min<int64_t>(blocks, (int64_t)c->n_cu * 8);
        const char* name = sp ? "k_bspline_coef_seed" : "k_bspline_coef";
        prof_begin(c, name, st);
#define GCV_COEF_CASE(PP, DD)                                                                                    \
    case 10 * PP + DD:                                                                                           \
        if (sp)                                                                                                  \
            hipLaunchKernelGGL((dev::k_bspline_coef_seed<PP, DD>), dim3(sblocks), dim3(256),                     \
                               dev::kMlpLdsBytes + 4 * dev::kWaveLds, st, B, list, r.count, t->ioff,             \
                               t->int_curve, t->int_scale, t->coef64, t->coef32, t->cstride, t->u0, t->h, sp->P, \
                               sp->O);                                                                           \
        else                                                                                                     \
            hipLaunchKernelGGL((dev::k_bspline_coef<PP, DD>), dim3(blocks), dim3(256), 0, st, B, list, r.count,  \
                               t->ioff, t->int_curve, t->int_scale, t->coef64, t->coef32, t->cstride, t->u0,     \
                               t->h);                                                                            \
        break;
        switch (bspline_class(r.p, r.D)) {
            GCV_COEF_CASE(1, 2)
            GCV_COEF_CASE(1, 3)
            GCV_COEF_CASE(2, 2)
            GCV_COEF_CASE(2, 3)
            GCV_COEF_CASE(3, 2)
            GCV_COEF_CASE(3, 3)
            GCV_COEF_CASE(4, 2)
            GCV_COEF_CASE(4, 3)
            GCV_COEF_CASE(5, 2)
            GCV_COEF_CASE(5, 3)
            default:
                e = hipErrorInvalidValue;
                break;
        }
#undef GCV_COEF_CASE
        if (e == hipSuccess) e = hipGetLastError();
        prof_end(c, st);
    }
    if (multi) {
        const hipError_t ej = aux_join(c);
        if (e == hipSuccess) e = ej;
    }
    return e;
}

// Cumulative scan of the interval lengths. Batches whose curves all have at
// most kWaveScanMax intervals, and at least 64 on average, scan a curve per
// wave (k_curve_scan: one read and one write per interval, no cross-block
// traffic; 63 us on 12M intervals of 50,000 curves against 124 us for the
// look-back pass); the others take the single-pass look-back scan, whose time
// follows the interval count alone (1.2M intervals as one curve: 31 us,
// where a wave per curve would run a 4,700-step serial chain; as 120,000
// curves of 10: 23 us, against 81 us for a wave per curve).
constexpr int64_t kWaveScanMax = 8192;
static hipError_t build_scan(gcv_table* t) {
    gcv_ctx* c = t->ctx;
    if (t->kind == KIND_ELLIPSE || t->n_curve == 0 || t->n_int == 0) return hipSuccess;
    const bool per_curve =
        t->scan_mode == 1 || (t->scan_mode == 0 && t->max_ipc <= kWaveScanMax && t->n_int >= 64 * (int64_t)t->n_curve);
    if (per_curve) {
        const int64_t blocks = std::min<int64_t>((t->n_curve + 3) / 4, 65536);  // a wave per curve
        LAUNCH_OR_RETURN("k_curve_scan", dev::k_curve_scan, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                         t->n_curve, t->ioff, t->int_cum, t->curve_len);
        return hipSuccess;
    }
    const int64_t n_tiles = (t->n_int + dev::kScanTile - 1) / dev::kScanTile;
    if (++t->scan_epoch == 0) t->scan_epoch = 1;  // epoch 0 marks never-written status words
    LAUNCH_OR_RETURN("k_fill_stats", dev::k_fill_u64, dim3(1), dim3(64), 0, c->stream, t->stats + dev::kScanTileWord,
                     2, 0ull);
    const dev::ScanTiles T{t->scan_status, t->scan_agg, t->scan_prefix};
    LAUNCH_OR_RETURN("k_seg_scan", dev::k_seg_scan, dim3((unsigned)n_tiles), dim3(dev::kScanThreads), 0, c->stream,
                     t->n_int,
                     (const int32_t*)t->int_curve, t->int_cum, t->curve_len, T, t->scan_epoch, t->stats);
    return hipSuccess;
}

// A scan whose look-back gave up (it never does when every block runs to
// completion) leaves the flag word set: the table's lengths are then wrong.
static int scan_status(gcv_table* t) {
    if (t->kind == KIND_ELLIPSE) return GCV_OK;
    unsigned long long fail = 0;
    if (int rc = gcv_copy_d2h(t->ctx, &fail, t->stats + dev::kScanFailWord, sizeof(fail))) return rc;
    return fail ? set_err(t->ctx, GCV_ERR_HIP, "cumulative scan stalled waiting for a tile") : GCV_OK;
}

// Level-0 tolerance of the plain error test, tol / n_subdivided (the device
// epsilon table of gcurval/base/constants.inl.h:8-18). Uniform and seeded
// builds apply it to every root whatever the root count, so their leaves meet
// at least the plain method's accuracy.
template <typename Dt>
static Dt eps_of(const gcv_table* t) {
    return (Dt)((Dt)t->o.tolerance / (Dt)(1 << t->log2_nsub));
}

// Task-stack entries per wave of the uniform and whole-interval subdivision
// launches (tiles that outgrow it are re-run with the full stack).
constexpr int kSmallStack = 256;

// Uniform build: every interval starts from 2^ls roots (the plain method when
// ls = log2 n_subdivided); optional per-interval quadrature counts.
template <typename Dt, int NGL>
static hipError_t build_uniform(gcv_table* t, int ls, int32_t* cost) {
    hipError_t e = build_setup<Dt, NGL>(t);
    if (e != hipSuccess) return e;
    const LaunchCfg L = cfg_of_table(t);
    dev::SubdivArgs<Dt> A = subdiv_args<Dt>(t);
    A.log2_nsub = ls;
    A.eps0 = eps_of<Dt>(t);
    A.max_depth = dev::effective_depth(t->o.max_depth, t->log2_nsub, ls, sizeof(Dt) == 4);
    // a small task stack (more resident waves: the full stack of 64 (depth +
    // 2) entries holds 4 waves per SIMD in LDS); the rare tile that outgrows
    // it is re-run with the full stack (c1 build 0.39 -> 0.32 ms, c3
    // subdivision 1.22 -> 1.07 ms)
    A.stack_cap = std::min(kSmallStack, 64 * (A.max_depth + 2));
    A.int_cost = cost;
    A.rerun = t->rerun_list;
    A.rerun_count = t->stats + dev::kStatWords;
    A.rerun_self = 1;
    A.ls_plain = t->log2_nsub;
    A.base_depth = t->o.max_depth;
    if (t->kind == KIND_ELLIPSE) {
        e = Launch<Dt, NGL>::subdivide_ellipse(L, t->a, t->b, t->o.normalize, A);
    } else {
        const BSplineTableDev T = bspline_dev(t);
        int ci = 0;
        for (const ClassRange& r : t->classes) {
            dev::SubdivArgs<Dt> Ar = A;
            if (t->class_list) {
                Ar.int_list = t->class_list + r.offset;
                Ar.n_slots = r.count;
            }
            Ar.rerun = t->rerun_list + r.offset;
            Ar.rerun_count = t->stats + dev::kStatWords + (int64_t)(ci++ % dev::kQueueClasses) * dev::kQueueStride;
            e = Launch<Dt, NGL>::subdivide_bspline(L, r.p, r.D, T, Ar);
            if (e != hipSuccess) return e;
        }
    }
    if (e != hipSuccess) return e;
    t->seeded = false;
    t->built_ls = ls;
    return build_scan(t);
}

// Seeded build: interval i starts from 2^k_i roots, k_i = max(k[i], safety
// floor). The bucketing sorts the intervals by (evaluator class, k) and
// writes each interval's first root (dense roots in interval order); per
// class, the bins kernel runs the 1..32-root tiles, the whole-interval kernel
// the 64-root intervals (bin ranges are read on the device), and the re-run
// kernel the tiles that outgrew the bins kernel's small stack. No host sync,
// nothing allocated, graph-capturable; classes are independent, so with
// tune_streams > 1 they run on the auxiliary streams.
template <typename Dt, int NGL>
static hipError_t build_seeded(gcv_table* t, const int8_t* kk, int32_t* cost, bool setup_done = false,
                               bool floor_done = false) {
    gcv_ctx* c = t->ctx;
    hipError_t e = setup_done ? hipSuccess : build_setup<Dt, NGL>(t);
    if (e != hipSuccess) return e;
    const int64_t n = t->n_int;
    if (n == 0) {  // an empty batch: no launches (an empty B-spline batch has no classes)
        t->seeded = true;
        return build_scan(t);
    }
    const int n_cls = t->kind == KIND_ELLIPSE ? 1 : (int)t->classes.size();
    if (n_cls > dev::kQueueClasses) return hipErrorInvalidValue;
    {
        const unsigned g1 = (unsigned)((n + 255) / 256);
        if (floor_done) {
            // the learned predictor wrote the floor with its features
        } else if (t->kind == KIND_ELLIPSE) {
            LAUNCH_OR_RETURN("k_seed_floor", dev::k_seed_floor_ellipse, dim3(g1), dim3(256), 0, c->stream, n, t->a,
                             t->b, t->log2_nsub, t->seed_floor);
        } else {
            LAUNCH_OR_RETURN("k_seed_floor", dev::k_seed_floor_bspline, dim3(g1), dim3(256), 0, c->stream, n,
                             t->int_curve, t->bat.degree, t->bat.dim, t->coef32, t->coef64, t->cstride, t->log2_nsub,
                             t->seed_floor);
        }
    }
    {
        const int n_bins = n_cls * dev::kSeedBins;
        dev::BinArgs BA{kk, t->seed_floor, t->cls_idx, t->kind == KIND_ELLIPSE ? nullptr : t->int_curve, t->root_k};
        // wave chunks of kBucketChunk intervals, longer when that would exceed kBucketMaxWaves
        const int64_t wchunk = std::max<int64_t>(
            dev::kBucketChunk, ((n + dev::kBucketMaxWaves - 1) / dev::kBucketMaxWaves + 63) / 64 * 64);
        const int nw = (int)((n + wchunk - 1) / wchunk);
        const unsigned gb = (unsigned)((nw + 3) / 4);
        LAUNCH_OR_RETURN("k_bucket_hist", dev::k_bucket_hist, dim3(gb), dim3(256), 0, c->stream, n, BA, n_bins,
                         wchunk, nw, t->bucket_bin, t->bucket_hist, t->bucket_root_chunk);
        LAUNCH_OR_RETURN("k_bucket_totals", dev::k_bucket_totals, dim3((unsigned)n_bins), dim3(1024), 0, c->stream,
                         nw, (const int32_t*)t->bucket_hist, t->bucket_totals);
        LAUNCH_OR_RETURN("k_bucket_rows", dev::k_bucket_rows, dim3((unsigned)n_bins + 1), dim3(1024), 0, c->stream,
                         nw, n_bins, t->bucket_hist, (const int32_t*)t->bucket_totals, t->bucket_ranges,
                         t->bucket_root_chunk);
        LAUNCH_OR_RETURN("k_bucket_place", dev::k_bucket_place, dim3(gb), dim3(256), 0, c->stream, n, n_bins,
                         wchunk, nw, (const int8_t*)t->bucket_bin, (const int32_t*)t->bucket_hist,
                         (const int64_t*)t->bucket_root_chunk, t->bucket_list, t->root_off);
    }
    const BSplineTableDev T = bspline_dev(t);
    const bool f32 = sizeof(Dt) == 4;
    const int n_streams = std::min(std::max(t->tune_streams, 1), kAuxStreams);
    const bool multi = n_streams > 1 && n_cls > 1;
    if (multi && (e = aux_fork(c)) != hipSuccess) return e;
    for (int ci = 0; ci < n_cls && e == hipSuccess; ci++) {
        LaunchCfg L = cfg_of_table(t);
        if (multi) L.stream = c->aux[ci % n_streams];
        const int64_t off = t->kind == KIND_ELLIPSE ? 0 : t->classes[ci].offset;
        const int64_t cnt = t->kind == KIND_ELLIPSE ? n : t->classes[ci].count;
        unsigned long long* q = t->stats + dev::kStatWords + (int64_t)ci * dev::kQueueStride;
        const int64_t* ranges = t->bucket_ranges + 2 * (ci * dev::kSeedBins);
        dev::SeededLaunch<Dt> S;
        S.bins = subdiv_args<Dt>(t);
        S.bins.n_slots = cnt;
        S.bins.int_list = t->bucket_list;
        S.bins.eps0 = eps_of<Dt>(t);
        S.bins.stack_cap = t->seed_stack_cap;
        S.bins.int_cost = cost;
        S.bins.root_off = t->root_off;
        S.run = dev::BinRun{ranges, t->log2_nsub, t->o.max_depth, t->rerun_list + off, q};
        S.whole = subdiv_args<Dt>(t);
        S.whole.n_slots = cnt;
        S.whole.int_list = t->bucket_list;
        S.whole.slot_range = ranges + 2 * dev::kMaxSeedK;
        S.whole.log2_nsub = 6;
        S.whole.eps0 = eps_of<Dt>(t);
        S.whole.max_depth = dev::effective_depth(t->o.max_depth, t->log2_nsub, 6, f32);
        // small stack; outgrown tiles join the bins' re-run list
        S.whole.stack_cap = std::min(kSmallStack, 64 * (S.whole.max_depth + 2));
        S.whole.rerun = t->rerun_list + off;
        S.whole.rerun_count = q;
        S.whole.rerun_self = 0;
        S.whole.int_cost = cost;
        S.whole.root_off = t->root_off;
        S.rerun = subdiv_args<Dt>(t);
        S.rerun.eps0 = eps_of<Dt>(t);
        S.rerun.stack_cap = 64 * (dev::effective_depth(t->o.max_depth, t->log2_nsub, 0, f32) + 2);
        S.rerun.int_cost = cost;
        S.rerun.root_off = t->root_off;
        S.root_k = t->root_k;
        if (t->kind == KIND_ELLIPSE) {
            e = Launch<Dt, NGL>::seeded_ellipse(L, t->a, t->b, t->o.normalize, S);
        } else {
            const ClassRange& r = t->classes[ci];
            e = Launch<Dt, NGL>::seeded_bspline(L, r.p, r.D, T, S);
        }
    }
    if (multi) {
        const hipError_t ej = aux_join(c);
        if (e == hipSuccess) e = ej;
    }
    if (e != hipSuccess) return e;
    t->seeded = true;
    return build_scan(t);
}

template <template <typename, int> class F, typename... Args>
static hipError_t dispatch_build(gcv_table* t, Args... args) {
    if (t->o.precision == GCV_FP32)
        return t->o.gl_order == 3 ? F<float, 3>::run(t, args...) : F<float, 5>::run(t, args...);
    return t->o.gl_order == 3 ? F<double, 3>::run(t, args...) : F<double, 5>::run(t, args...);
}
template <typename Dt, int NGL>
struct UniformBuild {
    static hipError_t run(gcv_table* t, int ls, int32_t* cost) { return build_uniform<Dt, NGL>(t, ls, cost); }
};
template <typename Dt, int NGL>
struct SeededBuild {
    static hipError_t run(gcv_table* t, const int8_t* kk, int32_t* cost) { return build_seeded<Dt, NGL>(t, kk, cost); }
};
template <typename Dt, int NGL>
struct SetupOnly {
    static hipError_t run(gcv_table* t, bool reset_stats) { return build_setup<Dt, NGL>(t, reset_stats); }
};
template <typename Dt, int NGL>
struct SetupWithSeed {
    static hipError_t run(gcv_table* t, bool reset_stats, const SeedPredict* sp) {
        return build_setup<Dt, NGL>(t, reset_stats, sp);
    }
};
template <typename Dt, int NGL>
struct SeededAfterSetup {
    static hipError_t run(gcv_table* t, const int8_t* kk, int32_t* cost) {
        return build_seeded<Dt, NGL>(t, kk, cost, true, true);
    }
};

extern "C" int gcv_table_build(gcv_table* t) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    hipError_t e = dispatch_build<UniformBuild>(t, t->log2_nsub, (int32_t*)nullptr);
    if (e != hipSuccess) return set_err(t->ctx, GCV_ERR_HIP, "table build launch failed: %s", hipGetErrorString(e));
    t->built = true;
    return GCV_OK;
}

extern "C" int gcv_table_build_uniform(gcv_table* t, int log2_roots, int32_t* d_cost) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (log2_roots < 0 || log2_roots > dev::kMaxSeedK)
        return set_err(t->ctx, GCV_ERR_INVALID, "log2_roots must be in [0, %d]", dev::kMaxSeedK);
    hipError_t e = dispatch_build<UniformBuild>(t, log2_roots, d_cost);
    if (e != hipSuccess) return set_err(t->ctx, GCV_ERR_HIP, "table build launch failed: %s", hipGetErrorString(e));
    t->built = true;
    return GCV_OK;
}

extern "C" int gcv_table_build_seeded(gcv_table* t, const int8_t* d_k, int32_t* d_cost) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (!d_k && t->n_int > 0) return set_err(t->ctx, GCV_ERR_INVALID, "null per-interval k array");
    hipError_t e = dispatch_build<SeededBuild>(t, d_k, d_cost);
    if (e != hipSuccess) return set_err(t->ctx, GCV_ERR_HIP, "seeded build launch failed: %s", hipGetErrorString(e));
    t->built = true;
    return GCV_OK;
}

extern "C" int gcv_table_destroy(gcv_table* t) {
    if (t) use_device(t->ctx);
    if (t && t->ctx && t->ctx->capturing)
        return set_err(t->ctx, GCV_ERR_STATE, "gcv_table_destroy synchronises the stream: not allowed during graph capture");
    table_free(t);
    return GCV_OK;
}

extern "C" int gcv_table_info(const gcv_table* t, int* n_curve, int64_t* n_intervals, int* n_subdivided) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (n_curve) *n_curve = t->n_curve;
    if (n_intervals) *n_intervals = t->n_int;
    if (n_subdivided) *n_subdivided = 1 << t->log2_nsub;
    return GCV_OK;
}

extern "C" int gcv_table_seed_floor(gcv_table* t, const int8_t** d_floor) {
    if (t) use_device(t->ctx);
    if (!t || !d_floor) return GCV_ERR_INVALID;
    *d_floor = t->seed_floor;
    return GCV_OK;
}

extern "C" int gcv_table_root_layout(gcv_table* t, int* seeded, int* log2_roots, const int8_t** d_root_k,
                                     const int64_t** d_root_off) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (seeded) *seeded = t->seeded ? 1 : 0;
    if (log2_roots) *log2_roots = t->built_ls;
    if (d_root_k) *d_root_k = t->root_k;
    if (d_root_off) *d_root_off = t->root_off;
    return GCV_OK;
}

extern "C" int gcv_table_lengths(gcv_table* t, double* host_out) {
    if (t) use_device(t->ctx);
    if (!t || (!host_out && t->n_curve > 0)) return GCV_ERR_INVALID;
    if (!t->built) return set_err(t->ctx, GCV_ERR_STATE, "table not built");
    if (int rc = gcv_copy_d2h(t->ctx, host_out, t->curve_len, sizeof(double) * t->n_curve)) return rc;
    return scan_status(t);
}

extern "C" int gcv_table_device_views(gcv_table* t, const double** int_cum, const double** root_prefix,
                                      const uint64_t** deep_mask, const double** curve_len,
                                      const int64_t** int_off) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (int_cum) *int_cum = t->int_cum;
    if (root_prefix) *root_prefix = t->root_prefix;
    if (deep_mask) *deep_mask = t->deep;
    if (curve_len) *curve_len = t->curve_len;
    if (int_off) *int_off = t->ioff;
    return GCV_OK;
}

extern "C" int gcv_points(gcv_table* t, const int64_t* d_sample_off, int64_t n_total, const double* d_u,
                          const int32_t* d_seg, int out_dim, double* d_out) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    if (n_total < 0) return set_err(c, GCV_ERR_INVALID, "n_total < 0");
    if (n_total == 0) return GCV_OK;
    if (!d_sample_off || !d_u || !d_out) return set_err(c, GCV_ERR_INVALID, "null sample offsets, parameters or output");
    if (t->n_curve == 0) return set_err(c, GCV_ERR_INVALID, "points requested from an empty batch");
    const int need = t->kind == KIND_ELLIPSE ? 2 : t->bat.ctrl_stride;
    if (out_dim < need || out_dim > 3) return set_err(c, GCV_ERR_INVALID, "out_dim must be in [%d, 3]", need);
    const unsigned blocks = (unsigned)((n_total + 255) / 256);
    if (t->kind == KIND_ELLIPSE) {
        LAUNCH_OR_RETURN_STATUS("k_points", dev::k_points_ellipse, dim3(blocks), dim3(256), 0, c->stream, t->n_curve,
                                n_total, d_sample_off, t->a, t->b, d_u, out_dim, d_out);
    } else {
        if (!d_seg) return set_err(c, GCV_ERR_INVALID, "B-spline points need the sampler's interval indices");
        LAUNCH_OR_RETURN_STATUS("k_points", dev::k_points_bspline, dim3(blocks), dim3(256), 0, c->stream, t->n_curve,
                                n_total, d_sample_off, t->bat.degree, t->bat.dim, t->bat.knot_off, t->bat.ctrl_off,
                                t->bat.knots, t->bat.ctrl, t->bat.ctrl_stride, d_u, d_seg, out_dim, d_out);
    }
    return GCV_OK;
}

// ===========================================================================
// Graphs: capture the context stream (builds, sampling, predictor) once and
// replay it with one launch.
// ===========================================================================
struct gcv_graph {
    gcv_ctx* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int gcv_graph_begin(gcv_ctx* c) {
    use_device(c);
    if (!c) return GCV_ERR_INVALID;
    if (c->capturing) return set_err(c, GCV_ERR_STATE, "graph capture already in progress");
    if (c->pending_on) return set_err(c, GCV_ERR_STATE, "a profiled launch is open");
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    return GCV_OK;
}

extern "C" int gcv_graph_end(gcv_ctx* c, gcv_graph** out) {
    use_device(c);
    if (!c || !out) return GCV_ERR_INVALID;
    if (!c->capturing) return set_err(c, GCV_ERR_STATE, "no graph capture in progress");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->capturing = false;
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return set_err(c, GCV_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    *out = new gcv_graph{c, g, x};
    return GCV_OK;
}

extern "C" int gcv_graph_launch(gcv_graph* g) {
    if (g) use_device(g->ctx);
    if (!g) return GCV_ERR_INVALID;
    if (g->ctx->capturing) return set_err(g->ctx, GCV_ERR_STATE, "graph launch during capture");
    HIPCHK(g->ctx, hipGraphLaunch(g->exec, g->ctx->stream));
    return GCV_OK;
}

extern "C" int gcv_graph_destroy(gcv_graph* g) {
    if (g) use_device(g->ctx);
    if (!g) return GCV_OK;
    if (int rc = capture_guard(g->ctx, "gcv_graph_destroy")) return rc;
    hipStreamSynchronize(g->ctx->stream);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return GCV_OK;
}

// ===========================================================================
// Autotuner: times the table build over the persistent subdivision grid's
// shape (waves per block x resident-grid fraction) and keeps the fastest.
// ===========================================================================
extern "C" int gcv_table_autotune(gcv_table* t, gcv_mlp* m, int reps, gcv_tune* out) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    if (int rc = capture_guard(c, "gcv_table_autotune")) return rc;
    if (reps < 1) reps = 3;
    static const int kWpb[3] = {4, 2, 1};
    static const int kGrid[3] = {100, 50, 25};
    const int keep_wpb = t->tune_wpb, keep_grid = t->tune_grid_pct, keep_streams = t->tune_streams;
    double best = 1e300, dflt = 0.0;
    int best_wpb = 4, best_grid = 100, best_streams = 1, tried = 0;
    auto build = [&]() { return m ? gcv_table_build_mlp(t, m, 0, nullptr) : gcv_table_build(t); };
    auto time_shape = [&](int wpb, int grid, int streams, double* ms) -> int {
        t->tune_wpb = wpb;
        t->tune_grid_pct = grid;
        t->tune_streams = streams;
        int rc = build();  // warm-up
        if (!rc) rc = gcv_timer_start(c);
        for (int r = 0; r < reps && !rc; r++) rc = build();
        if (!rc) rc = gcv_timer_stop(c, ms);
        if (rc) {
            t->tune_wpb = keep_wpb;
            t->tune_grid_pct = keep_grid;
            t->tune_streams = keep_streams;
            return rc;
        }
        *ms /= reps;
        tried++;
        return GCV_OK;
    };
    for (int wi = 0; wi < 3; wi++) {
        for (int gi = 0; gi < 3; gi++) {
            double ms = 0.0;
            if (int rc = time_shape(kWpb[wi], kGrid[gi], 1, &ms)) return rc;
            if (wi == 0 && gi == 0) dflt = ms;
            if (ms < best) {
                best = ms;
                best_wpb = kWpb[wi];
                best_grid = kGrid[gi];
            }
        }
    }
    if (m) {  // learned builds: the best shape with the bins spread over the auxiliary streams too
        double ms = 0.0;
        if (int rc = time_shape(best_wpb, best_grid, kAuxStreams, &ms)) return rc;
        if (ms < best) {
            best = ms;
            best_streams = kAuxStreams;
        }
    }
    t->tune_wpb = best_wpb;
    t->tune_grid_pct = best_grid;
    t->tune_streams = best_streams;
    if (out) {
        out->sub_waves_per_block = best_wpb;
        out->sub_grid_pct = best_grid;
        out->best_build_ms = best;
        out->default_build_ms = dflt;
        out->n_tried = tried;
        out->bin_streams = best_streams;
    }
    return build();
}
extern "C" int gcv_table_set_launch(gcv_table* t, int sub_waves_per_block, int sub_grid_pct, int bin_streams) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (sub_waves_per_block != 0 && sub_waves_per_block != 1 && sub_waves_per_block != 2 && sub_waves_per_block != 4)
        return set_err(t->ctx, GCV_ERR_INVALID, "waves per block must be 0 (default), 1, 2 or 4");
    if (sub_grid_pct < 1 || sub_grid_pct > 100) return set_err(t->ctx, GCV_ERR_INVALID, "grid percent in [1, 100]");
    if (bin_streams < 0 || bin_streams > kAuxStreams)
        return set_err(t->ctx, GCV_ERR_INVALID, "bin_streams must be in [0, %d]", kAuxStreams);
    t->tune_wpb = sub_waves_per_block;
    t->tune_grid_pct = sub_grid_pct;
    t->tune_streams = bin_streams > 0 ? bin_streams : default_streams(t);
    return GCV_OK;
}

extern "C" int gcv_table_set_newton_start(gcv_table* t, int mode) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (mode < GCV_START_PLAIN || mode > GCV_START_QUINTIC)
        return set_err(t->ctx, GCV_ERR_INVALID, "Newton start must be GCV_START_PLAIN, _LEARNED or _QUINTIC");
    t->newton_start = mode;
    return GCV_OK;
}

extern "C" int gcv_table_set_iters_out(gcv_table* t, int8_t* d_iters) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    t->iters_out = d_iters;
    return GCV_OK;
}

extern "C" int gcv_table_start_labels(gcv_table* t, float* d_out) {
    if (t) use_device(t->ctx);
    if (!t || (!d_out && t->n_int > 0)) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    if (!t->built) return set_err(c, GCV_ERR_STATE, "table not built");
    if (t->kind != KIND_BSPLINE || t->o.precision != GCV_FP32 || t->o.gl_order != 3)
        return set_err(c, GCV_ERR_INVALID, "start labels serve FP32 GL-3 B-spline tables");
    dev::SampleArgs A{};
    A.n_int = t->n_int;
    A.root_prefix = t->root_prefix;
    A.root_k = t->seeded ? t->root_k : nullptr;
    A.root_off = t->seeded ? t->root_off : nullptr;
    A.deep = t->deep;
    A.log2_nsub = t->built_ls;
    A.class_of_curve = t->class_of_curve;
    hipError_t e = Launch<float, 3>::start_labels(cfg_of(c), t->uniform_class, bspline_dev(t), t->int_curve, A, d_out);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "start-label launch failed: %s", hipGetErrorString(e));
    return GCV_OK;
}

extern "C" int gcv_table_set_scan(gcv_table* t, int mode) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (mode < 0 || mode > 2) return set_err(t->ctx, GCV_ERR_INVALID, "scan mode must be 0, 1 or 2");
    t->scan_mode = mode;
    return GCV_OK;
}

extern "C" int gcv_table_stats(gcv_table* t, uint64_t* quadratures, uint64_t* deep_roots, uint64_t* max_stack,
                               uint64_t* overflow, uint64_t* tasks, uint64_t* depth_capped, uint64_t* reruns,
                               uint64_t* tiles, uint64_t* passes) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    if (int rc = capture_guard(t->ctx, "gcv_table_stats")) return rc;
    if (!t->built) return set_err(t->ctx, GCV_ERR_STATE, "table not built");
    std::vector<unsigned long long> rows(dev::kStatWords);
    int rc = gcv_copy_d2h(t->ctx, rows.data(), t->stats, sizeof(unsigned long long) * dev::kStatWords);
    if (rc) return rc;
    unsigned long long s[dev::kStatCount] = {};
    for (int r = 0; r < dev::kStatSlots; r++) {
        const unsigned long long* row = rows.data() + r * dev::kStatStride;
        s[dev::kStatQuad] += row[dev::kStatQuad];
        s[dev::kStatDeepRoots] += row[dev::kStatDeepRoots];
        s[dev::kStatTasks] += row[dev::kStatTasks];
        s[dev::kStatDepthCap] += row[dev::kStatDepthCap];
        s[dev::kStatRerun] += row[dev::kStatRerun];
        s[dev::kStatTiles] += row[dev::kStatTiles];
        s[dev::kStatPasses] += row[dev::kStatPasses];
        s[dev::kStatMaxStack] = std::max(s[dev::kStatMaxStack], row[dev::kStatMaxStack]);
        s[dev::kStatOverflow] = std::max(s[dev::kStatOverflow], row[dev::kStatOverflow]);
    }
    if (quadratures) *quadratures = s[dev::kStatQuad];
    if (deep_roots) *deep_roots = s[dev::kStatDeepRoots];
    if (max_stack) *max_stack = s[dev::kStatMaxStack];
    if (overflow) *overflow = s[dev::kStatOverflow];
    if (tasks) *tasks = s[dev::kStatTasks];
    if (depth_capped) *depth_capped = s[dev::kStatDepthCap];
    if (reruns) *reruns = s[dev::kStatRerun];
    if (tiles) *tiles = s[dev::kStatTiles];
    if (passes) *passes = s[dev::kStatPasses];
    return GCV_OK;
}

// Arguments gcv_sample and gcv_sample_noalloc check alike (0: go on).
static int sample_check(gcv_table* t, const int64_t* d_soff, int64_t n_total, double* d_out_u) {
    gcv_ctx* c = t->ctx;
    if (!t->built) return set_err(c, GCV_ERR_STATE, "arclengthReverse before createArclengthTable");
    if (n_total < 0) return set_err(c, GCV_ERR_INVALID, "n_total < 0");
    if (n_total == 0) return GCV_OK;
    if (!d_soff || !d_out_u) return set_err(c, GCV_ERR_INVALID, "null sample offsets or output");
    if (t->n_curve == 0) return set_err(c, GCV_ERR_INVALID, "samples requested from an empty batch");
    return GCV_OK;
}

// The quintic Newton start serves FP32 GL-3 builds that take the FP64 step.
static bool f32_direct_ok(const gcv_table* t) {
    return t->o.precision == GCV_FP32 && t->o.gl_order == 3 && !t->o.skip_fp64_step;
}

// Sampling with the block maps in blk_curve / blk_int (n_total / kSampleBlock
// entries each, rounded up): the kernels of gcv_sample. `profiled`: the
// launches are timed when the context profiles (event creation included).
static int sample_launch(gcv_table* t, const int64_t* d_soff, int64_t n_total, const double* d_pos,
                         int uniform_closed, int is_relative, double* d_out_u, int32_t* d_out_seg,
                         int32_t* blk_curve, int32_t* blk_int, bool profiled) {
    gcv_ctx* c = t->ctx;
    const int64_t n_blk = (n_total + dev::kSampleBlock - 1) / dev::kSampleBlock;
    dev::SampleArgs A{};
    A.n_curve = t->n_curve;
    A.n_total = n_total;
    A.soff = d_soff;
    A.pos = d_pos;
    A.uniform_closed = uniform_closed;
    A.is_relative = is_relative;
    A.ioff = t->ioff;
    A.n_int = t->n_int;
    A.int_cum = t->int_cum;
    A.curve_len = t->curve_len;
    A.root_prefix = t->root_prefix;
    A.root_k = t->seeded ? t->root_k : nullptr;
    A.root_off = t->seeded ? t->root_off : nullptr;
    A.deep = t->deep;
    A.log2_nsub = t->built_ls;
    A.eps0 = t->o.precision == GCV_FP32 ? (double)eps_of<float>(t) : eps_of<double>(t);
    A.ls_plain = t->log2_nsub;
    A.max_depth = t->o.max_depth;
    A.newton_iters = t->o.newton_iters;
    A.fp64_step = t->o.skip_fp64_step ? 0 : 1;
    // Newton start of B-spline samples (FP32 builds with the FP64 step)
    A.start_mode = GCV_START_PLAIN;
    if (t->kind == KIND_BSPLINE && f32_direct_ok(t)) {
        if (t->newton_start == GCV_START_QUINTIC) A.start_mode = GCV_START_QUINTIC;
        else if (t->newton_start == GCV_START_LEARNED && t->has_start) A.start_mode = GCV_START_LEARNED;
    }
    A.start_ok = t->start_ok;
    A.out_iters = t->iters_out;
    A.out_u = d_out_u;
    A.out_seg = d_out_seg;
    A.class_of_curve = t->class_of_curve;
    LaunchCfg L = cfg_of(c);
    void* const prof = profiled ? (void*)c : nullptr;  // the profiling hooks skip a null owner
    L.owner = prof;
    hipError_t e;
    prof_begin(prof, "k_block_curve");
    hipLaunchKernelGGL(dev::k_block_curve, dim3((unsigned)((t->n_curve + 255) / 256)), dim3(256), 0, c->stream,
                       t->n_curve, d_soff, n_blk, blk_curve);
    e = hipGetLastError();
    prof_end(prof);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "block-curve launch failed: %s", hipGetErrorString(e));
    if (!d_pos && t->n_int > t->n_curve) {
        // uniform targets are monotone along each curve: per-block interval
        // bounds shorten every sample's interval search
        prof_begin(prof, "k_block_interval");
        hipLaunchKernelGGL(dev::k_block_interval, dim3((unsigned)((n_blk + 255) / 256)), dim3(256), 0, c->stream, A,
                           (const int32_t*)blk_curve, blk_int);
        e = hipGetLastError();
        prof_end(prof);
        if (e != hipSuccess)
            return set_err(c, GCV_ERR_HIP, "block-interval launch failed: %s", hipGetErrorString(e));
        A.blk_int = blk_int;
    }
    const bool f32 = t->o.precision == GCV_FP32;
    const bool g3 = t->o.gl_order == 3;
    const int32_t* bc = blk_curve;
    if (t->kind == KIND_ELLIPSE) {
        if (f32) e = g3 ? Launch<float, 3>::sample_ellipse(L, t->a, t->b, t->o.normalize, A, bc)
                        : Launch<float, 5>::sample_ellipse(L, t->a, t->b, t->o.normalize, A, bc);
        else e = g3 ? Launch<double, 3>::sample_ellipse(L, t->a, t->b, t->o.normalize, A, bc)
                    : Launch<double, 5>::sample_ellipse(L, t->a, t->b, t->o.normalize, A, bc);
    } else {
        const BSplineTableDev T = bspline_dev(t);
        if (f32) e = g3 ? Launch<float, 3>::sample_bspline(L, t->uniform_class, T, A, bc)
                        : Launch<float, 5>::sample_bspline(L, t->uniform_class, T, A, bc);
        else e = g3 ? Launch<double, 3>::sample_bspline(L, t->uniform_class, T, A, bc)
                    : Launch<double, 5>::sample_bspline(L, t->uniform_class, T, A, bc);
    }
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "sample launch failed: %s", hipGetErrorString(e));
    return GCV_OK;
}

extern "C" int gcv_sample(gcv_table* t, const int64_t* d_soff, int64_t n_total, const double* d_pos,
                          int uniform_closed, int is_relative, double* d_out_u, int32_t* d_out_seg) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    if (int rc = sample_check(t, d_soff, n_total, d_out_u)) return rc;
    if (n_total == 0) return GCV_OK;
    // block maps of the table (grown on demand; reused while large enough)
    const int64_t n_blk = (n_total + dev::kSampleBlock - 1) / dev::kSampleBlock;
    if (n_blk > t->blk_cap) {
        if (c->capturing)
            return set_err(c, GCV_ERR_STATE, "sampling %lld points needs a larger block map: run the call once "
                           "before capturing it", (long long)n_total);
        if (t->blk_curve) {
            hipStreamSynchronize(c->stream);
            gcv_free(c, t->blk_curve);
            gcv_free(c, t->blk_int);
            t->blk_curve = t->blk_int = nullptr;
            t->blk_cap = 0;
        }
        void *p = nullptr, *q = nullptr;
        int rc = gcv_alloc(c, sizeof(int32_t) * (size_t)n_blk, &p);
        if (rc) return rc;
        rc = gcv_alloc(c, sizeof(int32_t) * (size_t)n_blk, &q);
        if (rc) {
            gcv_free(c, p);
            return rc;
        }
        t->blk_curve = (int32_t*)p;
        t->blk_int = (int32_t*)q;
        t->blk_cap = n_blk;
    }
    return sample_launch(t, d_soff, n_total, d_pos, uniform_closed, is_relative, d_out_u, d_out_seg, t->blk_curve,
                         t->blk_int, true);
}

extern "C" int gcv_sample_noalloc(gcv_table* t, const int64_t* d_soff, int64_t n_total, const double* d_pos,
                                  int uniform_closed, int is_relative, double* d_out_u, int32_t* d_out_seg) {
    if (t) use_device(t->ctx);
    if (!t) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    if (int rc = sample_check(t, d_soff, n_total, d_out_u)) return rc;
    if (n_total == 0) return GCV_OK;
    const int64_t n_blk = (n_total + dev::kSampleBlock - 1) / dev::kSampleBlock;
    if (n_blk > kNoAllocBlocks)
        return set_err(c, GCV_ERR_INVALID,
                       "gcv_sample_noalloc: %lld samples exceed its static scratch (at most %lld samples)",
                       (long long)n_total, (long long)(kNoAllocBlocks * dev::kSampleBlock));
    if (!c->scratch_ptr) return set_err(c, GCV_ERR_STATE, "gcv_sample_noalloc: device scratch not resolved");
    // the device copy of `scratch`, never the host-side symbol
    int32_t* blk_curve = c->scratch_ptr;
    int32_t* blk_int = c->scratch_ptr + kNoAllocBlocks;
    return sample_launch(t, d_soff, n_total, d_pos, uniform_closed, is_relative, d_out_u, d_out_seg, blk_curve,
                         blk_int, false);
}

extern "C" int gcv_host_bspline_span_poly(int p, int D, int s, const double* knots, const double* ctrl,
                                          int ctrl_stride, double scale_inv, double* out, double* u0, double* h) {
    if (p < 1 || p > 5 || D < 2 || D > 3 || !knots || !ctrl || !out || ctrl_stride < D) return GCV_ERR_INVALID;
    double tmp[dev::kCoefStride];
    for (int i = 0; i < dev::kCoefStride; i++) tmp[i] = 0.0;
    int rc = dev::bspline_span_poly(p, D, s, knots, ctrl, ctrl_stride, scale_inv, tmp, u0, h);
    for (int i = 0; i < p * D; i++) out[i] = tmp[i];
    return rc ? GCV_ERR_INVALID : GCV_OK;
}

// ===========================================================================
// Learned seed (MLP)
// ===========================================================================
struct gcv_mlp {
    gcv_ctx* ctx = nullptr;
    float* p = nullptr;       // [kParams]
    float* mean = nullptr;    // [16]
    float* istd = nullptr;    // [16]
    uint16_t* w1b = nullptr;
    uint16_t* w2b = nullptr;
    uint16_t* w3b = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;
    int64_t step = 0;
    std::vector<void*> owned;
};

static dev::MlpParams mlp_dev(const gcv_mlp* m) {
    return dev::MlpParams{m->p, m->mean, m->istd, m->w1b, m->w2b, m->w3b};
}

template <typename T>
static int malloc_m(gcv_mlp* m, T** p, size_t count) {
    void* v = nullptr;
    int rc = gcv_alloc(m->ctx, sizeof(T) * count, &v);
    if (rc) return rc;
    m->owned.push_back(v);
    *p = (T*)v;
    return GCV_OK;
}

extern "C" int gcv_mlp_create(gcv_ctx* c, gcv_mlp** out) {
    use_device(c);
    if (!c || !out) return GCV_ERR_INVALID;
    gcv_mlp* m = new gcv_mlp;
    m->ctx = c;
    int rc;
    if ((rc = malloc_m(m, &m->p, dev::kParams)) || (rc = malloc_m(m, &m->mean, dev::kFeat)) ||
        (rc = malloc_m(m, &m->istd, dev::kFeat)) || (rc = malloc_m(m, &m->w1b, dev::kFeat * dev::kHid)) ||
        (rc = malloc_m(m, &m->w2b, dev::kHid * dev::kHid)) || (rc = malloc_m(m, &m->w3b, dev::kHid * 16)) ||
        (rc = malloc_m(m, &m->adam_m, dev::kParams)) || (rc = malloc_m(m, &m->adam_v, dev::kParams)) ||
        (rc = malloc_m(m, &m->grad, dev::kParams)) || (rc = malloc_m(m, &m->loss, 2))) {
        for (void* p : m->owned) gcv_free(c, p);
        delete m;
        return rc;
    }
    *out = m;
    return GCV_OK;
}

extern "C" int gcv_mlp_destroy(gcv_mlp* m) {
    if (m) use_device(m->ctx);
    if (!m) return GCV_OK;
    if (int rc = capture_guard(m->ctx, "gcv_mlp_destroy")) return rc;
    hipStreamSynchronize(m->ctx->stream);
    for (void* p : m->owned) gcv_free(m->ctx, p);
    delete m;
    return GCV_OK;
}

extern "C" int gcv_mlp_set_params(gcv_mlp* m, const float* params, const float* mean, const float* istd) {
    if (m) use_device(m->ctx);
    if (!m || !params || !mean || !istd) return GCV_ERR_INVALID;
    gcv_ctx* c = m->ctx;
    int rc;
    if ((rc = gcv_copy_h2d(c, m->p, params, sizeof(float) * dev::kParams)) ||
        (rc = gcv_copy_h2d(c, m->mean, mean, sizeof(float) * dev::kFeat)) ||
        (rc = gcv_copy_h2d(c, m->istd, istd, sizeof(float) * dev::kFeat)))
        return rc;
    std::vector<float> zero(dev::kParams, 0.f);
    if ((rc = gcv_copy_h2d(c, m->adam_m, zero.data(), sizeof(float) * dev::kParams)) ||
        (rc = gcv_copy_h2d(c, m->adam_v, zero.data(), sizeof(float) * dev::kParams)))
        return rc;
    m->step = 0;
    hipLaunchKernelGGL(dev::k_mlp_sync_bf16, dim3((dev::kParams + 255) / 256), dim3(256), 0, c->stream, mlp_dev(m));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GCV_OK;
}

extern "C" int gcv_mlp_get_params(gcv_mlp* m, float* params, float* mean, float* istd) {
    if (m) use_device(m->ctx);
    if (!m) return GCV_ERR_INVALID;
    gcv_ctx* c = m->ctx;
    int rc;
    if (params && (rc = gcv_copy_d2h(c, params, m->p, sizeof(float) * dev::kParams))) return rc;
    if (mean && (rc = gcv_copy_d2h(c, mean, m->mean, sizeof(float) * dev::kFeat))) return rc;
    if (istd && (rc = gcv_copy_d2h(c, istd, m->istd, sizeof(float) * dev::kFeat))) return rc;
    return GCV_OK;
}

static dev::FeatArgs feat_args(gcv_table* t) {
    dev::FeatArgs F{};
    F.n = t->n_int;
    F.kind = t->kind == KIND_ELLIPSE ? 0 : 1;
    F.a = t->a;
    F.b = t->b;
    F.normalize = t->o.normalize;
    F.int_curve = t->int_curve;
    F.degree = t->bat.degree;
    F.dim = t->bat.dim;
    F.ioff = t->ioff;
    F.ctrl_off = t->bat.ctrl_off;
    F.ctrl = t->bat.ctrl;
    F.ctrl_stride = t->bat.ctrl_stride;
    F.coef64 = t->coef64;
    F.coef32 = t->coef32;
    F.cstride = t->cstride;
    F.ls_plain = t->log2_nsub;
    F.h = t->h;
    F.tol0 = (float)(t->o.tolerance / (double)(1 << t->log2_nsub));
    return F;
}

extern "C" int gcv_mlp_features(gcv_table* t, float* d_out) {
    if (t) use_device(t->ctx);
    if (!t || (!d_out && t->n_int > 0)) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    hipError_t e = dispatch_build<SetupOnly>(t, false);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "setup launch failed: %s", hipGetErrorString(e));
    if (t->n_int == 0) return GCV_OK;
    prof_begin(c, "k_features");
    hipLaunchKernelGGL(dev::k_features, dim3((unsigned)((t->n_int + 255) / 256)), dim3(256), 0, c->stream,
                       feat_args(t), d_out);
    e = hipGetLastError();
    prof_end(c);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "features launch failed: %s", hipGetErrorString(e));
    return GCV_OK;
}

// Setup pass plus prediction: B-spline tables get their standardised feature
// rows and safety floor from the setup kernel and stream them through the
// network; ellipses compute features inside the prediction kernel.
static hipError_t setup_and_predict(gcv_mlp* m, gcv_table* t, int hedge, int8_t* d_k, float* d_raw,
                                    bool reset_stats, int8_t* d_start = nullptr) {
    gcv_ctx* c = t->ctx;
    const int64_t n = t->n_int;
    const dev::SeedOut O{d_k, d_raw, t->seed_floor, hedge, t->log2_nsub,
                         (float)(t->o.tolerance / (double)(1 << t->log2_nsub)), d_start};
    if (t->kind == KIND_BSPLINE) {
        const SeedPredict sp{mlp_dev(m), O};
        return dispatch_build<SetupWithSeed>(t, reset_stats, &sp);
    }
    hipError_t e = dispatch_build<SetupOnly>(t, reset_stats);
    if (e != hipSuccess || n == 0) return e;
    const int64_t waves = (n + 63) / 64;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((waves + 3) / 4, (int64_t)c->n_cu * 8));
    prof_begin(c, "k_mlp_predict");
    hipLaunchKernelGGL(dev::k_mlp_predict, dim3((unsigned)blocks), dim3(256), dev::kMlpLdsBytes + 4 * dev::kWaveLds,
                       c->stream,
                       feat_args(t), mlp_dev(m), O);
    e = hipGetLastError();
    prof_end(c);
    return e;
}

extern "C" int gcv_mlp_predict_k(gcv_mlp* m, gcv_table* t, int hedge, int8_t* d_k, float* d_raw) {
    if (m) use_device(m->ctx);
    if (!m || !t || (!d_k && t->n_int > 0)) return GCV_ERR_INVALID;
    if (m->ctx != t->ctx) return set_err(t->ctx, GCV_ERR_INVALID, "predictor and table belong to different contexts");
    hipError_t e = setup_and_predict(m, t, hedge, d_k, d_raw, false);
    if (e != hipSuccess) return set_err(t->ctx, GCV_ERR_HIP, "predict launch failed: %s", hipGetErrorString(e));
    return GCV_OK;
}

extern "C" int gcv_table_build_mlp(gcv_table* t, gcv_mlp* m, int hedge, int32_t* d_cost) {
    if (t) use_device(t->ctx);
    if (!t || !m) return GCV_ERR_INVALID;
    gcv_ctx* c = t->ctx;
    if (m->ctx != c) return set_err(c, GCV_ERR_INVALID, "predictor and table belong to different contexts");
    t->has_start = false;
    hipError_t e = setup_and_predict(m, t, hedge, t->seed_k, nullptr, true, t->start_ok);
    if (e == hipSuccess) e = dispatch_build<SeededAfterSetup>(t, (const int8_t*)t->seed_k, d_cost);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "learned-seed build failed: %s", hipGetErrorString(e));
    t->built = true;
    t->has_start = true;
    return GCV_OK;
}

extern "C" int gcv_table_seed_k(gcv_table* t, const int8_t** d_k) {
    if (t) use_device(t->ctx);
    if (!t || !d_k) return GCV_ERR_INVALID;
    *d_k = t->seed_k;
    return GCV_OK;
}

extern "C" int gcv_mlp_forward(gcv_mlp* m, const uint16_t* d_x, int64_t n, float* d_out) {
    if (m) use_device(m->ctx);
    if (!m || n < 0 || (n > 0 && (!d_x || !d_out))) return GCV_ERR_INVALID;
    gcv_ctx* c = m->ctx;
    if (n == 0) return GCV_OK;
    const int64_t tiles = (n + 15) / 16;
    const int64_t blocks = std::min<int64_t>(tiles, (int64_t)c->n_cu * 8);
    hipLaunchKernelGGL(dev::k_mlp_forward, dim3((unsigned)blocks), dim3(64), 0, c->stream, d_x, n, mlp_dev(m), d_out);
    HIPCHK(c, hipGetLastError());
    return GCV_OK;
}

static hipError_t launch_grad(gcv_mlp* m, const uint16_t* d_x, const float* d_y, int64_t n, float w_pass) {
    gcv_ctx* c = m->ctx;
    hipLaunchKernelGGL(dev::k_fill_u64, dim3((dev::kParams / 2 + 255) / 256), dim3(256), 0, c->stream,
                       (unsigned long long*)m->grad, dev::kParams / 2, 0ull);
    hipLaunchKernelGGL(dev::k_fill_u64, dim3(1), dim3(64), 0, c->stream, (unsigned long long*)m->loss, 1, 0ull);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t tiles = (n + 15) / 16;
    const int64_t blocks = std::min<int64_t>(tiles, (int64_t)c->n_cu * 8);
    prof_begin(c, "k_mlp_train");
    hipLaunchKernelGGL(dev::k_mlp_train, dim3((unsigned)blocks), dim3(64), dev::kTrainLds, c->stream, d_x, d_y, n,
                       mlp_dev(m), w_pass, m->grad, m->loss);
    e = hipGetLastError();
    prof_end(c);
    return e;
}

extern "C" int gcv_mlp_grad(gcv_mlp* m, const uint16_t* d_x, const float* d_y, int64_t n, float w_pass,
                            float* host_grad, float* host_loss) {
    if (m) use_device(m->ctx);
    if (!m || n <= 0 || !d_x || !d_y) return GCV_ERR_INVALID;
    hipError_t e = launch_grad(m, d_x, d_y, n, w_pass);
    if (e != hipSuccess) return set_err(m->ctx, GCV_ERR_HIP, "grad launch failed: %s", hipGetErrorString(e));
    int rc;
    if (host_grad && (rc = gcv_copy_d2h(m->ctx, host_grad, m->grad, sizeof(float) * dev::kParams))) return rc;
    if (host_loss && (rc = gcv_copy_d2h(m->ctx, host_loss, m->loss, sizeof(float)))) return rc;
    return GCV_OK;
}

extern "C" int gcv_mlp_train_step(gcv_mlp* m, const uint16_t* d_x, const float* d_y, int64_t n, float lr,
                                  float w_pass, float* host_loss) {
    if (m) use_device(m->ctx);
    if (!m || n <= 0 || !d_x || !d_y) return GCV_ERR_INVALID;
    gcv_ctx* c = m->ctx;
    hipError_t e = launch_grad(m, d_x, d_y, n, w_pass);
    if (e != hipSuccess) return set_err(c, GCV_ERR_HIP, "train launch failed: %s", hipGetErrorString(e));
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - powf(b1, (float)m->step), bc2 = 1.f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(dev::k_mlp_adam, dim3((dev::kParams + 255) / 256), dim3(256), 0, c->stream, mlp_dev(m),
                       m->adam_m, m->adam_v, m->grad, lr, b1, b2, 1e-8f, bc1, bc2);
    HIPCHK(c, hipGetLastError());
    if (host_loss) return gcv_copy_d2h(c, host_loss, m->loss, sizeof(float));
    return GCV_OK;
}
