// Context, storage and transfer entry points of the C ABI (include/rstsr_hip.h).
//
// Storage follows rstsr's ownership model (rstsr-core/src/storage/data.rs:11-43):
// the caller owns every buffer; the library frees only what rstsr_hip_free names.
// Every allocation made here is registered with its extent so that each kernel
// entry point can prove, on the host, that the layouts it is given stay inside
// the buffers they index before anything is launched.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <thread>
#include <vector>

#include "common.hpp"
#include "registry.hpp"

namespace rstsr {
namespace hip {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

/* ------------------------------ allocation registry ------------------------------ */

static std::mutex g_reg_mu;
static std::map<uintptr_t, size_t>& reg() {
    static std::map<uintptr_t, size_t> m;
    return m;
}

void registry_add(void* p, size_t bytes) {
    std::lock_guard<std::mutex> g(g_reg_mu);
    reg()[(uintptr_t)p] = bytes;
}
bool registry_remove(void* p) {
    std::lock_guard<std::mutex> g(g_reg_mu);
    return reg().erase((uintptr_t)p) > 0;
}

// The allocation containing address p: (base, bytes); bytes == 0 if none.
std::pair<uintptr_t, size_t> registry_find(const void* p) {
    std::lock_guard<std::mutex> g(g_reg_mu);
    auto& m = reg();
    auto it = m.upper_bound((uintptr_t)p);
    if (it == m.begin()) return {0, 0};
    --it;
    if ((uintptr_t)p >= it->first && (uintptr_t)p < it->first + std::max<size_t>(it->second, 1)) return *it;
    return {0, 0};
}

/* ------------------------------ block cache ------------------------------ */
// A buffer of rstsr_hip_malloc belongs to the context that allocated it (its owner)
// and has a capacity (its size class). rstsr_hip_free on the owner keeps the block in
// the owner's cache, and a later rstsr_hip_malloc of a size in the same class on that
// context reuses it, with no hipMalloc and no hipFree: each stalls the host (hipFree
// waits for the device, and a 512 MiB hipMalloc maps pages for about half a
// millisecond), which would otherwise sit between every allocating operation and the
// next. Reuse is ordered by the owner's stream: work queued before the free finishes
// before work queued after the reuse. Two things could break that order, and both are
// handled: a copy between contexts runs on one context's stream while the other owns
// a buffer (rstsr_hip_memcpy_d2d waits for such a copy before it returns), and a second
// context on the same device could still be reading a block (a free then waits for
// the whole device first). A buffer freed while the stream is captured into a graph
// is still used by every replay: it is held by that graph and freed by
// rstsr_hip_graph_destroy. Blocks are returned to the runtime when a hipMalloc fails
// (then retried), beyond the cache limit, by rstsr_hip_empty_cache and at
// rstsr_hip_ctx_destroy.
struct BlockInfo {
    rstsr_hip_ctx* owner;  // nullptr once the owner is destroyed
    size_t cap;
};
static std::map<void*, BlockInfo>& blocks() {  // guarded by g_reg_mu
    static std::map<void*, BlockInfo> m;
    return m;
}
// Buffers freed while a graph was being captured, held until that graph is destroyed
static std::map<void*, std::vector<void*>>& graph_blocks() {  // guarded by g_reg_mu
    static std::map<void*, std::vector<void*>> m;
    return m;
}
static std::map<int, int>& live_contexts() {  // device -> contexts alive; guarded by g_reg_mu
    static std::map<int, int> m;
    return m;
}

// Capacity class of a request: 256-byte multiples (vectorised tails never leave the
// allocation), 2 MiB multiples above 1 MiB.
static size_t size_class(size_t nbytes) {
    size_t nb = std::max<size_t>((nbytes + 255) & ~(size_t)255, 256);
    if (nb > ((size_t)1 << 20)) nb = (nb + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
    return nb;
}

// Return every cached block of ctx to the runtime.
static void release_cache(rstsr_hip_ctx* ctx) {
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> g(ctx->cache_mu);
        for (auto& kv : ctx->cache) drop.push_back(kv.second);
        ctx->cache.clear();
        ctx->cached_bytes = 0;
    }
    if (drop.empty()) return;
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (void* p : drop) {
        HIP_CHECK(hipFree(p));
        std::lock_guard<std::mutex> g(g_reg_mu);
        blocks().erase(p);
    }
}

static rstsr_hip_ctx* block_owner(const void* p) {
    auto f = registry_find(p);
    if (f.second == 0) return nullptr;
    std::lock_guard<std::mutex> g(g_reg_mu);
    auto it = blocks().find((void*)f.first);
    return it == blocks().end() ? nullptr : it->second.owner;
}

bool same_buffer(const void* x, const void* y) {
    auto a = registry_find(x), b = registry_find(y);
    if (a.second == 0 || b.second == 0) return true;
    return a.first == b.first;
}

void check_extent(const void* ptr, const Layout& l, size_t elem, const char* name) {
    if (l.size() == 0) return;
    if (!ptr) RSTSR_RAISE(InvalidValue, std::string("null device pointer for operand ") + name);
    auto b = l.bounds_index();
    auto a = registry_find(ptr);
    if (a.second == 0)
        RSTSR_RAISE(InvalidValue, std::string("operand ") + name + " is not inside a buffer from rstsr_hip_malloc");
    uintptr_t lo = (uintptr_t)ptr + b.first * elem;
    uintptr_t hi = (uintptr_t)ptr + b.second * elem;
    if (lo < a.first || hi > a.first + a.second)
        RSTSR_RAISE(ValueOutOfRange, std::string("layout of operand ") + name + " reaches outside its buffer");
}

void check_extent_rel(const void* ptr, size_t nd, const size_t* shape, const int64_t* stride, size_t elem,
                      const char* name) {
    int64_t lo = 0, hi = 0;
    for (size_t d = 0; d < nd; ++d) {
        if (shape[d] == 0) return;
        const int64_t span = (int64_t)(shape[d] - 1) * stride[d];
        (span < 0 ? lo : hi) += span;
    }
    if (!ptr) RSTSR_RAISE(InvalidValue, std::string("null device pointer for operand ") + name);
    auto a = registry_find(ptr);
    if (a.second == 0)
        RSTSR_RAISE(InvalidValue, std::string("operand ") + name + " is not inside a buffer from rstsr_hip_malloc");
    const intptr_t p0 = (intptr_t)ptr;
    if (p0 + lo * (intptr_t)elem < (intptr_t)a.first || p0 + (hi + 1) * (intptr_t)elem > (intptr_t)(a.first + a.second))
        RSTSR_RAISE(ValueOutOfRange, std::string("strides of operand ") + name + " reach outside its buffer");
}

void check_extent_bytes(const void* ptr, size_t bytes, const char* name) {
    if (bytes == 0) return;
    auto a = registry_find(ptr);
    if (a.second == 0)
        RSTSR_RAISE(InvalidValue, std::string("operand ") + name + " is not inside a buffer from rstsr_hip_malloc");
    if ((uintptr_t)ptr + bytes > a.first + a.second)
        RSTSR_RAISE(ValueOutOfRange, std::string("copy of operand ") + name + " reaches outside its buffer");
}

static void* grow(rstsr_hip_ctx* ctx, Workspace& w, size_t bytes) {
    const bool capturing = stream_capturing(ctx);
    // a kernel recorded now keeps this buffer's address: the graph takes it at graph_end
    if (capturing) w.captured = true;
    if (bytes <= w.bytes) return w.ptr;
    size_t nb = std::max<size_t>(bytes, w.bytes * 2);
    nb = (nb + 4095) & ~(size_t)4095;
    if (w.ptr) {
        if (capturing) {
            ctx->retired.push_back(w.ptr);
        } else {
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
            HIP_CHECK(hipFree(w.ptr));
        }
        w.ptr = nullptr;
        w.bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, nb);
    if (e != hipSuccess && !capturing) {
        // freed caller buffers cached on this context may be what is missing
        (void)hipGetLastError();
        release_cache(ctx);
        e = hipMalloc(&p, nb);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(ErrorKind::MemoryError, "workspace hipMalloc of " + std::to_string(nb) +
                                                " bytes failed: " + std::string(hipGetErrorString(e)));
    }
    w.ptr = p;
    w.bytes = nb;
    return p;
}

/* ---------------- host transfers ---------------- */

constexpr size_t STAGE_CHUNK = 32ull << 20;  // bytes per pinned chunk
constexpr size_t STAGE_MIN = 8ull << 20;     // smaller pageable copies go through the runtime directly

static bool host_is_pinned(const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

// Host memcpy split over a few threads (one thread cannot keep up with the DMA).
static void par_memcpy(void* d, const void* s, size_t n) {
    const size_t nt = n >= (8ull << 20) ? 4 : 1;
    if (nt == 1) {
        std::memcpy(d, s, n);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (size_t i = 0; i < nt; ++i) {
        const size_t o = i * per, len = o < n ? std::min(per, n - o) : 0;
        if (len) th.emplace_back([=] { std::memcpy((char*)d + o, (const char*)s + o, len); });
    }
    for (auto& t : th) t.join();
}

// Copy between pageable host memory and the device and return when it is done.
// mode 0 (default): a host pointer the runtime reports as pinned goes by direct
// DMA; a pageable copy of at least STAGE_MIN bytes runs through two pinned
// chunks, the host copy of one chunk overlapping the DMA of the other; smaller
// ones go through the runtime. Mode 1 forces the runtime path, mode 2 the
// staged path (for measurement).
static void host_transfer(rstsr_hip_ctx* ctx, void* dst, const void* src, size_t n, bool h2d, int mode = 0) {
    // refused before anything is enqueued: a synchronising copy would invalidate the capture
    if (stream_capturing(ctx)) RSTSR_RAISE(InvalidValue, "host transfers cannot be captured into a graph");
    const void* host = h2d ? src : dst;
    const bool staged = mode == 2 || (mode == 0 && n >= STAGE_MIN && !host_is_pinned(host));
    const hipMemcpyKind kind = h2d ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (!staged) {
        HIP_CHECK(hipMemcpyAsync(dst, src, n, kind, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return;
    }
    for (int i = 0; i < 2; ++i) {
        if (!ctx->stage[i]) HIP_CHECK(hipHostMalloc(&ctx->stage[i], STAGE_CHUNK, hipHostMallocDefault));
        if (!ctx->stage_ev[i]) HIP_CHECK(hipEventCreateWithFlags(&ctx->stage_ev[i], hipEventDisableTiming));
    }
    const size_t nchunk = (n + STAGE_CHUNK - 1) / STAGE_CHUNK;
    if (h2d) {
        for (size_t c = 0; c < nchunk; ++c) {
            const int b = (int)(c & 1);
            const size_t off = c * STAGE_CHUNK, len = std::min(STAGE_CHUNK, n - off);
            if (c >= 2) HIP_CHECK(hipEventSynchronize(ctx->stage_ev[b]));  // chunk b's previous DMA done
            par_memcpy(ctx->stage[b], (const char*)src + off, len);
            HIP_CHECK(hipMemcpyAsync((char*)dst + off, ctx->stage[b], len, kind, ctx->stream));
            HIP_CHECK(hipEventRecord(ctx->stage_ev[b], ctx->stream));
        }
    } else {
        // DMA of chunk c+1 runs while chunk c is copied out of its pinned buffer
        auto issue = [&](size_t c) {
            const int b = (int)(c & 1);
            const size_t off = c * STAGE_CHUNK, len = std::min(STAGE_CHUNK, n - off);
            HIP_CHECK(hipMemcpyAsync(ctx->stage[b], (const char*)src + off, len, kind, ctx->stream));
            HIP_CHECK(hipEventRecord(ctx->stage_ev[b], ctx->stream));
        };
        issue(0);
        for (size_t c = 0; c < nchunk; ++c) {
            const int b = (int)(c & 1);
            const size_t off = c * STAGE_CHUNK, len = std::min(STAGE_CHUNK, n - off);
            HIP_CHECK(hipEventSynchronize(ctx->stage_ev[b]));
            if (c + 1 < nchunk) issue(c + 1);
            par_memcpy((char*)dst + off, ctx->stage[b], len);
        }
    }
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

bool stream_capturing(rstsr_hip_ctx* ctx) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    HIP_CHECK(hipStreamIsCapturing(ctx->stream, &st));
    return st != hipStreamCaptureStatusNone;
}

static std::mutex g_tune_mu;
static std::map<std::string, int>& tune_table() {
    static std::map<std::string, int> t;
    static bool loaded = false;
    if (!loaded) {
        loaded = true;
        if (const char* f = std::getenv("RSTSR_HIP_TUNE_CACHE")) {
            if (FILE* fp = std::fopen(f, "r")) {
                char key[256];
                int v;
                while (std::fscanf(fp, "%255s %d", key, &v) == 2) t[key] = v;
                std::fclose(fp);
            }
        }
    }
    return t;
}

int tune_select(rstsr_hip_ctx* ctx, const std::string& key, int ncand, int heuristic, bool can_time,
                const std::function<void(int)>& run) {
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto& t = tune_table();
        auto it = t.find(key);
        if (it != t.end() && it->second < ncand) {
            run(it->second);
            return it->second;
        }
    }
    if (!can_time || ncand == 1 || stream_capturing(ctx) || std::getenv("RSTSR_HIP_NO_TUNE")) {
        run(heuristic);
        return heuristic;
    }
    // one untimed run first: the first touch of the operands (TLB and cache
    // warm-up) would otherwise be charged to whichever candidate runs first
    run(heuristic);
    float best = 1e30f;
    int bi = heuristic;
    // two rounds over the candidates, each candidate's best time counts
    for (int round = 0; round < 2; ++round) {
        for (int c = 0; c < ncand; ++c) {
            HIP_CHECK(hipEventRecord(ctx->events[14], ctx->stream));
            run(c);
            HIP_CHECK(hipEventRecord(ctx->events[15], ctx->stream));
            HIP_CHECK(hipEventSynchronize(ctx->events[15]));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, ctx->events[14], ctx->events[15]));
            if (ms < best) {
                best = ms;
                bi = c;
            }
        }
    }
    // the call's result comes from the configuration later calls will use
    if (bi != ncand - 1) run(bi);
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_table()[key] = bi;
    if (const char* f = std::getenv("RSTSR_HIP_TUNE_CACHE")) {
        if (FILE* fp = std::fopen(f, "a")) {
            std::fprintf(fp, "%s %d\n", key.c_str(), bi);
            std::fclose(fp);
        }
    }
    return bi;
}

void* workspace(rstsr_hip_ctx* ctx, size_t bytes) { return grow(ctx, ctx->ws, bytes); }
void* meta_stage(rstsr_hip_ctx* ctx, size_t bytes, const char* what) {
    if (stream_capturing(ctx)) RSTSR_RAISE(UnImplemented, std::string(what) + " cannot be captured into a graph");
    const int s = ctx->meta_slot;
    if (ctx->meta_ev[s]) HIP_CHECK(hipEventSynchronize(ctx->meta_ev[s]));  // its last copy has run
    if (ctx->meta_cap[s] < bytes) {
        if (ctx->meta_host[s]) HIP_CHECK(hipHostFree(ctx->meta_host[s]));
        ctx->meta_host[s] = nullptr;
        ctx->meta_cap[s] = 0;
        const size_t cap = (bytes + 65535) & ~(size_t)65535;
        HIP_CHECK(hipHostMalloc(&ctx->meta_host[s], cap, hipHostMallocDefault));
        ctx->meta_cap[s] = cap;
    }
    return ctx->meta_host[s];
}

void meta_upload(rstsr_hip_ctx* ctx, void* dst, size_t bytes) {
    const int s = ctx->meta_slot;
    HIP_CHECK(hipMemcpyAsync(dst, ctx->meta_host[s], bytes, hipMemcpyHostToDevice, ctx->stream));
    if (!ctx->meta_ev[s]) HIP_CHECK(hipEventCreateWithFlags(&ctx->meta_ev[s], hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(ctx->meta_ev[s], ctx->stream));
    ctx->meta_slot = s ^ 1;
}

void* aux_buffer(rstsr_hip_ctx* ctx, size_t bytes) { return grow(ctx, ctx->aux, bytes); }
void* zbuf_buffer(rstsr_hip_ctx* ctx, size_t bytes) { return grow(ctx, ctx->zbuf, bytes); }

}  // namespace hip
}  // namespace rstsr

using namespace rstsr;
using namespace rstsr::hip;

extern "C" {

const char* rstsr_hip_last_error(void) { return g_last_error.c_str(); }

int rstsr_hip_device_count(int* count) {
    return guarded([&] {
        if (!count) RSTSR_RAISE(InvalidValue, "null output");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = (e == hipSuccess) ? n : 0;
    });
}

int rstsr_hip_ctx_create(int device_id, rstsr_hip_ctx_t** out) {
    return guarded([&] {
        if (!out) RSTSR_RAISE(InvalidValue, "null output");
        *out = nullptr;
        int n = 0;
        HIP_CHECK(hipGetDeviceCount(&n));
        if (device_id < 0 || device_id >= n)
            RSTSR_RAISE(DeviceError, "device id " + std::to_string(device_id) + " not present (" + std::to_string(n) + " devices)");
        auto* ctx = new rstsr_hip_ctx();
        ctx->device = device_id;
        try {
            HIP_CHECK(hipSetDevice(device_id));
            HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            hipDeviceProp_t prop;
            HIP_CHECK(hipGetDeviceProperties(&prop, device_id));
            ctx->num_cu = prop.multiProcessorCount;
            reduce_scratch_setup(device_id);
            ctx->l2_bytes = prop.l2CacheSize;
            HIP_CHECK(hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocDefault));
            for (auto& ev : ctx->events) HIP_CHECK(hipEventCreate(&ev));
            // cache limit: RSTSR_HIP_CACHE_BYTES, else a quarter of the device memory
            ctx->cache_max = prop.totalGlobalMem / 4;
            if (const char* e = std::getenv("RSTSR_HIP_CACHE_BYTES")) ctx->cache_max = (size_t)std::strtoull(e, nullptr, 10);
        } catch (...) {
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
            if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
            delete ctx;
            throw;
        }
        {
            std::lock_guard<std::mutex> g(g_reg_mu);
            live_contexts()[device_id] += 1;
        }
        *out = ctx;
    });
}

int rstsr_hip_ctx_destroy(rstsr_hip_ctx_t* ctx) {
    return guarded([&] {
        if (!ctx) return;
        bind(ctx);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        release_cache(ctx);
        {
            // buffers still held by callers outlive the context: a later free returns them
            std::lock_guard<std::mutex> g(g_reg_mu);
            for (auto& kv : blocks())
                if (kv.second.owner == ctx) kv.second.owner = nullptr;
            live_contexts()[ctx->device] -= 1;
        }
        if (ctx->ws.ptr) HIP_CHECK(hipFree(ctx->ws.ptr));
        if (ctx->aux.ptr) HIP_CHECK(hipFree(ctx->aux.ptr));
        if (ctx->zbuf.ptr) HIP_CHECK(hipFree(ctx->zbuf.ptr));
        for (void* p : ctx->retired) HIP_CHECK(hipFree(p));
        for (void* p : ctx->deferred) HIP_CHECK(hipFree(p));
        for (auto& ev : ctx->events)
            if (ev) HIP_CHECK(hipEventDestroy(ev));
        if (ctx->order_ev) HIP_CHECK(hipEventDestroy(ctx->order_ev));
        if (ctx->host_pinned) HIP_CHECK(hipHostFree(ctx->host_pinned));
        for (int i = 0; i < 2; ++i) {
            if (ctx->stage[i]) HIP_CHECK(hipHostFree(ctx->stage[i]));
            if (ctx->stage_ev[i]) HIP_CHECK(hipEventDestroy(ctx->stage_ev[i]));
            if (ctx->meta_host[i]) HIP_CHECK(hipHostFree(ctx->meta_host[i]));
            if (ctx->meta_ev[i]) HIP_CHECK(hipEventDestroy(ctx->meta_ev[i]));
        }
        HIP_CHECK(hipStreamDestroy(ctx->stream));
        delete ctx;
    });
}

int rstsr_hip_same_device(const rstsr_hip_ctx_t* a, const rstsr_hip_ctx_t* b) {
    return (a && b && a->device == b->device) ? 1 : 0;
}

int rstsr_hip_synchronize(rstsr_hip_ctx_t* ctx) {
    return guarded([&] {
        bind(ctx);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int rstsr_hip_stream_wait(rstsr_hip_ctx_t* ctx, rstsr_hip_ctx_t* other) {
    return guarded([&] {
        if (!ctx || !other) RSTSR_RAISE(InvalidValue, "null context");
        if (ctx == other || ctx->stream == other->stream) return;
        if (stream_capturing(ctx)) {
            // an event recorded outside the capture cannot join the graph: wait on the host
            HIP_CHECK(hipSetDevice(other->device));
            HIP_CHECK(hipStreamSynchronize(other->stream));
            bind(ctx);
            return;
        }
        {
            std::lock_guard<std::mutex> g(other->mu);
            HIP_CHECK(hipSetDevice(other->device));
            if (!other->order_ev) HIP_CHECK(hipEventCreateWithFlags(&other->order_ev, hipEventDisableTiming));
            HIP_CHECK(hipEventRecord(other->order_ev, other->stream));
            bind(ctx);
            HIP_CHECK(hipStreamWaitEvent(ctx->stream, other->order_ev, 0));
        }
    });
}

int rstsr_hip_get_stream(rstsr_hip_ctx_t* ctx, void** stream) {
    return guarded([&] {
        if (!ctx || !stream) RSTSR_RAISE(InvalidValue, "null argument");
        *stream = (void*)ctx->stream;
    });
}

int rstsr_hip_malloc(rstsr_hip_ctx_t* ctx, size_t nbytes, void** ptr) {
    return guarded([&] {
        if (!ptr) RSTSR_RAISE(InvalidValue, "null output");
        bind(ctx);
        const size_t cls = size_class(nbytes);
        void* p = nullptr;
        {
            // a cached block of this class, or up to 1/8 larger
            std::lock_guard<std::mutex> g(ctx->cache_mu);
            auto it = ctx->cache.lower_bound(cls);
            if (it != ctx->cache.end() && it->first <= cls + cls / 8) {
                p = it->second;
                ctx->cached_bytes -= it->first;
                ctx->cache.erase(it);
            }
        }
        if (!p) {
            hipError_t e = hipMalloc(&p, cls);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                release_cache(ctx);
                e = hipMalloc(&p, cls);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(ErrorKind::MemoryError,
                            "hipMalloc of " + std::to_string(nbytes) + " bytes failed: " + hipGetErrorString(e));
            }
            std::lock_guard<std::mutex> g(g_reg_mu);
            blocks()[p] = BlockInfo{ctx, cls};
        }
        registry_add(p, nbytes);
        *ptr = p;
    });
}

int rstsr_hip_free(rstsr_hip_ctx_t* ctx, void* ptr) {
    return guarded([&] {
        if (!ptr) return;
        bind(ctx);
        if (!registry_remove(ptr)) RSTSR_RAISE(InvalidValue, "free of a pointer not returned by rstsr_hip_malloc");
        BlockInfo bi{nullptr, 0};
        int live = 0;
        {
            std::lock_guard<std::mutex> g(g_reg_mu);
            auto it = blocks().find(ptr);
            if (it != blocks().end()) bi = it->second;
            live = live_contexts()[ctx->device];
        }
        const bool capturing = stream_capturing(ctx);
        if (bi.owner == ctx && !capturing) {
            // another context on this device may still read the block
            if (live > 1) HIP_CHECK(hipDeviceSynchronize());
            std::lock_guard<std::mutex> g(ctx->cache_mu);
            if (ctx->cached_bytes + bi.cap <= ctx->cache_max) {
                ctx->cache.emplace(bi.cap, ptr);
                ctx->cached_bytes += bi.cap;
                return;
            }
        }
        if (capturing) {
            // the graph being captured still writes and reads this buffer at every replay:
            // it belongs to the graph from here on (rstsr_hip_graph_end), never to the cache
            ctx->deferred.push_back(ptr);
            return;
        }
        // returned to the runtime once every queued use has finished
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipFree(ptr));
        {
            std::lock_guard<std::mutex> g(g_reg_mu);
            blocks().erase(ptr);
        }
        for (void* p : ctx->deferred) {
            HIP_CHECK(hipFree(p));
            std::lock_guard<std::mutex> g(g_reg_mu);
            blocks().erase(p);
        }
        ctx->deferred.clear();
    });
}

int rstsr_hip_empty_cache(rstsr_hip_ctx_t* ctx) {
    return guarded([&] {
        bind(ctx);
        release_cache(ctx);
    });
}

int rstsr_hip_cached_bytes(rstsr_hip_ctx_t* ctx, size_t* bytes) {
    return guarded([&] {
        if (!ctx || !bytes) RSTSR_RAISE(InvalidValue, "null argument");
        std::lock_guard<std::mutex> g(ctx->cache_mu);
        *bytes = ctx->cached_bytes;
    });
}

int rstsr_hip_memcpy_h2d(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes) {
    return guarded([&] {
        if (nbytes == 0) return;
        if (!src) RSTSR_RAISE(InvalidValue, "null host source");
        bind(ctx);
        check_extent_bytes(dst, nbytes, "dst");
        std::lock_guard<std::mutex> g(ctx->mu);
        host_transfer(ctx, dst, src, nbytes, true);
    });
}

int rstsr_hip_memcpy_d2h(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes) {
    return guarded([&] {
        if (nbytes == 0) return;
        if (!dst) RSTSR_RAISE(InvalidValue, "null host destination");
        bind(ctx);
        check_extent_bytes(src, nbytes, "src");
        std::lock_guard<std::mutex> g(ctx->mu);
        host_transfer(ctx, dst, src, nbytes, false);
    });
}

int rstsr_hip_memcpy_h2d_mode(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes, int mode) {
    return guarded([&] {
        if (nbytes == 0) return;
        if (!src) RSTSR_RAISE(InvalidValue, "null host source");
        if (mode < 0 || mode > 2) RSTSR_RAISE(InvalidValue, "transfer mode must be 0, 1 or 2");
        bind(ctx);
        check_extent_bytes(dst, nbytes, "dst");
        std::lock_guard<std::mutex> g(ctx->mu);
        host_transfer(ctx, dst, src, nbytes, true, mode);
    });
}

int rstsr_hip_memcpy_d2h_mode(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes, int mode) {
    return guarded([&] {
        if (nbytes == 0) return;
        if (!dst) RSTSR_RAISE(InvalidValue, "null host destination");
        if (mode < 0 || mode > 2) RSTSR_RAISE(InvalidValue, "transfer mode must be 0, 1 or 2");
        bind(ctx);
        check_extent_bytes(src, nbytes, "src");
        std::lock_guard<std::mutex> g(ctx->mu);
        host_transfer(ctx, dst, src, nbytes, false, mode);
    });
}

int rstsr_hip_memcpy_d2d(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes) {
    return guarded([&] {
        if (nbytes == 0) return;
        bind(ctx);
        check_extent_bytes(dst, nbytes, "dst");
        check_extent_bytes(src, nbytes, "src");
        HIP_CHECK(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, ctx->stream));
        // a buffer of another context may be freed into that context's cache and
        // reused on its stream as soon as this returns: finish the copy first
        rstsr_hip_ctx* od = block_owner(dst);
        rstsr_hip_ctx* os = block_owner(src);
        if ((od && od != ctx) || (os && os != ctx)) {
            if (stream_capturing(ctx)) RSTSR_RAISE(UnImplemented, "copy between contexts during stream capture");
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

int rstsr_hip_event_record(rstsr_hip_ctx_t* ctx, int slot) {
    return guarded([&] {
        bind(ctx);
        if (slot < 0 || slot >= 16) RSTSR_RAISE(InvalidValue, "event slot out of range [0,16)");
        HIP_CHECK(hipEventRecord(ctx->events[slot], ctx->stream));
    });
}

int rstsr_hip_event_elapsed(rstsr_hip_ctx_t* ctx, int s0, int s1, float* ms) {
    return guarded([&] {
        bind(ctx);
        if (s0 < 0 || s0 >= 16 || s1 < 0 || s1 >= 16 || !ms) RSTSR_RAISE(InvalidValue, "bad event arguments");
        HIP_CHECK(hipEventSynchronize(ctx->events[s1]));
        HIP_CHECK(hipEventElapsedTime(ms, ctx->events[s0], ctx->events[s1]));
    });
}

int rstsr_hip_graph_begin(rstsr_hip_ctx_t* ctx) {
    return guarded([&] {
        bind(ctx);
        HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    });
}

// Buffers a capture leaves behind: caller buffers freed during it, workspaces replaced
// during it, and the workspaces its kernels were handed (their addresses are baked into
// the graph). The context forgets the workspaces; the next call starts fresh ones.
static std::vector<void*> take_capture_buffers(rstsr_hip_ctx* ctx) {
    std::lock_guard<std::mutex> g(ctx->mu);
    std::vector<void*> held;
    held.swap(ctx->deferred);
    held.insert(held.end(), ctx->retired.begin(), ctx->retired.end());
    ctx->retired.clear();
    for (Workspace* w : {&ctx->ws, &ctx->aux, &ctx->zbuf}) {
        if (!w->captured) continue;
        if (w->ptr) held.push_back(w->ptr);
        w->ptr = nullptr;
        w->bytes = 0;
        w->captured = false;
    }
    return held;
}

static void free_now(const std::vector<void*>& held) {
    if (held.empty()) return;
    HIP_CHECK(hipDeviceSynchronize());  // work queued before the capture may still use them
    for (void* p : held) {
        HIP_CHECK(hipFree(p));
        std::lock_guard<std::mutex> lg(g_reg_mu);
        blocks().erase(p);
    }
}

int rstsr_hip_graph_end(rstsr_hip_ctx_t* ctx, void** graph_exec) {
    return guarded([&] {
        bind(ctx);
        if (!graph_exec) RSTSR_RAISE(InvalidValue, "null output");
        *graph_exec = nullptr;
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        std::vector<void*> held = take_capture_buffers(ctx);
        if (ec != hipSuccess) {
            (void)hipGetLastError();
            if (g) (void)hipGraphDestroy(g);
            free_now(held);  // no graph uses them
            throw Error(ErrorKind::DeviceError, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        }
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            free_now(held);
            throw Error(ErrorKind::DeviceError, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        if (!held.empty()) {
            std::lock_guard<std::mutex> lg(g_reg_mu);
            graph_blocks()[(void*)ex] = std::move(held);
        }
        *graph_exec = (void*)ex;
    });
}

int rstsr_hip_graph_launch(rstsr_hip_ctx_t* ctx, void* graph_exec) {
    return guarded([&] {
        bind(ctx);
        if (!graph_exec) RSTSR_RAISE(InvalidValue, "null graph");
        HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, ctx->stream));
    });
}

int rstsr_hip_graph_destroy(void* graph_exec) {
    return guarded([&] {
        if (!graph_exec) return;
        std::vector<void*> held;
        {
            std::lock_guard<std::mutex> lg(g_reg_mu);
            auto it = graph_blocks().find(graph_exec);
            if (it != graph_blocks().end()) {
                held.swap(it->second);
                graph_blocks().erase(it);
            }
        }
        // a launch of the graph may still be running: it uses the buffers it holds
        if (!held.empty()) HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
        for (void* p : held) {
            HIP_CHECK(hipFree(p));
            std::lock_guard<std::mutex> lg(g_reg_mu);
            blocks().erase(p);
        }
    });
}

int rstsr_hip_reduce_out_layout(const rstsr_layout_t* la, size_t naxes, const ptrdiff_t* axes, size_t* ndim_out,
                                size_t* shape_out, ptrdiff_t* stride_out) {
    return guarded([&] {
        if (!ndim_out || (la && la->ndim && (!shape_out || !stride_out))) RSTSR_RAISE(InvalidValue, "null output");
        Layout l = from_abi(la);
        std::vector<ptrdiff_t> ax(axes, axes + naxes);
        auto sp = l.dim_split_axes(ax);
        Layout lo = layout_for_array_copy(sp.second, IterOrder::K);
        *ndim_out = lo.ndim();
        for (size_t i = 0; i < lo.ndim(); ++i) {
            shape_out[i] = lo.shape[i];
            stride_out[i] = lo.stride[i];
        }
    });
}

}  // extern "C"
