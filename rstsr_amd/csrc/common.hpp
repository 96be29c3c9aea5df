// Shared definitions of the gfx950 backend: context, error plumbing, element
// types, index helpers and launch utilities. Compiled only by hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rstsr_hip.h"
#include "layout.hpp"

namespace rstsr {
namespace hip {

constexpr int MAXD = 8;  // dimensions a kernel sees after axis collapsing

/* ------------------------------ errors ------------------------------ */

void set_last_error(const std::string& msg);

#define HIP_CHECK(expr)                                                                                 \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            throw ::rstsr::Error(::rstsr::ErrorKind::DeviceError,                                       \
                                 std::string(#expr) + " failed: " + hipGetErrorString(_e));             \
    } while (0)

// Run `body` and convert any exception into a status code + thread-local message.
template <class F>
inline int guarded(F&& body) {
    try {
        body();
        return RSTSR_OK;
    } catch (const ::rstsr::Error& e) {
        set_last_error(std::string(error_kind_name(e.kind)) + ": " + e.what());
        return (int)e.kind;
    } catch (const std::bad_alloc&) {
        set_last_error("MemoryError: host allocation failed");
        return RSTSR_ERR_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(std::string("RuntimeError: ") + e.what());
        return RSTSR_ERR_RUNTIME;
    }
}

/* ------------------------------ context ------------------------------ */

// Device workspace that grows on demand and is reused across calls on one
// context (partials of split reductions, scalar results, GEMM slices).
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool captured = false;  // handed to a kernel while the stream was being captured
};

}  // namespace hip
}  // namespace rstsr

struct rstsr_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t l2_bytes = 0;
    std::mutex mu;  // serialises workspace growth and host staging on one context
    rstsr::hip::Workspace ws;   // partials, residue planes, scalar results
    rstsr::hip::Workspace aux;  // small metadata (grouped GEMM descriptors), BLAS driver scratch
    rstsr::hip::Workspace zbuf;  // temporaries of the three-product complex split GEMM
    std::vector<void*> retired;  // workspaces replaced while the stream was being captured
    std::vector<void*> deferred;  // caller buffers freed while the stream was being captured
    hipEvent_t order_ev = nullptr;  // recorded on this stream when another context must wait for it
    void* host_pinned = nullptr;  // 4 KiB pinned staging for scalar results
    hipEvent_t events[16] = {};
    // two pinned chunks for host transfers of pageable memory (allocated on first use)
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // two pinned host slots for per-call problem tables (meta_stage / meta_upload)
    void* meta_host[2] = {nullptr, nullptr};
    size_t meta_cap[2] = {0, 0};
    hipEvent_t meta_ev[2] = {nullptr, nullptr};
    int meta_slot = 0;
    // freed caller buffers kept for reuse, by capacity (rstsr_hip_malloc / rstsr_hip_free)
    std::mutex cache_mu;
    std::multimap<size_t, void*> cache;
    size_t cached_bytes = 0;
    size_t cache_max = 0;  // bytes; frees beyond it return memory to the runtime
};

namespace rstsr {
namespace hip {

// Make the context's device current on this thread.
inline void bind(rstsr_hip_ctx* ctx) {
    if (!ctx) RSTSR_RAISE(InvalidValue, "null context");
    HIP_CHECK(hipSetDevice(ctx->device));
}

// Return a device buffer of at least `bytes` owned by the context. Growth is
// ordered after all work already enqueued on the context stream; during stream
// capture the old buffer is retired instead of freed. A workspace handed out
// during a capture belongs to the graph from rstsr_hip_graph_end on (the graph's
// kernels keep its address; it is freed with the graph), and the context starts
// a fresh one, so later eager growth never frees memory a replay writes.
void* workspace(rstsr_hip_ctx* ctx, size_t bytes);
void* aux_buffer(rstsr_hip_ctx* ctx, size_t bytes);
void* zbuf_buffer(rstsr_hip_ctx* ctx, size_t bytes);

// Resolve the calling device's fixed reduction scratch (reduce.hip) once per device; run by
// rstsr_hip_ctx_create with the context's device current.
void reduce_scratch_setup(int device);

// True while the context stream is being captured into a graph.
bool stream_capturing(rstsr_hip_ctx* ctx);

// Per-call host tables (problem descriptors, prefix sums) reach the device without the
// host waiting for the copy: meta_stage returns a pinned host region of `bytes` for the
// caller to fill, and meta_upload copies it to `dst` on the context stream. Two regions
// alternate, and each is refilled only after its previous copy has run (an event), so the
// host prepares the next call while the device works on this one. Not available while
// the stream is captured (a replay would read a region refilled since).
void* meta_stage(rstsr_hip_ctx* ctx, size_t bytes, const char* what);
void meta_upload(rstsr_hip_ctx* ctx, void* dst, size_t bytes);

// Autotuner: run(c) performs the whole call with candidate c (all candidates give
// identical results). The first call for `key` with can_time runs the heuristic
// once untimed (operand warm-up), then times every candidate twice with events on
// the context stream and keeps the fastest, which then runs once more, so that
// the tuning call's output is the chosen candidate's; later calls use the cached choice. No
// timing while the stream is captured, when can_time is false (the output is
// also an input) or under RSTSR_HIP_NO_TUNE: then the cached choice or the
// heuristic runs. RSTSR_HIP_TUNE_CACHE=<file> loads and appends choices.
int tune_select(rstsr_hip_ctx* ctx, const std::string& key, int ncand, int heuristic, bool can_time,
                const std::function<void(int)>& run);

/* ------------------------------ element types ------------------------------ */

struct c64 {
    float re, im;
};
struct c128 {
    double re, im;
};

inline size_t dtype_size(rstsr_dtype_t t) {
    switch (t) {
        case RSTSR_F32: return 4;
        case RSTSR_F64: return 8;
        case RSTSR_I32: return 4;
        case RSTSR_I64: return 8;
        case RSTSR_C64: return 8;
        case RSTSR_C128: return 16;
        case RSTSR_U64: return 8;
        case RSTSR_BOOL: return 1;
    }
    RSTSR_RAISE(InvalidValue, "unknown dtype");
}

/* ------------------------------ layouts across the ABI ------------------------------ */

inline Layout from_abi(const rstsr_layout_t* l) {
    if (!l) RSTSR_RAISE(InvalidValue, "null layout");
    if (l->ndim > 0 && (!l->shape || !l->stride)) RSTSR_RAISE(InvalidValue, "layout without shape/stride arrays");
    Shape s(l->shape, l->shape + l->ndim);
    Stride t(l->stride, l->stride + l->ndim);
    return Layout(s, t, l->offset);
}

/* ------------------------------ fast integer division ------------------------------ */

// q = n / d for 32-bit n via a multiply-high (Granlund-Montgomery round-up method).
struct FastDiv {
    uint32_t d = 1, m = 0, s = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t dv) : d(dv) {
        if (dv == 0) RSTSR_RAISE(InvalidValue, "division by zero extent");
        uint32_t sh = 0;
        while ((1ull << sh) < dv) ++sh;
        s = sh;
        m = (uint32_t)((((1ull << 32) * ((1ull << sh) - dv)) / dv) + 1);
    }
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    uint32_t t = __umulhi(n, f.m);
    return (uint32_t)(((uint64_t)t + n) >> f.s);
}

/* ------------------------------ streaming loads ------------------------------ */

// Load of an element the launch reads exactly once: the nontemporal hint lets it
// stream past the caches. Measured on MI355X, 2 GiB read-only sweeps sustain 7.1
// TB/s with it and 6.3 TB/s without. Operands that are re-read (broadcast axes)
// keep ordinary loads.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) {
    if constexpr (std::is_arithmetic<T>::value) {
        return __builtin_nontemporal_load(p);
    } else if constexpr (sizeof(T) == 16 && alignof(T) >= 16) {
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        u4 w = __builtin_nontemporal_load((const u4*)p);
        T v;
        __builtin_memcpy(&v, &w, 16);
        return v;
    } else if constexpr (sizeof(T) % 8 == 0 && alignof(T) >= 8) {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        T v;
        const u2* q = (const u2*)p;
#pragma unroll
        for (size_t i = 0; i < sizeof(T) / 8; ++i) {
            u2 w = __builtin_nontemporal_load(q + i);
            __builtin_memcpy((char*)&v + 8 * i, &w, 8);
        }
        return v;
    } else if constexpr (sizeof(T) % 4 != 0 || alignof(T) < 4) {
        return *p;  // byte-sized elements (bool): plain load
    } else {
        T v;
        const unsigned* q = (const unsigned*)p;
        unsigned* d = (unsigned*)&v;
#pragma unroll
        for (size_t i = 0; i < sizeof(T) / 4; ++i) d[i] = __builtin_nontemporal_load(q + i);
        return v;
    }
}

// Cached or streaming load chosen at compile time. The choice must stay a
// template argument: two runtime branches that differ only in the hint are
// merged by the compiler, and the merged load keeps no hint.
template <bool NT, class T>
__device__ __forceinline__ T ld(const T* p) {
    if constexpr (NT) return ld_nt(p);
    else return *p;
}

// Store of an output written once and not read back soon (large outputs): the
// nontemporal hint keeps it from displacing cached inputs.
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) {
    if constexpr (std::is_arithmetic<T>::value) {
        __builtin_nontemporal_store(v, p);
    } else if constexpr (sizeof(T) == 16 && alignof(T) >= 16) {
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        u4 w;
        __builtin_memcpy(&w, &v, 16);
        __builtin_nontemporal_store(w, (u4*)p);
    } else if constexpr (sizeof(T) % 8 == 0 && alignof(T) >= 8) {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        u2* q = (u2*)p;
#pragma unroll
        for (size_t i = 0; i < sizeof(T) / 8; ++i) {
            u2 w;
            __builtin_memcpy(&w, (const char*)&v + 8 * i, 8);
            __builtin_nontemporal_store(w, q + i);
        }
    } else if constexpr (sizeof(T) % 4 == 0 && alignof(T) >= 4) {
        unsigned* q = (unsigned*)p;
        const unsigned* s = (const unsigned*)&v;
#pragma unroll
        for (size_t i = 0; i < sizeof(T) / 4; ++i) __builtin_nontemporal_store(s[i], q + i);
    } else {
        *p = v;  // byte-sized elements: plain store
    }
}

// Cached or streaming store chosen at compile time.
template <bool NT, class T>
__device__ __forceinline__ void sto(T* p, T v) {
    if constexpr (NT) st_nt(p, v);
    else *p = v;
}

// Block id -> work index such that each of the 8 XCDs (blocks are dispatched to
// them round-robin) runs one contiguous range of work items: neighbouring tiles
// then share that XCD's L2 and DRAM pages. A bijection on [0, n) for any n.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t t, uint32_t n) {
    const uint32_t q = n / 8, r = n % 8, x = t % 8, i = t / 8;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

/* ------------------------------ launch helpers ------------------------------ */

inline void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(ErrorKind::DeviceError, std::string(what) + " launch failed: " + hipGetErrorString(e));
}

inline uint32_t ceil_div_u32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace hip
}  // namespace rstsr
