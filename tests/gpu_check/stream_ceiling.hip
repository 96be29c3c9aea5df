// Stream ceilings of one GPU's HBM, for the benchmark's roofline record (test infrastructure,
// not part of the library): a contiguous copy (16-byte lanes, one per thread, a thread per
// lane of the whole array), a read-only pass and a write-only pass, with ordinary or
// nontemporal accesses, each timed with events over repeated launches on buffers this file
// allocates and frees.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

typedef float f4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ void __launch_bounds__(256) k_copy(const f4* __restrict__ a, f4* __restrict__ c, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if constexpr (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), c + i);
    else c[i] = a[i];
}
template <bool NT>
__global__ void __launch_bounds__(256) k_read(const f4* __restrict__ a, f4* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    f4 v = f4{0, 0, 0, 0};
    if (i < n) v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    // a value no input here produces keeps the loads without a store per lane
    if (v.x == 1.5e38f && v.y == -1.5e38f) out[blockIdx.x] = v;
}
template <bool NT>
__global__ void __launch_bounds__(256) k_write(f4* __restrict__ c, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const f4 v = f4{1.0f, 2.0f, 3.0f, 4.0f};
    if constexpr (NT) __builtin_nontemporal_store(v, c + i);
    else c[i] = v;
}

// kind 0: copy, 1: read, 2: write; nt: nontemporal loads / stores. Over `bytes` (a multiple of
// 4096); best of `reps` launches after two warm-ups, in ms, to *best_ms. Returns a HIP error code
// (0 on success).
extern "C" int stream_ceiling(int kind, int nt, size_t bytes, int reps, float* best_ms) {
    const size_t n = bytes / 16;
    f4 *a = nullptr, *c = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&c, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const unsigned grid = (unsigned)((n + 255) / 256);
    std::vector<float> t;
    for (int r = 0; r < reps + 2 && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, 0);
        if (kind == 0) hipLaunchKernelGGL(nt ? k_copy<true> : k_copy<false>, dim3(grid), dim3(256), 0, 0, a, c, n);
        else if (kind == 1) hipLaunchKernelGGL(nt ? k_read<true> : k_read<false>, dim3(grid), dim3(256), 0, 0, a, c, n);
        else hipLaunchKernelGGL(nt ? k_write<true> : k_write<false>, dim3(grid), dim3(256), 0, 0, c, n);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (r >= 2) t.push_back(ms);
    }
    if (e == hipSuccess && !t.empty()) *best_ms = *std::min_element(t.begin(), t.end());
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (c) (void)hipFree(c);
    return (int)e;
}
