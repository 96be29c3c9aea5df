// BLAS-style driver entry points over the MFMA GEMM: GEMMDriverAPI::driver_gemm
// (rstsr-blas-traits/src/blas3/gemm.rs:4-20) and SYHEMMDriverAPI::driver_syhemm
// (blas3/syhemm.rs:4-19), with cblas flag values (FlagOrder C = 101 / F = 102,
// FlagTrans N = 111 / T = 112 / C = 113 / CN = 114, FlagUpLo U = 121 / L = 122,
// FlagSide L = 141 / R = 142; rstsr-common/src/flags.rs:62-202).
//
// Storage order and transposition become (row stride, column stride) pairs of
// op(A), op(B) and C for the strided GEMM, so no operand is copied, except:
//  - a conjugated complex operand (trans C or CN), copied conjugated into a
//    dense scratch matrix first (the GEMM tiles have no conjugate path);
//  - the symmetric / Hermitian matrix of syhemm, expanded from its stored
//    triangle into a dense scratch matrix (BLAS reads only that triangle, and
//    only the real part of a Hermitian diagonal).
// Scratch lives in the context's auxiliary buffer, which a single GEMM does not
// use.
#include "common.hpp"
#include "registry.hpp"
#include "gemm.hpp"

#include <cstring>
#include <type_traits>

namespace rstsr {
namespace hip {

void run_gemm(rstsr_hip_ctx* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, const GemmProblem& P0, const void* alpha,
              const void* beta);

namespace {

enum : int { ORD_C = 101, ORD_F = 102, TR_N = 111, TR_T = 112, TR_C = 113, TR_CN = 114 };
enum : int { UPLO_U = 121, UPLO_L = 122, SIDE_L = 141, SIDE_R = 142 };

template <class T>
__device__ __forceinline__ T conj_v(T x) {
    if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) return T{x.re, -x.im};
    else return x;
}

// dst (rows x cols, row-major, dense) = conj(src) read with strides (rs, cs)
template <class T>
__global__ void __launch_bounds__(256) k_conj_copy(T* __restrict__ dst, const T* __restrict__ src, uint32_t rows,
                                                   uint32_t cols, int64_t rs, int64_t cs) {
    const uint64_t n = (uint64_t)rows * cols;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (uint64_t)gridDim.x * 256) {
        const uint32_t i = (uint32_t)(e / cols), j = (uint32_t)(e - (uint64_t)i * cols);
        dst[e] = conj_v(src[(int64_t)i * rs + (int64_t)j * cs]);
    }
}

// dst (n x n, row-major) = the full symmetric / Hermitian matrix whose `lower`
// (else upper) triangle is stored at src with strides (rs, cs); a Hermitian
// diagonal keeps its real part only.
template <class T, bool HERMI>
__global__ void __launch_bounds__(256) k_sym_expand(T* __restrict__ dst, const T* __restrict__ src, uint32_t n,
                                                    int64_t rs, int64_t cs, int lower) {
    const uint64_t total = (uint64_t)n * n;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256) {
        const uint32_t i = (uint32_t)(e / n), j = (uint32_t)(e - (uint64_t)i * n);
        const bool stored = lower ? (i >= j) : (i <= j);
        T v = stored ? src[(int64_t)i * rs + (int64_t)j * cs] : src[(int64_t)j * rs + (int64_t)i * cs];
        if constexpr (HERMI) {
            if (!stored) v = conj_v(v);
            if (i == j) {
                if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) v.im = 0;
            }
        }
        dst[e] = v;
    }
}

uint32_t grid_for(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), 65536);
}

bool is_cplx(rstsr_dtype_t t) { return t == RSTSR_C64 || t == RSTSR_C128; }

template <class F>
void dispatch_gemm_dtype(rstsr_dtype_t t, F&& f) {
    switch (t) {
        case RSTSR_F32: f(float{}); break;
        case RSTSR_F64: f(double{}); break;
        case RSTSR_C64: f(c64{}); break;
        case RSTSR_C128: f(c128{}); break;
        default: RSTSR_RAISE(UnImplemented, "the BLAS drivers take f32, f64, c64 and c128");
    }
}

// Strides of a stored (rows x cols) matrix with leading dimension ld.
void stored_strides(int order, int64_t ld, int64_t& rs, int64_t& cs) {
    if (order == ORD_C) {
        rs = ld;
        cs = 1;
    } else {
        rs = 1;
        cs = ld;
    }
}

void check_ld(int order, size_t rows, size_t cols, size_t ld, const char* nm) {
    const size_t need = order == ORD_C ? cols : rows;
    if (ld < std::max<size_t>(need, 1))
        RSTSR_RAISE(InvalidValue, std::string("leading dimension of ") + nm + " is smaller than the stored extent");
}

void check_mat(const void* p, size_t rows, size_t cols, int64_t rs, int64_t cs, size_t es, const char* nm) {
    size_t s[2] = {rows, cols};
    int64_t t[2] = {rs, cs};
    check_extent_rel(p, 2, s, t, es, nm);
}

}  // namespace

// driver_gemm: C (m x n) = alpha op(A) op(B) + beta C.
static void driver_gemm(rstsr_hip_ctx* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int order, int transa,
                        int transb, size_t m, size_t n, size_t k, const void* alpha, const void* a, size_t lda,
                        const void* b, size_t ldb, const void* beta, void* c, size_t ldc) {
    if (order != ORD_C && order != ORD_F) RSTSR_RAISE(InvalidValue, "order must be 101 (C) or 102 (F)");
    dispatch_gemm_dtype(t, [](auto) {});  // BLAS types only
    for (int tr : {transa, transb})
        if (tr < TR_N || tr > TR_CN) RSTSR_RAISE(InvalidValue, "trans must be 111 (N), 112 (T), 113 (C) or 114 (CN)");
    const size_t es = dtype_size(t);
    const bool ta = transa == TR_T || transa == TR_C, tb = transb == TR_T || transb == TR_C;
    const bool ca = is_cplx(t) && (transa == TR_C || transa == TR_CN);
    const bool cb = is_cplx(t) && (transb == TR_C || transb == TR_CN);
    // stored shapes: A is (m x k) or (k x m) when transposed; B (k x n) or (n x k)
    const size_t ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
    check_ld(order, ar, ac, lda, "a");
    check_ld(order, br, bc, ldb, "b");
    check_ld(order, m, n, ldc, "c");
    int64_t srs, scs;
    GemmProblem P{};
    P.M = (int64_t)m;
    P.N = (int64_t)n;
    P.K = (int64_t)k;
    P.bd.nbd = 0;
    P.nbatch = 1;
    stored_strides(order, (int64_t)lda, srs, scs);
    P.a = a;
    P.rsa = ta ? scs : srs;
    P.csa = ta ? srs : scs;
    stored_strides(order, (int64_t)ldb, srs, scs);
    P.b = b;
    P.rsb = tb ? scs : srs;
    P.csb = tb ? srs : scs;
    stored_strides(order, (int64_t)ldc, srs, scs);
    P.c = c;
    P.rsc = srs;
    P.csc = scs;
    if (m == 0 || n == 0) return;
    if (k > 0) {
        check_mat(P.a, m, k, P.rsa, P.csa, es, "a");
        check_mat(P.b, k, n, P.rsb, P.csb, es, "b");
    }
    check_mat(P.c, m, n, P.rsc, P.csc, es, "c");
    std::lock_guard<std::mutex> g(ctx->mu);
    if ((ca || cb) && k > 0) {
        const size_t na = ca ? m * k : 0, nb = cb ? k * n : 0;
        char* buf = (char*)aux_buffer(ctx, (na + nb) * es + 256);
        dispatch_gemm_dtype(t, [&](auto tag) {
            using T = decltype(tag);
            if (ca) {
                T* d = (T*)buf;
                hipLaunchKernelGGL(k_conj_copy<T>, dim3(grid_for(na)), dim3(256), 0, ctx->stream, d, (const T*)P.a,
                                   (uint32_t)m, (uint32_t)k, P.rsa, P.csa);
                check_launch("conjugate copy");
                P.a = d;
                P.rsa = (int64_t)k;
                P.csa = 1;
            }
            if (cb) {
                T* d = (T*)(buf + ((na * es + 255) & ~(size_t)255));
                hipLaunchKernelGGL(k_conj_copy<T>, dim3(grid_for(nb)), dim3(256), 0, ctx->stream, d, (const T*)P.b,
                                   (uint32_t)k, (uint32_t)n, P.rsb, P.csb);
                check_launch("conjugate copy");
                P.b = d;
                P.rsb = (int64_t)n;
                P.csb = 1;
            }
        });
    }
    run_gemm(ctx, t, mode, P, alpha, beta);
}

// driver_syhemm: C (m x n) = alpha A B + beta C (side L, A m x m) or alpha B A + beta C
// (side R, A n x n); A symmetric (hermi 0) or Hermitian (hermi 1), its `uplo` triangle stored.
static void driver_syhemm(rstsr_hip_ctx* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int hermi, int order, int side,
                          int uplo, size_t m, size_t n, const void* alpha, const void* a, size_t lda, const void* b,
                          size_t ldb, const void* beta, void* c, size_t ldc) {
    if (order != ORD_C && order != ORD_F) RSTSR_RAISE(InvalidValue, "order must be 101 (C) or 102 (F)");
    if (side != SIDE_L && side != SIDE_R) RSTSR_RAISE(InvalidValue, "side must be 141 (L) or 142 (R)");
    if (uplo != UPLO_U && uplo != UPLO_L) RSTSR_RAISE(InvalidValue, "uplo must be 121 (U) or 122 (L)");
    dispatch_gemm_dtype(t, [](auto) {});  // BLAS types only
    const size_t es = dtype_size(t);
    const size_t na = side == SIDE_L ? m : n;
    check_ld(order, na, na, lda, "a");
    check_ld(order, m, n, ldb, "b");
    check_ld(order, m, n, ldc, "c");
    if (m == 0 || n == 0) return;
    int64_t ars, acs, brs, bcs, crs, ccs;
    stored_strides(order, (int64_t)lda, ars, acs);
    stored_strides(order, (int64_t)ldb, brs, bcs);
    stored_strides(order, (int64_t)ldc, crs, ccs);
    check_mat(a, na, na, ars, acs, es, "a");
    check_mat(b, m, n, brs, bcs, es, "b");
    check_mat(c, m, n, crs, ccs, es, "c");
    std::lock_guard<std::mutex> g(ctx->mu);
    void* full = aux_buffer(ctx, na * na * es + 256);
    dispatch_gemm_dtype(t, [&](auto tag) {
        using T = decltype(tag);
        auto go = [&](auto h) {
            hipLaunchKernelGGL((k_sym_expand<T, decltype(h)::value>), dim3(grid_for((uint64_t)na * na)), dim3(256), 0,
                               ctx->stream, (T*)full, (const T*)a, (uint32_t)na, ars, acs, uplo == UPLO_L ? 1 : 0);
        };
        if (hermi && is_cplx(t)) go(std::true_type{});
        else go(std::false_type{});
        check_launch("symmetric expand");
    });
    GemmProblem P{};
    P.M = (int64_t)m;
    P.N = (int64_t)n;
    P.K = (int64_t)na;
    P.bd.nbd = 0;
    P.nbatch = 1;
    if (side == SIDE_L) {
        P.a = full;
        P.rsa = (int64_t)na;
        P.csa = 1;
        P.b = b;
        P.rsb = brs;
        P.csb = bcs;
    } else {
        P.a = b;
        P.rsa = brs;
        P.csa = bcs;
        P.b = full;
        P.rsb = (int64_t)na;
        P.csb = 1;
    }
    P.c = c;
    P.rsc = crs;
    P.csc = ccs;
    run_gemm(ctx, t, mode, P, alpha, beta);
}

}  // namespace hip
}  // namespace rstsr

using namespace rstsr;
using namespace rstsr::hip;

extern "C" {

int rstsr_hip_driver_gemm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int order, int transa,
                          int transb, size_t m, size_t n, size_t k, const void* alpha, const void* a, size_t lda,
                          const void* b, size_t ldb, const void* beta, void* c, size_t ldc) {
    return guarded([&] {
        bind(ctx);
        driver_gemm(ctx, t, mode, order, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    });
}

int rstsr_hip_driver_syhemm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int hermi, int order,
                            int side, int uplo, size_t m, size_t n, const void* alpha, const void* a, size_t lda,
                            const void* b, size_t ldb, const void* beta, void* c, size_t ldc) {
    return guarded([&] {
        bind(ctx);
        driver_syhemm(ctx, t, mode, hermi, order, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
    });
}

}  // extern "C"
