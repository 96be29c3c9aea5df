// BLAS-style driver entry points over the MFMA GEMM: GEMMDriverAPI::driver_gemm
// (rstsr-blas-traits/src/blas3/gemm.rs:4-20), SYHEMMDriverAPI::driver_syhemm
// (blas3/syhemm.rs:4-19) and TRSMDriverAPI::driver_trsm (blas3/trsm.rs:4-19), with
// cblas flag values (FlagOrder C = 101 / F = 102, FlagTrans N = 111 / T = 112 /
// C = 113 / CN = 114, FlagUpLo U = 121 / L = 122, FlagDiag N = 131 / U = 132,
// FlagSide L = 141 / R = 142; rstsr-common/src/flags.rs:62-213).
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
enum : int { UPLO_U = 121, UPLO_L = 122, DIAG_N = 131, DIAG_U = 132, SIDE_L = 141, SIDE_R = 142 };

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

// complex-aware arithmetic of the triangular solve
template <class T>
__device__ __forceinline__ T t_mul(T a, T b) {
    if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value)
        return T{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    else return a * b;
}
template <class T>
__device__ __forceinline__ T t_sub(T a, T b) {
    if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) return T{a.re - b.re, a.im - b.im};
    else return a - b;
}
template <class T>
__device__ __forceinline__ T t_div(T a, T b) {
    if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) {
        // a conj(b) / |b|^2 with |b|^2 scaled by the larger part of b (no overflow for moderate b)
        const auto s = fmax(fabs(b.re), fabs(b.im));
        const auto br = b.re / s, bi = b.im / s, d = br * br + bi * bi;
        return T{(a.re * br + a.im * bi) / d / s, (a.im * br - a.re * bi) / d / s};
    } else {
        return a / b;
    }
}

constexpr int TRSM_NB = 64;  // rows of a diagonal block solved in LDS

// One diagonal block of the triangular solve, in place: X (kb x r, strides xrs / xcs) :=
// T^-1 X for the kb x kb triangle T (strides trs / tcs; lower or upper; unit: diagonal taken
// as one and not read). One thread per column of X (64 per block); T and the block's 64
// columns staged in LDS, each column by forward (lower) or backward substitution.
template <class T>
__global__ void __launch_bounds__(TRSM_NB) k_trsm_diag(const T* __restrict__ Tm, int64_t trs, int64_t tcs, int lower,
                                                      int unit, int kb, T* __restrict__ X, int64_t xrs, int64_t xcs,
                                                      int64_t r) {
    __shared__ T Ts[TRSM_NB][TRSM_NB + 1];
    __shared__ T Xs[TRSM_NB][TRSM_NB];
    const int j = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * TRSM_NB + j;
    for (int e = j; e < kb * kb; e += TRSM_NB) {
        const int i = e / kb, k = e - i * kb;
        // only the stored triangle is read, and a unit diagonal not at all
        const bool in = (lower ? k <= i : k >= i) && !(unit && k == i);
        Ts[i][k] = in ? Tm[(int64_t)i * trs + (int64_t)k * tcs] : T{};
    }
    for (int i = 0; i < kb; ++i) Xs[i][j] = col < r ? X[(int64_t)i * xrs + col * xcs] : T{};
    __syncthreads();
    if (col >= r) return;
    if (lower) {
        for (int i = 0; i < kb; ++i) {
            T acc = Xs[i][j];
            for (int k = 0; k < i; ++k) acc = t_sub(acc, t_mul(Ts[i][k], Xs[k][j]));
            Xs[i][j] = unit ? acc : t_div(acc, Ts[i][i]);
        }
    } else {
        for (int i = kb - 1; i >= 0; --i) {
            T acc = Xs[i][j];
            for (int k = i + 1; k < kb; ++k) acc = t_sub(acc, t_mul(Ts[i][k], Xs[k][j]));
            Xs[i][j] = unit ? acc : t_div(acc, Ts[i][i]);
        }
    }
    for (int i = 0; i < kb; ++i) X[(int64_t)i * xrs + col * xcs] = Xs[i][j];
}

// X (rows x cols, strides rs / cs) *= alpha, or := 0 when zero (BLAS: B is not read)
template <class T>
__global__ void __launch_bounds__(256) k_scale2d(T* __restrict__ X, uint32_t rows, uint32_t cols, int64_t rs,
                                                 int64_t cs, T alpha, int zero) {
    const uint64_t n = (uint64_t)rows * cols;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (uint64_t)gridDim.x * 256) {
        const uint32_t i = (uint32_t)(e / cols), j = (uint32_t)(e - (uint64_t)i * cols);
        T* p = X + (int64_t)i * rs + (int64_t)j * cs;
        *p = zero ? T{} : t_mul(alpha, *p);
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

// driver_trsm: op(A) X = alpha B (side L, A m x m) or X op(A) = alpha B (side R, A n x n), B (m x n)
// overwritten by X; A triangular (its `uplo` triangle read; diag U: unit diagonal, not read).
// Every case becomes T Y = alpha Z with T lower or upper: side R solves op(A)^T X^T = alpha B^T, and
// a transposition swaps T's strides and its triangle; a conjugated A (trans C or CN) is copied
// conjugated into scratch first. T Y = Z is solved recursively: halves of the rows, the
// trailing half updated by the GEMM (alpha -1, beta 1) with K the leading half's size, down to
// 64-row diagonal blocks solved in LDS (k_trsm_diag).
static void driver_trsm(rstsr_hip_ctx* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int order, int side, int uplo,
                        int transa, int diag, size_t m, size_t n, const void* alpha, const void* a, size_t lda, void* b,
                        size_t ldb) {
    if (order != ORD_C && order != ORD_F) RSTSR_RAISE(InvalidValue, "order must be 101 (C) or 102 (F)");
    if (side != SIDE_L && side != SIDE_R) RSTSR_RAISE(InvalidValue, "side must be 141 (L) or 142 (R)");
    if (uplo != UPLO_U && uplo != UPLO_L) RSTSR_RAISE(InvalidValue, "uplo must be 121 (U) or 122 (L)");
    if (diag != DIAG_N && diag != DIAG_U) RSTSR_RAISE(InvalidValue, "diag must be 131 (N) or 132 (U)");
    if (transa < TR_N || transa > TR_CN) RSTSR_RAISE(InvalidValue, "trans must be 111 (N), 112 (T), 113 (C) or 114 (CN)");
    dispatch_gemm_dtype(t, [](auto) {});  // BLAS types only
    const size_t es = dtype_size(t);
    const size_t s = side == SIDE_L ? m : n;  // order of A
    check_ld(order, s, s, lda, "a");
    check_ld(order, m, n, ldb, "b");
    if (m == 0 || n == 0) return;
    int64_t ars, acs, brs, bcs;
    stored_strides(order, (int64_t)lda, ars, acs);
    stored_strides(order, (int64_t)ldb, brs, bcs);
    check_mat(a, s, s, ars, acs, es, "a");
    check_mat(b, m, n, brs, bcs, es, "b");
    // T = op(A) (side L) or op(A)^T (side R); Y = X (rows s = m) or X^T (rows s = n)
    const bool trans_op = transa == TR_T || transa == TR_C;
    const bool tr = trans_op != (side == SIDE_R);
    const bool conj = is_cplx(t) && (transa == TR_C || transa == TR_CN);
    int64_t trs = tr ? acs : ars, tcs = tr ? ars : acs;
    const bool lower = (uplo == UPLO_L) != tr;
    const int64_t yrs = side == SIDE_L ? brs : bcs, ycs = side == SIDE_L ? bcs : brs;
    const int64_t r = (int64_t)(side == SIDE_L ? n : m);  // columns of Y
    std::lock_guard<std::mutex> g(ctx->mu);
    dispatch_gemm_dtype(t, [&](auto tag) {
        using T = decltype(tag);
        const T* Tm = (const T*)a;
        if (conj) {
            T* d = (T*)aux_buffer(ctx, s * s * es + 256);
            hipLaunchKernelGGL(k_conj_copy<T>, dim3(grid_for((uint64_t)s * s)), dim3(256), 0, ctx->stream, d, Tm,
                               (uint32_t)s, (uint32_t)s, trs, tcs);
            check_launch("conjugate copy");
            Tm = d;
            trs = (int64_t)s;
            tcs = 1;
        }
        T* Y = (T*)b;
        T al;
        std::memcpy(&al, alpha, sizeof(T));
        T one{}, minus_one{};
        if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) {
            one.re = 1;
            minus_one.re = -1;
        } else {
            one = 1;
            minus_one = -1;
        }
        const bool is_one = std::memcmp(&al, &one, sizeof(T)) == 0;
        const T zero{};
        const bool is_zero = std::memcmp(&al, &zero, sizeof(T)) == 0;
        if (!is_one) {
            hipLaunchKernelGGL(k_scale2d<T>, dim3(grid_for((uint64_t)s * r)), dim3(256), 0, ctx->stream, Y,
                               (uint32_t)s, (uint32_t)r, yrs, ycs, al, is_zero ? 1 : 0);
            check_launch("trsm scale");
            if (is_zero) return;
        }
        // rows [lo, hi) of the solve
        auto solve = [&](auto&& self, int64_t lo, int64_t hi) -> void {
            const int64_t len = hi - lo;
            if (len <= TRSM_NB) {
                hipLaunchKernelGGL(k_trsm_diag<T>, dim3((uint32_t)((r + TRSM_NB - 1) / TRSM_NB)), dim3(TRSM_NB), 0,
                                   ctx->stream, Tm + lo * trs + lo * tcs, trs, tcs, lower ? 1 : 0,
                                   diag == DIAG_U ? 1 : 0, (int)len, Y + lo * yrs, yrs, ycs, r);
                check_launch("trsm diagonal block");
                return;
            }
            const int64_t mid = lo + ((len / 2 + TRSM_NB - 1) / TRSM_NB) * TRSM_NB;
            // first the half the other depends on, then its update of the other half, then that half
            const int64_t f0 = lower ? lo : mid, f1 = lower ? mid : hi;  // solved first
            const int64_t s0 = lower ? mid : lo, s1 = lower ? hi : mid;  // updated, solved second
            self(self, f0, f1);
            GemmProblem P{};
            P.M = s1 - s0;
            P.N = r;
            P.K = f1 - f0;
            P.bd.nbd = 0;
            P.nbatch = 1;
            P.a = Tm + s0 * trs + f0 * tcs;
            P.rsa = trs;
            P.csa = tcs;
            P.b = Y + f0 * yrs;
            P.rsb = yrs;
            P.csb = ycs;
            P.c = Y + s0 * yrs;
            P.rsc = yrs;
            P.csc = ycs;
            run_gemm(ctx, t, mode, P, &minus_one, &one);
            self(self, s0, s1);
        };
        solve(solve, 0, (int64_t)s);
    });
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

int rstsr_hip_driver_trsm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int order, int side,
                          int uplo, int transa, int diag, size_t m, size_t n, const void* alpha, const void* a,
                          size_t lda, void* b, size_t ldb) {
    return guarded([&] {
        bind(ctx);
        driver_trsm(ctx, t, mode, order, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
    });
}

}  // extern "C"
