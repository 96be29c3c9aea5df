// GEMM on gfx950: FP64 MFMA DGEMM / ZGEMM, SIMT fallback for f32 and integers,
// and the matmul entry point with rstsr's batch-broadcast rules.
//
// Replaces DeviceMatMulAPI for DeviceFaer (rstsr-core/src/device_faer/matmul.rs:
// 19-239 -> faer) and the naive kernels of rstsr-native-impl/src/cpu_serial/
// matmul_naive.rs:78-122 (non-float types). Operand layouts map to (row stride,
// column stride) pairs directly, so no packing copy is made for any strided
// operand (the OpenBLAS device packs non-BLAS layouts, crates-device/
// rstsr-openblas/src/matmul_impl.rs:92-113; here the LDS tile load absorbs it).
//
// DGEMM tiling: 256 threads = 4 waves in a 2x2 grid; each wave owns a
// (16*WM) x (16*WN) patch of C built from v_mfma_f64_16x16x4_f64 tiles. A and
// B stage through LDS in BK=16 slices with a one-slice register prefetch, so
// global loads of slice k+1 overlap the MFMAs of slice k. Global loads walk the
// operand's unit-stride axis (template flags), so both row- and column-major
// operands read coalesced. ZGEMM uses the same kernel with separate real and
// imaginary LDS planes and four real MFMAs per complex tile (4M form).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>

#include "common.hpp"
#include "registry.hpp"
#include "gemm.hpp"

#include <algorithm>
#include <functional>

namespace rstsr {
namespace hip {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int GBK = 16;  // K slice held in LDS
constexpr int GPAD = 4;  // LDS row padding (doubles)

template <class T>
struct GemmArgs {
    int64_t M, N, K;
    const T* a;
    int64_t rsa, csa;
    const T* b;
    int64_t rsb, csb;
    T* c;
    int64_t rsc, csc;
    T alpha, beta;
    int beta_zero;
    uint32_t tiles_m, tiles_n;
    uint32_t batch0;  // first batch index of this launch
    int tri;          // symmetric product: only the tiles with tm >= tn
    BatchDesc bd;
};

template <class T>
struct CplxOf : std::false_type {};
template <>
struct CplxOf<c128> : std::true_type {};

__device__ __forceinline__ double re_of(double x) { return x; }
__device__ __forceinline__ double im_of(double) { return 0.0; }
__device__ __forceinline__ double re_of(c128 x) { return x.re; }
__device__ __forceinline__ double im_of(c128 x) { return x.im; }


// Operands of one C tile's problem.
template <class T>
struct TileArgs {
    int64_t M, N, K;
    const T* A;
    int64_t rsa, csa;
    const T* B;
    int64_t rsb, csb;
    T* C;
    int64_t rsc, csc;
};

// Block tile configuration: WVM x WVN waves, each owning FM x FN 16x16 output
// tiles, so a block covers BM x BN = (16 WVM FM) x (16 WVN FN).
// OCC: waves per SIMD the register budget is set for (the launch bound's second argument).
template <int WVM_, int WVN_, int FM_, int FN_, int OCC_ = 2>
struct MCfg {
    static constexpr int WVM = WVM_, WVN = WVN_, FM = FM_, FN = FN_;
    static constexpr int NT = 64 * WVM * WVN, BM = 16 * WVM * FM, BN = 16 * WVN * FN;
    static constexpr int OCC = OCC_;
};

// LDS planes of one block (re/im planes for complex). Real types double-buffer
// the K slices (one barrier per slice); complex tiles keep one buffer so that
// two blocks still fit in a CU's LDS.
template <class T, class CF>
struct MfmaSmem {
    static constexpr int NP = CplxOf<T>::value ? 2 : 1;
    static constexpr int NB = CplxOf<T>::value ? 1 : 2;
    double As[NB][NP][GBK][CF::BM + GPAD];
    double Bs[NB][NP][GBK][CF::BN + GPAD];
};

// One BM x BN C tile: C = alpha * A B + beta * C (beta_zero: C not read).
// A_MC: A's unit-stride axis is m (else k); B_KC: B's is k (else n). Each
// thread loads VW consecutive elements along that axis per slot (VEC: two
// doubles in one 16-byte load, which needs a +1 stride and 16-byte alignment).
template <class T, class CF, bool A_MC, bool B_KC, bool VEC>
__device__ __forceinline__ void mfma_tile(const TileArgs<T>& g, uint32_t tm, uint32_t tn, T alpha, T beta,
                                          int beta_zero, MfmaSmem<T, CF>& sm) {
    constexpr bool CX = CplxOf<T>::value;
    constexpr int NP = MfmaSmem<T, CF>::NP;
    constexpr int NT = CF::NT, BM = CF::BM, BN = CF::BN, FM = CF::FM, FN = CF::FN;
    constexpr int VW = VEC ? 2 : 1;
    static_assert(!VEC || !CX, "vector loads are for real operands");
    // A slots: along m (A_MC) rows of BM/VW threads, KRA k-rows per pass; else
    // along k rows of GBK/VW threads, RPA m-rows per pass. B likewise.
    constexpr int TPA = A_MC ? BM / VW : GBK / VW, PPA = NT / TPA, LA = (A_MC ? GBK : BM) / PPA;
    constexpr int TPB = B_KC ? GBK / VW : BN / VW, PPB = NT / TPB, LB = (B_KC ? BN : GBK) / PPB;
    static_assert(NT % TPA == 0 && NT % TPB == 0 && LA >= 1 && LB >= 1 && (A_MC ? GBK : BM) % PPA == 0 &&
                      (B_KC ? BN : GBK) % PPB == 0,
                  "loader shape");
    auto& As = sm.As;
    auto& Bs = sm.Bs;
    const T* A = g.A;
    const T* B = g.B;
    T* C = g.C;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // wave -> (row group, column group): waves w and w + 4 share a SIMD, so the column group is
    // rotated by w / 4 to give each SIMD waves of different column groups as well as rows (an edge
    // tile's padding then idles every SIMD alike)
    const int wm = wave / CF::WVN, wn = (wave % CF::WVN + wave / 4) % CF::WVN;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;

    dbl4 acc[NP][FM][FN];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[p][i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    // slot i of this thread covers A(mm0 + i*dam + v*[A_MC], ka0 + i*dak + v*[!A_MC]), v < VW
    constexpr int dam = A_MC ? 0 : PPA, dak = A_MC ? PPA : 0;
    constexpr int dbk = B_KC ? 0 : PPB, dbn = B_KC ? PPB : 0;
    const int mm0 = A_MC ? (tid % TPA) * VW : tid / TPA, ka0 = A_MC ? tid / TPA : (tid % TPA) * VW;
    const int kb0 = B_KC ? (tid % TPB) * VW : tid / TPB, nn0 = B_KC ? tid / TPB : (tid % TPB) * VW;
    const int64_t sa_i = dam * g.rsa + dak * g.csa, sb_i = dbk * g.rsb + dbn * g.csb;
    const int64_t sa_v = A_MC ? g.rsa : g.csa, sb_v = B_KC ? g.rsb : g.csb;
    const T* pa = A + (m0 + mm0) * g.rsa + ka0 * g.csa;
    const T* pb = B + kb0 * g.rsb + (n0 + nn0) * g.csb;
    // interior tiles skip the bounds tests except in a partial last K slice
    const bool interior = m0 + BM <= g.M && n0 + BN <= g.N;

    T ra[LA][VW], rb[LB][VW];
    auto load_slice = [&](int64_t k0) {
        const int64_t oa = k0 * g.csa, ob = k0 * g.rsb;
        if (interior && k0 + GBK <= g.K) {
#pragma unroll
            for (int i = 0; i < LA; ++i) {
                const T* p = pa + oa + i * sa_i;
                if constexpr (VEC) {
                    const dbl2 w = *(const dbl2*)p;
                    ra[i][0] = w[0];
                    ra[i][VW - 1] = w[1];
                } else {
                    ra[i][0] = *p;
                }
            }
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const T* p = pb + ob + i * sb_i;
                if constexpr (VEC) {
                    const dbl2 w = *(const dbl2*)p;
                    rb[i][0] = w[0];
                    rb[i][VW - 1] = w[1];
                } else {
                    rb[i][0] = *p;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < LA; ++i)
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    const bool ok = m0 + mm0 + i * dam + (A_MC ? v : 0) < g.M && k0 + ka0 + i * dak + (A_MC ? 0 : v) < g.K;
                    ra[i][v] = ok ? pa[oa + i * sa_i + v * sa_v] : T{};
                }
#pragma unroll
            for (int i = 0; i < LB; ++i)
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    const bool ok = k0 + kb0 + i * dbk + (B_KC ? v : 0) < g.K && n0 + nn0 + i * dbn + (B_KC ? 0 : v) < g.N;
                    rb[i][v] = ok ? pb[ob + i * sb_i + v * sb_v] : T{};
                }
        }
    };
    auto store_slice = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LA; ++i)
#pragma unroll
            for (int v = 0; v < VW; ++v) {
                const int kk = ka0 + i * dak + (A_MC ? 0 : v), mm = mm0 + i * dam + (A_MC ? v : 0);
                As[buf][0][kk][mm] = re_of(ra[i][v]);
                if constexpr (CX) As[buf][NP - 1][kk][mm] = im_of(ra[i][v]);
            }
#pragma unroll
        for (int i = 0; i < LB; ++i)
#pragma unroll
            for (int v = 0; v < VW; ++v) {
                const int kk = kb0 + i * dbk + (B_KC ? v : 0), nn = nn0 + i * dbn + (B_KC ? 0 : v);
                Bs[buf][0][kk][nn] = re_of(rb[i][v]);
                if constexpr (CX) Bs[buf][NP - 1][kk][nn] = im_of(rb[i][v]);
            }
    };

    // 16 x 16 fragments of this wave that hold any row < M and column < N (the wave index is
    // uniform, read through a scalar register so the tests below are scalar branches): edge
    // tiles skip the MFMAs of fragments wholly in the padding
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int64_t rows_left = g.M - (m0 + (wv / CF::WVN) * 16 * FM),
                  cols_left = g.N - (n0 + ((wv % CF::WVN + wv / 4) % CF::WVN) * 16 * FN);
    const int fmv = rows_left <= 0 ? 0 : (int)std::min<int64_t>(FM, (rows_left + 15) / 16);
    const int fnv = cols_left <= 0 ? 0 : (int)std::min<int64_t>(FN, (cols_left + 15) / 16);

    load_slice(0);
    store_slice(0);
    __syncthreads();
    const int fr = lane & 15, fk = lane >> 4;
    int buf = 0;
    for (int64_t k0 = 0; k0 < g.K; k0 += GBK) {
        const bool more = k0 + GBK < g.K;
        if (more) load_slice(k0 + GBK);
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            double fa[NP][FM], fb[NP][FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                fa[0][i] = As[buf][0][kk + fk][wm * 16 * FM + i * 16 + fr];
                if constexpr (CX) fa[NP - 1][i] = As[buf][NP - 1][kk + fk][wm * 16 * FM + i * 16 + fr];
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                fb[0][j] = Bs[buf][0][kk + fk][wn * 16 * FN + j * 16 + fr];
                if constexpr (CX) fb[NP - 1][j] = Bs[buf][NP - 1][kk + fk][wn * 16 * FN + j * 16 + fr];
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    if (i >= fmv || j >= fnv) continue;
                    if constexpr (!CX) {
                        acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[0][i], fb[0][j], acc[0][i][j], 0, 0, 0);
                    } else {
                        // re += ar*br - ai*bi ; im += ar*bi + ai*br
                        acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[0][i], fb[0][j], acc[0][i][j], 0, 0, 0);
                        acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-fa[1][i], fb[1][j], acc[0][i][j], 0, 0, 0);
                        acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[0][i], fb[1][j], acc[1][i][j], 0, 0, 0);
                        acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[1][i], fb[0][j], acc[1][i][j], 0, 0, 0);
                    }
                }
        }
        if constexpr (MfmaSmem<T, CF>::NB == 2) {
            // the other buffer was last read before the previous barrier
            if (more) store_slice(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        } else {
            __syncthreads();
            if (more) {
                store_slice(0);
                __syncthreads();
            }
        }
    }
    // epilogue: v_mfma_f64_16x16x4f64 leaves register r of lane l at row 4*r + l/16,
    // column l%16 of its 16x16 tile
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t gm = m0 + wm * 16 * FM + i * 16 + 4 * r + (lane >> 4);
                int64_t gn = n0 + wn * 16 * FN + j * 16 + (lane & 15);
                if (gm < g.M && gn < g.N) {
                    T* cp = C + gm * g.rsc + gn * g.csc;
                    if constexpr (!CX) {
                        double v = alpha * acc[0][i][j][r];
                        if (!beta_zero) v += beta * (*cp);
                        *cp = v;
                    } else {
                        double xr = acc[0][i][j][r], xi = acc[1][i][j][r];
                        c128 v{alpha.re * xr - alpha.im * xi, alpha.re * xi + alpha.im * xr};
                        if (!beta_zero) {
                            c128 o = *cp;
                            v.re += beta.re * o.re - beta.im * o.im;
                            v.im += beta.re * o.im + beta.im * o.re;
                        }
                        *cp = v;
                    }
                }
            }
}

// Index t of the lower triangle (diag = 0: i >= j, t = i(i+1)/2 + j; diag = 1:
// i > j, t = i(i-1)/2 + j) -> (i, j). The floating root is corrected exactly.
__device__ __forceinline__ void tri_coords(uint32_t t, int diag, uint32_t& i, uint32_t& j) {
    auto first = [&](uint64_t r) { return diag ? r * (r - 1) / 2 : r * (r + 1) / 2; };
    uint64_t r = (uint64_t)((diag ? 1.0 + sqrt(1.0 + 8.0 * t) : sqrt(1.0 + 8.0 * t) - 1.0) * 0.5);
    while (r > 0 && first(r) > t) --r;
    while (first(r + 1) <= t) ++r;
    i = (uint32_t)r;
    j = (uint32_t)(t - first(r));
}

// Symmetric product (A = B^T, same buffer): copy the transpose of every 64 x 64
// tile on and below the diagonal onto its mirror, through LDS so both the reads
// and the writes run along C's rows; only elements strictly above the diagonal
// are written, so C comes out exactly symmetric. Batch from blockIdx.y.
template <class T>
__global__ void __launch_bounds__(256) k_mirror(GemmArgs<T> g) {
    __shared__ T tile[64][65];
    uint32_t ti, tj;
    tri_coords(blockIdx.x, 0, ti, tj);
    int64_t oa, ob, oc;
    batch_offsets(g.bd, g.batch0 + blockIdx.y, oa, ob, oc);
    T* C = g.c + oc;
    const int tid = threadIdx.x, c = tid & 63, r0 = tid >> 6;
    const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int r = r0 + 4 * k;
        if (i0 + r < g.M && j0 + c < g.N) tile[r][c] = C[(i0 + r) * g.rsc + (j0 + c) * g.csc];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int r = r0 + 4 * k;  // row j0 + r of the upper tile, column i0 + c
        if (j0 + r < g.N && i0 + c < g.M && i0 + c > j0 + r) C[(j0 + r) * g.rsc + (i0 + c) * g.csc] = tile[c][r];
    }
}

// Strided-batched GEMM: one C tile per block, batch index from blockIdx.y.
// run_if (nullable): the whole grid exits unless *run_if != 0 (device-side
// selection between alternative paths without a host round trip).
template <class T, class CF, bool A_MC, bool B_KC, bool VEC>
__global__ void __launch_bounds__(CF::NT, CF::OCC) k_gemm_mfma(GemmArgs<T> g, const int* run_if) {
    if (run_if && *run_if == 0) return;
    __shared__ MfmaSmem<T, CF> sm;
    uint32_t tm, tn;
    if (g.tri) tri_coords(xcd_remap8(blockIdx.x, g.tiles_m * (g.tiles_m + 1) / 2), 0, tm, tn);
    else tile_coords<8>(blockIdx.x, g.tiles_m, g.tiles_n, tm, tn);
    int64_t boa, bob, boc;
    batch_offsets(g.bd, g.batch0 + blockIdx.y, boa, bob, boc);
    TileArgs<T> t{g.M, g.N, g.K, g.a + boa, g.rsa, g.csa, g.b + bob, g.rsb, g.csb, g.c + boc, g.rsc, g.csc};
    mfma_tile<T, CF, A_MC, B_KC, VEC>(t, tm, tn, g.alpha, g.beta, g.beta_zero, sm);
}

// Grouped (uneven) GEMMs in one persistent launch. Problem g owns tiles
// [prefix[g], prefix[g+1]); blocks stride over all tiles. flags bit0: A
// m-contiguous, bit1: B k-contiguous, bit2: 16-byte loads (selects the loader).
template <class T>
struct GroupProblem {
    TileArgs<T> t;
    uint32_t tiles_n;
    uint32_t flags;
};

// VM: bit f set for each loader form f (GroupProblem::flags) the launch holds; a kernel with fewer
// forms holds fewer inlined tile loops, so it needs fewer registers
template <class T, class CF, uint32_t VM = 31>
__global__ void __launch_bounds__(CF::NT, CF::OCC) k_gemm_grouped(const GroupProblem<T>* __restrict__ probs, int count,
                                                                  const uint32_t* __restrict__ prefix, uint32_t total,
                                                                  T alpha, T beta, int beta_zero,
                                                                  const int* __restrict__ run_if,
                                                                  uint32_t* __restrict__ queue) {
    if (run_if && *run_if == 0) return;
    __shared__ MfmaSmem<T, CF> sm;
    __shared__ uint32_t next_tile;
    constexpr bool RV = !CplxOf<T>::value;  // vector loads exist for real operands
    // tiles handed out in order: block b starts with tile b, then takes the next unclaimed one
    // from the queue counter (zero at launch), claimed while its current tile computes
    for (uint32_t tile = blockIdx.x; tile < total;) {
        uint32_t claimed = 0;
        if (threadIdx.x == 0) claimed = atomicAdd(queue, 1u) + gridDim.x;
        // problem of this tile: last g with prefix[g] <= tile
        int lo = 0, hi = count - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= tile) lo = mid;
            else hi = mid - 1;
        }
        const GroupProblem<T> gp = probs[lo];
        const uint32_t local = tile - prefix[lo];
        const uint32_t tm = local / gp.tiles_n, tn = local - (local / gp.tiles_n) * gp.tiles_n;
        switch (gp.flags & 7) {
            case 3:
                if constexpr ((VM & 8) != 0) mfma_tile<T, CF, true, true, false>(gp.t, tm, tn, alpha, beta, beta_zero, sm);
                break;
            case 1:
                if constexpr ((VM & 2) != 0) mfma_tile<T, CF, true, false, false>(gp.t, tm, tn, alpha, beta, beta_zero, sm);
                break;
            case 2:
                if constexpr ((VM & 4) != 0) mfma_tile<T, CF, false, true, false>(gp.t, tm, tn, alpha, beta, beta_zero, sm);
                break;
            case 4:
                if constexpr (RV && (VM & 16) != 0) mfma_tile<T, CF, false, false, RV>(gp.t, tm, tn, alpha, beta, beta_zero, sm);
                break;
            default:
                if constexpr ((VM & 1) != 0) mfma_tile<T, CF, false, false, false>(gp.t, tm, tn, alpha, beta, beta_zero, sm);
                break;
        }
        if (threadIdx.x == 0) next_tile = claimed;
        __syncthreads();  // LDS reuse by the next tile; next_tile visible
        tile = next_tile;
        __syncthreads();  // every thread has read next_tile before it is written again
    }
}

/* ---------------- SIMT fallback: f32, i32, i64, c64 ---------------- */

template <class T>
__device__ __forceinline__ T g_mul(T a, T b) {
    if constexpr (std::is_integral<T>::value) {
        using U = typename std::make_unsigned<T>::type;
        return (T)((U)a * (U)b);
    } else {
        return a * b;
    }
}
template <class T>
__device__ __forceinline__ T g_add(T a, T b) {
    if constexpr (std::is_integral<T>::value) {
        using U = typename std::make_unsigned<T>::type;
        return (T)((U)a + (U)b);
    } else {
        return a + b;
    }
}
__device__ __forceinline__ c64 g_mul(c64 a, c64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ c64 g_add(c64 a, c64 b) { return {a.re + b.re, a.im + b.im}; }

template <class T>
__global__ void __launch_bounds__(256) k_gemm_simt(GemmArgs<T> g) {
    constexpr int TS = 16;
    __shared__ T As[TS][TS + 1];
    __shared__ T Bs[TS][TS + 1];
    const uint32_t tile = blockIdx.x;
    const uint32_t tm = tile / g.tiles_n, tn = tile % g.tiles_n;
    int64_t boa, bob, boc;
    batch_offsets(g.bd, g.batch0 + blockIdx.y, boa, bob, boc);
    const T* A = g.a + boa;
    const T* B = g.b + bob;
    T* C = g.c + boc;
    const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
    const int64_t gm = (int64_t)tm * TS + ty, gn = (int64_t)tn * TS + tx;
    T acc{};
    for (int64_t k0 = 0; k0 < g.K; k0 += TS) {
        int64_t ka = k0 + tx, kb = k0 + ty;
        As[ty][tx] = (gm < g.M && ka < g.K) ? A[gm * g.rsa + ka * g.csa] : T{};
        Bs[ty][tx] = (kb < g.K && gn < g.N) ? B[kb * g.rsb + gn * g.csb] : T{};
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TS; ++k) acc = g_add(acc, g_mul(As[ty][k], Bs[k][tx]));
        __syncthreads();
    }
    if (gm < g.M && gn < g.N) {
        T* cp = C + gm * g.rsc + gn * g.csc;
        T v = g_mul(g.alpha, acc);
        if (!g.beta_zero) v = g_add(v, g_mul(g.beta, *cp));
        *cp = v;
    }
}

/* ---------------- host side ---------------- */

/* ---------------- tile configurations and the autotuner ---------------- */

// MFMA tile configurations: (WM, WN) -> block tile (32*WM) x (32*WN).
// Candidate tile configurations, in the autotuner's candidate order.
using D0 = MCfg<4, 2, 2, 4>;  // f64 128 x 128, 8 waves of 32 x 64
using D1 = MCfg<4, 4, 2, 2>;  // f64 128 x 128, 16 waves of 32 x 32
using D2 = MCfg<2, 2, 4, 4>;  // f64 128 x 128, 4 waves of 64 x 64
using D3 = MCfg<2, 2, 2, 2>;  // f64 64 x 64, 4 waves of 32 x 32 (small outputs)
using D4 = MCfg<4, 2, 2, 2, 4>;  // f64 128 x 64, 8 waves of 32 x 32, two blocks per CU (grouped only)
using Z0 = MCfg<2, 2, 2, 4>;  // c128 64 x 128, 4 waves of 32 x 64
using Z1 = MCfg<2, 2, 2, 2>;  // c128 64 x 64
using Z2 = MCfg<2, 4, 2, 2>;  // c128 64 x 128, 8 waves of 32 x 32
constexpr int NCAND_D = 4, NCAND_Z = 3;

// The autotuner (tune_select, api.cpp) times every candidate configuration once
// per problem class (sizes bucketed to powers of two, layout flags, batch bucket)
// on the real call and keeps the fastest. Timing runs only when C is write-only
// (beta == 0) and the stream is not being captured; otherwise the cached choice
// or the size heuristic is used. RSTSR_HIP_TUNE_CACHE=<file> persists choices.
static int bucket(int64_t x) {
    int b = 0;
    while ((1ll << b) < x) ++b;
    return b;
}

// run(c) performs the whole call with configuration c. Returns the chosen index.
template <class F>
static int tuned_run(rstsr_hip_ctx* ctx, const std::string& key, int ncand, int heuristic, bool can_time, F&& run) {
    return tune_select(ctx, key, ncand, heuristic, can_time, std::function<void(int)>(run));
}

template <class T, class CF>
static void launch_mfma_cfg(rstsr_hip_ctx* ctx, GemmArgs<T> g, uint32_t nb, bool amc, bool bkc, bool vec,
                            const int* run_if) {
    g.tiles_m = ceil_div_u32(g.M, CF::BM);
    g.tiles_n = ceil_div_u32(g.N, CF::BN);
    uint64_t nt = (uint64_t)g.tiles_m * g.tiles_n;
    if (g.tri) {
        if (CF::BM != CF::BN || g.M != g.N) RSTSR_RAISE(RuntimeError, "symmetric product needs square tiles");
        nt = (uint64_t)g.tiles_m * (g.tiles_m + 1) / 2;
    }
    if (nt >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "GEMM grid too large");
    dim3 grid((uint32_t)nt, nb), blk(CF::NT);
    auto go = [&](auto a_mc, auto b_kc, auto v) {
        hipLaunchKernelGGL((k_gemm_mfma<T, CF, decltype(a_mc)::value, decltype(b_kc)::value, decltype(v)::value>), grid,
                           blk, 0, ctx->stream, g, run_if);
    };
    using F = std::false_type;
    using Tr = std::true_type;
    if constexpr (!CplxOf<T>::value) {
        if (vec) {
            if (amc && bkc) go(Tr{}, Tr{}, Tr{});
            else if (amc) go(Tr{}, F{}, Tr{});
            else if (bkc) go(F{}, Tr{}, Tr{});
            else go(F{}, F{}, Tr{});
            check_launch("gemm");
            return;
        }
    }
    if (amc && bkc) go(Tr{}, Tr{}, F{});
    else if (amc) go(Tr{}, F{}, F{});
    else if (bkc) go(F{}, Tr{}, F{});
    else go(F{}, F{}, F{});
    check_launch("gemm");
}

template <class T>
static void launch_mfma_any(rstsr_hip_ctx* ctx, int c, const GemmArgs<T>& g, uint32_t nb, bool amc, bool bkc,
                            bool vec, const int* run_if) {
    if constexpr (std::is_same<T, double>::value) {
        switch (c) {
            case 0: return launch_mfma_cfg<T, D0>(ctx, g, nb, amc, bkc, vec, run_if);
            case 1: return launch_mfma_cfg<T, D1>(ctx, g, nb, amc, bkc, vec, run_if);
            case 2: return launch_mfma_cfg<T, D2>(ctx, g, nb, amc, bkc, vec, run_if);
            default: return launch_mfma_cfg<T, D3>(ctx, g, nb, amc, bkc, vec, run_if);
        }
    } else {
        switch (c) {
            case 0: return launch_mfma_cfg<T, Z0>(ctx, g, nb, amc, bkc, vec, run_if);
            case 1: return launch_mfma_cfg<T, Z1>(ctx, g, nb, amc, bkc, vec, run_if);
            default: return launch_mfma_cfg<T, Z2>(ctx, g, nb, amc, bkc, vec, run_if);
        }
    }
}

template <class T>
static int heuristic_cfg(int64_t m, int64_t n) {
    int64_t s = std::min(m, n);
    if constexpr (std::is_same<T, double>::value) return s >= 128 ? 0 : 3;
    else return s >= 128 ? 0 : 1;
}

// Two-double loads along an operand's unit-stride axis: the stride along it is
// +1, the other matrix stride and every batch stride are even, and the base is
// 16-byte aligned, so every tile-aligned element pair is one aligned 16-byte word.
static bool pair_loads_ok(const void* p, int64_t unit, int64_t other, const int64_t* bs, int nbd) {
    if (unit != 1 || other % 2 != 0 || ((uintptr_t)p & 15) != 0) return false;
    for (int d = 0; d < nbd; ++d)
        if (bs[d] % 2 != 0) return false;
    return true;
}

template <class T>
static void launch_gemm_t(rstsr_hip_ctx* ctx, const GemmProblem& P, T alpha, T beta, bool beta_zero,
                          const int* run_if = nullptr) {
    GemmArgs<T> g;
    g.M = P.M;
    g.N = P.N;
    g.K = P.K;
    g.a = (const T*)P.a;
    g.rsa = P.rsa;
    g.csa = P.csa;
    g.b = (const T*)P.b;
    g.rsb = P.rsb;
    g.csb = P.csb;
    g.c = (T*)P.c;
    g.rsc = P.rsc;
    g.csc = P.csc;
    g.alpha = alpha;
    g.beta = beta;
    g.beta_zero = beta_zero ? 1 : 0;
    g.tri = 0;
    g.bd = P.bd;
    const uint64_t YMAX = 65535;
    constexpr bool MF = std::is_same<T, double>::value || std::is_same<T, c128>::value;
    if constexpr (MF) {
        constexpr bool CX = std::is_same<T, c128>::value;
        const int ncand = CX ? NCAND_Z : NCAND_D;
        bool amc = std::llabs(P.rsa) == 1 && std::llabs(P.csa) != 1;
        bool bkc = std::llabs(P.rsb) == 1 && std::llabs(P.csb) != 1;
        const bool vec = !CX && pair_loads_ok(P.a, amc ? P.rsa : P.csa, amc ? P.csa : P.rsa, P.bd.sa, P.bd.nbd) &&
                         pair_loads_ok(P.b, bkc ? P.rsb : P.csb, bkc ? P.csb : P.rsb, P.bd.sb, P.bd.nbd);
        // SYRK shortcut (device_faer/matmul.rs:38-48): beta = 0 and A is B's transpose
        // in the same buffer, so C is symmetric. Only the tiles on and below the
        // diagonal are multiplied; the rest is mirrored. For real types C[i][j] and
        // C[j][i] sum the same products in the same K order, so the mirror equals the
        // full product bit for bit; for complex the imaginary part's two products
        // swap order, and the mirror makes the result exactly symmetric.
        bool syrk = beta_zero && run_if == nullptr && P.a == P.b && P.M == P.N && P.rsa == P.csb && P.csa == P.rsb;
        for (int d = 0; d < P.bd.nbd && syrk; ++d) syrk = P.bd.sa[d] == P.bd.sb[d];
        // square candidates only (complex: the 64 x 64 tile)
        const int ncand_s = syrk ? (CX ? 1 : ncand) : ncand;
        auto cand = [&](int c) { return syrk && CX ? 1 : c; };
        auto run = [&](int c) {
            for (uint64_t b0 = 0; b0 < P.nbatch; b0 += YMAX) {
                GemmArgs<T> gg = g;
                gg.batch0 = (uint32_t)b0;
                gg.tri = syrk ? 1 : 0;
                uint32_t nb = (uint32_t)std::min<uint64_t>(YMAX, P.nbatch - b0);
                launch_mfma_any<T>(ctx, cand(c), gg, nb, amc, bkc, vec, run_if);
                if (syrk) {
                    const uint64_t t64 = ceil_div_u32(P.M, 64);
                    const uint64_t nm = t64 * (t64 + 1) / 2;
                    if (nm >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "symmetric product too large");
                    if (nm > 0) {
                        hipLaunchKernelGGL(k_mirror<T>, dim3((uint32_t)nm, nb), dim3(256), 0, ctx->stream, gg);
                        check_launch("gemm mirror");
                    }
                }
            }
        };
        char key[160];
        std::snprintf(key, sizeof(key), "gemm:%s:m%d:n%d:k%d:b%d:f%d", CX ? "z" : "d", bucket(P.M), bucket(P.N),
                      bucket(P.K), bucket((int64_t)P.nbatch),
                      (amc ? 1 : 0) | (bkc ? 2 : 0) | (vec ? 4 : 0) | (syrk ? 8 : 0));
        tuned_run(ctx, key, ncand_s, syrk && CX ? 0 : heuristic_cfg<T>(P.M, P.N), beta_zero && run_if == nullptr, run);
    } else {
        for (uint64_t b0 = 0; b0 < P.nbatch; b0 += YMAX) {
            uint32_t nb = (uint32_t)std::min<uint64_t>(YMAX, P.nbatch - b0);
            g.batch0 = (uint32_t)b0;
            g.tiles_m = ceil_div_u32(P.M, 16);
            g.tiles_n = ceil_div_u32(P.N, 16);
            uint64_t nt = (uint64_t)g.tiles_m * g.tiles_n;
            if (nt >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "GEMM grid too large");
            hipLaunchKernelGGL(k_gemm_simt<T>, dim3((uint32_t)nt, nb), dim3(256), 0, ctx->stream, g);
            check_launch("gemm");
        }
    }
}

void plain_dgemm(rstsr_hip_ctx* ctx, const GemmProblem& P, double alpha, double beta, bool beta_zero,
                 const int* run_if) {
    launch_gemm_t<double>(ctx, P, alpha, beta, beta_zero, run_if);
}


/* ---------------- grouped GEMM ---------------- */

template <class T, class CF>
static void launch_grouped_cfg(rstsr_hip_ctx* ctx, const std::vector<TileArgs<T>>& probs, T alpha, T beta,
                               bool beta_zero, const int* run_if = nullptr) {
    std::vector<GroupProblem<T>> gp(probs.size());
    std::vector<uint32_t> prefix(probs.size() + 1, 0);
    // Tiles are handed out in problem order to blocks that stride over them: longest K
    // (the time of one tile) first, so that short tiles fill the end of the launch
    // instead of a few long ones trailing it. Each problem writes only its own C, so the
    // order changes no result.
    std::vector<size_t> order(probs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return probs[x].K > probs[y].K; });
    uint64_t total = 0;
    uint32_t forms = 0;  // bit f: some problem loads with form f
    for (size_t i = 0; i < probs.size(); ++i) {
        const auto& t = probs[order[i]];
        uint32_t tm = ceil_div_u32(t.M, CF::BM), tn = ceil_div_u32(t.N, CF::BN);
        gp[i].t = t;
        gp[i].tiles_n = tn;
        bool amc = std::llabs(t.rsa) == 1 && std::llabs(t.csa) != 1;
        bool bkc = std::llabs(t.rsb) == 1 && std::llabs(t.csb) != 1;
        bool vec = !CplxOf<T>::value && !amc && !bkc && pair_loads_ok(t.A, t.csa, t.rsa, nullptr, 0) &&
                   pair_loads_ok(t.B, t.csb, t.rsb, nullptr, 0);
        gp[i].flags = vec ? 4u : ((amc ? 1u : 0u) | (bkc ? 2u : 0u));
        forms |= 1u << gp[i].flags;
        prefix[i] = (uint32_t)total;
        total += (uint64_t)tm * tn;
    }
    prefix[probs.size()] = (uint32_t)total;
    if (total == 0) return;
    if (total >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "grouped GEMM has too many tiles");
    const size_t poff = (gp.size() * sizeof(GroupProblem<T>) + 255) & ~(size_t)255;
    const size_t qoff = poff + prefix.size() * sizeof(uint32_t);  // the tile queue counter
    const size_t bytes = qoff + sizeof(uint32_t);
    char* meta = (char*)aux_buffer(ctx, bytes + 256);
    char* host = (char*)meta_stage(ctx, bytes, "grouped GEMM");
    std::memcpy(host, gp.data(), gp.size() * sizeof(GroupProblem<T>));
    std::memcpy(host + poff, prefix.data(), prefix.size() * sizeof(uint32_t));
    const uint32_t zero = 0;
    std::memcpy(host + qoff, &zero, sizeof(zero));
    meta_upload(ctx, meta, bytes);
    auto launch = [&](auto vm) {
        constexpr uint32_t VM = decltype(vm)::value;
        // persistent grid: as many blocks as can be resident at once. The launch bound sets waves
        // per SIMD, not blocks per CU, so the count comes from the occupancy query: a block beyond
        // it would start only once the queue is empty and then run its first tile (one of the
        // longest) alone.
        static const int per_cu = [] {
            int v = 0;
            HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_gemm_grouped<T, CF, VM>, CF::NT, 0));
            return std::max(v, 1);
        }();
        uint32_t grid = (uint32_t)std::min<uint64_t>(total, (uint64_t)std::max(ctx->num_cu, 1) * per_cu);
        hipLaunchKernelGGL((k_gemm_grouped<T, CF, VM>), dim3(grid), dim3(CF::NT), 0, ctx->stream,
                           (const GroupProblem<T>*)meta, (int)gp.size(), (const uint32_t*)(meta + poff),
                           (uint32_t)total, alpha, beta, beta_zero ? 1 : 0, run_if, (uint32_t*)(meta + qoff));
        check_launch("grouped gemm");
    };
    // every problem row-major in both operands (forms 0 and 4): the kernel with those two only
    if ((forms & ~17u) == 0) launch(std::integral_constant<uint32_t, 17>{});
    else launch(std::integral_constant<uint32_t, 31>{});
}

template <class T>
static void grouped_t(rstsr_hip_ctx* ctx, const std::vector<TileArgs<T>>& probs, T alpha, T beta, bool beta_zero) {
    if (probs.empty()) return;
    constexpr bool CX = std::is_same<T, c128>::value;
    double mean = 0;
    for (auto& p : probs) mean += (double)std::min(p.M, p.N);
    mean /= (double)probs.size();
    int64_t tot_mnk = 0;
    for (auto& p : probs) tot_mnk += p.M * p.N * p.K;
    auto run = [&](int c) {
        if constexpr (CX) {
            if (c == 0) launch_grouped_cfg<T, Z2>(ctx, probs, alpha, beta, beta_zero);
            else launch_grouped_cfg<T, Z1>(ctx, probs, alpha, beta, beta_zero);
        } else {
            if (c == 0) launch_grouped_cfg<T, D0>(ctx, probs, alpha, beta, beta_zero);
            else if (c == 1) launch_grouped_cfg<T, D1>(ctx, probs, alpha, beta, beta_zero);
            else if (c == 2) launch_grouped_cfg<T, D3>(ctx, probs, alpha, beta, beta_zero);
            else launch_grouped_cfg<T, D4>(ctx, probs, alpha, beta, beta_zero);
        }
    };
    char key[160];
    std::snprintf(key, sizeof(key), "grouped:%s:n%d:mean%d:work%d", CX ? "z" : "d", bucket((int64_t)probs.size()),
                  bucket((int64_t)mean), bucket(tot_mnk));
    tuned_run(ctx, key, CX ? 2 : 4, mean >= 128 ? (CX ? 0 : 1) : (CX ? 1 : 2), beta_zero, run);
}

void plain_grouped_f64(rstsr_hip_ctx* ctx, const std::vector<GemmProblem>& probs, double alpha, double beta,
                       bool beta_zero, const int* run_if) {
    std::vector<TileArgs<double>> ta;
    ta.reserve(probs.size());
    for (auto& P : probs)
        ta.push_back(TileArgs<double>{P.M, P.N, P.K, (const double*)P.a, P.rsa, P.csa, (const double*)P.b, P.rsb, P.csb,
                                      (double*)P.c, P.rsc, P.csc});
    if (ta.empty()) return;
    // unconditional: the tuned configuration; conditional (a fallback inside a captured
    // sequence): the fixed one, since the autotuner cannot time a launch that may not run
    if (run_if) launch_grouped_cfg<double, D1>(ctx, ta, alpha, beta, beta_zero, run_if);
    else grouped_t<double>(ctx, ta, alpha, beta, beta_zero);
}

// Zero-fill C (k == 0 with beta == 0: crates-device/rstsr-openblas/src/matmul_impl.rs:78-81).
template <class T>
__global__ void k_scale_c(GemmArgs<T> g) {
    uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint64_t mn = (uint64_t)g.M * g.N;
    if (idx >= mn) return;
    int64_t oa, ob, oc;
    batch_offsets(g.bd, g.batch0 + blockIdx.y, oa, ob, oc);
    int64_t m = idx / g.N, n = idx - (idx / g.N) * g.N;
    T* cp = g.c + oc + m * g.rsc + n * g.csc;
    *cp = g.beta_zero ? T{} : g_mul(g.beta, *cp);
}
__device__ __forceinline__ c128 g_mul(c128 a, c128 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

template <class T>
static void scale_c(rstsr_hip_ctx* ctx, const GemmProblem& P, T beta, bool beta_zero) {
    GemmArgs<T> g{};
    g.M = P.M;
    g.N = P.N;
    g.c = (T*)P.c;
    g.rsc = P.rsc;
    g.csc = P.csc;
    g.beta = beta;
    g.beta_zero = beta_zero;
    g.bd = P.bd;
    uint64_t mn = (uint64_t)P.M * P.N;
    for (uint64_t b0 = 0; b0 < P.nbatch; b0 += 65535) {
        g.batch0 = (uint32_t)b0;
        uint32_t nb = (uint32_t)std::min<uint64_t>(65535, P.nbatch - b0);
        hipLaunchKernelGGL(k_scale_c<T>, dim3(ceil_div_u32(mn, 256), nb), dim3(256), 0, ctx->stream, g);
        check_launch("gemm scale");
    }
}

static bool is_zero_scalar(rstsr_dtype_t t, const void* v) {
    switch (t) {
        case RSTSR_F32: return *(const float*)v == 0.0f;
        case RSTSR_F64: return *(const double*)v == 0.0;
        case RSTSR_I32: return *(const int32_t*)v == 0;
        case RSTSR_I64: return *(const int64_t*)v == 0;
        case RSTSR_C64: return ((const float*)v)[0] == 0.0f && ((const float*)v)[1] == 0.0f;
        case RSTSR_C128: return ((const double*)v)[0] == 0.0 && ((const double*)v)[1] == 0.0;
        default: return false;
    }
}

// S (rows x cols, dense row-major) = Re X + Im X for a strided complex matrix X
__global__ void __launch_bounds__(256) k_zsum_parts(double* __restrict__ S, double* __restrict__ Xr,
                                                    double* __restrict__ Xi, const c128* __restrict__ X, int64_t rows,
                                                    int64_t cols, int64_t rs, int64_t cs) {
    // row blockIdx.y (grid-strided), columns along the block: S = re + im, and the real and
    // imaginary planes as contiguous rows (the split then reads them once, at unit stride)
    for (int64_t i = blockIdx.y; i < rows; i += gridDim.y) {
        const c128* xr = X + i * rs;
        const int64_t o = i * cols;
        for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < cols; j += (int64_t)gridDim.x * 256) {
            const c128 x = xr[j * cs];
            S[o + j] = x.re + x.im;
            Xr[o + j] = x.re;
            Xi[o + j] = x.im;
        }
    }
}

// C holds (a T1, a T3) with T1 = Ar Br, T3 = (Ar + Ai)(Br + Bi); T2 = a Ai Bi (dense M x N).
// C = (a T1 - a T2, a T3 - a T1 - a T2).
__global__ void __launch_bounds__(256) k_z3m_combine(c128* __restrict__ C, int64_t rsc, int64_t csc,
                                                     const double* __restrict__ T2, int64_t M, int64_t N) {
    const int64_t n = M * N;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / N, j = e - i * N;
        c128* p = C + i * rsc + j * csc;
        const c128 c = *p;
        const double t2 = T2[e];
        *p = c128{c.re - t2, (c.im - c.re) - t2};
    }
}

// Complex split GEMM from real split products on the real/imaginary planes
// (stride-2 double views). With beta = 0 and one matrix: three products (below);
// otherwise four: Cr = a ArBr - a AiBi + b Cr, Ci = a ArBi + a AiBr + b Ci.
static void ozaki_zgemm(rstsr_hip_ctx* ctx, const GemmProblem& P, c128 alpha, c128 beta, bool beta_zero) {
    if (alpha.im != 0.0 || beta.im != 0.0)
        RSTSR_RAISE(UnImplemented, "the complex split GEMM takes real alpha and beta");
    if (beta_zero && P.nbatch == 1 && P.bd.nbd == 0 && P.K > 0) {
        // Three real split products (Karatsuba form): T1 = Ar Br, T2 = Ai Bi,
        // T3 = (Ar + Ai)(Br + Bi); Re C = a (T1 - T2), Im C = a (T3 - T1 - T2).
        // Each product is accurate to the split GEMM's normwise bound, so the
        // result keeps it relative to (|Ar| + |Ai|)(|Br| + |Bi|).
        const int64_t M = P.M, N = P.N, K = P.K;
        const size_t nSA = (size_t)M * K, nSB = (size_t)K * N, nT = (size_t)M * N;
        auto al = [](size_t x) { return (x * sizeof(double) + 255) & ~(size_t)255; };
        // the sums Ar + Ai and Br + Bi, T2, and the four parts as contiguous planes
        char* z = (char*)zbuf_buffer(ctx, 3 * al(nSA) + 3 * al(nSB) + al(nT));
        double* SA = (double*)z;
        double* SB = (double*)(z + al(nSA));
        double* T2 = (double*)(z + al(nSA) + al(nSB));
        double* Ar = (double*)(z + al(nSA) + al(nSB) + al(nT));
        double* Ai = Ar + al(nSA) / sizeof(double);
        double* Br = Ai + al(nSA) / sizeof(double);
        double* Bi = Br + al(nSB) / sizeof(double);
        auto grid = [](int64_t rows, int64_t cols) {
            return dim3((uint32_t)std::min<int64_t>((cols + 255) / 256, 64), (uint32_t)std::min<int64_t>(rows, 65535));
        };
        hipLaunchKernelGGL(k_zsum_parts, grid(M, K), dim3(256), 0, ctx->stream, SA, Ar, Ai, (const c128*)P.a, M, K,
                           P.rsa, P.csa);
        hipLaunchKernelGGL(k_zsum_parts, grid(K, N), dim3(256), 0, ctx->stream, SB, Br, Bi, (const c128*)P.b, K, N,
                           P.rsb, P.csb);
        check_launch("complex split sums");
        const double a = alpha.re;
        GemmProblem Q = P;
        Q.bd.nbd = 0;
        Q.nbatch = 1;
        // T1 into Re C
        Q.a = Ar;
        Q.rsa = K;
        Q.csa = 1;
        Q.b = Br;
        Q.rsb = N;
        Q.csb = 1;
        Q.c = P.c;
        Q.rsc = 2 * P.rsc;
        Q.csc = 2 * P.csc;
        ozaki_gemm_f64(ctx, Q, a, 0.0, true);
        // T3 into Im C
        Q.a = SA;
        Q.rsa = K;
        Q.csa = 1;
        Q.b = SB;
        Q.rsb = N;
        Q.csb = 1;
        Q.c = (double*)P.c + 1;
        ozaki_gemm_f64(ctx, Q, a, 0.0, true);
        // T2 into the temporary
        Q.a = Ai;
        Q.rsa = K;
        Q.csa = 1;
        Q.b = Bi;
        Q.rsb = N;
        Q.csb = 1;
        Q.c = T2;
        Q.rsc = N;
        Q.csc = 1;
        ozaki_gemm_f64(ctx, Q, a, 0.0, true);
        hipLaunchKernelGGL(k_z3m_combine, dim3((uint32_t)std::min<size_t>((nT + 255) / 256, 65536)), dim3(256), 0,
                           ctx->stream, (c128*)P.c, P.rsc, P.csc, T2, M, N);
        check_launch("complex split combine");
        return;
    }
    auto part = [&](const void* base, int im) { return (const void*)((const double*)base + im); };
    GemmProblem R = P;
    R.rsa *= 2;
    R.csa *= 2;
    R.rsb *= 2;
    R.csb *= 2;
    R.rsc *= 2;
    R.csc *= 2;
    for (int d = 0; d < R.bd.nbd; ++d) {
        R.bd.sa[d] *= 2;
        R.bd.sb[d] *= 2;
        R.bd.sc[d] *= 2;
    }
    const double a = alpha.re, b = beta.re;
    auto prod = [&](int ai, int bi, int ci, double al, double be, bool bz) {
        GemmProblem Q = R;
        Q.a = part(P.a, ai);
        Q.b = part(P.b, bi);
        Q.c = (void*)((double*)P.c + ci);
        ozaki_gemm_f64(ctx, Q, al, be, bz);
    };
    prod(0, 0, 0, a, b, beta_zero);   // Cr  = a Ar Br + b Cr
    prod(1, 1, 0, -a, 1.0, false);    // Cr -= a Ai Bi
    prod(0, 1, 1, a, b, beta_zero);   // Ci  = a Ar Bi + b Ci
    prod(1, 0, 1, a, 1.0, false);     // Ci += a Ai Br
}

// Merge batch axes into the matrix: the innermost batch axis folds into M when
// B is the same for every batch and A's and C's batch strides continue their
// row strides (sa = M rsa, sc = M rsc), so batch b row m is row b M + m of one
// taller product; likewise into N when A is shared and B and C chain along
// columns. One large GEMM instead of many, and no per-batch tile padding: the
// 60 x (400 x 24000 x 1600) RI contraction becomes one 24000 x 24000 x 1600
// product, the broadcast 64 x 1024^3 batch one 65536 x 1024 x 1024.
static GemmProblem fold_batches(const GemmProblem& P0) {
    GemmProblem P = P0;
    while (P.bd.nbd > 0) {
        const int64_t s = P.bd.shape[0];
        if (P.bd.sb[0] == 0 && P.bd.sa[0] == P.M * P.rsa && P.bd.sc[0] == P.M * P.rsc &&
            (uint64_t)(P.M * s) < (1ull << 40)) {
            P.M *= s;
        } else if (P.bd.sa[0] == 0 && P.bd.sb[0] == P.N * P.csb && P.bd.sc[0] == P.N * P.csc &&
                   (uint64_t)(P.N * s) < (1ull << 40)) {
            P.N *= s;
        } else {
            break;
        }
        for (int d = 1; d < P.bd.nbd; ++d) {
            P.bd.shape[d - 1] = P.bd.shape[d];
            P.bd.fd[d - 1] = P.bd.fd[d];
            P.bd.sa[d - 1] = P.bd.sa[d];
            P.bd.sb[d - 1] = P.bd.sb[d];
            P.bd.sc[d - 1] = P.bd.sc[d];
        }
        P.bd.nbd -= 1;
        P.nbatch /= (uint64_t)s;
    }
    return P;
}

void run_gemm(rstsr_hip_ctx* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, const GemmProblem& P0, const void* alpha,
              const void* beta) {
    if (!alpha || !beta) RSTSR_RAISE(InvalidValue, "null alpha/beta");
    if (P0.M == 0 || P0.N == 0 || P0.nbatch == 0) return;
    const GemmProblem P = fold_batches(P0);
    bool bz = is_zero_scalar(t, beta);
    if (mode != RSTSR_GEMM_PLAIN && mode != RSTSR_GEMM_OZAKI_INT8) RSTSR_RAISE(InvalidValue, "unknown GEMM mode");
    if (mode == RSTSR_GEMM_OZAKI_INT8 && t != RSTSR_F64 && t != RSTSR_C128)
        RSTSR_RAISE(UnImplemented, "the Ozaki split GEMM is provided for f64 and c128");
    auto body = [&](auto tag) {
        using T = decltype(tag);
        T al, be;
        std::memcpy(&al, alpha, sizeof(T));
        std::memcpy(&be, beta, sizeof(T));
        if (P.K == 0) {
            scale_c<T>(ctx, P, be, bz);
            return;
        }
        if constexpr (std::is_same<T, double>::value) {
            if (mode == RSTSR_GEMM_OZAKI_INT8) {
                ozaki_gemm_f64(ctx, P, al, be, bz);
                return;
            }
        }
        if constexpr (std::is_same<T, c128>::value) {
            if (mode == RSTSR_GEMM_OZAKI_INT8) {
                ozaki_zgemm(ctx, P, al, be, bz);
                return;
            }
        }
        launch_gemm_t<T>(ctx, P, al, be, bz);
    };
    switch (t) {
        case RSTSR_F32: body(float{}); break;
        case RSTSR_F64: body(double{}); break;
        case RSTSR_I32: body(int32_t{}); break;
        case RSTSR_I64: body(int64_t{}); break;
        case RSTSR_C64: body(c64{}); break;
        case RSTSR_C128: body(c128{}); break;
        default: RSTSR_RAISE(UnImplemented, "matmul dtype");
    }
}

// Check that each operand's matrices of every batch stay inside its buffer.
static void check_gemm_extent(const GemmProblem& P, size_t es) {
    auto mk = [&](const void* p, int64_t r, int64_t cdim, int64_t rs, int64_t cs, const int64_t* bs, const char* nm) {
        size_t s[MAXD + 2];
        int64_t t[MAXD + 2];
        int nd = 0;
        for (int d = 0; d < P.bd.nbd; ++d, ++nd) {
            s[nd] = P.bd.shape[d];
            t[nd] = bs[d];
        }
        s[nd] = (size_t)r;
        t[nd++] = rs;
        s[nd] = (size_t)cdim;
        t[nd++] = cs;
        check_extent_rel(p, (size_t)nd, s, t, es, nm);
    };
    mk(P.a, P.M, P.K, P.rsa, P.csa, P.bd.sa, "a");
    mk(P.b, P.K, P.N, P.rsb, P.csb, P.bd.sb, "b");
    mk(P.c, P.M, P.N, P.rsc, P.csc, P.bd.sc, "c");
}

}  // namespace hip
}  // namespace rstsr

using namespace rstsr;
using namespace rstsr::hip;

extern "C" {

int rstsr_hip_gemm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, size_t m, size_t n, size_t k,
                   const void* alpha, const void* a, ptrdiff_t rsa, ptrdiff_t csa, const void* b, ptrdiff_t rsb,
                   ptrdiff_t csb, const void* beta, void* c, ptrdiff_t rsc, ptrdiff_t csc) {
    return guarded([&] {
        bind(ctx);
        GemmProblem P{};
        P.M = (int64_t)m;
        P.N = (int64_t)n;
        P.K = (int64_t)k;
        P.a = a;
        P.rsa = rsa;
        P.csa = csa;
        P.b = b;
        P.rsb = rsb;
        P.csb = csb;
        P.c = c;
        P.rsc = rsc;
        P.csc = csc;
        P.bd.nbd = 0;
        P.nbatch = 1;
        check_gemm_extent(P, dtype_size(t));
        std::lock_guard<std::mutex> g(ctx->mu);
        run_gemm(ctx, t, mode, P, alpha, beta);
    });
}

// rstsr-common/src/layout/matmul.rs:287-421 (layout_matmul_dyn_row_major_with_lc):
// split each operand into (batch rest, matrix) and broadcast the batch rests
// of A and B to C's; vectors become 1 x K / K x 1 matrices.
int rstsr_hip_matmul(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, void* c, const rstsr_layout_t* lc_, const void* a,
                     const rstsr_layout_t* la_, const void* b, const rstsr_layout_t* lb_, const void* alpha,
                     const void* beta, rstsr_gemm_mode_t mode) {
    return guarded([&] {
        bind(ctx);
        Layout lc = from_abi(lc_), la = from_abi(la_), lb = from_abi(lb_);
        size_t na = la.ndim(), nb = lb.ndim(), nc = lc.ndim();
        if (na == 0 || nb == 0) RSTSR_RAISE(InvalidValue, "In matmul, 0-dim is not allowed.");
        // matrices as (shape, stride) pairs
        int64_t M, N, K, rsa, csa, rsb, csb, rsc, csc;
        Layout ra, rb, rc;  // batch rests
        auto split = [](const Layout& l, size_t nm, Layout& rest) {
            Shape s(l.shape.begin(), l.shape.end() - nm);
            Stride st(l.stride.begin(), l.stride.end() - nm);
            rest = Layout(s, st, 0);
        };
        if (na == 1 && nb == 1) {  // rule 1: inner dot -> 1 x 1
            if (la.shape != lb.shape || nc != 0) RSTSR_RAISE(InvalidLayout, "matmul rule 1 shape mismatch");
            M = 1;
            N = 1;
            K = la.shape[0];
            rsa = 0;
            csa = la.stride[0];
            rsb = lb.stride[0];
            csb = 0;
            rsc = 0;
            csc = 0;
            ra = rb = rc = Layout({}, {}, 0);
        } else if (na == 1) {  // rule 3: K @ (..., K, N) -> (..., N)
            if (nb != nc + 1 || la.shape[0] != lb.shape[nb - 2] || lc.shape[nc - 1] != lb.shape[nb - 1])
                RSTSR_RAISE(InvalidLayout, "matmul rule 3 shape mismatch");
            M = 1;
            K = la.shape[0];
            N = lb.shape[nb - 1];
            rsa = 0;
            csa = la.stride[0];
            rsb = lb.stride[nb - 2];
            csb = lb.stride[nb - 1];
            rsc = 0;
            csc = lc.stride[nc - 1];
            split(lb, 2, rb);
            split(lc, 1, rc);
            ra = Layout(rc.shape, Stride(rc.ndim(), 0), 0);
        } else if (nb == 1) {  // rule 4: (..., M, K) @ K -> (..., M)
            if (na != nc + 1 || lb.shape[0] != la.shape[na - 1] || lc.shape[nc - 1] != la.shape[na - 2])
                RSTSR_RAISE(InvalidLayout, "matmul rule 4 shape mismatch");
            M = la.shape[na - 2];
            K = la.shape[na - 1];
            N = 1;
            rsa = la.stride[na - 2];
            csa = la.stride[na - 1];
            rsb = lb.stride[0];
            csb = 0;
            rsc = lc.stride[nc - 1];
            csc = 0;
            split(la, 2, ra);
            split(lc, 1, rc);
            rb = Layout(rc.shape, Stride(rc.ndim(), 0), 0);
        } else {  // rules 2, 5, 6, 7
            if (nc < 2) RSTSR_RAISE(InvalidLayout, "matmul output rank");
            M = la.shape[na - 2];
            K = la.shape[na - 1];
            N = lb.shape[nb - 1];
            if (lb.shape[nb - 2] != (size_t)K || lc.shape[nc - 2] != (size_t)M || lc.shape[nc - 1] != (size_t)N)
                RSTSR_RAISE(InvalidLayout, "matmul shape mismatch");
            rsa = la.stride[na - 2];
            csa = la.stride[na - 1];
            rsb = lb.stride[nb - 2];
            csb = lb.stride[nb - 1];
            rsc = lc.stride[nc - 2];
            csc = lc.stride[nc - 1];
            split(la, 2, ra);
            split(lb, 2, rb);
            split(lc, 2, rc);
        }
        // broadcast batch rests of A and B to C's (broadcast_layout_to_first)
        ra = broadcast_layout_to_shape(ra, rc.shape, FlagOrder::RowMajor);
        rb = broadcast_layout_to_shape(rb, rc.shape, FlagOrder::RowMajor);
        GemmProblem P{};
        P.M = M;
        P.N = N;
        P.K = K;
        P.a = (const char*)a + la.offset * dtype_size(t);
        P.b = (const char*)b + lb.offset * dtype_size(t);
        P.c = (char*)c + lc.offset * dtype_size(t);
        P.rsa = rsa;
        P.csa = csa;
        P.rsb = rsb;
        P.csb = csb;
        P.rsc = rsc;
        P.csc = csc;
        // batch dims: drop size 1, innermost first for the unravel
        std::vector<size_t> bd;
        for (size_t i = 0; i < rc.ndim(); ++i)
            if (rc.shape[i] != 1) bd.push_back(i);
        if ((int)bd.size() > MAXD) RSTSR_RAISE(UnImplemented, "too many batch axes");
        P.bd.nbd = (int)bd.size();
        uint64_t nbat = 1;
        for (int d = 0; d < MAXD; ++d) {
            P.bd.shape[d] = 1;
            P.bd.fd[d] = FastDiv(1);
            P.bd.sa[d] = P.bd.sb[d] = P.bd.sc[d] = 0;
        }
        for (int d = 0; d < P.bd.nbd; ++d) {
            size_t ax = bd[bd.size() - 1 - d];
            if (rc.shape[ax] >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "batch extent too large");
            P.bd.shape[d] = (uint32_t)rc.shape[ax];
            P.bd.fd[d] = FastDiv((uint32_t)rc.shape[ax]);
            P.bd.sa[d] = ra.stride[ax];
            P.bd.sb[d] = rb.stride[ax];
            P.bd.sc[d] = rc.stride[ax];
            nbat *= rc.shape[ax];
        }
        if (nbat >= (1ull << 32)) RSTSR_RAISE(UnImplemented, "too many batches");
        P.nbatch = nbat;
        if (lc.size() == 0) return;
        // extents (relative to the operand base pointers, offsets folded in)
        {
            size_t es = dtype_size(t);
            check_extent(a, la, es, "a");
            check_extent(b, lb, es, "b");
            check_extent(c, lc, es, "c");
        }
        std::lock_guard<std::mutex> g(ctx->mu);
        run_gemm(ctx, t, mode, P, alpha, beta);
    });
}

int rstsr_hip_gemm_grouped(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, size_t count,
                           const rstsr_gemm_desc_t* descs, const void* alpha, const void* beta) {
    return guarded([&] {
        bind(ctx);
        if (count && !descs) RSTSR_RAISE(InvalidValue, "null descriptors");
        if (!alpha || !beta) RSTSR_RAISE(InvalidValue, "null alpha/beta");
        size_t es = dtype_size(t);
        std::vector<GemmProblem> probs;
        for (size_t i = 0; i < count; ++i) {
            const auto& d = descs[i];
            GemmProblem P{};
            P.M = (int64_t)d.m;
            P.N = (int64_t)d.n;
            P.K = (int64_t)d.k;
            P.a = d.a;
            P.rsa = d.rsa;
            P.csa = d.csa;
            P.b = d.b;
            P.rsb = d.rsb;
            P.csb = d.csb;
            P.c = d.c;
            P.rsc = d.rsc;
            P.csc = d.csc;
            P.bd.nbd = 0;
            P.nbatch = 1;
            check_gemm_extent(P, es);
            if (P.M && P.N) probs.push_back(P);
        }
        std::lock_guard<std::mutex> g(ctx->mu);
        bool plain_mfma = mode == RSTSR_GEMM_PLAIN && (t == RSTSR_F64 || t == RSTSR_C128);
        std::vector<GemmProblem> rest;
        if (plain_mfma) {
            auto go = [&](auto tag) {
                using T = decltype(tag);
                T al, be;
                std::memcpy(&al, alpha, sizeof(T));
                std::memcpy(&be, beta, sizeof(T));
                std::vector<TileArgs<T>> ta;
                for (auto& P : probs) {
                    if (P.K == 0) {
                        rest.push_back(P);
                        continue;
                    }
                    ta.push_back(TileArgs<T>{P.M, P.N, P.K, (const T*)P.a, P.rsa, P.csa, (const T*)P.b, P.rsb, P.csb,
                                             (T*)P.c, P.rsc, P.csc});
                }
                grouped_t<T>(ctx, ta, al, be, is_zero_scalar(t, beta));
            };
            if (t == RSTSR_F64) go(double{});
            else go(c128{});
        } else if (mode == RSTSR_GEMM_OZAKI_INT8 && t == RSTSR_F64) {
            double al, be;
            std::memcpy(&al, alpha, sizeof(double));
            std::memcpy(&be, beta, sizeof(double));
            std::vector<GemmProblem> oz;
            for (auto& P : probs) (P.K == 0 ? rest : oz).push_back(P);
            ozaki_grouped_f64(ctx, oz, al, be, is_zero_scalar(t, beta));
        } else {
            rest = probs;
        }
        // complex split mode and non-MFMA dtypes: one call per problem, in order
        for (auto& P : rest) run_gemm(ctx, t, mode, P, alpha, beta);
    });
}

}  // extern "C"
