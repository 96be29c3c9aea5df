// Shared GEMM problem description (gemm.hip, ozaki.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace rstsr {
namespace hip {

// Batch axes of a strided-batched GEMM, innermost first.
struct BatchDesc {
    int nbd = 0;
    uint32_t shape[MAXD];
    FastDiv fd[MAXD];
    int64_t sa[MAXD], sb[MAXD], sc[MAXD];
};

// One strided-batched GEMM: C[m,n] = alpha * sum_k A[m,k] B[k,n] + beta * C[m,n]
// per batch; matrices by (row stride, column stride) in elements.
struct GemmProblem {
    int64_t M, N, K;
    const void* a;
    int64_t rsa, csa;
    const void* b;
    int64_t rsb, csb;
    void* c;
    int64_t rsc, csc;
    BatchDesc bd;
    uint64_t nbatch = 1;
};

__device__ __forceinline__ uint32_t xcd_remap8(uint32_t t, uint32_t ntiles) {
    if (ntiles % 8 != 0) return t;
    return (t % 8) * (ntiles / 8) + t / 8;
}

// Block id -> output tile. Ids are round-robined over the 8 XCDs, so each XCD
// first gets a contiguous range of tiles; inside it, tiles run down groups of
// GM tile rows, so the blocks resident on one XCD at a time share GM row panels
// of A and a few column panels of B in that XCD's L2.
template <int GM>
__device__ __forceinline__ void tile_coords(uint32_t id, uint32_t tiles_m, uint32_t tiles_n, uint32_t& tm,
                                            uint32_t& tn) {
    const uint32_t t = xcd_remap8(id, tiles_m * tiles_n);
    const uint32_t per = GM * tiles_n, grp = t / per, first = grp * GM;
    const uint32_t gsz = min((uint32_t)GM, tiles_m - first), r = t - grp * per;
    tm = first + r % gsz;
    tn = r / gsz;
}

// Device offsets of batch `bi` (innermost batch axis first).
__device__ __forceinline__ void batch_offsets(const BatchDesc& bd, uint32_t bi, int64_t& oa, int64_t& ob, int64_t& oc) {
    oa = ob = oc = 0;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        if (d >= bd.nbd) break;
        uint32_t q = fdiv(bi, bd.fd[d]);
        uint32_t r = bi - q * bd.shape[d];
        oa += (int64_t)r * bd.sa[d];
        ob += (int64_t)r * bd.sb[d];
        oc += (int64_t)r * bd.sc[d];
        bi = q;
    }
}

// Host offsets of batch `bi` (outermost-last unravel of bd).
inline void batch_offsets_host(const BatchDesc& bd, uint64_t bi, int64_t& oa, int64_t& ob, int64_t& oc) {
    oa = ob = oc = 0;
    for (int d = 0; d < bd.nbd; ++d) {
        uint64_t r = bi % bd.shape[d];
        bi /= bd.shape[d];
        oa += (int64_t)r * bd.sa[d];
        ob += (int64_t)r * bd.sb[d];
        oc += (int64_t)r * bd.sc[d];
    }
}

// Ozaki-II split GEMM for f64 (ozaki.hip). Falls back to the plain kernel for
// non-finite inputs.
void ozaki_gemm_f64(rstsr_hip_ctx* ctx, const GemmProblem& P, double alpha, double beta, bool beta_zero);
// Plain FP64 MFMA DGEMM of one problem (gemm.hip). run_if (device, nullable):
// the kernel exits unless *run_if != 0.
void plain_dgemm(rstsr_hip_ctx* ctx, const GemmProblem& P, double alpha, double beta, bool beta_zero,
                 const int* run_if = nullptr);
// Grouped split GEMM of unbatched f64 problems with M, N, K > 0 (ozaki.hip).
// Outputs must not overlap.
void ozaki_grouped_f64(rstsr_hip_ctx* ctx, const std::vector<GemmProblem>& probs, double alpha, double beta,
                       bool beta_zero);
// Plain grouped FP64 GEMM in one persistent launch (gemm.hip): the autotuned tile
// shape, or with run_if (as for plain_dgemm) a fixed one.
void plain_grouped_f64(rstsr_hip_ctx* ctx, const std::vector<GemmProblem>& probs, double alpha, double beta,
                       bool beta_zero, const int* run_if);

}  // namespace hip
}  // namespace rstsr
