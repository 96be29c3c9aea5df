// GPU error checker for the GEMM legs (test infrastructure: loaded only by tests/ and
// bench.py's check leg, never by the product library).
//
// k_dot2 computes C* = A B in Dot2 (Ogita, Rump, Oishi, "Accurate sum and dot product",
// SIAM J. Sci. Comput. 26(6), 2005): every product split exactly into p + e (e = fma(a, b, -p)),
// p summed with TwoSum, the rounding errors of both accumulated in a second double. Each
// element runs k = 0, 1, ..., K-1 in order with the same operations as the CPU oracle's
// o_gemm_dot2_f64, so the two agree bit for bit (the oracle's Dekker TwoProduct yields the
// same exact e). Built with -ffp-contract=off so that no step is fused.
//
// k_normwise reduces max|C - C*| and max|C*| per block (NaN in C counts as infinity).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace {

constexpr int TM = 64, TN = 64, TK = 16, NT = 256;

__global__ void __launch_bounds__(NT) k_dot2(int64_t m, int64_t n, int64_t k, const double* __restrict__ a,
                                              int64_t rsa, int64_t csa, const double* __restrict__ b, int64_t rsb,
                                              int64_t csb, double* __restrict__ c, int64_t rsc, int64_t csc) {
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int64_t i0 = (int64_t)blockIdx.y * TM, j0 = (int64_t)blockIdx.x * TN;
    double s[4][4], cc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) s[r][q] = cc[r][q] = 0.0;
    for (int64_t k0 = 0; k0 < k; k0 += TK) {
        for (int l = threadIdx.x; l < TM * TK; l += NT) {
            const int ii = l / TK, kk = l % TK;
            const int64_t gi = i0 + ii, gk = k0 + kk;
            As[kk][ii] = (gi < m && gk < k) ? a[gi * rsa + gk * csa] : 0.0;
        }
        for (int l = threadIdx.x; l < TK * TN; l += NT) {
            const int kk = l / TN, jj = l % TN;
            const int64_t gk = k0 + kk, gj = j0 + jj;
            Bs[kk][jj] = (gk < k && gj < n) ? b[gk * rsb + gj * csb] : 0.0;
        }
        __syncthreads();
        const int kmax = (int)(k - k0 < TK ? k - k0 : TK);
        for (int kk = 0; kk < kmax; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = As[kk][ty + 16 * r];
#pragma unroll
            for (int q = 0; q < 4; ++q) bv[q] = Bs[kk][tx + 16 * q];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double p = av[r] * bv[q];
                    const double e = fma(av[r], bv[q], -p);
                    const double t = s[r][q] + p;
                    const double bp = t - s[r][q];
                    const double err = (s[r][q] - (t - bp)) + (p - bp);
                    s[r][q] = t;
                    cc[r][q] += err + e;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t gi = i0 + ty + 16 * r, gj = j0 + tx + 16 * q;
            if (gi < m && gj < n) c[gi * rsc + gj * csc] = s[r][q] + cc[r][q];
        }
}

__global__ void __launch_bounds__(NT) k_normwise(int64_t m, int64_t n, const double* __restrict__ got, int64_t rsg,
                                                  int64_t csg, const double* __restrict__ ref, int64_t rsr,
                                                  int64_t csr, double* __restrict__ blk) {
    __shared__ double se[NT], sr[NT];
    double me = 0.0, mr = 0.0;
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < total; t += (int64_t)gridDim.x * NT) {
        const int64_t i = t / n, j = t - i * n;
        const double g = got[i * rsg + j * csg], r = ref[i * rsr + j * csr];
        double d = fabs(g - r);
        if (d != d) d = INFINITY;
        const double ar = fabs(r);
        me = d > me ? d : me;
        mr = ar > mr ? ar : mr;
    }
    se[threadIdx.x] = me;
    sr[threadIdx.x] = mr;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] = se[threadIdx.x + w] > se[threadIdx.x] ? se[threadIdx.x + w] : se[threadIdx.x];
            sr[threadIdx.x] = sr[threadIdx.x + w] > sr[threadIdx.x] ? sr[threadIdx.x + w] : sr[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        blk[2 * blockIdx.x] = se[0];
        blk[2 * blockIdx.x + 1] = sr[0];
    }
}

}  // namespace

extern "C" {

// C* (m x n, strides rsc/csc, device memory) = A B in Dot2; A, B device f64 with element strides.
// Runs on the null stream and waits for it. Returns 0 or the HIP error code.
int dot2_gemm(int64_t m, int64_t n, int64_t k, const double* a, int64_t rsa, int64_t csa, const double* b,
              int64_t rsb, int64_t csb, double* c, int64_t rsc, int64_t csc) {
    if (m <= 0 || n <= 0) return 0;
    const int64_t gx = (n + TN - 1) / TN, gy = (m + TM - 1) / TM;
    if (gx >= (1ll << 31) || gy >= 65536) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dot2, dim3((unsigned)gx, (unsigned)gy), dim3(NT), 0, 0, m, n, k, a, rsa, csa, b, rsb, csb,
                       c, rsc, csc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    return (int)hipDeviceSynchronize();
}

// out[0] = max|got - ref|, out[1] = max|ref| over an m x n pair of device matrices (host out).
int dot2_normwise(int64_t m, int64_t n, const double* got, int64_t rsg, int64_t csg, const double* ref,
                  int64_t rsr, int64_t csr, double* out) {
    out[0] = out[1] = 0.0;
    if (m <= 0 || n <= 0) return 0;
    const int nb = 2048;
    double* blk = nullptr;
    hipError_t e = hipMalloc((void**)&blk, sizeof(double) * 2 * nb);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_normwise, dim3(nb), dim3(NT), 0, 0, m, n, got, rsg, csg, ref, rsr, csr, blk);
    e = hipGetLastError();
    double h[2 * nb];
    if (e == hipSuccess) e = hipMemcpy(h, blk, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(blk);
    if (e != hipSuccess) return (int)e;
    for (int i = 0; i < nb; ++i) {
        out[0] = h[2 * i] > out[0] ? h[2 * i] : out[0];
        out[1] = h[2 * i + 1] > out[1] ? h[2 * i + 1] : out[1];
    }
    return 0;
}

}  // extern "C"
