// Ozaki scheme II split-precision DGEMM on gfx950 INT8 MFMA.
//
// The reference multiplies f64 matrices with faer's FP64 GEMM
// (rstsr-core/src/device_faer/matmul_impl.rs:19-102); this is an additional
// arithmetic for the same DeviceMatMulAPI call, selected by
// RSTSR_GEMM_OZAKI_INT8. Method: Ozaki, Uchino, Imamura, "Ozaki Scheme II: A
// GEMM-oriented emulation of floating-point matrix multiplication using an
// integer modular technique" (2025), restated here:
//
//  1. Scale: row i of A by 2^e_i and column j of B by 2^f_j, chosen from the
//     2-norms so that ||A_i 2^e_i||_2 <= 2^g and ||B_j 2^f_j||_2 <= 2^g; round
//     to integers Ab = rint(A 2^e), Bb = rint(B 2^f). By Cauchy-Schwarz every
//     |(Ab Bb)_ij| <= 2^(2g) (1+d)^2 < M/2, where M = prod m_l (pairwise coprime
//     moduli m_l <= 256), so Ab Bb is recovered exactly from its residues.
//  2. Residues: Ab mod m_l and Bb mod m_l in [-128, 127] (int8), one plane per
//     modulus; B planes are stored transposed (K contiguous) for the MFMA.
//  3. One INT8 GEMM per modulus with INT32 accumulation (|partial| <=
//     K * 2^14 < 2^31 for K <= 2^16 per accumulation chunk), reduced mod m_l
//     into a uint8 plane.
//  4. CRT: X = sum_l r_l w_l mod M with w_l = (M/m_l) * ((M/m_l)^-1 mod m_l),
//     in exact 128-bit integer arithmetic, mapped to (-M/2, M/2), converted to
//     f64 and scaled by 2^-(e_i + f_j); then alpha/beta.
// The only rounding errors are the integer roundings of step 1 (relative
// 2^-g per element against the vector norm) and the final conversion, so the
// error is set by g: with 14 moduli M ~ 2^111.5 and g = 55.
//
// Planes live in the context workspace: 14 x (M + N) x K int8 + 14 x M x N
// uint8 (2.7 GB at 8192^3).
#include <cfloat>
#include <cmath>
#include <cstdlib>

#include "gemm.hpp"

namespace rstsr {
namespace hip {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned __int128 u128;
typedef __int128 i128;

constexpr int OZ_MAXMOD = 14;  // X = sum r_l w_l < 14 * 256 * M < 2^128 needs M < 2^116
// largest pairwise coprime moduli <= 256
static const uint32_t OZ_MODULI[16] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193};

struct CrtConst {
    int n;
    uint32_t mod[OZ_MAXMOD];
    u128 w[OZ_MAXMOD];
    u128 M, Mhalf;
    double Md;
};

/* ---------------- vector norms -> scale exponents ---------------- */

// Exponent e with ||x 2^e||_2 <= 2^g given amax = max|x| and ss = sum (x 2^-E)^2,
// E = ilogb(amax). Zero vectors get e = 0.
__device__ __forceinline__ int scale_exponent(double amax, double ss, int g) {
    if (!(amax > 0.0)) return 0;
    int E = ilogb(amax);
    double l2 = (double)E + 0.5 * log2(ss);  // log2 of the norm
    return g - (int)ceil(l2 + 1e-12);
}

// One block per vector (vector elements strided by st): two passes, max then
// scaled sum of squares. flag[0] |= 1 on a non-finite element.
__global__ void __launch_bounds__(256) k_vnorm_block(const double* __restrict__ x, int64_t nvec, int64_t len,
                                                     int64_t sv, int64_t st, int g, int* __restrict__ ev,
                                                     int* __restrict__ flag) {
    __shared__ double red[256];
    const int64_t v = blockIdx.x;
    const double* p = x + v * sv;
    double m = 0.0;
    bool bad = false;
    for (int64_t t = threadIdx.x; t < len; t += 256) {
        double a = p[t * st];
        if (!isfinite(a)) bad = true;
        m = fmax(m, fabs(a));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double amax = red[0];
    __syncthreads();
    double ss = 0.0;
    if (amax > 0.0 && isfinite(amax)) {
        int E = ilogb(amax);
        for (int64_t t = threadIdx.x; t < len; t += 256) {
            double a = ldexp(p[t * st], -E);
            ss += a * a;
        }
    }
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (bad) atomicOr(flag, 1);
    if (threadIdx.x == 0) ev[v] = isfinite(amax) ? scale_exponent(amax, red[0], g) : 0;
}

// Vectors adjacent in memory (|sv| small): thread per vector, partial max over
// chunks of `chunk` elements into pm[c][v].
__global__ void __launch_bounds__(256) k_vnorm_cols_max(const double* __restrict__ x, int64_t nvec, int64_t len,
                                                        int64_t sv, int64_t st, int64_t chunk,
                                                        double* __restrict__ pm, int* __restrict__ flag) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (v >= nvec) return;
    const int64_t t0 = c * chunk, t1 = min(len, t0 + chunk);
    double m = 0.0;
    bool bad = false;
    for (int64_t t = t0; t < t1; ++t) {
        double a = x[v * sv + t * st];
        if (!isfinite(a)) bad = true;
        m = fmax(m, fabs(a));
    }
    if (bad) atomicOr(flag, 1);
    pm[c * nvec + v] = m;
}

__global__ void __launch_bounds__(256) k_vnorm_cols_ss(const double* __restrict__ x, int64_t nvec, int64_t len,
                                                       int64_t sv, int64_t st, int64_t chunk, int64_t nchunk,
                                                       const double* __restrict__ pm, double* __restrict__ ps) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (v >= nvec) return;
    double amax = 0.0;
    for (int64_t q = 0; q < nchunk; ++q) amax = fmax(amax, pm[q * nvec + v]);
    double ss = 0.0;
    if (amax > 0.0 && isfinite(amax)) {
        int E = ilogb(amax);
        const int64_t t0 = c * chunk, t1 = min(len, t0 + chunk);
        for (int64_t t = t0; t < t1; ++t) {
            double a = ldexp(x[v * sv + t * st], -E);
            ss += a * a;
        }
    }
    ps[c * nvec + v] = ss;
}

__global__ void __launch_bounds__(256) k_vnorm_cols_fin(int64_t nvec, int64_t nchunk, const double* __restrict__ pm,
                                                        const double* __restrict__ ps, int g, int* __restrict__ ev) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= nvec) return;
    double amax = 0.0, ss = 0.0;
    for (int64_t q = 0; q < nchunk; ++q) amax = fmax(amax, pm[q * nvec + v]);
    for (int64_t q = 0; q < nchunk; ++q) ss += ps[q * nvec + v];
    ev[v] = isfinite(amax) ? scale_exponent(amax, ss, g) : 0;
}

/* ---------------- split into residue planes ---------------- */

struct ModSet {
    int n;
    double m[OZ_MAXMOD];
    double inv[OZ_MAXMOD];
};

// symmetric residue of an integer-valued double x (|x| < 2^60) modulo m <= 256
__device__ __forceinline__ int sym_residue(double x, double m, double inv) {
    double q = floor(x * inv);
    double r = fma(-q, m, x);  // exact: the true value is a small integer
    if (r < 0.0) r += m;
    if (r >= m) r -= m;
    int ri = (int)r;
    return ri > 127 ? ri - (int)m : ri;
}

// Source S(v, t) = s[v*sv + t*st] for v < nv, t < len; planes P[l][v][t] for
// v < nvp, t < lenp (zero padded). 64 x 64 tiles staged through LDS so that
// reads follow the source's unit-stride axis and plane writes follow t.
__global__ void __launch_bounds__(256) k_split(const double* __restrict__ s, int64_t nv, int64_t len, int64_t sv,
                                               int64_t st, const int* __restrict__ ev, int64_t nvp, int64_t lenp,
                                               int8_t* __restrict__ planes, ModSet ms, int read_along_t,
                                               const int* __restrict__ skip, int64_t plane) {
    if (*skip) return;  // non-finite input: the plain kernel computes this product
    __shared__ double tile[64][65];
    const int64_t v0 = (int64_t)blockIdx.y * 64, t0 = (int64_t)blockIdx.x * 64;
    const int tid = threadIdx.x;
    // phase 1: load (scaled, rounded) values
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        int e = k * 256 + tid;
        int a, b;  // a fastest
        if (read_along_t) {
            a = e & 63;
            b = e >> 6;
        } else {
            b = e & 63;
            a = e >> 6;
        }
        // a indexes t when reading along t, else b indexes t
        int tt = read_along_t ? a : b, vv = read_along_t ? b : a;
        int64_t gv = v0 + vv, gt = t0 + tt;
        double x = 0.0;
        if (gv < nv && gt < len) x = rint(ldexp(s[gv * sv + gt * st], ev[gv]));
        tile[vv][tt] = x;
    }
    __syncthreads();
    // phase 2: each thread packs 4 consecutive t of one v, for every modulus
    const int vv = tid >> 2;       // 0..63
    const int tq = (tid & 3) * 16;  // 4 threads x 16 t = 64 t per row
    const int64_t gv = v0 + vv;
    if (gv >= nvp) return;
#pragma unroll
    for (int c = 0; c < 16; c += 4) {
        int64_t gt = t0 + tq + c;
        if (gt >= lenp) continue;
        double x0 = tile[vv][tq + c], x1 = tile[vv][tq + c + 1], x2 = tile[vv][tq + c + 2], x3 = tile[vv][tq + c + 3];
        for (int l = 0; l < ms.n; ++l) {
            uint32_t b0 = (uint8_t)(int8_t)sym_residue(x0, ms.m[l], ms.inv[l]);
            uint32_t b1 = (uint8_t)(int8_t)sym_residue(x1, ms.m[l], ms.inv[l]);
            uint32_t b2 = (uint8_t)(int8_t)sym_residue(x2, ms.m[l], ms.inv[l]);
            uint32_t b3 = (uint8_t)(int8_t)sym_residue(x3, ms.m[l], ms.inv[l]);
            uint32_t packed = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
            *(uint32_t*)(planes + l * plane + (size_t)gv * lenp + gt) = packed;
        }
    }
}

/* ---------------- INT8 MFMA GEMM per modulus ---------------- */

constexpr int I8_BM = 128, I8_BN = 128, I8_BK = 128, I8_PAD = 16;

__device__ __forceinline__ uint32_t xcd_remap8(uint32_t t, uint32_t ntiles) {
    if (ntiles % 8 != 0) return t;
    return (t % 8) * (ntiles / 8) + t / 8;
}

// A8: [Mp][Kp], B8: [Np][Kp] (both K contiguous, zero padded); R: [M][ldr] uint8
// residues of (A8 B8^T) mod `mod`. 256 threads = 4 waves (2x2), 64x64 per wave
// from 4x4 v_mfma_i32_16x16x64_i8 tiles; LDS double buffer, one barrier per slice.
// Batches on blockIdx.y: plane offsets sA / sB (0 for a broadcast operand) and sR.
template <bool ROW4G>
__global__ void __launch_bounds__(256) k_i8gemm_mod(const int8_t* __restrict__ A, const int8_t* __restrict__ B,
                                                    uint8_t* __restrict__ R, int64_t M, int64_t N, int64_t Kp,
                                                    int64_t ldr, int mod, uint32_t tiles_n, uint32_t ntiles,
                                                    int64_t sA, int64_t sB, int64_t sR,
                                                    const int* __restrict__ skip) {
    if (*skip) return;
    A += (int64_t)blockIdx.y * sA;
    B += (int64_t)blockIdx.y * sB;
    R += (int64_t)blockIdx.y * sR;
    __shared__ __attribute__((aligned(16))) int8_t As[2][I8_BM][I8_BK + I8_PAD];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2][I8_BN][I8_BK + I8_PAD];
    const uint32_t tile = xcd_remap8(blockIdx.x, ntiles);
    const uint32_t tm = tile / tiles_n, tn = tile - (tile / tiles_n) * tiles_n;
    const int8_t* Ab = A + (size_t)tm * I8_BM * Kp;
    const int8_t* Bb = B + (size_t)tn * I8_BN * Kp;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = tid >> 3, lc = (tid & 7) * 16;  // loader: 8 threads x 16 B per 128 B row, 32 rows per pass

    v4i ra[4], rb[4];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ra[i] = *(const v4i*)(Ab + (size_t)(lr + 32 * i) * Kp + k0 + lc);
            rb[i] = *(const v4i*)(Bb + (size_t)(lr + 32 * i) * Kp + k0 + lc);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            *(v4i*)&As[buf][lr + 32 * i][lc] = ra[i];
            *(v4i*)&Bs[buf][lr + 32 * i][lc] = rb[i];
        }
    };
    v4i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4i{0, 0, 0, 0};

    load(0);
    store(0);
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 16;
    int buf = 0;
    for (int64_t k0 = 0; k0 < Kp; k0 += I8_BK) {
        const bool more = k0 + I8_BK < Kp;
        if (more) load(k0 + I8_BK);
#pragma unroll
        for (int kk = 0; kk < I8_BK; kk += 64) {
            v4i af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = *(const v4i*)&As[buf][wm * 64 + i * 16 + fr][kk + fk];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = *(const v4i*)&Bs[buf][wn * 64 + j * 16 + fr][kk + fk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        // keep |acc| < 2^31: reduce every 2^16 of K (each step adds < 128*128*128 per element)
        if (((k0 + I8_BK) & 65535) == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[i][j][r] %= mod;
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int lrow = ROW4G ? (4 * (lane >> 4) + r) : (4 * r + (lane >> 4));
                int64_t gm = (int64_t)tm * I8_BM + wm * 64 + i * 16 + lrow;
                int64_t gn = (int64_t)tn * I8_BN + wn * 64 + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    int v = acc[i][j][r] % mod;
                    if (v < 0) v += mod;
                    R[gm * ldr + gn] = (uint8_t)v;
                }
            }
}

// zero n ints (a kernel rather than a memset node, so the sequence can be graph-captured)
__global__ void k_zero_i32(int* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

/* ---------------- CRT reconstruction ---------------- */

__device__ __forceinline__ double u128_to_double(u128 x) {
    uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
    return ldexp((double)hi, 64) + (double)lo;
}

// R planes: [l][batch][Mp][N] (plane stride pl); eA [nbA][Mp], fB [nbB][Np].
__global__ void __launch_bounds__(256) k_crt(const uint8_t* __restrict__ R, int64_t M, int64_t N, int64_t Mp,
                                             int64_t Np, uint64_t nbatch, int64_t pl, int a_batched, int b_batched,
                                             BatchDesc bd, CrtConst cc, const int* __restrict__ eA,
                                             const int* __restrict__ fB, double* C, int64_t rsc, int64_t csc,
                                             double alpha, double beta, int beta_zero, const int* __restrict__ skip) {
    const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t mn = (uint64_t)M * N;
    if (idx >= mn * nbatch || *skip) return;
    const uint64_t b = idx / mn;
    const uint64_t e = idx - b * mn;
    const int64_t i = (int64_t)(e / N), j = (int64_t)(e - (e / N) * N);
    const uint64_t r_off = (b * Mp + i) * N + j;
    u128 X = 0;
#pragma unroll
    for (int l = 0; l < OZ_MAXMOD; ++l) {
        if (l >= cc.n) break;
        X += (u128)R[l * pl + r_off] * cc.w[l];
    }
    // X < n * 256 * M < 2^128: reduce mod M with a floating quotient estimate
    double qd = floor(u128_to_double(X) / cc.Md);
    if (qd < 0.0) qd = 0.0;
    u128 qM = (u128)(uint64_t)qd * cc.M;
    i128 Y = (i128)(X - qM);  // X - qM in (-M, 2M)
    if (qM > X) Y = -(i128)(qM - X);
    if (Y < 0) Y += (i128)cc.M;
    if (Y >= (i128)cc.M) Y -= (i128)cc.M;
    if (Y >= (i128)cc.M - Y) Y -= (i128)cc.M;  // symmetric range (-M/2, M/2]
    double v = Y < 0 ? -u128_to_double((u128)(-Y)) : u128_to_double((u128)Y);
    v = ldexp(v, -(eA[(a_batched ? b : 0) * Mp + i] + fB[(b_batched ? b : 0) * Np + j]));
    int64_t oa, ob, oc;
    batch_offsets(bd, (uint32_t)b, oa, ob, oc);
    double* cp = C + oc + i * rsc + j * csc;
    double out = alpha * v;
    if (!beta_zero) out += beta * (*cp);
    *cp = out;
}

/* ---------------- host ---------------- */

static u128 mul_mod(u128 a, uint64_t b, u128 m) {
    // (a * b) mod m for a < m < 2^127, b < 2^64: double-and-add over b's bits
    u128 r = 0;
    a %= m;
    while (b) {
        if (b & 1) {
            r += a;
            if (r >= m) r -= m;
        }
        a <<= 1;
        if (a >= m) a -= m;
        b >>= 1;
    }
    return r;
}

static uint64_t inv_mod(uint64_t a, uint64_t m) {
    // extended Euclid; a and m coprime
    int64_t t = 0, nt = 1, r = (int64_t)m, nr = (int64_t)(a % m);
    while (nr != 0) {
        int64_t q = r / nr;
        int64_t tmp = t - q * nt;
        t = nt;
        nt = tmp;
        tmp = r - q * nr;
        r = nr;
        nr = tmp;
    }
    if (t < 0) t += (int64_t)m;
    return (uint64_t)t;
}

static CrtConst make_crt(int n) {
    CrtConst cc{};
    cc.n = n;
    u128 M = 1;
    for (int l = 0; l < n; ++l) {
        cc.mod[l] = OZ_MODULI[l];
        M *= OZ_MODULI[l];
    }
    cc.M = M;
    cc.Mhalf = M / 2;
    cc.Md = (double)(uint64_t)(M >> 64) * 18446744073709551616.0 + (double)(uint64_t)M;
    for (int l = 0; l < n; ++l) {
        u128 Ml = M / OZ_MODULI[l];
        uint64_t r = (uint64_t)(Ml % OZ_MODULI[l]);
        uint64_t inv = inv_mod(r, OZ_MODULI[l]);
        cc.w[l] = mul_mod(Ml, inv, M);
    }
    return cc;
}

static int ozaki_moduli_count() {
    const char* s = std::getenv("RSTSR_HIP_OZAKI_MODULI");
    int n = s ? std::atoi(s) : OZ_MAXMOD;
    if (n < 2 || n > OZ_MAXMOD) RSTSR_RAISE(InvalidValue, "RSTSR_HIP_OZAKI_MODULI must be in [2, 14]");
    return n;
}

// Output row layout of v_mfma_i32_16x16x64_i8 (true: register r of lane l holds
// row 4*(l/16) + r).
static constexpr bool I8_ROW4G = true;

void ozaki_gemm_f64(rstsr_hip_ctx* ctx, const GemmProblem& P, double alpha, double beta, bool beta_zero) {
    const int n = ozaki_moduli_count();
    CrtConst cc = make_crt(n);
    const double log2M = std::log2(cc.Md);
    // Cauchy-Schwarz: (2^g (1 + d))^2 < M/2 with a 1/2-bit margin; exactness of x = rint(A 2^e) as an
    // integer-valued double and of the residue arithmetic needs g <= 60
    int g = (int)std::floor((log2M - 1.5) / 2.0);
    if (g > 60) g = 60;
    ModSet ms{};
    ms.n = n;
    for (int l = 0; l < n; ++l) {
        ms.m[l] = (double)cc.mod[l];
        ms.inv[l] = 1.0 / (double)cc.mod[l];
    }
    const int64_t M = P.M, N = P.N, K = P.K;
    if (K > (int64_t)1 << 40) RSTSR_RAISE(UnImplemented, "K too large for the split GEMM");
    const int64_t Mp = (M + I8_BM - 1) / I8_BM * I8_BM, Np = (N + I8_BN - 1) / I8_BN * I8_BN;
    const int64_t Kp = (K + I8_BK - 1) / I8_BK * I8_BK;
    const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(64, K / 256));
    const int64_t chunk = (K + nchunk - 1) / nchunk;
    const uint64_t nb = P.nbatch;
    if (nb >= 65536) RSTSR_RAISE(UnImplemented, "too many batches for the split GEMM");
    // an operand whose batch strides are all zero is split once and shared by every batch
    bool a_b = false, b_b = false;
    for (int d = 0; d < P.bd.nbd; ++d) {
        if (P.bd.sa[d] != 0) a_b = true;
        if (P.bd.sb[d] != 0) b_b = true;
    }
    const uint64_t nbA = a_b ? nb : 1, nbB = b_b ? nb : 1;
    // workspace layout: flag | eA [nbA][Mp] | fB [nbB][Np] | norm partials | A8 [l][nbA][Mp][Kp] |
    // B8 [l][nbB][Np][Kp] | R [l][nb][Mp][N]
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_flag = 0;
    off = sizeof(int);
    const size_t o_eA = off;
    off += nbA * Mp * sizeof(int);
    const size_t o_fB = off;
    off += nbB * Np * sizeof(int);
    const size_t n_zero = off / sizeof(int);
    off = (off + 255) & ~(size_t)255;
    const size_t o_pm = take((size_t)nchunk * std::max(Mp, Np) * sizeof(double));
    const size_t o_ps = take((size_t)nchunk * std::max(Mp, Np) * sizeof(double));
    const size_t planeA = (size_t)nbA * Mp * Kp, planeB = (size_t)nbB * Np * Kp, planeR = (size_t)nb * Mp * N;
    const size_t o_A8 = take((size_t)n * planeA);
    const size_t o_B8 = take((size_t)n * planeB);
    const size_t o_R = take((size_t)n * planeR);
    char* ws = (char*)workspace(ctx, off);
    int* flag = (int*)(ws + o_flag);
    int* eA = (int*)(ws + o_eA);
    int* fB = (int*)(ws + o_fB);
    double* pm = (double*)(ws + o_pm);
    double* ps = (double*)(ws + o_ps);
    int8_t* A8 = (int8_t*)(ws + o_A8);
    int8_t* B8 = (int8_t*)(ws + o_B8);
    uint8_t* R = (uint8_t*)(ws + o_R);
    const uint32_t tiles_n = (uint32_t)(Np / I8_BN), ntiles = (uint32_t)((Mp / I8_BM) * tiles_n);

    hipLaunchKernelGGL(k_zero_i32, dim3(ceil_div_u32(n_zero, 256)), dim3(256), 0, ctx->stream, flag, (int64_t)n_zero);
    check_launch("ozaki zero");
    // vector norms: rows of A (elements along k: stride csa), columns of B (along k: stride rsb)
    auto vnorm = [&](const double* x, int64_t nvec, int64_t len, int64_t sv, int64_t st, int* ev) {
        if (std::llabs(st) == 1 || std::llabs(sv) > std::llabs(st)) {
            hipLaunchKernelGGL(k_vnorm_block, dim3((uint32_t)nvec), dim3(256), 0, ctx->stream, x, nvec, len, sv, st, g,
                               ev, flag);
        } else {
            dim3 grid(ceil_div_u32(nvec, 256), (uint32_t)nchunk);
            hipLaunchKernelGGL(k_vnorm_cols_max, grid, dim3(256), 0, ctx->stream, x, nvec, len, sv, st, chunk, pm, flag);
            hipLaunchKernelGGL(k_vnorm_cols_ss, grid, dim3(256), 0, ctx->stream, x, nvec, len, sv, st, chunk, nchunk, pm,
                               ps);
            hipLaunchKernelGGL(k_vnorm_cols_fin, dim3(ceil_div_u32(nvec, 256)), dim3(256), 0, ctx->stream, nvec, nchunk,
                               pm, ps, g, ev);
        }
        check_launch("ozaki norms");
    };
    for (uint64_t b = 0; b < nbA; ++b) {
        int64_t oa, ob, oc;
        batch_offsets_host(P.bd, b, oa, ob, oc);
        const double* A = (const double*)P.a + oa;
        vnorm(A, M, K, P.rsa, P.csa, eA + b * Mp);
        dim3 ga((uint32_t)(Kp / 64), (uint32_t)(Mp / 64));
        hipLaunchKernelGGL(k_split, ga, dim3(256), 0, ctx->stream, A, M, K, P.rsa, P.csa, eA + b * Mp, Mp, Kp,
                           A8 + b * Mp * Kp, ms, std::llabs(P.csa) == 1 ? 1 : 0, flag, (int64_t)planeA);
        check_launch("ozaki split A");
    }
    for (uint64_t b = 0; b < nbB; ++b) {
        int64_t oa, ob, oc;
        batch_offsets_host(P.bd, b, oa, ob, oc);
        const double* B = (const double*)P.b + ob;
        vnorm(B, N, K, P.csb, P.rsb, fB + b * Np);
        dim3 gb((uint32_t)(Kp / 64), (uint32_t)(Np / 64));
        hipLaunchKernelGGL(k_split, gb, dim3(256), 0, ctx->stream, B, N, K, P.csb, P.rsb, fB + b * Np, Np, Kp,
                           B8 + b * Np * Kp, ms, std::llabs(P.rsb) == 1 ? 1 : 0, flag, (int64_t)planeB);
        check_launch("ozaki split B");
    }
    // Non-finite inputs set *flag: the split kernels exit at once and the plain
    // FP64 kernel (which runs only when *flag != 0) computes C with IEEE special
    // values. No host round trip, so the sequence can be graph-captured.
    plain_dgemm(ctx, P, alpha, beta, beta_zero, flag);
    for (int l = 0; l < n; ++l) {
        hipLaunchKernelGGL(k_i8gemm_mod<I8_ROW4G>, dim3(ntiles, (uint32_t)nb), dim3(256), 0, ctx->stream,
                           A8 + (size_t)l * planeA, B8 + (size_t)l * planeB, R + (size_t)l * planeR, Mp, N, Kp, N,
                           (int)cc.mod[l], tiles_n, ntiles, a_b ? Mp * Kp : (int64_t)0, b_b ? Np * Kp : (int64_t)0,
                           Mp * N, flag);
        check_launch("ozaki int8 gemm");
    }
    uint64_t tot = (uint64_t)M * N * nb;
    if (tot >= (1ull << 40)) RSTSR_RAISE(UnImplemented, "split GEMM output too large");
    hipLaunchKernelGGL(k_crt, dim3(ceil_div_u32(tot, 256)), dim3(256), 0, ctx->stream, R, M, N, Mp, Np, nb,
                       (int64_t)planeR, a_b ? 1 : 0, b_b ? 1 : 0, P.bd, cc, eA, fB, (double*)P.c, P.rsc, P.csc, alpha,
                       beta, beta_zero ? 1 : 0, flag);
    check_launch("ozaki crt");
}

}  // namespace hip
}  // namespace rstsr
