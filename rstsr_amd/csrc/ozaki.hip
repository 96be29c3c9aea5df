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
// error is set by g: with 14 moduli M ~ 2^110.2 and g = 54.
//
// Planes live in the context workspace: 14 x (M + N) x K int8 + 14 x M x N
// uint8 (2.7 GB at 8192^3).
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "gemm.hpp"

namespace rstsr {
namespace hip {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef unsigned __int128 u128;
typedef __int128 i128;

constexpr int OZ_MAXMOD = 14;  // X = sum r_l w_l < 14 * 256 * M < 2^128 needs M < 2^116
// largest pairwise coprime moduli <= 256
static const uint32_t OZ_MODULI[16] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193};
constexpr uint32_t OZ_MODULI_C[OZ_MAXMOD] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199};

// CRT constants: w_l = (M/m_l) ((M/m_l)^-1 mod m_l) and M = prod m_l, both
// also in 32-bit limbs (least significant first) for the device sums.
struct CrtConst {
    int n;
    uint32_t mod[OZ_MAXMOD];
    double wl[OZ_MAXMOD][3];  // w_l = wl[0] 2^76 + wl[1] 2^38 + wl[2], limbs < 2^38
    double Ml[3];             // M in the same limbs
    u128 M;
    double Md, invM;
};

// The CRT limbs of the default 14 moduli as compile-time constants: kernels that
// use them keep no constant table in registers (the 42 limbs would not fit the
// scalar registers next to the rest).
struct CrtLimbs {
    double wl[OZ_MAXMOD][3];
    double Ml[3];
    double invM;
};
constexpr CrtLimbs make_crt_limbs14() {
    CrtLimbs k{};
    u128 M = 1;
    for (int l = 0; l < OZ_MAXMOD; ++l) M *= OZ_MODULI_C[l];
    const u128 L38 = ((u128)1 << 38) - 1;
    for (int l = 0; l < OZ_MAXMOD; ++l) {
        const uint32_t m = OZ_MODULI_C[l];
        const u128 Mm = M / m;
        const uint32_t rm = (uint32_t)(Mm % m);
        uint32_t inv = 1;
        while ((rm * inv) % m != 1) ++inv;
        const u128 w = Mm * inv % M;  // Mm * inv < 2^8 M < 2^128
        k.wl[l][0] = (double)(uint64_t)(w >> 76);
        k.wl[l][1] = (double)(uint64_t)((w >> 38) & L38);
        k.wl[l][2] = (double)(uint64_t)(w & L38);
    }
    k.Ml[0] = (double)(uint64_t)(M >> 76);
    k.Ml[1] = (double)(uint64_t)((M >> 38) & L38);
    k.Ml[2] = (double)(uint64_t)(M & L38);
    k.invM = 1.0 / ((double)(uint64_t)(M >> 64) * 18446744073709551616.0 + (double)(uint64_t)M);
    return k;
}
constexpr CrtLimbs CRT14 = make_crt_limbs14();

/* ---------------- vector norms -> scale exponents ---------------- */

// Running (max |x|, E = ilogb(max), ss = sum (x 2^-E)^2) in one pass: when the
// max moves to a new binade the sum is rescaled by a power of two (exact, or
// the old terms underflow and are negligible). Non-finite elements are skipped
// and reported.
struct NormAcc {
    double m;
    int E;
    double ss;
};

__device__ __forceinline__ void norm_add(NormAcc& a, double x, bool& bad) {
    if (!isfinite(x)) {
        bad = true;
        return;
    }
    const double ax = fabs(x);
    if (ax > a.m) {
        const int En = ilogb(ax);
        if (a.m > 0.0) a.ss = ldexp(a.ss, 2 * (a.E - En));
        a.E = En;
        a.m = ax;
    }
    const double t = ldexp(x, -a.E);
    a.ss += t * t;
}

// Commutative, so every lane of a butterfly ends with the same value.
__device__ __forceinline__ void norm_merge(NormAcc& a, const NormAcc& b) {
    if (!(b.m > 0.0)) return;
    if (!(a.m > 0.0)) {
        a = b;
        return;
    }
    const int E = max(a.E, b.E);
    a.ss = ldexp(a.ss, 2 * (a.E - E)) + ldexp(b.ss, 2 * (b.E - E));
    a.E = E;
    a.m = fmax(a.m, b.m);
}

__device__ __forceinline__ NormAcc wave_norm(NormAcc a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        NormAcc b{__shfl_xor(a.m, o), __shfl_xor(a.E, o), __shfl_xor(a.ss, o)};
        norm_merge(a, b);
    }
    return a;
}

// Exponent e with ||x 2^e||_2 <= 2^g; zero vectors get e = 0.
__device__ __forceinline__ int scale_exponent(const NormAcc& a, int g) {
    if (!(a.m > 0.0)) return 0;
    const double l2 = (double)a.E + 0.5 * log2(a.ss);  // log2 of the norm
    return g - (int)ceil(l2 + 1e-12);
}

// Scale exponents of nvec vectors of length len (element t of vector v at
// x[v*sv + t*st]) for every batch b = blockIdx.y of one operand (`which`: 0 = A,
// 1 = B offsets of bd); exponents to ev[b*evs + v]. One wave per vector: for
// vectors whose elements are contiguous.
__global__ void __launch_bounds__(256) k_norm_rows(const double* __restrict__ x, int64_t nvec, int64_t len,
                                                   int64_t sv, int64_t st, BatchDesc bd, int which, int64_t evs, int g,
                                                   int* __restrict__ ev, int* __restrict__ flag) {
    const int lane = threadIdx.x & 63;
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t b = blockIdx.y;
    if (v >= nvec) return;  // uniform across the wave
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const double* xv = x + (which ? ob : oa) + v * sv;
    NormAcc a{0.0, 0, 0.0};
    bool bad = false;
    for (int64_t t = lane; t < len; t += 64) norm_add(a, xv[t * st], bad);
    a = wave_norm(a);
    if (bad) atomicOr(flag, 1);
    if (lane == 0) ev[(int64_t)b * evs + v] = scale_exponent(a, g);
}

// The same for vectors side by side: a lane per vector (64 per block), the four
// waves and gridDim.y chunks of `chunk` elements split the length, batch on
// blockIdx.z. With several chunks the partials go to part[(b*nchunk + c)*nvec + v].
__global__ void __launch_bounds__(256) k_norm_cols(const double* __restrict__ x, int64_t nvec, int64_t len,
                                                   int64_t sv, int64_t st, int64_t chunk, BatchDesc bd, int which,
                                                   int64_t evs, int g, NormAcc* __restrict__ part,
                                                   int* __restrict__ ev, int* __restrict__ flag) {
    __shared__ NormAcc red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t v = (int64_t)blockIdx.x * 64 + lane;
    const uint32_t c = blockIdx.y, b = blockIdx.z;
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const double* xb = x + (which ? ob : oa);
    NormAcc a{0.0, 0, 0.0};
    bool bad = false;
    if (v < nvec) {
        // four independent accumulators over rows t, t + 4, t + 8, t + 12 (four loads in flight,
        // four norm chains instead of one), merged in a fixed order
        const int64_t t1 = min(len, (int64_t)(c + 1) * chunk);
        const double* xp = xb + v * sv;
        NormAcc a1{0.0, 0, 0.0}, a2{0.0, 0, 0.0}, a3{0.0, 0, 0.0};
        int64_t t = (int64_t)c * chunk + w;
        for (; t + 12 < t1; t += 16) {
            const double x0 = xp[t * st], x1 = xp[(t + 4) * st], x2 = xp[(t + 8) * st], x3 = xp[(t + 12) * st];
            norm_add(a, x0, bad);
            norm_add(a1, x1, bad);
            norm_add(a2, x2, bad);
            norm_add(a3, x3, bad);
        }
        for (; t < t1; t += 4) norm_add(a, xp[t * st], bad);
        norm_merge(a, a1);
        norm_merge(a, a2);
        norm_merge(a, a3);
    }
    red[w][lane] = a;
    __syncthreads();
    if (bad) atomicOr(flag, 1);
    if (w != 0 || v >= nvec) return;
#pragma unroll
    for (int q = 1; q < 4; ++q) norm_merge(a, red[q][lane]);
    if (gridDim.y == 1) ev[(int64_t)b * evs + v] = scale_exponent(a, g);
    else part[((int64_t)b * gridDim.y + c) * nvec + v] = a;
}

__global__ void __launch_bounds__(256) k_norm_fin(const NormAcc* __restrict__ part, int64_t nvec, int nchunk,
                                                  int64_t evs, int g, int* __restrict__ ev) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t b = blockIdx.y;
    if (v >= nvec) return;
    NormAcc a = part[(int64_t)b * nchunk * nvec + v];
    for (int c = 1; c < nchunk; ++c) norm_merge(a, part[((int64_t)b * nchunk + c) * nvec + v]);
    ev[(int64_t)b * evs + v] = scale_exponent(a, g);
}

/* ---------------- split into residue planes ---------------- */

struct ModSet {
    int n;
    double m[OZ_MAXMOD];
    uint32_t mi[OZ_MAXMOD];     // m
    uint32_t c16[OZ_MAXMOD];    // 2^16 mod m, for the INT8 epilogue
    float inv[OZ_MAXMOD];       // f32(1/m) (1 - 2^-22), for the INT8 epilogue
    uint32_t clo[OZ_MAXMOD];    // bytes 256^k mod m, k = 0..3
    uint32_t chi[OZ_MAXMOD];    // bytes 256^k mod m, k = 4..7
    uint32_t off[OZ_MAXMOD];    // m - (2^56 mod m) + 128
    uint32_t mg20[OZ_MAXMOD];   // ceil(2^20 / m)
};

// Per-modulus split constants (the ModSet fields above) of one modulus.
struct ResConst {
    uint32_t clo, chi, off, mg20, m;
};
constexpr ResConst make_res_const(uint32_t m) {
    ResConst c{0, 0, 0, 0, m};
    uint32_t pw = 1 % m;  // 256^k mod m
    for (int k = 0; k < 8; ++k) {
        if (k < 4) c.clo |= pw << (8 * k);
        else c.chi |= pw << (8 * (k - 4));
        pw = (pw * 256) % m;
    }
    // pw = 256^8 mod m here; 2^56 mod m = 256^7 mod m
    uint32_t p56 = 1 % m;
    for (int k = 0; k < 7; ++k) p56 = (p56 * 256) % m;
    c.off = m - p56 + 128;
    c.mg20 = (uint32_t)(((1u << 20) + m - 1) / m);
    return c;
}
struct ResTable {
    ResConst c[OZ_MAXMOD];
};
constexpr ResTable make_res_table14() {
    ResTable t{};
    for (int l = 0; l < OZ_MAXMOD; ++l) t.c[l] = make_res_const(OZ_MODULI_C[l]);
    return t;
}
// the default 14 moduli's constants at compile time
constexpr ResTable RES14 = make_res_table14();

// A split integer x (|x| <= ||x||_2 <= 2^g, g <= 55) as u = x + 2^56 in [2^55, 3 2^55]:
// eight bytes in two words, u = hi 2^32 + lo.
struct Digits {
    uint32_t lo, hi;
};
__device__ __forceinline__ Digits to_digits(double x) {
    const double h = floor(x * 0x1p-32);  // in [-2^23, 2^23]
    const double l = fma(-h, 0x1p32, x);  // exact, in [0, 2^32)
    return {(uint32_t)l, (uint32_t)((int)h + (1 << 24))};
}

// Residue byte of x modulo m (res_reduce of the sum s that pack_residues forms from the digits
// of u = x + 2^56): t in [0, m) with t = x + 128 (mod m), so that t - 128 in
// [-128, m - 128) is an int8 representative of x mod m, stored as the byte t ^ 0x80.
// s = sum_k b_k (256^k mod m) + m - (2^56 mod m) + 128 = u - 2^56 + 128 = x + 128 (mod m), with
// s < 8 * 255^2 + 384 < 2^19 (two byte dot products, the offset as their accumulator). The
// quotient from 24-bit products: s * ceil(2^20 / m) < 2^32, and its floor / 2^20 exceeds
// floor(s / m) by at most one (the error s (ceil(2^20/m) m - 2^20) / (m 2^20) < s / 2^20 < 1/2),
// so t = s - q m lies in [-m, m) and one correction ends it: min(t, t + m) as unsigned words
// (a negative t wraps above t + m). q < 2^12 is the bit field [20, 32) of the product, and
// t = s + q (-m) one 24-bit multiply-add.
__device__ __forceinline__ uint32_t res_reduce(uint32_t s, const ResConst& c) {
    const int q = (int)__builtin_amdgcn_ubfe(__umul24(s, c.mg20), 20, 12);
    const int t = __mul24(q, -(int)c.m) + (int)s;
    return min((uint32_t)t, (uint32_t)(t + (int)c.m));
}

template <bool K14>
__device__ __forceinline__ ResConst res_const(const ModSet& ms, int l) {
    if constexpr (K14) return RES14.c[l];
    else return ResConst{ms.clo[l], ms.chi[l], ms.off[l], ms.mg20[l], ms.mi[l]};
}

// four residue bytes in one word, each stage over the four before the next (independent
// dot products back to back instead of dependent ones, which wait on each other)
__device__ __forceinline__ uint32_t pack_residues(const Digits* d, const ResConst& c) {
    uint32_t s[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = __builtin_amdgcn_udot4(d[i].lo, c.clo, c.off, false);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = __builtin_amdgcn_udot4(d[i].hi, c.chi, s[i], false);
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) w |= res_reduce(s[i], c) << (8 * i);
    return w ^ 0x80808080u;
}

// f(l) for every modulus: unrolled with compile-time constants for the default 14
template <bool K14, class F>
__device__ __forceinline__ void for_moduli(const ModSet& ms, F&& f) {
    if constexpr (K14) {
#pragma unroll
        for (int l = 0; l < OZ_MAXMOD; ++l) f(l);
    } else {
        for (int l = 0; l < ms.n; ++l) f(l);
    }
}

// Tile (v0, t0) of the split: source S(v, t) = s[v*sv + t*st] for v < nv,
// t < len; planes P[l][v][t] for v < nvp, t < lenp (zero padded; lenp a
// multiple of 64). The 64 x 64 tile is staged through LDS (`tile`) so that
// reads run along the source's unit-stride axis and each thread writes 16
// consecutive plane bytes per modulus. split_load fetches a thread's 16 source
// elements (zero outside the source), split_store scales them into the tile and
// writes the residues; split_tile is the two in turn.
__device__ __forceinline__ void split_load(const double* __restrict__ s, int64_t nv, int64_t len, int64_t sv,
                                           int64_t st, bool read_along_t, int64_t v0, int64_t t0, double (&x)[16]) {
    // element k of this thread: the fast index tid % 64 along the unit-stride axis, the slow one
    // 4k + tid / 64; its address is a uniform base plus a per-lane offset plus k steps of a
    // uniform stride (no per-element 64-bit address kept live)
    const int tid = threadIdx.x, fast = tid & 63, slow0 = tid >> 6;
    const int64_t lane = read_along_t ? (int64_t)slow0 * sv + (int64_t)fast * st : (int64_t)fast * sv + (int64_t)slow0 * st;
    const int64_t kstep = 4 * (read_along_t ? sv : st);
    const double* base = s + v0 * sv + t0 * st;
    if (v0 + 64 <= nv && t0 + 64 <= len) {
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = base[k * kstep + lane];
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int tt = read_along_t ? fast : 4 * k + slow0, vv = read_along_t ? 4 * k + slow0 : fast;
            x[k] = v0 + vv < nv && t0 + tt < len ? base[k * kstep + lane] : 0.0;
        }
    }
}

template <bool K14>
__device__ __forceinline__ void split_store(const double (&x)[16], int64_t nv, const int* __restrict__ ev, int64_t nvp,
                                            int64_t lenp, int8_t* __restrict__ planes, const ModSet& ms,
                                            bool read_along_t, int64_t plane, int64_t v0, int64_t t0,
                                            double (&tile)[64][65]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = k * 256 + tid, fast = e & 63, slow = e >> 6;
        const int tt = read_along_t ? fast : slow, vv = read_along_t ? slow : fast;
        const int64_t gv = v0 + vv;
        tile[vv][tt] = gv < nv ? rint(ldexp(x[k], ev[gv])) : 0.0;
    }
    __syncthreads();
    const int vv = tid >> 2, tq = (tid & 3) * 16;
    const int64_t gv = v0 + vv;
    if (gv < nvp) {
        Digits dg[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) dg[i] = to_digits(tile[vv][tq + i]);
        int8_t* dst = planes + gv * lenp + t0 + tq;
        for_moduli<K14>(ms, [&](int l) {
            const ResConst c = res_const<K14>(ms, l);
            const uint4 w{pack_residues(dg, c), pack_residues(dg + 4, c), pack_residues(dg + 8, c),
                          pack_residues(dg + 12, c)};
            *(uint4*)(dst + l * plane) = w;
        });
    }
}

template <bool K14>
__device__ __forceinline__ void split_tile(const double* __restrict__ s, int64_t nv, int64_t len, int64_t sv,
                                           int64_t st, const int* __restrict__ ev, int64_t nvp, int64_t lenp,
                                           int8_t* __restrict__ planes, const ModSet& ms, bool read_along_t,
                                           int64_t plane, int64_t v0, int64_t t0, double (&tile)[64][65]) {
    double x[16];
    split_load(s, nv, len, sv, st, read_along_t, v0, t0, x);
    split_store<K14>(x, nv, ev, nvp, lenp, planes, ms, read_along_t, plane, v0, t0, tile);
}

// Batch b = blockIdx.z of one operand (`which`: 0 = A, 1 = B offsets of bd):
// exponents ev + b*nvp, planes + b*nvp*lenp. A block splits tpb tiles along t,
// the next tile's source loads in flight while it computes the residues of the
// one before: t tile group blockIdx.x and vector tile blockIdx.y, or the other
// way round when `swap`.
template <bool K14>
__global__ void __launch_bounds__(256) k_split(const double* __restrict__ s, int64_t nv, int64_t len, int64_t sv,
                                               int64_t st, BatchDesc bd, int which, const int* __restrict__ ev,
                                               int64_t nvp, int64_t lenp, int8_t* __restrict__ planes, ModSet ms,
                                               int read_along_t, const int* __restrict__ skip, int64_t plane,
                                               int tpb, int swap) {
    if (*skip) return;  // non-finite input: the plain kernel computes this product
    __shared__ double tile[64][65];
    const uint32_t b = blockIdx.z;
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const double* sb = s + (which ? ob : oa);
    const int* evb = ev + (int64_t)b * nvp;
    int8_t* pb = planes + (int64_t)b * nvp * lenp;
    const bool along = read_along_t != 0;
    const uint32_t bx = swap ? blockIdx.y : blockIdx.x, by = swap ? blockIdx.x : blockIdx.y;
    const int64_t v0 = (int64_t)by * 64, ta = (int64_t)bx * tpb * 64;
    const int64_t te = min(lenp, ta + (int64_t)tpb * 64);
    double x[16];
    split_load(sb, nv, len, sv, st, along, v0, ta, x);
    for (int64_t t0 = ta; t0 < te; t0 += 64) {
        double xn[16];
        const bool more = t0 + 64 < te;
        if (more) split_load(sb, nv, len, sv, st, along, v0, t0 + 64, xn);
        split_store<K14>(x, nv, evb, nvp, lenp, pb, ms, along, plane, v0, t0, tile);
        __syncthreads();  // the tile is rewritten for the next t0
        if (more) {
#pragma unroll
            for (int k = 0; k < 16; ++k) x[k] = xn[k];
        }
    }
}

// Norm and split fused, for vectors whose elements are contiguous (st = 1): one wave per
// vector v < nvp of batch b = blockIdx.y. Pass 1 computes the scale exponent (ev[b*nvp + v])
// exactly as k_norm_rows does; pass 2 re-reads the vector (from cache) and writes its 16-byte
// residue runs to every plane like split_tile. Padding rows v >= nv and positions t >= len
// are zero.
template <bool K14>
__global__ void __launch_bounds__(256) k_normsplit_rows(const double* __restrict__ x, int64_t nv, int64_t len,
                                                        int64_t sv, BatchDesc bd, int which, int64_t nvp,
                                                        int64_t lenp, int g, int* __restrict__ ev,
                                                        int8_t* __restrict__ planes, int64_t plane, ModSet ms,
                                                        int* __restrict__ flag) {
    const int lane = threadIdx.x & 63;
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t b = blockIdx.y;
    if (v >= nvp) return;  // uniform across the wave
    int8_t* pv = planes + ((int64_t)b * nvp + v) * lenp;
    if (v >= nv) {
        for (int64_t t0 = (int64_t)lane * 16; t0 < lenp; t0 += 1024)
            for (int l = 0; l < ms.n; ++l) *(uint4*)(pv + l * plane + t0) = uint4{0, 0, 0, 0};
        return;
    }
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const double* xv = x + (which ? ob : oa) + v * sv;
    NormAcc a{0.0, 0, 0.0};
    bool bad = false;
    // the elements of k_norm_rows's lane walk in the same order, four loads in flight
    int64_t t = lane;
    for (; t + 192 < len; t += 256) {
        const double x0 = xv[t], x1 = xv[t + 64], x2 = xv[t + 128], x3 = xv[t + 192];
        norm_add(a, x0, bad);
        norm_add(a, x1, bad);
        norm_add(a, x2, bad);
        norm_add(a, x3, bad);
    }
    for (; t < len; t += 64) norm_add(a, xv[t], bad);
    a = wave_norm(a);
    if (bad) atomicOr(flag, 1);
    const int e = scale_exponent(a, g);
    if (lane == 0) ev[(int64_t)b * nvp + v] = e;
    for (int64_t t0 = (int64_t)lane * 16; t0 < lenp; t0 += 1024) {
        Digits dg[16];
        if (t0 + 16 <= len) {
#pragma unroll
            for (int i = 0; i < 16; ++i) dg[i] = to_digits(rint(ldexp(xv[t0 + i], e)));
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) dg[i] = to_digits(t0 + i < len ? rint(ldexp(xv[t0 + i], e)) : 0.0);
        }
        for_moduli<K14>(ms, [&](int l) {
            const ResConst c = res_const<K14>(ms, l);
            const uint4 w{pack_residues(dg, c), pack_residues(dg + 4, c), pack_residues(dg + 8, c),
                          pack_residues(dg + 12, c)};
            *(uint4*)(pv + l * plane + t0) = w;
        });
    }
}

// The same with the vector held in registers, read once: RPB vectors per block
// (v = RPB blockIdx.x + t / TPR, TPR = 256 / RPB threads each) of batch b = blockIdx.y,
// lenp <= 16 TPR CH. Thread t holds the 16-element chunks t % TPR + TPR i (i < CH) of its
// vector; the norm combines over the vector's waves (through LDS in a fixed order when a
// block holds one vector), then each chunk is split from registers. The wave-per-vector
// kernel above reads each vector twice, and rows of 16 KiB and more miss the cache the
// second time.
template <bool K14, int RPB, int CH>
__global__ void __launch_bounds__(256) k_normsplit_rows_reg(const double* __restrict__ x, int64_t nv, int64_t len,
                                                            int64_t sv, BatchDesc bd, int which, int64_t nvp,
                                                            int64_t lenp, int g, int* __restrict__ ev,
                                                            int8_t* __restrict__ planes, int64_t plane, ModSet ms,
                                                            int* __restrict__ flag) {
    static_assert(RPB == 1 || RPB == 4, "a vector per block or per wave");
    constexpr int TPR = 256 / RPB;
    __shared__ NormAcc red[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, tr = tid % TPR;
    const int64_t v = (int64_t)blockIdx.x * RPB + tid / TPR;  // uniform across each wave
    const uint32_t b = blockIdx.y;
    if (v >= nvp) return;
    int8_t* pv = planes + ((int64_t)b * nvp + v) * lenp;
    if (v >= nv) {  // padding vector: zero planes (uniform across the vector's threads)
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int64_t t0 = ((int64_t)tr + TPR * i) * 16;
            if (t0 < lenp)
                for (int l = 0; l < ms.n; ++l) *(uint4*)(pv + l * plane + t0) = uint4{0, 0, 0, 0};
        }
        return;
    }
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const double* xv = x + (which ? ob : oa) + v * sv;
    const bool pairs = ((uintptr_t)xv & 15) == 0;  // 16-byte loads when aligned
    double xs[CH][16];
    NormAcc a{0.0, 0, 0.0};
    bool bad = false;
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        const int64_t t0 = ((int64_t)tr + TPR * i) * 16;
        if (t0 + 16 <= len) {
            if (pairs) {
#pragma unroll
                for (int e = 0; e < 16; e += 2) {
                    const double2 p2 = *(const double2*)(xv + t0 + e);
                    xs[i][e] = p2.x;
                    xs[i][e + 1] = p2.y;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) xs[i][e] = xv[t0 + e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) xs[i][e] = t0 + e < len ? xv[t0 + e] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) norm_add(a, xs[i][e], bad);
    }
    a = wave_norm(a);
    if (bad) atomicOr(flag, 1);
    if constexpr (RPB == 1) {
        if (lane == 0) red[w] = a;
        __syncthreads();
        a = red[0];
#pragma unroll
        for (int q = 1; q < 4; ++q) norm_merge(a, red[q]);
    }
    const int e = scale_exponent(a, g);
    if (tr == 0) ev[(int64_t)b * nvp + v] = e;
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        const int64_t t0 = ((int64_t)tr + TPR * i) * 16;
        if (t0 >= lenp) continue;
        Digits dg[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) dg[k] = to_digits(rint(ldexp(xs[i][k], e)));
        for_moduli<K14>(ms, [&](int l) {
            const ResConst c = res_const<K14>(ms, l);
            const uint4 wd{pack_residues(dg, c), pack_residues(dg + 4, c), pack_residues(dg + 8, c),
                           pack_residues(dg + 12, c)};
            *(uint4*)(pv + l * plane + t0) = wd;
        });
    }
}

/* ---------------- INT8 MFMA GEMM per modulus ---------------- */

// INT8 tile configurations: a BM x BN block tile from WVM x WVN waves, K slices
// of BK bytes double-buffered in LDS (rows padded by PAD bytes against bank
// conflicts of the 16-byte fragment reads).
// FR: the MFMA's output block, 16 (v_mfma_i32_16x16x64_i8) or 32 (v_mfma_i32_32x32x32_i8).
// D: K slices whose global loads are in flight at once (register sets).
// ST: 0 for loads staged through registers; otherwise the slices go straight from global memory
// into a ring of ST LDS buffers (global_load_lds_dwordx4, unpadded swizzled rows), ST - 1 ahead.
// SG: the K steps of a slice scheduled as fragment reads interleaved with the MFMAs before them.
template <int BM_, int BN_, int BK_, int WVM_, int WVN_, int OCC_, int PAD_, int FR_ = 16, int D_ = 1, int ST_ = 0,
          int SG_ = 0>
struct I8Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WVM = WVM_, WVN = WVN_, OCC = OCC_, PAD = PAD_, FR = FR_,
                         D = D_, ST = ST_, SG = SG_;
    static constexpr int NT = 64 * WVM * WVN;
    static constexpr int NBUF = ST > 2 ? ST : 2;
};
// products more than 128 wide: eight waves of 64 x 64, two blocks per CU
using I8Wide8 = I8Cfg<128, 256, 64, 2, 4, 2, 32>;
// small and grouped products: 64 x 64 per wave, two blocks per CU
using I8Sq = I8Cfg<128, 128, 128, 2, 2, 2, 16>;
// large products: 256 x 256 block tiles from eight waves of 64 x 128 (a third fewer LDS
// fragment bytes per MFMA than 64 x 64 waves), 128-byte K slices in XOR-swizzled rows loaded
// straight into two LDS buffers, fragment reads interleaved with the MFMAs, one block per CU
using I8Big = I8Cfg<256, 256, 128, 4, 2, 1, 0, 16, 1, 2, 1>;
// planes: rows padded to a multiple of the tile's BM (I8_MPAD for all but the 256-row tiles),
// K to I8_KPAD (every BK divides it)
constexpr int I8_MPAD = 128, I8_KPAD = 128;
// one launch sums at most I8_KCHUNK products per output: |sum| <= 2^16 * 2^14 < 2^31
constexpr int64_t I8_KCHUNK = 65536;
static_assert(I8Sq::BM == I8_MPAD && I8Sq::BN == 128, "grouped problems pad to the square tile");

// x mod m in [0, m) for |x| <= 2^30 and 199 <= m <= 256 in full-rate operations (no 32-bit
// multiplies, which issue at a quarter of the rate): x = xh 2^16 + xl (xh = x >> 16, xl the low
// 16 bits); s = xh c16 + xl + m 2^14 with c16 = 2^16 mod m is x mod m and lies in [2^14, 2^24)
// (|xh| <= 2^14, c16 < m), one 24-bit multiply-add; q = trunc(f32(s) inv) with
// inv = f32(1/m) (1 - 2^-22) is floor(s / m) or one less (checked for every s < 2^24 and every
// modulus), so t = s - q m (24-bit multiply-add) lies in [0, 2m) and min(t, t - m) as unsigned
// words is the residue.
__device__ __forceinline__ uint32_t mod_res(int x, uint32_t m, uint32_t c16, float inv) {
    const int s = __mul24(x >> 16, (int)c16) + (x & 0xFFFF) + (int)(m << 14);
    const int q = (int)((float)s * inv);
    const int t = __mul24(q, -(int)m) + s;
    return min((uint32_t)t, (uint32_t)(t - (int)m));
}

// x mod m in [0, m) for |x| < 2^31, m <= 256, inv = 1/m: x * inv is within
// 2^-21 of x / m, so its floor is the quotient or (x a multiple of m rounding
// below) one less; the remainder is then in [0, 2m) and corrected.
__device__ __forceinline__ int mod_small(int x, int m, double inv) {
    const int q = (int)floor((double)x * inv);
    int r = x - q * m;
    if (r < 0) r += m;
    if (r >= m) r -= m;
    return r;
}

// this wave's vector memory operations down to at most N outstanding and its LDS operations
// done, then the workgroup barrier; a compiler fence for memory on both sides
template <int N>
__device__ __forceinline__ void vm_wait_barrier() {
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

// Moduli of one launch
struct ModList {
    int m[OZ_MAXMOD];
    uint32_t c16[OZ_MAXMOD];  // 2^16 mod m (mod_res)
    float inv[OZ_MAXMOD];     // f32(1/m) (1 - 2^-22) (mod_res)
};

template <class CF>
struct I8Smem {
    int8_t As[CF::NBUF][CF::BM][CF::BK + CF::PAD];
    int8_t Bs[CF::NBUF][CF::BN][CF::BK + CF::PAD];
};

// Output tile (tm, tn) of one residue product over a K range of Kc <= I8_KCHUNK
// bytes: A8 [>= (tm+1) BM][ldk] and B8 [>= (tn+1) BN][ldk] (K contiguous, zero
// padded); R [M][ldr] gets the uint8 residues of (A8 B8^T) mod `mod` for rows
// < M and columns < N, added to the residues already there when `accum`. Each
// wave owns (BM / WVM) x (BN / WVN) outputs from FR x FR MFMA blocks
// (v_mfma_i32_16x16x64_i8 or v_mfma_i32_32x32x32_i8). K runs in slices of BK
// bytes through LDS buffers: with CF::ST == 0 global loads run CF::D slices ahead
// in registers (D register sets) and two LDS buffers take them in turn, so a load
// has D slices of MFMA work to arrive; with CF::ST > 0 the slices go straight
// into a ring of ST buffers, ST - 1 slices ahead.
// TOUT: each product block is formed transposed (B fragment as the MFMA's row
// operand), so a lane ends with runs of four consecutive columns of one row of
// R: one 4-byte store per run (N % 4 == 0), the residue by mod_res. The 32 x 32
// blocks always use it.
// MM (direct loads, TOUT, no accum; R holds every row and column of the tile grid, ldr a
// multiple of 4): the tile's products for nmod moduli (planes pA, pB, pR apart; moduli
// ml->m[l0 ..]) in one call, the residues of padding rows and columns written too.
template <class CF, bool ROW4G, bool TOUT = false, bool MM = false>
__device__ __forceinline__ void i8_tile(const int8_t* __restrict__ A, const int8_t* __restrict__ B,
                                        uint8_t* __restrict__ R, int64_t M, int64_t N, int64_t ldk, int64_t Kc,
                                        int64_t ldr, int mod, bool accum, uint32_t tm, uint32_t tn, I8Smem<CF>& sm,
                                        uint32_t c16 = 0, float inv = 0.f, int nmod = 1, int64_t pA = 0, int64_t pB = 0,
                                        int64_t pR = 0, const ModList* ml = nullptr, int l0 = 0) {
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT, FR = CF::FR, D = CF::D;
    constexpr bool T = TOUT || FR == 32;
    constexpr int FK = FR == 16 ? 64 : 32;  // bytes of K per MFMA
    constexpr int NREG = FR == 16 ? 4 : 16;  // output registers per lane
    using Acc = std::conditional_t<FR == 16, v4i, v16i>;
    constexpr int WTM = BM / CF::WVM, WTN = BN / CF::WVN;
    constexpr int FM = WTM / FR, FN = WTN / FR;
    constexpr int CPR = BK / 16, RPP = NT / CPR, LA = BM / RPP, LB = BN / RPP;
    static_assert(BM % RPP == 0 && BN % RPP == 0 && BK % FK == 0 && I8_KPAD % BK == 0 && FM > 0 && FN > 0 &&
                      (FR == 16 || FR == 32) && D >= 1,
                  "tile shape");
    auto& As = sm.As;
    auto& Bs = sm.Bs;
    const int8_t* Ab = A + (size_t)tm * BM * ldk;
    const int8_t* Bb = B + (size_t)tn * BN * ldk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / CF::WVN, wn = wave % CF::WVN;
    const int lr = tid / CPR, lc = (tid % CPR) * 16;  // loader: CPR threads x 16 B per row, RPP rows per pass

    v4i ra[D][LA], rb[D][LB];
    auto load = [&](int set, int64_t k0) {
#pragma unroll
        for (int i = 0; i < LA; ++i) ra[set][i] = *(const v4i*)(Ab + (size_t)(lr + RPP * i) * ldk + k0 + lc);
#pragma unroll
        for (int i = 0; i < LB; ++i) rb[set][i] = *(const v4i*)(Bb + (size_t)(lr + RPP * i) * ldk + k0 + lc);
    };
    // PAD == 0: rows unpadded, 16-byte chunk c of row r stored at chunk c ^ swz(r), which puts
    // any 8 consecutive rows' chunk c (a fragment read by 8 lanes) and any 2 rows' 4 chunks (a
    // loader store by 8 lanes) in 8 distinct 16-byte bank groups of 128 bytes
    static_assert(CF::PAD != 0 || BK == 64 || BK == 128, "swizzle for 64 or 128-byte rows");
    auto swz = [](int r) { return CF::PAD != 0 ? 0 : (BK == 64 ? (r >> 1) & 3 : r & 7); };
    const int lcs = ((lc >> 4) ^ swz(lr)) << 4;  // RPP is a multiple of 8: one swizzle for all passes
    auto store = [&](int set, int buf) {
#pragma unroll
        for (int i = 0; i < LA; ++i) *(v4i*)&As[buf][lr + RPP * i][lcs] = ra[set][i];
#pragma unroll
        for (int i = 0; i < LB; ++i) *(v4i*)&Bs[buf][lr + RPP * i][lcs] = rb[set][i];
    };
    Acc acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < NREG; ++r) acc[i][j][r] = 0;

    const int fr = lane % FR, fk = (lane / FR) * 16;
    static_assert(RPP % 8 == 0 && FR % 8 == 0 && WTM % 8 == 0 && WTN % 8 == 0, "swizzle period");
    const int sr = swz(fr);  // fragment rows start at multiples of 8
    auto compute = [&](int buf) {
        if constexpr (CF::SG) __builtin_amdgcn_sched_barrier(0);  // the slice's reads and MFMAs alone
#pragma unroll
        for (int kk = 0; kk < BK; kk += FK) {
            v4i af[FM], bf[FN];
            const int kc = (((kk + fk) >> 4) ^ sr) << 4;
            // program order a0, b0 .. b(FN-1), a1 .. a(FM-1): the first row of MFMAs needs the first reads
            af[0] = *(const v4i*)&As[buf][wm * WTM + fr][kc];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = *(const v4i*)&Bs[buf][wn * WTN + j * FR + fr][kc];
#pragma unroll
            for (int i = 1; i < FM; ++i) af[i] = *(const v4i*)&As[buf][wm * WTM + i * FR + fr][kc];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    if constexpr (FR == 32)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bf[j], af[i], acc[i][j], 0, 0, 0);
                    else if constexpr (T)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[j], af[i], acc[i][j], 0, 0, 0);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
                }
        }
        if constexpr (CF::SG) {
            // instruction order for the scheduler: the first K step's fragment reads, then each
            // later step's reads spread over the MFMAs of the step before it (Q MFMAs per read),
            // so fragments arrive while the matrix cores run
            constexpr int NR = FM + FN, NMF = FM * FN, NS = BK / FK, Q = NMF / NR >= 2 ? 2 : 1;
            static_assert(Q * NR <= NMF && NS >= 2, "reads fit between the MFMAs");
            __builtin_amdgcn_sched_group_barrier(0x100, NR, 0);
#pragma unroll
            for (int s = 0; s + 1 < NS; ++s) {
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    __builtin_amdgcn_sched_group_barrier(0x008, Q, 0);
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                }
                __builtin_amdgcn_sched_group_barrier(0x008, NMF - Q * NR, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, NMF, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    // residues of the accumulators mod modm into Rm (added to those there when accum)
    auto epilogue = [&](uint8_t* __restrict__ Rm, int modm, uint32_t c16m, float invm) {
        if constexpr (T) {
            static_assert(ROW4G, "transposed output assumes rows 4 (lane / 16) + r of the 16 x 16 MFMA result");
            // block (i, j), register quad q: lane holds Rm[row lane % FR][4 consecutive columns from cq]
            const bool vec = !accum && (ldr % 4) == 0;
    #pragma unroll
            for (int i = 0; i < FM; ++i)
    #pragma unroll
                for (int j = 0; j < FN; ++j)
    #pragma unroll
                    for (int q = 0; q < NREG / 4; ++q) {
                        const int cq = FR == 16 ? 4 * (lane >> 4) : 8 * q + 4 * (lane >> 5);
                        const int64_t gm = (int64_t)tm * BM + wm * WTM + i * FR + (lane % FR);
                        const int64_t gn = (int64_t)tn * BN + wn * WTN + j * FR + cq;
                        if (gm >= M) continue;
                        uint32_t w = 0;
    #pragma unroll
                        for (int r = 0; r < 4; ++r) w |= mod_res(acc[i][j][4 * q + r], (uint32_t)modm, c16m, invm) << (8 * r);
                        uint8_t* rp = Rm + gm * ldr + gn;
                        if (vec && gn + 4 <= N) {
                            *(uint32_t*)rp = w;
                        } else {
                            for (int r = 0; r < 4 && gn + r < N; ++r) {
                                uint32_t v = (w >> (8 * r)) & 255u;
                                if (accum) {
                                    v += rp[r];
                                    if (v >= (uint32_t)modm) v -= modm;
                                }
                                rp[r] = (uint8_t)v;
                            }
                        }
                    }
        } else {
            const double inv = 1.0 / (double)modm;
    #pragma unroll
            for (int i = 0; i < FM; ++i)
    #pragma unroll
                for (int j = 0; j < FN; ++j)
    #pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int lrow = ROW4G ? (4 * (lane >> 4) + r) : (4 * r + (lane >> 4));
                        int64_t gm = (int64_t)tm * BM + wm * WTM + i * 16 + lrow;
                        int64_t gn = (int64_t)tn * BN + wn * WTN + j * 16 + (lane & 15);
                        if (gm < M && gn < N) {
                            int v = mod_small(acc[i][j][r], modm, inv);
                            if (accum) {
                                v += Rm[gm * ldr + gn];
                                if (v >= modm) v -= modm;
                            }
                            Rm[gm * ldr + gn] = (uint8_t)v;
                        }
                    }
        }
    };
    auto zero_acc = [&]() {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < NREG; ++r) acc[i][j][r] = 0;
    };

    if constexpr (CF::ST > 0) {
        // direct loads: one global_load_lds_dwordx4 fills 1 KiB of LDS (lane i at byte 16 i), RW
        // whole rows of one wave; lane i takes chunk (i % CPR) ^ swz(row) of its row from global
        // memory, so the LDS rows come out swizzled as the fragment reads expect (RW is a multiple
        // of the swizzle period: swz(row) = swz(i / CPR)). Slice t goes to buffer t % ST, issued
        // ST - 1 slices ahead into the buffer every wave finished reading at the last barrier.
        constexpr int S = CF::ST, RW = 1024 / BK, RP = (NT / 64) * RW, NI = BM / RP + BN / RP;
        static_assert(CF::PAD == 0 && RW % 8 == 0 && BM % RP == 0 && BN % RP == 0 && (S - 2) * NI < 64 && S >= 2,
                      "direct-load tile shape");
        const int wv = __builtin_amdgcn_readfirstlane(wave);
        const int dr = lane / CPR, dch = ((lane % CPR) ^ swz(lane / CPR)) * 16;
        // the buffers are written only by the direct loads, which the compiler does not count as
        // stores: their address escapes here, so no read of them can be folded away as undefined
        asm volatile("" ::"v"(&sm) : "memory");
        // the lane's part of each load address (row dr of the wave's rows, its chunk): one 32-bit
        // offset (dr < 8, ldk < 2^28: the host keeps longer K off this tile) on top of a
        // wave-uniform base
        const uint32_t loff = (uint32_t)(dr * ldk + dch);
        // The block's nmod moduli (planes pA, pB, pR apart) run as one sequence of nmod * ns
        // slices: the next modulus's first slices load during the last products and the
        // residue epilogue of the one before.
        int64_t da = 0, db = 0, dk = 0;  // the next slice to load: plane offsets, K offset
        auto dma_next = [&](int st) {
#pragma unroll
            for (int i = 0; i < BM / RP; ++i) {
                const int r0 = i * RP + wv * RW;
                const int8_t* base = Ab + da + (size_t)r0 * ldk + dk;
                __builtin_amdgcn_global_load_lds(base + loff, &As[st][r0][0], 16, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < BN / RP; ++i) {
                const int r0 = i * RP + wv * RW;
                const int8_t* base = Bb + db + (size_t)r0 * ldk + dk;
                __builtin_amdgcn_global_load_lds(base + loff, &Bs[st][r0][0], 16, 0, 0);
            }
            dk += BK;
            if (dk == Kc) {
                dk = 0;
                da += pA;
                db += pB;
            }
        };
        const int64_t nv = Kc / BK * nmod;
#pragma unroll
        for (int s = 0; s < S - 1; ++s)
            if (s < nv) dma_next(s);
        // slice 0 landed (the slices after it may still be in flight)
        if (nv >= S - 1) vm_wait_barrier<(S - 2) * NI>();
        else vm_wait_barrier<0>();
        int st = 0, sn = S - 1;
        int64_t v = 0;
        auto slices = [&](int64_t cnt) {
            for (int64_t t = 0; t < cnt; ++t, ++v) {
                const bool more = v + S - 1 < nv;
                if (more) dma_next(sn);
                compute(st);
                // slice v + 1 landed, and no wave reads buffer st any more
                if (more) vm_wait_barrier<(S - 2) * NI>();
                else vm_wait_barrier<0>();
                st = st + 1 == S ? 0 : st + 1;
                sn = sn + 1 == S ? 0 : sn + 1;
            }
        };
        if constexpr (MM) {
            // rows of R at least tiles x BM, ldr at least tiles x BN (a multiple of 4), nothing to
            // add to: unguarded word stores (guards or a byte-wise path inside this loop would
            // cost the slices' schedule registers)
            static_assert(T && CF::ST > 0, "word stores from the transposed blocks");
            for (int g = 0; g < nmod; ++g) {
                slices(Kc / BK);
                const int l = l0 + g;
                const uint32_t modm = (uint32_t)ml->m[l], c16m = ml->c16[l];
                const float invm = ml->inv[l];
                uint8_t* Rm = R + (int64_t)g * pR;
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
#pragma unroll
                        for (int q = 0; q < NREG / 4; ++q) {
                            const int cq = FR == 16 ? 4 * (lane >> 4) : 8 * q + 4 * (lane >> 5);
                            const int64_t gm = (int64_t)tm * BM + wm * WTM + i * FR + (lane % FR);
                            const int64_t gn = (int64_t)tn * BN + wn * WTN + j * FR + cq;
                            uint32_t w = 0;
#pragma unroll
                            for (int r = 0; r < 4; ++r) w |= mod_res(acc[i][j][4 * q + r], modm, c16m, invm) << (8 * r);
                            *(uint32_t*)(Rm + gm * ldr + gn) = w;
                        }
                zero_acc();
            }
        } else {
            slices(nv);
        }
    } else {
        // slices 0 .. D-1 in flight, slice 0 in LDS; slice t uses register set t % D and LDS buffer t % 2
#pragma unroll
        for (int s = 0; s < D; ++s)
            if ((int64_t)s * BK < Kc) load(s, (int64_t)s * BK);
        store(0, 0);
        __syncthreads();
        int buf = 0;
        for (int64_t kb = 0; kb < Kc; kb += (int64_t)D * BK) {
#pragma unroll
            for (int s = 0; s < D; ++s) {
                const int64_t k0 = kb + (int64_t)s * BK;
                if (k0 < Kc) {
                    if (k0 + (int64_t)D * BK < Kc) load(s, k0 + (int64_t)D * BK);
                    compute(buf);
                    if (k0 + BK < Kc) store((s + 1) % D, buf ^ 1);
                    __syncthreads();
                    buf ^= 1;
                }
            }
        }
    }
    if constexpr (!MM) epilogue(R, mod, c16, inv);
}

// Batches on blockIdx.y: plane offsets sA / sB (0 for a broadcast operand) and sR.

// Tiles touching the lower triangle of a square product with BN = R BM: tile row tm
// holds floor(tm / R) + 1 of them. Index t -> (tm, tn): for R = 1 and 2 in groups of
// tile rows walked down the columns, otherwise row-major; the floating root estimates
// are corrected exactly.
__host__ __device__ inline uint64_t lower_tile_prefix(uint64_t r, uint64_t R) {
    const uint64_t q = r / R, m = r - q * R;
    return R * q * (q + 1) / 2 + m * (q + 1);
}
__device__ __forceinline__ void lower_tile(uint32_t t, uint32_t R, uint32_t tiles_m, uint32_t& tm, uint32_t& tn) {
    if (R == 1) {
        // groups of four tile rows 4g .. 4g+3 (the last group k <= 4 rows), walked down the
        // columns so that blocks resident together share A and B panels in L2: columns
        // 0 .. 4g-1 with all k rows, then the triangle (column 4g + c, rows 4g + c .. 4g + k-1).
        // Group g starts at 8g^2 + 2g (full groups of 16h + 10 tiles before it).
        uint64_t g = (uint64_t)((sqrt(4.0 + 32.0 * (double)t) - 2.0) * 0.0625);
        while (g > 0 && 8 * g * g + 2 * g > t) --g;
        while (8 * (g + 1) * (g + 1) + 2 * (g + 1) <= t) ++g;
        uint32_t u = (uint32_t)(t - (8 * g * g + 2 * g));
        const uint32_t r0 = (uint32_t)(4 * g), k = min(4u, tiles_m - r0);
        if (u < k * r0) {
            tn = u / k;
            tm = r0 + u % k;
        } else {
            u -= k * r0;
            uint32_t c = 0;
            while (u >= k - c) {
                u -= k - c;
                ++c;
            }
            tn = r0 + c;
            tm = r0 + c + u;
        }
        return;
    }
    if (R == 2) {
        // groups of four tile rows 4g .. 4g+3 (counts 2g+1, 2g+1, 2g+2, 2g+2), walked down
        // the columns so that blocks resident together share A and B panels in L2:
        // columns 0 .. 2g for all four rows, then column 2g+1 for rows 4g+2 and 4g+3.
        // Group g starts at 4g^2 + 2g and holds 8g + 6 tiles; a last group of two rows
        // (tiles_m = 2 mod 4) holds 4g + 2 tiles, column by column.
        uint64_t g = (uint64_t)((sqrt(1.0 + 16.0 * (double)t) - 1.0) * 0.125);
        while (g > 0 && 4 * g * g + 2 * g > t) --g;
        while (4 * (g + 1) * (g + 1) + 2 * (g + 1) <= t) ++g;
        const uint32_t u = (uint32_t)(t - (4 * g * g + 2 * g));
        const uint32_t r0 = (uint32_t)(4 * g);
        if (r0 + 2 >= tiles_m) {  // the two-row last group
            tn = u / 2;
            tm = r0 + u % 2;
        } else if (u < 4 * (2 * g + 1)) {
            tn = u / 4;
            tm = r0 + u % 4;
        } else {
            tn = (uint32_t)(2 * g + 1);
            tm = r0 + 2 + (u - 4 * (uint32_t)(2 * g + 1));
        }
        return;
    }
    uint64_t r = (uint64_t)sqrt(2.0 * R * (double)t);
    while (r > 0 && lower_tile_prefix(r, R) > t) --r;
    while (lower_tile_prefix(r + 1, R) <= t) ++r;
    tm = (uint32_t)r;
    tn = (uint32_t)(t - lower_tile_prefix(r, R));
}

template <class CF, bool ROW4G, bool TOUT, bool MM>
__global__ void __launch_bounds__(CF::NT, CF::OCC) k_i8gemm_mod(const int8_t* __restrict__ A,
                                                                const int8_t* __restrict__ B, uint8_t* __restrict__ R,
                                                                int64_t M, int64_t N, int64_t ldk, int64_t Kc,
                                                                int64_t ldr, ModList mods, int accum, uint32_t tiles_m,
                                                                uint32_t tiles_n, int64_t sA, int64_t sB, int64_t sR,
                                                                int64_t pA, int64_t pB, int64_t pR,
                                                                const int* __restrict__ skip, int lower, int gmod,
                                                                int nmods) {
    if (*skip) return;
    // moduli l0 .. l0 + gmod - 1 (l0 = blockIdx.z gmod, at most nmods in all): every modulus's
    // product in one launch (no tail between them); gmod > 1 only with MM
    const int l0 = (int)blockIdx.z * gmod, g = min(gmod, nmods - l0);
    A += l0 * pA;
    B += l0 * pB;
    R += l0 * pR;
    const int mod = mods.m[l0];
    __shared__ __attribute__((aligned(16))) I8Smem<CF> sm;
    uint32_t tm, tn;
    // symmetric product: the grid enumerates only the tiles touching the lower triangle,
    // each XCD one contiguous range of them (an equal share of the work)
    if (lower) lower_tile(xcd_chunk(blockIdx.x, gridDim.x), CF::BN / CF::BM, tiles_m, tm, tn);
    else tile_coords<4>(blockIdx.x, tiles_m, tiles_n, tm, tn);
    i8_tile<CF, ROW4G, TOUT, MM>(A + (int64_t)blockIdx.y * sA, B + (int64_t)blockIdx.y * sB,
                             R + (int64_t)blockIdx.y * sR, M, N, ldk, Kc, ldr, mod, accum != 0, tm, tn, sm,
                             mods.c16[l0], mods.inv[l0], g, pA, pB, pR, &mods, l0);
}

// zero n ints (a kernel rather than a memset node, so the sequence can be graph-captured)
__global__ void k_zero_i32(int* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

/* ---------------- CRT reconstruction ---------------- */

// The integer Z in (-M/2, M/2) with residues r[l] mod m_l, as a double (the
// product's integer lies well inside: |Z| <= 2^(2g) (1+d)^2 < 0.23 M for g = 54).
// X = sum_l r_l w_l is summed per 38-bit limb of w in f64 (each limb sum
// < 14 * 2^8 * 2^38 < 2^50: exact). X < 14 * 2^8 * M, so X / M is off an integer
// Q by |Z| / M < 0.23, and the f64 estimate of X / M is within 2^-38 of it:
// rint gives Q exactly, and Z = X - Q M limb by limb (each exact, |.| < 2^51).
// Carries then bring the two low limbs into [0, 2^38), and the value is
// assembled with two roundings (a result below 2^53 is exact). crt_finish takes the
// limb sums; crt_from forms them (K14: the default 14 moduli's constants from CRT14,
// otherwise those of cc).
__device__ __forceinline__ double crt_finish(double Sa, double Sb, double Sc, double invM, const double (&Ml)[3]) {
    constexpr double T38 = 274877906944.0, I38 = 1.0 / 274877906944.0, T76 = T38 * T38;
    const double Xd = fma(Sa, T76, fma(Sb, T38, Sc));
    const double Q = rint(Xd * invM);
    double Za = fma(-Q, Ml[0], Sa), Zb = fma(-Q, Ml[1], Sb), Zc = fma(-Q, Ml[2], Sc);
    const double c1 = floor(Zc * I38);
    Zc = fma(-c1, T38, Zc);
    Zb += c1;
    const double c2 = floor(Zb * I38);
    Zb = fma(-c2, T38, Zb);
    Za += c2;
    return fma(fma(Za, T38, Zb), T38, Zc);
}

template <bool K14>
__device__ __forceinline__ double crt_from(const uint32_t (&r)[OZ_MAXMOD], const CrtConst& cc) {
    // a full unrolled sum (constants past cc.n are zero): no runtime register indexing
    double Sa = 0.0, Sb = 0.0, Sc = 0.0;
#pragma unroll
    for (int l = 0; l < OZ_MAXMOD; ++l) {
        const double rl = (double)r[l];
        Sa = fma(rl, K14 ? CRT14.wl[l][0] : cc.wl[l][0], Sa);
        Sb = fma(rl, K14 ? CRT14.wl[l][1] : cc.wl[l][1], Sb);
        Sc = fma(rl, K14 ? CRT14.wl[l][2] : cc.wl[l][2], Sc);
    }
    return K14 ? crt_finish(Sa, Sb, Sc, CRT14.invM, CRT14.Ml) : crt_finish(Sa, Sb, Sc, cc.invM, cc.Ml);
}

// The same from residue bytes p[l * pl]: with the runtime constants a rolled loop that
// accumulates as it loads (each limb triple read by scalar loads in its iteration; the
// unrolled form would hold all 42 limbs in scalar registers, which do not fit).
template <bool K14>
__device__ __forceinline__ double crt_value(const uint8_t* __restrict__ p, int64_t pl, const CrtConst& cc) {
    if constexpr (K14) {
        uint32_t r[OZ_MAXMOD];
#pragma unroll
        for (int l = 0; l < OZ_MAXMOD; ++l) r[l] = p[l * pl];
        return crt_from<true>(r, cc);
    } else {
        double Sa = 0.0, Sb = 0.0, Sc = 0.0;
#pragma unroll 2
        for (int l = 0; l < cc.n; ++l) {
            const double rl = (double)p[l * pl];
            Sa = fma(rl, cc.wl[l][0], Sa);
            Sb = fma(rl, cc.wl[l][1], Sb);
            Sc = fma(rl, cc.wl[l][2], Sc);
        }
        return crt_finish(Sa, Sb, Sc, cc.invM, cc.Ml);
    }
}

// R planes: [l][batch][Mp][ldr] (plane stride pl, rows ldr >= Np bytes); eA [nbA][Mp], fB [nbB][Np].
// C = alpha 2^-(eA_i + fB_j) CRT(R) + beta C: columns along x (V consecutive
// per thread, read as one V-byte word per plane), rows along y, batches along z.
template <int V, bool K14>  // columns per thread: 1, 4 (one 4-byte word per plane) or 16 (one 16-byte word)
__global__ void __launch_bounds__(256) k_crt(const uint8_t* __restrict__ R, int64_t M, int64_t N, int64_t Mp,
                                             int64_t Np, int64_t ldr, int64_t pl, int a_batched, int b_batched, BatchDesc bd,
                                             CrtConst cc, const int* __restrict__ eA, const int* __restrict__ fB,
                                             double* C, int64_t rsc, int64_t csc, double alpha, double beta,
                                             int beta_zero, const int* __restrict__ skip, int tpr) {
    static_assert(V == 1 || V == 4 || V == 16, "word size");
    if (*skip) return;
    // tpr threads (a power of two <= 256) per row, 256 / tpr rows of a block at a time
    const int rpb = 256 / tpr;
    const int64_t j0 = ((int64_t)blockIdx.x * tpr + (threadIdx.x & (tpr - 1))) * V;
    if (j0 >= N) return;
    const uint32_t b = blockIdx.z;
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const int* eAb = eA + (a_batched ? (int64_t)b * Mp : 0);
    const int* fBb = fB + (b_batched ? (int64_t)b * Np : 0);
    if constexpr (V == 4) {
        // one 4-byte word per plane and row; the next row's words load while this row's
        // CRT runs, and aligned unit-stride rows store two doubles at a time
        const int nm = K14 ? OZ_MAXMOD : cc.n;  // a constant for the default moduli: no branches
        const int64_t step = (int64_t)gridDim.y * rpb;
        int64_t i = (int64_t)blockIdx.y * rpb + threadIdx.x / tpr;
        auto load = [&](int64_t row, uint32_t (&w)[OZ_MAXMOD]) {
            const uint8_t* rp = R + ((int64_t)b * Mp + row) * ldr + j0;
#pragma unroll
            for (int l = 0; l < OZ_MAXMOD; ++l) w[l] = l < nm ? *(const uint32_t*)(rp + l * pl) : 0u;
        };
        uint32_t w[OZ_MAXMOD], wn[OZ_MAXMOD];
        if (i < M) load(i, w);
        int fj[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) fj[k] = fBb[j0 + k];  // fB holds Np >= j0 + 4 entries
        for (; i < M; i += step) {
            const bool more = i + step < M;
            if (more) load(i + step, wn);
            const int ei = eAb[i];
            double v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t r[OZ_MAXMOD];
#pragma unroll
                for (int l = 0; l < OZ_MAXMOD; ++l) r[l] = (w[l] >> (8 * k)) & 255u;
                v[k] = alpha * ldexp(crt_from<K14>(r, cc), -(ei + fj[k]));
            }
            double* cp = C + oc + i * rsc + j0 * csc;
            if (csc == 1 && beta_zero && j0 + 4 <= N && ((uintptr_t)cp & 15) == 0) {
                *(double2*)cp = double2{v[0], v[1]};
                *(double2*)(cp + 2) = double2{v[2], v[3]};
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (j0 + k < N) {
                        double* ck = cp + k * csc;
                        *ck = beta_zero ? v[k] : v[k] + beta * (*ck);
                    }
            }
            if (more) {
#pragma unroll
                for (int l = 0; l < OZ_MAXMOD; ++l) w[l] = wn[l];
            }
        }
        return;
    }
    for (int64_t i = (int64_t)blockIdx.y * rpb + threadIdx.x / tpr; i < M; i += (int64_t)gridDim.y * rpb) {
        const uint8_t* rp = R + ((int64_t)b * Mp + i) * ldr + j0;
        double* cp = C + oc + i * rsc + j0 * csc;
        const int ei = eAb[i];
        if constexpr (V == 1) {
            const double v = ldexp(crt_value<K14>(rp, pl, cc), -(ei + fBb[j0]));
            *cp = beta_zero ? alpha * v : alpha * v + beta * (*cp);
        } else {
            constexpr int NW = V / 4;  // 4-byte words per plane
            const int nm = K14 ? OZ_MAXMOD : cc.n;  // a constant for the default moduli: no branches
            uint32_t w[OZ_MAXMOD][NW];
#pragma unroll
            for (int l = 0; l < OZ_MAXMOD; ++l) {
                if constexpr (NW == 1) {
                    w[l][0] = l < nm ? *(const uint32_t*)(rp + l * pl) : 0u;
                } else {
                    const uint4 q = l < nm ? *(const uint4*)(rp + l * pl) : uint4{0, 0, 0, 0};
                    w[l][0] = q.x;
                    w[l][1] = q.y;
                    w[l][2] = q.z;
                    w[l][3] = q.w;
                }
            }
            // every column computed (fB holds Np >= j0 + V entries), the stores guarded
#pragma unroll
            for (int k = 0; k < V; ++k) {
                uint32_t r[OZ_MAXMOD];
#pragma unroll
                for (int l = 0; l < OZ_MAXMOD; ++l) r[l] = (w[l][k / 4] >> (8 * (k % 4))) & 255u;
                const double v = ldexp(crt_from<K14>(r, cc), -(ei + fBb[j0 + k]));
                if (j0 + k < N) {
                    double* ck = cp + k * csc;
                    *ck = beta_zero ? alpha * v : alpha * v + beta * (*ck);
                }
            }
        }
    }
}

/* ---------------- grouped (uneven) split GEMM ---------------- */

// One problem of a grouped split GEMM, with its offsets into the group's shared
// arrays: oA / oB / oR elements into one A8 / B8 / R plane, oe / of into eA / fB.
struct OzProb {
    int64_t M, N, K, Mp, Np, Kp;
    const double* a;
    int64_t rsa, csa;
    const double* b;
    int64_t rsb, csb;
    double* c;
    int64_t rsc, csc;
    int64_t oA, oB, oR, oe, of;
    FastDiv fdN;  // division by N (M N < 2^32)
};

// Units of a grouped launch are numbered problem by problem: unit u belongs to
// the last problem p with pre[p] <= u.
__device__ __forceinline__ int find_prob(const uint32_t* __restrict__ pre, int count, uint32_t u) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Shape of the CRT tiles: TC columns (64 or 128), TPR = TC / 4 threads per row (four
// columns each), RPP = 256 / TPR rows per pass, TR = 4 RPP rows (four passes): 64 x 64 or
// 32 x 128. Host and kernel take the shape from here.
template <int TC_>
struct CrtTile {
    static_assert(TC_ == 64 || TC_ == 128, "tile width");
    static constexpr int TC = TC_, TPR = TC / 4, RPP = 256 / TPR, TR = 4 * RPP;
};

// CRT over TR x TC tiles of C (CrtTile): thread t computes rows i0 + t / TPR + RPP k
// (k < 4) at columns j0 + 4 (t % TPR) .. + 3, from one 4-byte word per plane and row
// (planes of Mp x ldr bytes, batch b at row b Mp; the grid's tiles lie within Mp x Np).
// Symmetric product (LOWER): only the tiles touching the lower triangle run (index
// blockIdx.x over them), and each also writes its transpose onto the mirror tile through
// LDS (both stores run along C's rows; only elements above the diagonal), so no separate
// mirror pass reads C back.
template <bool K14, bool LOWER, int TC>
__global__ void __launch_bounds__(256) k_crt_tile(const uint8_t* __restrict__ R, int64_t M, int64_t N, int64_t Mp,
                                                  int64_t Np, int64_t ldr, int64_t pl, int a_batched, int b_batched,
                                                  BatchDesc bd, CrtConst cc, const int* __restrict__ eA,
                                                  const int* __restrict__ fB, double* C, int64_t rsc, int64_t csc,
                                                  double alpha, double beta, int beta_zero, const int* __restrict__ skip,
                                                  uint32_t tiles_m, uint32_t tiles_n) {
    constexpr int TPR = CrtTile<TC>::TPR, RPP = CrtTile<TC>::RPP, TR = CrtTile<TC>::TR;
    if (*skip) return;
    __shared__ double tile[LOWER ? TR : 1][TC + 1];
    uint32_t ti, tj;
    if constexpr (LOWER) lower_tile(blockIdx.x, TC / TR, tiles_m, ti, tj);
    else {
        ti = blockIdx.x / tiles_n;
        tj = blockIdx.x - ti * tiles_n;
    }
    const uint32_t b = blockIdx.y;
    int64_t oa, ob, oc;
    batch_offsets(bd, b, oa, ob, oc);
    const int* eAb = eA + (a_batched ? (int64_t)b * Mp : 0);
    const int* fBb = fB + (b_batched ? (int64_t)b * Np : 0);
    double* Cb = C + oc;
    const int tid = threadIdx.x, cg = tid % TPR, rg = tid / TPR;
    const int64_t i0 = (int64_t)ti * TR, j0 = (int64_t)tj * TC, jc = j0 + 4 * cg;
    const int nm = K14 ? OZ_MAXMOD : cc.n;
    uint32_t w[4][OZ_MAXMOD];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = i0 + rg + RPP * k;  // i < Mp: rows of R exist up to the padded extent
        const uint8_t* rp = R + ((int64_t)b * Mp + i) * ldr + jc;
#pragma unroll
        for (int l = 0; l < OZ_MAXMOD; ++l) w[k][l] = l < nm ? *(const uint32_t*)(rp + l * pl) : 0u;
    }
    int fj[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) fj[c] = fBb[jc + c];  // fB holds Np >= jc + 4 entries
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = i0 + rg + RPP * k;
        const int ei = eAb[i];
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uint32_t r[OZ_MAXMOD];
#pragma unroll
            for (int l = 0; l < OZ_MAXMOD; ++l) r[l] = (w[k][l] >> (8 * c)) & 255u;
            v[c] = alpha * ldexp(crt_from<K14>(r, cc), -(ei + fj[c]));
        }
        if (i < M) {
            double* cp = Cb + i * rsc + jc * csc;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (jc + c < N && !(LOWER && jc + c > i)) {
                    double* ck = cp + c * csc;
                    *ck = beta_zero ? v[c] : v[c] + beta * (*ck);
                }
        }
        if constexpr (LOWER) {
#pragma unroll
            for (int c = 0; c < 4; ++c) tile[rg + RPP * k][4 * cg + c] = v[c];
        }
    }
    if constexpr (LOWER) {
        __syncthreads();
        // mirror: row j0 + r of C (r < TC) gets columns i0 .. i0 + TR - 1 from the tile's column r
        constexpr int MPR = 256 / TR;  // mirror rows per pass
        const int c = tid % TR, r0 = tid / TR;
#pragma unroll
        for (int k = 0; k < TC / MPR; ++k) {
            const int r = r0 + MPR * k;
            if (j0 + r < N && i0 + c < M && i0 + c > j0 + r) Cb[(j0 + r) * rsc + (i0 + c) * csc] = tile[c][r];
        }
    }
}

// Scale exponents for a group: one block per 64 rows of an A or 64 columns of a
// B. A vector whose elements are contiguous is read by one wave along the
// vector; vectors side by side are read a lane per vector, the four waves
// splitting the elements. flag[0] |= 1 on a non-finite element.
__global__ void __launch_bounds__(256) k_gnorm(const OzProb* __restrict__ probs, int count,
                                               const uint32_t* __restrict__ pre, int g, int* __restrict__ eA,
                                               int* __restrict__ fB, int* __restrict__ flag) {
    __shared__ NormAcc red[4][64];
    const uint32_t u = blockIdx.x;
    const int p = find_prob(pre, count, u);
    const OzProb& P = probs[p];
    const uint32_t local = u - pre[p];
    const uint32_t ua = (uint32_t)((P.M + 63) / 64);
    const bool isA = local < ua;
    const double* x = isA ? P.a : P.b;
    const int64_t nvec = isA ? P.M : P.N, sv = isA ? P.rsa : P.csb, st = isA ? P.csa : P.rsb, len = P.K;
    int* ev = isA ? eA + P.oe : fB + P.of;
    const int64_t v0 = (int64_t)(isA ? local : local - ua) * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    bool bad = false;
    if ((st == 1 || st == -1) && !(sv == 1 || sv == -1)) {
        for (int q = w; q < 64; q += 4) {
            const int64_t v = v0 + q;
            if (v >= nvec) break;  // uniform across the wave
            const double* xv = x + v * sv;
            NormAcc a{0.0, 0, 0.0};
            for (int64_t t = lane; t < len; t += 64) norm_add(a, xv[t * st], bad);
            a = wave_norm(a);
            if (lane == 0) ev[v] = scale_exponent(a, g);
        }
    } else {
        const int64_t v = v0 + lane;
        NormAcc a{0.0, 0, 0.0};
        if (v < nvec)
            for (int64_t t = w; t < len; t += 4) norm_add(a, x[v * sv + t * st], bad);
        red[w][lane] = a;
        __syncthreads();
        if (w == 0 && v < nvec) {
#pragma unroll
            for (int q = 1; q < 4; ++q) norm_merge(a, red[q][lane]);
            ev[v] = scale_exponent(a, g);
        }
    }
    if (bad) atomicOr(flag, 1);
}

// Residue planes for a group: one block per 64 x 64 tile of an A8 or B8 plane.
template <bool K14>
__global__ void __launch_bounds__(256) k_gsplit(const OzProb* __restrict__ probs, int count,
                                                const uint32_t* __restrict__ pre, const int* __restrict__ eA,
                                                const int* __restrict__ fB, int8_t* __restrict__ A8,
                                                int8_t* __restrict__ B8, int64_t planeA, int64_t planeB, ModSet ms,
                                                const int* __restrict__ skip) {
    if (*skip) return;
    __shared__ double tile[64][65];
    const uint32_t u = blockIdx.x;
    const int p = find_prob(pre, count, u);
    const OzProb& P = probs[p];
    const uint32_t local = u - pre[p];
    const uint32_t kt = (uint32_t)(P.Kp / 64);
    const uint32_t ta = (uint32_t)(P.Mp / 64) * kt;
    if (local < ta) {
        const uint32_t vt = local / kt, tt = local - vt * kt;
        split_tile<K14>(P.a, P.M, P.K, P.rsa, P.csa, eA + P.oe, P.Mp, P.Kp, A8 + P.oA, ms, P.csa == 1 || P.csa == -1,
                        planeA,
                   (int64_t)vt * 64, (int64_t)tt * 64, tile);
    } else {
        const uint32_t lb = local - ta;
        const uint32_t vt = lb / kt, tt = lb - vt * kt;
        split_tile<K14>(P.b, P.N, P.K, P.csb, P.rsb, fB + P.of, P.Np, P.Kp, B8 + P.oB, ms, P.rsb == 1 || P.rsb == -1,
                        planeB,
                   (int64_t)vt * 64, (int64_t)tt * 64, tile);
    }
}

// Residue products for a group: blockIdx.x runs over every output tile of every
// problem, blockIdx.y over the moduli.
template <class CF, bool ROW4G>
__global__ void __launch_bounds__(CF::NT, CF::OCC) k_gi8gemm(const OzProb* __restrict__ probs, int count,
                                                 const uint32_t* __restrict__ pre, uint32_t total,
                                                 const int8_t* __restrict__ A8, const int8_t* __restrict__ B8,
                                                 uint8_t* __restrict__ R, int64_t planeA, int64_t planeB,
                                                 int64_t planeR, ModSet ms, const int* __restrict__ skip) {
    if (*skip) return;
    __shared__ __attribute__((aligned(16))) I8Smem<CF> sm;
    const uint32_t u = xcd_remap8(blockIdx.x, total);
    const int p = find_prob(pre, count, u);
    const OzProb& P = probs[p];
    const uint32_t local = u - pre[p];
    const uint32_t tiles_n = (uint32_t)(P.Np / CF::BN);
    const uint32_t tm = local / tiles_n, tn = local - tm * tiles_n;
    const int64_t l = blockIdx.y;
    i8_tile<CF, ROW4G>(A8 + l * planeA + P.oA, B8 + l * planeB + P.oB, R + l * planeR + P.oR, P.Mp, P.N, P.Kp, P.Kp,
                       P.N, (int)ms.m[l], false, tm, tn, sm);
}

// CRT for a group: one thread per output element, ceil(M N / 256) blocks per problem.
template <bool K14>
__global__ void __launch_bounds__(256) k_gcrt(const OzProb* __restrict__ probs, int count,
                                              const uint32_t* __restrict__ pre, const uint8_t* __restrict__ R,
                                              int64_t planeR, CrtConst cc, const int* __restrict__ eA,
                                              const int* __restrict__ fB, double alpha, double beta, int beta_zero,
                                              const int* __restrict__ skip) {
    if (*skip) return;
    const uint32_t u = blockIdx.x;
    const int p = find_prob(pre, count, u);
    const OzProb& P = probs[p];
    const uint32_t idx = (u - pre[p]) * 256 + threadIdx.x;
    if ((uint64_t)idx >= (uint64_t)P.M * (uint64_t)P.N) return;
    const uint32_t iq = fdiv(idx, P.fdN);
    const int64_t i = iq, j = (int64_t)(idx - iq * (uint32_t)P.N);
    double v = crt_value<K14>(R + P.oR + i * P.N + j, planeR, cc);
    v = ldexp(v, -(eA[P.oe + i] + fB[P.of + j]));
    double* cp = P.c + i * P.rsc + j * P.csc;
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
    const u128 L38 = ((u128)1 << 38) - 1;
    auto limbs = [&](u128 v, double* out) {
        out[0] = (double)(uint64_t)(v >> 76);
        out[1] = (double)(uint64_t)((v >> 38) & L38);
        out[2] = (double)(uint64_t)(v & L38);
    };
    limbs(M, cc.Ml);
    cc.Md = (double)(uint64_t)(M >> 64) * 18446744073709551616.0 + (double)(uint64_t)M;
    cc.invM = 1.0 / cc.Md;
    for (int l = 0; l < n; ++l) {
        u128 Ml = M / OZ_MODULI[l];
        uint64_t r = (uint64_t)(Ml % OZ_MODULI[l]);
        uint64_t inv = inv_mod(r, OZ_MODULI[l]);
        const u128 w = mul_mod(Ml, inv, M);
        limbs(w, cc.wl[l]);
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

// Moduli, CRT constants and the scale target g for the configured modulus count.
struct OzSetup {
    int n;
    CrtConst cc;
    int g;
    ModSet ms;
};

static OzSetup ozaki_setup() {
    OzSetup su{};
    su.n = ozaki_moduli_count();
    su.cc = make_crt(su.n);
    const double log2M = std::log2(su.cc.Md);
    // Cauchy-Schwarz: (2^g (1 + d))^2 < M/2 with a 1/2-bit margin; the residue digits (x + 2^56 in
    // eight bytes) need |x| <= 2^55
    su.g = (int)std::floor((log2M - 1.5) / 2.0);
    if (su.g > 55) su.g = 55;
    su.ms.n = su.n;
    for (int l = 0; l < su.n; ++l) {
        const uint32_t m = su.cc.mod[l];
        const ResConst c = make_res_const(m);
        su.ms.m[l] = (double)m;
        su.ms.mi[l] = m;
        su.ms.c16[l] = (1u << 16) % m;
        su.ms.inv[l] = (float)(1.0 / (double)m) * (1.0f - 0x1p-22f);
        su.ms.clo[l] = c.clo;
        su.ms.chi[l] = c.chi;
        su.ms.off[l] = c.off;
        su.ms.mg20[l] = c.mg20;
    }
    return su;
}

// Tile shape of the residue products. Measured (14 moduli): at 8192^3 the 256 x 256
// tile takes 6.2-6.4 ms (6.8-7.1 with loads staged through registers and the compiler's
// read order) against 8.0 ms for 128 x 256 from eight 64 x 64 waves (7.8 with direct loads,
// two blocks per CU); 128 x 256 from four 64 x 128 waves, 256 x 256 from four 128 x 128 waves
// (which spill), 32 x 32 x 32 MFMAs (11 % slower on the direct-load tile), global loads two
// slices ahead and 64-byte slices in three or four LDS buffers were no faster. Products at most 256 long on a side
// keep the smaller tiles (less padding); outputs at most 128 wide the 128 x 128 tile.
// RSTSR_HIP_OZAKI_TILE (square, wide8, big) overrides the choice.
enum class I8Tile { Sq, Wide8, Big };

static I8Tile pick_i8_tile(int64_t M, int64_t N) {
    const char* s = std::getenv("RSTSR_HIP_OZAKI_TILE");
    if (!s || !*s) return N <= 128 ? I8Tile::Sq : (M > 256 && N > 256 ? I8Tile::Big : I8Tile::Wide8);
    if (std::strcmp(s, "square") == 0) return I8Tile::Sq;
    if (std::strcmp(s, "wide8") == 0) return I8Tile::Wide8;
    if (std::strcmp(s, "big") == 0) return I8Tile::Big;
    RSTSR_RAISE(InvalidValue, "RSTSR_HIP_OZAKI_TILE must be square, wide8 or big");
}

// f(cfg) with a value of the configuration type of tile t
template <class F>
static void with_i8_tile(I8Tile t, F&& f) {
    switch (t) {
        case I8Tile::Wide8: f(I8Wide8{}); break;
        case I8Tile::Big: f(I8Big{}); break;
        default: f(I8Sq{}); break;
    }
}

void ozaki_gemm_f64(rstsr_hip_ctx* ctx, const GemmProblem& P, double alpha, double beta, bool beta_zero) {
    const OzSetup su = ozaki_setup();
    const int n = su.n;
    const CrtConst& cc = su.cc;
    const int g = su.g;
    const ModSet& ms = su.ms;
    const int64_t M = P.M, N = P.N, K = P.K;
    if (K > (int64_t)1 << 40) RSTSR_RAISE(UnImplemented, "K too large for the split GEMM");
    I8Tile tile = pick_i8_tile(M, N);
    // the direct-load tile forms its lanes' load offsets in 32 bits (row < 8, ldk < 2^28)
    if (tile == I8Tile::Big && K >= ((int64_t)1 << 28) - I8_KPAD) tile = I8Tile::Wide8;
    int64_t bm = 0, bn = 0;
    with_i8_tile(tile, [&](auto c) {
        bm = decltype(c)::BM;
        bn = decltype(c)::BN;
    });
    // Symmetric product (the SYRK condition of device_faer/matmul.rs:38-48: beta = 0, A is B's
    // transpose in the same buffer): B's columns are A's rows, so B's scales and residue planes are
    // A's; only the INT8 tiles touching the lower triangle run, and the CRT (k_crt_tile) fills i >= j
    // and mirrors each tile onto j > i. Residue sums are exact integers, so the result is the full
    // product bit for bit.
    bool sym = beta_zero && P.a == P.b && M == N && P.rsa == P.csb && P.csa == P.rsb;
    for (int d = 0; d < P.bd.nbd && sym; ++d) sym = P.bd.sa[d] == P.bd.sb[d];
    const int64_t Np = (N + bn - 1) / bn * bn;
    const int64_t Mp = sym ? Np : (M + bm - 1) / bm * bm;  // one padded extent when shared
    const int64_t Kp = (K + I8_KPAD - 1) / I8_KPAD * I8_KPAD;
    const uint64_t nb = P.nbatch;
    if (nb >= 65536) RSTSR_RAISE(UnImplemented, "too many batches for the split GEMM");
    if (Mp / 64 > 65535 || Np / 64 > 65535) RSTSR_RAISE(UnImplemented, "matrix too tall for the split GEMM");
    // an operand whose batch strides are all zero is split once and shared by every batch
    bool a_b = false, b_b = false;
    for (int d = 0; d < P.bd.nbd; ++d) {
        if (P.bd.sa[d] != 0) a_b = true;
        if (P.bd.sb[d] != 0) b_b = true;
    }
    const uint64_t nbA = a_b ? nb : 1, nbB = b_b ? nb : 1;
    // Vector norms, one launch per operand over all its batches: rows of A (elements
    // along k at stride csa), columns of B (along k at stride rsb). A wave per vector
    // when the elements are the closer ones in memory, else a lane per vector with
    // the length split into chunks until the grid has ~2048 blocks.
    struct NormPlan {
        const double* x;
        int64_t nvec, sv, st, evs;
        uint64_t nbx;
        int which;
        bool rows;
        int64_t nchunk, chunk;
    };
    auto plan_norm = [&](const double* x, int64_t nvec, int64_t sv, int64_t st, int64_t evs, uint64_t nbx, int which) {
        NormPlan q{x, nvec, sv, st, evs, nbx, which, std::llabs(st) <= std::llabs(sv), 1, K};
        if (!q.rows) {
            const int64_t blocks = (nvec + 63) / 64 * (int64_t)nbx;
            q.nchunk = std::max<int64_t>(1, std::min<int64_t>({(2048 + blocks - 1) / blocks, K / 64, 4096}));
            q.chunk = (K + q.nchunk - 1) / q.nchunk;
            q.nchunk = (K + q.chunk - 1) / q.chunk;
        }
        return q;
    };
    const NormPlan qa = plan_norm((const double*)P.a, M, P.rsa, P.csa, Mp, nbA, 0);
    const NormPlan qb = plan_norm((const double*)P.b, N, P.csb, P.rsb, Np, nbB, 1);
    auto part_bytes = [](const NormPlan& q) {
        return q.nchunk > 1 ? (size_t)q.nbx * q.nchunk * q.nvec * sizeof(NormAcc) : (size_t)0;
    };
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
    const size_t o_part = take(std::max(part_bytes(qa), part_bytes(qb)));
    // residue rows padded to Np bytes: every tile's 4-byte residue words land inside their row
    const size_t planeA = (size_t)nbA * Mp * Kp, planeB = (size_t)nbB * Np * Kp, planeR = (size_t)nb * Mp * Np;
    const size_t o_A8 = take((size_t)n * planeA);
    const size_t o_B8 = sym ? o_A8 : take((size_t)n * planeB);
    const size_t o_R = take((size_t)n * planeR);
    char* ws = (char*)workspace(ctx, off);
    int* flag = (int*)(ws + o_flag);
    int* eA = (int*)(ws + o_eA);
    int* fB = sym ? eA : (int*)(ws + o_fB);
    NormAcc* part = (NormAcc*)(ws + o_part);
    int8_t* A8 = (int8_t*)(ws + o_A8);
    int8_t* B8 = (int8_t*)(ws + o_B8);
    uint8_t* R = (uint8_t*)(ws + o_R);
    const uint32_t tiles_n = (uint32_t)(Np / bn);
    const uint32_t ntiles = sym ? (uint32_t)lower_tile_prefix((uint64_t)(Mp / bm), (uint64_t)(bn / bm))
                                : (uint32_t)((Mp / bm) * tiles_n);

    hipLaunchKernelGGL(k_zero_i32, dim3(ceil_div_u32(n_zero, 256)), dim3(256), 0, ctx->stream, flag, (int64_t)n_zero);
    check_launch("ozaki zero");
    auto run_norm = [&](const NormPlan& q, int* ev) {
        if (q.rows) {
            hipLaunchKernelGGL(k_norm_rows, dim3(ceil_div_u32(q.nvec, 4), (uint32_t)q.nbx), dim3(256), 0, ctx->stream,
                               q.x, q.nvec, K, q.sv, q.st, P.bd, q.which, q.evs, g, ev, flag);
        } else {
            hipLaunchKernelGGL(k_norm_cols, dim3(ceil_div_u32(q.nvec, 64), (uint32_t)q.nchunk, (uint32_t)q.nbx),
                               dim3(256), 0, ctx->stream, q.x, q.nvec, K, q.sv, q.st, q.chunk, P.bd, q.which, q.evs, g,
                               part, ev, flag);
            if (q.nchunk > 1)
                hipLaunchKernelGGL(k_norm_fin, dim3(ceil_div_u32(q.nvec, 256), (uint32_t)q.nbx), dim3(256), 0,
                                   ctx->stream, part, q.nvec, (int)q.nchunk, q.evs, g, ev);
        }
        check_launch("ozaki norms");
    };
    // vectors with contiguous elements: norm and split in one pass (k_normsplit_rows_reg, or
    // k_normsplit_rows past 8192 elements)
    auto norm_split = [&](const NormPlan& q, const double* x, int64_t nv, int64_t nvp, int64_t sv, int64_t st,
                          int which, int* ev, int8_t* planes, size_t plane, const char* what) {
        if (st == 1 && Kp <= 8192) {
            // vectors held in registers, read once: four per block (one per wave) up to 2048
            // elements, one per block up to 8192 (measured: 8192^2 A 380-404 -> 310-333 us,
            // 4096^2 102-120 -> 80-102 us, 65536 x 1024 317-378 -> 284-332 us)
            auto go = [&](auto rpb, auto ch) {
                constexpr int RPB = decltype(rpb)::value, CH = decltype(ch)::value;
                auto kern = n == OZ_MAXMOD ? k_normsplit_rows_reg<true, RPB, CH> : k_normsplit_rows_reg<false, RPB, CH>;
                hipLaunchKernelGGL(kern, dim3(ceil_div_u32(nvp, RPB), (uint32_t)q.nbx), dim3(256), 0, ctx->stream, x,
                                   nv, K, sv, P.bd, which, nvp, Kp, g, ev, planes, (int64_t)plane, ms, flag);
            };
            using I1 = std::integral_constant<int, 1>;
            using I2 = std::integral_constant<int, 2>;
            using I4 = std::integral_constant<int, 4>;
            if (Kp <= 1024) go(I4{}, I1{});
            else if (Kp <= 2048) go(I4{}, I2{});
            else if (Kp <= 4096) go(I1{}, I1{});
            else go(I1{}, I2{});
            check_launch(what);
            return;
        }
        if (st == 1) {
            hipLaunchKernelGGL(n == OZ_MAXMOD ? k_normsplit_rows<true> : k_normsplit_rows<false>,
                               dim3(ceil_div_u32(nvp, 4), (uint32_t)q.nbx), dim3(256), 0,
                               ctx->stream, x, nv, K, sv, P.bd, which, nvp, Kp, g, ev, planes, (int64_t)plane, ms,
                               flag);
            check_launch(what);
            return;
        }
        run_norm(q, ev);
        // tpb tiles along K per block while the grid keeps at least eight blocks per CU
        const uint64_t tiles = (uint64_t)(Kp / 64) * (uint64_t)(nvp / 64) * q.nbx;
        // Sources strided along t (B's columns of a row-major B): consecutive blocks take
        // neighbouring vector tiles, so that the rows they read lie side by side (measured on
        // 8192^3: 400-425 us against 460-540 with the t tiles consecutive and one per block)
        int tpb = 1;
        const int swap = std::llabs(st) == 1 ? 0 : 1;
        while (tpb < 8 && tiles / (uint64_t)(2 * tpb) >= (uint64_t)std::max(ctx->num_cu, 1) * 8) tpb *= 2;
        const uint32_t gx = (uint32_t)((Kp / 64 + tpb - 1) / tpb), gy = (uint32_t)(nvp / 64);
        hipLaunchKernelGGL(n == OZ_MAXMOD ? k_split<true> : k_split<false>,
                           dim3(swap ? gy : gx, swap ? gx : gy, (uint32_t)q.nbx),
                           dim3(256), 0, ctx->stream, x, nv, K, sv, st, P.bd, which, ev, nvp, Kp, planes, ms,
                           std::llabs(st) == 1 ? 1 : 0, flag, (int64_t)plane, tpb, swap);
        check_launch(what);
    };
    norm_split(qa, (const double*)P.a, M, Mp, P.rsa, P.csa, 0, eA, A8, planeA, "ozaki split A");
    if (!sym)  // the B norms reuse the partials buffer after the A norms (stream order)
        norm_split(qb, (const double*)P.b, N, Np, P.csb, P.rsb, 1, fB, B8, planeB, "ozaki split B");
    // Non-finite inputs set *flag: the split kernels exit at once and the plain
    // FP64 kernel (which runs only when *flag != 0) computes C with IEEE special
    // values. No host round trip, so the sequence can be graph-captured.
    plain_dgemm(ctx, P, alpha, beta, beta_zero, flag);
    ModList mods{};
    for (int l = 0; l < n; ++l) {
        mods.m[l] = (int)cc.mod[l];
        mods.c16[l] = ms.c16[l];
        mods.inv[l] = ms.inv[l];
    }
    // K beyond I8_KCHUNK: further launches add their residues into R
    for (int64_t k0 = 0; k0 < Kp; k0 += I8_KCHUNK) {
        const int64_t Kc = std::min<int64_t>(I8_KCHUNK, Kp - k0);
        auto go = [&](auto cfg) {
            using CF = decltype(cfg);
            // the 256 x 256 tile stores its residues four columns at a time
            constexpr bool TOUT = CF::BM == 256;
            // direct-load tiles over one K chunk run gmod moduli per block (a divisor of the
            // modulus count): as many as keep a block within 96 K slices and the launch at four
            // blocks per CU or more. Measured: the 65536 x 1024 x 1024 batch (8 slices per
            // modulus) 1.28 ms at 7 or 14 moduli per block against 1.38-1.44 ms at one; 8192^3
            // (64 slices) the same at one or two.
            int gmod = 1;
            const bool mm = CF::ST > 0 && TOUT && Kp <= I8_KCHUNK;
            if (mm) {
                const uint64_t blocks = (uint64_t)ntiles * nb;
                const int64_t ns = Kp / CF::BK;
                for (int g = 1; g <= n; ++g)
                    if (n % g == 0 && g * ns <= 96 && blocks * (uint64_t)(n / g) >= 1024) gmod = g;
            }
            auto kern = mm ? k_i8gemm_mod<CF, I8_ROW4G, TOUT, CF::ST != 0 && TOUT>
                           : k_i8gemm_mod<CF, I8_ROW4G, TOUT, false>;
            hipLaunchKernelGGL(kern, dim3(ntiles, (uint32_t)nb, (uint32_t)((n + gmod - 1) / gmod)), dim3(CF::NT), 0,
                               ctx->stream, A8 + k0, B8 + k0, R, Mp, N, Kp, Kc, Np, mods, k0 > 0 ? 1 : 0,
                               (uint32_t)(Mp / bm), tiles_n, a_b ? Mp * Kp : (int64_t)0,
                               b_b ? Np * Kp : (int64_t)0, Mp * Np, (int64_t)planeA, (int64_t)planeB,
                               (int64_t)planeR, flag, sym ? 1 : 0, gmod, n);
        };
        with_i8_tile(tile, go);
        check_launch("ozaki int8 gemm");
    }
    // four columns per thread when every row and plane of R starts on a 4-byte boundary
    auto crt = [&](auto v) {
        constexpr int V = decltype(v)::value;
        // a row's column groups over tpr threads (no idle lanes for short rows); each block
        // walks about four row groups (rows i, i + gridDim.y * rpb, ...): fewer, longer blocks
        const int64_t ncg = (N + V - 1) / V;
        int tpr = 1;
        while (tpr < 256 && tpr < ncg) tpr *= 2;
        const int64_t rpb = 256 / tpr;
        const int64_t gy = std::min<int64_t>(std::max<int64_t>(1, (M + 4 * rpb - 1) / (4 * rpb)), 65535);
        auto kern = n == OZ_MAXMOD ? k_crt<V, true> : k_crt<V, false>;
        hipLaunchKernelGGL(kern, dim3(ceil_div_u32(ncg, tpr), (uint32_t)gy, (uint32_t)nb), dim3(256), 0,
                           ctx->stream, R, M, N, Mp, Np, Np, (int64_t)planeR, a_b ? 1 : 0, b_b ? 1 : 0, P.bd, cc, eA, fB,
                           (double*)P.c, P.rsc, P.csc, alpha, beta, beta_zero ? 1 : 0, flag, tpr);
    };
    if (!sym) {
        // R rows are Np bytes long (a multiple of 128): four columns per thread, one 4-byte word
        // per plane (16 columns per thread measured slower: each lane's 16 doubles land 128 bytes
        // apart; 32 x 128 tiles as fast, 64 x 64 tiles a third slower)
        crt(std::integral_constant<int, 4>{});
        check_launch("ozaki crt");
    } else {
        // symmetric product: the lower tiles' CRT writes each tile's mirror too (RI 24000^2:
        // 1.70-1.86 ms against 1.77-1.95 ms for the row CRT plus 1.07-1.09 ms for k_mirror)
        using CT = CrtTile<128>;
        const uint64_t tmt = (uint64_t)((M + CT::TR - 1) / CT::TR), tnt = (uint64_t)((N + CT::TC - 1) / CT::TC);
        const uint64_t ntl = lower_tile_prefix(tmt, (uint64_t)(CT::TC / CT::TR));
        if (ntl >= (1ull << 31) || nb > 65535) RSTSR_RAISE(UnImplemented, "split GEMM output too large");
        // the kernel reads residue rows and exponents of every tile row and column it runs
        if ((int64_t)(tmt * CT::TR) > Mp || (int64_t)(tnt * CT::TC) > Np)
            RSTSR_RAISE(RuntimeError, "CRT tiles exceed the padded planes");
        auto kern = n == OZ_MAXMOD ? k_crt_tile<true, true, CT::TC> : k_crt_tile<false, true, CT::TC>;
        hipLaunchKernelGGL(kern, dim3((uint32_t)ntl, (uint32_t)nb), dim3(256), 0, ctx->stream, R, M, N, Mp, Np, Np,
                           (int64_t)planeR, a_b ? 1 : 0, b_b ? 1 : 0, P.bd, cc, eA, fB, (double*)P.c, P.rsc, P.csc,
                           alpha, beta, beta_zero ? 1 : 0, flag, (uint32_t)tmt, (uint32_t)tnt);
        check_launch("ozaki crt");
    }
}

// Grouped split GEMM routing (RSTSR_HIP_OZAKI_GROUPED: "auto", the default, or "split").
// The split pays for its memory passes (per input element an f64 read and 14 plane
// bytes written, per output element 14 plane bytes read and an f64 written) only
// when the 14 INT8 products outrun one FP64 product; with the rates this build
// measures on MI355X (FP64 MFMA ~60 TFLOP/s, INT8 2.0-2.3 POP/s on large tiles,
// ~5 TB/s on the memory passes) that needs K and the output large enough to fill
// the GPU by themselves. "auto" sends such problems through the single-problem
// split path and computes the rest in one plain grouped FP64 launch (the reference
// arithmetic, so no accuracy is lost); for the short products of a typical group
// (sides 32-1024) the grouped split kernels are slower than FP64 at every size.
// "split" forces every problem through the grouped split kernels.
static bool ozaki_grouped_forced() {
    const char* s = std::getenv("RSTSR_HIP_OZAKI_GROUPED");
    if (!s || !*s || std::strcmp(s, "auto") == 0) return false;
    if (std::strcmp(s, "split") == 0) return true;
    RSTSR_RAISE(InvalidValue, "RSTSR_HIP_OZAKI_GROUPED must be auto or split");
}
static bool split_pays(const GemmProblem& P) {
    const double M = (double)P.M, N = (double)P.N, K = (double)P.K;
    if (M * N < 4194304.0) return false;  // fewer than 64 256 x 256 tiles per modulus
    const double t_plain = 2.0 * M * N * K / 60e12;
    const double t_split = 22.0 * (M * K + N * K + M * N) / 5e12 + 28.0 * M * N * K / 2.0e15;
    return t_split < t_plain;
}

void ozaki_grouped_f64(rstsr_hip_ctx* ctx, const std::vector<GemmProblem>& probs, double alpha, double beta,
                       bool beta_zero) {
    if (!ozaki_grouped_forced()) {
        std::vector<GemmProblem> plain;
        for (const GemmProblem& P : probs) {
            if (split_pays(P)) ozaki_gemm_f64(ctx, P, alpha, beta, beta_zero);
            else plain.push_back(P);
        }
        plain_grouped_f64(ctx, plain, alpha, beta, beta_zero, nullptr);
        return;
    }
    // a K longer than one INT8 launch can sum goes through the single-problem path
    std::vector<GemmProblem> keep;
    for (const GemmProblem& P : probs) {
        if ((P.K + I8_KPAD - 1) / I8_KPAD * I8_KPAD > I8_KCHUNK) ozaki_gemm_f64(ctx, P, alpha, beta, beta_zero);
        else keep.push_back(P);
    }
    if (keep.size() != probs.size()) return ozaki_grouped_f64(ctx, keep, alpha, beta, beta_zero);
    if (probs.empty()) return;
    const OzSetup su = ozaki_setup();
    const int n = su.n;
    const int count = (int)probs.size();
    if (probs.size() >= (1u << 30)) RSTSR_RAISE(UnImplemented, "too many problems for the grouped split GEMM");
    std::vector<OzProb> op(probs.size());
    std::vector<uint32_t> pre((size_t)4 * (count + 1));
    uint32_t* pn = pre.data();
    uint32_t* psp = pn + (count + 1);
    uint32_t* pg = psp + (count + 1);
    uint32_t* pc = pg + (count + 1);
    uint64_t tn = 0, tsp = 0, tg = 0, tc = 0;
    int64_t sA = 0, sB = 0, sR = 0, se = 0, sf = 0;
    for (int i = 0; i < count; ++i) {
        const GemmProblem& P = probs[i];
        if (P.M <= 0 || P.N <= 0 || P.K <= 0 || P.nbatch != 1)
            RSTSR_RAISE(InvalidValue, "grouped split GEMM needs non-empty unbatched problems");
        if (P.K > (int64_t)1 << 40) RSTSR_RAISE(UnImplemented, "K too large for the split GEMM");
        if ((uint64_t)P.M * (uint64_t)P.N + 255 >= (1ull << 32))
            RSTSR_RAISE(UnImplemented, "grouped split GEMM problem too large; run it on its own");
        OzProb& o = op[i];
        o.fdN = FastDiv((uint32_t)P.N);
        o.M = P.M;
        o.N = P.N;
        o.K = P.K;
        o.Mp = (P.M + I8_MPAD - 1) / I8_MPAD * I8_MPAD;
        o.Np = (P.N + I8Sq::BN - 1) / I8Sq::BN * I8Sq::BN;
        o.Kp = (P.K + I8_KPAD - 1) / I8_KPAD * I8_KPAD;
        o.a = (const double*)P.a;
        o.rsa = P.rsa;
        o.csa = P.csa;
        o.b = (const double*)P.b;
        o.rsb = P.rsb;
        o.csb = P.csb;
        o.c = (double*)P.c;
        o.rsc = P.rsc;
        o.csc = P.csc;
        o.oA = sA;
        o.oB = sB;
        o.oR = sR;
        o.oe = se;
        o.of = sf;
        sA += o.Mp * o.Kp;
        sB += o.Np * o.Kp;
        sR += o.Mp * o.N;
        se += o.Mp;
        sf += o.Np;
        pn[i] = (uint32_t)tn;
        tn += (uint64_t)((P.M + 63) / 64 + (P.N + 63) / 64);
        psp[i] = (uint32_t)tsp;
        tsp += (uint64_t)((o.Mp + o.Np) / 64) * (uint64_t)(o.Kp / 64);
        pg[i] = (uint32_t)tg;
        tg += (uint64_t)(o.Mp / I8Sq::BM) * (uint64_t)(o.Np / I8Sq::BN);
        pc[i] = (uint32_t)tc;
        tc += ((uint64_t)P.M * (uint64_t)P.N + 255) / 256;
        if (tn >= (1ull << 31) || tsp >= (1ull << 31) || tg >= (1ull << 31) || tc >= (1ull << 31))
            RSTSR_RAISE(UnImplemented, "grouped split GEMM too large for one launch");
    }
    pn[count] = (uint32_t)tn;
    psp[count] = (uint32_t)tsp;
    pg[count] = (uint32_t)tg;
    pc[count] = (uint32_t)tc;
    // workspace: flag | eA [se] | fB [sf] | problems | prefix tables | A8 [l][sA] | B8 [l][sB] | R [l][sR]
    size_t off = sizeof(int);
    const size_t o_eA = off;
    off += (size_t)se * sizeof(int);
    const size_t o_fB = off;
    off += (size_t)sf * sizeof(int);
    const size_t n_zero = off / sizeof(int);
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        size_t o = off;
        off += bytes;
        return o;
    };
    const size_t o_pr = take(op.size() * sizeof(OzProb));
    const size_t o_pre = take(pre.size() * sizeof(uint32_t));
    const size_t o_A8 = take((size_t)n * sA);
    const size_t o_B8 = take((size_t)n * sB);
    const size_t o_R = take((size_t)n * sR);
    char* ws = (char*)workspace(ctx, off);
    int* flag = (int*)ws;
    int* eA = (int*)(ws + o_eA);
    int* fB = (int*)(ws + o_fB);
    const OzProb* dprobs = (const OzProb*)(ws + o_pr);
    const uint32_t* dpre = (const uint32_t*)(ws + o_pre);
    int8_t* A8 = (int8_t*)(ws + o_A8);
    int8_t* B8 = (int8_t*)(ws + o_B8);
    uint8_t* R = (uint8_t*)(ws + o_R);
    {
        // problems and prefix tables: one upload of [o_pr, o_pre + prefix bytes)
        const size_t span = o_pre - o_pr + pre.size() * sizeof(uint32_t);
        char* host = (char*)meta_stage(ctx, span, "grouped split GEMM");
        std::memcpy(host, op.data(), op.size() * sizeof(OzProb));
        std::memcpy(host + (o_pre - o_pr), pre.data(), pre.size() * sizeof(uint32_t));
        meta_upload(ctx, ws + o_pr, span);
    }
    hipLaunchKernelGGL(k_zero_i32, dim3(ceil_div_u32(n_zero, 256)), dim3(256), 0, ctx->stream, flag, (int64_t)n_zero);
    check_launch("grouped ozaki zero");
    hipLaunchKernelGGL(k_gnorm, dim3((uint32_t)tn), dim3(256), 0, ctx->stream, dprobs, count, dpre + 0 * (count + 1),
                       su.g, eA, fB, flag);
    check_launch("grouped ozaki norms");
    hipLaunchKernelGGL(n == OZ_MAXMOD ? k_gsplit<true> : k_gsplit<false>, dim3((uint32_t)tsp), dim3(256), 0,
                       ctx->stream, dprobs, count, dpre + 1 * (count + 1),
                       eA, fB, A8, B8, (int64_t)sA, (int64_t)sB, su.ms, flag);
    check_launch("grouped ozaki split");
    // non-finite inputs anywhere in the group: the plain grouped kernel computes every product
    plain_grouped_f64(ctx, probs, alpha, beta, beta_zero, flag);
    hipLaunchKernelGGL((k_gi8gemm<I8Sq, I8_ROW4G>), dim3((uint32_t)tg, (uint32_t)n), dim3(I8Sq::NT), 0, ctx->stream,
                       dprobs, count,
                       dpre + 2 * (count + 1), (uint32_t)tg, A8, B8, R, (int64_t)sA, (int64_t)sB, (int64_t)sR, su.ms,
                       flag);
    check_launch("grouped ozaki int8 gemm");
    // (constants from the kernel arguments: one thread per element has no loop to hoist them out of,
    // and materializing the 42 limbs per element measured twice as slow)
    hipLaunchKernelGGL(k_gcrt<false>, dim3((uint32_t)tc), dim3(256), 0, ctx->stream,
                       dprobs, count, dpre + 3 * (count + 1), R,
                       (int64_t)sR, su.cc, eA, fB, alpha, beta, beta_zero ? 1 : 0, flag);
    check_launch("grouped ozaki crt");
}

}  // namespace hip
}  // namespace rstsr
