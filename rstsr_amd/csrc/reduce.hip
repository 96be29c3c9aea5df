// Axis and all-element reductions for gfx950: sum, max, min, argmax, argmin.
//
// Replaces rstsr-native-impl/src/cpu_rayon/reduction.rs:20-631 (reduce_all,
// reduce_axes, the arg variants) behind the traits of
// rstsr-core/src/operators/reduction.rs:3-95.
//
// Semantics kept from the reference:
//  - sum: T::zero + x ... (auto_impl/reduction.rs:14-40). Floating sums are
//    accumulated here in FP64 with TwoSum error compensation (f32 inputs too),
//    so the result is the exact sum rounded once, up to O(n u^2) terms;
//    integers wrap like Rust release builds.
//  - max/min: start from T::MIN / T::MAX and use IEEE maxNum/minNum, so NaN is
//    ignored (rstsr-dtype-traits/src/ext_real.rs:70-87, auto_impl/reduction.rs:94-128).
//  - argmax/argmin: the first extreme in row-major order over the reduced axes
//    in the order given; a NaN at logical position 0 wins, later NaNs are
//    ignored (cpu_serial/reduction.rs:425-580). Indices are u64 raveled over the
//    reduced shape. Every combine below uses (value, smaller logical index), a
//    total order, so the result does not depend on how work is split.
//
// Kernels (host planning in reduce_plan):
//  - row: one 64-lane wave per (output, split); lanes stride a unit-stride
//    reduced axis (coalesced 512 B per load wave), wave shuffles combine;
//  - col: one thread per (output, split) with outputs along the input's
//    unit-stride kept axis (coalesced across lanes), sequential over the
//    reduced range;
//  - finalize: combine the split partials of each output in fixed order.
// Built with -ffp-contract=off so that TwoSum is not fused or reassociated.
#include <array>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "common.hpp"
#include "registry.hpp"

namespace rstsr {
namespace hip {

/* =============================== accumulators =============================== */

struct DD {
    double s, c;
};
__device__ __forceinline__ void dd_add(DD& a, double x) {
    double t = a.s + x;
    double bp = t - a.s;
    double e = (a.s - (t - bp)) + (x - bp);
    a.s = t;
    a.c += e;
}
__device__ __forceinline__ DD dd_combine(DD a, const DD& b) {
    dd_add(a, b.s);
    a.c += b.c;
    return a;
}
__device__ __forceinline__ double shfl_d(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int m) {
    return (uint64_t)__shfl_xor((long long)v, m, 64);
}

template <class T>
struct SumF {  // f32, f64
    using Val = T;
    using Out = T;
    using S = DD;
    __device__ static S init() { return {0.0, 0.0}; }
    __device__ static void add(S& s, T x, uint64_t) { dd_add(s, (double)x); }
    __device__ static S combine(S a, S b) { return dd_combine(a, b); }
    __device__ static S shfl(S s, int m) { return {shfl_d(s.s, m), shfl_d(s.c, m)}; }
    __device__ static Out fin(S s, T, uint64_t) { return (T)(s.s + s.c); }
};

template <class T, class R>  // complex: T = c64/c128, R = float/double
struct SumC {
    using Val = T;
    using Out = T;
    struct S {
        DD re, im;
    };
    __device__ static S init() { return {{0.0, 0.0}, {0.0, 0.0}}; }
    __device__ static void add(S& s, T x, uint64_t) {
        dd_add(s.re, (double)x.re);
        dd_add(s.im, (double)x.im);
    }
    __device__ static S combine(S a, S b) { return {dd_combine(a.re, b.re), dd_combine(a.im, b.im)}; }
    __device__ static S shfl(S s, int m) {
        return {{shfl_d(s.re.s, m), shfl_d(s.re.c, m)}, {shfl_d(s.im.s, m), shfl_d(s.im.c, m)}};
    }
    __device__ static Out fin(S s, T, uint64_t) { return T{(R)(s.re.s + s.re.c), (R)(s.im.s + s.im.c)}; }
};

template <class T>
struct SumI {  // i32, i64: wrapping
    using Val = T;
    using Out = T;
    using U = typename std::make_unsigned<T>::type;
    using S = U;
    __device__ static S init() { return 0; }
    __device__ static void add(S& s, T x, uint64_t) { s = (U)(s + (U)x); }
    __device__ static S combine(S a, S b) { return (U)(a + b); }
    __device__ static S shfl(S s, int m) {
        if constexpr (sizeof(T) == 8) return (U)shfl_u64((uint64_t)s, m);
        else return (U)__shfl_xor((int)s, m, 64);
    }
    __device__ static Out fin(S s, T, uint64_t) { return (T)s; }
};

template <class T>
__device__ __forceinline__ T shfl_val(T v, int m) {
    if constexpr (sizeof(T) == 8) {
        if constexpr (std::is_floating_point<T>::value) return __shfl_xor(v, m, 64);
        else return (T)__shfl_xor((long long)v, m, 64);
    } else {
        if constexpr (std::is_floating_point<T>::value) return __shfl_xor(v, m, 64);
        else return (T)__shfl_xor((int)v, m, 64);
    }
}

template <class T>
__device__ __forceinline__ T lowest_v() {
    if constexpr (std::is_same<T, float>::value) return -FLT_MAX;
    if constexpr (std::is_same<T, double>::value) return -DBL_MAX;
    if constexpr (std::is_same<T, int32_t>::value) return INT_MIN;
    if constexpr (std::is_same<T, int64_t>::value) return LLONG_MIN;
}
template <class T>
__device__ __forceinline__ T highest_v() {
    if constexpr (std::is_same<T, float>::value) return FLT_MAX;
    if constexpr (std::is_same<T, double>::value) return DBL_MAX;
    if constexpr (std::is_same<T, int32_t>::value) return INT_MAX;
    if constexpr (std::is_same<T, int64_t>::value) return LLONG_MAX;
}

template <class T, bool MAX>
struct Ext {  // max / min, NaN ignored
    using Val = T;
    using Out = T;
    using S = T;
    __device__ static S init() { return MAX ? lowest_v<T>() : highest_v<T>(); }
    __device__ static void add(S& s, T x, uint64_t) {
        if (MAX ? (x > s) : (x < s)) s = x;
    }
    __device__ static S combine(S a, S b) { return (MAX ? (b > a) : (b < a)) ? b : a; }
    __device__ static S shfl(S s, int m) { return shfl_val<T>(s, m); }
    __device__ static Out fin(S s, T, uint64_t) { return s; }
};

template <class T, bool MAX>
struct Arg {  // argmax / argmin on (value, logical index)
    static constexpr bool ARG = true;
    using Val = T;
    using Out = uint64_t;
    struct S {
        T v;
        uint64_t i;  // UINT64_MAX: no candidate yet
    };
    __device__ static S init() { return {T{}, ~0ull}; }
    __device__ static bool better(T x, uint64_t li, const S& s) {
        if (s.i == ~0ull) return true;
        if (MAX ? (x > s.v) : (x < s.v)) return true;
        return x == s.v && li < s.i;
    }
    __device__ static void add(S& s, T x, uint64_t li) {
        if constexpr (std::is_floating_point<T>::value)
            if (x != x) return;
        if (better(x, li, s)) {
            s.v = x;
            s.i = li;
        }
    }
    __device__ static S combine(S a, S b) {
        if (b.i == ~0ull) return a;
        return better(b.v, b.i, a) ? b : a;
    }
    __device__ static S shfl(S s, int m) { return {shfl_val<T>(s.v, m), shfl_u64(s.i, m)}; }
    // Lane-local scan of a run whose logical indices increase with position:
    // strictly better values replace, so ties keep the earliest position; NaN is
    // skipped as in add.
    __device__ static void scan(T x, uint32_t e, T& bv, uint32_t& be) {
        if constexpr (std::is_floating_point<T>::value)
            if (x != x) return;
        if (be == ~0u || (MAX ? (x > bv) : (x < bv))) {
            bv = x;
            be = e;
        }
    }
    // x0: the element at logical index 0 of this output's reduced space
    __device__ static Out fin(S s, T x0, uint64_t) {
        if constexpr (std::is_floating_point<T>::value)
            if (x0 != x0) return 0;
        return s.i == ~0ull ? 0 : s.i;
    }
};

// prod (auto_impl/reduction.rs:131-165): init one, acc * x. Floating products
// run in f64 (complex: (ac-bd, ad+bc) in f64); integers wrap.
template <class T>
struct ProdF {
    using Val = T;
    using Out = T;
    using S = double;
    __device__ static S init() { return 1.0; }
    __device__ static void add(S& s, T x, uint64_t) { s *= (double)x; }
    __device__ static S combine(S a, S b) { return a * b; }
    __device__ static S shfl(S s, int m) { return shfl_d(s, m); }
    __device__ static Out fin(S s, T, uint64_t) { return (T)s; }
};

template <class T, class R>
struct ProdC {
    using Val = T;
    using Out = T;
    struct S {
        double re, im;
    };
    __device__ static S init() { return {1.0, 0.0}; }
    __device__ static S mul(S a, S b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
    __device__ static void add(S& s, T x, uint64_t) { s = mul(s, S{(double)x.re, (double)x.im}); }
    __device__ static S combine(S a, S b) { return mul(a, b); }
    __device__ static S shfl(S s, int m) { return {shfl_d(s.re, m), shfl_d(s.im, m)}; }
    __device__ static Out fin(S s, T, uint64_t) { return T{(R)s.re, (R)s.im}; }
};

template <class T>
struct ProdI {
    using Val = T;
    using Out = T;
    using U = typename std::make_unsigned<T>::type;
    using S = U;
    __device__ static S init() { return 1; }
    __device__ static void add(S& s, T x, uint64_t) { s = (U)(s * (U)x); }
    __device__ static S combine(S a, S b) { return (U)(a * b); }
    __device__ static S shfl(S s, int m) {
        if constexpr (sizeof(T) == 8) return (U)shfl_u64((uint64_t)s, m);
        else return (U)__shfl_xor((int)s, m, 64);
    }
    __device__ static Out fin(S s, T, uint64_t) { return (T)s; }
};

// all / any of a bool array (auto_impl/reduction.rs:466-519): fold with && / || from true /
// false. Elements are read as bytes (nonzero is true).
template <bool ALL>
struct BoolFold {
    using Val = uint8_t;
    using Out = uint8_t;
    using S = uint32_t;
    __device__ static S init() { return ALL ? 1u : 0u; }
    __device__ static void add(S& s, uint8_t x, uint64_t) {
        if (ALL) s &= (x != 0);
        else s |= (x != 0);
    }
    __device__ static S combine(S a, S b) { return ALL ? (a & b) : (a | b); }
    __device__ static S shfl(S s, int m) { return (S)__shfl_xor((int)s, m, 64); }
    __device__ static Out fin(S s, uint8_t, uint64_t) { return (Out)s; }
};

// count_nonzero (auto_impl/reduction.rs:521-551): elements != T::zero() counted into a usize;
// a bool array's sum (OpSumBoolAPI, :678-715) is the same count. A complex element is nonzero
// when either part is; NaN is nonzero, -0.0 is zero.
template <class T>
struct CountNZ {
    using Val = T;
    using Out = uint64_t;
    using S = uint64_t;
    __device__ static bool nz(T x) {
        if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) return x.re != 0 || x.im != 0;
        else return x != 0;
    }
    __device__ static S init() { return 0; }
    __device__ static void add(S& s, T x, uint64_t) { s += nz(x) ? 1 : 0; }
    __device__ static S combine(S a, S b) { return a + b; }
    __device__ static S shfl(S s, int m) { return shfl_u64(s, m); }
    __device__ static Out fin(S s, T, uint64_t) { return s; }
};

// mean (auto_impl/reduction.rs:167-205): the compensated sum divided by the
// number of reduced elements.
template <class T>
struct MeanF : SumF<T> {
    using S = typename SumF<T>::S;
    __device__ static T fin(S s, T, uint64_t n) { return (T)((s.s + s.c) / (double)n); }
};

template <class T, class R>
struct MeanC : SumC<T, R> {
    using S = typename SumC<T, R>::S;
    __device__ static T fin(S s, T, uint64_t n) {
        return T{(R)((s.re.s + s.re.c) / (double)n), (R)((s.im.s + s.im.c) / (double)n)};
    }
};

// Exact square into a double-double: x*x plus its rounding error.
__device__ __forceinline__ void dd_add_sq(DD& a, double x) {
    const double p = x * x;
    dd_add(a, p);
    a.c += fma(x, x, -p);
}

// var / std (auto_impl/reduction.rs:207-315) and l2_norm (:317-354) on the
// real type R: compensated sums of x and |x|^2, then the reference's formulas
// var = sum|x|^2 / n - |sum x / n|^2, std = sqrt(var), l2 = sqrt(sum |x|^2).
// KIND 0: var, 1: std, 2: l2_norm. T real (R = T) or complex (R its part type).
template <class T, class R, int KIND>
struct Moments {
    static constexpr bool CX = !std::is_same<T, R>::value;
    using Val = T;
    using Out = R;
    struct S {
        DD re, im, q;
    };
    __device__ static S init() { return {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}; }
    __device__ static void add(S& s, T x, uint64_t) {
        if constexpr (CX) {
            if constexpr (KIND != 2) {
                dd_add(s.re, (double)x.re);
                dd_add(s.im, (double)x.im);
            }
            dd_add_sq(s.q, (double)x.re);
            dd_add_sq(s.q, (double)x.im);
        } else {
            if constexpr (KIND != 2) dd_add(s.re, (double)x);
            dd_add_sq(s.q, (double)x);
        }
    }
    __device__ static S combine(S a, S b) { return {dd_combine(a.re, b.re), dd_combine(a.im, b.im), dd_combine(a.q, b.q)}; }
    __device__ static S shfl(S s, int m) {
        return {{shfl_d(s.re.s, m), shfl_d(s.re.c, m)}, {shfl_d(s.im.s, m), shfl_d(s.im.c, m)},
                {shfl_d(s.q.s, m), shfl_d(s.q.c, m)}};
    }
    __device__ static Out fin(S s, T, uint64_t n) {
        const double q = s.q.s + s.q.c;
        if constexpr (KIND == 2) {
            return (R)sqrt(q);
        } else {
            const double dn = (double)n;
            const double mr = (s.re.s + s.re.c) / dn, mi = (s.im.s + s.im.c) / dn;
            const double var = q / dn - (mr * mr + mi * mi);
            return KIND == 0 ? (R)var : (R)sqrt(var);
        }
    }
};

template <class A, class = void>
struct IsArg : std::false_type {};
template <class A>
struct IsArg<A, std::void_t<decltype(A::ARG)>> : std::true_type {};
template <class A>
__device__ constexpr bool is_arg() {
    return IsArg<A>::value;
}

/* =============================== plan =============================== */

struct RedPlan {
    int nk = 0, nr = 0;
    uint32_t kshape[MAXD];
    FastDiv kfd[MAXD];
    int64_t kin[MAXD], kout[MAXD];
    uint64_t rshape[MAXD];
    int64_t rin[MAXD], rlog[MAXD];
    int64_t in_off = 0, out_off = 0;
    uint32_t M = 1;   // outputs
    uint64_t R = 1;   // reduced elements per output
    uint32_t nsplit = 1;
    FastDiv fsplit;
    // row kernel segmentation of the reduced space: units = rows(rest) x nseg0
    uint32_t seg0 = 1, nseg0 = 1;
    uint64_t units = 1, upw = 1;  // units per split
    int stream = 0;               // every input element is read once: streaming loads
};

__device__ __forceinline__ void kept_offsets(const RedPlan& p, uint32_t o, int64_t& oin, int64_t& oout) {
    oin = p.in_off;
    oout = p.out_off;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
        if (d >= p.nk) break;
        uint32_t q = fdiv(o, p.kfd[d]);
        uint32_t r = o - q * p.kshape[d];
        oin += (int64_t)r * p.kin[d];
        oout += (int64_t)r * p.kout[d];
        o = q;
    }
}

// offsets of reduced "row" u (dims 1..nr-1) : input offset and logical index
__device__ __forceinline__ void row_offsets(const RedPlan& p, uint64_t u, int64_t& oin, uint64_t& li) {
    oin = 0;
    li = 0;
#pragma unroll
    for (int d = 1; d < MAXD; ++d) {
        if (d >= p.nr) break;
        uint64_t q = u / p.rshape[d];
        uint64_t r = u - q * p.rshape[d];
        oin += (int64_t)r * p.rin[d];
        li += r * (uint64_t)p.rlog[d];
        u = q;
    }
}

constexpr int RB = 256;  // threads per block (4 waves)

// Reduced positions loaded per iteration by the column kernel, from the bytes one lane
// loads per position: 16 (i32/f32 pairs) or 8 (f64/i64 pairs, c128) positions, so that a
// wave keeps about 8 KiB in flight, and 4 for narrower loads. The deeper forms run with
// half as many waves (col_waves_per_cu): fewer, longer walks keep DRAM pages open along
// the reduced axis (measured on the strided column reductions: i32 argmax 51 -> 41 us).
__host__ __device__ constexpr int col_unroll(int lane_bytes) {
    return lane_bytes == 8 ? 16 : (lane_bytes >= 16 ? 8 : 4);
}
__host__ __device__ constexpr int col_waves_per_cu(int lane_bytes) { return col_unroll(lane_bytes) > 4 ? 4 : 8; }

// VEC consecutive elements loaded as one aligned access
template <class T, int VEC>
struct alignas(sizeof(T) * VEC) Pack {
    T v[VEC];
};

// One wave per (output, split) over rows of the unit-stride reduced axis; each
// lane loads VEC consecutive elements per access (16-byte loads when the host
// proved alignment), four accesses in flight per iteration.
template <class A, class T, int VEC>
__global__ void __launch_bounds__(RB) k_red_row(const T* __restrict__ a, typename A::Out* out,
                                                typename A::S* part, RedPlan p) {
    using S = typename A::S;
    using vt = Pack<T, VEC>;
    const uint32_t wave = blockIdx.x * (RB / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwork = (uint64_t)p.M * p.nsplit;
    if (wave >= nwork) return;  // whole wave exits together; no block barrier below
    const uint32_t o = fdiv(wave, p.fsplit);
    const uint32_t sp = wave - o * p.nsplit;
    int64_t oin, oout;
    kept_offsets(p, o, oin, oout);
    S acc = A::init();
    const uint64_t u0 = (uint64_t)sp * p.upw;
    const uint64_t u1 = min(p.units, u0 + p.upw);
    const int64_t s0 = p.rin[0];
    const uint64_t l0 = (uint64_t)p.rlog[0];
    const uint32_t L0 = (uint32_t)p.rshape[0];
    constexpr uint32_t W = 64 * VEC;  // elements per wave access
    auto sweep = [&](auto nt) {
        constexpr bool NT = decltype(nt)::value;
        auto get = [&](const T* base, uint32_t e) -> vt {
            if constexpr (VEC == 1) {
                vt x;
                x.v[0] = ld<NT>(base + (int64_t)e * s0);
                return x;
            } else {
                return ld<NT>((const vt*)(base + e));  // s0 == 1 on this path
            }
        };
        for (uint64_t u = u0; u < u1; ++u) {
            uint64_t row = u / p.nseg0;
            uint32_t seg = (uint32_t)(u - row * p.nseg0);
            int64_t roff;
            uint64_t rli;
            row_offsets(p, row, roff, rli);
            const T* base = a + oin + roff;
            uint32_t e = seg * p.seg0 + lane * VEC;
            uint32_t eend = min(L0, (seg + 1) * p.seg0);
            if constexpr (is_arg<A>()) {
                // positions increase with the logical index inside a row (l0 > 0)
                T bv{};
                uint32_t be = ~0u;
                for (; e + 3 * W < eend; e += 4 * W) {
                    vt x0 = get(base, e), x1 = get(base, e + W), x2 = get(base, e + 2 * W), x3 = get(base, e + 3 * W);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::scan(x0.v[v], e + v, bv, be);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::scan(x1.v[v], e + W + v, bv, be);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::scan(x2.v[v], e + 2 * W + v, bv, be);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::scan(x3.v[v], e + 3 * W + v, bv, be);
                }
                for (; e < eend; e += W) {
                    vt x = get(base, e);
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::scan(x.v[v], e + v, bv, be);
                }
                if (be != ~0u) A::add(acc, bv, rli + (uint64_t)be * l0);
                continue;
            }
            for (; e + 3 * W < eend; e += 4 * W) {
                vt x0 = get(base, e), x1 = get(base, e + W), x2 = get(base, e + 2 * W), x3 = get(base, e + 3 * W);
#pragma unroll
                for (int v = 0; v < VEC; ++v) A::add(acc, x0.v[v], rli + (uint64_t)(e + v) * l0);
#pragma unroll
                for (int v = 0; v < VEC; ++v) A::add(acc, x1.v[v], rli + (uint64_t)(e + W + v) * l0);
#pragma unroll
                for (int v = 0; v < VEC; ++v) A::add(acc, x2.v[v], rli + (uint64_t)(e + 2 * W + v) * l0);
#pragma unroll
                for (int v = 0; v < VEC; ++v) A::add(acc, x3.v[v], rli + (uint64_t)(e + 3 * W + v) * l0);
            }
            for (; e < eend; e += W) {
                vt x = get(base, e);
#pragma unroll
                for (int v = 0; v < VEC; ++v) A::add(acc, x.v[v], rli + (uint64_t)(e + v) * l0);
            }
        }
    };
    // streaming loads pay off for 8-byte and wider elements; narrower ones (also as
    // 16-byte vectors) measured slower
    if (sizeof(T) >= 8 && p.stream) sweep(std::true_type{});
    else sweep(std::false_type{});
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        S other = A::shfl(acc, m);
        // lower lane's partial precedes the higher lane's
        acc = (lane & m) ? A::combine(other, acc) : A::combine(acc, other);
    }
    if (lane == 0) {
        if (p.nsplit == 1) out[oout] = A::fin(acc, a[oin], p.R);
        else part[(uint64_t)o * p.nsplit + sp] = acc;
    }
}

// VEC adjacent outputs per thread along the unit-stride kept axis (VEC * sizeof(T)
// = 16 B vector loads when the host proved alignment), sequential over the reduced range.
template <class A, class T, int VEC>
__global__ void __launch_bounds__(RB) k_red_col(const T* __restrict__ a, typename A::Out* out,
                                                typename A::S* part, RedPlan p) {
    using S = typename A::S;
    using vt = Pack<T, VEC>;
    constexpr int COLU = col_unroll((int)(sizeof(T) * VEC));
    const uint64_t t = (uint64_t)blockIdx.x * RB + threadIdx.x;
    const uint32_t MV = p.M / VEC;
    if (t >= (uint64_t)MV * p.nsplit) return;
    const uint32_t sp = (uint32_t)(t / MV);
    const uint32_t o = (uint32_t)(t - (uint64_t)sp * MV) * VEC;
    int64_t oin, oout;
    kept_offsets(p, o, oin, oout);
    S acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] = A::init();
    const uint64_t chunk = p.upw;  // reduced elements per split
    uint64_t r0 = (uint64_t)sp * chunk;
    uint64_t r1 = min(p.R, r0 + chunk);
    if (r0 < r1) {
        // unravel the start position over the reduced dims
        uint64_t coord[MAXD];
        int64_t off = 0;
        uint64_t li = 0;
        {
            uint64_t u = r0;
#pragma unroll
            for (int d = 0; d < MAXD; ++d) {
                coord[d] = 0;
                if (d >= p.nr) continue;
                uint64_t q = u / p.rshape[d];
                coord[d] = u - q * p.rshape[d];
                off += (int64_t)coord[d] * p.rin[d];
                li += coord[d] * (uint64_t)p.rlog[d];
                u = q;
            }
        }
        const T* base = a + oin;
        uint64_t remaining = r1 - r0;
        const int64_t s0 = p.rin[0];
        const uint64_t l0 = (uint64_t)p.rlog[0];
        auto run_row = [&](auto nt, uint64_t n) {
            constexpr bool NT = decltype(nt)::value;
            const T* rowp = base + off;
            uint64_t j = 0;
            if constexpr (is_arg<A>()) {
                // j increases with the logical index along reduced axis 0 (l0 > 0)
                T bv[VEC];
                uint32_t be[VEC];
#pragma unroll
                for (int v = 0; v < VEC; ++v) be[v] = ~0u;
                for (; j + COLU <= n; j += COLU) {
                    vt x[COLU];
#pragma unroll
                    for (int q = 0; q < COLU; ++q) x[q] = ld<NT>((const vt*)(rowp + (int64_t)(j + q) * s0));
#pragma unroll
                    for (int q = 0; q < COLU; ++q)
#pragma unroll
                        for (int v = 0; v < VEC; ++v) A::scan(x[q].v[v], (uint32_t)(j + q), bv[v], be[v]);
                }
                for (; j < n; ++j) {
                    vt x = ld<NT>((const vt*)(rowp + (int64_t)j * s0));
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::scan(x.v[v], (uint32_t)j, bv[v], be[v]);
                }
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                    if (be[v] != ~0u) A::add(acc[v], bv[v], li + (uint64_t)be[v] * l0);
                return;
            }
            for (; j + COLU <= n; j += COLU) {
                vt x[COLU];
#pragma unroll
                for (int q = 0; q < COLU; ++q) x[q] = ld<NT>((const vt*)(rowp + (int64_t)(j + q) * s0));
#pragma unroll
                for (int q = 0; q < COLU; ++q)
#pragma unroll
                    for (int v = 0; v < VEC; ++v) A::add(acc[v], x[q].v[v], li + (j + q) * l0);
            }
            for (; j < n; ++j) {
                vt x = ld<NT>((const vt*)(rowp + (int64_t)j * s0));
#pragma unroll
                for (int v = 0; v < VEC; ++v) A::add(acc[v], x.v[v], li + j * l0);
            }
        };
        while (remaining > 0) {
            uint64_t n = min(remaining, p.rshape[0] - coord[0]);
            if (sizeof(T) >= 8 && p.stream) run_row(std::true_type{}, n);
            else run_row(std::false_type{}, n);
            remaining -= n;
            if (remaining == 0) break;
            // carry into dims >= 1
            off -= (int64_t)coord[0] * s0;
            li -= coord[0] * l0;
            coord[0] = 0;
#pragma unroll
            for (int d = 1; d < MAXD; ++d) {
                if (d >= p.nr) break;
                coord[d] += 1;
                off += p.rin[d];
                li += (uint64_t)p.rlog[d];
                if (coord[d] < p.rshape[d]) break;
                off -= (int64_t)coord[d] * p.rin[d];
                li -= coord[d] * (uint64_t)p.rlog[d];
                coord[d] = 0;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        // kept axis 0 has input stride +1 when VEC > 1
        if (p.nsplit == 1) out[oout + v * p.kout[0]] = A::fin(acc[v], a[oin + v], p.R);
        else part[(uint64_t)(o + v) * p.nsplit + sp] = acc[v];
    }
}

// One wave per output: lane l combines partials l, l+64, ... in order, then a
// fixed shuffle tree combines the lanes (deterministic for a given split).
template <class A, class T>
__global__ void __launch_bounds__(RB) k_red_fin(const T* __restrict__ a, typename A::Out* out,
                                                const typename A::S* part, RedPlan p) {
    using S = typename A::S;
    const uint32_t o = blockIdx.x * (RB / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (o >= p.M) return;  // whole wave exits together
    const S* pp = part + (uint64_t)o * p.nsplit;
    S acc = A::init();
    for (uint32_t s = lane; s < p.nsplit; s += 64) acc = A::combine(acc, pp[s]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        S other = A::shfl(acc, m);
        acc = (lane & m) ? A::combine(other, acc) : A::combine(acc, other);
    }
    if (lane == 0) {
        int64_t oin, oout;
        kept_offsets(p, o, oin, oout);
        out[oout] = A::fin(acc, a[oin], p.R);
    }
}

/* =============================== host planning =============================== */

enum class RedKind { Row, Col };

struct RedSetup {
    RedPlan p;
    RedKind kind = RedKind::Row;
    bool empty_out = false;     // no outputs
    bool empty_reduce = false;  // outputs exist, reduced extent 0
};

// la: input layout; axes: reduced axes (given order); lo: output layout (kept shape)
// Outputs per thread of the column kernel: two for the real 4- and 8-byte dtypes (one 8- or
// 16-byte load per position), one otherwise. Matches run_reduce's VC when the host proves
// alignment; the unaligned fallback keeps the same split count.
static int col_vec(rstsr_dtype_t t) {
    return (t == RSTSR_F32 || t == RSTSR_F64 || t == RSTSR_I32 || t == RSTSR_I64 || t == RSTSR_U64) ? 2 : 1;
}

// la: input layout; axes: reduced axes (given order); lo: output layout (kept shape);
// t: element dtype, which sets the column kernel's vector width, unroll and split count
static RedSetup reduce_plan(rstsr_hip_ctx* ctx, const Layout& la, const std::vector<ptrdiff_t>& axes, const Layout& lo,
                            rstsr_dtype_t t) {
    const size_t esize = dtype_size(t);
    RedSetup s;
    auto sp = la.dim_split_axes(axes);
    const Layout& ls = sp.first;
    const Layout& lm = sp.second;
    if (lo.shape != lm.shape) RSTSR_RAISE(InvalidLayout, "output layout shape does not match the kept axes");
    if (lm.size() == 0) {
        s.empty_out = true;
        return s;
    }
    if (ls.size() == 0) {
        s.empty_reduce = true;
        return s;
    }
    // logical (row-major over the reduced axes in the given order) strides
    Layout lsc = c_layout(ls.shape);
    // reduced dims: drop size 1, sort by |in stride|, merge when both strides chain
    struct RD {
        uint64_t n;
        int64_t in, lg;
    };
    std::vector<RD> rd;
    for (size_t i = 0; i < ls.ndim(); ++i)
        if (ls.shape[i] != 1) rd.push_back({ls.shape[i], ls.stride[i], lsc.stride[i]});
    std::stable_sort(rd.begin(), rd.end(), [](const RD& x, const RD& y) {
        if (std::llabs(x.in) != std::llabs(y.in)) return std::llabs(x.in) < std::llabs(y.in);
        return std::llabs(x.lg) < std::llabs(y.lg);
    });
    std::vector<RD> rm;
    for (auto& d : rd) {
        if (!rm.empty() && d.in == rm.back().in * (int64_t)rm.back().n && d.lg == rm.back().lg * (int64_t)rm.back().n) {
            rm.back().n *= d.n;
            continue;
        }
        rm.push_back(d);
    }
    if (rm.empty()) rm.push_back({1, 0, 0});
    // The kernels walk reduced dim 0 with 32-bit positions: an extent of 2^31 or more is
    // factored into n = f * g (f < 2^31 the inner dim, g the next; logical strides chain, so
    // the logical index order is unchanged). An extent without such a factor pair stays
    // unsupported.
    if (rm[0].n >= (1ull << 31)) {
        const uint64_t n = rm[0].n;
        uint64_t f = 0;
        for (uint64_t g = 2; g * g <= n && g < (1ull << 31); ++g)
            if (n % g == 0 && n / g < (1ull << 31)) {
                f = n / g;
                break;
            }
        if (f == 0) RSTSR_RAISE(UnImplemented, "a reduced extent of 2^31 or more without a factor below 2^31");
        const RD d0 = rm[0];
        rm[0] = {f, d0.in, d0.lg};
        rm.insert(rm.begin() + 1, RD{n / f, d0.in * (int64_t)f, d0.lg * (int64_t)f});
    }
    // kept dims: sort by |in stride| then |out stride|; merge when both chain
    struct KD {
        uint64_t n;
        int64_t in, out;
    };
    std::vector<KD> kd;
    for (size_t i = 0; i < lm.ndim(); ++i)
        if (lm.shape[i] != 1) kd.push_back({lm.shape[i], lm.stride[i], lo.stride[i]});
    std::stable_sort(kd.begin(), kd.end(), [](const KD& x, const KD& y) {
        if (std::llabs(x.in) != std::llabs(y.in)) return std::llabs(x.in) < std::llabs(y.in);
        return std::llabs(x.out) < std::llabs(y.out);
    });
    std::vector<KD> km;
    for (auto& d : kd) {
        if (!km.empty() && d.in == km.back().in * (int64_t)km.back().n && d.out == km.back().out * (int64_t)km.back().n) {
            km.back().n *= d.n;
            continue;
        }
        km.push_back(d);
    }
    if ((int)rm.size() > MAXD || (int)km.size() > MAXD)
        RSTSR_RAISE(UnImplemented, "reductions over more than 8 non-mergeable axes are not supported by the HIP device");
    RedPlan& p = s.p;
    p.nr = (int)rm.size();
    p.nk = (int)km.size();
    for (int i = 0; i < MAXD; ++i) {
        p.kshape[i] = 1;
        p.kfd[i] = FastDiv(1);
        p.kin[i] = p.kout[i] = 0;
        p.rshape[i] = 1;
        p.rin[i] = p.rlog[i] = 0;
    }
    uint64_t M = 1, R = 1;
    for (int i = 0; i < p.nk; ++i) {
        if (km[i].n >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "kept extent too large");
        p.kshape[i] = (uint32_t)km[i].n;
        p.kfd[i] = FastDiv((uint32_t)km[i].n);
        p.kin[i] = km[i].in;
        p.kout[i] = km[i].out;
        M *= km[i].n;
    }
    for (int i = 0; i < p.nr; ++i) {
        p.rshape[i] = rm[i].n;
        p.rin[i] = rm[i].in;
        p.rlog[i] = rm[i].lg;
        R *= rm[i].n;
    }
    if (M >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "too many outputs for one reduction launch");
    p.M = (uint32_t)M;
    p.R = R;
    p.in_off = (int64_t)la.offset;
    p.out_off = (int64_t)lo.offset;
    // a zero stride on an axis longer than one means repeated reads: keep those cached
    p.stream = 1;
    for (int i = 0; i < p.nk; ++i)
        if (p.kshape[i] > 1 && p.kin[i] == 0) p.stream = 0;
    for (int i = 0; i < p.nr; ++i)
        if (p.rshape[i] > 1 && p.rin[i] == 0) p.stream = 0;

    const uint64_t target_waves = (uint64_t)std::max(ctx->num_cu, 1) * 32;
    bool row_unit = std::llabs(p.rin[0]) == 1 && p.rshape[0] >= 16;
    bool col_unit = p.nk >= 1 && std::llabs(p.kin[0]) <= 1 && M >= 32;
    if (row_unit || !(col_unit || M >= 4096)) {
        s.kind = RedKind::Row;
        if (p.rshape[0] >= (1ull << 31)) RSTSR_RAISE(UnImplemented, "reduced extent too large");
        uint32_t L0 = (uint32_t)p.rshape[0];
        p.seg0 = std::min<uint32_t>(L0, 8192);
        p.nseg0 = (L0 + p.seg0 - 1) / p.seg0;
        p.units = (R / L0) * p.nseg0;
        uint64_t want = std::max<uint64_t>(1, target_waves / M);
        uint64_t nsplit = std::min<uint64_t>(want, p.units);
        nsplit = std::min<uint64_t>(nsplit, 65536);
        p.upw = (p.units + nsplit - 1) / nsplit;
        nsplit = (p.units + p.upw - 1) / p.upw;
        p.nsplit = (uint32_t)nsplit;
    } else {
        s.kind = RedKind::Col;
        // argmin/argmax track positions along reduced axis 0 in 32 bits
        if (p.rshape[0] >= 0xffffffffull) RSTSR_RAISE(UnImplemented, "reduced extent too large");
        // 4 or 8 waves per CU by the unroll depth (measured: more splits lose DRAM
        // locality along the reduced walk, fewer lose latency hiding); a thread covers
        // vec outputs
        const uint64_t vec = (uint64_t)col_vec(t);
        const uint64_t wpc = (uint64_t)col_waves_per_cu((int)(esize * vec));
        uint64_t want = std::max<uint64_t>(1, ((uint64_t)std::max(ctx->num_cu, 1) * wpc * 64 * vec) / M);
        uint64_t nsplit = std::min<uint64_t>(want, std::max<uint64_t>(1, R / 64));
        nsplit = std::min<uint64_t>(nsplit, 65536);
        p.upw = (R + nsplit - 1) / nsplit;
        nsplit = (R + p.upw - 1) / p.upw;
        p.nsplit = (uint32_t)nsplit;
    }
    if ((uint64_t)p.M * p.nsplit >= (1ull << 32)) RSTSR_RAISE(UnImplemented, "reduction grid too large");
    p.fsplit = FastDiv(p.nsplit);
    return s;
}

// Where a split reduction keeps its partials: the context workspace (grown on demand), or a
// fixed caller-supplied device region of `cap` bytes (no allocation; a plan that needs more
// is refused before anything is enqueued).
struct PartBuf {
    void* ptr = nullptr;
    size_t cap = 0;
};

template <class A, class T>
static void run_reduce(rstsr_hip_ctx* ctx, const RedSetup& s, const T* a, typename A::Out* out, PartBuf fixed) {
    const RedPlan& p = s.p;
    using S = typename A::S;
    S* part = nullptr;
    if (p.nsplit > 1) {
        const size_t need = (size_t)p.M * p.nsplit * sizeof(S);
        if (fixed.ptr) {
            if (need > fixed.cap)
                RSTSR_RAISE(ValueOutOfRange, "reduction partials need " + std::to_string(need) +
                                                 " bytes, more than the fixed scratch of " +
                                                 std::to_string(fixed.cap) + " bytes");
            part = (S*)fixed.ptr;
        } else {
            part = (S*)workspace(ctx, need);
        }
    }
    // 16-byte loads when every address a thread touches is 16-byte aligned
    constexpr int V16 = (std::is_arithmetic<T>::value && (sizeof(T) == 8 || sizeof(T) == 4)) ? (int)(16 / sizeof(T)) : 1;
    if (s.kind == RedKind::Row) {
        uint64_t nwaves = (uint64_t)p.M * p.nsplit;
        uint32_t nb = ceil_div_u32(nwaves, RB / 64);
        // rows start 16-byte aligned and run in whole 16-byte words
        bool vec = V16 > 1 && p.rin[0] == 1 && p.rshape[0] % V16 == 0 && p.seg0 % V16 == 0 &&
                   (p.in_off % V16) == 0 && ((uintptr_t)a % 16) == 0;
        for (int d = 0; d < p.nk && vec; ++d) vec = (p.kin[d] % V16) == 0;
        for (int d = 1; d < p.nr && vec; ++d) vec = (p.rin[d] % V16) == 0;
        if (vec)
            hipLaunchKernelGGL((k_red_row<A, T, V16>), dim3(nb), dim3(RB), 0, ctx->stream, a, out, part, p);
        else
            hipLaunchKernelGGL((k_red_row<A, T, 1>), dim3(nb), dim3(RB), 0, ctx->stream, a, out, part, p);
    } else {
        // column kernel: two outputs per thread for 4-byte elements too (8-byte loads), so a
        // split covers the same number of lanes as for 8-byte elements
        constexpr int VC = V16 > 2 ? 2 : V16;
        bool vec = VC > 1 && p.nk >= 1 && p.kin[0] == 1 && p.kshape[0] % VC == 0 && p.M % VC == 0 &&
                   (p.in_off % VC) == 0 && ((uintptr_t)a % (VC * sizeof(T))) == 0;
        for (int d = 1; d < p.nk && vec; ++d) vec = (p.kin[d] % VC) == 0;
        for (int d = 0; d < p.nr && vec; ++d) vec = (p.rin[d] % VC) == 0;
        if (vec) {
            uint64_t nthr = (uint64_t)(p.M / VC) * p.nsplit;
            hipLaunchKernelGGL((k_red_col<A, T, VC>), dim3(ceil_div_u32(nthr, RB)), dim3(RB), 0, ctx->stream, a, out,
                               part, p);
        } else {
            uint64_t nthr = (uint64_t)p.M * p.nsplit;
            hipLaunchKernelGGL((k_red_col<A, T, 1>), dim3(ceil_div_u32(nthr, RB)), dim3(RB), 0, ctx->stream, a, out,
                               part, p);
        }
    }
    check_launch("reduce");
    if (p.nsplit > 1) {
        hipLaunchKernelGGL((k_red_fin<A, T>), dim3(ceil_div_u32(p.M, RB / 64)), dim3(RB), 0, ctx->stream, a, out, part,
                           p);
        check_launch("reduce finalize");
    }
}

template <class T>
__global__ void k_fill_value(T* out, RedPlan p, T v) {
    uint32_t o = blockIdx.x * RB + threadIdx.x;
    if (o >= p.M) return;
    int64_t oin, oout;
    kept_offsets(p, o, oin, oout);
    out[oout] = v;
}

// Dispatch (op, dtype) to an accumulator and run it.
static void reduce_dispatch(rstsr_hip_ctx* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a, const RedSetup& s,
                            void* out, PartBuf fixed = PartBuf{}) {
    auto go = [&](auto acc_tag, auto val_tag) {
        using A = decltype(acc_tag);
        using T = decltype(val_tag);
        run_reduce<A, T>(ctx, s, (const T*)a, (typename A::Out*)out, fixed);
    };
    switch (op) {
        case RSTSR_RED_ALL:
        case RSTSR_RED_ANY:
            if (t == RSTSR_BOOL) {
                if (op == RSTSR_RED_ALL) go(BoolFold<true>{}, uint8_t{});
                else go(BoolFold<false>{}, uint8_t{});
                return;
            }
            break;
        case RSTSR_RED_COUNT_NONZERO:
            switch (t) {
                case RSTSR_F32: go(CountNZ<float>{}, float{}); return;
                case RSTSR_F64: go(CountNZ<double>{}, double{}); return;
                case RSTSR_I32: go(CountNZ<int32_t>{}, int32_t{}); return;
                case RSTSR_I64: go(CountNZ<int64_t>{}, int64_t{}); return;
                case RSTSR_C64: go(CountNZ<c64>{}, c64{}); return;
                case RSTSR_C128: go(CountNZ<c128>{}, c128{}); return;
                case RSTSR_BOOL: go(CountNZ<uint8_t>{}, uint8_t{}); return;
                default: break;
            }
            break;
        case RSTSR_RED_SUM:
            switch (t) {
                case RSTSR_BOOL: go(CountNZ<uint8_t>{}, uint8_t{}); return;
                case RSTSR_F32: go(SumF<float>{}, float{}); return;
                case RSTSR_F64: go(SumF<double>{}, double{}); return;
                case RSTSR_I32: go(SumI<int32_t>{}, int32_t{}); return;
                case RSTSR_I64: go(SumI<int64_t>{}, int64_t{}); return;
                case RSTSR_C64: go(SumC<c64, float>{}, c64{}); return;
                case RSTSR_C128: go(SumC<c128, double>{}, c128{}); return;
                default: break;
            }
            break;
        case RSTSR_RED_MAX:
        case RSTSR_RED_MIN: {
            bool mx = op == RSTSR_RED_MAX;
            switch (t) {
                case RSTSR_F32: mx ? go(Ext<float, true>{}, float{}) : go(Ext<float, false>{}, float{}); return;
                case RSTSR_F64: mx ? go(Ext<double, true>{}, double{}) : go(Ext<double, false>{}, double{}); return;
                case RSTSR_I32: mx ? go(Ext<int32_t, true>{}, int32_t{}) : go(Ext<int32_t, false>{}, int32_t{}); return;
                case RSTSR_I64: mx ? go(Ext<int64_t, true>{}, int64_t{}) : go(Ext<int64_t, false>{}, int64_t{}); return;
                default: break;
            }
            break;
        }
        case RSTSR_RED_PROD:
            switch (t) {
                case RSTSR_F32: go(ProdF<float>{}, float{}); return;
                case RSTSR_F64: go(ProdF<double>{}, double{}); return;
                case RSTSR_I32: go(ProdI<int32_t>{}, int32_t{}); return;
                case RSTSR_I64: go(ProdI<int64_t>{}, int64_t{}); return;
                case RSTSR_C64: go(ProdC<c64, float>{}, c64{}); return;
                case RSTSR_C128: go(ProdC<c128, double>{}, c128{}); return;
                default: break;
            }
            break;
        case RSTSR_RED_MEAN:
            switch (t) {
                case RSTSR_F32: go(MeanF<float>{}, float{}); return;
                case RSTSR_F64: go(MeanF<double>{}, double{}); return;
                case RSTSR_C64: go(MeanC<c64, float>{}, c64{}); return;
                case RSTSR_C128: go(MeanC<c128, double>{}, c128{}); return;
                default: break;
            }
            break;
        case RSTSR_RED_VAR:
        case RSTSR_RED_STD:
        case RSTSR_RED_L2NORM: {
            auto mom = [&](auto kind) {
                constexpr int K = decltype(kind)::value;
                switch (t) {
                    case RSTSR_F32: go(Moments<float, float, K>{}, float{}); return true;
                    case RSTSR_F64: go(Moments<double, double, K>{}, double{}); return true;
                    case RSTSR_C64: go(Moments<c64, float, K>{}, c64{}); return true;
                    case RSTSR_C128: go(Moments<c128, double, K>{}, c128{}); return true;
                    default: return false;
                }
            };
            bool ok = op == RSTSR_RED_VAR   ? mom(std::integral_constant<int, 0>{})
                      : op == RSTSR_RED_STD ? mom(std::integral_constant<int, 1>{})
                                            : mom(std::integral_constant<int, 2>{});
            if (ok) return;
            break;
        }
        case RSTSR_RED_ARGMAX:
        case RSTSR_RED_ARGMIN: {
            bool mx = op == RSTSR_RED_ARGMAX;
            switch (t) {
                case RSTSR_F32: mx ? go(Arg<float, true>{}, float{}) : go(Arg<float, false>{}, float{}); return;
                case RSTSR_F64: mx ? go(Arg<double, true>{}, double{}) : go(Arg<double, false>{}, double{}); return;
                case RSTSR_I32: mx ? go(Arg<int32_t, true>{}, int32_t{}) : go(Arg<int32_t, false>{}, int32_t{}); return;
                case RSTSR_I64: mx ? go(Arg<int64_t, true>{}, int64_t{}) : go(Arg<int64_t, false>{}, int64_t{}); return;
                default: break;
            }
            break;
        }
    }
    RSTSR_RAISE(UnImplemented, "reduction not provided for this dtype");
}

static size_t out_size_of(rstsr_redop_t op, rstsr_dtype_t t) {
    if (op == RSTSR_RED_ARGMAX || op == RSTSR_RED_ARGMIN) return 8;
    // all / any: bool; count_nonzero and the sum of a bool array: usize
    if (op == RSTSR_RED_ALL || op == RSTSR_RED_ANY) return 1;
    if (op == RSTSR_RED_COUNT_NONZERO || (op == RSTSR_RED_SUM && t == RSTSR_BOOL)) return 8;
    // var/std/l2_norm of a complex array are real (T::Real)
    if ((op == RSTSR_RED_VAR || op == RSTSR_RED_STD || op == RSTSR_RED_L2NORM) && (t == RSTSR_C64 || t == RSTSR_C128))
        return dtype_size(t) / 2;
    return dtype_size(t);
}

static void check_op(rstsr_redop_t op, bool arg) {
    if ((int)op < 0 || (int)op > 12) RSTSR_RAISE(InvalidValue, "unknown reduction op");
    bool is_arg = op == RSTSR_RED_ARGMAX || op == RSTSR_RED_ARGMIN;
    if (is_arg != arg) RSTSR_RAISE(InvalidValue, arg ? "expected argmax/argmin" : "argmax/argmin need the arg entry point");
}

// Common path of *_axes entry points; la, lo already validated.
// Value of a reduction over zero elements: the accumulator's f_out(f_init())
// in the reference (sum 0, prod 1, l2_norm 0, mean/var/std 0/0 = NaN); max/min
// and argmax/argmin over nothing are errors there (auto_impl/reduction.rs:95-97).
static bool empty_value_op(rstsr_redop_t op) {
    return op == RSTSR_RED_SUM || op == RSTSR_RED_PROD || op == RSTSR_RED_MEAN || op == RSTSR_RED_VAR ||
           op == RSTSR_RED_STD || op == RSTSR_RED_L2NORM || op == RSTSR_RED_ALL || op == RSTSR_RED_ANY ||
           op == RSTSR_RED_COUNT_NONZERO;
}

static void empty_value(rstsr_redop_t op, rstsr_dtype_t t, void* v) {
    const size_t es = out_size_of(op, t);
    std::memset(v, 0, 16);
    // all of nothing is true; any, count_nonzero and a bool sum of nothing are zero
    if (op == RSTSR_RED_ALL) {
        *(uint8_t*)v = 1;
        return;
    }
    if (op == RSTSR_RED_ANY || op == RSTSR_RED_COUNT_NONZERO || (op == RSTSR_RED_SUM && t == RSTSR_BOOL)) return;
    const bool nan = op == RSTSR_RED_MEAN || op == RSTSR_RED_VAR || op == RSTSR_RED_STD;
    auto put = [&](double re, double im) {
        if (t == RSTSR_F32 || ((t == RSTSR_C64) && es == 4)) *(float*)v = (float)re;
        else if (t == RSTSR_F64 || (t == RSTSR_C128 && es == 8)) *(double*)v = re;
        else if (t == RSTSR_C64) {
            ((float*)v)[0] = (float)re;
            ((float*)v)[1] = (float)im;
        } else if (t == RSTSR_C128) {
            ((double*)v)[0] = re;
            ((double*)v)[1] = im;
        } else if (t == RSTSR_I32) *(int32_t*)v = (int32_t)re;
        else if (t == RSTSR_I64) *(int64_t*)v = (int64_t)re;
    };
    if (op == RSTSR_RED_PROD) put(1.0, 0.0);
    else if (nan) {
        if (t == RSTSR_I32 || t == RSTSR_I64) RSTSR_RAISE(UnImplemented, "mean/var/std of integers");
        put(std::nan(""), (t == RSTSR_C64 || t == RSTSR_C128) && op == RSTSR_RED_MEAN ? std::nan("") : 0.0);
    }
}

static void reduce_axes_impl(rstsr_hip_ctx* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a, const Layout& la,
                             const std::vector<ptrdiff_t>& axes, void* out, const Layout& lo,
                             PartBuf fixed = PartBuf{}) {
    if (!empty_value_op(op) && la.size() == 0) {
        if (op == RSTSR_RED_MAX || op == RSTSR_RED_MIN)
            RSTSR_RAISE(InvalidValue, "zero-size array is not supported for max/min");
        RSTSR_RAISE(InvalidLayout, "empty sequence is not allowed for reduce_arg.");
    }
    // Launches address at most 2^31 - 1 outputs: 2^31 or more are reduced in pieces of the
    // largest kept axis (2^30 positions each), input and output views narrowed alike.
    if (lo.size() >= (1ull << 31)) {
        auto ax = normalize_axes_index(axes, la.ndim(), false);
        std::vector<size_t> kept;
        for (size_t i = 0; i < la.ndim(); ++i)
            if (std::find(ax.begin(), ax.end(), (ptrdiff_t)i) == ax.end()) kept.push_back(i);
        size_t j = 0;  // position in lo (the kept axes in order) of the largest kept axis
        for (size_t q = 1; q < kept.size(); ++q)
            if (la.shape[kept[q]] > la.shape[kept[j]]) j = q;
        const size_t ia = kept[j];
        const uint64_t n = la.shape[ia], step = 1ull << 30;
        for (uint64_t s0 = 0; s0 < n; s0 += step) {
            const uint64_t cnt = std::min(step, n - s0);
            Layout la2 = la, lo2 = lo;
            la2.shape[ia] = cnt;
            lo2.shape[j] = cnt;
            la2.offset = (size_t)((int64_t)la.offset + (int64_t)s0 * la.stride[ia]);
            lo2.offset = (size_t)((int64_t)lo.offset + (int64_t)s0 * lo.stride[j]);
            reduce_axes_impl(ctx, op, t, a, la2, axes, out, lo2, fixed);
        }
        return;
    }
    RedSetup s = reduce_plan(ctx, la, axes, lo, t);
    if (s.empty_out) return;
    if (s.empty_reduce) {
        // outputs exist but each reduces zero elements: every output is the empty value
        if (!empty_value_op(op)) RSTSR_RAISE(InvalidValue, "zero-size reduction");
        alignas(16) unsigned char ev[16];
        empty_value(op, t, ev);
        // build a plan over the kept axes only to address outputs
        Layout lm = la.dim_split_axes(axes).second;
        std::vector<ptrdiff_t> none;
        Layout la_k(lm.shape, lm.stride, la.offset);
        // outputs via a trivial reduced space of extent 1 at a harmless address
        RedPlan p;
        p.nk = 0;
        std::vector<std::pair<uint64_t, int64_t>> kd;
        for (size_t i = 0; i < lo.ndim(); ++i)
            if (lo.shape[i] != 1) kd.push_back({lo.shape[i], lo.stride[i]});
        if ((int)kd.size() > MAXD) RSTSR_RAISE(UnImplemented, "too many output axes");
        p.nk = (int)kd.size();
        uint64_t M = 1;
        for (int i = 0; i < MAXD; ++i) {
            p.kshape[i] = 1;
            p.kfd[i] = FastDiv(1);
            p.kin[i] = p.kout[i] = 0;
        }
        for (int i = 0; i < p.nk; ++i) {
            p.kshape[i] = (uint32_t)kd[i].first;
            p.kfd[i] = FastDiv((uint32_t)kd[i].first);
            p.kout[i] = kd[i].second;
            M *= kd[i].first;
        }
        p.M = (uint32_t)M;
        p.out_off = (int64_t)lo.offset;
        auto z = [&](auto tag) {
            using T = decltype(tag);
            T v;
            std::memcpy(&v, ev, sizeof(T));
            hipLaunchKernelGGL((k_fill_value<T>), dim3(ceil_div_u32(M, RB)), dim3(RB), 0, ctx->stream, (T*)out, p, v);
            check_launch("fill");
        };
        // the output element, stored by its width (the value's bytes are already in ev)
        switch (out_size_of(op, t)) {
            case 1: z(uint8_t{}); break;
            case 4: z(uint32_t{}); break;
            case 8: z(uint64_t{}); break;
            case 16: z(c128{}); break;
            default: RSTSR_RAISE(UnImplemented, "dtype");
        }
        return;
    }
    reduce_dispatch(ctx, op, t, a, s, out, fixed);
}

/* ======================= fixed scratch (no-allocation entry point) ======================= */

// Partials of split reductions for rstsr_hip_reduce_axes_noalloc. Each GPU the library runs
// on holds its own copy until the process ends. Its host-side name is not a device address:
// host code uses only the per-device address hipGetSymbolAddress returns (scratch_addr).
// On 256 CUs the largest plan needs 131072 partials of at most 48 bytes (6 MiB).
constexpr size_t kNoallocScratchBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, void*>& scratch_table() {  // device -> address of its scratch; g_scratch_mu
    static std::map<int, void*> t;
    return t;
}

// Resolve this device's scratch address once per device, with that device current.
void reduce_scratch_setup(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (scratch_table().count(device)) return;
    HIP_CHECK(hipSetDevice(device));
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        RSTSR_RAISE(DeviceError, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    scratch_table()[device] = ptr;
}

static void* scratch_addr(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = scratch_table().find(device);
    if (it == scratch_table().end()) RSTSR_RAISE(DeviceError, "reduction scratch was not set up for this device");
    return it->second;
}

}  // namespace hip
}  // namespace rstsr

using namespace rstsr;
using namespace rstsr::hip;

static std::vector<ptrdiff_t> all_axes(size_t nd) {
    std::vector<ptrdiff_t> v(nd);
    for (size_t i = 0; i < nd; ++i) v[i] = (ptrdiff_t)i;
    return v;
}

extern "C" {

int rstsr_hip_reduce_axes(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                          const rstsr_layout_t* la_, size_t naxes, const ptrdiff_t* axes, void* out,
                          const rstsr_layout_t* lo_) {
    return guarded([&] {
        bind(ctx);
        check_op(op, false);
        Layout la = from_abi(la_), lo = from_abi(lo_);
        if (naxes && !axes) RSTSR_RAISE(InvalidValue, "null axes");
        std::vector<ptrdiff_t> ax(axes, axes + naxes);
        check_extent(a, la, dtype_size(t), "a");
        check_extent(out, lo, out_size_of(op, t), "out");
        std::lock_guard<std::mutex> g(ctx->mu);
        reduce_axes_impl(ctx, op, t, a, la, ax, out, lo);
    });
}

int rstsr_hip_reduce_axes_noalloc(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                                  const rstsr_layout_t* la_, size_t naxes, const ptrdiff_t* axes, void* out,
                                  const rstsr_layout_t* lo_) {
    return guarded([&] {
        bind(ctx);
        check_op(op, false);
        Layout la = from_abi(la_), lo = from_abi(lo_);
        if (naxes && !axes) RSTSR_RAISE(InvalidValue, "null axes");
        std::vector<ptrdiff_t> ax(axes, axes + naxes);
        check_extent(a, la, dtype_size(t), "a");
        check_extent(out, lo, out_size_of(op, t), "out");
        PartBuf fixed{scratch_addr(ctx->device), kNoallocScratchBytes};
        std::lock_guard<std::mutex> g(ctx->mu);
        reduce_axes_impl(ctx, op, t, a, la, ax, out, lo, fixed);
    });
}

int rstsr_hip_argreduce_axes(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                             const rstsr_layout_t* la_, size_t naxes, const ptrdiff_t* axes, uint64_t* out,
                             const rstsr_layout_t* lo_) {
    return guarded([&] {
        bind(ctx);
        check_op(op, true);
        Layout la = from_abi(la_), lo = from_abi(lo_);
        if (naxes && !axes) RSTSR_RAISE(InvalidValue, "null axes");
        std::vector<ptrdiff_t> ax(axes, axes + naxes);
        check_extent(a, la, dtype_size(t), "a");
        check_extent(out, lo, 8, "out");
        std::lock_guard<std::mutex> g(ctx->mu);
        reduce_axes_impl(ctx, op, t, a, la, ax, out, lo);
    });
}

static void reduce_all_common(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                              const rstsr_layout_t* la_, void* host_out, size_t out_bytes) {
    bind(ctx);
    if (!host_out) RSTSR_RAISE(InvalidValue, "null output");
    if (stream_capturing(ctx))
        RSTSR_RAISE(UnImplemented, "a reduction over all elements hands its value to the host and cannot be captured");
    Layout la = from_abi(la_);
    check_extent(a, la, dtype_size(t), "a");
    std::lock_guard<std::mutex> g(ctx->mu);
    // the scalar result lives in the last 256 bytes of the workspace; partials
    // come first, so reserve room for both in one allocation.
    RedSetup s;
    bool trivial_sum = false;
    if (la.size() == 0) {
        if (op == RSTSR_RED_MAX || op == RSTSR_RED_MIN)
            RSTSR_RAISE(InvalidValue, "zero-size array is not supported for max/min");
        if (!empty_value_op(op)) RSTSR_RAISE(InvalidLayout, "empty sequence is not allowed for reduce_arg.");
        trivial_sum = true;
    }
    if (trivial_sum) {
        alignas(16) unsigned char ev[16];
        empty_value(op, t, ev);
        std::memcpy(host_out, ev, out_bytes);
        return;
    }
    Layout lo({}, {}, 0);
    s = reduce_plan(ctx, la, all_axes(la.ndim()), lo, t);
    size_t part_bytes = 0;
    if (s.p.nsplit > 1) part_bytes = (size_t)s.p.M * s.p.nsplit * 64;  // upper bound of sizeof(S)
    char* ws = (char*)workspace(ctx, part_bytes + 256);
    void* dout = ws + part_bytes;
    // run_reduce asks for the same workspace size again, which is a no-op
    reduce_dispatch(ctx, op, t, a, s, dout);
    HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, dout, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::memcpy(host_out, ctx->host_pinned, out_bytes);
}

int rstsr_hip_reduce_all(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                         const rstsr_layout_t* la, void* host_out) {
    return guarded([&] {
        check_op(op, false);
        reduce_all_common(ctx, op, t, a, la, host_out, out_size_of(op, t));
    });
}

int rstsr_hip_argreduce_all(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                            const rstsr_layout_t* la, uint64_t* host_out) {
    return guarded([&] {
        check_op(op, true);
        reduce_all_common(ctx, op, t, a, la, host_out, 8);
    });
}

}  // extern "C"
