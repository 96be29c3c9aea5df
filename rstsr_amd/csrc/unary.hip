// Unary elementwise operations for gfx950: the 35 OpXxxAPI traits of
// rstsr-core/src/operators/ops/op_binary_common.rs:3-39 plus negation and
// bitwise/logical not (operators/ops/op_binary_arithmetic.rs:90-109), over the
// strided planner and kernels of ew.hpp.
//
// Element semantics follow the reference's functors
// (device_cpu_serial/operators/op_binary_common.rs:10-237 and
// feature_rayon/auto_impl/op_binary_common.rs, identical bodies):
//  - T::FloatType ops convert the input with DTypeIntoFloatAPI (i32/i64 -> f64)
//    and apply the num::Float / num::complex::ComplexFloat method; for real
//    floats these are the C math library's functions, here the device's;
//  - complex functions are num-complex 0.4's formulas (Complex::exp, ln, sqrt,
//    sin, cos, tan, sinh, cosh, tanh, asin, acos, atan, asinh, acosh, atanh,
//    finv), restated below with num-complex's own complex multiply/divide;
//  - sign, abs, real, imag: rstsr-dtype-traits/src/ext_num.rs (ExtNum);
//  - signbit is Signed::is_positive: x > 0 for integers, the sign bit clear for
//    floats (so +0.0 and +NaN give true);
//  - integer negation, square and abs wrap (Rust release arithmetic).
// Built with -ffp-contract=off: every product and sum rounds on its own, as
// in Rust.
//
// Integer inputs to T::FloatType ops run in two passes: a cast into the f64
// output, then the f64 operation in place on it.
#include "ew.hpp"

#include <type_traits>

namespace rstsr {
namespace hip {

/* =============================== real math =============================== */

#define RSTSR_RM1(NAME, FNAME)                                                   \
    __device__ __forceinline__ double m_##NAME(double x) { return ::NAME(x); } \
    __device__ __forceinline__ float m_##NAME(float x) { return ::FNAME(x); }
RSTSR_RM1(acos, acosf)
RSTSR_RM1(acosh, acoshf)
RSTSR_RM1(asin, asinf)
RSTSR_RM1(asinh, asinhf)
RSTSR_RM1(atan, atanf)
RSTSR_RM1(atanh, atanhf)
RSTSR_RM1(ceil, ceilf)
RSTSR_RM1(cos, cosf)
RSTSR_RM1(cosh, coshf)
RSTSR_RM1(exp, expf)
RSTSR_RM1(expm1, expm1f)
RSTSR_RM1(floor, floorf)
RSTSR_RM1(log, logf)
RSTSR_RM1(log1p, log1pf)
RSTSR_RM1(log2, log2f)
RSTSR_RM1(log10, log10f)
RSTSR_RM1(round, roundf)
RSTSR_RM1(sin, sinf)
RSTSR_RM1(sinh, sinhf)
RSTSR_RM1(sqrt, sqrtf)
RSTSR_RM1(tan, tanf)
RSTSR_RM1(tanh, tanhf)
RSTSR_RM1(trunc, truncf)
RSTSR_RM1(fabs, fabsf)
#undef RSTSR_RM1
__device__ __forceinline__ double m_hypot(double x, double y) { return ::hypot(x, y); }
__device__ __forceinline__ float m_hypot(float x, float y) { return ::hypotf(x, y); }
__device__ __forceinline__ double m_atan2(double y, double x) { return ::atan2(y, x); }
__device__ __forceinline__ float m_atan2(float y, float x) { return ::atan2f(y, x); }

// num::Float methods on a real float (the T::FloatType ops that have a real form)
template <int OP, class R>
__device__ __forceinline__ R real_fn(R x) {
    if constexpr (OP == RSTSR_UN_ACOS) return m_acos(x);
    else if constexpr (OP == RSTSR_UN_ACOSH) return m_acosh(x);
    else if constexpr (OP == RSTSR_UN_ASIN) return m_asin(x);
    else if constexpr (OP == RSTSR_UN_ASINH) return m_asinh(x);
    else if constexpr (OP == RSTSR_UN_ATAN) return m_atan(x);
    else if constexpr (OP == RSTSR_UN_ATANH) return m_atanh(x);
    else if constexpr (OP == RSTSR_UN_CEIL) return m_ceil(x);
    else if constexpr (OP == RSTSR_UN_CONJ) return x;
    else if constexpr (OP == RSTSR_UN_COS) return m_cos(x);
    else if constexpr (OP == RSTSR_UN_COSH) return m_cosh(x);
    else if constexpr (OP == RSTSR_UN_EXP) return m_exp(x);
    else if constexpr (OP == RSTSR_UN_EXPM1) return m_expm1(x);
    else if constexpr (OP == RSTSR_UN_FLOOR) return m_floor(x);
    else if constexpr (OP == RSTSR_UN_INV || OP == RSTSR_UN_RECIPROCAL) return R(1) / x;
    else if constexpr (OP == RSTSR_UN_LOG) return m_log(x);
    else if constexpr (OP == RSTSR_UN_LOG1P) return m_log1p(x);
    else if constexpr (OP == RSTSR_UN_LOG2) return m_log2(x);
    else if constexpr (OP == RSTSR_UN_LOG10) return m_log10(x);
    else if constexpr (OP == RSTSR_UN_ROUND) return m_round(x);
    else if constexpr (OP == RSTSR_UN_SIN) return m_sin(x);
    else if constexpr (OP == RSTSR_UN_SINH) return m_sinh(x);
    else if constexpr (OP == RSTSR_UN_SQRT) return m_sqrt(x);
    else if constexpr (OP == RSTSR_UN_TAN) return m_tan(x);
    else if constexpr (OP == RSTSR_UN_TANH) return m_tanh(x);
    else if constexpr (OP == RSTSR_UN_TRUNC) return m_trunc(x);
    else static_assert(OP < 0, "not a real float op");
}

/* =============================== complex (num-complex 0.4) =============================== */

template <class R>
struct CplxOf;
template <>
struct CplxOf<float> {
    using type = c64;
};
template <>
struct CplxOf<double> {
    using type = c128;
};

template <class R>
struct Cx {
    using C = typename CplxOf<R>::type;
    static __device__ __forceinline__ C mk(R re, R im) { return C{re, im}; }
    // Mul: (a.re b.re - a.im b.im, a.re b.im + a.im b.re)
    static __device__ __forceinline__ C mul(C a, C b) { return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
    // Div: ((a.re c + a.im d) / n, (a.im c - a.re d) / n), n = c*c + d*d
    static __device__ __forceinline__ C div(C a, C b) {
        R n = b.re * b.re + b.im * b.im;
        R re = a.re * b.re + a.im * b.im;
        R im = a.im * b.re - a.re * b.im;
        return C{re / n, im / n};
    }
    static __device__ __forceinline__ C add(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
    static __device__ __forceinline__ C sub(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
    static __device__ __forceinline__ C neg(C a) { return C{-a.re, -a.im}; }
    static __device__ __forceinline__ C unscale(C a, R t) { return C{a.re / t, a.im / t}; }
    static __device__ __forceinline__ R norm(C a) { return m_hypot(a.re, a.im); }
    static __device__ __forceinline__ R arg(C a) { return m_atan2(a.im, a.re); }
    static __device__ __forceinline__ C from_polar(R r, R t) { return C{r * m_cos(t), r * m_sin(t)}; }
    static __device__ __forceinline__ bool eq(C a, C b) { return a.re == b.re && a.im == b.im; }

    static __device__ __forceinline__ C exp(C z) {
        R re = z.re, im = z.im;
        if (isinf(re)) {
            if (re < R(0)) {
                if (!isfinite(im)) return C{R(0), R(0)};
            } else if (im == R(0) || !isfinite(im)) {
                if (isinf(im)) im = R(NAN);
                return C{re, im};
            }
        } else if (isnan(re) && im == R(0)) {
            return z;
        }
        return from_polar(m_exp(re), im);
    }
    static __device__ __forceinline__ C ln(C z) { return C{m_log(norm(z)), arg(z)}; }
    static __device__ __forceinline__ C sqrt(C z) {
        if (z.im == R(0)) {
            if (!signbit(z.re)) return C{m_sqrt(z.re), z.im};
            R im = m_sqrt(-z.re);
            return signbit(z.im) ? C{R(0), -im} : C{R(0), im};
        }
        if (z.re == R(0)) {
            R x = m_sqrt(m_fabs(z.im) / R(2));
            return signbit(z.im) ? C{x, -x} : C{x, x};
        }
        return from_polar(m_sqrt(norm(z)), arg(z) / R(2));
    }
    static __device__ __forceinline__ C sin(C z) { return C{m_sin(z.re) * m_cosh(z.im), m_cos(z.re) * m_sinh(z.im)}; }
    static __device__ __forceinline__ C cos(C z) { return C{m_cos(z.re) * m_cosh(z.im), -m_sin(z.re) * m_sinh(z.im)}; }
    static __device__ __forceinline__ C tan(C z) {
        R a = z.re + z.re, b = z.im + z.im;
        return unscale(C{m_sin(a), m_sinh(b)}, m_cos(a) + m_cosh(b));
    }
    static __device__ __forceinline__ C sinh(C z) { return C{m_sinh(z.re) * m_cos(z.im), m_cosh(z.re) * m_sin(z.im)}; }
    static __device__ __forceinline__ C cosh(C z) { return C{m_cosh(z.re) * m_cos(z.im), m_sinh(z.re) * m_sin(z.im)}; }
    static __device__ __forceinline__ C tanh(C z) {
        R a = z.re + z.re, b = z.im + z.im;
        return unscale(C{m_sinh(a), m_sin(b)}, m_cosh(a) + m_cos(b));
    }
    // arcsin(z) = -i ln(sqrt(1 - z^2) + i z)
    static __device__ __forceinline__ C asin(C z) {
        const C i{R(0), R(1)}, one{R(1), R(0)};
        return mul(neg(i), ln(add(sqrt(sub(one, mul(z, z))), mul(i, z))));
    }
    // arccos(z) = -i ln(i sqrt(1 - z^2) + z)
    static __device__ __forceinline__ C acos(C z) {
        const C i{R(0), R(1)}, one{R(1), R(0)};
        return mul(neg(i), ln(add(mul(i, sqrt(sub(one, mul(z, z)))), z)));
    }
    // arctan(z) = (ln(1 + i z) - ln(1 - i z)) / (2i); z = +-i map to (0, +-inf)
    static __device__ __forceinline__ C atan(C z) {
        const C i{R(0), R(1)}, one{R(1), R(0)};
        if (eq(z, i)) return C{R(0), R(INFINITY)};
        if (eq(z, neg(i))) return C{R(0), -R(INFINITY)};
        const C two = add(one, one);
        return div(sub(ln(add(one, mul(i, z))), ln(sub(one, mul(i, z)))), mul(two, i));
    }
    // arcsinh(z) = ln(z + sqrt(1 + z^2))
    static __device__ __forceinline__ C asinh(C z) {
        const C one{R(1), R(0)};
        return ln(add(z, sqrt(add(one, mul(z, z)))));
    }
    // arccosh(z) = 2 ln(sqrt((z + 1)/2) + sqrt((z - 1)/2))
    static __device__ __forceinline__ C acosh(C z) {
        const C one{R(1), R(0)};
        const C two = add(one, one);
        return mul(two, ln(add(sqrt(div(add(z, one), two)), sqrt(div(sub(z, one), two)))));
    }
    // arctanh(z) = (ln(1 + z) - ln(1 - z)) / 2; z = +-1 map to (+-inf, 0)
    static __device__ __forceinline__ C atanh(C z) {
        const C one{R(1), R(0)};
        if (eq(z, one)) return C{R(INFINITY), R(0)};
        if (eq(z, neg(one))) return C{-R(INFINITY), R(0)};
        const C two = add(one, one);
        return div(sub(ln(add(one, z)), ln(sub(one, z))), two);
    }
    // ComplexFloat::recip = Complex::finv: conj(z) / |z| / |z|
    static __device__ __forceinline__ C recip(C z) {
        R n = norm(z);
        return unscale(unscale(C{z.re, -z.im}, n), n);
    }
};

template <int OP, class C>
__device__ __forceinline__ C cplx_fn(C z) {
    using R = decltype(z.re);
    using X = Cx<R>;
    if constexpr (OP == RSTSR_UN_ACOS) return X::acos(z);
    else if constexpr (OP == RSTSR_UN_ACOSH) return X::acosh(z);
    else if constexpr (OP == RSTSR_UN_ASIN) return X::asin(z);
    else if constexpr (OP == RSTSR_UN_ASINH) return X::asinh(z);
    else if constexpr (OP == RSTSR_UN_ATAN) return X::atan(z);
    else if constexpr (OP == RSTSR_UN_ATANH) return X::atanh(z);
    else if constexpr (OP == RSTSR_UN_CONJ) return C{z.re, -z.im};
    else if constexpr (OP == RSTSR_UN_COS) return X::cos(z);
    else if constexpr (OP == RSTSR_UN_COSH) return X::cosh(z);
    else if constexpr (OP == RSTSR_UN_EXP) return X::exp(z);
    else if constexpr (OP == RSTSR_UN_INV || OP == RSTSR_UN_RECIPROCAL) return X::recip(z);
    else if constexpr (OP == RSTSR_UN_LOG) return X::ln(z);
    // log2 / log10: ln(z) / LN_2, ln(z) / LN_10 (componentwise division by the real constant)
    else if constexpr (OP == RSTSR_UN_LOG2) return X::unscale(X::ln(z), R(0.693147180559945309417232121458176568));
    else if constexpr (OP == RSTSR_UN_LOG10) return X::unscale(X::ln(z), R(2.30258509299404568401799145468436421));
    else if constexpr (OP == RSTSR_UN_SIN) return X::sin(z);
    else if constexpr (OP == RSTSR_UN_SINH) return X::sinh(z);
    else if constexpr (OP == RSTSR_UN_SQRT) return X::sqrt(z);
    else if constexpr (OP == RSTSR_UN_TAN) return X::tan(z);
    else if constexpr (OP == RSTSR_UN_TANH) return X::tanh(z);
    else static_assert(OP < 0, "not a complex float op");
}

/* =============================== element functor =============================== */

template <class T>
constexpr bool is_cplx_v = std::is_same<T, c64>::value || std::is_same<T, c128>::value;

// op(x) for input type TA with output type TC (TC = uint8_t for bool outputs)
template <int OP, class TC, class TA>
__device__ __forceinline__ TC un_apply(TA x) {
    if constexpr (OP == RSTSR_UN_SQUARE) {
        if constexpr (is_cplx_v<TA>) return Cx<decltype(x.re)>::mul(x, x);
        else if constexpr (std::is_integral<TA>::value) {
            using U = typename std::make_unsigned<TA>::type;
            return (TA)((U)x * (U)x);
        } else return x * x;
    } else if constexpr (OP == RSTSR_UN_NEG) {
        if constexpr (is_cplx_v<TA>) return TA{-x.re, -x.im};
        else if constexpr (std::is_integral<TA>::value) {
            using U = typename std::make_unsigned<TA>::type;
            return (TA)(U(0) - (U)x);
        } else return -x;
    } else if constexpr (OP == RSTSR_UN_NOT) {
        if constexpr (std::is_same<TA, uint8_t>::value) return (uint8_t)(x ? 0 : 1);
        else return (TA)~x;
    } else if constexpr (OP == RSTSR_UN_ABS) {
        if constexpr (is_cplx_v<TA>) return Cx<decltype(x.re)>::norm(x);
        else if constexpr (std::is_integral<TA>::value) {
            using U = typename std::make_unsigned<TA>::type;
            return x < 0 ? (TA)(U(0) - (U)x) : x;
        } else return m_fabs(x);
    } else if constexpr (OP == RSTSR_UN_REAL) {
        if constexpr (is_cplx_v<TA>) return x.re;
        else return x;
    } else if constexpr (OP == RSTSR_UN_IMAG) {
        if constexpr (is_cplx_v<TA>) return x.im;
        else return TC(0);
    } else if constexpr (OP == RSTSR_UN_SIGN) {
        if constexpr (is_cplx_v<TA>) {
            // z / |z|, and 0 for |z| == 0
            using R = decltype(x.re);
            R a = Cx<R>::norm(x);
            if (a == R(0)) return TA{R(0), R(0)};
            return Cx<R>::unscale(x, a);
        } else if constexpr (std::is_integral<TA>::value) {
            return x > 0 ? TA(1) : (x < 0 ? TA(-1) : TA(0));
        } else {
            if (x != x) return x;
            return x > TA(0) ? TA(1) : (x < TA(0) ? TA(-1) : TA(0));
        }
    } else if constexpr (OP == RSTSR_UN_SIGNBIT) {
        // Signed::is_positive
        if constexpr (std::is_integral<TA>::value) return (uint8_t)(x > 0);
        else return (uint8_t)(!signbit(x));
    } else if constexpr (OP == RSTSR_UN_IS_FINITE) {
        if constexpr (is_cplx_v<TA>) return (uint8_t)(isfinite(x.re) && isfinite(x.im));
        else return (uint8_t)isfinite(x);
    } else if constexpr (OP == RSTSR_UN_IS_INF) {
        if constexpr (is_cplx_v<TA>) return (uint8_t)(!(isnan(x.re) || isnan(x.im)) && (isinf(x.re) || isinf(x.im)));
        else return (uint8_t)isinf(x);
    } else if constexpr (OP == RSTSR_UN_IS_NAN) {
        if constexpr (is_cplx_v<TA>) return (uint8_t)(isnan(x.re) || isnan(x.im));
        else return (uint8_t)isnan(x);
    } else if constexpr (is_cplx_v<TA>) {
        return cplx_fn<OP, TA>(x);
    } else {
        static_assert(std::is_same<TA, TC>::value, "T::FloatType ops run on float inputs");
        return real_fn<OP, TA>(x);
    }
}

template <int OP, class TCv, class TAv>
struct UnBody {
    using TA = TAv;
    using TB = None;
    TCv* c;
    const TAv* a;
    static constexpr int OPS = 1;
    static constexpr size_t OUT_SIZE = sizeof(TCv);
    static constexpr bool NT_JOINT = true;
    template <bool NT = false>
    __device__ __forceinline__ TAv la(int64_t o) const { return ld<NT>(a + o); }
    template <bool NT = false>
    __device__ __forceinline__ None lb(int64_t) const { return None{}; }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t o, TAv x, None) const { sto<NT>(c + o, un_apply<OP, TCv, TAv>(x)); }
};

// integer -> f64 conversion (DTypeIntoFloatAPI::into_float = `as f64`)
template <class TAv>
struct ToF64Body {
    using TA = TAv;
    using TB = None;
    double* c;
    const TAv* a;
    static constexpr int OPS = 1;
    static constexpr size_t OUT_SIZE = sizeof(double);
    static constexpr bool NT_JOINT = true;
    template <bool NT = false>
    __device__ __forceinline__ TAv la(int64_t o) const { return ld<NT>(a + o); }
    template <bool NT = false>
    __device__ __forceinline__ None lb(int64_t) const { return None{}; }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t o, TAv x, None) const { sto<NT>(c + o, (double)x); }
};

/* =============================== host side =============================== */

enum class UnKind { Float, FloatReal, Square, Bool, BoolCplx, SignBit, Abs, Sign, Neg, Not };

static UnKind un_kind(int op) {
    switch (op) {
        case RSTSR_UN_CEIL:
        case RSTSR_UN_EXPM1:
        case RSTSR_UN_FLOOR:
        case RSTSR_UN_LOG1P:
        case RSTSR_UN_ROUND:
        case RSTSR_UN_TRUNC: return UnKind::FloatReal;
        case RSTSR_UN_IS_FINITE:
        case RSTSR_UN_IS_INF:
        case RSTSR_UN_IS_NAN: return UnKind::BoolCplx;
        case RSTSR_UN_SIGNBIT: return UnKind::SignBit;
        case RSTSR_UN_SQUARE: return UnKind::Square;
        case RSTSR_UN_ABS:
        case RSTSR_UN_IMAG:
        case RSTSR_UN_REAL: return UnKind::Abs;
        case RSTSR_UN_SIGN: return UnKind::Sign;
        case RSTSR_UN_NEG: return UnKind::Neg;
        case RSTSR_UN_NOT: return UnKind::Not;
        default: return UnKind::Float;
    }
}

static bool is_int_t(rstsr_dtype_t t) { return t == RSTSR_I32 || t == RSTSR_I64; }
static bool is_cplx_t(rstsr_dtype_t t) { return t == RSTSR_C64 || t == RSTSR_C128; }
static bool is_num_t(rstsr_dtype_t t) {
    return t == RSTSR_F32 || t == RSTSR_F64 || is_int_t(t) || is_cplx_t(t);
}

// The trait's TOut for input dtype ta; raises when the reference has no implementation.
static rstsr_dtype_t unary_out(int op, rstsr_dtype_t ta) {
    if (op < 0 || op > RSTSR_UN_NOT) RSTSR_RAISE(InvalidValue, "unknown unary op");
    auto none = [&]() -> rstsr_dtype_t {
        RSTSR_RAISE(UnImplemented, "this unary op has no implementation for the input dtype");
    };
    switch (un_kind(op)) {
        case UnKind::Float:
            if (!is_num_t(ta)) return none();
            return is_int_t(ta) ? RSTSR_F64 : ta;
        case UnKind::FloatReal:
            if (!is_num_t(ta) || is_cplx_t(ta)) return none();
            return is_int_t(ta) ? RSTSR_F64 : ta;
        case UnKind::BoolCplx:
            if (!is_num_t(ta) || is_int_t(ta)) return none();
            return RSTSR_BOOL;
        case UnKind::SignBit:
            if (!is_num_t(ta) || is_cplx_t(ta)) return none();
            return RSTSR_BOOL;
        case UnKind::Abs:
            if (!is_num_t(ta)) return none();
            return ta == RSTSR_C64 ? RSTSR_F32 : (ta == RSTSR_C128 ? RSTSR_F64 : ta);
        case UnKind::Square:
        case UnKind::Sign:
        case UnKind::Neg:
            if (!is_num_t(ta)) return none();
            return ta;
        case UnKind::Not:
            if (!(is_int_t(ta) || ta == RSTSR_BOOL)) return none();
            return ta;
        case UnKind::Bool: break;
    }
    return none();
}

template <class F>
static void dispatch_op(int op, F&& f) {
    switch (op) {
#define RSTSR_UC(X) \
    case X: f(std::integral_constant<int, X>{}); break;
        RSTSR_UC(RSTSR_UN_ACOS)
        RSTSR_UC(RSTSR_UN_ACOSH)
        RSTSR_UC(RSTSR_UN_ASIN)
        RSTSR_UC(RSTSR_UN_ASINH)
        RSTSR_UC(RSTSR_UN_ATAN)
        RSTSR_UC(RSTSR_UN_ATANH)
        RSTSR_UC(RSTSR_UN_CEIL)
        RSTSR_UC(RSTSR_UN_CONJ)
        RSTSR_UC(RSTSR_UN_COS)
        RSTSR_UC(RSTSR_UN_COSH)
        RSTSR_UC(RSTSR_UN_EXP)
        RSTSR_UC(RSTSR_UN_EXPM1)
        RSTSR_UC(RSTSR_UN_FLOOR)
        RSTSR_UC(RSTSR_UN_IS_FINITE)
        RSTSR_UC(RSTSR_UN_IS_INF)
        RSTSR_UC(RSTSR_UN_IS_NAN)
        RSTSR_UC(RSTSR_UN_LOG)
        RSTSR_UC(RSTSR_UN_LOG1P)
        RSTSR_UC(RSTSR_UN_LOG2)
        RSTSR_UC(RSTSR_UN_LOG10)
        RSTSR_UC(RSTSR_UN_ROUND)
        RSTSR_UC(RSTSR_UN_SIGNBIT)
        RSTSR_UC(RSTSR_UN_SIN)
        RSTSR_UC(RSTSR_UN_SINH)
        RSTSR_UC(RSTSR_UN_SQUARE)
        RSTSR_UC(RSTSR_UN_SQRT)
        RSTSR_UC(RSTSR_UN_TAN)
        RSTSR_UC(RSTSR_UN_TANH)
        RSTSR_UC(RSTSR_UN_TRUNC)
        RSTSR_UC(RSTSR_UN_ABS)
        RSTSR_UC(RSTSR_UN_IMAG)
        RSTSR_UC(RSTSR_UN_REAL)
        RSTSR_UC(RSTSR_UN_SIGN)
        RSTSR_UC(RSTSR_UN_NEG)
        RSTSR_UC(RSTSR_UN_NOT)
#undef RSTSR_UC
        // inv and reciprocal share one functor (ComplexFloat::recip)
        case RSTSR_UN_INV:
        case RSTSR_UN_RECIPROCAL: f(std::integral_constant<int, RSTSR_UN_RECIPROCAL>{}); break;
        default: RSTSR_RAISE(InvalidValue, "unknown unary op");
    }
}

template <class F>
static void dispatch_in(rstsr_dtype_t t, F&& f) {
    switch (t) {
        case RSTSR_F32: f(float{}); break;
        case RSTSR_F64: f(double{}); break;
        case RSTSR_I32: f(int32_t{}); break;
        case RSTSR_I64: f(int64_t{}); break;
        case RSTSR_C64: f(c64{}); break;
        case RSTSR_C128: f(c128{}); break;
        case RSTSR_BOOL: f(uint8_t{}); break;
        default: RSTSR_RAISE(UnImplemented, "dtype not supported for this operation");
    }
}

template <class T>
struct RealOfT {
    using type = T;
};
template <>
struct RealOfT<c64> {
    using type = float;
};
template <>
struct RealOfT<c128> {
    using type = double;
};

// Device-side output type of op OP on input TA, mirroring unary_out(); void when none.
template <int OP, class TA>
struct UnOut {
    static constexpr bool I = std::is_integral<TA>::value && !std::is_same<TA, uint8_t>::value;
    static constexpr bool B = std::is_same<TA, uint8_t>::value;
    static constexpr bool X = is_cplx_v<TA>;
    using type = typename std::conditional<
        OP == RSTSR_UN_NOT, typename std::conditional<(I || B), TA, void>::type,
        typename std::conditional<
            B, void,
            typename std::conditional<
                OP == RSTSR_UN_IS_FINITE || OP == RSTSR_UN_IS_INF || OP == RSTSR_UN_IS_NAN,
                typename std::conditional<I, void, uint8_t>::type,
                typename std::conditional<
                    OP == RSTSR_UN_SIGNBIT, typename std::conditional<X, void, uint8_t>::type,
                    typename std::conditional<
                        OP == RSTSR_UN_SQUARE || OP == RSTSR_UN_SIGN || OP == RSTSR_UN_NEG, TA,
                        typename std::conditional<
                            OP == RSTSR_UN_ABS || OP == RSTSR_UN_REAL || OP == RSTSR_UN_IMAG,
                            typename RealOfT<TA>::type,
                            typename std::conditional<
                                (OP == RSTSR_UN_CEIL || OP == RSTSR_UN_EXPM1 || OP == RSTSR_UN_FLOOR ||
                                 OP == RSTSR_UN_LOG1P || OP == RSTSR_UN_ROUND || OP == RSTSR_UN_TRUNC) &&
                                    X,
                                void, typename std::conditional<I, double, TA>::type>::type>::type>::type>::
                    type>::type>::type>::type;
};

static void unary_impl(rstsr_hip_ctx* ctx, int op, rstsr_dtype_t tc, void* c, const Layout& lc, rstsr_dtype_t ta,
                       const void* a, const Layout& la) {
    if (lc.shape != la.shape) RSTSR_RAISE(InvalidLayout, "unary op requires equal shapes");
    rstsr_dtype_t want = unary_out(op, ta);
    if (tc != want) RSTSR_RAISE(InvalidValue, "output dtype differs from the op's output dtype for this input");
    check_extent(c, lc, dtype_size(tc), "c");
    check_extent(a, la, dtype_size(ta), "a");
    const Stride* s[3] = {&lc.stride, &la.stride, nullptr};
    size_t off[3] = {lc.offset, la.offset, 0};
    Dims d = collapse3(lc.shape, s, off);
    EwChoice ch = choose(d, true, false);
    const UnKind k = un_kind(op);
    const bool two_pass = is_int_t(ta) && (k == UnKind::Float || k == UnKind::FloatReal);
    if (two_pass) {
        // into_float, then the f64 op in place on the output
        dispatch_in(ta, [&](auto atag) {
            using TA = decltype(atag);
            if constexpr (std::is_same<TA, int32_t>::value || std::is_same<TA, int64_t>::value) {
                ToF64Body<TA> body{(double*)c, (const TA*)a};
                launch_ew<ToF64Body<TA>, false>(ctx, body, d, ch);
            }
        });
        const Stride* s2[3] = {&lc.stride, &lc.stride, nullptr};
        size_t off2[3] = {lc.offset, lc.offset, 0};
        Dims d2 = collapse3(lc.shape, s2, off2);
        EwChoice ch2 = choose(d2, true, false);
        dispatch_op(op, [&](auto opc) {
            constexpr int OP = decltype(opc)::value;
            if constexpr (std::is_same<typename UnOut<OP, double>::type, double>::value) {
                UnBody<OP, double, double> body{(double*)c, (const double*)c};
                launch_ew<UnBody<OP, double, double>, false>(ctx, body, d2, ch2);
            }
        });
        return;
    }
    dispatch_in(ta, [&](auto atag) {
        using TA = decltype(atag);
        dispatch_op(op, [&](auto opc) {
            constexpr int OP = decltype(opc)::value;
            using TC = typename UnOut<OP, TA>::type;
            constexpr bool int_float = std::is_integral<TA>::value && std::is_same<TC, double>::value;
            if constexpr (!std::is_void<TC>::value && !int_float) {
                UnBody<OP, TC, TA> body{(TC*)c, (const TA*)a};
                launch_ew<UnBody<OP, TC, TA>, false>(ctx, body, d, ch);
            } else {
                RSTSR_RAISE(UnImplemented, "this unary op has no implementation for the input dtype");
            }
        });
    });
}

}  // namespace hip
}  // namespace rstsr

using namespace rstsr;
using namespace rstsr::hip;

extern "C" {

int rstsr_hip_unary_out_dtype(rstsr_unop_t op, rstsr_dtype_t ta, rstsr_dtype_t* tc) {
    return guarded([&] {
        if (!tc) RSTSR_RAISE(InvalidValue, "null output");
        *tc = unary_out((int)op, ta);
    });
}

int rstsr_hip_unary(rstsr_hip_ctx_t* ctx, rstsr_unop_t op, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc_,
                    rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la_) {
    return guarded([&] {
        bind(ctx);
        if (!c || !a) RSTSR_RAISE(InvalidValue, "null operand");
        Layout lc = from_abi(lc_), la = from_abi(la_);
        if (c == a && !(lc.shape == la.shape && lc.stride == la.stride && lc.offset == la.offset))
            RSTSR_RAISE(InvalidLayout, "an in-place unary op needs identical input and output layouts");
        unary_impl(ctx, (int)op, tc, c, lc, ta, a, la);
    });
}

int rstsr_hip_unary_inplace(rstsr_hip_ctx_t* ctx, rstsr_unop_t op, rstsr_dtype_t t, void* c,
                            const rstsr_layout_t* lc_) {
    return guarded([&] {
        bind(ctx);
        if (!c) RSTSR_RAISE(InvalidValue, "null operand");
        Layout lc = from_abi(lc_);
        if (unary_out((int)op, t) != t)
            RSTSR_RAISE(InvalidValue, "in-place unary op needs an output dtype equal to the input dtype");
        unary_impl(ctx, (int)op, t, c, lc, t, c, lc);
    });
}

}  // extern "C"
