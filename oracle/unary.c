/*
 * unary.c — CPU restatement of the reference's unary elementwise functors, used ONLY as
 * test infrastructure (parity checker for the HIP unary kernels); the product library
 * never links or calls this file.
 *
 * Follows rstsr-core/src/device_cpu_serial/operators/op_binary_common.rs:10-237
 * (same-type and T::FloatType ops, boolean classifiers, abs/imag/real/sign) and
 * op_binary_arithmetic.rs:96 (neg, not), over the serial row-major odometer of oracle.c:
 *  - real floats call the C math library, which is what Rust's f32/f64 methods (and
 *    num-traits' Float with its std feature) call on Linux;
 *  - integers convert with `as f64` (DTypeIntoFloatAPI) before a T::FloatType op;
 *  - complex functions restate num-complex 0.4 (a crates.io dependency of the reference,
 *    `num = "0.4"` in Cargo.toml:59, not vendored; no lock file pins the patch release):
 *    Complex::{exp, ln, sqrt, sin, cos, tan, sinh, cosh, tanh, asin, acos, atan, asinh,
 *    acosh, atanh, finv} with num-complex's own Mul (ac - bd, ad + bc) and Div
 *    ((ac + bd)/n, (bc - ad)/n, n = c^2 + d^2);
 *  - sign/abs/real/imag: rstsr-dtype-traits/src/ext_num.rs:60-330;
 *  - signbit is num::Signed::is_positive (x > 0 for integers, sign bit clear for floats).
 * Integer arithmetic wraps (Rust release builds). Built with -ffp-contract=off.
 *
 * Pinning (tests/test_oracle.py): the reference's own known answers
 * (rstsr-core/tests/core_func/math/test_unary_math.rs, doc_draft/math/test_math_doc.rs),
 * NumPy/cmath on the same inputs (real functions within 2 ulp; complex functions on
 * principal branches away from the cuts), exact identities (floor/ceil/trunc/round on
 * halves, sign and signbit on signed zeros and NaN) and planted mistakes.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    size_t ndim;
    const size_t* shape;
    const ptrdiff_t* stride;
    size_t offset;
} ulayout;

enum { F32 = 0, F64 = 1, I32 = 2, I64 = 3, C64 = 4, C128 = 5, U64 = 6, BOOL = 7 };
enum {
    ACOS, ACOSH, ASIN, ASINH, ATAN, ATANH, CEIL, CONJ, COS, COSH, EXP, EXPM1, FLOOR, INV, IS_FINITE,
    IS_INF, IS_NAN, LOG, LOG1P, LOG2, LOG10, RECIPROCAL, ROUND, SIGNBIT, SIN, SINH, SQUARE, SQRT, TAN,
    TANH, TRUNC, ABS, IMAG, REAL, SIGN, NEG, NOT, NUM_UNARY
};

typedef struct { float re, im; } cf_t;
typedef struct { double re, im; } cd_t;

/* ---------------- real functions (num::Float on f64 / f32) ---------------- */

#define REAL_FN(R, S)                                                      \
    static int real_##R(int op, R x, R* y) {                              \
        switch (op) {                                                      \
            case ACOS: *y = acos##S(x); break;                             \
            case ACOSH: *y = acosh##S(x); break;                           \
            case ASIN: *y = asin##S(x); break;                             \
            case ASINH: *y = asinh##S(x); break;                           \
            case ATAN: *y = atan##S(x); break;                             \
            case ATANH: *y = atanh##S(x); break;                           \
            case CEIL: *y = ceil##S(x); break;                             \
            case CONJ: *y = x; break;                                      \
            case COS: *y = cos##S(x); break;                               \
            case COSH: *y = cosh##S(x); break;                             \
            case EXP: *y = exp##S(x); break;                               \
            case EXPM1: *y = expm1##S(x); break;                           \
            case FLOOR: *y = floor##S(x); break;                           \
            case INV:                                                      \
            case RECIPROCAL: *y = (R)1 / x; break;                         \
            case LOG: *y = log##S(x); break;                               \
            case LOG1P: *y = log1p##S(x); break;                           \
            case LOG2: *y = log2##S(x); break;                             \
            case LOG10: *y = log10##S(x); break;                           \
            case ROUND: *y = round##S(x); break;                           \
            case SIN: *y = sin##S(x); break;                               \
            case SINH: *y = sinh##S(x); break;                             \
            case SQRT: *y = sqrt##S(x); break;                             \
            case TAN: *y = tan##S(x); break;                               \
            case TANH: *y = tanh##S(x); break;                             \
            case TRUNC: *y = trunc##S(x); break;                           \
            default: return 1;                                             \
        }                                                                  \
        return 0;                                                          \
    }
REAL_FN(double, )
REAL_FN(float, f)

/* ---------------- complex functions (num-complex 0.4) ---------------- */

#define CPLX_FNS(C, R, S)                                                                          \
    static C C##_mk(R re, R im) { C z; z.re = re; z.im = im; return z; }                            \
    static C C##_mul(C a, C b) { return C##_mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); } \
    static C C##_div(C a, C b) {                                                                   \
        R n = b.re * b.re + b.im * b.im;                                                           \
        R re = a.re * b.re + a.im * b.im;                                                          \
        R im = a.im * b.re - a.re * b.im;                                                          \
        return C##_mk(re / n, im / n);                                                             \
    }                                                                                              \
    static C C##_add(C a, C b) { return C##_mk(a.re + b.re, a.im + b.im); }                       \
    static C C##_sub(C a, C b) { return C##_mk(a.re - b.re, a.im - b.im); }                       \
    static C C##_neg(C a) { return C##_mk(-a.re, -a.im); }                                        \
    static C C##_unscale(C a, R t) { return C##_mk(a.re / t, a.im / t); }                         \
    static R C##_norm(C a) { return hypot##S(a.re, a.im); }                                        \
    static R C##_arg(C a) { return atan2##S(a.im, a.re); }                                         \
    static C C##_from_polar(R r, R t) { return C##_mk(r * cos##S(t), r * sin##S(t)); }            \
    static int C##_eq(C a, C b) { return a.re == b.re && a.im == b.im; }                           \
    static C C##_exp(C z) {                                                                        \
        R re = z.re, im = z.im;                                                                    \
        if (isinf(re)) {                                                                           \
            if (re < 0) {                                                                          \
                if (!isfinite(im)) return C##_mk(0, 0);                                            \
            } else if (im == 0 || !isfinite(im)) {                                                 \
                if (isinf(im)) im = (R)NAN;                                                        \
                return C##_mk(re, im);                                                             \
            }                                                                                      \
        } else if (isnan(re) && im == 0) {                                                         \
            return z;                                                                              \
        }                                                                                          \
        return C##_from_polar(exp##S(re), im);                                                     \
    }                                                                                              \
    static C C##_ln(C z) { return C##_mk(log##S(C##_norm(z)), C##_arg(z)); }                      \
    static C C##_sqrt(C z) {                                                                       \
        if (z.im == 0) {                                                                           \
            if (!signbit(z.re)) return C##_mk(sqrt##S(z.re), z.im);                                \
            R im = sqrt##S(-z.re);                                                                 \
            return signbit(z.im) ? C##_mk(0, -im) : C##_mk(0, im);                                 \
        }                                                                                          \
        if (z.re == 0) {                                                                           \
            R x = sqrt##S(fabs##S(z.im) / (R)2);                                                   \
            return signbit(z.im) ? C##_mk(x, -x) : C##_mk(x, x);                                   \
        }                                                                                          \
        return C##_from_polar(sqrt##S(C##_norm(z)), C##_arg(z) / (R)2);                            \
    }                                                                                              \
    static C C##_fn(int op, C z, int* bad) {                                                       \
        const C i = C##_mk(0, 1), one = C##_mk(1, 0);                                              \
        const C two = C##_add(one, one);                                                           \
        R a, b;                                                                                    \
        *bad = 0;                                                                                  \
        switch (op) {                                                                              \
            case ACOS: /* -i ln(i sqrt(1 - z^2) + z) */                                            \
                return C##_mul(C##_neg(i),                                                         \
                               C##_ln(C##_add(C##_mul(i, C##_sqrt(C##_sub(one, C##_mul(z, z)))), z))); \
            case ACOSH: /* 2 ln(sqrt((z + 1)/2) + sqrt((z - 1)/2)) */                              \
                return C##_mul(two, C##_ln(C##_add(C##_sqrt(C##_div(C##_add(z, one), two)),        \
                                                   C##_sqrt(C##_div(C##_sub(z, one), two)))));     \
            case ASIN: /* -i ln(sqrt(1 - z^2) + i z) */                                            \
                return C##_mul(C##_neg(i),                                                         \
                               C##_ln(C##_add(C##_sqrt(C##_sub(one, C##_mul(z, z))), C##_mul(i, z)))); \
            case ASINH: /* ln(z + sqrt(1 + z^2)) */                                                \
                return C##_ln(C##_add(z, C##_sqrt(C##_add(one, C##_mul(z, z)))));                 \
            case ATAN: /* (ln(1 + i z) - ln(1 - i z)) / (2i) */                                    \
                if (C##_eq(z, i)) return C##_mk(0, (R)INFINITY);                                   \
                if (C##_eq(z, C##_neg(i))) return C##_mk(0, -(R)INFINITY);                         \
                return C##_div(C##_sub(C##_ln(C##_add(one, C##_mul(i, z))),                        \
                                       C##_ln(C##_sub(one, C##_mul(i, z)))),                       \
                               C##_mul(two, i));                                                   \
            case ATANH: /* (ln(1 + z) - ln(1 - z)) / 2 */                                          \
                if (C##_eq(z, one)) return C##_mk((R)INFINITY, 0);                                 \
                if (C##_eq(z, C##_neg(one))) return C##_mk(-(R)INFINITY, 0);                       \
                return C##_div(C##_sub(C##_ln(C##_add(one, z)), C##_ln(C##_sub(one, z))), two);    \
            case CONJ: return C##_mk(z.re, -z.im);                                                 \
            case COS: return C##_mk(cos##S(z.re) * cosh##S(z.im), -sin##S(z.re) * sinh##S(z.im)); \
            case COSH: return C##_mk(cosh##S(z.re) * cos##S(z.im), sinh##S(z.re) * sin##S(z.im)); \
            case EXP: return C##_exp(z);                                                           \
            case INV:                                                                              \
            case RECIPROCAL: { /* finv: conj(z) / |z| / |z| */                                    \
                R n = C##_norm(z);                                                                 \
                return C##_unscale(C##_unscale(C##_mk(z.re, -z.im), n), n);                        \
            }                                                                                      \
            case LOG: return C##_ln(z);                                                            \
            case LOG2: return C##_unscale(C##_ln(z), (R)0.693147180559945309417232121458176568);   \
            case LOG10: return C##_unscale(C##_ln(z), (R)2.30258509299404568401799145468436421);   \
            case SIN: return C##_mk(sin##S(z.re) * cosh##S(z.im), cos##S(z.re) * sinh##S(z.im));  \
            case SINH: return C##_mk(sinh##S(z.re) * cos##S(z.im), cosh##S(z.re) * sin##S(z.im)); \
            case SQRT: return C##_sqrt(z);                                                         \
            case TAN:                                                                              \
                a = z.re + z.re;                                                                   \
                b = z.im + z.im;                                                                   \
                return C##_unscale(C##_mk(sin##S(a), sinh##S(b)), cos##S(a) + cosh##S(b));        \
            case TANH:                                                                             \
                a = z.re + z.re;                                                                   \
                b = z.im + z.im;                                                                   \
                return C##_unscale(C##_mk(sinh##S(a), sin##S(b)), cosh##S(a) + cos##S(b));         \
            case SQUARE: return C##_mul(z, z);                                                     \
            case NEG: return C##_neg(z);                                                           \
            case SIGN: {                                                                           \
                R n = C##_norm(z);                                                                 \
                if (n == 0) return C##_mk(0, 0);                                                   \
                return C##_unscale(z, n);                                                          \
            }                                                                                      \
            default: *bad = 1; return z;                                                           \
        }                                                                                          \
    }
CPLX_FNS(cd_t, double, )
CPLX_FNS(cf_t, float, f)

/* ---------------- output dtype (the trait's TOut) ---------------- */

static int is_int(int t) { return t == I32 || t == I64; }
static int is_cplx(int t) { return t == C64 || t == C128; }
static int is_num(int t) { return t == F32 || t == F64 || is_int(t) || is_cplx(t); }

/* returns -1 when the reference has no implementation for this input type */
int o_unary_out_dtype(int op, int ta) {
    switch (op) {
        case CEIL: case EXPM1: case FLOOR: case LOG1P: case ROUND: case TRUNC:
            if (!is_num(ta) || is_cplx(ta)) return -1;
            return is_int(ta) ? F64 : ta;
        case IS_FINITE: case IS_INF: case IS_NAN:
            if (!is_num(ta) || is_int(ta)) return -1;
            return BOOL;
        case SIGNBIT:
            if (!is_num(ta) || is_cplx(ta)) return -1;
            return BOOL;
        case ABS: case IMAG: case REAL:
            if (!is_num(ta)) return -1;
            return ta == C64 ? F32 : (ta == C128 ? F64 : ta);
        case SQUARE: case SIGN: case NEG:
            if (!is_num(ta)) return -1;
            return ta;
        case NOT:
            if (!(is_int(ta) || ta == BOOL)) return -1;
            return ta;
        default:
            if (op < 0 || op >= NUM_UNARY || !is_num(ta)) return -1;
            return is_int(ta) ? F64 : ta;
    }
}

static size_t usize_of(int t) {
    switch (t) {
        case F32: case I32: return 4;
        case F64: case I64: case C64: case U64: return 8;
        case C128: return 16;
        case BOOL: return 1;
    }
    return 0;
}

/* one element: y = op(x); tc is o_unary_out_dtype(op, ta) */
static int unary_elem(int op, int tc, void* y, int ta, const void* x) {
    int bad = 0;
    if (op == IS_FINITE || op == IS_INF || op == IS_NAN) {
        double re, im = 0;
        if (ta == F32) re = *(const float*)x;
        else if (ta == F64) re = *(const double*)x;
        else if (ta == C64) { re = ((const float*)x)[0]; im = ((const float*)x)[1]; }
        else { re = ((const double*)x)[0]; im = ((const double*)x)[1]; }
        uint8_t r;
        if (op == IS_FINITE) r = isfinite(re) && isfinite(im);
        else if (op == IS_NAN) r = isnan(re) || isnan(im);
        else r = !(isnan(re) || isnan(im)) && (isinf(re) || isinf(im));
        *(uint8_t*)y = r;
        return 0;
    }
    if (op == SIGNBIT) {
        uint8_t r;
        switch (ta) {
            case I32: r = *(const int32_t*)x > 0; break;
            case I64: r = *(const int64_t*)x > 0; break;
            case F32: r = !signbit(*(const float*)x); break;
            default: r = !signbit(*(const double*)x); break;
        }
        *(uint8_t*)y = r;
        return 0;
    }
    if (ta == BOOL) { /* not */
        *(uint8_t*)y = *(const uint8_t*)x ? 0 : 1;
        return 0;
    }
    if (is_int(ta)) {
        int64_t v = ta == I32 ? *(const int32_t*)x : *(const int64_t*)x;
        int64_t r;
        switch (op) {
            case SQUARE: r = (int64_t)((uint64_t)v * (uint64_t)v); break;
            case NEG: r = (int64_t)(0 - (uint64_t)v); break;
            case ABS: r = v < 0 ? (int64_t)(0 - (uint64_t)v) : v; break;
            case REAL: r = v; break;
            case IMAG: r = 0; break;
            case SIGN: r = v > 0 ? 1 : (v < 0 ? -1 : 0); break;
            case NOT: r = ~v; break;
            default: { /* T::FloatType op: `as f64`, then the f64 function */
                double d = (double)v, o;
                if (real_double(op, d, &o)) return 1;
                *(double*)y = o;
                return 0;
            }
        }
        if (ta == I32) {
            /* 32-bit wrap of the 64-bit result */
            uint32_t u = (uint32_t)(uint64_t)r;
            int32_t w;
            if (op == SQUARE) u = (uint32_t)((uint32_t)v * (uint32_t)v);
            memcpy(&w, &u, 4);
            *(int32_t*)y = w;
        } else {
            *(int64_t*)y = r;
        }
        return 0;
    }
    if (ta == F64 || ta == F32) {
#define REAL_ELEM(R, FN, FABS)                                                          \
        {                                                                               \
            R v = *(const R*)x, r;                                                      \
            switch (op) {                                                               \
                case SQUARE: r = v * v; break;                                          \
                case NEG: r = -v; break;                                                \
                case ABS: r = FABS(v); break;                                           \
                case REAL: r = v; break;                                                \
                case IMAG: r = 0; break;                                                \
                case SIGN: r = (v != v) ? v : (v > 0 ? (R)1 : (v < 0 ? (R)-1 : (R)0)); break; \
                default: if (FN(op, v, &r)) return 1;                                   \
            }                                                                           \
            *(R*)y = r;                                                                 \
            return 0;                                                                   \
        }
        if (ta == F64) REAL_ELEM(double, real_double, fabs)
        else REAL_ELEM(float, real_float, fabsf)
#undef REAL_ELEM
    }
    if (ta == C128) {
        cd_t z = *(const cd_t*)x;
        if (op == ABS) { *(double*)y = cd_t_norm(z); return 0; }
        if (op == REAL) { *(double*)y = z.re; return 0; }
        if (op == IMAG) { *(double*)y = z.im; return 0; }
        cd_t r = cd_t_fn(op, z, &bad);
        if (bad) return 1;
        *(cd_t*)y = r;
        return 0;
    }
    if (ta == C64) {
        cf_t z = *(const cf_t*)x;
        if (op == ABS) { *(float*)y = cf_t_norm(z); return 0; }
        if (op == REAL) { *(float*)y = z.re; return 0; }
        if (op == IMAG) { *(float*)y = z.im; return 0; }
        cf_t r = cf_t_fn(op, z, &bad);
        if (bad) return 1;
        *(cf_t*)y = r;
        return 0;
    }
    (void)tc;
    return 1;
}

/* op_muta_refb (op_binary_common.rs:45-57): c = op(a) over two layouts of one shape,
 * in logical row-major order. Returns 1 for an op/dtype pair without implementation. */
int o_unary(int op, int tc, void* c, const ulayout* lc, int ta, const void* a, const ulayout* la) {
    if (o_unary_out_dtype(op, ta) != tc || tc < 0) return 1;
    if (lc->ndim != la->ndim) return 1;
    size_t n = 1;
    for (size_t i = 0; i < lc->ndim; ++i) {
        if (lc->shape[i] != la->shape[i]) return 1;
        n *= lc->shape[i];
    }
    if (n == 0) return 0;
    size_t ec = usize_of(tc), ea = usize_of(ta), nd = lc->ndim;
    size_t idx[32];
    if (nd > 32) return 1;
    memset(idx, 0, sizeof(idx));
    for (size_t k = 0; k < n; ++k) {
        ptrdiff_t oc = (ptrdiff_t)lc->offset, oa = (ptrdiff_t)la->offset;
        for (size_t d = 0; d < nd; ++d) {
            oc += (ptrdiff_t)idx[d] * lc->stride[d];
            oa += (ptrdiff_t)idx[d] * la->stride[d];
        }
        if (unary_elem(op, tc, (char*)c + oc * (ptrdiff_t)ec, ta, (const char*)a + oa * (ptrdiff_t)ea)) return 1;
        for (size_t d = nd; d-- > 0;) {
            if (++idx[d] < lc->shape[d]) break;
            idx[d] = 0;
        }
    }
    return 0;
}
