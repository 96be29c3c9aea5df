/*
 * oracle.c — CPU restatement of the RSTSR (v0.7.10) CPU path for the
 * elementwise / copy / reduction / matmul operations, used ONLY as test
 * infrastructure: by tests/ as the parity checker, by __graft_entry__.smoke()
 * as its checker, and by bench.py's cpu_baseline leg as the timed CPU baseline.
 * The product library never links or calls this file.
 *
 * Two families:
 *  - o_* : serial, logical row-major odometer semantics of
 *          rstsr-native-impl/src/cpu_serial (the DeviceCpuSerial device).
 *  - b_* : OpenMP baseline following the control flow of
 *          rstsr-native-impl/src/cpu_rayon (the DeviceFaer device): greedy
 *          (K) iteration order, a contiguous inner run when present, and a
 *          parallel split over outer iterations.
 *
 * Built with -ffp-contract=off so that a*b+c is never fused (Rust never fuses).
 *
 * Pinning (see tests/test_oracle.py): every function is checked against the
 * reference's own known answers (rstsr-core/src/tensor/reduction.rs:417-870,
 * tensor/operators/op_binary_arithmetic.rs:994-1070, core_func tests) and
 * against mathematically fixed results: NumPy's IEEE add/mul on the same views
 * (bit-exact), math.fsum (sums), numpy argmax/argmin first-occurrence,
 * Python's complex product, exact integer GEMMs and a longdouble GEMM.
 */
#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>

typedef struct {
    size_t ndim;
    const size_t* shape;
    const ptrdiff_t* stride;
    size_t offset;
} olayout;

enum { F32 = 0, F64 = 1, I32 = 2, I64 = 3, C64 = 4, C128 = 5, U64 = 6, BOOL = 7 };
/* operators of auto_impl/op_ternary_arithmetic.rs:3-14: Add Sub Mul Div Rem BitAnd BitOr BitXor Shl Shr */
enum { OP_ADD = 0, OP_SUB = 1, OP_MUL = 2, OP_DIV = 3, OP_REM = 4, OP_AND = 5, OP_OR = 6, OP_XOR = 7, OP_SHL = 8,
       OP_SHR = 9 };
enum { R_SUM = 0, R_MAX = 1, R_MIN = 2, R_ARGMAX = 3, R_ARGMIN = 4, R_PROD = 5, R_MEAN = 6, R_VAR = 7, R_STD = 8, R_L2 = 9,
       R_ALL = 10, R_ANY = 11, R_COUNT = 12 };

typedef struct { float re, im; } c64_t;
typedef struct { double re, im; } c128_t;

#define MAXND 32

static size_t lsize(const olayout* l) {
    size_t n = 1;
    for (size_t i = 0; i < l->ndim; ++i) n *= l->shape[i];
    return n;
}

/* ---------------- odometer (rstsr-common/src/layout/iterator.rs:23-343, row-major) ---------------- */

typedef struct {
    size_t nd;
    size_t idx[MAXND];
    ptrdiff_t off;
} odo_t;

static void odo_init(odo_t* o, const olayout* l) {
    o->nd = l->ndim;
    memset(o->idx, 0, sizeof(o->idx));
    o->off = (ptrdiff_t)l->offset;
}

/* advance in row-major order; returns 0 when exhausted */
static int odo_next(odo_t* o, const olayout* l) {
    for (size_t ii = l->ndim; ii-- > 0;) {
        o->idx[ii] += 1;
        o->off += l->stride[ii];
        if (o->idx[ii] < l->shape[ii]) return 1;
        o->off -= (ptrdiff_t)l->shape[ii] * l->stride[ii];
        o->idx[ii] = 0;
    }
    return 0;
}

/* ---------------- scalar helpers ---------------- */

static c128_t cmul128(c128_t a, c128_t b) {
    c128_t r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static c64_t cmul64(c64_t a, c64_t b) {
    c64_t r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* num-complex 0.4 Div: ((a.re c + a.im d) / n, (a.im c - a.re d) / n), n = c*c + d*d; and Rem:
 * q = a / b, q0 = (q.re - q.re % 1, q.im - q.im % 1) (q truncated toward zero), a - b * q0. */
static c128_t cdiv128(c128_t a, c128_t b) {
    double n = b.re * b.re + b.im * b.im;
    c128_t r;
    r.re = (a.re * b.re + a.im * b.im) / n;
    r.im = (a.im * b.re - a.re * b.im) / n;
    return r;
}
static c64_t cdiv64(c64_t a, c64_t b) {
    float n = b.re * b.re + b.im * b.im;
    c64_t r;
    r.re = (a.re * b.re + a.im * b.im) / n;
    r.im = (a.im * b.re - a.re * b.im) / n;
    return r;
}
static c128_t crem128(c128_t a, c128_t b) {
    c128_t q = cdiv128(a, b), q0, p, r;
    q0.re = q.re - fmod(q.re, 1.0);
    q0.im = q.im - fmod(q.im, 1.0);
    p = cmul128(b, q0);
    r.re = a.re - p.re;
    r.im = a.im - p.im;
    return r;
}
static c64_t crem64(c64_t a, c64_t b) {
    c64_t q = cdiv64(a, b), q0, p, r;
    q0.re = q.re - fmodf(q.re, 1.0f);
    q0.im = q.im - fmodf(q.im, 1.0f);
    p = cmul64(b, q0);
    r.re = a.re - p.re;
    r.im = a.im - p.im;
    return r;
}

/* Integer ops as Rust release builds compute them: add/sub/mul wrap; / and % truncate toward
 * zero; << and >> take the shift amount modulo the bit width, >> is arithmetic. Rust panics on
 * a zero divisor and on MIN / -1 (MIN % -1); the values given here for those cases (0, and the
 * wrapped MIN / -1 = MIN, MIN % -1 = 0) are the device's documented substitutes. */
#define INT_OP(T, U, BITS)                                                              \
    static T int_op_##T(int op, T x, T y) {                                             \
        U ux = (U)x, uy = (U)y;                                                         \
        switch (op) {                                                                   \
            case OP_ADD: return (T)(ux + uy);                                           \
            case OP_SUB: return (T)(ux - uy);                                           \
            case OP_MUL: return (T)(ux * uy);                                           \
            case OP_DIV: return y == 0 ? (T)0 : y == -1 ? (T)((U)0 - ux) : (T)(x / y);  \
            case OP_REM: return (y == 0 || y == -1) ? (T)0 : (T)(x % y);                \
            case OP_AND: return (T)(ux & uy);                                           \
            case OP_OR: return (T)(ux | uy);                                            \
            case OP_XOR: return (T)(ux ^ uy);                                           \
            case OP_SHL: return (T)(ux << (uy & (BITS - 1)));                           \
            default: return (T)(x >> (uy & (BITS - 1)));                                \
        }                                                                               \
    }
INT_OP(int32_t, uint32_t, 32)
INT_OP(int64_t, uint64_t, 64)

/* apply op on element pointers (operands of one dtype; every combination the reference's
 * types implement: arithmetic for numbers, Rem with fmod for floats, bit ops for integers and
 * bool, shifts for integers) */
static void apply(int op, int t, void* c, const void* a, const void* b) {
    switch (t) {
        case F32: {
            float x = *(const float*)a, y = *(const float*)b, r;
            r = op == OP_ADD ? x + y : op == OP_SUB ? x - y : op == OP_MUL ? x * y : op == OP_DIV ? x / y : fmodf(x, y);
            *(float*)c = r;
            break;
        }
        case F64: {
            double x = *(const double*)a, y = *(const double*)b, r;
            r = op == OP_ADD ? x + y : op == OP_SUB ? x - y : op == OP_MUL ? x * y : op == OP_DIV ? x / y : fmod(x, y);
            *(double*)c = r;
            break;
        }
        case I32: *(int32_t*)c = int_op_int32_t(op, *(const int32_t*)a, *(const int32_t*)b); break;
        case I64: *(int64_t*)c = int_op_int64_t(op, *(const int64_t*)a, *(const int64_t*)b); break;
        case BOOL: {
            uint8_t x = *(const uint8_t*)a != 0, y = *(const uint8_t*)b != 0;
            *(uint8_t*)c = op == OP_AND ? (x & y) : op == OP_OR ? (x | y) : (x ^ y);
            break;
        }
        case C64: {
            c64_t x = *(const c64_t*)a, y = *(const c64_t*)b, r;
            if (op == OP_MUL) r = cmul64(x, y);
            else if (op == OP_ADD) { r.re = x.re + y.re; r.im = x.im + y.im; }
            else if (op == OP_SUB) { r.re = x.re - y.re; r.im = x.im - y.im; }
            else if (op == OP_DIV) r = cdiv64(x, y);
            else r = crem64(x, y);
            *(c64_t*)c = r;
            break;
        }
        case C128: {
            c128_t x = *(const c128_t*)a, y = *(const c128_t*)b, r;
            if (op == OP_MUL) r = cmul128(x, y);
            else if (op == OP_ADD) { r.re = x.re + y.re; r.im = x.im + y.im; }
            else if (op == OP_SUB) { r.re = x.re - y.re; r.im = x.im - y.im; }
            else if (op == OP_DIV) r = cdiv128(x, y);
            else r = crem128(x, y);
            *(c128_t*)c = r;
            break;
        }
    }
}

static size_t esize(int t) {
    switch (t) {
        case F32: return 4;
        case F64: return 8;
        case I32: return 4;
        case I64: return 8;
        case C64: return 8;
        case C128: return 16;
        case U64: return 8;
        case BOOL: return 1;
    }
    return 0;
}

/* ================= serial oracle (cpu_serial semantics) ================= */

/* op_mutc_refa_refb_func_cpu_serial (cpu_serial/op_with_func.rs:10-45): c = a op b over
 * three layouts of one shape. b_scalar: b is one value (op_mutc_refa_numb). */
int o_binary(int op, int t, void* c, const olayout* lc, const void* a, const olayout* la, const void* b,
             const olayout* lb) {
    size_t n = lsize(lc), es = esize(t);
    if (n == 0) return 0;
    odo_t oc, oa, ob;
    odo_init(&oc, lc);
    odo_init(&oa, la);
    if (lb) odo_init(&ob, lb);
    for (size_t i = 0; i < n; ++i) {
        const char* pb = lb ? (const char*)b + ob.off * (ptrdiff_t)es : (const char*)b;
        apply(op, t, (char*)c + oc.off * (ptrdiff_t)es, (const char*)a + oa.off * (ptrdiff_t)es, pb);
        odo_next(&oc, lc);
        odo_next(&oa, la);
        if (lb) odo_next(&ob, lb);
    }
    return 0;
}

/* c = a op b with one operand complex and the other its real type (num-complex 0.4's
 * `Complex<T> op T` and `T op Complex<T>` behind auto_impl/op_ternary_arithmetic.rs:3-14):
 * Complex op t = (re + t, im), (re - t, im), (re t, im t), (re / t, im / t);
 * t op Complex = (t + re, im), (t - re, 0 - im), (t re, t im), (t c / n, (-t) d / n), n = c c + d d.
 * lb == NULL: b is one value. */
static void mixed1(int op, int tc, int a_cplx, char* c, const char* a, const char* b) {
    if (tc == C128) {
        double re, im, t;
        double* o = (double*)c;
        if (a_cplx) { re = ((const double*)a)[0]; im = ((const double*)a)[1]; t = *(const double*)b; }
        else { re = ((const double*)b)[0]; im = ((const double*)b)[1]; t = *(const double*)a; }
        if (a_cplx) {
            if (op == OP_ADD) { o[0] = re + t; o[1] = im; }
            else if (op == OP_SUB) { o[0] = re - t; o[1] = im; }
            else if (op == OP_MUL) { o[0] = re * t; o[1] = im * t; }
            else { o[0] = re / t; o[1] = im / t; }
        } else {
            if (op == OP_ADD) { o[0] = t + re; o[1] = im; }
            else if (op == OP_SUB) { o[0] = t - re; o[1] = 0.0 - im; }
            else if (op == OP_MUL) { o[0] = t * re; o[1] = t * im; }
            else { double n = re * re + im * im; o[0] = t * re / n; o[1] = -t * im / n; }
        }
    } else {
        float re, im, t;
        float* o = (float*)c;
        if (a_cplx) { re = ((const float*)a)[0]; im = ((const float*)a)[1]; t = *(const float*)b; }
        else { re = ((const float*)b)[0]; im = ((const float*)b)[1]; t = *(const float*)a; }
        if (a_cplx) {
            if (op == OP_ADD) { o[0] = re + t; o[1] = im; }
            else if (op == OP_SUB) { o[0] = re - t; o[1] = im; }
            else if (op == OP_MUL) { o[0] = re * t; o[1] = im * t; }
            else { o[0] = re / t; o[1] = im / t; }
        } else {
            if (op == OP_ADD) { o[0] = t + re; o[1] = im; }
            else if (op == OP_SUB) { o[0] = t - re; o[1] = 0.0f - im; }
            else if (op == OP_MUL) { o[0] = t * re; o[1] = t * im; }
            else { float n = re * re + im * im; o[0] = t * re / n; o[1] = -t * im / n; }
        }
    }
}

int o_binary_mixed(int op, int tc, void* c, const olayout* lc, int ta, const void* a, const olayout* la, int tb,
                   const void* b, const olayout* lb) {
    size_t n = lsize(lc), ec = esize(tc), ea = esize(ta), eb = esize(tb);
    int a_cplx = ta == tc;
    if (n == 0) return 0;
    odo_t oc, oa, ob;
    odo_init(&oc, lc);
    odo_init(&oa, la);
    if (lb) odo_init(&ob, lb);
    for (size_t i = 0; i < n; ++i) {
        const char* pb = lb ? (const char*)b + ob.off * (ptrdiff_t)eb : (const char*)b;
        mixed1(op, tc, a_cplx, (char*)c + oc.off * (ptrdiff_t)ec, (const char*)a + oa.off * (ptrdiff_t)ea, pb);
        odo_next(&oc, lc);
        odo_next(&oa, la);
        if (lb) odo_next(&ob, lb);
    }
    return 0;
}

static void cast_elem(int tc, void* c, int ta, const void* a) {
    double re = 0, im = 0;
    int64_t iv = 0;
    int is_int = 0;
    switch (ta) {
        case F32: re = *(const float*)a; break;
        case F64: re = *(const double*)a; break;
        case I32: iv = *(const int32_t*)a; is_int = 1; re = (double)iv; break;
        case I64: iv = *(const int64_t*)a; is_int = 1; re = (double)iv; break;
        case C64: re = ((const float*)a)[0]; im = ((const float*)a)[1]; break;
        case C128: re = ((const double*)a)[0]; im = ((const double*)a)[1]; break;
    }
    switch (tc) {
        case F32: *(float*)c = (ta == F32) ? *(const float*)a : (is_int ? (float)iv : (float)re); break;
        case F64: *(double*)c = is_int ? (double)iv : re; break;
        case I32: *(int32_t*)c = is_int ? (int32_t)iv : (int32_t)re; break;
        case I64: *(int64_t*)c = is_int ? iv : (int64_t)re; break;
        case C64:
            ((float*)c)[0] = (ta == F32 || ta == C64) ? (ta == F32 ? *(const float*)a : ((const float*)a)[0]) : (float)re;
            ((float*)c)[1] = (ta == C64) ? ((const float*)a)[1] : (float)im;
            break;
        case C128: ((double*)c)[0] = re; ((double*)c)[1] = im; break;
    }
}

/* assign_arbitary_uninit_promote (cpu_serial/assignment.rs:6-67): copy in logical row-major
 * order of each operand's own shape (equal element counts). */
int o_assign(int tc, void* c, const olayout* lc, int ta, const void* a, const olayout* la) {
    size_t n = lsize(lc);
    if (n != lsize(la)) return 1;
    if (n == 0) return 0;
    size_t ec = esize(tc), ea = esize(ta);
    odo_t oc, oa;
    odo_init(&oc, lc);
    odo_init(&oa, la);
    for (size_t i = 0; i < n; ++i) {
        if (tc == ta) memcpy((char*)c + oc.off * (ptrdiff_t)ec, (const char*)a + oa.off * (ptrdiff_t)ea, ec);
        else cast_elem(tc, (char*)c + oc.off * (ptrdiff_t)ec, ta, (const char*)a + oa.off * (ptrdiff_t)ea);
        odo_next(&oc, lc);
        odo_next(&oa, la);
    }
    return 0;
}

/* ---------------- reductions ---------------- */

/* double-double accumulation (TwoSum); exact sum rounded once up to O(n u^2) */
typedef struct { double s, c; } dd_t;
static void dd_add(dd_t* a, double x) {
    double t = a->s + x;
    double bp = t - a->s;
    double e = (a->s - (t - bp)) + (x - bp);
    a->s = t;
    a->c += e;
}

/* Split la into (reduced, kept) by axes in the given order (indexer.rs:453-476). */
static int split_axes(const olayout* la, size_t naxes, const ptrdiff_t* axes, size_t* rs, ptrdiff_t* rt, size_t* nr,
                      size_t* ks, ptrdiff_t* kt, size_t* nk) {
    int used[MAXND] = {0};
    *nr = 0;
    *nk = 0;
    for (size_t i = 0; i < naxes; ++i) {
        ptrdiff_t x = axes[i] < 0 ? axes[i] + (ptrdiff_t)la->ndim : axes[i];
        if (x < 0 || x >= (ptrdiff_t)la->ndim || used[x]) return 1;
        used[x] = 1;
        rs[*nr] = la->shape[x];
        rt[*nr] = la->stride[x];
        (*nr)++;
    }
    for (size_t i = 0; i < la->ndim; ++i) {
        if (used[i]) continue;
        ks[*nk] = la->shape[i];
        kt[*nk] = la->stride[i];
        (*nk)++;
    }
    return 0;
}

static int is_nan_t(int t, const void* p) {
    if (t == F32) return isnan(*(const float*)p);
    if (t == F64) return isnan(*(const double*)p);
    return 0;
}
static double as_d(int t, const void* p) {
    switch (t) {
        case F32: return *(const float*)p;
        case F64: return *(const double*)p;
        case I32: return *(const int32_t*)p;
        case I64: return (double)*(const int64_t*)p;
    }
    return 0;
}
/* x > y (1), equal (0), less (-1), unordered (2) on the native type */
static int cmp_t(int t, const void* x, const void* y) {
    switch (t) {
        case F32: { float a = *(const float*)x, b = *(const float*)y; if (a > b) return 1; if (a == b) return 0; if (a < b) return -1; return 2; }
        case F64: { double a = *(const double*)x, b = *(const double*)y; if (a > b) return 1; if (a == b) return 0; if (a < b) return -1; return 2; }
        case I32: { int32_t a = *(const int32_t*)x, b = *(const int32_t*)y; return a > b ? 1 : a == b ? 0 : -1; }
        case I64: { int64_t a = *(const int64_t*)x, b = *(const int64_t*)y; return a > b ? 1 : a == b ? 0 : -1; }
    }
    return 2;
}

/* One reduction over a (reduced) layout `lr` with offset base; writes result to out (dtype of op).
 *  sum: T::zero + ... ; f32/f64 accumulate in double-double (the oracle's exact-ish sum)
 *  max/min: init T::MIN/T::MAX, IEEE maxNum (NaN ignored)  (ext_real.rs:70-87)
 *  arg*: first extreme in row-major order; NaN at position 0 wins, later NaN ignored
 *        (cpu_serial/reduction.rs:425-500); result raveled row-major over lr's shape. */
static void reduce_one(int op, int t, const void* a, const olayout* lr, void* out) {
    size_t n = lsize(lr), es = esize(t);
    odo_t o;
    odo_init(&o, lr);
    if (op == R_ALL || op == R_ANY) {
        /* all / any (feature_rayon/auto_impl/reduction.rs:466-519): fold && from true, || from false */
        int acc = op == R_ALL;
        for (size_t i = 0; i < n; ++i) {
            int x = *((const uint8_t*)a + o.off) != 0;
            acc = op == R_ALL ? (acc && x) : (acc || x);
            odo_next(&o, lr);
        }
        *(uint8_t*)out = (uint8_t)acc;
        return;
    }
    if (op == R_COUNT || (op == R_SUM && t == BOOL)) {
        /* count_nonzero (:521-551): elements != zero (complex: either part) into a usize; the sum
         * of a bool array (OpSumBoolAPI :678-715) counts the true elements */
        uint64_t cnt = 0;
        for (size_t i = 0; i < n; ++i) {
            const char* p = (const char*)a + o.off * (ptrdiff_t)es;
            int nz;
            switch (t) {
                case F32: nz = *(const float*)p != 0.0f; break;
                case F64: nz = *(const double*)p != 0.0; break;
                case I32: nz = *(const int32_t*)p != 0; break;
                case I64: nz = *(const int64_t*)p != 0; break;
                case C64: nz = ((const float*)p)[0] != 0.0f || ((const float*)p)[1] != 0.0f; break;
                case C128: nz = ((const double*)p)[0] != 0.0 || ((const double*)p)[1] != 0.0; break;
                default: nz = *(const uint8_t*)p != 0; break;
            }
            cnt += (uint64_t)nz;
            odo_next(&o, lr);
        }
        *(uint64_t*)out = cnt;
        return;
    }
    if (op == R_SUM) {
        if (t == F32 || t == F64) {
            dd_t acc = {0, 0};
            for (size_t i = 0; i < n; ++i) { dd_add(&acc, as_d(t, (const char*)a + o.off * (ptrdiff_t)es)); odo_next(&o, lr); }
            if (t == F32) *(float*)out = (float)(acc.s + acc.c);
            else *(double*)out = acc.s + acc.c;
        } else if (t == I32) {
            uint32_t s = 0;
            for (size_t i = 0; i < n; ++i) { s += (uint32_t) * (const int32_t*)((const char*)a + o.off * 4); odo_next(&o, lr); }
            *(int32_t*)out = (int32_t)s;
        } else if (t == I64) {
            uint64_t s = 0;
            for (size_t i = 0; i < n; ++i) { s += (uint64_t) * (const int64_t*)((const char*)a + o.off * 8); odo_next(&o, lr); }
            *(int64_t*)out = (int64_t)s;
        } else if (t == C128 || t == C64) {
            dd_t r = {0, 0}, im = {0, 0};
            for (size_t i = 0; i < n; ++i) {
                const char* p = (const char*)a + o.off * (ptrdiff_t)es;
                if (t == C128) { dd_add(&r, ((const double*)p)[0]); dd_add(&im, ((const double*)p)[1]); }
                else { dd_add(&r, ((const float*)p)[0]); dd_add(&im, ((const float*)p)[1]); }
                odo_next(&o, lr);
            }
            if (t == C128) { ((double*)out)[0] = r.s + r.c; ((double*)out)[1] = im.s + im.c; }
            else { ((float*)out)[0] = (float)(r.s + r.c); ((float*)out)[1] = (float)(im.s + im.c); }
        }
        return;
    }
    if (op == R_PROD) {
        /* prod (feature_rayon/auto_impl/reduction.rs:131-165): serial fold acc * x from one,
         * in the element type (complex (ac-bd, ad+bc) unfused); integers wrap */
        unsigned char acc[16] = {0};
        switch (t) {
            case F32: *(float*)acc = 1.0f; break;
            case F64: *(double*)acc = 1.0; break;
            case I32: *(int32_t*)acc = 1; break;
            case I64: *(int64_t*)acc = 1; break;
            case C64: ((float*)acc)[0] = 1.0f; break;
            case C128: ((double*)acc)[0] = 1.0; break;
        }
        for (size_t i = 0; i < n; ++i) {
            const void* p = (const char*)a + o.off * (ptrdiff_t)es;
            unsigned char r[16];
            if (t == I32) *(int32_t*)r = (int32_t)((uint32_t) * (int32_t*)acc * (uint32_t) * (const int32_t*)p);
            else if (t == I64) *(int64_t*)r = (int64_t)((uint64_t) * (int64_t*)acc * (uint64_t) * (const int64_t*)p);
            else apply(OP_MUL, t, r, acc, p);
            memcpy(acc, r, es);
            odo_next(&o, lr);
        }
        memcpy(out, acc, es);
        return;
    }
    if (op == R_MEAN || op == R_VAR || op == R_STD || op == R_L2) {
        /* mean :167-205 = sum / n; var :207-259 = sum|x|^2/n - |sum x/n|^2; std :261-315 = sqrt(var);
         * l2_norm :317-354 = sqrt(sum |x|^2). Sums in double-double, squares exact (fma). */
        dd_t sr = {0, 0}, si = {0, 0}, q = {0, 0};
        int cx = t == C64 || t == C128;
        for (size_t i = 0; i < n; ++i) {
            const char* p = (const char*)a + o.off * (ptrdiff_t)es;
            double re, im = 0.0;
            if (t == C128) { re = ((const double*)p)[0]; im = ((const double*)p)[1]; }
            else if (t == C64) { re = ((const float*)p)[0]; im = ((const float*)p)[1]; }
            else re = as_d(t, p);
            dd_add(&sr, re);
            dd_add(&si, im);
            double pr = re * re, pi = im * im;
            dd_add(&q, pr);
            q.c += fma(re, re, -pr);
            dd_add(&q, pi);
            q.c += fma(im, im, -pi);
            odo_next(&o, lr);
        }
        double dn = (double)n, qs = q.s + q.c, mr = (sr.s + sr.c) / dn, mi = (si.s + si.c) / dn;
        double r = 0.0;
        if (op == R_L2) r = sqrt(qs);
        else if (op == R_VAR) r = qs / dn - (mr * mr + mi * mi);
        else if (op == R_STD) r = sqrt(qs / dn - (mr * mr + mi * mi));
        if (op == R_MEAN) {
            if (t == F32) *(float*)out = (float)mr;
            else if (t == F64) *(double*)out = mr;
            else if (t == C64) { ((float*)out)[0] = (float)mr; ((float*)out)[1] = (float)mi; }
            else { ((double*)out)[0] = mr; ((double*)out)[1] = mi; }
        } else {
            if (t == F32 || t == C64) *(float*)out = (float)r;
            else *(double*)out = r;
        }
        (void)cx;
        return;
    }
    if (op == R_MAX || op == R_MIN) {
        int mx = op == R_MAX;
        unsigned char acc[16];
        switch (t) {
            case F32: *(float*)acc = mx ? -FLT_MAX : FLT_MAX; break;
            case F64: *(double*)acc = mx ? -DBL_MAX : DBL_MAX; break;
            case I32: *(int32_t*)acc = mx ? INT_MIN : INT_MAX; break;
            case I64: *(int64_t*)acc = mx ? LLONG_MIN : LLONG_MAX; break;
        }
        for (size_t i = 0; i < n; ++i) {
            const void* p = (const char*)a + o.off * (ptrdiff_t)es;
            int c = cmp_t(t, p, acc);
            if ((mx && c == 1) || (!mx && c == -1)) memcpy(acc, p, es);
            odo_next(&o, lr);
        }
        memcpy(out, acc, es);
        return;
    }
    /* arg */
    int mx = op == R_ARGMAX;
    unsigned char best[16];
    uint64_t bi = 0;
    for (size_t i = 0; i < n; ++i) {
        const void* p = (const char*)a + o.off * (ptrdiff_t)es;
        if (i == 0) {
            memcpy(best, p, es);
            bi = 0;
        } else {
            int c = cmp_t(t, p, best);
            if ((mx && c == 1) || (!mx && c == -1)) { memcpy(best, p, es); bi = i; }
            /* equal: keep the smaller (earlier) index; unordered: keep */
        }
        odo_next(&o, lr);
    }
    *(uint64_t*)out = bi;
}

/* reduce_axes_cpu_serial / reduce_axes_arg_cpu_serial: one output per kept position, written
 * at lo (the output layout, same shape as the kept axes). axes in the given order. */
int o_reduce_axes(int op, int t, const void* a, const olayout* la, size_t naxes, const ptrdiff_t* axes, void* out,
                  const olayout* lo) {
    size_t rs[MAXND], ks[MAXND], nr, nk;
    ptrdiff_t rt[MAXND], kt[MAXND];
    if (split_axes(la, naxes, axes, rs, rt, &nr, ks, kt, &nk)) return 1;
    olayout lk = {nk, ks, kt, la->offset};
    size_t m = lsize(&lk);
    size_t eo = (op == R_ARGMAX || op == R_ARGMIN) ? 8 : esize(t);
    if ((op == R_VAR || op == R_STD || op == R_L2) && (t == C64 || t == C128)) eo = esize(t) / 2;
    if (op == R_ALL || op == R_ANY) eo = 1;
    if (op == R_COUNT || (op == R_SUM && t == BOOL)) eo = 8;
    olayout lr = {nr, rs, rt, 0};
    if (lsize(&lr) == 0 && (op == R_MAX || op == R_MIN || op == R_ARGMAX || op == R_ARGMIN)) return 2;
    odo_t ok, oo;
    odo_init(&ok, &lk);
    odo_init(&oo, lo);
    for (size_t i = 0; i < m; ++i) {
        lr.offset = (size_t)ok.off;
        reduce_one(op, t, a, &lr, (char*)out + oo.off * (ptrdiff_t)eo);
        odo_next(&ok, &lk);
        odo_next(&oo, lo);
    }
    return 0;
}

int o_reduce_all(int op, int t, const void* a, const olayout* la, void* out) {
    if (lsize(la) == 0 && (op == R_MAX || op == R_MIN || op == R_ARGMAX || op == R_ARGMIN)) return 2;
    reduce_one(op, t, a, la, out);
    return 0;
}

/* allclose_all (feature_rayon/auto_impl/reduction.rs:725-750): every element pair isclose
 * (rstsr-dtype-traits/src/isclose.rs:92-106): diff = a.ext_abs_diff(b) and |b| = b.ext_abs() in
 * the element type, both then cast to f64; close when diff <= atol + rtol * |b|, or when
 * equal_nan and both are NaN. ext_abs_diff (ext_num.rs:76-214): floats |a - b| in the float
 * type; signed integers (a >= b ? a - b : b - a) in the integer type (wrapping, as Rust release
 * builds); complex |a - b| (hypot of the parts); ext_abs: floats fabs, integers a.abs()
 * (wrapping: MIN stays MIN), complex hypot. A NaN diff (inf - inf) compares false. la and lb
 * already broadcast to one shape. Returns 1 (close), 0 (not close), -1 (zero-size: an error in
 * the reference). */
static int o_isclose1(int t, const char* pa, const char* pb, double rtol, double atol, int equal_nan) {
    double diff = 0.0, ab = 0.0;
    int nan_a = 0, nan_b = 0;
    switch (t) {
        case F64: {
            double x = *(const double*)pa, y = *(const double*)pb;
            diff = fabs(x - y); ab = fabs(y); nan_a = isnan(x); nan_b = isnan(y);
            break;
        }
        case F32: {
            float x = *(const float*)pa, y = *(const float*)pb;
            diff = (double)fabsf(x - y); ab = (double)fabsf(y); nan_a = isnan(x); nan_b = isnan(y);
            break;
        }
        case I32: {
            int32_t x = *(const int32_t*)pa, y = *(const int32_t*)pb;
            int32_t d = x >= y ? (int32_t)((uint32_t)x - (uint32_t)y) : (int32_t)((uint32_t)y - (uint32_t)x);
            int32_t m = y < 0 ? (int32_t)(0u - (uint32_t)y) : y;
            diff = (double)d; ab = (double)m;
            break;
        }
        case I64: {
            int64_t x = *(const int64_t*)pa, y = *(const int64_t*)pb;
            int64_t d = x >= y ? (int64_t)((uint64_t)x - (uint64_t)y) : (int64_t)((uint64_t)y - (uint64_t)x);
            int64_t m = y < 0 ? (int64_t)(0ull - (uint64_t)y) : y;
            diff = (double)d; ab = (double)m;
            break;
        }
        case C128: {
            const double* x = (const double*)pa; const double* y = (const double*)pb;
            diff = hypot(x[0] - y[0], x[1] - y[1]); ab = hypot(y[0], y[1]);
            nan_a = isnan(x[0]) || isnan(x[1]); nan_b = isnan(y[0]) || isnan(y[1]);
            break;
        }
        case C64: {
            const float* x = (const float*)pa; const float* y = (const float*)pb;
            diff = (double)hypotf(x[0] - y[0], x[1] - y[1]); ab = (double)hypotf(y[0], y[1]);
            nan_a = isnan(x[0]) || isnan(x[1]); nan_b = isnan(y[0]) || isnan(y[1]);
            break;
        }
    }
    int comp = diff <= atol + rtol * ab;
    return comp || (equal_nan && nan_a && nan_b);
}

int o_allclose(int t, const void* a, const olayout* la, const void* b, const olayout* lb, double rtol, double atol,
               int equal_nan) {
    size_t n = lsize(la), es = esize(t);
    if (n == 0 || lsize(lb) == 0) return -1;
    odo_t oa, ob;
    odo_init(&oa, la);
    odo_init(&ob, lb);
    int all = 1;
    for (size_t i = 0; i < n; ++i) {
        all = all && o_isclose1(t, (const char*)a + oa.off * (ptrdiff_t)es, (const char*)b + ob.off * (ptrdiff_t)es,
                                rtol, atol, equal_nan);
        odo_next(&oa, la);
        odo_next(&ob, lb);
    }
    return all;
}

/* gemm_naive_cpu_serial (cpu_serial/matmul_naive.rs:78-122): c = beta*c, then for k in order
 * c += alpha*(a*b). Strided 2-D operands (rs, cs) in elements. f64/f32/i32/i64/c128/c64. */
int o_gemm_naive(int t, size_t m, size_t n, size_t k, const void* alpha, const void* a, ptrdiff_t rsa, ptrdiff_t csa,
                 const void* b, ptrdiff_t rsb, ptrdiff_t csb, const void* beta, void* c, ptrdiff_t rsc, ptrdiff_t csc) {
    size_t es = esize(t);
    for (size_t i = 0; i < m; ++i) {
        for (size_t j = 0; j < n; ++j) {
            char* pc = (char*)c + ((ptrdiff_t)i * rsc + (ptrdiff_t)j * csc) * (ptrdiff_t)es;
            unsigned char tmp[16];
            apply(OP_MUL, t, tmp, beta, pc);
            memcpy(pc, tmp, es);
        }
        for (size_t q = 0; q < k; ++q) {
            const char* pa = (const char*)a + ((ptrdiff_t)i * rsa + (ptrdiff_t)q * csa) * (ptrdiff_t)es;
            for (size_t j = 0; j < n; ++j) {
                const char* pb = (const char*)b + ((ptrdiff_t)q * rsb + (ptrdiff_t)j * csb) * (ptrdiff_t)es;
                char* pc = (char*)c + ((ptrdiff_t)i * rsc + (ptrdiff_t)j * csc) * (ptrdiff_t)es;
                unsigned char ab[16], aab[16], s[16];
                apply(OP_MUL, t, ab, pa, pb);
                apply(OP_MUL, t, aab, alpha, ab);
                apply(OP_ADD, t, s, aab, pc);
                memcpy(pc, s, es);
            }
        }
    }
    return 0;
}

/* Accurate f64 GEMM (the error oracle): each c[i,j] = sum_k a[i,k] b[k,j] in Dot2 (TwoProduct by
 * Dekker splitting + TwoSum, Ogita-Rump-Oishi 2005), parallel over rows. Accurate to about
 * u * |result| + u^2 * sum|a||b|. */
static void two_prod(double a, double b, double* p, double* e) {
    const double split = 134217729.0; /* 2^27 + 1 */
    double x = a * b;
    double ca = split * a, ah = ca - (ca - a), al = a - ah;
    double cb = split * b, bh = cb - (cb - b), bl = b - bh;
    *p = x;
    *e = al * bl - (((x - ah * bh) - al * bh) - ah * bl);
}
int o_gemm_dot2_f64(size_t m, size_t n, size_t k, const double* a, ptrdiff_t rsa, ptrdiff_t csa, const double* b,
                    ptrdiff_t rsb, ptrdiff_t csb, double* c, ptrdiff_t rsc, ptrdiff_t csc, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (ptrdiff_t i = 0; i < (ptrdiff_t)m; ++i) {
        for (size_t j = 0; j < n; ++j) {
            double s = 0, cc = 0;
            for (size_t q = 0; q < k; ++q) {
                double p, e;
                two_prod(a[i * rsa + (ptrdiff_t)q * csa], b[(ptrdiff_t)q * rsb + (ptrdiff_t)j * csb], &p, &e);
                double t = s + p, bp = t - s, err = (s - (t - bp)) + (p - bp);
                s = t;
                cc += err + e;
            }
            c[i * rsc + (ptrdiff_t)j * csc] = s + cc;
        }
    }
    return 0;
}

/* ================= OpenMP baseline (cpu_rayon control flow) ================= */

/* greedy iteration order of one layout (rearrangement.rs:36-113, keep_shape=true): axes sorted
 * with broadcast/size-1 first, then by |stride| ascending; returns the permutation. */
static void greedy_perm(const olayout* l, size_t* perm) {
    size_t nd = l->ndim;
    for (size_t i = 0; i < nd; ++i) perm[i] = i;
    for (size_t i = 1; i < nd; ++i) { /* stable insertion sort */
        size_t x = perm[i];
        size_t j = i;
        while (j > 0) {
            size_t y = perm[j - 1];
            int bx = l->shape[x] == 1 || l->stride[x] == 0, by = l->shape[y] == 1 || l->stride[y] == 0;
            int less;
            if (bx && by) less = x < y;
            else if (bx && !by) less = 1;
            else if (!bx && by) less = 0;
            else less = labs(l->stride[x]) < labs(l->stride[y]);
            if (!less) break;
            perm[j] = y;
            --j;
        }
        perm[j] = x;
    }
}

/* Collapse up to three operands in column-major greedy order: returns inner contiguous size
 * (translate_to_col_major_with_contig, rearrangement.rs:297-321) and outer dims. */
typedef struct {
    size_t nd;
    size_t shape[MAXND];
    ptrdiff_t st[3][MAXND];
    ptrdiff_t off[3];
    size_t contig;  /* size of the common f-contiguous prefix (0: none) */
} bplan_t;

static void bplan(bplan_t* p, int nops, const olayout* const* ls, int ref) {
    size_t perm[MAXND];
    greedy_perm(ls[ref], perm);
    size_t nd = ls[0]->ndim;
    p->nd = 0;
    for (int k = 0; k < 3; ++k) p->off[k] = k < nops ? (ptrdiff_t)ls[k]->offset : 0;
    /* f-contiguous prefix common to all operands, in greedy order */
    size_t acc[3] = {1, 1, 1};
    size_t i = 0, contig = 1;
    int any = 0;
    for (; i < nd; ++i) {
        size_t ax = perm[i];
        int ok = 1;
        for (int k = 0; k < nops; ++k)
            if (ls[0]->shape[ax] != 1 && ls[k]->stride[ax] != (ptrdiff_t)acc[k]) ok = 0;
        if (!ok) break;
        for (int k = 0; k < nops; ++k) acc[k] *= ls[0]->shape[ax];
        contig *= ls[0]->shape[ax];
        any = 1;
    }
    p->contig = any ? contig : 0;
    for (; i < nd; ++i) {
        size_t ax = perm[i];
        p->shape[p->nd] = ls[0]->shape[ax];
        for (int k = 0; k < 3; ++k) p->st[k][p->nd] = k < nops ? ls[k]->stride[ax] : 0;
        p->nd++;
    }
}

static size_t prod_shape(const bplan_t* p) {
    size_t n = 1;
    for (size_t i = 0; i < p->nd; ++i) n *= p->shape[i];
    return n;
}

/* outer position -> offsets (column-major over p->shape) */
static void bunravel(const bplan_t* p, size_t q, ptrdiff_t* o) {
    for (int k = 0; k < 3; ++k) o[k] = p->off[k];
    for (size_t d = 0; d < p->nd; ++d) {
        size_t r = q % p->shape[d];
        q /= p->shape[d];
        for (int k = 0; k < 3; ++k) o[k] += (ptrdiff_t)r * p->st[k][d];
    }
}

/* op_mutc_refa_refb_func_cpu_rayon (cpu_rayon/op_with_func.rs:13-80): K order from the largest
 * non-broadcast operand (last of equals), contiguous inner loop when the prefix is >= 16. */
int b_binary(int op, int t, void* c, const olayout* lc, const void* a, const olayout* la, const void* b,
             const olayout* lb, int nthreads) {
    const olayout* ls[3] = {lc, la, lb};
    size_t sz[3];
    for (int k = 0; k < 3; ++k) {
        size_t s = 1;
        for (size_t i = 0; i < ls[k]->ndim; ++i)
            if (ls[k]->stride[i] != 0) s *= ls[k]->shape[i];
        sz[k] = s;
    }
    int ref = 0;
    if (!(sz[0] == sz[1] && sz[1] == sz[2])) {
        size_t best = 0;
        for (int k = 0; k < 3; ++k)
            if (sz[k] >= best) { best = sz[k]; ref = k; }
    }
    bplan_t p;
    bplan(&p, 3, ls, ref);
    size_t es = esize(t);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    size_t contig = p.contig >= 16 ? p.contig : 0;
    if (contig == 0 && p.contig > 0) {
        /* short prefix: fold it back into the outer loop as one axis of unit stride */
        memmove(&p.shape[1], &p.shape[0], p.nd * sizeof(size_t));
        for (int k = 0; k < 3; ++k) memmove(&p.st[k][1], &p.st[k][0], p.nd * sizeof(ptrdiff_t));
        p.shape[0] = p.contig;
        for (int k = 0; k < 3; ++k) p.st[k][0] = 1;
        p.nd++;
    }
    size_t outer = prod_shape(&p);
    size_t inner = contig ? contig : 1;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t q = 0; q < (ptrdiff_t)outer; ++q) {
        ptrdiff_t o[3];
        bunravel(&p, (size_t)q, o);
        char* pc = (char*)c + o[0] * (ptrdiff_t)es;
        const char* pa = (const char*)a + o[1] * (ptrdiff_t)es;
        const char* pb = (const char*)b + o[2] * (ptrdiff_t)es;
        if (t == F64 && op == OP_MUL) {
            double* x = (double*)pc;
            const double *y = (const double*)pa, *z = (const double*)pb;
            for (size_t i = 0; i < inner; ++i) x[i] = y[i] * z[i];
        } else if (t == F64 && op == OP_ADD) {
            double* x = (double*)pc;
            const double *y = (const double*)pa, *z = (const double*)pb;
            for (size_t i = 0; i < inner; ++i) x[i] = y[i] + z[i];
        } else {
            for (size_t i = 0; i < inner; ++i) apply(op, t, pc + i * es, pa + i * es, pb + i * es);
        }
    }
    return 0;
}

/* assign_arbitary (cpu_rayon/assignment.rs:14-93), same-shape case: both sides iterated in
 * logical C order, parallel over outer rows. */
int b_assign(int t, void* c, const olayout* lc, const void* a, const olayout* la, int nthreads) {
    size_t nd = lc->ndim, es = esize(t);
    if (nd == 0) { memcpy((char*)c + lc->offset * es, (const char*)a + la->offset * es, es); return 0; }
    size_t inner = lc->shape[nd - 1];
    size_t outer = lsize(lc) / (inner ? inner : 1);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t q = 0; q < (ptrdiff_t)outer; ++q) {
        ptrdiff_t oc = (ptrdiff_t)lc->offset, oa = (ptrdiff_t)la->offset;
        size_t r = (size_t)q;
        for (size_t d = nd - 1; d-- > 0;) {
            size_t x = r % lc->shape[d];
            r /= lc->shape[d];
            oc += (ptrdiff_t)x * lc->stride[d];
            oa += (ptrdiff_t)x * la->stride[d];
        }
        ptrdiff_t sc = lc->stride[nd - 1], sa = la->stride[nd - 1];
        if (es == 8) {
            uint64_t* pc = (uint64_t*)c + oc;
            const uint64_t* pa = (const uint64_t*)a + oa;
            for (size_t i = 0; i < inner; ++i) pc[(ptrdiff_t)i * sc] = pa[(ptrdiff_t)i * sa];
        } else {
            for (size_t i = 0; i < inner; ++i)
                memcpy((char*)c + (oc + (ptrdiff_t)i * sc) * (ptrdiff_t)es, (const char*)a + (oa + (ptrdiff_t)i * sa) * (ptrdiff_t)es, es);
        }
    }
    return 0;
}

/* unrolled_reduce (cpu_serial/reduction.rs:44-83): 8 partial sums, combined (p0+p4)...(p3+p7) */
static double unrolled_sum_d(const double* x, size_t n, ptrdiff_t s) {
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0}, acc = 0;
    size_t i = 0;
    if (s == 1) {
        for (; i + 8 <= n; i += 8)
            for (int k = 0; k < 8; ++k) p[k] += x[i + k];
    }
    acc = acc + (p[0] + p[4]);
    acc = acc + (p[1] + p[5]);
    acc = acc + (p[2] + p[6]);
    acc = acc + (p[3] + p[7]);
    for (; i < n; ++i) acc += x[(ptrdiff_t)i * s];
    return acc;
}

/* reduce_axes_cpu_rayon (cpu_rayon/reduction.rs:109-328), f64 sum/max and argmax, parallel over
 * outputs; branch (i) unrolled_reduce when the reduced set has a unit-stride run, branch (ii)
 * chunked vector accumulate when the kept set does. One baseline for the timed comparison. */
int b_reduce_axes(int op, int t, const void* a, const olayout* la, size_t naxes, const ptrdiff_t* axes, void* out,
                  const olayout* lo, int nthreads) {
    size_t rs[MAXND], ks[MAXND], nr, nk;
    ptrdiff_t rt[MAXND], kt[MAXND];
    if (split_axes(la, naxes, axes, rs, rt, &nr, ks, kt, &nk)) return 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    olayout lk = {nk, ks, kt, la->offset};
    size_t m = lsize(&lk);
    /* unit-stride reduced axis? (branch i) */
    int ru = -1, ku = -1;
    for (size_t i = 0; i < nr; ++i) if (rt[i] == 1) ru = (int)i;
    for (size_t i = 0; i < nk; ++i) if (kt[i] == 1) ku = (int)i;
    size_t eo = (op == R_ARGMAX || op == R_ARGMIN) ? 8 : esize(t);
    if (ru >= 0 || ku < 0 || op != R_SUM || t != F64) {
        /* per-output fold; the unit-stride reduced axis (if any) innermost */
        size_t rr[MAXND];
        ptrdiff_t rtt[MAXND];
        size_t n2 = 0;
        for (size_t i = 0; i < nr; ++i) if ((int)i != ru) { rr[n2] = rs[i]; rtt[n2] = rt[i]; n2++; }
        size_t inner = ru >= 0 ? rs[ru] : 1;
#pragma omp parallel for schedule(static)
        for (ptrdiff_t q = 0; q < (ptrdiff_t)m; ++q) {
            ptrdiff_t oi = (ptrdiff_t)la->offset, oo = (ptrdiff_t)lo->offset;
            size_t r = (size_t)q;
            for (size_t d = nk; d-- > 0;) {
                size_t x = r % ks[d];
                r /= ks[d];
                oi += (ptrdiff_t)x * kt[d];
                oo += (ptrdiff_t)x * lo->stride[d];
            }
            if (op == R_SUM && t == F64 && ru >= 0) {
                olayout lr2 = {n2, rr, rtt, (size_t)oi};
                size_t nouter = lsize(&lr2);
                double acc = 0;
                odo_t o;
                odo_init(&o, &lr2);
                for (size_t u = 0; u < nouter; ++u) {
                    acc += unrolled_sum_d((const double*)a + o.off, inner, 1);
                    odo_next(&o, &lr2);
                }
                *(double*)((char*)out + oo * 8) = acc;
            } else {
                olayout lr = {nr, rs, rt, (size_t)oi};
                reduce_one(op, t, a, &lr, (char*)out + oo * (ptrdiff_t)eo);
            }
        }
        return 0;
    }
    /* branch (ii): kept unit-stride axis; vector accumulate over reduced positions in chunks */
    size_t kk[MAXND];
    ptrdiff_t ktt[MAXND], kot[MAXND];
    size_t n3 = 0;
    for (size_t i = 0; i < nk; ++i) if ((int)i != ku) { kk[n3] = ks[i]; ktt[n3] = kt[i]; kot[n3] = lo->stride[i]; n3++; }
    size_t vlen = ks[ku];
    ptrdiff_t vout = lo->stride[ku];
    size_t nout = 1;
    for (size_t i = 0; i < n3; ++i) nout *= kk[i];
    size_t nchunk = (vlen + 63) / 64;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t q = 0; q < (ptrdiff_t)(nout * nchunk); ++q) {
        size_t ch = (size_t)q % nchunk, r = (size_t)q / nchunk;
        ptrdiff_t oi = (ptrdiff_t)la->offset, oo = (ptrdiff_t)lo->offset;
        for (size_t d = n3; d-- > 0;) {
            size_t x = r % kk[d];
            r /= kk[d];
            oi += (ptrdiff_t)x * ktt[d];
            oo += (ptrdiff_t)x * kot[d];
        }
        size_t c0 = ch * 64, c1 = c0 + 64 < vlen ? c0 + 64 : vlen;
        double acc[64];
        for (size_t i = 0; i < 64; ++i) acc[i] = 0;
        olayout lr = {nr, rs, rt, (size_t)oi};
        odo_t o;
        odo_init(&o, &lr);
        size_t nred = lsize(&lr);
        for (size_t u = 0; u < nred; ++u) {
            const double* p = (const double*)a + o.off;
            for (size_t i = c0; i < c1; ++i) acc[i - c0] += p[i];
            odo_next(&o, &lr);
        }
        for (size_t i = c0; i < c1; ++i) *(double*)((char*)out + (oo + (ptrdiff_t)i * vout) * 8) = acc[i - c0];
    }
    return 0;
}

int o_max_threads(void) { return omp_get_max_threads(); }
