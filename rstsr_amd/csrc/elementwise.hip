// Strided n-d elementwise operations for gfx950: binary arithmetic, cast/copy,
// fill and allclose, over layouts already broadcast to one shape. The kernels
// and the planner are in ew.hpp.
//
// Replaces the layout-iterator loops of rstsr-native-impl/src/cpu_rayon/
// op_with_func.rs:13-390 (binary ops), assignment.rs:14-225 (assign, fill) and
// reduction.rs:334 (allclose).
// IEEE add/sub/mul/div are single correctly rounded operations; complex
// multiply is (ac - bd, ad + bc) as num-complex computes it. This file is built
// with -ffp-contract=off so that no product is fused into an FMA.
#include "ew.hpp"

#include <array>
#include <type_traits>

namespace rstsr {
namespace hip {

/* =============================== element ops =============================== */

template <class T>
__device__ __forceinline__ T op_add(T a, T b) { return a + b; }
template <class T>
__device__ __forceinline__ T op_sub(T a, T b) { return a - b; }
template <class T>
__device__ __forceinline__ T op_mul(T a, T b) { return a * b; }
template <class T>
__device__ __forceinline__ T op_div(T a, T b) { return a / b; }

// integers wrap (Rust release semantics); computed on the unsigned type
template <>
__device__ __forceinline__ int32_t op_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
template <>
__device__ __forceinline__ int32_t op_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
template <>
__device__ __forceinline__ int32_t op_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
template <>
__device__ __forceinline__ int64_t op_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
template <>
__device__ __forceinline__ int64_t op_sub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
template <>
__device__ __forceinline__ int64_t op_mul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }

// num-complex: Add/Sub componentwise; Mul (re*re' - im*im', re*im' + im*re')
template <>
__device__ __forceinline__ c128 op_add(c128 a, c128 b) { return {a.re + b.re, a.im + b.im}; }
template <>
__device__ __forceinline__ c128 op_sub(c128 a, c128 b) { return {a.re - b.re, a.im - b.im}; }
template <>
__device__ __forceinline__ c128 op_mul(c128 a, c128 b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <>
__device__ __forceinline__ c64 op_add(c64 a, c64 b) { return {a.re + b.re, a.im + b.im}; }
template <>
__device__ __forceinline__ c64 op_sub(c64 a, c64 b) { return {a.re - b.re, a.im - b.im}; }
template <>
__device__ __forceinline__ c64 op_mul(c64 a, c64 b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// num-complex Div: ((a.re c + a.im d) / n, (a.im c - a.re d) / n) with n = c*c + d*d
template <>
__device__ __forceinline__ c128 op_div(c128 a, c128 b) {
    double n = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
}
template <>
__device__ __forceinline__ c64 op_div(c64 a, c64 b) {
    float n = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
}

// Integer / and % truncate toward zero, as Rust's. Rust panics on a zero divisor and on
// MIN / -1 (MIN % -1); a kernel cannot stop there, so the device gives 0 for a zero divisor,
// the wrapped MIN for MIN / -1 and 0 for MIN % -1.
template <class T>
__device__ __forceinline__ T int_div(T a, T b) {
    using U = typename std::make_unsigned<T>::type;
    if (b == 0) return 0;
    if (b == -1) return (T)((U)0 - (U)a);
    return a / b;
}
template <class T>
__device__ __forceinline__ T int_rem(T a, T b) {
    return (b == 0 || b == -1) ? (T)0 : (T)(a % b);
}

// Rem: integers as above; floats fmod (Rust's f64 % f64); complex as num-complex:
// q = a / b, q0 = (q.re - q.re % 1, q.im - q.im % 1), a - b * q0
template <class T>
__device__ __forceinline__ T op_rem(T a, T b) {
    if constexpr (std::is_same<T, double>::value) return fmod(a, b);
    else if constexpr (std::is_same<T, float>::value) return fmodf(a, b);
    else if constexpr (std::is_same<T, c128>::value) {
        c128 q = op_div<c128>(a, b);
        c128 q0{q.re - fmod(q.re, 1.0), q.im - fmod(q.im, 1.0)};
        return op_sub<c128>(a, op_mul<c128>(b, q0));
    } else if constexpr (std::is_same<T, c64>::value) {
        c64 q = op_div<c64>(a, b);
        c64 q0{q.re - fmodf(q.re, 1.0f), q.im - fmodf(q.im, 1.0f)};
        return op_sub<c64>(a, op_mul<c64>(b, q0));
    } else {
        return int_rem<T>(a, b);
    }
}

// Shifts as Rust release builds: the amount is taken modulo the bit width; >> of a signed
// integer is arithmetic. Bit ops on bool are logical.
template <class T>
__device__ __forceinline__ T op_shl(T a, T b) {
    using U = typename std::make_unsigned<T>::type;
    return (T)((U)a << ((U)b & (U)(8 * sizeof(T) - 1)));
}
template <class T>
__device__ __forceinline__ T op_shr(T a, T b) {
    using U = typename std::make_unsigned<T>::type;
    return (T)(a >> ((U)b & (U)(8 * sizeof(T) - 1)));
}

template <int OP, class T>
__device__ __forceinline__ T apply_op(T a, T b) {
    if constexpr (OP == RSTSR_OP_ADD) return op_add<T>(a, b);
    if constexpr (OP == RSTSR_OP_SUB) return op_sub<T>(a, b);
    if constexpr (OP == RSTSR_OP_MUL) return op_mul<T>(a, b);
    if constexpr (OP == RSTSR_OP_DIV) {
        if constexpr (std::is_integral<T>::value) return int_div<T>(a, b);
        else return op_div<T>(a, b);
    }
    if constexpr (OP == RSTSR_OP_REM) return op_rem<T>(a, b);
    if constexpr (OP == RSTSR_OP_BITAND) return (T)(a & b);
    if constexpr (OP == RSTSR_OP_BITOR) return (T)(a | b);
    if constexpr (OP == RSTSR_OP_BITXOR) return (T)(a ^ b);
    if constexpr (OP == RSTSR_OP_SHL) return op_shl<T>(a, b);
    if constexpr (OP == RSTSR_OP_SHR) return op_shr<T>(a, b);
}

// Which (op, element type) pairs exist in the reference (the std / num-complex operator impls
// behind auto_impl/op_ternary_arithmetic.rs:3-14): arithmetic and Rem for every number type,
// bit ops for integers and bool, shifts for integers.
template <int OP, class T>
constexpr bool op_defined() {
    constexpr bool num = !std::is_same<T, bool>::value;
    constexpr bool integ = std::is_integral<T>::value && num;
    if constexpr (OP <= RSTSR_OP_REM) return num;
    else if constexpr (OP <= RSTSR_OP_BITXOR) return std::is_integral<T>::value;
    else return integ;
}

// num-complex 0.4's operators between a complex number and its real type:
//   Complex op t: (re + t, im), (re - t, im), (re t, im t), (re / t, im / t);
//   t op Complex: (t + re, im), (t - re, 0 - im), (t re, t im), (t c / n, (-t) d / n), n = c c + d d.
template <int OP, class C, class R>
__device__ __forceinline__ C mix_cr(C a, R t) {
    if constexpr (OP == RSTSR_OP_ADD) return C{a.re + t, a.im};
    if constexpr (OP == RSTSR_OP_SUB) return C{a.re - t, a.im};
    if constexpr (OP == RSTSR_OP_MUL) return C{a.re * t, a.im * t};
    if constexpr (OP == RSTSR_OP_DIV) return C{a.re / t, a.im / t};
}
template <int OP, class C, class R>
__device__ __forceinline__ C mix_rc(R t, C b) {
    if constexpr (OP == RSTSR_OP_ADD) return C{t + b.re, b.im};
    if constexpr (OP == RSTSR_OP_SUB) return C{t - b.re, (R)0 - b.im};
    if constexpr (OP == RSTSR_OP_MUL) return C{t * b.re, t * b.im};
    if constexpr (OP == RSTSR_OP_DIV) {
        R n = b.re * b.re + b.im * b.im;
        return C{t * b.re / n, -t * b.im / n};
    }
}

// value conversion for assign with promotion (Rust `as`-free DTypeCastAPI subset)
template <class TC, class TA>
__device__ __forceinline__ TC cast_to(TA a) {
    if constexpr (std::is_same<TC, TA>::value) return a;
    else if constexpr (std::is_same<TC, c128>::value) {
        if constexpr (std::is_same<TA, c64>::value) return c128{(double)a.re, (double)a.im};
        else return c128{(double)a, 0.0};
    } else if constexpr (std::is_same<TC, c64>::value) {
        if constexpr (std::is_same<TA, c128>::value) return c64{(float)a.re, (float)a.im};
        else return c64{(float)a, 0.0f};
    } else {
        return (TC)a;
    }
}

/* =============================== bodies =============================== */
// A body loads operand A and B at element offsets and stores C; kernels only do
// index arithmetic. TA/TB = None marks an absent operand.

template <int OP, class T, bool AS, bool BS>
struct BinBody {
    using TA = T;
    using TB = T;
    T* c;
    const T* a;
    const T* b;
    T sa, sb;
    static constexpr int OPS = (AS ? 0 : 1) | (BS ? 0 : 2);
    static constexpr size_t OUT_SIZE = sizeof(T);
    static constexpr bool NT_JOINT = false;
    template <bool NT = false>
    __device__ __forceinline__ T la(int64_t o) const { return AS ? sa : ld<NT>(a + o); }
    template <bool NT = false>
    __device__ __forceinline__ T lb(int64_t o) const { return BS ? sb : ld<NT>(b + o); }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t o, T x, T y) const { sto<NT>(c + o, apply_op<OP, T>(x, y)); }
};

// c = a op b with one operand complex (C) and the other its real type; AS / BS: host scalar.
template <int OP, class C, class TAv, class TBv, bool AS, bool BS>
struct MixBody {
    using TA = TAv;
    using TB = TBv;
    C* c;
    const TAv* a;
    const TBv* b;
    TAv sa;
    TBv sb;
    static constexpr int OPS = (AS ? 0 : 1) | (BS ? 0 : 2);
    static constexpr size_t OUT_SIZE = sizeof(C);
    static constexpr bool NT_JOINT = false;
    template <bool NT = false>
    __device__ __forceinline__ TAv la(int64_t o) const { return AS ? sa : ld<NT>(a + o); }
    template <bool NT = false>
    __device__ __forceinline__ TBv lb(int64_t o) const { return BS ? sb : ld<NT>(b + o); }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t o, TAv x, TBv y) const {
        if constexpr (std::is_same<TAv, C>::value) sto<NT>(c + o, mix_cr<OP, C>(x, y));
        else sto<NT>(c + o, mix_rc<OP, C>(x, y));
    }
};

template <class TCv, class TAv>
struct CastBody {
    using TA = TAv;
    using TB = None;
    TCv* c;
    const TAv* a;
    static constexpr int OPS = 1;
    static constexpr size_t OUT_SIZE = sizeof(TCv);
    static constexpr bool NT_JOINT = false;
    template <bool NT = false>
    __device__ __forceinline__ TAv la(int64_t o) const { return ld<NT>(a + o); }
    template <bool NT = false>
    __device__ __forceinline__ None lb(int64_t) const { return None{}; }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t o, TAv x, None) const { sto<NT>(c + o, cast_to<TCv, TAv>(x)); }
};

template <class T>
struct FillBody {
    using TA = None;
    using TB = None;
    T* c;
    T v;
    static constexpr int OPS = 0;
    static constexpr size_t OUT_SIZE = sizeof(T);
    static constexpr bool NT_JOINT = false;
    template <bool NT = false>
    __device__ __forceinline__ None la(int64_t) const { return None{}; }
    template <bool NT = false>
    __device__ __forceinline__ None lb(int64_t) const { return None{}; }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t o, None, None) const { sto<NT>(c + o, v); }
};

// ExtNum::ext_abs_diff / ext_abs (rstsr-dtype-traits/src/ext_num.rs:72-214) in the element
// type, then as f64: floats |a - b|; signed integers (a >= b ? a - b : b - a) and a.abs()
// wrapping like Rust release builds (MIN.abs() stays MIN); complex the norm (hypot).
template <class T>
__device__ __forceinline__ double abs_diff_d(T a, T b) {
    if constexpr (std::is_same<T, c128>::value) return hypot(a.re - b.re, a.im - b.im);
    else if constexpr (std::is_same<T, c64>::value) return (double)hypotf(a.re - b.re, a.im - b.im);
    else if constexpr (std::is_integral<T>::value) {
        using U = typename std::make_unsigned<T>::type;
        return (double)(a >= b ? (T)((U)a - (U)b) : (T)((U)b - (U)a));
    } else return (double)fabs(a - b);
}
template <class T>
__device__ __forceinline__ double abs_d(T a) {
    if constexpr (std::is_same<T, c128>::value) return hypot(a.re, a.im);
    else if constexpr (std::is_same<T, c64>::value) return (double)hypotf(a.re, a.im);
    else if constexpr (std::is_integral<T>::value) {
        using U = typename std::make_unsigned<T>::type;
        return (double)(a < 0 ? (T)((U)0 - (U)a) : a);
    } else return (double)fabs(a);
}
template <class T>
__device__ __forceinline__ bool is_nan_v(T a) {
    if constexpr (std::is_same<T, c128>::value || std::is_same<T, c64>::value) return a.re != a.re || a.im != a.im;
    else if constexpr (std::is_floating_point<T>::value) return a != a;
    else return false;
}

// isclose (rstsr-dtype-traits/src/isclose.rs:92-106): |a-b| <= atol + rtol*|b|,
// or both NaN when equal_nan. A failing element sets *flag (vector atomic).
template <class T>
struct CloseBody {
    using TA = T;
    using TB = T;
    const T* a;
    const T* b;
    int* flag;
    double rtol, atol;
    int equal_nan;
    static constexpr int OPS = 3;
    static constexpr size_t OUT_SIZE = 0;
    static constexpr bool NT_JOINT = false;
    template <bool NT = false>
    __device__ __forceinline__ T la(int64_t o) const { return ld<NT>(a + o); }
    template <bool NT = false>
    __device__ __forceinline__ T lb(int64_t o) const { return ld<NT>(b + o); }
    template <bool NT = false>
    __device__ __forceinline__ void st(int64_t, T x, T y) const {
        bool ok = abs_diff_d<T>(x, y) <= atol + rtol * abs_d<T>(y);
        if (equal_nan && is_nan_v<T>(x) && is_nan_v<T>(y)) ok = true;
        if (!ok) atomicOr(flag, 1);
    }
};


// Binary op with each of a / b an array or a host scalar (nullptr layout).
static void binary_impl(rstsr_hip_ctx* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c, const rstsr_layout_t* lc_,
                        const void* a, const rstsr_layout_t* la_, const void* b, const rstsr_layout_t* lb_) {
    bind(ctx);
    if ((int)op < 0 || (int)op > 9) RSTSR_RAISE(InvalidValue, "unknown binary op");
    Layout lc = from_abi(lc_);
    bool has_a = la_ != nullptr, has_b = lb_ != nullptr;
    Layout la = has_a ? from_abi(la_) : lc;
    Layout lb = has_b ? from_abi(lb_) : lc;
    if (la.shape != lc.shape || lb.shape != lc.shape)
        RSTSR_RAISE(InvalidLayout, "operands of an elementwise op must be broadcast to one shape before the device call");
    if (!a || !b) RSTSR_RAISE(InvalidValue, "null operand");
    size_t es = dtype_size(t);
    check_extent(c, lc, es, "c");
    if (has_a) check_extent(a, la, es, "a");
    if (has_b) check_extent(b, lb, es, "b");
    const Stride* s[3] = {&lc.stride, has_a ? &la.stride : nullptr, has_b ? &lb.stride : nullptr};
    size_t off[3] = {lc.offset, has_a ? la.offset : 0, has_b ? lb.offset : 0};
    Dims d = collapse3(lc.shape, s, off);
    EwChoice ch = choose(d, has_a, has_b);
    // tile-order tuning reruns the op: only when the output is not also an input
    const bool can_time = (!has_a || !same_buffer(c, a)) && (!has_b || !same_buffer(c, b));
    auto by_type = [&](auto tag) {
        using T = decltype(tag);
        auto run = [&](auto opc) {
            constexpr int OP = decltype(opc)::value;
            if constexpr (!op_defined<OP, T>()) {
                RSTSR_RAISE(UnImplemented, "this binary operator is not defined for this dtype");
            } else {
                // add/sub/mul and floating division are the hot path: every kernel form, streaming
                // variants included; the other operators run the generic and row kernels
                constexpr bool HOT = OP <= RSTSR_OP_MUL || (OP == RSTSR_OP_DIV && !std::is_integral<T>::value);
                T sa{}, sb{};
                if (!has_a) std::memcpy(&sa, a, sizeof(T));
                if (!has_b) std::memcpy(&sb, b, sizeof(T));
                auto go = [&](auto body, bool timed) {
                    launch_ew<decltype(body), HOT, HOT>(ctx, body, d, ch, timed);
                };
                if (has_a && has_b) go(BinBody<OP, T, false, false>{(T*)c, (const T*)a, (const T*)b, sa, sb}, can_time);
                else if (has_a) go(BinBody<OP, T, false, true>{(T*)c, (const T*)a, nullptr, sa, sb}, can_time);
                else if (has_b) go(BinBody<OP, T, true, false>{(T*)c, nullptr, (const T*)b, sa, sb}, can_time);
                else go(BinBody<OP, T, true, true>{(T*)c, nullptr, nullptr, sa, sb}, false);
            }
        };
        switch (op) {
            case RSTSR_OP_ADD: run(std::integral_constant<int, RSTSR_OP_ADD>{}); break;
            case RSTSR_OP_SUB: run(std::integral_constant<int, RSTSR_OP_SUB>{}); break;
            case RSTSR_OP_MUL: run(std::integral_constant<int, RSTSR_OP_MUL>{}); break;
            case RSTSR_OP_DIV: run(std::integral_constant<int, RSTSR_OP_DIV>{}); break;
            case RSTSR_OP_REM: run(std::integral_constant<int, RSTSR_OP_REM>{}); break;
            case RSTSR_OP_BITAND: run(std::integral_constant<int, RSTSR_OP_BITAND>{}); break;
            case RSTSR_OP_BITOR: run(std::integral_constant<int, RSTSR_OP_BITOR>{}); break;
            case RSTSR_OP_BITXOR: run(std::integral_constant<int, RSTSR_OP_BITXOR>{}); break;
            case RSTSR_OP_SHL: run(std::integral_constant<int, RSTSR_OP_SHL>{}); break;
            case RSTSR_OP_SHR: run(std::integral_constant<int, RSTSR_OP_SHR>{}); break;
        }
    };
    if (t == RSTSR_BOOL) by_type(bool{});
    else dispatch_dtype(t, by_type);
}

/* ---------------- assign / copy ---------------- */

// Collapse one layout in logical row-major order (merge axis k into k+1 when
// stride[k] == stride[k+1] * shape[k+1]); size-1 axes dropped.
static Layout collapse_rowmajor(const Layout& l) {
    Shape s;
    Stride t;
    for (size_t i = 0; i < l.ndim(); ++i) {
        if (l.shape[i] == 1) continue;
        if (!s.empty() && t.back() == l.stride[i] * (ptrdiff_t)l.shape[i]) {
            s.back() *= l.shape[i];
            t.back() = l.stride[i];
            continue;
        }
        s.push_back(l.shape[i]);
        t.push_back(l.stride[i]);
    }
    return Layout(s, t, l.offset);
}

// Two layouts with equal element count but different shapes, refined to one
// shape when the row-major split points of both coincide; returns false otherwise.
static bool common_refinement(const Layout& x, const Layout& y, Layout& xo, Layout& yo) {
    // walk from the innermost axis, splitting the larger extent (always valid for a strided axis)
    Shape sx, sy;
    Stride tx, ty;
    int i = (int)x.ndim() - 1, j = (int)y.ndim() - 1;
    uint64_t rx = i >= 0 ? x.shape[i] : 1, ry = j >= 0 ? y.shape[j] : 1;
    ptrdiff_t stx = i >= 0 ? x.stride[i] : 0, sty = j >= 0 ? y.stride[j] : 0;
    while (i >= 0 && j >= 0) {
        uint64_t g;
        if (rx == ry) g = rx;
        else if (rx % ry == 0) g = ry;
        else if (ry % rx == 0) g = rx;
        else return false;
        sx.push_back(g);
        tx.push_back(stx);
        sy.push_back(g);
        ty.push_back(sty);
        rx /= g;
        ry /= g;
        stx *= (ptrdiff_t)g;
        sty *= (ptrdiff_t)g;
        if (rx == 1) {
            --i;
            if (i >= 0) {
                rx = x.shape[i];
                stx = x.stride[i];
            }
        }
        if (ry == 1) {
            --j;
            if (j >= 0) {
                ry = y.shape[j];
                sty = y.stride[j];
            }
        }
    }
    if (i >= 0 || j >= 0) return false;
    std::reverse(sx.begin(), sx.end());
    std::reverse(tx.begin(), tx.end());
    std::reverse(sy.begin(), sy.end());
    std::reverse(ty.begin(), ty.end());
    xo = Layout(sx, tx, x.offset);
    yo = Layout(sy, ty, y.offset);
    return true;
}

template <class F>
static void dispatch_pair(rstsr_dtype_t tc, rstsr_dtype_t ta, F&& f) {
    dispatch_dtype(tc, [&](auto ctag) {
        dispatch_dtype(ta, [&](auto atag) {
            using TC = decltype(ctag);
            using TA = decltype(atag);
            constexpr bool c_cplx = std::is_same<TC, c64>::value || std::is_same<TC, c128>::value;
            constexpr bool a_cplx = std::is_same<TA, c64>::value || std::is_same<TA, c128>::value;
            if constexpr (a_cplx && !c_cplx) RSTSR_RAISE(InvalidValue, "cannot assign complex into a real tensor");
            else f(ctag, atag);
        });
    });
}

static void assign_impl(rstsr_hip_ctx* ctx, rstsr_dtype_t tc, void* c, const Layout& lc, rstsr_dtype_t ta, const void* a,
                        const Layout& la) {
    const Stride* s[3] = {&lc.stride, &la.stride, nullptr};
    size_t off[3] = {lc.offset, la.offset, 0};
    Dims d = collapse3(lc.shape, s, off);
    EwChoice ch = choose(d, true, false);
    const bool can_time = !same_buffer(c, a);
    dispatch_pair(tc, ta, [&](auto ctag, auto atag) {
        using TC = decltype(ctag);
        using TA = decltype(atag);
        CastBody<TC, TA> body{(TC*)c, (const TA*)a};
        launch_ew(ctx, body, d, ch, can_time);
    });
}

}  // namespace hip
}  // namespace rstsr

using namespace rstsr;
using namespace rstsr::hip;

extern "C" {

// Mixed complex / real operands; a or b may be a host scalar (null layout).
static void mixed_impl(rstsr_hip_ctx* ctx, rstsr_binop_t op, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc_,
                       rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la_, rstsr_dtype_t tb, const void* b,
                       const rstsr_layout_t* lb_) {
    bind(ctx);
    if ((int)op < 0 || (int)op > 3) RSTSR_RAISE(UnImplemented, "mixed complex/real operands take add, sub, mul, div");
    const bool is128 = tc == RSTSR_C128, is64 = tc == RSTSR_C64;
    const rstsr_dtype_t tr = is128 ? RSTSR_F64 : RSTSR_F32;
    if (!(is128 || is64) || !((ta == tc && tb == tr) || (ta == tr && tb == tc)))
        RSTSR_RAISE(UnImplemented, "mixed operands are a complex dtype and its real type, with a complex output");
    Layout lc = from_abi(lc_);
    const bool has_a = la_ != nullptr, has_b = lb_ != nullptr;
    if (!has_a && !has_b) RSTSR_RAISE(InvalidValue, "at least one operand must be an array");
    Layout la = has_a ? from_abi(la_) : lc, lb = has_b ? from_abi(lb_) : lc;
    if (la.shape != lc.shape || lb.shape != lc.shape)
        RSTSR_RAISE(InvalidLayout, "operands of an elementwise op must be broadcast to one shape before the device call");
    if (!a || !b) RSTSR_RAISE(InvalidValue, "null operand");
    check_extent(c, lc, dtype_size(tc), "c");
    if (has_a) check_extent(a, la, dtype_size(ta), "a");
    if (has_b) check_extent(b, lb, dtype_size(tb), "b");
    const Stride* s[3] = {&lc.stride, has_a ? &la.stride : nullptr, has_b ? &lb.stride : nullptr};
    size_t off[3] = {lc.offset, has_a ? la.offset : 0, has_b ? lb.offset : 0};
    Dims d = collapse3(lc.shape, s, off);
    EwChoice ch = choose(d, has_a, has_b);
    auto by_types = [&](auto ctag, auto atag, auto btag) {
        using C = decltype(ctag);
        using TA = decltype(atag);
        using TB = decltype(btag);
        auto run = [&](auto opc) {
            constexpr int OP = decltype(opc)::value;
            TA sa{};
            TB sb{};
            if (!has_a) std::memcpy(&sa, a, sizeof(TA));
            if (!has_b) std::memcpy(&sb, b, sizeof(TB));
            auto go = [&](auto body) { launch_ew<decltype(body), false, false>(ctx, body, d, ch); };
            if (has_a && has_b) go(MixBody<OP, C, TA, TB, false, false>{(C*)c, (const TA*)a, (const TB*)b, sa, sb});
            else if (has_a) go(MixBody<OP, C, TA, TB, false, true>{(C*)c, (const TA*)a, nullptr, sa, sb});
            else go(MixBody<OP, C, TA, TB, true, false>{(C*)c, nullptr, (const TB*)b, sa, sb});
        };
        switch (op) {
            case RSTSR_OP_ADD: run(std::integral_constant<int, RSTSR_OP_ADD>{}); break;
            case RSTSR_OP_SUB: run(std::integral_constant<int, RSTSR_OP_SUB>{}); break;
            case RSTSR_OP_MUL: run(std::integral_constant<int, RSTSR_OP_MUL>{}); break;
            default: run(std::integral_constant<int, RSTSR_OP_DIV>{}); break;
        }
    };
    if (is128) {
        if (ta == tc) by_types(c128{}, c128{}, double{});
        else by_types(c128{}, double{}, c128{});
    } else {
        if (ta == tc) by_types(c64{}, c64{}, float{});
        else by_types(c64{}, float{}, c64{});
    }
}

int rstsr_hip_op_mutc_refa_refb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                                const rstsr_layout_t* lc, const void* a, const rstsr_layout_t* la, const void* b,
                                const rstsr_layout_t* lb) {
    return guarded([&] {
        if (!la || !lb) RSTSR_RAISE(InvalidValue, "null layout");
        binary_impl(ctx, op, t, c, lc, a, la, b, lb);
    });
}

int rstsr_hip_op_mutc_refa_numb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                                const rstsr_layout_t* lc, const void* a, const rstsr_layout_t* la, const void* b) {
    return guarded([&] {
        if (!la) RSTSR_RAISE(InvalidValue, "null layout");
        binary_impl(ctx, op, t, c, lc, a, la, b, nullptr);
    });
}

int rstsr_hip_op_mutc_numa_refb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                                const rstsr_layout_t* lc, const void* a, const void* b, const rstsr_layout_t* lb) {
    return guarded([&] {
        if (!lb) RSTSR_RAISE(InvalidValue, "null layout");
        binary_impl(ctx, op, t, c, lc, a, nullptr, b, lb);
    });
}

int rstsr_hip_op_mutc_mixed(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t tc, void* c,
                            const rstsr_layout_t* lc, rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la,
                            rstsr_dtype_t tb, const void* b, const rstsr_layout_t* lb) {
    return guarded([&] { mixed_impl(ctx, op, tc, c, lc, ta, a, la, tb, b, lb); });
}

int rstsr_hip_op_muta_refb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                           const rstsr_layout_t* lc, const void* b, const rstsr_layout_t* lb) {
    return guarded([&] {
        if (!lb) RSTSR_RAISE(InvalidValue, "null layout");
        binary_impl(ctx, op, t, c, lc, c, lc, b, lb);
    });
}

int rstsr_hip_op_muta_numb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                           const rstsr_layout_t* lc, const void* b) {
    return guarded([&] { binary_impl(ctx, op, t, c, lc, c, lc, b, nullptr); });
}

int rstsr_hip_assign(rstsr_hip_ctx_t* ctx, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc_, rstsr_dtype_t ta,
                     const void* a, const rstsr_layout_t* la_) {
    return guarded([&] {
        bind(ctx);
        Layout lc = from_abi(lc_), la = from_abi(la_);
        if (lc.shape != la.shape) RSTSR_RAISE(InvalidLayout, "assign requires equal shapes");
        check_extent(c, lc, dtype_size(tc), "c");
        check_extent(a, la, dtype_size(ta), "a");
        assign_impl(ctx, tc, c, lc, ta, a, la);
    });
}

int rstsr_hip_assign_arbitary(rstsr_hip_ctx_t* ctx, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc_,
                              rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la_) {
    return guarded([&] {
        bind(ctx);
        Layout lc = from_abi(lc_), la = from_abi(la_);
        if (lc.size() != la.size()) RSTSR_RAISE(InvalidLayout, "assign_arbitary requires equal element counts");
        check_extent(c, lc, dtype_size(tc), "c");
        check_extent(a, la, dtype_size(ta), "a");
        if (lc.size() == 0) return;
        if (lc.shape == la.shape) {
            assign_impl(ctx, tc, c, lc, ta, a, la);
            return;
        }
        Layout xc = collapse_rowmajor(lc), xa = collapse_rowmajor(la), rc, ra;
        if (!common_refinement(xc, xa, rc, ra))
            RSTSR_RAISE(UnImplemented, "assign_arbitary between shapes without a common row-major refinement");
        assign_impl(ctx, tc, c, rc, ta, a, ra);
    });
}

int rstsr_hip_fill(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, void* c, const rstsr_layout_t* lc_, const void* value) {
    return guarded([&] {
        bind(ctx);
        if (!value) RSTSR_RAISE(InvalidValue, "null fill value");
        Layout lc = from_abi(lc_);
        check_extent(c, lc, dtype_size(t), "c");
        const Stride* s[3] = {&lc.stride, nullptr, nullptr};
        size_t off[3] = {lc.offset, 0, 0};
        Dims d = collapse3(lc.shape, s, off);
        EwChoice ch = choose(d, false, false);
        dispatch_dtype(t, [&](auto tag) {
            using T = decltype(tag);
            FillBody<T> body{(T*)c, T{}};
            std::memcpy(&body.v, value, sizeof(T));
            launch_ew<FillBody<T>, false>(ctx, body, d, ch);
        });
    });
}

int rstsr_hip_allclose(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, const void* a, const rstsr_layout_t* la_, const void* b,
                       const rstsr_layout_t* lb_, double rtol, double atol, int equal_nan, int* host_out) {
    return guarded([&] {
        bind(ctx);
        if (!host_out) RSTSR_RAISE(InvalidValue, "null output");
        // refused before anything is enqueued (its flag is cleared with a memset)
        if (stream_capturing(ctx)) RSTSR_RAISE(UnImplemented, "allclose hands its value to the host and cannot be captured");
        Layout la = from_abi(la_), lb = from_abi(lb_);
        if (la.size() == 0 || lb.size() == 0) RSTSR_RAISE(InvalidValue, "zero-size array is not supported for allclose");
        auto bl = broadcast_layout(la, lb, FlagOrder::RowMajor);
        la = bl.first;
        lb = bl.second;
        size_t es = dtype_size(t);
        check_extent(a, la, es, "a");
        check_extent(b, lb, es, "b");
        std::lock_guard<std::mutex> g(ctx->mu);
        int* flag = (int*)workspace(ctx, 256);
        HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
        const Stride* s[3] = {&la.stride, &la.stride, &lb.stride};
        size_t off[3] = {la.offset, la.offset, lb.offset};
        Dims d = collapse3(la.shape, s, off);
        EwChoice ch = choose(d, true, true);
        dispatch_dtype(t, [&](auto tag) {
            using T = decltype(tag);
            CloseBody<T> body{(const T*)a, (const T*)b, flag, rtol, atol, equal_nan};
            launch_ew<CloseBody<T>, false>(ctx, body, d, ch);
        });
        HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        int bad = 0;
        std::memcpy(&bad, ctx->host_pinned, sizeof(int));
        *host_out = bad ? 0 : 1;
    });
}

}  // extern "C"
