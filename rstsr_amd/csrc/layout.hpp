// Strided n-d layout algebra: shape / stride (elements, signed) / offset.
//
// Host-only C++17, header-only; shared by the HIP library (kernel planning,
// reduction output layouts) and the C++ tensor front end. Each routine
// restates the rstsr-common function cited beside it, so that layouts chosen
// here are the ones the reference's tensor layer and device kernels choose.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rstsr {

// Error kinds of rstsr-common/src/error.rs:15-48 that this path can raise.
enum class ErrorKind : int {
    Ok = 0,
    InvalidLayout = 1,
    InvalidValue = 2,
    ValueOutOfRange = 3,
    DeviceMismatch = 4,
    DeviceError = 5,
    MemoryError = 6,
    UnImplemented = 7,
    IndexError = 8,
    RuntimeError = 9,
};

struct Error : public std::runtime_error {
    ErrorKind kind;
    Error(ErrorKind k, const std::string& msg) : std::runtime_error(msg), kind(k) {}
};

inline const char* error_kind_name(ErrorKind k) {
    switch (k) {
        case ErrorKind::Ok: return "Ok";
        case ErrorKind::InvalidLayout: return "InvalidLayout";
        case ErrorKind::InvalidValue: return "InvalidValue";
        case ErrorKind::ValueOutOfRange: return "ValueOutOfRange";
        case ErrorKind::DeviceMismatch: return "DeviceMismatch";
        case ErrorKind::DeviceError: return "DeviceError";
        case ErrorKind::MemoryError: return "MemoryError";
        case ErrorKind::UnImplemented: return "UnImplemented";
        case ErrorKind::IndexError: return "IndexError";
        case ErrorKind::RuntimeError: return "RuntimeError";
    }
    return "Unknown";
}

#define RSTSR_RAISE(kind, msg) throw ::rstsr::Error(::rstsr::ErrorKind::kind, std::string(msg))
#define RSTSR_ASSERT(cond, kind, msg)                                                       \
    do {                                                                                    \
        if (!(cond)) throw ::rstsr::Error(::rstsr::ErrorKind::kind, std::string(msg));     \
    } while (0)

// rstsr-common/src/flags.rs:62-86
enum class FlagOrder : int { RowMajor = 0, ColMajor = 1 };
// rstsr-common/src/flags.rs:94-136 (default K)
enum class IterOrder : int { C = 0, F = 1, A = 2, K = 3, G = 4, B = 5 };

using Shape = std::vector<size_t>;
using Stride = std::vector<ptrdiff_t>;

struct Layout {
    Shape shape;
    Stride stride;
    size_t offset = 0;

    Layout() = default;
    Layout(Shape s, Stride t, size_t off) : shape(std::move(s)), stride(std::move(t)), offset(off) {
        if (shape.size() != stride.size()) RSTSR_RAISE(InvalidLayout, "shape and stride ranks differ");
    }

    size_t ndim() const { return shape.size(); }
    size_t size() const {
        size_t n = 1;
        for (auto d : shape) n *= d;
        return n;
    }
    bool operator==(const Layout& o) const {
        return shape == o.shape && stride == o.stride && offset == o.offset;
    }
    bool operator!=(const Layout& o) const { return !(*this == o); }

    // rstsr-common/src/layout/layoutbase.rs:148-164
    size_t ndim_of_f_contig() const {
        if (ndim() == 0 || size() == 0) return ndim();
        ptrdiff_t acc = 1;
        for (size_t i = 0; i < ndim(); ++i) {
            if (shape[i] != 1 && stride[i] != acc) return i;
            acc *= (ptrdiff_t)shape[i];
        }
        return ndim();
    }
    // rstsr-common/src/layout/layoutbase.rs:170-186
    size_t ndim_of_c_contig() const {
        if (ndim() == 0 || size() == 0) return ndim();
        ptrdiff_t acc = 1;
        size_t cnt = 0;
        for (size_t ii = ndim(); ii-- > 0; ++cnt) {
            if (shape[ii] != 1 && stride[ii] != acc) return cnt;
            acc *= (ptrdiff_t)shape[ii];
        }
        return ndim();
    }
    bool f_contig() const { return ndim() == ndim_of_f_contig(); }
    bool c_contig() const { return ndim() == ndim_of_c_contig(); }

    // rstsr-common/src/layout/layoutbase.rs:89-116: strides non-decreasing over
    // non-trivial axes, first one unit
    bool f_prefer() const {
        if (ndim() == 0 || size() == 0) return true;
        ptrdiff_t last = 0;
        for (size_t i = 0; i < ndim(); ++i) {
            if (shape[i] != 1) {
                if (stride[i] < last) return false;
                if (last == 0 && stride[i] != 1) return false;
                last = stride[i];
            } else if (last == 0) {
                last = 1;
            }
        }
        return true;
    }
    // rstsr-common/src/layout/layoutbase.rs:119-146 (mirror of f_prefer from the last axis)
    bool c_prefer() const {
        if (ndim() == 0 || size() == 0) return true;
        ptrdiff_t last = 0;
        for (size_t ii = ndim(); ii-- > 0;) {
            if (shape[ii] != 1) {
                if (stride[ii] < last) return false;
                if (last == 0 && stride[ii] != 1) return false;
                last = stride[ii];
            } else if (last == 0) {
                last = 1;
            }
        }
        return true;
    }

    // rstsr-common/src/layout/layoutbase.rs:237-262, [min, max)
    std::pair<size_t, size_t> bounds_index() const {
        if (ndim() == 0) return {offset, offset + 1};
        ptrdiff_t mn = (ptrdiff_t)offset, mx = (ptrdiff_t)offset;
        for (size_t i = 0; i < ndim(); ++i) {
            if (shape[i] == 0) return {offset, offset};
            if (stride[i] > 0)
                mx += stride[i] * ((ptrdiff_t)shape[i] - 1);
            else
                mn += stride[i] * ((ptrdiff_t)shape[i] - 1);
        }
        if (mn < 0) RSTSR_RAISE(ValueOutOfRange, "layout reaches a negative memory index");
        return {(size_t)mn, (size_t)mx + 1};
    }

    // number of elements whose stride is non-zero (broadcast.rs size_non_broadcast)
    size_t size_non_broadcast() const {
        if (size() == 0) return 0;
        size_t n = 1;
        for (size_t i = 0; i < ndim(); ++i)
            if (stride[i] != 0) n *= shape[i];
        return n;
    }

    Layout transpose(const std::vector<ptrdiff_t>& perm) const {
        RSTSR_ASSERT(perm.size() == ndim(), InvalidLayout, "transpose: permutation rank mismatch");
        std::vector<bool> seen(ndim(), false);
        Shape s(ndim());
        Stride t(ndim());
        for (size_t i = 0; i < ndim(); ++i) {
            ptrdiff_t p = perm[i];
            if (p < 0) p += (ptrdiff_t)ndim();
            RSTSR_ASSERT(p >= 0 && (size_t)p < ndim() && !seen[p], InvalidLayout, "transpose: invalid permutation");
            seen[p] = true;
            s[i] = shape[p];
            t[i] = stride[p];
        }
        return Layout(s, t, offset);
    }
    Layout reverse_axes() const {
        Shape s(shape.rbegin(), shape.rend());
        Stride t(stride.rbegin(), stride.rend());
        return Layout(s, t, offset);
    }
    Layout swapaxes(ptrdiff_t a, ptrdiff_t b) const {
        a = norm_axis(a);
        b = norm_axis(b);
        Layout l = *this;
        std::swap(l.shape[a], l.shape[b]);
        std::swap(l.stride[a], l.stride[b]);
        return l;
    }
    ptrdiff_t norm_axis(ptrdiff_t a) const {
        ptrdiff_t n = (ptrdiff_t)ndim();
        ptrdiff_t r = a < 0 ? a + n : a;
        if (r < 0 || r >= n)
            RSTSR_RAISE(InvalidValue, "axis " + std::to_string(a) + " out of range for ndim " + std::to_string(n));
        return r;
    }

    // Slice one axis (rstsr-common/src/layout/indexer.rs:120-197). `has_*` false
    // means None. Out-of-range bounds clamp as the reference does; note that an
    // explicit stop of -1 with a negative step means "through index 0" there.
    Layout dim_narrow(ptrdiff_t axis, bool has_start, ptrdiff_t start, bool has_stop, ptrdiff_t stop,
                      ptrdiff_t step) const {
        axis = norm_axis(axis);
        if (!has_start && !has_stop && step == 1) return *this;
        RSTSR_ASSERT(step != 0, InvalidValue, "slice step must not be zero");
        ptrdiff_t len = (ptrdiff_t)shape[axis];
        if (len == 0) return *this;
        Layout l = *this;
        if (step > 0) {
            ptrdiff_t b = has_start ? start : 0, e = has_stop ? stop : len;
            if (b < 0) b = std::max<ptrdiff_t>(len + b, 0);
            if (e < 0) e = std::max<ptrdiff_t>(len + e, 0);
            if (b > len || b > e) {
                b = 0;
                e = 0;
            } else if (e > len) {
                e = len;
            }
            l.offset = (size_t)((ptrdiff_t)offset + stride[axis] * b);
            l.shape[axis] = (size_t)std::max<ptrdiff_t>((e - b + step - 1) / step, 0);
        } else {
            ptrdiff_t b = has_start ? start : len - 1, e = has_stop ? stop : -1;
            if (b < 0) b = std::max<ptrdiff_t>(len + b, 0);
            if (e < -1) e = std::max<ptrdiff_t>(len + e, -1);
            if (e > len - 1 || e > b) {
                b = 0;
                e = 0;
            } else if (b > len - 1) {
                b = len - 1;
            }
            l.offset = (size_t)((ptrdiff_t)offset + stride[axis] * b);
            l.shape[axis] = (size_t)std::max<ptrdiff_t>((e - b + step + 1) / step, 0);
        }
        l.stride[axis] = stride[axis] * step;
        return l;
    }

    // rstsr-common/src/layout/indexer.rs:453-476; axes kept in the given order
    std::pair<Layout, Layout> dim_split_axes(const std::vector<ptrdiff_t>& axes) const;
};

inline Layout c_layout(const Shape& shape, size_t offset = 0) {
    Stride t(shape.size());
    ptrdiff_t acc = 1;
    for (size_t ii = shape.size(); ii-- > 0;) {
        t[ii] = acc;
        acc *= (ptrdiff_t)std::max<size_t>(shape[ii], 1);
    }
    return Layout(shape, t, offset);
}
inline Layout f_layout(const Shape& shape, size_t offset = 0) {
    Stride t(shape.size());
    ptrdiff_t acc = 1;
    for (size_t i = 0; i < shape.size(); ++i) {
        t[i] = acc;
        acc *= (ptrdiff_t)std::max<size_t>(shape[i], 1);
    }
    return Layout(shape, t, offset);
}

// rstsr-common/src/axis_index.rs:379-411 (no duplicates, keep order unless sort)
inline std::vector<ptrdiff_t> normalize_axes_index(const std::vector<ptrdiff_t>& axes, size_t ndim, bool sort) {
    std::vector<ptrdiff_t> v;
    for (auto a : axes) {
        ptrdiff_t r = a < 0 ? a + (ptrdiff_t)ndim : a;
        if (r < 0 || r >= (ptrdiff_t)ndim)
            RSTSR_RAISE(InvalidValue, "axis " + std::to_string(a) + " out of range for ndim " + std::to_string(ndim));
        v.push_back(r);
    }
    std::vector<ptrdiff_t> s = v;
    std::sort(s.begin(), s.end());
    for (size_t i = 1; i < s.size(); ++i)
        if (s[i] == s[i - 1]) RSTSR_RAISE(InvalidValue, "Duplicate axes are not allowed.");
    return sort ? s : v;
}

inline std::pair<Layout, Layout> Layout::dim_split_axes(const std::vector<ptrdiff_t>& axes) const {
    auto ax = normalize_axes_index(axes, ndim(), false);
    Shape s1, s2;
    Stride t1, t2;
    for (auto a : ax) {
        s1.push_back(shape[a]);
        t1.push_back(stride[a]);
    }
    for (size_t i = 0; i < ndim(); ++i) {
        if (std::find(ax.begin(), ax.end(), (ptrdiff_t)i) != ax.end()) continue;
        s2.push_back(shape[i]);
        t2.push_back(stride[i]);
    }
    return {Layout(s1, t1, offset), Layout(s2, t2, offset)};
}

/* ---------------- broadcasting: rstsr-common/src/layout/broadcast.rs ---------------- */

enum class BroadcastType { Upcast, Expand, Preserve, Undefined };

// broadcast.rs:21-101
inline Shape broadcast_shape(const Shape& s1_in, const Shape& s2_in, FlagOrder order, std::vector<BroadcastType>& tp1,
                             std::vector<BroadcastType>& tp2) {
    Shape s1 = s1_in, s2 = s2_in;
    if (order == FlagOrder::ColMajor) {
        std::reverse(s1.begin(), s1.end());
        std::reverse(s2.begin(), s2.end());
    }
    size_t n1 = s1.size(), n2 = s2.size(), n = std::max(n1, n2);
    Shape shape(n, 0);
    tp1.assign(n, BroadcastType::Undefined);
    tp2.assign(n, BroadcastType::Undefined);
    for (size_t ii = n; ii-- > 0;) {
        ptrdiff_t in1 = (ptrdiff_t)(n1 + ii) - (ptrdiff_t)n;
        ptrdiff_t in2 = (ptrdiff_t)(n2 + ii) - (ptrdiff_t)n;
        size_t d1 = in1 >= 0 ? s1[in1] : 1;
        size_t d2 = in2 >= 0 ? s2[in2] : 1;
        if (d1 == 1 && d2 == 1) {
            tp1[ii] = tp2[ii] = BroadcastType::Preserve;
            shape[ii] = 1;
        } else if (d1 != 1 && d2 == 1) {
            tp1[ii] = BroadcastType::Preserve;
            tp2[ii] = BroadcastType::Upcast;
            shape[ii] = d1;
        } else if (d1 == 1 && d2 != 1) {
            tp1[ii] = BroadcastType::Upcast;
            tp2[ii] = BroadcastType::Preserve;
            shape[ii] = d2;
        } else {
            if (d1 != d2) RSTSR_RAISE(InvalidLayout, "Broadcasting failed.");
            tp1[ii] = tp2[ii] = BroadcastType::Preserve;
            shape[ii] = d1;
        }
        if (in1 < 0) tp1[ii] = BroadcastType::Expand;
        if (in2 < 0) tp2[ii] = BroadcastType::Expand;
    }
    if (order == FlagOrder::ColMajor) {
        std::reverse(shape.begin(), shape.end());
        std::reverse(tp1.begin(), tp1.end());
        std::reverse(tp2.begin(), tp2.end());
    }
    return shape;
}

// broadcast.rs:207-239
inline Layout update_layout_by_shape(const Layout& layout, const Shape& shape, const std::vector<BroadcastType>& tp,
                                     FlagOrder order) {
    if (order == FlagOrder::ColMajor) {
        Shape s(shape.rbegin(), shape.rend());
        std::vector<BroadcastType> t(tp.rbegin(), tp.rend());
        return update_layout_by_shape(layout.reverse_axes(), s, t, FlagOrder::RowMajor).reverse_axes();
    }
    size_t n_old = layout.ndim(), n = shape.size();
    Stride stride(n, 0);
    for (size_t i = 0; i < n_old; ++i) stride[n - n_old + i] = layout.stride[i];
    for (size_t i = 0; i < n; ++i)
        if (tp[i] == BroadcastType::Expand || tp[i] == BroadcastType::Upcast) stride[i] = 0;
    return Layout(shape, stride, layout.offset);
}

// broadcast.rs:166-182
inline std::pair<Layout, Layout> broadcast_layout(const Layout& l1, const Layout& l2, FlagOrder order) {
    std::vector<BroadcastType> tp1, tp2;
    Shape shape = broadcast_shape(l1.shape, l2.shape, order, tp1, tp2);
    return {update_layout_by_shape(l1, shape, tp1, order), update_layout_by_shape(l2, shape, tp2, order)};
}

// broadcast a layout to a given (larger) shape, as broadcast_layout_to_first with
// the first operand being a pure shape.
inline Layout broadcast_layout_to_shape(const Layout& l, const Shape& shape, FlagOrder order) {
    std::vector<BroadcastType> tp1, tp2;
    Shape s = broadcast_shape(shape, l.shape, order, tp1, tp2);
    if (s != shape) RSTSR_RAISE(InvalidLayout, "Broadcasting failed: operand cannot be broadcast to target shape.");
    return update_layout_by_shape(l, shape, tp2, order);
}

/* ---------------- rearrangement: rstsr-common/src/layout/rearrangement.rs ---------------- */

// rearrangement.rs:36-113; returns (layout, permutation index)
inline std::pair<Layout, std::vector<ptrdiff_t>> greedy_layout(const Layout& layout_in, bool keep_shape) {
    Layout layout = layout_in;
    size_t nd = layout.ndim();
    std::vector<ptrdiff_t> index(nd);
    std::iota(index.begin(), index.end(), 0);
    if (layout.size() == 0) return {layout, index};
    if (keep_shape) {
        for (size_t n = 0; n < nd; ++n)
            if (layout.stride[n] < 0) layout = layout.dim_narrow((ptrdiff_t)n, false, 0, false, 0, -1);
    }
    const Shape& sh = layout.shape;
    const Stride& st = layout.stride;
    auto bcast = [&](ptrdiff_t i) { return sh[i] == 1 || st[i] == 0; };
    // std::stable_sort with a strict weak order equivalent to Rust's stable sort_by
    std::stable_sort(index.begin(), index.end(), [&](ptrdiff_t i1, ptrdiff_t i2) {
        bool b1 = bcast(i1), b2 = bcast(i2);
        if (b1 && b2) return i1 < i2;
        if (b1 && !b2) return keep_shape;   // keep: broadcast first; else broadcast last
        if (!b1 && b2) return !keep_shape;
        return std::llabs((long long)st[i1]) < std::llabs((long long)st[i2]);
    });
    Layout out = layout.transpose(index);
    if (!keep_shape) {
        for (size_t i = 0; i < nd; ++i) {
            if (out.shape[i] == 1 || out.stride[i] == 0) {
                out.shape[i] = 1;
                out.stride[i] = 0;
            }
        }
    }
    return {out, index};
}

// rearrangement.rs:116-123
inline std::vector<ptrdiff_t> reversed_permute(const std::vector<ptrdiff_t>& idx) {
    std::vector<ptrdiff_t> r(idx.size());
    for (size_t i = 0; i < idx.size(); ++i) r[idx[i]] = (ptrdiff_t)i;
    return r;
}

// rearrangement.rs:125-152
inline Layout layout_for_array_copy(const Layout& layout, IterOrder order, FlagOrder dflt = FlagOrder::RowMajor) {
    switch (order) {
        case IterOrder::C: return c_layout(layout.shape);
        case IterOrder::F: return f_layout(layout.shape);
        case IterOrder::A:
            if (layout.c_contig()) return c_layout(layout.shape);
            if (layout.f_contig()) return f_layout(layout.shape);
            return dflt == FlagOrder::RowMajor ? c_layout(layout.shape) : f_layout(layout.shape);
        case IterOrder::K: {
            auto g = greedy_layout(layout, true);
            Layout f = f_layout(g.first.shape);
            return f.transpose(reversed_permute(g.second));
        }
        default: RSTSR_RAISE(InvalidValue, "Iter order for copy only accepts CFAK.");
    }
}

// rearrangement.rs:335-372: (singleton, broadcast, contiguous, discontiguous)
struct AxesComposition {
    std::vector<size_t> i, b, c, d;
};
inline AxesComposition get_axes_composition(const Layout& layout) {
    AxesComposition r;
    std::vector<size_t> remain;
    for (size_t i = 0; i < layout.ndim(); ++i) {
        if (layout.shape[i] == 1)
            r.i.push_back(i);
        else if (layout.stride[i] == 0)
            r.b.push_back(i);
        else
            remain.push_back(i);
    }
    std::stable_sort(remain.begin(), remain.end(), [&](size_t a, size_t b) {
        return std::llabs((long long)layout.stride[a]) < std::llabs((long long)layout.stride[b]);
    });
    ptrdiff_t cur = 1;
    for (auto i : remain) {
        if (layout.stride[i] == cur) {
            r.c.push_back(i);
            cur *= (ptrdiff_t)layout.shape[i];
        }
    }
    for (auto i : remain)
        if (std::find(r.c.begin(), r.c.end(), i) == r.c.end()) r.d.push_back(i);
    return r;
}

// rearrangement.rs:394-459
inline Layout get_layout_for_binary_op(const Layout& la, const Layout& lb, FlagOrder order) {
    if (la.shape != lb.shape) RSTSR_RAISE(InvalidLayout, "Shape of two layouts must be the same for this function.");
    size_t nd = la.ndim();
    auto ca = get_axes_composition(la);
    auto cb = get_axes_composition(lb);
    std::vector<size_t> a0_o, ac_o, ad_o;
    for (auto i : ca.b)
        if (std::find(cb.b.begin(), cb.b.end(), i) != cb.b.end()) a0_o.push_back(i);
    for (size_t k = 0; k < std::min(ca.c.size(), cb.c.size()); ++k) {
        if (ca.c[k] == cb.c[k])
            ac_o.push_back(ca.c[k]);
        else
            break;
    }
    auto has = [](const std::vector<size_t>& v, size_t x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    for (size_t i = 0; i < nd; ++i)
        if (!has(a0_o, i) && !has(ac_o, i) && !has(ca.i, i)) ad_o.push_back(i);
    if (order == FlagOrder::RowMajor) std::reverse(ad_o.begin(), ad_o.end());
    Stride stride(nd, 0);
    ptrdiff_t cur = 1;
    for (auto i : ac_o) {
        stride[i] = cur;
        cur *= (ptrdiff_t)la.shape[i];
    }
    for (auto i : ad_o) {
        stride[i] = cur;
        cur *= (ptrdiff_t)la.shape[i];
    }
    for (auto i : ca.i) {
        ptrdiff_t s = 1;
        if (order == FlagOrder::RowMajor) {
            for (size_t k = i; k < nd; ++k)
                if (stride[k] != 0) {
                    s = stride[k];
                    break;
                }
        } else {
            for (size_t k = i; k-- > 0;)
                if (stride[k] != 0) {
                    s = stride[k];
                    break;
                }
        }
        stride[i] = s;
    }
    return Layout(la.shape, stride, 0);
}

/* ---------------- matmul rules: rstsr-common/src/layout/matmul.rs ---------------- */

enum class MatMulType { InnerDot, GEMM22, GEVM, GEMV, GEMM2X, GEMMX2, GEMMXX };

// Output layout of `a % b` (matmul.rs:121-280, row-major; col-major by the
// reverse-and-swap identity of matmul.rs:423-455).
inline Layout matmul_output_layout(const Layout& la, const Layout& lb, FlagOrder order) {
    if (order == FlagOrder::ColMajor) {
        return matmul_output_layout(lb.reverse_axes(), la.reverse_axes(), FlagOrder::RowMajor).reverse_axes();
    }
    size_t na = la.ndim(), nb = lb.ndim();
    if (na == 0 || nb == 0) RSTSR_RAISE(InvalidValue, "In matmul, 0-dim is not allowed.");
    if (na == 1 && nb == 1) {
        if (la.shape != lb.shape) RSTSR_RAISE(InvalidLayout, "matmul: inner dimensions differ");
        return Layout({}, {}, 0);
    }
    if (na == 2 && nb == 2) {
        if (la.shape[1] != lb.shape[0]) RSTSR_RAISE(InvalidLayout, "matmul: inner dimensions differ");
        return c_layout({la.shape[0], lb.shape[1]});
    }
    if (na == 1) {  // rule 3
        if (la.shape[0] != lb.shape[nb - 2]) RSTSR_RAISE(InvalidLayout, "matmul: inner dimensions differ");
        Shape s(lb.shape.begin(), lb.shape.end() - 2);
        s.push_back(lb.shape[nb - 1]);
        return c_layout(s);
    }
    if (nb == 1) {  // rule 4
        if (lb.shape[0] != la.shape[na - 1]) RSTSR_RAISE(InvalidLayout, "matmul: inner dimensions differ");
        Shape s(la.shape.begin(), la.shape.end() - 2);
        s.push_back(la.shape[na - 2]);
        return c_layout(s);
    }
    if (la.shape[na - 1] != lb.shape[nb - 2]) RSTSR_RAISE(InvalidLayout, "matmul: inner dimensions differ");
    Shape ra(la.shape.begin(), la.shape.end() - 2), rb(lb.shape.begin(), lb.shape.end() - 2);
    std::vector<BroadcastType> t1, t2;
    Shape rc = broadcast_shape(ra, rb, FlagOrder::RowMajor, t1, t2);
    rc.push_back(la.shape[na - 2]);
    rc.push_back(lb.shape[nb - 1]);
    return c_layout(rc);
}

}  // namespace rstsr
