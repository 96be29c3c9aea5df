// C++ host front end over the C ABI: the DeviceHIP device and a tensor type
// whose operations follow rstsr-core's tensor layer (the code above the device
// trait line), so that parity tests read like the reference's own tests.
//
//   DeviceHIP        rstsr-core/src/device_faer/device.rs:12-147 (a device value:
//                    one GPU + one stream + default order)
//   Tensor           rstsr-core/src/tensorbase.rs:5-24 (storage + layout)
//   binary ops       tensor/operators/op_binary_arithmetic.rs:188-211 (broadcast,
//                    get_layout_for_binary_op, uninit, device op)
//   scalar ops       op_binary_arithmetic.rs:905-911 (layout_for_array_copy(K))
//   reductions       tensor/reduction.rs:3-133
//   to_contig        tensor/manipulation/to_contig.rs:8-23, to_layout.rs:8-38
//   matmul (%)       tensor/linalg/matmul.rs:45-97
#pragma once

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rstsr_hip.h"
#include "layout.hpp"

namespace rstsr {
namespace front {

inline void check(int st) {
    if (st != 0) throw Error((ErrorKind)st, rstsr_hip_last_error());
}

inline size_t dtype_size(rstsr_dtype_t t) {
    switch (t) {
        case RSTSR_F32: return 4;
        case RSTSR_F64: return 8;
        case RSTSR_I32: return 4;
        case RSTSR_I64: return 8;
        case RSTSR_C64: return 8;
        case RSTSR_C128: return 16;
        case RSTSR_U64: return 8;
        case RSTSR_BOOL: return 1;
    }
    RSTSR_RAISE(InvalidValue, "unknown dtype");
}

struct Ctx {
    rstsr_hip_ctx_t* ctx = nullptr;
    int device = 0;
    explicit Ctx(int dev) : device(dev) { check(rstsr_hip_ctx_create(dev, &ctx)); }
    ~Ctx() {
        if (ctx) rstsr_hip_ctx_destroy(ctx);
    }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
};

// Layout <-> ABI view (the arrays must outlive the call)
struct AbiLayout {
    rstsr_layout_t l;
    explicit AbiLayout(const Layout& x) {
        l.ndim = x.ndim();
        l.shape = x.shape.data();
        l.stride = x.stride.data();
        l.offset = x.offset;
    }
    const rstsr_layout_t* operator&() const { return &l; }
};

class DeviceHIP {
   public:
    std::shared_ptr<Ctx> ctx;
    FlagOrder order = FlagOrder::RowMajor;

    explicit DeviceHIP(int dev = 0) : ctx(std::make_shared<Ctx>(dev)) {}
    rstsr_hip_ctx_t* raw() const { return ctx->ctx; }
    bool same_device(const DeviceHIP& o) const { return rstsr_hip_same_device(raw(), o.raw()) == 1; }
    FlagOrder default_order() const { return order; }
    void set_default_order(FlagOrder o) { order = o; }
    void synchronize() const { check(rstsr_hip_synchronize(raw())); }
    // Work this device queues next runs after what `o` has queued (same GPU, other
    // stream): an operand another context just wrote is read whole, and one it still
    // reads is not overwritten early.
    void order_after(const DeviceHIP& o) const {
        if (!same_device(o)) RSTSR_RAISE(DeviceMismatch, "operands live on different devices");
        if (o.raw() != raw()) check(rstsr_hip_stream_wait(raw(), o.raw()));
    }
};

struct Storage {
    DeviceHIP dev;
    void* ptr = nullptr;
    size_t nbytes = 0;
    Storage(const DeviceHIP& d, size_t nb) : dev(d), nbytes(nb) { check(rstsr_hip_malloc(dev.raw(), nb, &ptr)); }
    ~Storage() {
        if (ptr) rstsr_hip_free(dev.raw(), ptr);
    }
    Storage(const Storage&) = delete;
    Storage& operator=(const Storage&) = delete;
};

class Tensor {
   public:
    std::shared_ptr<Storage> storage;
    Layout layout;
    rstsr_dtype_t dtype = RSTSR_F64;

    Tensor() = default;
    Tensor(std::shared_ptr<Storage> s, Layout l, rstsr_dtype_t t) : storage(std::move(s)), layout(std::move(l)), dtype(t) {}

    const DeviceHIP& device() const { return storage->dev; }
    void* ptr() const { return storage->ptr; }
    size_t esize() const { return dtype_size(dtype); }
    size_t ndim() const { return layout.ndim(); }
    size_t size() const { return layout.size(); }

    // uninit_impl(bounds_index().1) (storage/creation.rs:26-30)
    static Tensor uninit(const DeviceHIP& dev, const Layout& l, rstsr_dtype_t t) {
        size_t n = l.bounds_index().second;
        auto s = std::make_shared<Storage>(dev, std::max<size_t>(n, 1) * dtype_size(t));
        return Tensor(s, l, t);
    }

    // outof_cpu_vec (storage/creation.rs:11) with a c-contiguous layout
    static Tensor from_host(const DeviceHIP& dev, const void* data, const Shape& shape, rstsr_dtype_t t) {
        Tensor r = uninit(dev, c_layout(shape), t);
        check(rstsr_hip_memcpy_h2d(dev.raw(), r.ptr(), data, r.size() * r.esize()));
        return r;
    }

    // bytes of [bounds_min, bounds_max) and the offset of this layout inside it
    std::pair<std::string, size_t> raw_range() const {
        if (size() == 0) return {std::string(), 0};
        auto b = layout.bounds_index();
        std::string out((b.second - b.first) * esize(), '\0');
        check(rstsr_hip_memcpy_d2h(device().raw(), &out[0], (char*)ptr() + b.first * esize(), out.size()));
        return {out, layout.offset - b.first};
    }

    Tensor view_with(const Layout& l) const { return Tensor(storage, l, dtype); }

    // DeviceChangeAPI::to_device (storage/conversion.rs:3-21, tensor/device_conversion.rs:36-58):
    // the storage range the view covers moves to `dev` and the layout is kept (offset rebased).
    // Work queued on the source stream finishes first; the copy is ordered on the target stream.
    Tensor to_device(const DeviceHIP& dev) const {
        if (size() == 0) return uninit(dev, layout, dtype);
        auto b = layout.bounds_index();
        const size_t n = (b.second - b.first) * esize();
        auto st = std::make_shared<Storage>(dev, std::max<size_t>(n, 1));
        check(rstsr_hip_synchronize(device().raw()));
        check(rstsr_hip_memcpy_d2d(dev.raw(), st->ptr, (char*)ptr() + b.first * esize(), n));
        Layout l = layout;
        l.offset = layout.offset - b.first;
        return Tensor(st, l, dtype);
    }

    /* ---------- assignment ---------- */
    void assign(const Tensor& a) {
        device().order_after(a.device());
        Layout la = broadcast_layout_to_shape(a.layout, layout.shape, device().default_order());
        AbiLayout x(layout), y(la);
        check(rstsr_hip_assign(device().raw(), dtype, ptr(), &x, a.dtype, a.ptr(), &y));
    }
    void fill(const void* value) {
        AbiLayout x(layout);
        check(rstsr_hip_fill(device().raw(), dtype, ptr(), &x, value));
    }
    // change_layout_f (tensor/manipulation/to_layout.rs:8-38): no copy when the layout is
    // unchanged, else assign_arbitary into the new layout. assign_arbitary pairs elements in
    // the device's default order (cpu_rayon/assignment.rs:59-70): C order for RowMajor, F order
    // for ColMajor, which is C order over both layouts with their axes reversed.
    Tensor to_layout(const Layout& lnew) const {
        if (lnew.size() != size()) RSTSR_RAISE(InvalidLayout, "to_layout: element counts differ");
        if (lnew == layout) return *this;
        Tensor r = uninit(device(), lnew, dtype);
        const bool f = device().default_order() == FlagOrder::ColMajor;
        Layout lc = f ? lnew.reverse_axes() : lnew, la = f ? layout.reverse_axes() : layout;
        AbiLayout x(lc), y(la);
        check(rstsr_hip_assign_arbitary(device().raw(), dtype, r.ptr(), &x, dtype, ptr(), &y));
        return r;
    }
    Tensor to_contig(FlagOrder o) const {
        return to_layout(o == FlagOrder::RowMajor ? c_layout(layout.shape) : f_layout(layout.shape));
    }
    Tensor astype(rstsr_dtype_t t) const {
        Layout lc = layout_for_array_copy(layout, IterOrder::K);
        Tensor r = uninit(device(), lc, t);
        AbiLayout x(lc), y(layout);
        check(rstsr_hip_assign(device().raw(), t, r.ptr(), &x, dtype, ptr(), &y));
        return r;
    }

    /* ---------- elementwise ---------- */
    Tensor binary(rstsr_binop_t op, const Tensor& b) const {
        if (dtype != b.dtype) RSTSR_RAISE(InvalidValue, "operands must share one dtype");
        device().order_after(b.device());
        FlagOrder ord = device().default_order();
        auto bl = broadcast_layout(layout, b.layout, ord);
        Layout lc = get_layout_for_binary_op(bl.first, bl.second, ord);
        Tensor c = uninit(device(), lc, dtype);
        AbiLayout xc(lc), xa(bl.first), xb(bl.second);
        check(rstsr_hip_op_mutc_refa_refb(device().raw(), op, dtype, c.ptr(), &xc, ptr(), &xa, b.ptr(), &xb));
        return c;
    }
    // A complex and a real operand (Complex<T> op T, T op Complex<T>): a complex result laid out
    // as for same-type operands.
    Tensor binary_mixed(rstsr_binop_t op, const Tensor& b) const {
        device().order_after(b.device());
        FlagOrder ord = device().default_order();
        auto bl = broadcast_layout(layout, b.layout, ord);
        Layout lc = get_layout_for_binary_op(bl.first, bl.second, ord);
        rstsr_dtype_t tc = (dtype == RSTSR_C64 || dtype == RSTSR_C128) ? dtype : b.dtype;
        Tensor c = uninit(device(), lc, tc);
        AbiLayout xc(lc), xa(bl.first), xb(bl.second);
        check(rstsr_hip_op_mutc_mixed(device().raw(), op, tc, c.ptr(), &xc, dtype, ptr(), &xa, b.dtype, b.ptr(), &xb));
        return c;
    }
    // tensor op scalar (or scalar op tensor) with the scalar of dtype ts: complex tensor and real
    // scalar, or real tensor and complex scalar
    Tensor binary_scalar_mixed(rstsr_binop_t op, const void* s, rstsr_dtype_t ts, bool scalar_left) const {
        Layout lc = layout_for_array_copy(layout, IterOrder::K);
        rstsr_dtype_t tc = (dtype == RSTSR_C64 || dtype == RSTSR_C128) ? dtype : ts;
        Tensor c = uninit(device(), lc, tc);
        AbiLayout xc(lc), xa(layout);
        if (scalar_left)
            check(rstsr_hip_op_mutc_mixed(device().raw(), op, tc, c.ptr(), &xc, ts, s, nullptr, dtype, ptr(), &xa));
        else
            check(rstsr_hip_op_mutc_mixed(device().raw(), op, tc, c.ptr(), &xc, dtype, ptr(), &xa, ts, s, nullptr));
        return c;
    }
    Tensor binary_scalar(rstsr_binop_t op, const void* s, bool scalar_left) const {
        Layout lc = layout_for_array_copy(layout, IterOrder::K);
        Tensor c = uninit(device(), lc, dtype);
        AbiLayout xc(lc), xa(layout);
        if (scalar_left)
            check(rstsr_hip_op_mutc_numa_refb(device().raw(), op, dtype, c.ptr(), &xc, s, ptr(), &xa));
        else
            check(rstsr_hip_op_mutc_refa_numb(device().raw(), op, dtype, c.ptr(), &xc, ptr(), &xa, s));
        return c;
    }
    void binary_inplace(rstsr_binop_t op, const Tensor& b) {
        if (dtype != b.dtype) RSTSR_RAISE(InvalidValue, "operands must share one dtype");
        device().order_after(b.device());
        Layout lb = broadcast_layout_to_shape(b.layout, layout.shape, device().default_order());
        AbiLayout xc(layout), xb(lb);
        check(rstsr_hip_op_muta_refb(device().raw(), op, dtype, ptr(), &xc, b.ptr(), &xb));
    }
    void binary_inplace_scalar(rstsr_binop_t op, const void* s) {
        AbiLayout xc(layout);
        check(rstsr_hip_op_muta_numb(device().raw(), op, dtype, ptr(), &xc, s));
    }

    /* ---------- unary (tensor/operators/op_unary_common.rs:141-190) ---------- */
    // &tensor path: a new output laid out as layout_for_array_copy(self, K)
    Tensor unary(rstsr_unop_t op) const {
        rstsr_dtype_t to;
        check(rstsr_hip_unary_out_dtype(op, dtype, &to));
        Layout lc = layout_for_array_copy(layout, IterOrder::K);
        Tensor c = uninit(device(), lc, to);
        AbiLayout xc(lc), xa(layout);
        check(rstsr_hip_unary(device().raw(), op, to, c.ptr(), &xc, dtype, ptr(), &xa));
        return c;
    }
    // op_muta_refb into a caller-given output view (any layout of the same shape)
    void unary_to(rstsr_unop_t op, Tensor& out) const {
        device().order_after(out.device());
        AbiLayout xc(out.layout), xa(layout);
        check(rstsr_hip_unary(device().raw(), op, out.dtype, out.ptr(), &xc, dtype, ptr(), &xa));
    }
    // owned-tensor path: op_muta on the tensor's own storage (same-type ops only)
    void unary_inplace(rstsr_unop_t op) {
        AbiLayout xc(layout);
        check(rstsr_hip_unary_inplace(device().raw(), op, dtype, ptr(), &xc));
    }

    /* ---------- reductions ---------- */
    Layout reduce_out_layout(const std::vector<ptrdiff_t>& axes) const {
        auto sp = layout.dim_split_axes(axes);
        return layout_for_array_copy(sp.second, IterOrder::K);
    }
    // Element type of a reduction's result: u64 for argmax/argmin, the real part
    // type for var/std/l2_norm of complex (T::Real), else T.
    rstsr_dtype_t reduce_out_dtype(rstsr_redop_t op) const {
        if (op == RSTSR_RED_ARGMAX || op == RSTSR_RED_ARGMIN) return RSTSR_U64;
        if (op == RSTSR_RED_ALL || op == RSTSR_RED_ANY) return RSTSR_BOOL;
        if (op == RSTSR_RED_COUNT_NONZERO || (op == RSTSR_RED_SUM && dtype == RSTSR_BOOL)) return RSTSR_U64;
        if (op == RSTSR_RED_VAR || op == RSTSR_RED_STD || op == RSTSR_RED_L2NORM) {
            if (dtype == RSTSR_C64) return RSTSR_F32;
            if (dtype == RSTSR_C128) return RSTSR_F64;
        }
        return dtype;
    }
    Tensor reduce_axes(rstsr_redop_t op, const std::vector<ptrdiff_t>& axes) const {
        Layout lo = reduce_out_layout(axes);
        bool arg = op == RSTSR_RED_ARGMAX || op == RSTSR_RED_ARGMIN;
        rstsr_dtype_t to = reduce_out_dtype(op);
        Tensor out = uninit(device(), lo, to);
        AbiLayout xa(layout), xo(lo);
        if (arg)
            check(rstsr_hip_argreduce_axes(device().raw(), op, dtype, ptr(), &xa, axes.size(), axes.data(),
                                           (uint64_t*)out.ptr(), &xo));
        else
            check(rstsr_hip_reduce_axes(device().raw(), op, dtype, ptr(), &xa, axes.size(), axes.data(), out.ptr(), &xo));
        return out;
    }
    // host value bytes (dtype-sized; u64 for arg)
    std::string reduce_all(rstsr_redop_t op) const {
        AbiLayout xa(layout);
        bool arg = op == RSTSR_RED_ARGMAX || op == RSTSR_RED_ARGMIN;
        std::string out(dtype_size(reduce_out_dtype(op)), '\0');
        if (arg)
            check(rstsr_hip_argreduce_all(device().raw(), op, dtype, ptr(), &xa, (uint64_t*)&out[0]));
        else
            check(rstsr_hip_reduce_all(device().raw(), op, dtype, ptr(), &xa, &out[0]));
        return out;
    }
    bool allclose(const Tensor& b, double rtol, double atol, bool equal_nan) const {
        if (dtype != b.dtype) RSTSR_RAISE(InvalidValue, "operands must share one dtype");
        device().order_after(b.device());
        // allclose_all_f (tensor/reduction.rs:324-351) broadcasts in the device's default order
        auto bl = broadcast_layout(layout, b.layout, device().default_order());
        AbiLayout xa(bl.first), xb(bl.second);
        int ok = 0;
        check(rstsr_hip_allclose(device().raw(), dtype, ptr(), &xa, b.ptr(), &xb, rtol, atol, equal_nan ? 1 : 0, &ok));
        return ok != 0;
    }

    /* ---------- matmul ---------- */
    // The output layout follows the device's default order (layout/matmul.rs:423-463); the
    // device call for ColMajor reverses every layout's axes and swaps a and b, as DeviceFaer's
    // matmul does (device_faer/matmul.rs:227-237).
    Tensor matmul(const Tensor& b, rstsr_gemm_mode_t mode) const {
        if (dtype != b.dtype) RSTSR_RAISE(InvalidValue, "operands must share one dtype");
        device().order_after(b.device());
        const FlagOrder ord = device().default_order();
        Layout lc = matmul_output_layout(layout, b.layout, ord);
        Tensor c = uninit(device(), lc, dtype);
        const bool f = ord == FlagOrder::ColMajor;
        const Layout lcd = f ? lc.reverse_axes() : lc;
        const Layout lad = f ? b.layout.reverse_axes() : layout;
        const Layout lbd = f ? layout.reverse_axes() : b.layout;
        const void* pa = f ? b.ptr() : ptr();
        const void* pb = f ? ptr() : b.ptr();
        AbiLayout xc(lcd), xa(lad), xb(lbd);
        unsigned char one[16] = {0}, zero[16] = {0};
        switch (dtype) {
            case RSTSR_F32: { float v = 1; std::memcpy(one, &v, 4); break; }
            case RSTSR_F64: { double v = 1; std::memcpy(one, &v, 8); break; }
            case RSTSR_I32: { int32_t v = 1; std::memcpy(one, &v, 4); break; }
            case RSTSR_I64: { int64_t v = 1; std::memcpy(one, &v, 8); break; }
            case RSTSR_C64: { float v[2] = {1, 0}; std::memcpy(one, v, 8); break; }
            case RSTSR_C128: { double v[2] = {1, 0}; std::memcpy(one, v, 16); break; }
            default: RSTSR_RAISE(UnImplemented, "matmul dtype");
        }
        check(rstsr_hip_matmul(device().raw(), dtype, c.ptr(), &xc, pa, &xa, pb, &xb, one, zero, mode));
        return c;
    }
};

}  // namespace front
}  // namespace rstsr
