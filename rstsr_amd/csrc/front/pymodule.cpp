// pybind11 module `rstsr_amd._rstsr`: exposes the C++ front end (tensor.hpp)
// and the host layout algebra (layout.hpp) to Python.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "tensor.hpp"

namespace py = pybind11;
using namespace rstsr;
using namespace rstsr::front;

using PyLayout = std::tuple<std::vector<size_t>, std::vector<ptrdiff_t>, size_t>;

static Layout L(const PyLayout& t) { return Layout(std::get<0>(t), std::get<1>(t), std::get<2>(t)); }
static PyLayout P(const Layout& l) { return PyLayout(l.shape, l.stride, l.offset); }

static IterOrder iter_order(const std::string& s) {
    if (s == "C") return IterOrder::C;
    if (s == "F") return IterOrder::F;
    if (s == "A") return IterOrder::A;
    if (s == "K") return IterOrder::K;
    if (s == "G") return IterOrder::G;
    if (s == "B") return IterOrder::B;
    RSTSR_RAISE(InvalidValue, "unknown iteration order " + s);
}
static FlagOrder flag_order(const std::string& s) {
    if (s == "C" || s == "RowMajor") return FlagOrder::RowMajor;
    if (s == "F" || s == "ColMajor") return FlagOrder::ColMajor;
    RSTSR_RAISE(InvalidValue, "unknown order " + s);
}

PYBIND11_MODULE(_rstsr, m) {
    m.doc() = "RSTSR MI355X device backend: C++ front end over the C ABI";
    static py::exception<rstsr::Error> exc(m, "RSTSRError");
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const rstsr::Error& e) {
            py::tuple args = py::make_tuple(std::string(error_kind_name(e.kind)), std::string(e.what()));
            PyErr_SetObject(exc.ptr(), args.ptr());
        }
    });

    m.def("device_count", [] {
        int n = 0;
        check(rstsr_hip_device_count(&n));
        return n;
    });

    py::class_<DeviceHIP>(m, "DeviceHIP")
        .def(py::init<int>(), py::arg("device_id") = 0)
        .def("synchronize", &DeviceHIP::synchronize)
        .def("same_device", &DeviceHIP::same_device)
        .def_property_readonly("device_id", [](const DeviceHIP& d) { return d.ctx->device; })
        // DeviceBaseAPI (storage/device.rs:3-7): a clone shares the GPU context and carries its own
        // default order; tensors keep the order of the device value they were created with
        .def("clone", [](const DeviceHIP& d) { return d; })
        .def("set_default_order", [](DeviceHIP& d, const std::string& o) { d.set_default_order(flag_order(o)); })
        .def_property_readonly("default_order", [](const DeviceHIP& d) {
            return std::string(d.default_order() == FlagOrder::RowMajor ? "RowMajor" : "ColMajor");
        })
        .def("_ctx_ptr", [](const DeviceHIP& d) { return (uintptr_t)d.raw(); });

    py::class_<Tensor>(m, "Tensor")
        .def_property_readonly("shape", [](const Tensor& t) { return t.layout.shape; })
        .def_property_readonly("stride", [](const Tensor& t) { return t.layout.stride; })
        .def_property_readonly("offset", [](const Tensor& t) { return t.layout.offset; })
        .def_property_readonly("layout", [](const Tensor& t) { return P(t.layout); })
        .def_property_readonly("dtype_code", [](const Tensor& t) { return (int)t.dtype; })
        .def_property_readonly("ndim", &Tensor::ndim)
        .def_property_readonly("size", &Tensor::size)
        .def_property_readonly("device", [](const Tensor& t) { return t.device(); })
        .def("_data_ptr", [](const Tensor& t) { return (uintptr_t)t.ptr(); })
        .def("_nbytes", [](const Tensor& t) { return t.storage->nbytes; })
        .def("to_device", &Tensor::to_device)
        .def("_raw_range", [](const Tensor& t) {
            auto r = t.raw_range();
            return py::make_tuple(py::bytes(r.first), r.second);
        })
        .def("_with_layout", [](const Tensor& t, const PyLayout& l) {
            Layout x = L(l);
            (void)x.bounds_index();
            return t.view_with(x);
        })
        .def("assign", &Tensor::assign)
        .def("fill", [](Tensor& t, py::bytes v) {
            std::string s = v;
            if (s.size() != t.esize()) RSTSR_RAISE(InvalidValue, "fill value size");
            t.fill(s.data());
        })
        .def("to_layout", [](const Tensor& t, const PyLayout& l) { return t.to_layout(L(l)); })
        .def("to_contig", [](const Tensor& t, const std::string& o) { return t.to_contig(flag_order(o)); },
             py::arg("order") = "C")
        .def("astype", [](const Tensor& t, int code) { return t.astype((rstsr_dtype_t)code); })
        .def("binary", [](const Tensor& a, int op, const Tensor& b) { return a.binary((rstsr_binop_t)op, b); })
        .def("binary_mixed",
             [](const Tensor& a, int op, const Tensor& b) { return a.binary_mixed((rstsr_binop_t)op, b); })
        .def("binary_scalar_mixed",
             [](const Tensor& a, int op, py::bytes v, int ts, bool left) {
                 std::string s = v;
                 return a.binary_scalar_mixed((rstsr_binop_t)op, s.data(), (rstsr_dtype_t)ts, left);
             })
        .def("binary_scalar",
             [](const Tensor& a, int op, py::bytes s, bool left) {
                 std::string v = s;
                 if (v.size() != a.esize()) RSTSR_RAISE(InvalidValue, "scalar size");
                 return a.binary_scalar((rstsr_binop_t)op, v.data(), left);
             })
        .def("binary_inplace", [](Tensor& a, int op, const Tensor& b) { a.binary_inplace((rstsr_binop_t)op, b); })
        .def("binary_inplace_scalar",
             [](Tensor& a, int op, py::bytes s) {
                 std::string v = s;
                 if (v.size() != a.esize()) RSTSR_RAISE(InvalidValue, "scalar size");
                 a.binary_inplace_scalar((rstsr_binop_t)op, v.data());
             })
        .def("reduce_axes",
             [](const Tensor& a, int op, const std::vector<ptrdiff_t>& axes) {
                 return a.reduce_axes((rstsr_redop_t)op, axes);
             })
        .def("reduce_all", [](const Tensor& a, int op) { return py::bytes(a.reduce_all((rstsr_redop_t)op)); })
        .def("unary", [](const Tensor& a, int op) { return a.unary((rstsr_unop_t)op); })
        .def("unary_inplace", [](Tensor& a, int op) { a.unary_inplace((rstsr_unop_t)op); })
        .def("unary_to", [](const Tensor& a, int op, Tensor& out) { a.unary_to((rstsr_unop_t)op, out); })
        .def("allclose", &Tensor::allclose)
        .def("matmul", [](const Tensor& a, const Tensor& b, int mode) { return a.matmul(b, (rstsr_gemm_mode_t)mode); });

    m.def("from_host", [](const DeviceHIP& dev, py::bytes data, const std::vector<size_t>& shape, int code) {
        std::string s = data;
        size_t n = 1;
        for (auto d : shape) n *= d;
        if (s.size() != n * dtype_size((rstsr_dtype_t)code)) RSTSR_RAISE(InvalidValue, "host buffer size does not match shape");
        return Tensor::from_host(dev, s.data(), shape, (rstsr_dtype_t)code);
    });
    m.def("empty", [](const DeviceHIP& dev, const PyLayout& l, int code) {
        return Tensor::uninit(dev, L(l), (rstsr_dtype_t)code);
    });

    // ---- host layout algebra (no GPU needed) ----
    m.def("greedy_layout", [](const PyLayout& l, bool keep) {
        auto r = greedy_layout(L(l), keep);
        return py::make_tuple(P(r.first), r.second);
    });
    m.def("layout_for_array_copy",
          [](const PyLayout& l, const std::string& o) { return P(layout_for_array_copy(L(l), iter_order(o))); });
    m.def("get_layout_for_binary_op", [](const PyLayout& a, const PyLayout& b, const std::string& o) {
        return P(get_layout_for_binary_op(L(a), L(b), flag_order(o)));
    });
    m.def("broadcast_layout", [](const PyLayout& a, const PyLayout& b, const std::string& o) {
        auto r = broadcast_layout(L(a), L(b), flag_order(o));
        return py::make_tuple(P(r.first), P(r.second));
    });
    m.def("broadcast_layout_to_shape", [](const PyLayout& l, const std::vector<size_t>& shape, const std::string& o) {
        return P(broadcast_layout_to_shape(L(l), shape, flag_order(o)));
    });
    m.def("matmul_output_layout", [](const PyLayout& a, const PyLayout& b, const std::string& o) {
        return P(matmul_output_layout(L(a), L(b), flag_order(o)));
    });
    m.def("dim_narrow", [](const PyLayout& l, ptrdiff_t axis, py::object start, py::object stop, ptrdiff_t step) {
        bool hs = !start.is_none(), he = !stop.is_none();
        return P(L(l).dim_narrow(axis, hs, hs ? start.cast<ptrdiff_t>() : 0, he, he ? stop.cast<ptrdiff_t>() : 0, step));
    });
    m.def("bounds_index", [](const PyLayout& l) { return L(l).bounds_index(); });
    m.def("c_contig", [](const PyLayout& l) { return L(l).c_contig(); });
    m.def("f_contig", [](const PyLayout& l) { return L(l).f_contig(); });
    m.def("c_prefer", [](const PyLayout& l) { return L(l).c_prefer(); });
    m.def("f_prefer", [](const PyLayout& l) { return L(l).f_prefer(); });
    m.def("axes_composition", [](const PyLayout& l) {
        auto r = get_axes_composition(L(l));
        return py::make_tuple(r.i, r.b, r.c, r.d);
    });
    m.def("dim_split_axes", [](const PyLayout& l, const std::vector<ptrdiff_t>& axes) {
        auto r = L(l).dim_split_axes(axes);
        return py::make_tuple(P(r.first), P(r.second));
    });
    m.def("c_layout", [](const std::vector<size_t>& s) { return P(c_layout(s)); });
    m.def("f_layout", [](const std::vector<size_t>& s) { return P(f_layout(s)); });
}
