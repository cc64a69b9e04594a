// py_bindings.cpp — Python module `lattica_hw` for the MI355X backend.
//
// The names and argument orders are those the reference's module exposes
// (example_impl/py_bindings.cpp:87-133: <op>_32 / <op>_64 / <op>_float64 and
// the classes DeviceTensor32 / DeviceTensor64 / DeviceTensorfloat64), so its
// Python dispatcher can load this module unchanged.  Added: mod_tt / mod_tc /
// mod_ct (declared by the reference API but not bound there), the transcript
// operations the reference runtime names without defining (modneg, abs,
// modmul_axis_sum, set_const_val, moveaxis, get_slice, pad_single_axis), and
// backend controls (set_stream, synchronize, plan_cache_size, capture records).
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <torch/extension.h>

#include "include/lattica_hw_api.h"

namespace py = pybind11;
using namespace lattica_hw_api;

template <typename T>
static void bind_tensor_class(py::module_& m, const std::string& suffix) {
    using DT = DeviceTensor<T>;
    py::class_<DT, std::shared_ptr<DT>>(m, ("DeviceTensor" + suffix).c_str())
        .def("print", &DT::print)
        .def("print_metadata", &DT::print_metadata)
        .def("reshape", &DT::reshape)
        .def_property_readonly("shape", [](const DT& t) { return t.dims; })
        .def_property_readonly("strides", [](const DT& t) { return t.strides; })
        .def_property_readonly("offset", [](const DT& t) { return t.offset; })
        .def_property_readonly("buffer_id", [](const DT& t) { return t.buffer ? t.buffer->id : uint64_t(0); })
        .def_property_readonly("buffer_version", [](const DT& t) { return t.buffer ? t.buffer->version : uint64_t(0); })
        .def("is_contiguous", &DT::is_contiguous);
}

template <typename T>
static void bind_memory(py::module_& m, const std::string& s) {
    m.def(("allocate_on_hardware_" + s).c_str(), &allocate_on_hardware<T>, py::arg("dims"));
    m.def(("host_to_device_" + s).c_str(), &host_to_device<T>, py::arg("tensor"));
    m.def(("device_to_host_" + s).c_str(), &device_to_host<T>, py::arg("device_mem"));
    m.def(("expand_" + s).c_str(), &expand<T>, py::arg("tensor"), py::arg("axis"), py::arg("repeats"));
    m.def(("squeeze_" + s).c_str(), &squeeze<T>, py::arg("tensor"), py::arg("axis"));
    m.def(("unsqueeze_" + s).c_str(), &unsqueeze<T>, py::arg("tensor"), py::arg("axis"));
    m.def(("make_contiguous_" + s).c_str(), &make_contiguous<T>, py::arg("tensor"));
    // transcript operations without a reference definition (see lattica_hw_api.h)
    m.def(("set_const_val_" + s).c_str(), &set_const_val<T>, py::arg("tensor"), py::arg("value"));
    m.def(("moveaxis_" + s).c_str(), &moveaxis<T>, py::arg("tensor"), py::arg("source"), py::arg("destination"));
    m.def(("pad_single_axis_" + s).c_str(), &pad_single_axis<T>, py::arg("tensor"), py::arg("axis"),
          py::arg("before"), py::arg("after"));
    m.def(("alias_view_" + s).c_str(), &alias_view<T>, py::arg("tensor"), py::arg("dims"), py::arg("strides"),
          py::arg("offset"));
    m.def(("copy_from_host_" + s).c_str(), &copy_from_host<T>, py::arg("device_tensor"), py::arg("tensor"));
    // spec: one (start, stop, step, is_index) per leading dim, already normalised
    m.def(
        ("get_slice_" + s).c_str(),
        [](const std::shared_ptr<DeviceTensor<T>>& t, const std::vector<std::tuple<int64_t, int64_t, int64_t, bool>>& spec) {
            std::vector<SliceSpec> v;
            for (const auto& e : spec) {
                SliceSpec sp;
                sp.start = std::get<0>(e);
                sp.stop = std::get<1>(e);
                sp.step = std::get<2>(e);
                sp.index = std::get<3>(e);
                v.push_back(sp);
            }
            return get_slice<T>(t, v);
        },
        py::arg("tensor"), py::arg("spec"));
}

template <typename T>
static void bind_compute(py::module_& m, const std::string& s) {
    m.def(("modmul_ttt_" + s).c_str(), &modmul_ttt<T>);
    m.def(("modmul_ttc_" + s).c_str(), &modmul_ttc<T>);
    m.def(("modmul_tct_" + s).c_str(), &modmul_tct<T>);
    m.def(("modmul_tcc_" + s).c_str(), &modmul_tcc<T>);
    m.def(("modsum_ttt_" + s).c_str(), &modsum_ttt<T>);
    m.def(("modsum_ttc_" + s).c_str(), &modsum_ttc<T>);
    m.def(("modsum_tct_" + s).c_str(), &modsum_tct<T>);
    m.def(("modsum_tcc_" + s).c_str(), &modsum_tcc<T>);
    m.def(("mod_tt_" + s).c_str(), &mod_tt<T>);
    m.def(("mod_tc_" + s).c_str(), &mod_tc<T>);
    m.def(("mod_ct_" + s).c_str(), &mod_ct<T>);
    m.def(("axis_modsum_" + s).c_str(), &axis_modsum<T>);
    m.def(("g_decomposition_" + s).c_str(), &g_decomposition<T>, py::arg("a"), py::arg("result"), py::arg("power"),
          py::arg("base_bits"));
    m.def(("ntt_" + s).c_str(), &ntt<T>);
    m.def(("intt_" + s).c_str(), &intt<T>);
    m.def(("permute_" + s).c_str(), &permute<T>);
    // transcript operations without a reference definition (see lattica_hw_api.h)
    m.def(("modneg_tt_" + s).c_str(), &modneg_tt<T>);
    m.def(("modneg_tc_" + s).c_str(), &modneg_tc<T>);
    m.def(("modmul_axis_sum_" + s).c_str(), &modmul_axis_sum<T>, py::arg("a"), py::arg("b"), py::arg("p"),
          py::arg("result"), py::arg("axis"));
    m.def(("abs_" + s).c_str(), &abs_values<T>);
}

PYBIND11_MODULE(lattica_hw, m) {
    m.doc() = "HEAL hardware API on MI355X (gfx950)";
    bind_tensor_class<int32_t>(m, "32");
    bind_tensor_class<int64_t>(m, "64");
    bind_tensor_class<double>(m, "float64");
    bind_memory<int32_t>(m, "32");
    bind_memory<int64_t>(m, "64");
    bind_memory<double>(m, "float64");
    bind_compute<int32_t>(m, "32");
    bind_compute<int64_t>(m, "64");
    m.def("set_stream", [](uintptr_t s) { set_stream(reinterpret_cast<void*>(s)); }, py::arg("stream"));
    m.def("get_stream", []() { return reinterpret_cast<uintptr_t>(get_stream()); });
    m.def("synchronize", &synchronize);
    m.def("plan_cache_size", &plan_cache_size);
    // graph capture: the record returned by capture_end keeps the plans a
    // captured graph uses; hold it as long as the graph
    struct PinRecord {
        std::shared_ptr<CapturePins> pins;
    };
    py::class_<PinRecord>(m, "CapturePins");
    m.def("capture_begin", &capture_begin);
    m.def("capture_end", []() { return PinRecord{capture_end()}; });
    m.def("captured_plan_count", &captured_plan_count);
}
