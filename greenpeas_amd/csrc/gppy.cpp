// gppy.cpp -- pybind11 module greenpeas_amd._gppy.
//
// The reference's extension module _gppy (source/GreenPeasPy.cu:8-17) has an
// error_analysis submodule (source/QEC/ErrorAnalysis/ErrorAnalysis.pybind.cu:
// 47-79) holding CorrelationLevel, Driver and get_driver; this module has the
// same submodule, names, overloads and defaults, over the C++ facade
// (include/greenpeas_amd.hpp).  Differences: a circuit is stim circuit text
// or any object whose str() is (a stim.Circuit, a pathlib.Path to a .stim
// file), and the DEM is a greenpeas_amd.DetectorErrorModel (stim text via
// str(), to_stim() where stim is installed).  Compiles release the GIL; a
// Driver is used by one thread at a time.  std::invalid_argument arrives as
// ValueError and std::runtime_error as RuntimeError, as in the reference.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <string>
#include <utility>

#include "../../include/greenpeas_amd.hpp"

namespace py = pybind11;
using greenpeas_amd::CorrelationLevel;
using greenpeas_amd::DetectorErrorModel;
using greenpeas_amd::Driver;

namespace {

std::string circuitFromPython(const py::object &circuit) {
  if (py::isinstance<py::str>(circuit) || py::isinstance<py::bytes>(circuit)) return circuit.cast<std::string>();
  if (py::isinstance(circuit, py::module_::import("os").attr("PathLike"))) {
    py::object f = py::module_::import("builtins").attr("open")(circuit);
    std::string text = f.attr("read")().cast<std::string>();
    f.attr("close")();
    return text;
  }
  return py::str(circuit).cast<std::string>();
}

template <typename T>
py::array_t<T> toArray(std::vector<T> &&v) {
  auto *owned = new std::vector<T>(std::move(v));
  py::capsule free(owned, [](void *p) { delete static_cast<std::vector<T> *>(p); });
  return py::array_t<T>({owned->size()}, {sizeof(T)}, owned->data(), free);
}

py::object demToPython(DetectorErrorModel &&d) {
  py::object cls = py::module_::import("greenpeas_amd.dem").attr("DetectorErrorModel");
  const uint32_t nd = d.numDetectors, no = d.numObservables;
  return cls(toArray(std::move(d.probs)), toArray(std::move(d.offsets)), toArray(std::move(d.targets)), nd, no);
}

}  // namespace

PYBIND11_MODULE(_gppy, m) {
  m.doc() = "greenpeas_amd extension module (MI355X DEM compiler)";
  auto ea = m.def_submodule("error_analysis", "Error analysis driver.");

  py::enum_<CorrelationLevel>(ea, "CorrelationLevel")
      .value("L0", CorrelationLevel::L0)
      .value("L1", CorrelationLevel::L1)
      .value("L2", CorrelationLevel::L2)
      .export_values();

  py::class_<Driver>(ea, "Driver")
      .def(
          "compile_detector_error_model",
          [](Driver &d, const py::object &circuit) {
            const std::string text = circuitFromPython(circuit);
            DetectorErrorModel dem;
            {
              py::gil_scoped_release nogil;
              dem = d.compileDetectorErrorModel(text);
            }
            return demToPython(std::move(dem));
          },
          py::arg("circuit"))
      .def("compile_detector_error_model", [](Driver &d) {
        DetectorErrorModel dem;
        {
          py::gil_scoped_release nogil;
          dem = d.compileDetectorErrorModel();
        }
        return demToPython(std::move(dem));
      });

  ea.def(
      "get_driver",
      [](const py::object &circuit, CorrelationLevel correlation_level) {
        const std::string text = circuitFromPython(circuit);
        py::gil_scoped_release nogil;
        return Driver::fromStimCircuit(text, correlation_level);
      },
      py::arg("circuit"), py::arg("correlation_level") = CorrelationLevel::L2);
}
