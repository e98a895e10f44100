/*
 * greenpeas_amd.hpp -- C++ facade over the C-ABI in greenpeas_amd.h.
 *
 * Mirrors the reference's C++ driver (include/GreenPeas/QEC/ErrorAnalysis/
 * Driver.hpp:37-219): a Driver is sized once from a "max" circuit
 * (fromStimCircuit, :216-218) and compiles any circuit that fits
 * (compileDetectorErrorModel(circuit), :199-204) or the circuit it holds
 * (compileDetectorErrorModel(), :208-211), reusing its device buffers.
 * Errors come back as the reference's exception classes:
 * std::invalid_argument (unsupported gates, bad arguments) and
 * std::runtime_error (capacity, flow-rule violations, device errors).
 *
 * Circuits are stim circuit text (stim.h is not needed: str() of a
 * stim::Circuit is that text); the DEM comes back as CSR arrays in the
 * reference's error order, with str() giving stim DEM text.
 *
 * Header-only; link against libgreenpeas_amd.so.
 */
#ifndef GREENPEAS_AMD_HPP
#define GREENPEAS_AMD_HPP

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "greenpeas_amd.h"

namespace greenpeas_amd {

/* CorrelationLevel (Mixer.hpp:13-20). */
enum class CorrelationLevel : int { L0 = GP_L0, L1 = GP_L1, L2 = GP_L2 };

/* Throws the reference's exception class for a C-ABI status code. */
inline void check(int rc) {
  if (rc == GP_OK) return;
  if (rc == GP_INVALID_ARGUMENT) throw std::invalid_argument(gp_last_error());
  throw std::runtime_error(gp_last_error());
}

/* A detector error model: error e has probability probs[e] and targets
 * targets[offsets[e] .. offsets[e + 1]); a target t < numDetectors is
 * detector t, otherwise observable t - numDetectors. */
struct DetectorErrorModel {
  std::vector<double> probs;
  std::vector<uint32_t> offsets{0};
  std::vector<uint32_t> targets;
  uint32_t numDetectors = 0, numObservables = 0;

  size_t numErrors() const { return probs.size(); }

  /* stim DEM text: "error(p) D.. L.." per error. */
  std::string str() const {
    std::string out;
    char buf[64];
    for (size_t e = 0; e < probs.size(); ++e) {
      int n = std::snprintf(buf, sizeof buf, "error(%.17g)", probs[e]);
      out.append(buf, static_cast<size_t>(n));
      for (uint32_t k = offsets[e]; k < offsets[e + 1]; ++k) {
        const uint32_t t = targets[k];
        n = t < numDetectors ? std::snprintf(buf, sizeof buf, " D%u", t)
                             : std::snprintf(buf, sizeof buf, " L%u", t - numDetectors);
        out.append(buf, static_cast<size_t>(n));
      }
      out.push_back('\n');
    }
    return out;
  }
};

class Driver {
 public:
  /* Driver::fromStimCircuit (Driver.hpp:216-218): parses the circuit and
   * sizes every device buffer for it.  `stream` (a hipStream_t, nullptr =
   * the driver's own) queues the driver's work on a caller's stream. */
  static Driver fromStimCircuit(const std::string &circuit, CorrelationLevel level = CorrelationLevel::L2,
                                int device = 0, void *stream = nullptr) {
    gp_driver *h = nullptr;
    check(gp_get_driver_on_stream(circuit.c_str(), static_cast<int>(level), device, stream, &h));
    return Driver(h);
  }

  Driver(Driver &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  Driver &operator=(Driver &&o) noexcept {
    if (this != &o) {
      if (h_) gp_driver_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }
  Driver(const Driver &) = delete;
  Driver &operator=(const Driver &) = delete;
  ~Driver() {
    if (h_) gp_driver_destroy(h_);
  }

  /* Driver::compileDetectorErrorModel(const stim::Circuit &)
   * (Driver.hpp:199-204). */
  DetectorErrorModel compileDetectorErrorModel(const std::string &circuit) {
    check(gp_compile_detector_error_model(h_, circuit.c_str()));
    return dem();
  }

  /* Driver::compileDetectorErrorModel() (Driver.hpp:208-211): the circuit
   * or STEPG the driver holds. */
  DetectorErrorModel compileDetectorErrorModel() {
    check(gp_compile_detector_error_model(h_, nullptr));
    return dem();
  }

  /* A caller's STEPG in the reference's layout (gp_driver_load_stepg),
   * compiled by the next compileDetectorErrorModel(). */
  void loadStepg(const gp_stepg &stepg) { check(gp_driver_load_stepg(h_, &stepg)); }

  /* Later work on `stream` (nullptr: the driver's own). */
  void setStream(void *stream) { check(gp_driver_set_stream(h_, stream)); }

  gp_driver *handle() const { return h_; }

 private:
  explicit Driver(gp_driver *h) : h_(h) {}

  DetectorErrorModel dem() const {
    uint32_t ne = 0, nd = 0, no = 0;
    uint64_t nt = 0;
    check(gp_dem_shape(h_, &ne, &nt, &nd, &no));
    DetectorErrorModel d;
    d.probs.resize(ne);
    d.offsets.resize(static_cast<size_t>(ne) + 1);
    d.targets.resize(nt);
    d.numDetectors = nd;
    d.numObservables = no;
    check(gp_dem_copy(h_, d.probs.data(), d.offsets.data(), d.targets.data(), nullptr));
    return d;
  }

  gp_driver *h_ = nullptr;
};

/* get_driver(circuit, correlation_level=L2) (ErrorAnalysis.pybind.cu:72-78). */
inline Driver getDriver(const std::string &circuit, CorrelationLevel level = CorrelationLevel::L2,
                        int device = 0) {
  return Driver::fromStimCircuit(circuit, level, device);
}

}  // namespace greenpeas_amd

#endif /* GREENPEAS_AMD_HPP */
