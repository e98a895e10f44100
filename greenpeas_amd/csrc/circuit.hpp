// circuit.hpp -- stim circuit text -> flat instruction list.
//
// The reference receives a parsed stim::Circuit and walks it with
// stim's for_each_operation (Circuit.hpp:471-478), dispatching on the gate
// type (Circuit.hpp:425-462).  stim is not part of this library, so the
// text is parsed here.  Gate subset = the reference's (TICK, R, M(p), CX, H,
// DEPOLARIZE1, DEPOLARIZE2, X_ERROR, DETECTOR, OBSERVABLE_INCLUDE) plus
// aliases (RZ, MZ, CNOT, ZCX, H_XZ), MR/MRZ (= M then R), REPEAT blocks
// (flattened, as for_each_operation does) and the coordinate annotations
// QUBIT_COORDS / SHIFT_COORDS (ignored).  Anything else raises
// InvalidArgument, like Circuit.hpp:457-460.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace gpa {

// std::invalid_argument / std::runtime_error carry the reference's error
// classes through the C-ABI (GP_INVALID_ARGUMENT / GP_RUNTIME_ERROR).
struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum class OpKind : uint8_t { Tick, R, M, CX, H, Dep1, Dep2, XErr, Detector, Observable };

struct Op {
  OpKind kind;
  double arg;   // probability; observable index for Observable
  uint32_t t0;  // first target in Circuit::targets
  uint32_t nt;  // number of targets
};

// Flattened circuit.  Qubit targets are stored as non-negative ints, record
// targets rec[-k] as -k.
struct Circuit {
  std::vector<Op> ops;
  std::vector<int32_t> targets;
  uint32_t numQubits = 0;
  uint32_t numTicks = 0;
  uint32_t numMeasurements = 0;
  uint32_t numDetectors = 0;
  uint32_t numObservables = 0;
  // One per DEPOLARIZE1 / X_ERROR / M target and one per DEPOLARIZE2 pair.
  uint32_t numLocations = 0;

  static Circuit parse(std::string_view text);
};

// Sizes read off circuit text in one pass, without parsing it (for memory
// admission): layers (TICK count + 1, REPEAT bodies counted per
// repetition), qubits (highest qubit target + 1), detectors (per
// repetition) and observables (highest OBSERVABLE_INCLUDE index + 1).
struct CircuitSummary {
  uint64_t layers = 1, qubits = 0, detectors = 0, observables = 0;
};
CircuitSummary summarize(std::string_view text);

}  // namespace gpa
