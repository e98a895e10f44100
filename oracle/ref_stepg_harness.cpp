// ref_stepg_harness.cpp -- test infrastructure only.
//
// Builds the reference's own stim-free headers (Mixer.hpp, STEPG.hpp,
// Graph.hpp, Vector.hpp, Storage/Host.hpp, Array.hpp) in place from
// /root/reference/include, with a minimal line reader of ours standing in
// for stim's parser (the reference's Circuit.hpp needs stim.h, which is not
// available).  The resulting library oracle/_ref/libgp_refstepg.so lets the
// tests check the oracle's STEPG restatement (adjacency, node probabilities,
// measurement keys) and its FNV-1a-64 against the reference code itself.
//
// Gate subset: TICK, R, M(p), CX, H, DEPOLARIZE1, DEPOLARIZE2, X_ERROR;
// DETECTOR / OBSERVABLE_INCLUDE / QUBIT_COORDS lines are skipped (the
// detector bookkeeping lives in Circuit.hpp).  No REPEAT blocks.

#include <cstdint>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "GreenPeas/Core/Array.hpp"
#include "GreenPeas/QEC/ErrorAnalysis/Mixer.hpp"
#include "GreenPeas/QEC/ErrorAnalysis/STEPG.hpp"

namespace {

struct Op {
  std::string name;
  double arg = 0;
  std::vector<uint32_t> targets;
};

std::string g_err;

std::vector<Op> readOps(const char *text, uint32_t &numQubits, uint32_t &numTicks) {
  std::vector<Op> ops;
  std::istringstream in(text);
  std::string line;
  numQubits = 0;
  numTicks = 0;
  while (std::getline(in, line)) {
    auto hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ls(line);
    std::string head;
    if (!(ls >> head)) continue;
    Op op;
    auto paren = head.find('(');
    op.name = head.substr(0, paren);
    if (paren != std::string::npos) op.arg = std::stod(head.substr(paren + 1));
    if (op.name == "DETECTOR" || op.name == "OBSERVABLE_INCLUDE" || op.name == "QUBIT_COORDS" ||
        op.name == "SHIFT_COORDS")
      continue;
    if (op.name == "TICK") numTicks++;
    std::string tok;
    while (ls >> tok) {
      uint32_t q = static_cast<uint32_t>(std::stoul(tok));
      op.targets.push_back(q);
      if (q + 1 > numQubits) numQubits = q + 1;
    }
    ops.push_back(std::move(op));
  }
  return ops;
}

template <gp::CorrelationLevel Level>
void build(const std::vector<Op> &ops, gp::STEPG &stepg, std::vector<uint32_t> &keys,
           uint32_t numQubits, uint32_t numLayers) {
  using M = gp::Mixer<Level>;
  M::initialise(stepg, numQubits, numLayers);
  uint32_t layer = 0;
  for (const auto &op : ops) {
    const auto &t = op.targets;
    if (op.name == "TICK") {
      layer++;
    } else if (op.name == "R") {
      for (auto q : t) M::applyR(stepg, q, layer);
    } else if (op.name == "M") {
      for (auto q : t) keys.push_back(M::applyXError(stepg, q, layer, op.arg));
    } else if (op.name == "CX") {
      for (size_t i = 0; i + 1 < t.size(); i += 2) M::applyCX(stepg, t[i], t[i + 1], layer);
    } else if (op.name == "H") {
      for (auto q : t) M::applyH(stepg, q, layer);
    } else if (op.name == "DEPOLARIZE1") {
      for (auto q : t) M::applyDepolarize1(stepg, q, layer, op.arg);
    } else if (op.name == "DEPOLARIZE2") {
      for (size_t i = 0; i + 1 < t.size(); i += 2)
        M::applyDepolarize2(stepg, t[i], t[i + 1], layer, op.arg);
    } else if (op.name == "X_ERROR") {
      for (auto q : t) M::applyXError(stepg, q, layer, op.arg);
    } else {
      throw std::invalid_argument("Unsupported operation: " + op.name);
    }
  }
}

struct RefStepg {
  uint32_t numLayers, numNodesPerLayer, numNodes;
  std::vector<uint64_t> adjacency;
  std::vector<double> probs;
  std::vector<uint32_t> keys;
};

}  // namespace

extern "C" {

const char *gpr_last_error() { return g_err.c_str(); }

// Builds the STEPG of `text` with the reference's Mixer at `level`.
// Returns an opaque handle or nullptr (see gpr_last_error).
void *gpr_build_stepg(const char *text, int level) {
  try {
    uint32_t nq, nt;
    auto ops = readOps(text, nq, nt);
    const uint32_t numLayers = nt + 1;
    const uint32_t npq = level == 0 ? 3 : level == 1 ? 5 : 8;
    gp::STEPG stepg(numLayers, nq * npq);
    auto *out = new RefStepg;
    switch (level) {
    case 0: build<gp::CorrelationLevel::L0>(ops, stepg, out->keys, nq, numLayers); break;
    case 1: build<gp::CorrelationLevel::L1>(ops, stepg, out->keys, nq, numLayers); break;
    default: build<gp::CorrelationLevel::L2>(ops, stepg, out->keys, nq, numLayers); break;
    }
    out->numLayers = stepg.numLayers;
    out->numNodesPerLayer = stepg.numNodesPerLayer;
    out->numNodes = stepg.numNodes;
    out->adjacency.resize(stepg.numNodes);
    out->probs.resize(stepg.numNodes);
    for (uint32_t n = 0; n < stepg.numNodes; ++n) {
      out->adjacency[n] = stepg.graph[n];
      out->probs[n] = stepg.probs[n];
    }
    return out;
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

// dims[0..3] = layers, nodes per layer, nodes, measurements
void gpr_dims(void *h, uint32_t *dims) {
  auto *s = static_cast<RefStepg *>(h);
  dims[0] = s->numLayers;
  dims[1] = s->numNodesPerLayer;
  dims[2] = s->numNodes;
  dims[3] = static_cast<uint32_t>(s->keys.size());
}

void gpr_copy(void *h, uint64_t *adjacency, double *probs, uint32_t *keys) {
  auto *s = static_cast<RefStepg *>(h);
  std::memcpy(adjacency, s->adjacency.data(), s->adjacency.size() * 8);
  std::memcpy(probs, s->probs.data(), s->probs.size() * 8);
  std::memcpy(keys, s->keys.data(), s->keys.size() * 4);
}

void gpr_free(void *h) { delete static_cast<RefStepg *>(h); }

// The reference's Array<uint32_t, 32>::hash64 over 32 slots.
uint64_t gpr_hash64_slots32(const uint32_t *slots) {
  gp::Array<uint32_t, 32> a;
  for (int i = 0; i < 32; ++i) a[i] = slots[i];
  return a.hash64();
}

double gpr_split_dep1(double p) {
  gp::STEPG stepg(2, 8);
  gp::Mixer<gp::CorrelationLevel::L2>::applyDepolarize1(stepg, 0, 1, p);
  return stepg.probs[0 * 8 + 3];  // Y node of qubit 0, layer 0
}

}  // extern "C"
