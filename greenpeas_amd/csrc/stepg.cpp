// stepg.cpp -- STEPG builder (see stepg.hpp).
#include "stepg.hpp"

#include "workers.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <cstring>
#include <unordered_map>
#include <cmath>
#include <string>

namespace gpa {
namespace {

// Node slots within a qubit (Mixer.hpp:59-78).
enum Slot : uint32_t { kX = 0, kZ = 1, kXXZZ = 2, kY = 3, kXZZX = 4, kXYYX = 5, kYY = 6, kYZZY = 7 };

// One correlated two-qubit Pauli component, as data: the auxiliary node sits
// on qubit `host` (0 = first, 1 = second of the pair) in `slot`, `back`
// layers before the channel, and flows into two targets given as (qubit of
// the pair, slot, layer offset from the channel layer: 0 or -1).  The target
// order is the reference's edge insertion order (Mixer.hpp:305-533), which
// fixes which child lands in the low/high half of the adjacency word.
struct PairRule {
  uint8_t pa, pb;  // Pauli on q0, q1: 1 X, 2 Y, 3 Z
  uint8_t minLevel;
  uint8_t host, slot, back;
  uint8_t t0q, t0slot;
  int8_t t0dl;
  uint8_t t1q, t1slot;
  int8_t t1dl;
  const char *name;
};

constexpr PairRule kPairRules[] = {
    {1, 1, 0, 0, kXXZZ, 1, 0, kX, 0, 1, kX, 0, "XX"},
    {3, 3, 0, 1, kXXZZ, 1, 0, kZ, 0, 1, kZ, 0, "ZZ"},
    {1, 3, 1, 0, kXZZX, 1, 0, kX, 0, 1, kZ, 0, "XZ"},
    {3, 1, 1, 1, kXZZX, 1, 0, kZ, 0, 1, kX, 0, "ZX"},
    {1, 2, 2, 0, kXYYX, 2, 0, kX, 0, 1, kY, -1, "XY"},
    {2, 1, 2, 1, kXYYX, 2, 0, kY, -1, 1, kX, 0, "YX"},
    {2, 2, 2, 0, kYY, 2, 0, kY, -1, 1, kY, -1, "YY"},
    {2, 3, 2, 1, kYZZY, 2, 0, kY, -1, 1, kZ, 0, "YZ"},
    {3, 2, 2, 0, kYZZY, 2, 0, kZ, 0, 1, kY, -1, "ZY"},
};

// kPairRules index of each (pa, pb) with both non-identity, at 4 * pa + pb.
struct PairRuleIndex {
  uint8_t at[16] = {};
  constexpr PairRuleIndex() {
    for (uint8_t i = 0; i < sizeof(kPairRules) / sizeof(kPairRules[0]); ++i)
      at[4 * kPairRules[i].pa + kPairRules[i].pb] = i;
  }
};
constexpr PairRuleIndex kPairRuleIndex{};
constexpr uint32_t pairRuleOf(int pa, int pb) { return kPairRuleIndex.at[4 * pa + pb]; }
constexpr bool pairRuleIndexComplete() {
  for (int pa = 1; pa < 4; ++pa)
    for (int pb = 1; pb < 4; ++pb) {
      const PairRule &r = kPairRules[pairRuleOf(pa, pb)];
      if (r.pa != pa || r.pb != pb) return false;
    }
  return true;
}
static_assert(pairRuleIndexComplete(), "a two-qubit Pauli without its rule");

struct Builder {
  Stepg &g;
  bool record;

  uint32_t node(uint32_t q, uint32_t slot, uint32_t l) const { return l * g.npl + q * g.k + slot; }

  // Graph::addEdge / STEPG::addFlow (Graph.hpp:132-156, STEPG.hpp:86-101).
  void addFlow(uint32_t src, uint32_t sl, uint32_t dst, uint32_t dl) {
    if (dl <= sl) throw RuntimeError("STEPG: flows must move forward in time.");
    if (sl >= g.numLayers || dl >= g.numLayers)
      throw RuntimeError("STEPG: graph.addEdge failed with status 3.");
    uint64_t &a = g.adjacency[src];
    const uint32_t lo = static_cast<uint32_t>(a), hi = static_cast<uint32_t>(a >> 32);
    if (lo == dst || hi == dst) throw RuntimeError("STEPG: graph.addEdge failed with status 1.");
    if (lo == kNone) a = (a & 0xFFFFFFFF00000000ull) | dst;
    else if (hi == kNone) a = (a & 0xFFFFFFFFull) | (static_cast<uint64_t>(dst) << 32);
    else throw RuntimeError("STEPG: graph.addEdge failed with status 2.");
  }

  // Graph::removeEdge / STEPG::removeFlow (Graph.hpp:162-182, STEPG.hpp:108-123).
  void removeFlow(uint32_t src, uint32_t sl, uint32_t dst, uint32_t dl) {
    if (dl <= sl) throw RuntimeError("STEPG: flows must move forward in time.");
    if (sl >= g.numLayers || dl >= g.numLayers)
      throw RuntimeError("STEPG: graph.removeEdge failed with status 3.");
    uint64_t &a = g.adjacency[src];
    if (static_cast<uint32_t>(a) == dst) a |= 0xFFFFFFFFull;
    else if (static_cast<uint32_t>(a >> 32) == dst) a |= 0xFFFFFFFF00000000ull;
    else throw RuntimeError("STEPG: graph.removeEdge failed with status 4.");
  }

  // An auxiliary node already wired to exactly these two children (the same
  // Pauli component from another channel in the same layer) is reused: its
  // probability is merged and no flow is added.  The reference would throw
  // on the duplicate edge (Graph.hpp:141-143); any other overlap still does.
  bool wiredTo(uint32_t aux, uint32_t c0, uint32_t c1) const {
    const uint64_t a = g.adjacency[aux];
    return static_cast<uint32_t>(a) == c0 && static_cast<uint32_t>(a >> 32) == c1;
  }

  // STEPG::mergeProbabilities (STEPG.hpp:128-131), same operand order so the
  // node probabilities are bit-identical to the reference's.
  void merge(uint32_t n, double p, uint32_t loc, uint8_t kind) {
    double &old = g.probs[n];
    old = (1 - p) * old + (1 - old) * p;
    if (record) {
      g.compNode.push_back(n);
      g.compLoc.push_back(loc);
      g.compKind.push_back(kind);
    }
  }

  void single(int pauli, uint32_t q, uint32_t l, double p, uint32_t loc, uint8_t kind) {
    if (pauli == 1) {
      merge(node(q, kX, l), p, loc, kind);
    } else if (pauli == 3) {
      merge(node(q, kZ, l), p, loc, kind);
    } else {  // Y (Mixer.hpp:222-235), dropped at L0
      if (g.level < 1) return;
      if (l < 1) throw RuntimeError("Mixer: cannot apply Y error in layer l < 1");
      const uint32_t y = node(q, kY, l - 1);
      if (!wiredTo(y, node(q, kX, l), node(q, kZ, l))) {
        addFlow(y, l - 1, node(q, kX, l), l);
        addFlow(y, l - 1, node(q, kZ, l), l);
      }
      merge(y, p, loc, kind);
    }
  }

  void pair(int pa, int pb, uint32_t q0, uint32_t q1, uint32_t l, double p, uint32_t loc) {
    if (pa == 0) return single(pb, q1, l, p, loc, 2);
    if (pb == 0) return single(pa, q0, l, p, loc, 2);
    {
      const PairRule &r = kPairRules[pairRuleOf(pa, pb)];
      if (g.level < r.minLevel) return;
      if (l < r.back)
        throw RuntimeError(std::string("Mixer: cannot apply ") + r.name + " error in layer l < " +
                           std::to_string(r.back));
      const uint32_t qs[2] = {q0, q1};
      const uint32_t al = l - r.back;
      const uint32_t aux = node(qs[r.host], r.slot, al);
      const uint32_t c0 = node(qs[r.t0q], r.t0slot, l + r.t0dl);
      const uint32_t c1 = node(qs[r.t1q], r.t1slot, l + r.t1dl);
      if (!wiredTo(aux, c0, c1)) {
        addFlow(aux, al, c0, l + r.t0dl);
        addFlow(aux, al, c1, l + r.t1dl);
      }
      merge(aux, p, loc, 2);
      return;
    }
  }
};

}  // namespace

namespace {

// Position in the op list: the layer, and the measurements, detectors and
// fault locations met so far.
struct OpCursor {
  uint32_t layer = 0, meas = 0, det = 0, loc = 0;
};

// Advances the cursor over one op without touching the graph.
void countOp(const Op &o, OpCursor &cur) {
  switch (o.kind) {
  case OpKind::Tick: cur.layer++; break;
  case OpKind::M: cur.meas += o.nt; cur.loc += o.nt; break;
  case OpKind::XErr:
  case OpKind::Dep1: cur.loc += o.nt; break;
  case OpKind::Dep2: cur.loc += o.nt / 2; break;
  case OpKind::Detector: cur.det++; break;
  default: break;
  }
}

// Applies ops [begin, end) from cursor `cur` (Mixer.hpp:97-533,
// Circuit.hpp:290-328); DETECTOR / OBSERVABLE_INCLUDE records only when
// `records`.
void applyOps(const Circuit &c, size_t begin, size_t end, OpCursor cur, Builder &b, const double *theta,
              bool records) {
  Stepg &g = b.g;
  uint32_t &layer = cur.layer, &meas = cur.meas, &det = cur.det, &loc = cur.loc;
  // per-component rates of DEPOLARIZE1/2 (Mixer.hpp:164, 184), memoized on
  // the channel rate (circuits repeat few rates)
  double last1 = -1, last1s = 0, last2 = -1, last2s = 0;
  auto split1 = [&](double p) {
    if (p != last1) {
      last1 = p;
      last1s = 0.5 - 0.5 * std::sqrt(1 - (4 * p) / 3);
    }
    return last1s;
  };
  auto split2 = [&](double p) {
    if (p != last2) {
      last2 = p;
      last2s = 0.5 - 0.5 * std::pow(1 - (16 * p) / 15, 0.125);
    }
    return last2s;
  };
  auto rate = [&](double p) {
    if (!theta) return p;
    return theta[loc];
  };
  for (size_t oi = begin; oi < end; ++oi) {
    const Op &o = c.ops[oi];
    const int32_t *t = c.targets.data() + o.t0;
    switch (o.kind) {
    case OpKind::Tick:
      layer++;
      break;
    case OpKind::R:  // Mixer.hpp:114-120
      // Gates in the last layer have no later layer to flow into: no-op
      // (the reference throws there, Graph.hpp:163-165).
      if (layer + 1 >= g.numLayers) break;
      for (uint32_t j = 0; j < o.nt; ++j) {
        const uint32_t q = t[j];
        b.removeFlow(b.node(q, kX, layer), layer, b.node(q, kX, layer + 1), layer + 1);
        b.removeFlow(b.node(q, kZ, layer), layer, b.node(q, kZ, layer + 1), layer + 1);
      }
      break;
    case OpKind::CX:  // Mixer.hpp:127-136
      if (layer + 1 >= g.numLayers) break;
      for (uint32_t j = 0; j < o.nt; j += 2) {
        const uint32_t cq = t[j], tq = t[j + 1];
        b.addFlow(b.node(cq, kX, layer), layer, b.node(tq, kX, layer + 1), layer + 1);
        b.addFlow(b.node(tq, kZ, layer), layer, b.node(cq, kZ, layer + 1), layer + 1);
      }
      break;
    case OpKind::H:  // Mixer.hpp:142-150
      if (layer + 1 >= g.numLayers) break;
      for (uint32_t j = 0; j < o.nt; ++j) {
        const uint32_t q = t[j];
        b.removeFlow(b.node(q, kX, layer), layer, b.node(q, kX, layer + 1), layer + 1);
        b.removeFlow(b.node(q, kZ, layer), layer, b.node(q, kZ, layer + 1), layer + 1);
        b.addFlow(b.node(q, kX, layer), layer, b.node(q, kZ, layer + 1), layer + 1);
        b.addFlow(b.node(q, kZ, layer), layer, b.node(q, kX, layer + 1), layer + 1);
      }
      break;
    case OpKind::M:  // Circuit.hpp:316-319: X error before the measurement + key
      for (uint32_t j = 0; j < o.nt; ++j) {
        const double p = rate(o.arg);
        b.single(1, t[j], layer, p, loc, 0);
        g.keys[meas++] = b.node(t[j], kX, layer);
        loc++;
      }
      break;
    case OpKind::XErr:  // Mixer.hpp:209-214
      for (uint32_t j = 0; j < o.nt; ++j) {
        b.single(1, t[j], layer, rate(o.arg), loc, 0);
        loc++;
      }
      break;
    case OpKind::Dep1:  // Mixer.hpp:158-169: X, Z, Y with the per-channel split
      for (uint32_t j = 0; j < o.nt; ++j) {
        const double p = rate(o.arg);
        if (!(p >= 0 && p <= 0.75)) throw RuntimeError("Mixer: invalid DEP1 p > 0.75.");
        const double sp = split1(p);
        b.single(1, t[j], layer, sp, loc, 1);
        b.single(3, t[j], layer, sp, loc, 1);
        b.single(2, t[j], layer, sp, loc, 1);
        loc++;
      }
      break;
    case OpKind::Dep2:  // Mixer.hpp:178-201: 15 components in IX..ZZ order
      for (uint32_t j = 0; j < o.nt; j += 2) {
        const double p = rate(o.arg);
        if (!(p >= 0 && p <= 0.9375)) throw RuntimeError("Mixer: invalid DEP2 p > 0.9375.");
        const double sp = split2(p);
        for (int pa = 0; pa < 4; ++pa)
          for (int pb = 0; pb < 4; ++pb)
            if (pa || pb) b.pair(pa, pb, t[j], t[j + 1], layer, sp, loc);
        loc++;
      }
      break;
    case OpKind::Detector:
    case OpKind::Observable: {
      // Detectors take bits 0..D-1 in order; OBSERVABLE_INCLUDE(k) takes bit
      // D + k.  Records are XORed (a repeated record cancels).
      const uint32_t bit =
          o.kind == OpKind::Detector ? det++ : g.numDetectors + static_cast<uint32_t>(o.arg);
      if (!records) break;
      for (uint32_t j = 0; j < o.nt; ++j) {
        const uint32_t row = meas + static_cast<uint32_t>(t[j]);  // t[j] = -k
        g.measBits[static_cast<size_t>(row) * g.numCols + bit / 32] ^= 1u << (bit % 32);
      }
      break;
    }
    }
  }
}

// Circuits below this many target entries are built on one thread.
constexpr size_t kParallelBuildMinTargets = 20000;

}  // namespace

void buildStepg(const Circuit &c, int level, Stepg &g, const double *theta, bool recordComponents,
                bool parallel) {
  g.level = level;
  g.k = nodesPerQubit(level);
  g.numLayers = c.numTicks + 1;
  g.numQubits = c.numQubits;
  g.npl = g.numQubits * g.k;
  const uint64_t nn = static_cast<uint64_t>(g.numLayers) * g.npl;
  if (nn >= (1ull << 31)) throw RuntimeError("circuit too large: more than 2^31 STEPG nodes");
  if (g.npl >= (1u << 30)) throw RuntimeError("circuit too large: too many nodes per layer");
  g.numNodes = static_cast<uint32_t>(nn);
  g.numMeasurements = c.numMeasurements;
  g.numDetectors = c.numDetectors;
  g.numObservables = c.numObservables;
  g.numLocations = c.numLocations;
  g.numCols = (g.numDetectors + g.numObservables + 31) / 32;
  // node arrays: no child, no probability (filled in parallel when large)
  g.adjacency.resize(g.numNodes);
  g.probs.resize(g.numNodes);
  {
    const uint32_t fparts = parallel && g.numNodes >= (1u << 18) ? 4 * parallelWidth() : 1u;
    parallelFor(fparts, [&](uint32_t q) {
      const size_t a = static_cast<size_t>(g.numNodes) * q / fparts, e = static_cast<size_t>(g.numNodes) * (q + 1) / fparts;
      std::fill(g.adjacency.begin() + a, g.adjacency.begin() + e, ~0ull);
      std::fill(g.probs.begin() + a, g.probs.begin() + e, 0.0);
    });
  }
  g.keys.assign(g.numMeasurements, 0);
  g.measBits.assign(static_cast<size_t>(g.numMeasurements) * g.numCols, 0u);
  g.compNode.clear();
  g.compLoc.clear();
  g.compKind.clear();

  // Every node is written by the ops of one layer only: a frame node (X, Z)
  // by its own layer's gates and errors, an auxiliary node by the layer its
  // slot's offset points back from (Y one layer, the pair slots a fixed 0
  // or 2).  So layer ranges are built in parallel, each node's merges
  // keeping their order, with the same results as one pass; detector
  // records (which reach back to earlier measurements) follow in one pass.
  // The component list of a plan keeps circuit order: one pass.
  const bool split = parallel && !recordComponents && c.targets.size() >= kParallelBuildMinTargets &&
                     parallelWidth() > 1 && g.numLayers > 8;
  const uint32_t L = g.numLayers, Q = g.numQubits;
  if (!split) {
    Builder b{g, recordComponents};
    // persistence flows X->X, Z->Z between consecutive layers (Mixer.hpp:97-108)
    for (uint32_t l = 0; l + 1 < L; ++l)
      for (uint32_t q = 0; q < Q; ++q) {
        b.addFlow(b.node(q, kX, l), l, b.node(q, kX, l + 1), l + 1);
        b.addFlow(b.node(q, kZ, l), l, b.node(q, kZ, l + 1), l + 1);
      }
    applyOps(c, 0, c.ops.size(), OpCursor{}, b, theta, true);
    return;
  }
  const uint32_t chunks = std::min<uint32_t>(4 * parallelWidth(), L);
  parallelFor(chunks, [&](uint32_t k) {
    Builder b{g, false};
    const uint32_t l0 = static_cast<uint32_t>(static_cast<uint64_t>(L - 1) * k / chunks);
    const uint32_t l1 = static_cast<uint32_t>(static_cast<uint64_t>(L - 1) * (k + 1) / chunks);
    for (uint32_t l = l0; l < l1; ++l)
      for (uint32_t q = 0; q < Q; ++q) {
        b.addFlow(b.node(q, kX, l), l, b.node(q, kX, l + 1), l + 1);
        b.addFlow(b.node(q, kZ, l), l, b.node(q, kZ, l + 1), l + 1);
      }
  });
  // op ranges of about equal target counts, cut at TICKs, with the cursor
  // at each cut
  std::vector<size_t> cut{0};
  std::vector<OpCursor> at{OpCursor{}};
  {
    const size_t per = c.targets.size() / chunks + 1;
    OpCursor cur;
    size_t acc = 0;
    for (size_t oi = 0; oi < c.ops.size(); ++oi) {
      const Op &o = c.ops[oi];
      if (o.kind == OpKind::Tick && acc >= per) {
        cut.push_back(oi);
        at.push_back(cur);
        acc = 0;
      }
      acc += o.nt;
      countOp(o, cur);
    }
    cut.push_back(c.ops.size());
  }
  parallelFor(static_cast<uint32_t>(cut.size() - 1), [&](uint32_t k) {
    Builder b{g, false};
    applyOps(c, cut[k], cut[k + 1], at[k], b, theta, false);
  });
  // detector and observable records
  {
    OpCursor cur;
    for (const Op &o : c.ops) {
      if (o.kind == OpKind::Detector || o.kind == OpKind::Observable) {
        const uint32_t bit = o.kind == OpKind::Detector ? cur.det : g.numDetectors + static_cast<uint32_t>(o.arg);
        const int32_t *t = c.targets.data() + o.t0;
        for (uint32_t j = 0; j < o.nt; ++j) {
          const uint32_t row = cur.meas + static_cast<uint32_t>(t[j]);
          g.measBits[static_cast<size_t>(row) * g.numCols + bit / 32] ^= 1u << (bit % 32);
        }
      }
      countOp(o, cur);
    }
  }
}

namespace {
std::atomic<uint32_t> gSweepMinRpt{1};
}

void setSweepMinRpt(uint32_t rpt) {
  if (rpt != 1 && rpt != 2 && rpt != 4 && rpt != 8)
    throw InvalidArgument("sweep_rpt must be 1, 2, 4 or 8");
  gSweepMinRpt.store(rpt);
}
uint32_t sweepMinRpt() { return gSweepMinRpt.load(); }

std::atomic<uint32_t> gProbTable{1};
void setProbTable(bool on) { gProbTable.store(on ? 1u : 0u); }
bool probTable() { return gProbTable.load() != 0; }

namespace {
// knob index: bucket_fill, bucket_cap, sweep_rpt
int knobIndex(const char *name) {
  const std::string n = name ? name : "";
  if (n == "bucket_fill") return 0;
  if (n == "bucket_cap") return 1;
  if (n == "sweep_rpt") return 2;
  throw InvalidArgument("no per-size-class tuning for knob " + n);
}
std::mutex gClassMu;
std::map<uint32_t, std::array<uint32_t, 3>> gClassKnobs;  // level << 8 | size class -> values (0 = unset)
}  // namespace

uint32_t sizeClassOf(uint64_t numNodes) {
  uint32_t c = 0;
  while (c < 63 && (numNodes >> (c + 1)) != 0) ++c;
  return c;
}

void setClassTuning(int level, uint32_t sizeClass, const char *name, uint32_t value) {
  const int k = knobIndex(name);
  nodesPerQubit(level);  // validates the level
  if (sizeClass > 63) throw InvalidArgument("size class must be at most 63");
  if (value) {
    if (k == 0 && (value < 1 || value > 4096)) throw InvalidArgument("bucket_fill must be in 1..4096");
    if (k == 1 && value != 64 && value != 128 && value != 256)
      throw InvalidArgument("bucket_cap must be 64, 128 or 256");
    if (k == 2 && value != 1 && value != 2 && value != 4 && value != 8)
      throw InvalidArgument("sweep_rpt must be 1, 2, 4 or 8");
  }
  std::lock_guard<std::mutex> lk(gClassMu);
  auto &e = gClassKnobs[static_cast<uint32_t>(level) << 8 | sizeClass];
  e[k] = value;
  if (!e[0] && !e[1] && !e[2]) gClassKnobs.erase(static_cast<uint32_t>(level) << 8 | sizeClass);
}

void clearClassTuning() {
  std::lock_guard<std::mutex> lk(gClassMu);
  gClassKnobs.clear();
}

uint32_t classTuning(const char *name, int level, uint64_t numNodes, uint32_t fallback) {
  const int k = knobIndex(name);
  std::lock_guard<std::mutex> lk(gClassMu);
  if (gClassKnobs.empty()) return fallback;
  auto it = gClassKnobs.find(static_cast<uint32_t>(level) << 8 | sizeClassOf(numNodes));
  return it != gClassKnobs.end() && it->second[k] ? it->second[k] : fallback;
}

namespace {

// Node-indexed scratch of buildImage, kept per thread between builds (a
// bb12 image needs ~4 MB of it; fresh pages each time cost more than the
// passes that use them).
struct ImageScratch {
  std::vector<uint8_t> active, frameSlot;
  std::vector<uint32_t> seedOf, actIdx;
};
thread_local ImageScratch tScratch;
thread_local std::vector<uint32_t> tLocalOdd, tLocalEven;

// Splits [0, n) into `parts` contiguous ranges; range k is [lo, hi).
inline void rangeOf(uint64_t n, uint32_t parts, uint32_t k, uint64_t &lo, uint64_t &hi) {
  lo = n * k / parts;
  hi = n * (k + 1) / parts;
}

}  // namespace

void buildImage(const Stepg &g, StepgImage &img, bool parallel) {
  const uint32_t N = g.numNodes, npl = g.npl, L = g.numLayers, C = g.numCols;
  const uint32_t k = nodesPerQubit(g.level);
  img.numLayers = L;
  img.npl = npl;
  img.numCols = C;
  img.numDetectors = g.numDetectors;
  img.numObservables = g.numObservables;
  img.numLocations = g.numLocations;
  img.level = g.level;
  img.rowSpan = 0;
  ImageScratch &ws = tScratch;
  // work split over the worker pool (a small circuit takes one part)
  const uint32_t parts =
      parallel && N >= 65536 ? std::min<uint32_t>(4 * parallelWidth(), std::max<uint32_t>(L, 1)) : 1u;
  // slot j of a layer is a frame node (X or Z of its qubit)
  std::vector<uint8_t> &frameSlot = ws.frameSlot;
  frameSlot.resize(npl);
  for (uint32_t j = 0; j < npl; ++j) frameSlot[j] = (j % k) < 2;

  // ranks (p > 0 nodes, node order): counted and listed by node range
  std::vector<uint32_t> partR(parts + 1, 0);
  parallelFor(parts, [&](uint32_t q) {
    uint64_t lo, hi;
    rangeOf(N, parts, q, lo, hi);
    uint32_t r = 0;
    for (uint64_t n = lo; n < hi; ++n) r += g.probs[n] > 0;
    partR[q + 1] = r;
  });
  for (uint32_t q = 0; q < parts; ++q) partR[q + 1] += partR[q];
  const uint32_t R = partR[parts];
  img.numRanks = R;
  img.rankNode.resize(R);
  img.rankLayer.resize(R);
  img.rankProb.resize(R);
  parallelFor(parts, [&](uint32_t q) {
    uint64_t lo, hi;
    rangeOf(N, parts, q, lo, hi);
    uint32_t r = partR[q];
    for (uint64_t n = lo; n < hi; ++n) {
      const double p = g.probs[n];
      if (p > 0) {
        img.rankNode[r] = static_cast<uint32_t>(n);
        img.rankLayer[r] = static_cast<uint32_t>(n / npl);
        img.rankProb[r] = p;
        ++r;
      }
    }
  });
  // active nodes: ranks and everything their flows reach.  Children lie in
  // the next one or two layers, so each range of at least two layers takes
  // its own closure in node order (in parallel), sending the marks that
  // leave it to the next range; then, range by range in order, marks that
  // arrive at a node not yet active are carried on from there.
  std::vector<uint8_t> &active = ws.active;
  active.resize(N);
  const uint32_t aparts = std::max<uint32_t>(1, std::min<uint32_t>(parts, L / 2));
  std::vector<uint64_t> aEnd(aparts);  // first node past each range
  std::vector<std::vector<uint32_t>> outMarks(aparts);
  parallelFor(aparts, [&](uint32_t q) {
    uint64_t l0, l1;
    rangeOf(L, aparts, q, l0, l1);
    const uint64_t n0 = l0 * npl, n1 = l1 * npl;
    aEnd[q] = n1;
    std::fill(active.begin() + n0, active.begin() + n1, uint8_t{0});
  });
  parallelFor(aparts, [&](uint32_t q) {
    const uint64_t n0 = q ? aEnd[q - 1] : 0, n1 = aEnd[q];
    std::vector<uint32_t> &out = outMarks[q];
    for (uint64_t n = n0; n < n1; ++n) {
      if (!active[n] && !(g.probs[n] > 0)) continue;
      active[n] = 1;
      const uint64_t a = g.adjacency[n];
      for (uint32_t c : {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)}) {
        if (c == kNone) continue;
        if (c < n1) active[c] = 1;
        else out.push_back(c);
      }
    }
  });
  for (uint32_t q = 1; q < aparts; ++q) {
    const uint64_t n1 = aEnd[q];
    std::vector<uint32_t> work;
    for (uint32_t c : outMarks[q - 1])
      if (!active[c]) {
        active[c] = 1;
        work.push_back(c);
      }
    while (!work.empty()) {
      const uint32_t n = work.back();
      work.pop_back();
      const uint64_t a = g.adjacency[n];
      for (uint32_t c : {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)}) {
        if (c == kNone) continue;
        if (c >= n1) {
          outMarks[q].push_back(c);
        } else if (!active[c]) {
          active[c] = 1;
          work.push_back(c);
        }
      }
    }
  }

  // seeds: XOR of the detector bits of every measurement on an active node
  // (seedOf holds kNone everywhere between builds: only seeded nodes are
  // set, and reset before returning)
  std::vector<uint32_t> &seedOf = ws.seedOf;
  if (seedOf.size() < N) seedOf.resize(N, kNone);
  std::vector<uint32_t> seedNode;
  struct ResetSeeds {  // on every exit, errors included
    std::vector<uint32_t> &of;
    const std::vector<uint32_t> &nodes;
    ~ResetSeeds() {
      for (uint32_t n : nodes) of[n] = kNone;
    }
  } resetSeeds{seedOf, seedNode};
  std::vector<uint32_t> seedRows;  // numSeeds x C, row-major
  seedNode.reserve(g.numMeasurements);
  seedRows.reserve(static_cast<size_t>(g.numMeasurements) * C);
  for (uint32_t m = 0; m < g.numMeasurements; ++m) {
    const uint32_t n = g.keys[m];
    if (!active[n]) continue;
    const uint32_t *bits = g.measBits.data() + static_cast<size_t>(m) * C;
    bool any = false;
    for (uint32_t c = 0; c < C; ++c) any |= bits[c] != 0;
    if (!any) continue;
    if (seedOf[n] == kNone) {
      if (!frameSlot[n % npl]) throw RuntimeError("STEPG: a measurement seeds an auxiliary node");
      seedOf[n] = static_cast<uint32_t>(seedNode.size());
      seedNode.push_back(n);
      seedRows.resize(seedRows.size() + C, 0u);
    }
    uint32_t *row = seedRows.data() + static_cast<size_t>(seedOf[n]) * C;
    for (uint32_t c = 0; c < C; ++c) row[c] ^= bits[c];
  }
  const uint32_t S = static_cast<uint32_t>(seedNode.size());
  {
    std::vector<uint8_t> seeded(L, 0);
    for (uint32_t n : seedNode) seeded[n / npl] = 1;
    uint32_t last = 0, gap = 0;
    for (uint32_t l = 0; l < L; ++l)
      if (seeded[l]) {
        gap = std::max(gap, l - last);
        last = l;
      }
    img.seedGap = gap;
  }
  img.colHi.assign(C, -1);
  img.colLow.assign(C, -1);
  for (uint32_t s = 0; s < S; ++s) {
    const int32_t l = static_cast<int32_t>(seedNode[s] / npl);
    const uint32_t *row = seedRows.data() + static_cast<size_t>(s) * C;
    for (uint32_t c = 0; c < C; ++c) {
      if (!row[c]) continue;
      if (img.colHi[c] < l) img.colHi[c] = l;
      if (img.colLow[c] < 0 || img.colLow[c] > l) img.colLow[c] = l;
    }
  }

  // positions of the active frame nodes in their layers, and the structure
  // the sweep relies on: frame nodes flow only into frame nodes of the next
  // layer (layer by layer, in parallel)
  std::vector<uint32_t> &actIdx = ws.actIdx;
  actIdx.assign(N, kNone);
  img.layerAct.assign(L + 1, 0);
  const uint32_t lparts = std::min<uint32_t>(parts, std::max<uint32_t>(L, 1));
  parallelFor(lparts, [&](uint32_t q) {
    uint64_t l0, l1;
    rangeOf(L, lparts, q, l0, l1);
    for (uint64_t l = l0; l < l1; ++l) {
      uint32_t cnt = 0;
      const uint64_t next0 = (l + 1) * npl;
      for (uint32_t j = 0; j < npl; ++j) {
        const uint64_t n = l * npl + j;
        if (!active[n] || !frameSlot[j]) continue;
        actIdx[n] = cnt++;
        const uint64_t a = g.adjacency[n];
        for (uint32_t ch : {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)}) {
          if (ch == kNone) continue;
          if (ch < next0 || ch >= next0 + npl) throw RuntimeError("STEPG: a frame flow skips a layer");
          if (!frameSlot[ch - next0]) throw RuntimeError("STEPG: a frame node flows into an auxiliary node");
        }
      }
      img.layerAct[l + 1] = cnt;
    }
  });
  uint32_t total = 0, maxAct = 0;
  for (uint32_t l = 0; l < L; ++l) {
    const uint32_t cnt = img.layerAct[l + 1];
    img.layerAct[l] = total;
    total += cnt;
    maxAct = std::max(maxAct, cnt);
  }
  img.layerAct[L] = total;
  img.numActive = total;
  img.maxActive = maxAct;

  // sweep shape: the fewest records per thread (a power of two) that fit
  // the layer in one workgroup, then the fewest 64-lane waves
  uint32_t rpt = 1;
  const uint32_t minRpt = classTuning("sweep_rpt", g.level, N, sweepMinRpt());
  while (rpt < minRpt) rpt <<= 1;
  while (rpt < 8 && static_cast<uint64_t>(rpt) * kSweepMaxThreads < maxAct) rpt <<= 1;
  if (static_cast<uint64_t>(rpt) * kSweepMaxThreads < maxAct)
    throw RuntimeError("circuit has too many active STEPG nodes per layer for the LDS sweep (" +
                       std::to_string(maxAct) + " > " + std::to_string(8 * kSweepMaxThreads) + ")");
  const uint32_t threads = std::max<uint32_t>(64, ((maxAct + rpt - 1) / rpt + 63) / 64 * 64);
  const uint32_t P = threads * rpt;
  img.sweepThreads = threads;
  img.sweepRpt = rpt;
  img.layerStride = P;
  const uint32_t numPairs = (L + 1) / 2;
  img.numPairs = numPairs;
  const uint64_t rowsPerCol = 2ull * (numPairs + kSweepLeadPad);
  if (rowsPerCol * P * 4 >= (1ull << 32) || P > (1u << kRowPosBits) || L >= (1u << (32 - kRowPosBits)))
    throw RuntimeError("circuit too large for the sweep: too many layers x frame nodes");

  // records and seeds, pair by pair (in parallel; each pair's seed list is
  // its own, joined in pair order afterwards)
  const uint32_t zeroW = 2u * P;
  img.act.resize(static_cast<size_t>(kSweepLeadPad + numPairs) * (P + 1) * 4);
  for (uint32_t blk = 0; blk < kSweepLeadPad; ++blk) {
    uint32_t *rec = img.act.data() + static_cast<size_t>(blk) * (P + 1) * 4;
    for (uint32_t j = 0; j < P; ++j, rec += 4) {
      rec[0] = rec[1] = rec[2] = zeroW | (zeroW << 16);
      rec[3] = 0xFFFFFFFFu;
    }
    rec[0] = rec[1] = rec[2] = rec[3] = 0u;  // header
  }
  std::vector<std::vector<uint32_t>> pairRows(numPairs);  // each pair's seeds' column words
  parallelFor(numPairs, [&](uint32_t m) {
    std::vector<uint32_t> &newRows = pairRows[m];
    std::vector<uint32_t> &localOdd = tLocalOdd, &localEven = tLocalEven;  // a seeded node's index in the pair's list
    localOdd.resize(npl);
    localEven.resize(npl);
    const uint32_t le = 2 * m, lo = 2 * m + 1;
    const uint32_t e0 = le * npl, o0 = lo * npl;  // first nodes of layers 2m, 2m+1
    uint32_t *rec = img.act.data() + static_cast<size_t>(kSweepLeadPad + m) * (P + 1) * 4;
    for (uint32_t j = 0; j < P; ++j) {
      rec[4 * j] = rec[4 * j + 1] = rec[4 * j + 2] = zeroW | (zeroW << 16);
      rec[4 * j + 3] = 0xFFFFFFFFu;
    }
    uint32_t cnt = 0;
    // the seeds of the two layers, in node order
    for (uint32_t j = 0; j < npl; ++j) {
      const uint32_t n = e0 + j;
      localEven[j] = 0xFFFFu;
      if (actIdx[n] == kNone || seedOf[n] == kNone) continue;
      localEven[j] = cnt++;
      const uint32_t *row = seedRows.data() + static_cast<size_t>(seedOf[n]) * C;
      newRows.insert(newRows.end(), row, row + C);
    }
    for (uint32_t j = 0; j < npl; ++j) {
      localOdd[j] = 0xFFFFu;
      if (lo >= L) continue;
      const uint32_t n = o0 + j;
      if (actIdx[n] == kNone || seedOf[n] == kNone) continue;
      localOdd[j] = cnt++;
      const uint32_t *row = seedRows.data() + static_cast<size_t>(seedOf[n]) * C;
      newRows.insert(newRows.end(), row, row + C);
    }
    rec[4 * P + 1] = rec[4 * P + 2] = rec[4 * P + 3] = 0u;  // header; [4P] = first seed, below
    const uint32_t slotBase = ((m + 1) & 1u) * P;
    auto word = [&](uint32_t ch) { return ch == kNone ? zeroW : slotBase + actIdx[ch]; };
    // odd layer: children in layer 2m + 2
    if (lo < L)
      for (uint32_t j = 0; j < npl; ++j) {
        const uint32_t n = o0 + j;
        if (actIdx[n] == kNone) continue;
        uint32_t *r = rec + 4 * actIdx[n];
        const uint64_t a = g.adjacency[n];
        r[0] = word(static_cast<uint32_t>(a)) | (word(static_cast<uint32_t>(a >> 32)) << 16);
        r[3] = (r[3] & 0xFFFF0000u) | localOdd[j];
      }
    // even layer: grandchildren in layer 2m + 2 and the combined seed
    for (uint32_t j = 0; j < npl; ++j) {
      const uint32_t n = e0 + j;
      if (actIdx[n] == kNone) continue;
      const uint64_t a = g.adjacency[n];
      const uint32_t ch[2] = {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)};
      uint32_t gc[4], ng = 0;
      uint32_t parts3[3], np = 0;  // seeded nodes among n and its children (counted mod 2)
      auto toggle = [](uint32_t *set, uint32_t &cn, uint32_t x, uint32_t cap) {
        for (uint32_t i = 0; i < cn; ++i)
          if (set[i] == x) {
            set[i] = set[--cn];
            return true;
          }
        if (cn == cap) return false;
        set[cn++] = x;
        return true;
      };
      if (localEven[j] != 0xFFFFu) toggle(parts3, np, n, 3);
      for (uint32_t c : ch) {
        if (c == kNone) continue;
        if (localOdd[c - o0] != 0xFFFFu) toggle(parts3, np, c, 3);
        const uint64_t b2 = g.adjacency[c];
        for (uint32_t gch : {static_cast<uint32_t>(b2), static_cast<uint32_t>(b2 >> 32)})
          if (gch != kNone && !toggle(gc, ng, gch, 4))
            throw RuntimeError("STEPG: a node reaches more than four nodes two layers on");
      }
      uint32_t w[4] = {zeroW, zeroW, zeroW, zeroW};
      for (uint32_t i = 0; i < ng; ++i) w[i] = word(gc[i]);
      uint32_t *r = rec + 4 * actIdx[n];
      r[1] = w[0] | (w[1] << 16);
      r[2] = w[2] | (w[3] << 16);
      uint32_t es = 0xFFFFu;
      if (np == 1) {
        es = parts3[0] < o0 ? localEven[parts3[0] - e0] : localOdd[parts3[0] - o0];
      } else if (np > 1) {
        es = cnt++;
        newRows.resize(newRows.size() + C, 0u);
        uint32_t *row = newRows.data() + newRows.size() - C;
        for (uint32_t i = 0; i < np; ++i) {
          const uint32_t *src = seedRows.data() + static_cast<size_t>(seedOf[parts3[i]]) * C;
          for (uint32_t c = 0; c < C; ++c) row[c] ^= src[c];
        }
      }
      r[3] = (r[3] & 0x0000FFFFu) | (es << 16);
    }
    if (cnt >= 0xFFFFu) throw RuntimeError("circuit too large for the sweep: too many seeds in two layers");
  });
  // each pair's first seed in the joined list, and the seeds column-major
  uint32_t S2 = 0;
  for (uint32_t m = 0; m < numPairs; ++m) {
    img.act[(static_cast<size_t>(kSweepLeadPad + m) * (P + 1) + P) * 4] = S2;
    S2 += C ? static_cast<uint32_t>(pairRows[m].size() / C) : 0u;
  }
  img.numSeeds = S2;
  img.seedVals.resize(static_cast<size_t>(C) * S2);
  for (uint32_t m = 0, s2 = 0; m < numPairs; ++m) {
    const uint32_t ns = C ? static_cast<uint32_t>(pairRows[m].size() / C) : 0u;
    for (uint32_t i = 0; i < ns; ++i, ++s2)
      for (uint32_t c = 0; c < C; ++c)
        img.seedVals[static_cast<size_t>(c) * S2 + s2] = pairRows[m][static_cast<size_t>(i) * C + c];
  }

  // each rank's frame rows (itself, or the frame nodes its flows reach
  // through at most one auxiliary node, each counted modulo 2), encoded
  // compactly (see kSpillBase), and its probability's index in the table of
  // distinct values: by rank range in parallel, each range with its own
  // spill entries and distinct values (in first-use order), joined in range
  // order so the image is the one a single pass makes
  img.rankCode.resize(2ull * R);
  img.rankSpill.clear();
  img.probTab.clear();
  img.probIdx.clear();
  const bool useTable = probTable();
  if (useTable) img.probIdx.resize(R);
  const uint32_t rparts = R >= 4096 ? parts : 1u;
  struct RankPart {
    std::vector<uint32_t> spill;    // 4 codes per spilled rank
    std::vector<uint32_t> spilled;  // those ranks
    std::vector<double> values;     // distinct probabilities, first-use order
    uint32_t rowSpan = 0;
  };
  std::vector<RankPart> rp(rparts);
  parallelFor(rparts, [&](uint32_t q) {
    RankPart &part = rp[q];
    uint64_t r0, r1;
    rangeOf(R, rparts, q, r0, r1);
    std::unordered_map<uint64_t, uint32_t> seen;
    // a direct-mapped cache of recent values in front of the map (circuits
    // use few distinct rates, in repeating patterns)
    constexpr uint32_t kCache = 64;
    uint64_t cacheBits[kCache];
    uint32_t cacheIdx[kCache];
    for (uint32_t i = 0; i < kCache; ++i) cacheIdx[i] = kNone;
    for (uint64_t r = r0; r < r1; ++r) {
      const uint32_t n = img.rankNode[r], rl = img.rankLayer[r];
      uint32_t frames[8], nf = 0;
      auto toggle = [&](uint32_t f) {
        for (uint32_t i = 0; i < nf; ++i)
          if (frames[i] == f) {
            frames[i] = frames[--nf];
            return;
          }
        if (nf == 8) throw RuntimeError("STEPG: an auxiliary node reaches too many frame nodes");
        frames[nf++] = f;
      };
      if (frameSlot[n - rl * npl]) {
        frames[nf++] = n;
      } else {
        const uint64_t a = g.adjacency[n];
        for (uint32_t c : {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)}) {
          if (c == kNone) continue;
          if (frameSlot[c % npl]) {
            toggle(c);
            continue;
          }
          const uint64_t b2 = g.adjacency[c];
          for (uint32_t gch : {static_cast<uint32_t>(b2), static_cast<uint32_t>(b2 >> 32)}) {
            if (gch == kNone) continue;
            if (!frameSlot[gch % npl]) throw RuntimeError("STEPG: auxiliary nodes nested more than two deep");
            toggle(gch);
          }
        }
      }
      if (nf > 4) throw RuntimeError("STEPG: an auxiliary node reaches more than four frame nodes");
      for (uint32_t i = 1; i < nf; ++i)  // insertion sort (at most four)
        for (uint32_t x = i; x > 0 && frames[x - 1] > frames[x]; --x) std::swap(frames[x - 1], frames[x]);
      uint32_t rows[4] = {kNone, kNone, kNone, kNone};
      for (uint32_t i = 0; i < nf; ++i) {
        const uint32_t fl = frames[i] == n ? rl : frames[i] / npl;
        if (fl < rl) throw RuntimeError("STEPG: a flow runs backwards in time");
        part.rowSpan = std::max(part.rowSpan, fl - rl);
        rows[i] = (fl << kRowPosBits) | actIdx[frames[i]];
      }
      // compact code, or a spill entry (indexed within the range for now)
      const uint32_t base = rows[0] == kNone ? 0u : rows[0] >> kRowPosBits;
      uint32_t c16[3] = {kRowNone16, kRowNone16, kRowNone16};
      bool fits = rows[3] == kNone && base < kSpillBase;
      for (int i = 0; i < 3 && fits; ++i) {
        if (rows[i] == kNone) break;
        const uint32_t dl = (rows[i] >> kRowPosBits) - base;
        const uint32_t v = (dl << kRowPosBits) | (rows[i] & ((1u << kRowPosBits) - 1));
        if (dl >= (1u << (16 - kRowPosBits)) || v == kRowNone16) fits = false;
        c16[i] = v;
      }
      if (fits) {
        img.rankCode[2ull * r] = (base << 16) | c16[0];
        img.rankCode[2ull * r + 1] = (c16[2] << 16) | c16[1];
      } else {
        part.spilled.push_back(static_cast<uint32_t>(r));
        part.spill.insert(part.spill.end(), rows, rows + 4);
      }
      // probability: index among the range's distinct values (exact: the
      // table holds the doubles themselves)
      if (useTable) {
        uint64_t bits;
        std::memcpy(&bits, &img.rankProb[r], 8);
        const uint32_t slot = static_cast<uint32_t>((bits * 0x9E3779B97F4A7C15ull) >> 58);
        uint32_t idx;
        if (cacheIdx[slot] != kNone && cacheBits[slot] == bits) {
          idx = cacheIdx[slot];
        } else {
          auto it = seen.find(bits);
          if (it == seen.end()) {
            it = seen.emplace(bits, static_cast<uint32_t>(part.values.size())).first;
            part.values.push_back(img.rankProb[r]);
          }
          idx = it->second;
          cacheBits[slot] = bits;
          cacheIdx[slot] = idx;
        }
        // a range-local index for now (ranges with more distinct values
        // than the table holds are caught when joining)
        img.probIdx[r] = static_cast<uint16_t>(std::min<uint32_t>(idx, 0xFFFFu));
      }
    }
  });
  // join: spill entries in rank order, the distinct values in first-use
  // order over all ranks
  bool tableOk = useTable;
  std::vector<std::vector<uint16_t>> remap(rparts);
  std::unordered_map<uint64_t, uint32_t> global;
  for (uint32_t q = 0; q < rparts; ++q) {
    RankPart &part = rp[q];
    img.rowSpan = std::max(img.rowSpan, part.rowSpan);
    for (size_t i = 0; i < part.spilled.size(); ++i) {
      const uint64_t idx = img.rankSpill.size() / 4;
      img.rankSpill.insert(img.rankSpill.end(), part.spill.begin() + 4 * i, part.spill.begin() + 4 * i + 4);
      const uint64_t r = part.spilled[i];
      img.rankCode[2ull * r] = (kSpillBase << 16) | static_cast<uint32_t>(idx >> 32);
      img.rankCode[2ull * r + 1] = static_cast<uint32_t>(idx);
    }
    if (!tableOk) continue;
    if (part.values.size() > kProbDictMax) {
      tableOk = false;
      continue;
    }
    remap[q].resize(part.values.size());
    for (size_t v = 0; v < part.values.size(); ++v) {
      uint64_t bits;
      std::memcpy(&bits, &part.values[v], 8);
      auto it = global.find(bits);
      if (it == global.end()) {
        if (global.size() == kProbDictMax) {
          tableOk = false;
          break;
        }
        it = global.emplace(bits, static_cast<uint32_t>(img.probTab.size())).first;
        img.probTab.push_back(part.values[v]);
      }
      remap[q][v] = static_cast<uint16_t>(it->second);
    }
  }
  if (tableOk) {
    parallelFor(rparts, [&](uint32_t q) {
      uint64_t r0, r1;
      rangeOf(R, rparts, q, r0, r1);
      for (uint64_t r = r0; r < r1; ++r) img.probIdx[r] = remap[q][img.probIdx[r]];
    });
  } else {
    img.probTab.clear();
    img.probIdx.clear();
  }
}

}  // namespace gpa
