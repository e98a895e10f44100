// stepg.cpp -- STEPG builder (see stepg.hpp).
#include "stepg.hpp"

#include <algorithm>
#include <atomic>
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

void buildStepg(const Circuit &c, int level, Stepg &g, const double *theta, bool recordComponents) {
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
  g.adjacency.assign(g.numNodes, ~0ull);
  g.probs.assign(g.numNodes, 0.0);
  g.keys.assign(g.numMeasurements, 0);
  g.measBits.assign(static_cast<size_t>(g.numMeasurements) * g.numCols, 0u);
  g.compNode.clear();
  g.compLoc.clear();
  g.compKind.clear();

  Builder b{g, recordComponents};
  // persistence flows X->X, Z->Z between consecutive layers (Mixer.hpp:97-108)
  for (uint32_t l = 0; l + 1 < g.numLayers; ++l)
    for (uint32_t q = 0; q < g.numQubits; ++q) {
      b.addFlow(b.node(q, kX, l), l, b.node(q, kX, l + 1), l + 1);
      b.addFlow(b.node(q, kZ, l), l, b.node(q, kZ, l + 1), l + 1);
    }

  uint32_t layer = 0, meas = 0, det = 0, loc = 0;
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
  for (const Op &o : c.ops) {
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
        const double s = split1(p);
        b.single(1, t[j], layer, s, loc, 1);
        b.single(3, t[j], layer, s, loc, 1);
        b.single(2, t[j], layer, s, loc, 1);
        loc++;
      }
      break;
    case OpKind::Dep2:  // Mixer.hpp:178-201: 15 components in IX..ZZ order
      for (uint32_t j = 0; j < o.nt; j += 2) {
        const double p = rate(o.arg);
        if (!(p >= 0 && p <= 0.9375)) throw RuntimeError("Mixer: invalid DEP2 p > 0.9375.");
        const double s = split2(p);
        for (int pa = 0; pa < 4; ++pa)
          for (int pb = 0; pb < 4; ++pb)
            if (pa || pb) b.pair(pa, pb, t[j], t[j + 1], layer, s, loc);
        loc++;
      }
      break;
    case OpKind::Detector:
    case OpKind::Observable: {
      // Detectors take bits 0..D-1 in order; OBSERVABLE_INCLUDE(k) takes bit
      // D + k.  Records are XORed (a repeated record cancels).
      const uint32_t bit =
          o.kind == OpKind::Detector ? det++ : g.numDetectors + static_cast<uint32_t>(o.arg);
      for (uint32_t j = 0; j < o.nt; ++j) {
        const uint32_t row = meas + static_cast<uint32_t>(t[j]);  // t[j] = -k
        g.measBits[static_cast<size_t>(row) * g.numCols + bit / 32] ^= 1u << (bit % 32);
      }
      break;
    }
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

void buildImage(const Stepg &g, StepgImage &img) {
  const uint32_t N = g.numNodes, npl = g.npl, L = g.numLayers, C = g.numCols;
  const uint32_t k = nodesPerQubit(g.level);
  img.numLayers = L;
  img.npl = npl;
  img.numCols = C;
  img.numDetectors = g.numDetectors;
  img.numObservables = g.numObservables;
  img.numLocations = g.numLocations;
  img.rowSpan = 0;
  // slot j of a layer is a frame node (X or Z of its qubit)
  std::vector<uint8_t> frameSlot(npl);
  for (uint32_t j = 0; j < npl; ++j) frameSlot[j] = (j % k) < 2;

  // ranks (p > 0 nodes, node order) and active nodes (ranks and everything
  // their flows reach; children lie in later layers)
  uint32_t R = 0;
  for (uint32_t n = 0; n < N; ++n) R += g.probs[n] > 0;
  img.numRanks = R;
  img.rankNode.resize(R);
  img.rankLayer.resize(R);
  img.rankProb.resize(R);
  std::vector<uint8_t> active(N, 0);
  for (uint32_t l = 0, n = 0, r = 0; l < L; ++l)
    for (uint32_t j = 0; j < npl; ++j, ++n) {
      const double p = g.probs[n];
      if (p > 0) {
        img.rankNode[r] = n;
        img.rankLayer[r] = l;
        img.rankProb[r] = p;
        ++r;
        active[n] = 1;
      }
      if (!active[n]) continue;
      const uint64_t a = g.adjacency[n];
      const uint32_t c0 = static_cast<uint32_t>(a), c1 = static_cast<uint32_t>(a >> 32);
      if (c0 != kNone) active[c0] = 1;
      if (c1 != kNone) active[c1] = 1;
    }

  // seeds: XOR of the detector bits of every measurement on an active node
  std::vector<uint32_t> seedOf(N, kNone);
  std::vector<uint32_t> seedNode;
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
  // layer
  std::vector<uint32_t> actIdx(N, kNone);
  img.layerAct.assign(L + 1, 0);
  uint32_t total = 0, maxAct = 0;
  for (uint32_t l = 0, n = 0; l < L; ++l) {
    img.layerAct[l] = total;
    uint32_t cnt = 0;
    const uint32_t next0 = (l + 1) * npl;
    for (uint32_t j = 0; j < npl; ++j, ++n) {
      if (!active[n] || !frameSlot[j]) continue;
      actIdx[n] = cnt++;
      const uint64_t a = g.adjacency[n];
      for (uint32_t ch : {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)}) {
        if (ch == kNone) continue;
        if (ch < next0 || ch >= next0 + npl) throw RuntimeError("STEPG: a frame flow skips a layer");
        if (!frameSlot[ch - next0]) throw RuntimeError("STEPG: a frame node flows into an auxiliary node");
      }
    }
    total += cnt;
    maxAct = std::max(maxAct, cnt);
  }
  img.layerAct[L] = total;
  img.numActive = total;
  img.maxActive = maxAct;

  // sweep shape: the fewest records per thread (a power of two) that fit
  // the layer in one workgroup, then the fewest 64-lane waves
  uint32_t rpt = 1;
  while (rpt < sweepMinRpt()) rpt <<= 1;
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

  // records and seeds, pair by pair
  const uint32_t zeroW = 2u * P;
  img.act.resize(static_cast<size_t>(kSweepLeadPad + numPairs) * (P + 1) * 4);
  {
    uint32_t *rec = img.act.data();
    for (size_t j = 0, e = static_cast<size_t>(kSweepLeadPad + numPairs) * (P + 1); j < e; ++j, rec += 4) {
      rec[0] = rec[1] = rec[2] = zeroW | (zeroW << 16);
      rec[3] = 0xFFFFFFFFu;
    }
    for (uint32_t blk = 0; blk < kSweepLeadPad; ++blk) {
      uint32_t *h = img.act.data() + (static_cast<size_t>(blk) * (P + 1) + P) * 4;
      h[0] = h[1] = h[2] = h[3] = 0u;
    }
  }
  std::vector<uint32_t> newRows;  // seeds' column words, row-major, grouped by pair
  newRows.reserve(seedRows.size() + static_cast<size_t>(C) * 64);
  std::vector<uint32_t> localOdd(npl), localEven(npl);  // a seeded node's index in its pair's list
  for (uint32_t m = 0; m < numPairs; ++m) {
    const uint32_t first = C ? static_cast<uint32_t>(newRows.size() / C) : 0u;
    const uint32_t le = 2 * m, lo = 2 * m + 1;
    const uint32_t e0 = le * npl, o0 = lo * npl, n0 = (lo + 1) * npl;  // first nodes of layers 2m, 2m+1, 2m+2
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
    uint32_t *rec = img.act.data() + static_cast<size_t>(kSweepLeadPad + m) * (P + 1) * 4;
    rec[4 * P] = first;
    rec[4 * P + 1] = rec[4 * P + 2] = rec[4 * P + 3] = 0u;
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
      uint32_t parts[3], np = 0;  // seeded nodes among n and its children (counted mod 2)
      auto toggle = [](uint32_t *set, uint32_t &cnt, uint32_t x, uint32_t cap) {
        for (uint32_t i = 0; i < cnt; ++i)
          if (set[i] == x) {
            set[i] = set[--cnt];
            return true;
          }
        if (cnt == cap) return false;
        set[cnt++] = x;
        return true;
      };
      if (localEven[j] != 0xFFFFu) toggle(parts, np, n, 3);
      for (uint32_t c : ch) {
        if (c == kNone) continue;
        if (localOdd[c - o0] != 0xFFFFu) toggle(parts, np, c, 3);
        const uint64_t b2 = g.adjacency[c];
        for (uint32_t gch : {static_cast<uint32_t>(b2), static_cast<uint32_t>(b2 >> 32)})
          if (gch != kNone && !toggle(gc, ng, gch, 4))
            throw RuntimeError("STEPG: a node reaches more than four nodes two layers on");
      }
      (void)n0;
      uint32_t w[4] = {zeroW, zeroW, zeroW, zeroW};
      for (uint32_t i = 0; i < ng; ++i) w[i] = word(gc[i]);
      uint32_t *r = rec + 4 * actIdx[n];
      r[1] = w[0] | (w[1] << 16);
      r[2] = w[2] | (w[3] << 16);
      uint32_t es = 0xFFFFu;
      if (np == 1) {
        es = parts[0] < o0 ? localEven[parts[0] - e0] : localOdd[parts[0] - o0];
      } else if (np > 1) {
        es = cnt++;
        newRows.resize(newRows.size() + C, 0u);
        uint32_t *row = newRows.data() + newRows.size() - C;
        for (uint32_t i = 0; i < np; ++i) {
          const uint32_t *src = seedRows.data() + static_cast<size_t>(seedOf[parts[i]]) * C;
          for (uint32_t c = 0; c < C; ++c) row[c] ^= src[c];
        }
      }
      r[3] = (r[3] & 0x0000FFFFu) | (es << 16);
    }
    if (cnt >= 0xFFFFu) throw RuntimeError("circuit too large for the sweep: too many seeds in two layers");
  }
  const uint32_t S2 = C ? static_cast<uint32_t>(newRows.size() / C) : 0u;
  img.numSeeds = S2;
  img.seedVals.resize(static_cast<size_t>(C) * S2);
  for (uint32_t s2 = 0; s2 < S2; ++s2)
    for (uint32_t c = 0; c < C; ++c)
      img.seedVals[static_cast<size_t>(c) * S2 + s2] = newRows[static_cast<size_t>(s2) * C + c];

  // each rank's frame rows (itself, or the frame nodes its flows reach
  // through at most one auxiliary node, each counted modulo 2), encoded
  // compactly (see kSpillBase); and the probability table
  img.rankCode.resize(2ull * R);
  img.rankSpill.clear();
  img.probTab.clear();
  img.probIdx.clear();
  const bool useTable = probTable();
  if (useTable) img.probIdx.resize(R);
  std::unordered_map<uint64_t, uint32_t> seen;
  // a direct-mapped cache of recent values in front of the map (circuits
  // use few distinct rates, in repeating patterns)
  constexpr uint32_t kCache = 64;
  uint64_t cacheBits[kCache];
  uint32_t cacheIdx[kCache];
  for (uint32_t i = 0; i < kCache; ++i) cacheIdx[i] = kNone;
  bool tableOk = useTable;
  for (uint32_t r = 0; r < R; ++r) {
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
      for (uint32_t q = i; q > 0 && frames[q - 1] > frames[q]; --q) std::swap(frames[q - 1], frames[q]);
    uint32_t rows[4] = {kNone, kNone, kNone, kNone};
    for (uint32_t i = 0; i < nf; ++i) {
      const uint32_t fl = frames[i] == n ? rl : frames[i] / npl;
      if (fl < rl) throw RuntimeError("STEPG: a flow runs backwards in time");
      img.rowSpan = std::max(img.rowSpan, fl - rl);
      rows[i] = (fl << kRowPosBits) | actIdx[frames[i]];
    }
    // compact code, or a spill entry
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
      const uint64_t idx = img.rankSpill.size() / 4;
      img.rankSpill.insert(img.rankSpill.end(), rows, rows + 4);
      img.rankCode[2ull * r] = (kSpillBase << 16) | static_cast<uint32_t>(idx >> 32);
      img.rankCode[2ull * r + 1] = static_cast<uint32_t>(idx);
    }
    // probability table (exact: it holds the doubles themselves)
    if (tableOk) {
      uint64_t bits;
      std::memcpy(&bits, &img.rankProb[r], 8);
      const uint32_t slot = static_cast<uint32_t>((bits * 0x9E3779B97F4A7C15ull) >> 58);
      if (cacheIdx[slot] != kNone && cacheBits[slot] == bits) {
        img.probIdx[r] = static_cast<uint16_t>(cacheIdx[slot]);
      } else {
        auto it = seen.find(bits);
        if (it == seen.end()) {
          if (seen.size() == kProbDictMax) {
            tableOk = false;
          } else {
            it = seen.emplace(bits, static_cast<uint32_t>(img.probTab.size())).first;
            img.probTab.push_back(img.rankProb[r]);
          }
        }
        if (tableOk) {
          img.probIdx[r] = static_cast<uint16_t>(it->second);
          cacheBits[slot] = bits;
          cacheIdx[slot] = it->second;
        }
      }
    }
  }
  if (!tableOk) {
    img.probTab.clear();
    img.probIdx.clear();
  }
}

uint64_t imageDeviceBytes(const StepgImage &img) {
  const uint64_t R = img.numRanks;
  // sort buckets at the default first-digit fill (64 nodes per bucket),
  // 256 slots each, plus an arena of R slots for oversized buckets
  uint64_t B = 1;
  while (B * 64 < R) B <<= 1;
  const uint64_t slots = B * 256 + R;
  uint64_t b = 0;
  b += static_cast<uint64_t>(kSweepLeadPad + img.numPairs) * (img.layerStride + 1) * 16ull + img.seedVals.size() * 4ull;
  b += img.numCols * 16ull + R * 16ull + img.rankSpill.size() * 4ull + img.probTab.size() * 8ull;  // column bands, probabilities, frame rows
  b += 2ull * (img.numPairs + kSweepLeadPad) * img.layerStride * img.numCols * 4ull;  // frame sensitivities
  b += R * (16 + 64 + 16 + 24);                 // node records, target rows, overflow list, sort workspace
  b += slots * (8 + 4 + 4 + 8 + 8 + 4 + 4);     // bucket slots: keys, ranks, run arrays
  b += (B + 1) * 4 * 9;                         // per-bucket counters, counts and offsets
  b += R * (8 + 4 + 8 + 4) + R * 16 * 4;        // outputs (targets sized 16 per class at first)
  return b;
}

}  // namespace gpa
