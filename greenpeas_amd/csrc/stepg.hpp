// stepg.hpp -- space-time error-propagation graph (STEPG) builder and the
// device image the GPU pipeline consumes.
//
// STEPG semantics follow the reference (STEPG.hpp:40-132, Mixer.hpp:89-534,
// Circuit.hpp:290-478): per layer (TICK) and qubit there are 3/5/8 error
// nodes (L0/L1/L2); gates add "flows" (edges) from layer l to l+1; each node
// has at most two children; noise components XOR-merge their probability
// into a node, and correlated components live on auxiliary nodes one or two
// layers earlier whose flows reproduce the Pauli pair.  A node's detector
// sensitivity is the XOR of its measurement seed and its children's
// sensitivities, computed backwards in time on the GPU.
//
// The device image re-lays that graph out for the MI355X sweep:
//  * only "active" nodes are kept (nodes with p > 0 and every node reachable
//    from one along flows; the others can never influence an output);
//  * only the frame nodes (slots X and Z of each qubit) are swept: gates
//    flow frame nodes into frame nodes, measurements seed frame nodes, and
//    the auxiliary nodes of correlated components flow into frame nodes (or
//    into an auxiliary Y node that does) but never receive a flow from one.
//    An auxiliary node's sensitivity is therefore the XOR of at most four
//    frame nodes' sensitivities, which the later stages form on the fly from
//    its frame row codes; the sequential sweep touches frame nodes only;
//  * the sweep takes layers in pairs (2m + 1, 2m), one step per pair: frame
//    flows always go from layer l to l + 1, so a node of the odd layer 2m + 1
//    is the XOR of its seed and its two children in layer 2m + 2, and a node
//    of the even layer 2m is the XOR of its own and its children's seeds and
//    of its grandchildren in layer 2m + 2 (at most four after cancellation).
//    Only even layers are read back: the sweep keeps layer 2m + 2 in an LDS
//    ring of two slots (slot (m + 1) mod 2, `layerStride` words each),
//    followed by a zero word (what absent nodes and padding read) and the
//    column's seed words;
//  * each pair is a fixed-stride block of `layerStride` 16-byte records plus
//    a header, position j holding the odd layer's j-th and the even layer's
//    j-th active node (node order): {odd node's children, even node's
//    grandchildren 0-1, grandchildren 2-3 (16-bit LDS word indices), odd
//    seed | even seed << 16 (indices in the pair's seed list, 0xFFFF =
//    none)}; the header holds the pair's first seed.  A node's sensitivity
//    goes to row l * layerStride + position of its column of S;
//  * measurement seeds are stored per pair (the seeds of the pair's two
//    layers, then the XOR of the seeds each even node combines, where it
//    combines more than one), column-major over 32-bit detector words
//    ("columns");
//  * per column, the highest and lowest layer holding a seed bound the time
//    band in which the column can be non-zero;
//  * the p > 0 nodes ("ranks", ascending node order) are the items that are
//    hashed, sorted and merged; each lists the frame rows whose XOR is its
//    sensitivity (rankCode).
#pragma once

#include <cstdint>
#include <vector>

#include "circuit.hpp"

namespace gpa {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// largest sweep workgroup; a layer is split over sweepThreads <= this many
// threads with sweepRpt records each
constexpr uint32_t kSweepMaxThreads = 1024;
// all-padding record blocks (pairs) in front of pair 0: what the sweep's
// prefetches load for the pairs below 0 that its trips run into; a column of
// S has rows for their layers after its 2 numPairs layers
constexpr uint32_t kSweepLeadPad = 7;
// frame row codes: layer << kRowPosBits | position in the layer
constexpr uint32_t kRowPosBits = 14;
// Compact frame rows (rankCode, two words per rank, what the device reads):
// x = base << 16 | row 0, y = row 2 << 16 | row 1, each row (layer - base)
// << kRowPosBits | position, kRowNone16 = none, base = the first row's
// layer.  A rank with four rows, a base >= kSpillBase or a row that would
// read as kRowNone16 is spilled instead: x = kSpillBase << 16 | (index >>
// 32), y = the low 32 bits of index, and rankSpill[4 index .. 4 index + 3]
// holds its four frame row codes (kNone-padded).
constexpr uint32_t kRowNone16 = 0xFFFFu, kSpillBase = 0xFFFFu;
// probabilities are sent as 16-bit indices into a per-circuit table when a
// circuit has at most this many distinct values (else as doubles)
constexpr uint32_t kProbDictMax = 65536;

inline uint32_t nodesPerQubit(int level) {
  switch (level) {
  case 0: return 3;
  case 1: return 5;
  case 2: return 8;
  }
  throw InvalidArgument("Invalid correlation level.");
}

// Host STEPG in the reference's layout: node = layer * npl + qubit * k + slot.
struct Stepg {
  int level = 2;
  uint32_t k = 8, numLayers = 0, numQubits = 0, npl = 0, numNodes = 0;
  uint32_t numMeasurements = 0, numDetectors = 0, numObservables = 0, numLocations = 0;
  uint32_t numCols = 0;  // 32-bit detector words: ceil((D + O) / 32)
  std::vector<uint64_t> adjacency;  // child0 | child1 << 32, kNone = no child
  std::vector<double> probs;
  std::vector<uint32_t> keys;       // measurement -> node
  std::vector<uint32_t> measBits;   // measurement x numCols (row-major) detector bits
  // noise components (for the differentiable path)
  std::vector<uint32_t> compNode, compLoc;
  std::vector<uint8_t> compKind;    // 0 rate itself, 1 DEPOLARIZE1 split, 2 DEPOLARIZE2 split
};

// Builds the STEPG of `c` at `level`.  `theta` (if non-null) overrides the
// rate of every fault location.  Throws RuntimeError on flow-rule violations
// and invalid rates (the reference's std::runtime_error cases).  `parallel`
// lets a large circuit's layer ranges be built on the worker pool (same
// result).
void buildStepg(const Circuit &c, int level, Stepg &g, const double *theta = nullptr,
                bool recordComponents = false, bool parallel = false);

// Device-ready arrays of one circuit (see the file comment).
struct StepgImage {
  int level = 2;
  uint32_t numLayers = 0, npl = 0, numCols = 0, numDetectors = 0, numObservables = 0;
  uint32_t numRanks = 0, numSeeds = 0, numActive = 0, numLocations = 0, maxActive = 0;
  // sweep shape: record j*sweepThreads + t of a layer belongs to thread t
  // (j < sweepRpt); layerStride = sweepThreads * sweepRpt >= maxActive
  uint32_t sweepThreads = 64, sweepRpt = 1, layerStride = 64;
  // largest (frame row layer - rank layer) over all ranks' frame rows
  uint32_t rowSpan = 0;
  uint32_t numPairs = 0;           // (numLayers + 1) / 2 sweep steps
  // largest distance in layers between consecutive layers holding seeds
  // (from layer 0; a memory experiment's round length)
  uint32_t seedGap = 0;
  std::vector<uint32_t> layerAct;  // numLayers + 1 offsets of each layer's first active node
  std::vector<uint32_t> act;       // 4 * (kSweepLeadPad + numPairs) * (layerStride + 1): padding
                                   // blocks, then one block per pair (see the file comment)
  std::vector<uint32_t> seedVals;  // numCols x numSeeds (column-major; grouped by pair)
  std::vector<int32_t> colHi, colLow;  // numCols: seed layer range (-1 = no seed)
  std::vector<uint32_t> rankLayer; // numRanks (host only)
  std::vector<double> rankProb;    // numRanks (sent when probTab is empty)
  std::vector<uint32_t> rankNode;  // numRanks (host only: node index of each rank)
  std::vector<uint32_t> rankCode;  // 2 * numRanks: compact frame rows
  std::vector<uint32_t> rankSpill; // 4 per spilled rank: its frame row codes
  std::vector<double> probTab;     // distinct probabilities (empty: too many)
  std::vector<uint16_t> probIdx;   // numRanks indices into probTab
};

// `parallel` lets a large image be built on the worker pool (same image).
void buildImage(const Stepg &g, StepgImage &img, bool parallel = false);

// Tuning knob: the fewest sweep records per thread an image is built with
// (1, 2, 4 or 8; more when a layer does not fit 1024 threads).  Fewer waves
// with more records each trade barrier participants for per-wave work.
void setSweepMinRpt(uint32_t rpt);
uint32_t sweepMinRpt();
// Tuning knob: whether images may send probabilities as table indices (on
// by default; off, every rank's double is sent).  Read when a circuit is
// parsed.
void setProbTable(bool on);
bool probTable();

// Tuned knobs per size class: circuits at `level` whose STEPG has n nodes,
// size class floor(log2 n), take `value` for knob `name` ("bucket_fill",
// "bucket_cap" or "sweep_rpt"; value 0 removes the entry) instead of the
// process-wide setting.  Read where the process-wide knob is read.
void setClassTuning(int level, uint32_t sizeClass, const char *name, uint32_t value);
void clearClassTuning();
// The knob's value for such a circuit: its class entry, else `fallback`.
uint32_t classTuning(const char *name, int level, uint64_t numNodes, uint32_t fallback);
uint32_t sizeClassOf(uint64_t numNodes);

}  // namespace gpa
