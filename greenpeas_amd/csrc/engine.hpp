// engine.hpp -- the GPU compile pipeline for one or many circuits.
//
// Stages (one HIP stream, five launches, all on device; kernels in
// engine.hip):
//   sweep      backward sensitivity sweep, one workgroup per (circuit,
//              32-bit detector column), three-layer LDS ring, per-column
//              time band (replaces CUDA.hpp:26-52 + CUDAHandlers.hpp:38-62);
//   classify   per p>0 node: set bits -> FNV-1a-64 over 32 padded slots
//              (CUDAHandlers.hpp:86-121, Array.hpp:52-65), second hash for
//              collision checks, a row of its first 16 targets, and a slot
//              in its sort bucket (first radix digit = top b hash bits);
//   buckets    one wave64 per bucket: register bitonic sort on (hash, rank)
//              and segmented XOR-merge scan (replaces the CUB radix sort and
//              reduce-by-key, CUDA.hpp:275-343); oversized buckets are
//              sorted by their workgroup;
//   scan       per-bucket class, target and member offsets;
//   emit       compact CSR output: probability, offset and targets per class
//              (replaces the full N x 32 gather + copy, CUDA.hpp:488-497).
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "stepg.hpp"

namespace gpa {

struct VariantResult {
  uint32_t numErrors = 0;
  uint64_t numTargets = 0;
  uint32_t numDetectors = 0, numObservables = 0;
  const double *probs = nullptr;     // numErrors
  const uint32_t *offsets = nullptr; // numErrors + 1, relative to targets
  const uint32_t *targets = nullptr; // numTargets (nullptr when targets16 holds them)
  const uint16_t *targets16 = nullptr;  // numTargets, 16 bits each (setNarrowTargets)
  const uint64_t *hashes = nullptr;  // numErrors (nullptr when not downloaded)
  std::vector<uint32_t> offsetsRel;  // storage for offsets
  uint32_t target(uint64_t k) const { return targets16 ? targets16[k] : targets[k]; }
  void copyTargets(uint32_t *dst) const {
    if (targets16) for (uint64_t k = 0; k < numTargets; ++k) dst[k] = targets16[k];
    else if (numTargets) std::copy(targets, targets + numTargets, dst);
  }
};

// Pinned host memory holding one download's results (defined in
// engine.hip).  A holder of a handle keeps the block's arrays valid; the
// pipeline downloads into another block while one is held.
struct ResultBlock;
using ResultHandle = std::shared_ptr<const ResultBlock>;

struct StageTimes {
  double upload = 0, sweep = 0, classify = 0, partition = 0, buckets = 0, emit = 0, download = 0,
         total = 0;
};

class Pipeline {
 public:
  // `growable`: buffers grow to fit any load; otherwise load() throws
  // RuntimeError when a circuit exceeds the capacity fixed by reserve().
  Pipeline(int device, bool growable);
  ~Pipeline();
  Pipeline(const Pipeline &) = delete;
  Pipeline &operator=(const Pipeline &) = delete;

  // Fixes the capacity to what `imgs` need (non-growable pipelines).
  void reserve(const std::vector<const StepgImage *> &imgs);
  // Packs the images into the pinned input arena, prepares the launch plan
  // and uploads the arena (host -> device, stream-ordered).
  void load(const std::vector<const StepgImage *> &imgs);
  // Uploads the packed arena of the last load again (what each compile of
  // an unchanged circuit does).
  void upload();
  // Enqueues the device pipeline; `timed` records per-stage events.
  void run(bool timed);
  // Enqueues the pipeline as one HIP graph (captured on first use after a
  // load, replayed afterwards).
  void runGraph();
  // Waits for the device, copies the results to host.
  void download();
  // run() `iters` times and returns mean device microseconds per run.
  double benchDevice(uint32_t iters);

  uint32_t numVariants() const { return static_cast<uint32_t>(variants_.size()); }
  void result(uint32_t v, VariantResult &out) const;
  // The block the last download wrote (the arrays result() points into).
  ResultHandle resultBlock() const;
  // Whether download() copies the class hashes too (default on; off,
  // result() gives no hashes and copyHashes() reads them from the device).
  void setDownloadHashes(bool on);
  // Whether circuits of at most 65536 targets get 16-bit target arrays
  // (default off; halves the target download).  Read at the next load.
  void setNarrowTargets(bool on);
  // Circuit v's class hashes of the last run, from the device (valid until
  // the next run).
  void copyHashes(uint32_t v, uint64_t *dst) const;
  const StageTimes &times() const { return times_; }
  // Items the last run sorted (p > 0 nodes with a non-empty class, over all
  // circuits; synchronizes).  For measurement.
  uint64_t sortedItems() const;
  uint64_t deviceBytes() const;
  // Device bytes a fresh pipeline allocates to load and compile `imgs`
  // (growable: with the headroom a growing pipeline takes), including a
  // time-chunked sweep's buffers at their largest: an upper bound of what
  // deviceBytes() reports after such a load and its compiles.
  static uint64_t deviceBytesFor(const std::vector<const StepgImage *> &imgs, bool growable);
  // Device bytes this pipeline holds after loading and compiling `imgs`:
  // its buffers where the load fits them, grown (with headroom) where not.
  // An upper bound of its footprint during that load and its compiles too.
  uint64_t deviceBytesAfter(const std::vector<const StepgImage *> &imgs) const;
  // Times a run of this pipeline found a column band below its stored rows
  // (and the load then stored more rows), and S's words now.
  uint32_t bandRelayouts() const;
  uint64_t sWords() const;
  // Sort buckets of the load and slots per bucket.
  uint32_t numBuckets() const;
  uint32_t bucketCap() const;
  int device() const { return device_; }
  void *stream() const;
  // Queues later work on `stream` (a hipStream_t of the pipeline's device;
  // nullptr: the pipeline's own stream again), after the work queued so far
  // has completed.  Uploads still overlap the sweep on a second stream of
  // the pipeline's own.
  void setStream(void *stream);

  // Parity/debug access (after download or run + sync).
  void debugSensitivities(uint32_t v, std::vector<uint32_t> &cols32) const;  // ranks x cols
  // The load's time-chunked sweep plan: {split columns, blocks the first
  // sweep launch adds for them, blocks of the second launch}.
  void sweepPlan(uint32_t out[3]) const;
  // Device pointers for the differentiable path (single variant, after a
  // run): the p>0 nodes of non-empty classes, grouped by class in output
  // order (compacted on each call), and each class's first member in it.
  const uint32_t *devSortedRanks();
  const uint32_t *devClassMember() const;

  struct Impl;

 private:
  struct VariantInfo {
    uint32_t numRanks, numCols, numDetectors, numObservables, bucketBase, numBuckets;
  };
  int device_;
  bool growable_;
  std::vector<VariantInfo> variants_;
  StageTimes times_;
  Impl *impl_;
};

// Autotuning knob: mean number of items per sort bucket the first radix
// digit aims at (the digit width is chosen per circuit from it); 0 (the
// default): 64, or 192 for loads of more than 2^22 error nodes.
void setBucketFill(uint32_t fill);
uint32_t bucketFill();
// Whether classify stores each node's first targets for emit (on by
// default; off, emit gathers every class's targets from S -- the path taken
// for circuits with more than 65536 targets).  Read at each run.
void setStash(bool on);
bool stash();
// Whether classify may build per-workgroup column-mask tables in LDS (on
// by default; off, every node tests every column's band -- the path of
// workgroups whose table would not fit).  Read at each run.
void setMaskTable(bool on);
bool maskTable();
// Whether the bucket stage may sort on the packed 64-bit key (hash bits
// below the bucket digit, circuit-local rank, slot) -- on by default; off,
// every bucket takes the exact (hash, rank) sort that the packed path falls
// back to.  Same results either way.  Read at each run.
void setPackedSort(bool on);
bool packedSort();
// How growable pipelines (batches, serving) lay out S: 1 (default) each
// detector column stores the layers its band can reach (from its lowest
// seed layer less twice the circuit's seed gap), 0 every layer, 2 only
// from its lowest seed layer (bands reach below that: exercises the re-run
// with more rows; for tests).  Same results in every mode.  Read at load.
void setSBands(uint32_t mode);
uint32_t sBands();
// Slots per sort bucket (64, 128 or 256; default 256): classify writes a
// bucket's first `cap` nodes to its slots (the rest to an overflow list
// sorted by the workgroup); slot storage is numBuckets * cap.
// Same results for every value.  Read at load (a size class's tuned value
// first).
void setBucketCap(uint32_t cap);
uint32_t bucketCap();
// Target slots per error node a growable load's output buffer starts with
// (1..16, default 4); a DEM with more targets grows it and runs again.
void setTargetFill(uint32_t n);
uint32_t targetFill();
// How growable loads store the bucket stage's sorted items and runs: 1
// (default) packed by bucket -- at the prefix sums of the bucket counts,
// numRanks entries in all, with the outputs in the space of classify's slot
// records -- when a load has more than 2^22 error nodes, else at each
// bucket's slots (numBuckets * cap entries plus an arena of numRanks); 2
// packed for every growable load (for tests); 0 never.  Same results in
// every mode.  Read at load.
void setCompactItems(uint32_t mode);
uint32_t compactItems();
// Entries a growable load's sort overflow list (nodes past their bucket's
// slots; at least numRanks / 16) and oversized-bucket workspace start with
// (default 2^20); a run that needs more grows them and runs again.  Drivers
// hold the worst case.  Same results for every value.  Read at load.
void setOverflowStart(uint32_t n);
uint32_t overflowStart();
// Whether a load's long columns may be swept in time chunks from its second
// run on (planned from the bands of its first): 1 (default) where the
// estimated critical path gains, 2 every column whose band reaches the
// bottom with the most chunks that fit (for tests), 0 never (every column
// swept whole).  Read when the plan is made.
void setSweepSplit(uint32_t mode);
uint32_t sweepSplit();

// Segmented XOR-merge kernels (differentiable path); pointers are device
// pointers, stream is a hipStream_t.
// Node probabilities from location rates (the split components of each
// node XOR-merged) and their backward: per-component gradients (scratch,
// one per component), then one sum per location.  Device pointers; stream
// is a hipStream_t.
void nodeRatesForward(const double *theta, const uint32_t *compLoc, const uint8_t *compKind,
                      const uint32_t *nodeStart, const uint32_t *nodeComp, uint32_t numNodes, double *q,
                      void *stream);
void nodeRatesBackward(const double *theta, const uint32_t *compLoc, const uint8_t *compKind,
                       const uint32_t *nodeStart, const uint32_t *nodeComp, uint32_t numNodes, const double *gradQ,
                       double *gradComp, const uint32_t *locStart, const uint32_t *locComp, uint32_t numLocations,
                       double *gradTheta, void *stream);
void xorMergeForward(const double *values, const uint32_t *index, const uint32_t *segStart,
                     uint32_t numSegments, double *out, void *stream);
// Components the allocation-free backward can hold: its per-component
// gradients live in a fixed device array of this many doubles (8 MiB) per GPU.
constexpr uint32_t kNoAllocMaxComponents = 1u << 20;
// nodeRatesBackward without a gradComp buffer: the fixed array takes its
// place (numComponents above kNoAllocMaxComponents throws InvalidArgument;
// calls overlapping on one GPU share the array and race).  Allocates nothing.
void nodeRatesBackwardNoAlloc(const double *theta, const uint32_t *compLoc, const uint8_t *compKind,
                              const uint32_t *nodeStart, const uint32_t *nodeComp, uint32_t numNodes,
                              uint32_t numComponents, const double *gradQ, const uint32_t *locStart,
                              const uint32_t *locComp, uint32_t numLocations, double *gradTheta, void *stream);
void xorMergeBackward(const double *values, const uint32_t *index, const uint32_t *segStart,
                      uint32_t numSegments, const double *gradOut, double *gradValues,
                      void *stream);

}  // namespace gpa
