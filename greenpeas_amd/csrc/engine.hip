// engine.hip -- gfx950 kernels and host orchestration of the DEM compile
// pipeline (see engine.hpp for the stage list).
//
// Conventions: one wavefront = 64 lanes; "rank" = index of a p>0 STEPG node
// (ascending node order) across all circuits of a load; "column" = one
// 32-bit word of detector/observable bits; bit t of the concatenated words is
// detector t (t < D) or observable t - D.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <cstring>
#include <string>

#include "engine.hpp"

namespace gpa {

#define GPA_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t err_ = (call);                                                          \
    if (err_ != hipSuccess)                                                            \
      throw RuntimeError(std::string("HIP error: ") + hipGetErrorString(err_) + " at " + \
                         __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;  // Array.hpp:54
constexpr uint64_t kFnvPrime = 0x100000001b3ull;       // Array.hpp:56
constexpr int kClassBlock = 256;
constexpr int kScanBlock = 1024;
constexpr int kScanTile = 4 * kScanBlock;
constexpr int kBucketItemsPerWave = 256;  // wave-register sort capacity (4 per lane)
std::atomic<uint32_t> gBucketFill{0};     // mean items per bucket the digit width aims at (0: by load size)
std::atomic<uint32_t> gBucketCap{256};    // slots per sort bucket (64, 128 or 256)
std::atomic<uint32_t> gTargetFill{4};     // target slots per error node a growable load starts with
std::atomic<uint32_t> gOverflowStart{1u << 20};  // start entries of a growable load's overflow list and workspace
std::atomic<uint32_t> gCompactItems{1};   // growable loads' sorted items and runs packed by bucket: 1 large loads, 2 all, 0 none
std::atomic<uint32_t> gStash{1};          // target rows on (1) or off (0)
std::atomic<uint32_t> gMaskTable{1};      // classify's LDS mask tables allowed (1) or not (0)
std::atomic<uint32_t> gSweepSplit{1};     // long columns swept in time chunks (1; 2: always) or whole (0)
std::atomic<uint32_t> gPackedSort{1};     // bucket sort on the packed 64-bit key (1) or always exact (0)
std::atomic<uint32_t> gSBands{1};         // growable loads' S by column band (1), whole (0), tight (2)

// Per-circuit descriptor (device and host).
struct VDesc {
  uint32_t numLayers, npl, numCols, numDetectors;
  uint32_t numObservables, numRanks, numSeeds, bucketBits;
  uint32_t bucketBase, numBuckets, maskWords, layerStride;
  uint32_t sweepThreads, sweepRpt, rowSpan, probDict;  // rowSpan: see StepgImage
  uint32_t numPairs;
  uint32_t colRows;  // rows of every column's region of S (0: regions differ, see colAdj)
  uint32_t pad1, pad2;
  uint64_t actOff, seedOff, colOff, rankOff;
  uint64_t sOff;       // the circuit's first word of S (its columns' regions follow)
  uint64_t spillOff;   // first spill entry of the circuit's ranks
  uint64_t probOff;    // first probability (table entry if probDict, else the ranks' own)
};

// A rank's four frame row codes from its compact code (the format is given
// with kSpillBase); `spill` = the circuit's spill entries.
__host__ __device__ inline uint4 expandRankCode(const uint2 c, const uint4 *spill) {
  if ((c.x >> 16) == kSpillBase) return spill[(static_cast<uint64_t>(c.x & 0xFFFFu) << 32) | c.y];
  const uint32_t base = c.x >> 16;
  auto full = [base](uint32_t v) {
    return v == kRowNone16 ? 0xFFFFFFFFu
                           : ((base + (v >> kRowPosBits)) << kRowPosBits) | (v & ((1u << kRowPosBits) - 1));
  };
  return make_uint4(full(c.x & 0xFFFFu), full(c.y & 0xFFFFu), full(c.y >> 16), 0xFFFFFFFFu);
}

// A rank's probability: table entry or its own value.
__device__ __forceinline__ double rankProbability(const VDesc &d, const double *probVal, const uint16_t *probIdx,
                                                  uint64_t gi) {
  return probVal[d.probOff + (d.probDict ? probIdx[gi] : gi - d.rankOff)];
}

// Counter block zeroed at the start of every run: [flags | bucket counts].
// Per-run flag words, zeroed by the sweep kernel: [0] hash collision,
// [1] target overflow, [3] overflow-list length, [5] oversized-bucket arena fill,
// [kFlagSplitWait] a time-chunked sweep's wait for its inputs timed out;
// and [kFlagBandOverflow], zeroed by the host: a column's band reached below
// the rows its region of S stores
constexpr uint32_t kFlagWords = 16;
constexpr uint32_t kFlagBandOverflow = kFlagWords - 1;
constexpr uint32_t kFlagSplitWait = 7;
constexpr uint32_t kFlagScratchNeed = 6;  // workspace entries the oversized buckets asked for

// S holds, per detector column, the rows of layers base .. top (the rows
// its band can reach) at a region of its own, then two rows that the
// sweep's stores for layers below base go to (nothing reads them):
// colRegion = {region offset (words) low, high, base (even), rows stored
// before those two}; colAdj = (region offset - the circuit's sOff) /
// layerStride - base (rows), so the word of layer l >= base, position j is
// S[sOff + (colAdj + l) * layerStride + j].  A circuit's S stays below 2^32
// words, so those offsets are 32-bit.

// --------------------------------------------------------------------------
// wave64 helpers
// --------------------------------------------------------------------------

__device__ __forceinline__ uint64_t lanemaskLe(uint32_t lane) {
  return lane == 63 ? ~0ull : ((2ull << lane) - 1);
}

__device__ __forceinline__ uint32_t waveInclusiveSum(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d);
    if (lane >= static_cast<uint32_t>(d)) v += o;
  }
  return v;
}

// p (+) q = p + q - 2pq: the XOR of two independent events.  Equal to
// (1 - (1 - 2p)(1 - 2q)) / 2, the product form, evaluated without forming
// 1 - 2p so tiny probabilities keep their relative precision.
__device__ __forceinline__ double xmerge(double a, double b) { return a + b * (1.0 - 2.0 * a); }

__device__ __forceinline__ bool keyLess(uint64_t ha, uint32_t ra, uint64_t hb, uint32_t rb) {
  return ha < hb || (ha == hb && ra < rb);
}

// Second, independent 32-bit hash of a class: used only to detect FNV
// collisions between different classes inside one equal-hash run (the class
// size is compared too).  Each step is a bijection of h for a fixed target
// (xor, multiply by an odd constant, xorshift), so two lists of one length
// that differ in a single target always get different hashes.
__device__ __forceinline__ uint32_t mix2(uint32_t h, uint32_t t) {
  h = (h ^ t) * 0x9E3779B1u;
  return h ^ (h >> 16);
}

// Bitonic sort of 64*R (hash, rank) keys held R per lane, each with a
// payload word that moves with it; element index e = s * 64 + lane.
// Distances below 64 exchange through lane shuffles, larger ones inside a
// lane's registers.  (Measured on bb12: DPP row shifts and gfx950's
// permlane swaps in place of the shuffles cost 1.4 us more -- the extra
// VALU work outweighs the LDS round trips at 4 waves per SIMD.)  Used where
// the packed 64-bit key below cannot order a bucket.
template <int R>
__device__ __forceinline__ void waveBitonicSort(uint64_t (&h)[R], uint32_t (&r)[R], uint32_t (&x)[R],
                                                uint32_t lane) {
  constexpr uint32_t N = 64 * R;
#pragma unroll
  for (uint32_t k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const uint32_t js = j >> 6;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int sp = s ^ static_cast<int>(js);
          if (sp > s) {
            const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
            const bool asc = (e & k) == 0;
            // one comparison (distinct keys, or identical padding)
            const bool sw = keyLess(h[sp], r[sp], h[s], r[s]) == asc;
            if (sw) {
              uint64_t th = h[s];
              h[s] = h[sp];
              h[sp] = th;
              uint32_t tr = r[s];
              r[s] = r[sp];
              r[sp] = tr;
              uint32_t tx = x[s];
              x[s] = x[sp];
              x[sp] = tx;
            }
          }
        }
      } else {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
          const uint64_t ho = __shfl_xor(h[s], static_cast<int>(j));
          const uint32_t ro = __shfl_xor(r[s], static_cast<int>(j));
          const uint32_t xo = __shfl_xor(x[s], static_cast<int>(j));
          const bool lower = (lane & j) == 0;
          const bool asc = (e & k) == 0;
          const bool keepMin = lower == asc;
          // keys are distinct (ranks are) except the identical padding
          // entries, whose exchange changes nothing: one comparison decides
          const bool take = keyLess(ho, ro, h[s], r[s]) == keepMin;
          if (take) {
            h[s] = ho;
            r[s] = ro;
            x[s] = xo;
          }
        }
      }
    }
  }
}

// State carried from one 64-element chunk of a bucket to the next: the open
// run's merged probability and second hash, the last key, run and target
// counts so far.
struct RunCarry {
  double p;
  uint64_t prevHash;
  uint32_t h2;
  uint32_t runs;
  uint32_t nnzSum;
  uint32_t havePrev;
  uint32_t openStart;  // bucket position of the first member of the run still open
};

// The XOR-merge of a run's members a(0..L-1) (in sorted order) in the
// association of a segmented Hillis-Steele scan over the whole run:
// V(j, 0) = a(j), V(j, k) = xmerge(V(j - 2^(k-1), k - 1), V(j, k - 1)) where
// j >= 2^(k-1) (else V(j, k - 1)); the run's value is V(L - 1, K), 2^K >= L.
// A wave's 64-lane scan gives exactly this to every run that lies within
// one row of 64 sorted positions; a run across rows of at most 64 members
// gets it from a second scan of its members alone (scanRuns), a longer one
// from the evaluation below, so a class's probability does not depend on
// where its run falls in its bucket (nor on the bucket layout).  Counted
// from the run's end (q = L - 1 - j), the blocks are the aligned blocks of
// a binary tree, each the merge of its upper half (earlier members) into
// its lower half: built bottom-up like a binary counter, with one pending
// block per level in `slots`.
template <typename A>
__device__ __forceinline__ double hsRunMerge(const A &a, uint32_t first, uint32_t L) {
  double *slots = a.slots;
  for (uint32_t q = 0; q < L; ++q) {
    double v = a.get(first + L - 1 - q);
    uint32_t k = 0;
    for (; (q >> k) & 1u; ++k) v = xmerge(v, slots[k]);
    slots[k] = v;
  }
  double acc = 0.0;
  bool have = false;
  for (uint32_t k = 0; k < 32 && (L >> k) != 0; ++k)
    if ((L >> k) & 1u) {
      acc = have ? xmerge(acc, slots[k]) : slots[k];
      have = true;
    }
  return acc;
}

// Members' values by bucket position: a wave's sorted values in LDS, or
// read from the node records of the bucket's sorted ranks; `slots`: 32
// doubles of LDS for hsRunMerge.
struct LdsRunValues {
  const double *p;
  double *slots;
  __device__ double get(uint32_t i) const { return p[i]; }
};
struct NodeRunValues {
  const uint32_t *rank;
  const uint4 *info;
  double *slots;
  __device__ double get(uint32_t i) const;
};

// Bucket slot written by classify, one per (node, bucket slot), 16 bytes:
// {hash low, hash high, rank, bucketBits << 24 | circuit-local rank}; a local
// rank of kLocalRankNone marks circuits too large for the packed sort key.
// Beside it, a node record per rank, written in rank order: {probability
// low, high, second hash, class size}.  The bucket stage keeps both in LDS as
// one 32-byte SlotRec {slot, node record}.
struct SlotRec {
  uint4 a, b;
};
constexpr uint32_t kLocalRankBits = 24;
constexpr uint32_t kLocalRankNone = (1u << kLocalRankBits) - 1;

__device__ __forceinline__ uint4 nodeInfo(double p, uint32_t h2, uint32_t nnz) {
  const uint64_t b = __double_as_longlong(p);
  return make_uint4(static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32), h2, nnz);
}
__device__ __forceinline__ double infoProb(const uint4 &x) {
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(x.y) << 32) | x.x));
}

__device__ double NodeRunValues::get(uint32_t i) const { return infoProb(info[rank[i]]); }

struct RunOut {
  uint32_t *sortedRank;    // bucket-sorted ranks (in place over the bucket slots)
  uint32_t *runRank;       // per run: representative (smallest) rank
  double *runP;            // per run: merged probability
  uint64_t *runHash;       // per run: FNV hash
  uint32_t *runTgtPre;     // per run: target offset within the bucket
  uint32_t *runMember;     // per run: position of its first member within the bucket
  uint32_t *flags;         // [0] hash collision
};

// Segmented XOR-merge of one chunk of sorted keys (cnt valid elements,
// positions chunkBase.. within bucket starting at bs) continuing the runs in
// `cy`.  Runs are maximal sequences of equal hash; each run's probability is
// the XOR-merge of its members, its representative the first member.  Each
// element comes with its probability, second hash and class size.
template <int R, typename MV>
__device__ __forceinline__ void scanRuns(const uint64_t (&h)[R], const uint32_t (&rk)[R], const double (&pv)[R],
                                         const uint32_t (&h2v)[R], const uint32_t (&nzv)[R], uint32_t cnt,
                                         uint32_t chunkBase, uint32_t bs, RunCarry &cy, uint32_t lane,
                                         const RunOut &o, const MV &members) {
  bool f[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    // every lane executes every shuffle; lane 0 then takes the previous
    // slot's last key (or the carried key of the previous chunk)
    uint64_t ph = __shfl_up(h[s], 1);
    const uint64_t prevSlotLast = __shfl(h[s > 0 ? s - 1 : 0], 63);
    if (lane == 0) ph = s > 0 ? prevSlotLast : cy.prevHash;
    const bool hasPrev = (s > 0 || lane > 0) ? true : (cy.havePrev != 0);
    f[s] = e < cnt && (!hasPrev || h[s] != ph);
  }
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    if (static_cast<uint32_t>(s) * 64 >= cnt) break;
    const bool valid = e < cnt;
    const uint32_t r = rk[s];
    const uint32_t ownH2 = h2v[s];
    double v = pv[s];
    uint32_t hf = ownH2;
    int fl = f[s] ? 1 : 0;
    const uint32_t nz = f[s] ? nzv[s] : 0u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double pvs = __shfl_up(v, d);
      const uint32_t ph2 = __shfl_up(hf, d);
      const int pf = __shfl_up(fl, d);
      if (lane >= static_cast<uint32_t>(d)) {
        if (!fl) {
          v = xmerge(pvs, v);
          hf = ph2;
        }
        fl |= pf;
      }
    }
    const bool across = !fl;  // the run began before this row
    if (across) {
      v = xmerge(cy.p, v);
      hf = cy.h2;
    }
    const uint64_t bal = __ballot(f[s]);
    // A run begun in an earlier row: its merge in the association of one
    // scan over the run (its members right-aligned in a row of their own;
    // longer runs by hsRunMerge), for the lane where it ends in this row
    // (or, continuing, the row's last; a later row then writes it again).
    double canon = 0.0;
    if (!(bal & 1ull)) {
      const uint32_t rowValid = cnt - static_cast<uint32_t>(s) * 64 < 64 ? cnt - static_cast<uint32_t>(s) * 64 : 64u;
      const uint32_t endLane = bal ? static_cast<uint32_t>(__builtin_ctzll(bal)) - 1 : rowValid - 1;
      const uint32_t L = chunkBase + static_cast<uint32_t>(s) * 64 + endLane - cy.openStart + 1;
      if (L <= 64) {
        const uint32_t first = 64 - L;  // lane of the run's first member
        double w = lane >= first ? members.get(cy.openStart + lane - first) : 0.0;
        int wf = lane == first ? 1 : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const double pw = __shfl_up(w, d);
          const int pf = __shfl_up(wf, d);
          if (lane >= static_cast<uint32_t>(d)) {
            if (!wf) w = xmerge(pw, w);
            wf |= pf;
          }
        }
        canon = __shfl(w, 63);
      } else {
        canon = hsRunMerge(members, cy.openStart, L);
      }
    }
    const uint32_t before = __popcll(bal & lanemaskLe(lane));
    const uint32_t runIdx = cy.runs + before - 1;
    const uint32_t incl = waveInclusiveSum(nz, lane);
    // next element's flag (run end detection)
    int nf = __shfl_down(f[s] ? 1 : 0, 1);
    const int nextSlotFirst = __shfl(f[s + 1 < R ? s + 1 : s] ? 1 : 0, 0);
    if (lane == 63) nf = (s + 1 < R) ? nextSlotFirst : 0;
    if (valid) {
      o.sortedRank[bs + chunkBase + e] = r;
      if (f[s]) {
        o.runRank[bs + runIdx] = r;
        o.runHash[bs + runIdx] = h[s];
        o.runTgtPre[bs + runIdx] = cy.nnzSum + incl - nz;
        o.runMember[bs + runIdx] = chunkBase + e;
      }
      if (hf != ownH2) atomicOr(&o.flags[0], 1u);
      if (e + 1 == cnt || nf) {
        o.runP[bs + runIdx] = across ? canon : v;
      }
    }
    cy.p = __shfl(v, 63);
    cy.h2 = __shfl(hf, 63);
    if (bal) cy.openStart = chunkBase + static_cast<uint32_t>(s) * 64 + 63 - static_cast<uint32_t>(__clzll(bal));
    cy.runs += __popcll(bal);
    cy.nnzSum += __shfl(incl, 63);
  }
  if (cnt > 0) {
    const uint32_t last = cnt - 1;
    uint64_t lh = 0;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const uint64_t x = __shfl(h[s], static_cast<int>(last & 63));
      if (static_cast<uint32_t>(s) == (last >> 6)) lh = x;
    }
    cy.prevHash = lh;
    cy.havePrev = 1;
  }
}

// --------------------------------------------------------------------------
// kernels
// --------------------------------------------------------------------------

// LDS-only workgroup barrier: orders LDS accesses but lets global loads
// (the prefetched records of the next layers) stay in flight.
__device__ __forceinline__ void ldsBarrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The sweep addresses LDS by raw byte address (its kernel has no static LDS,
// so its dynamic LDS starts at address 0): the addresses stored in the
// records feed the LDS instructions with no base added.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ lds_u32 *ldsAt(uint32_t byteAddr) {
  return reinterpret_cast<lds_u32 *>(static_cast<uintptr_t>(byteAddr));
}

// Loads this thread's records of one pair block and the block's header
// word: buffer loads with the block's byte offset in the scalar offset, so
// the address costs no vector work.
template <int RPT>
__device__ __forceinline__ void loadBlock(uint4 (&r)[RPT], uint32_t &head, __amdgpu_buffer_rsrc_t rs,
                                          uint32_t blockOff, const uint32_t (&off)[RPT], uint32_t headOff) {
  blockOff = __builtin_amdgcn_readfirstlane(blockOff);
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off[k], blockOff, 0);
    r[k] = make_uint4(v[0], v[1], v[2], v[3]);
  }
  head = __builtin_amdgcn_raw_buffer_load_b32(rs, headOff, blockOff, 0);
}

// Backward sensitivity sweep of one 32-bit column of one circuit, two
// layers (2m + 1, 2m) per step and one LDS-only barrier per step.  Layer
// 2m + 2 is held in an LDS ring of two slots (slot (m + 1) mod 2); a record
// (one position of a pair, see StepgImage) gives the odd node's value as
// the XOR of two ring words and the even node's as the XOR of four, each
// plus a seed word; the even value goes to the ring, both to S.  Rows of S
// are addressed by layer (rows of the padding layers below 0 come after the
// circuit's rows), so a record needs no S address.  The loop runs trips of D
// steps (D even, see sweepTrip), unrolled, so each position in a trip has a
// fixed ring slot: the ring addresses a thread writes are per-column
// constants, and the record prefetch (D steps ahead, into D rotating
// register sets) advances one scalar offset.  The early-termination vote
// runs once per trip: a trip whose sensitivities are all zero, at or below
// the column's lowest seed, ends the band (the ring then holds zeros and no
// seed lies below).  Each trip ORs its values into one of two LDS words (by
// trip parity) before its last barrier; the next trip reads the word in its
// first step, behind that step's own LDS reads, and clears it after its
// first barrier.  Waves beyond the circuit's sweepThreads (a batch launches
// with its widest circuit's shape) only take part in the barriers and
// votes.  On exit the column's band [lo, hi] is recorded ({lo, hi};
// {INT_MAX, -1} for a column without seeds; {0, hi} for a split column).
// Block 0 also zeroes the run's flag words (read and written only by the
// later stages; a second, split launch zeroes them again); the grid has at
// least one block.  Split columns (SplitDesc) run chunks of whole trips with
// no early end; see there.
// Steps per trip = prefetch distance: four while the register sets fit (up
// to two records per thread), else two.
__host__ __device__ constexpr int sweepTrip(int rpt) { return rpt <= 2 ? 4 : 2; }
static_assert(2 * sweepTrip(1) - 1 <= static_cast<int>(kSweepLeadPad),
              "trip length: prefetches must stay within the padding blocks");

template <int RPT>
struct SweepState {
  __amdgpu_buffer_rsrc_t ac;  // the circuit's record blocks
  __amdgpu_buffer_rsrc_t sc;  // the column of S
  const uint32_t *gseeds;     // the column's seeds in global memory (GSEEDS)
  bool seeds;                 // false: the block sweeps with no seeds (a transfer map's)
  uint32_t blockBytes;        // 16 * (layerStride + 1)
  uint32_t headOff;           // 16 * layerStride: the header within a block
  uint32_t zeroB, seedB;      // LDS byte addresses of the zero word and the first seed
  uint32_t rowBytes;          // 4 * layerStride
  int32_t base;               // first layer the column's region stores
  uint32_t trash;             // its row for the even layers below base (the odd ones: the next)
  int32_t snapPair;           // the pair whose even layer is copied to snapB (a split chunk's bottom)
  uint32_t snapB;
  uint32_t off[RPT];          // byte offsets of this thread's records in a block
  uint32_t pos4[RPT];         // byte offsets of this thread's positions in a row or slot
};

// cache policy of a buffer store: sc0 | sc1 (system scope, written through)
constexpr int kStoreThrough = 1 | 16;
// and of a load that must see what other compute units wrote through in the
// same launch (system scope: not served from this XCD's L2)
constexpr int kCoherent = 1 | 16;
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt not waited for)
constexpr int kWaitVm0 = 0x0F70;

// LDS byte address of the ring word named by a 16-bit index.
__device__ __forceinline__ uint32_t ringLo(uint32_t x) { return (x & 0xFFFFu) << 2; }
__device__ __forceinline__ uint32_t ringHi(uint32_t x) { return (x >> 14) & 0x3FFFCu; }

// One step (pair m, trip position u): consume `r` (this thread's records),
// write the even values to ring slot `slotB` and both values to S, OR them
// into `any`, then refill `r` and `head` with the block one trip below
// (`loadOff`, advanced here).
template <int RPT, bool GSEEDS>
__device__ __forceinline__ void sweepStep(const SweepState<RPT> &st, uint4 (&r)[RPT], uint32_t &head,
                                          uint32_t slotB, int32_t m, uint32_t &loadOff, uint32_t &any) {
  // all of the wave's ring reads first, so the records' round trips overlap;
  // seeds are read only by waves holding a seeded record (seeded nodes are
  // rare: measurement nodes and their parents)
  uint32_t a[RPT], b[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    a[k] = *ldsAt(ringLo(r[k].x)) ^ *ldsAt(ringHi(r[k].x));
    b[k] = (*ldsAt(ringLo(r[k].y)) ^ *ldsAt(ringHi(r[k].y))) ^ (*ldsAt(ringLo(r[k].z)) ^ *ldsAt(ringHi(r[k].z)));
  }
  bool seeded = false;
#pragma unroll
  for (int k = 0; k < RPT; ++k) seeded |= r[k].w != 0xFFFFFFFFu;
  if (st.seeds && __ballot(seeded) != 0ull) {
    const uint32_t first = __builtin_amdgcn_readfirstlane(head);
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const uint32_t so = r[k].w & 0xFFFFu, se = r[k].w >> 16;
      if constexpr (GSEEDS) {
        a[k] ^= so != 0xFFFFu ? st.gseeds[first + so] : 0u;
        b[k] ^= se != 0xFFFFu ? st.gseeds[first + se] : 0u;
      } else {
        a[k] ^= *ldsAt(so != 0xFFFFu ? st.seedB + 4 * (first + so) : st.zeroB);
        b[k] ^= *ldsAt(se != 0xFFFFu ? st.seedB + 4 * (first + se) : st.zeroB);
      }
    }
  }
  // rows of layers 2m (even) and 2m + 1 (odd) in the column's region (base
  // is even); layers below base (padding layers below 0 among them) go to
  // the two rows after the stored ones
  const uint32_t evenRow = 2 * m < st.base ? st.trash : static_cast<uint32_t>(2 * m - st.base);
  const uint32_t evenB = evenRow * st.rowBytes, oddB = evenB + st.rowBytes;
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    *ldsAt(slotB + st.pos4[k]) = b[k];
    if (m == st.snapPair) *ldsAt(st.snapB + st.pos4[k]) = b[k];
    // write-through (system scope): S leaves no dirty lines for the end of
    // the kernel to write back (measured: 1 us less per compile on bb12)
    __builtin_amdgcn_raw_buffer_store_b32(a[k], st.sc, st.pos4[k], oddB, kStoreThrough);
    __builtin_amdgcn_raw_buffer_store_b32(b[k], st.sc, st.pos4[k], evenB, kStoreThrough);
    any |= a[k] | b[k];
  }
  loadOff -= st.blockBytes;
  loadBlock<RPT>(r, head, st.ac, loadOff, st.off, st.headOff);
}

// Time-chunked sweep of a long column.  The sweep is linear over GF(2): the
// values of a run of pairs are a linear function of the ring state entering
// it (the even layer above) plus the contribution of its own seeds, and of
// that state only the ring words its first pair reads count (the "live"
// positions of the boundary: where a memory round resets its ancillas, only
// the data qubits' frame nodes).  A column split into chunks 0..K-1 (chunk
// k = pairs bound[k] .. bound[k+1]-1) is swept by one launch:
//   producers  the top chunk, from the column's top seed (its state out: the
//              even layer of pair bound[K-1] at the live positions of that
//              boundary), and for each middle chunk k (1 <= k <= K-2) its
//              transfer map T_k -- one sweep per 32 live positions of its
//              top boundary, with unit vectors injected there, no seeds, no
//              S stores -- and its particular part p_k (zero state, its
//              seeds, no S stores), read out at the live positions of its
//              bottom boundary.  A map's sweeps also list each output's set
//              bits (atomics on per-output counts); a row of more than
//              kRowBits bits marks the map dense.  Each producer then counts
//              itself done (release);
//   resumed    chunks 0..K-2, launched after every producer (the hardware
//              dispatches workgroups in order, so a producer never waits
//              for a slot a waiting chunk holds), wait until the column's
//              producers are all done (acquire; a bounded wait: past it the
//              run is flagged and the host raises), form their entering
//              state as s = T_j s ^ p_j for j = K-2 .. k+1 from the top
//              chunk's state out (from the row lists, or a 4-bit lookup
//              table for a dense map), and sweep, storing S as the plain
//              sweep does.  The last of them to stage its inputs clears the
//              column's counts for the next run.
// A chunk runs whole trips from its top pair.  A trip that passes the
// chunk's bottom pair goes on below it: with the column's true values in
// the top and resumed chunks (they store the very words the chunk below
// stores), or with values nothing reads (maps and parts, which store no S).
// The state out is a snapshot of the ring taken at the bottom pair's step.
// n_k = live positions at boundary bound[k], listed ascending in
// SplitIo::live from liveOff[k].  State area of a split column (words, each
// part starting on a multiple of 4): the top chunk's state out [n_{K-1}] at
// stateOff; then, for each middle chunk k in ascending order, at tOff[k]:
// T_k [ceil(n_{k+1} / 32)][n_k] (word g * n_k + o: bit b = bit of output o
// for input 32g + b), p_k [n_k], the map's dense flag (4 words), its row
// counts [n_k] and row lists (kRowBits 16-bit inputs per output); stateEnd
// = one past the last.
constexpr int kSplitMaxChunks = 8;
constexpr uint32_t kSplitMaxColumns = 4;  // split columns per load
constexpr uint32_t kRowBits = 8;          // set bits a row list holds
constexpr uint32_t kSplitSpinLimit = 1u << 21;  // polls of a resumed chunk's wait (~2 s)
struct SplitDesc {
  uint32_t v, c, K, stateOff;
  int32_t bound[kSplitMaxChunks + 1];  // first pair of chunk k; bound[K] = top pair + 1
  uint32_t nLive[kSplitMaxChunks + 1];
  uint32_t liveOff[kSplitMaxChunks + 1];
  uint32_t tOff[kSplitMaxChunks + 1];
  uint32_t stateEnd;
  uint32_t producers;  // workgroups that count themselves done per run
};
static_assert(sizeof(SplitDesc) == 4 * (6 + 4 * (kSplitMaxChunks + 1)), "SplitDesc: 32-bit words, no padding");
// A split column's descriptor held in one register of each wave (lane l =
// word l; one load per wave instead of a memory round trip per field read)
struct SplitView {
  uint32_t w;
  __device__ uint32_t at(uint32_t i) const { return static_cast<uint32_t>(__builtin_amdgcn_readlane(w, i)); }
  __device__ uint32_t K() const { return at(2); }
  __device__ uint32_t stateOff() const { return at(3); }
  __device__ int32_t bound(uint32_t k) const { return static_cast<int32_t>(at(4 + k)); }
  __device__ uint32_t nLive(uint32_t k) const { return at(5 + kSplitMaxChunks + k); }
  __device__ uint32_t liveOff(uint32_t k) const { return at(6 + 2 * kSplitMaxChunks + k); }
  __device__ uint32_t tOff(uint32_t k) const { return at(7 + 3 * kSplitMaxChunks + k); }
  __device__ uint32_t stateEnd() const { return at(8 + 4 * kSplitMaxChunks); }
  __device__ uint32_t producers() const { return at(9 + 4 * kSplitMaxChunks); }
};
enum : uint32_t { kSplitNone = 0, kSplitTop = 1, kSplitBasis = 2, kSplitPart = 3, kSplitResume = 4 };
struct SplitIo {
  // after the launch's plain items: the maps' and parts' items, then the
  // resumed chunks' {split, mode | chunk << 8 | group << 16, circuit, column}
  const uint4 *items;
  const SplitDesc *descs;
  uint32_t *state;
  const uint32_t *live;
  uint32_t *sync;  // per split column: producers done, resumed chunks staged
  uint32_t numItems;
};
// A plain sweep item is {circuit, column, top seed layer, lowest seed
// layer}; a split column's top chunk is {circuit, column | kItemSplit, top
// seed layer, split}.
constexpr uint32_t kItemSplit = 0x80000000u;
__host__ __device__ inline uint32_t quad(uint32_t n) { return (n + 3) & ~3u; }
// words of T_k for n_{k+1} inputs and n_k outputs
__host__ __device__ inline uint32_t splitTWords(uint32_t nIn, uint32_t nOut) { return quad((nIn + 31) / 32 * nOut); }
// a middle chunk's block of the state area: T_k, p_k, dense flag, row
// counts, row lists; and where its flag starts
__host__ __device__ inline uint32_t splitListOff(uint32_t nIn, uint32_t nOut) {
  return splitTWords(nIn, nOut) + quad(nOut);
}
__host__ __device__ inline uint32_t splitMapWords(uint32_t nIn, uint32_t nOut) {
  return splitListOff(nIn, nOut) + 4 + quad(nOut) + 4 * quad(nOut);
}
// LDS words a resumed chunk's entering-state chain adds to the sweep's own
// (P = layer stride): two state vectors, the live positions it injects at,
// a zero word and a flag word (and two spare), a 4-bit lookup table (16
// entries per 4 inputs, P inputs at most) and the column's state area
// (mapWords: the top chunk's state out and every map's block, copied in at
// once).
__host__ __device__ inline uint32_t splitChainWords(uint32_t P, uint32_t mapWords) {
  return 3 * quad(P) + 4 + 128 * ((P + 31) / 32) + mapWords + 4;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u4;
__device__ __forceinline__ lds_u4 *lds4At(uint32_t byteAddr) {
  return reinterpret_cast<lds_u4 *>(static_cast<uintptr_t>(byteAddr));
}

// y = T x ^ p over GF(2), all in LDS (x [nIn] at xB; T [ceil(nIn/32)][nOut]
// at TB, p [nOut] at pB; y [nOut] to yB), with a 4-bit lookup table at tabB
// (for each group of 4 inputs, the XOR of every subset of their words;
// entry 0 of a group is 0, so a zero word of T looks up nothing).  L
// adjacent lanes (a power of two, at most 8) share an output, each taking
// every L-th word of its row of T, four words and their 32 lookups at a
// time; shuffles combine them.  All lanes run the same number of passes.
__device__ __forceinline__ void splitMatvec(uint32_t nIn, uint32_t nOut, uint32_t xB, uint32_t TB, uint32_t pB,
                                            uint32_t yB, uint32_t tabB) {
  const uint32_t G = (nIn + 31) / 32, tid = threadIdx.x, nt = blockDim.x;
  for (uint32_t e = tid; e < 128 * G; e += nt) {
    const uint32_t q = e >> 4, xs = e & 15;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
      const uint32_t j = 4 * q + b;
      const uint32_t x = *ldsAt(xB + 4 * (j < nIn ? j : 0u));
      v ^= (xs >> b & 1u) && j < nIn ? x : 0u;
    }
    *ldsAt(tabB + 4 * e) = v;
  }
  __syncthreads();
  uint32_t lg = 0;
  while (lg < 3 && (2u << lg) * nOut <= nt) ++lg;
  const uint32_t L = 1u << lg, lane = tid & (L - 1);
  for (uint32_t base = 0; base < (nOut << lg); base += nt) {
    const uint32_t o = (base + tid) >> lg;
    uint32_t v = 0;
    if (o < nOut) {
      for (uint32_t g0 = lane; g0 < G; g0 += 4 * L) {
        uint32_t w[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t g = g0 + u * L;
          const uint32_t x = *ldsAt(TB + 4 * ((g < G ? g : 0u) * nOut + o));
          w[u] = g < G ? x : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t g = g0 + u * L, gg = g < G ? g : 0u;
          uint32_t a = 0;
#pragma unroll
          for (uint32_t n = 0; n < 8; ++n) a ^= *ldsAt(tabB + 4 * (128 * gg + 16 * n + ((w[u] >> (4 * n)) & 15u)));
          v ^= a;
        }
      }
    }
    for (uint32_t sh = 1; sh < L; sh <<= 1) v ^= __shfl_xor(v, static_cast<int>(sh));
    if (o < nOut && lane == 0) *ldsAt(yB + 4 * o) = v ^ *ldsAt(pB + 4 * o);
  }
  __syncthreads();
}

// y = T x ^ p from a map's row lists (cB: counts, iB: kRowBits 16-bit
// inputs per output; every count <= kRowBits): one lane per output; the
// lists' reads, then the inputs', each in one batch (empty entries read the
// zero word zB).  A row of a memory circuit's round holds one or two bits,
// so this replaces a scan of every row.  Waves with no output skip it.
__device__ __forceinline__ void listMatvec(uint32_t nOut, uint32_t xB, uint32_t cB, uint32_t iB, uint32_t pB,
                                           uint32_t yB, uint32_t zB) {
  const uint32_t nt = blockDim.x;
  for (uint32_t base = 0; base < nOut; base += nt) {
    if (base + (threadIdx.x & ~63u) >= nOut) continue;  // (the whole wave)
    const uint32_t o = base + threadIdx.x, oc = o < nOut ? o : 0u;
    const uint32_t c = *ldsAt(cB + 4 * oc), pv = *ldsAt(pB + 4 * oc);
    const u32x4 pk = *lds4At(iB + 16 * oc);
    const uint32_t pw[4] = {pk.x, pk.y, pk.z, pk.w};
    uint32_t v = pv;
#pragma unroll
    for (uint32_t k = 0; k < kRowBits; ++k) {
      const uint32_t in = (pw[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
      v ^= *ldsAt(k < c ? xB + 4 * in : zB);
    }
    if (o < nOut) *ldsAt(yB + 4 * o) = v;
  }
  __syncthreads();
}

// A resumed chunk's state inputs to LDS: the column's state area (nq
// 16-byte words through `st`, loaded system-coherent: other compute units
// wrote them through in this launch; the top chunk's state out, then every
// middle chunk's block) to stB.  Each pass issues all of a thread's U loads
// (clamped indices: no branches) before its stores, so one pass costs one
// memory round trip.
template <uint32_t U>
__device__ void stageChain(__amdgpu_buffer_rsrc_t st, uint32_t nq, uint32_t stB) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  for (uint32_t r = 0; r * U * nt < nq; ++r) {
    u32x4 x[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(st, 16 * min((r * U + u) * nt + tid, nq - 1), 0, kCoherent);
      x[u] = u32x4{v[0], v[1], v[2], v[3]};
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t q = (r * U + u) * nt + tid;
      if (q < nq) *lds4At(stB + 16 * q) = x[u];
    }
  }
  __syncthreads();
}

template <int RPT, bool GSEEDS>
__global__ __launch_bounds__(kSweepMaxThreads) void sweepKernel(
    const VDesc *__restrict__ desc, const uint4 *__restrict__ items, const uint4 *__restrict__ act,
    const uint32_t *__restrict__ seedVals, const uint4 *__restrict__ colRegion, int2 *__restrict__ band,
    uint32_t *__restrict__ S, uint32_t numItems, uint32_t *__restrict__ flags, SplitIo sp, uint32_t oneCircuit) {
  // dynamic LDS only, addressed from 0:
  // ring[2P] | zero word | seeds (unless GSEEDS) | vote[2] | (resumed
  // chunks: the entering-state chain's words)
  const uint32_t tid = threadIdx.x;
  if (blockIdx.x == 0 && tid < kFlagBandOverflow) flags[tid] = 0u;
  if (blockIdx.x >= numItems + sp.numItems) return;
  // the item, then the circuit's descriptor (and the split's): two round
  // trips before the seed, state and record loads -- one in a load of one
  // circuit, whose descriptor is requested with the item
  const VDesc d0 = desc[0];
  const bool plain = blockIdx.x < numItems;
  const uint4 it = plain ? items[blockIdx.x] : sp.items[blockIdx.x - numItems];
  uint32_t v, c, mode, chunk = 0, group = 0, si;
  int32_t hi = 0, lowSeed = INT32_MIN;  // split chunks run their whole range (no early end)
  if (plain) {
    v = __builtin_amdgcn_readfirstlane(it.x);
    c = __builtin_amdgcn_readfirstlane(it.y & ~kItemSplit);
    hi = static_cast<int32_t>(__builtin_amdgcn_readfirstlane(it.z));
    mode = (it.y & kItemSplit) ? kSplitTop : kSplitNone;
    si = __builtin_amdgcn_readfirstlane(it.w);
    if (mode == kSplitNone) lowSeed = static_cast<int32_t>(si);
  } else {
    si = __builtin_amdgcn_readfirstlane(it.x);
    mode = __builtin_amdgcn_readfirstlane(it.y & 0xFFu);
    chunk = __builtin_amdgcn_readfirstlane((it.y >> 8) & 0xFFu);
    group = __builtin_amdgcn_readfirstlane(it.y >> 16);
    v = __builtin_amdgcn_readfirstlane(it.z);
    c = __builtin_amdgcn_readfirstlane(it.w);
  }
  mode = __builtin_amdgcn_readfirstlane(mode);
  SplitView sd{0u};  // loaded in split modes only
  if (mode != kSplitNone) sd.w = (tid & 63) < sizeof(SplitDesc) / 4 ? reinterpret_cast<const uint32_t *>(sp.descs + si)[tid & 63] : 0u;
  VDesc d = d0;
  if (!oneCircuit) d = desc[v];
  if (mode == kSplitNone && hi < 0) {
    if (tid == 0) band[d.colOff + c] = make_int2(0x7FFFFFFF, -1);
    return;
  }
  const uint32_t P = d.layerStride, T = d.sweepThreads;
  int32_t start = hi / 2, stop = 0;
  if (mode != kSplitNone) {
    if (mode == kSplitTop) {
      stop = sd.bound(sd.K() - 1);
    } else {
      start = sd.bound(chunk + 1) - 1;
      stop = sd.bound(chunk);
    }
  }
  const bool seeds = mode != kSplitBasis, stores = mode != kSplitBasis && mode != kSplitPart;
  // a producer's live positions of its bottom boundary (where its state out
  // is read from the snapshot), loaded now, off the path to its signal:
  // position tid + k * blockDim in lvOut[k] (at most RPT per thread, as
  // n <= layerStride)
  const bool producer = mode == kSplitTop || mode == kSplitBasis || mode == kSplitPart;
  uint32_t lvOut[RPT] = {};
  if (producer) {
    const uint32_t kb = mode == kSplitTop ? sd.K() - 1 : chunk, n = sd.nLive(kb);
    const uint32_t *lv = sp.live + sd.liveOff(kb);
#pragma unroll
    for (int k = 0; k < RPT; ++k) lvOut[k] = n ? lv[min(tid + k * blockDim.x, n - 1)] : 0u;
  }
  const uint32_t *sv = seedVals + d.seedOff + static_cast<uint64_t>(c) * d.numSeeds;
  uint32_t voteB = 4 * (2 * P + 1);
  if constexpr (!GSEEDS) voteB += 4 * d.numSeeds;

  // Both loops below run D-step trips with the vote steps at the same
  // points and decide on leaving at the end of a trip, from the same vote
  // words, so all waves meet the same number of barriers.  A single exit
  // per trip keeps the prefetches in flight (an exit inside would make the
  // compiler drain them at the loop head).  Pairs below 0 process padding;
  // the trip after a zero trip recomputes zeros below the band, which
  // nothing reads.
  constexpr int D = sweepTrip(RPT);
  int32_t mt = start;   // top pair of the current trip
  int32_t lo = -1;      // lowest layer of the band, once known
  uint32_t trip = 0;    // vote word parity
  // whole waves: sweepThreads is a multiple of 64
  const bool active = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(tid >> 6)) < (T >> 6);
  SweepState<RPT> st;
  uint32_t slotB[D];
  uint4 rs[D][RPT];
  uint32_t head[D];
  uint32_t loadOff = 0;
  // seeds to LDS, several loads in flight per thread (fewer beside many
  // prefetched records)
  constexpr uint32_t U = RPT >= 4 ? 2 : 8;
  // a resumed chunk's state inputs (16-byte words) per pass
  constexpr uint32_t US = RPT >= 4 ? 2 : 8;
  if (!GSEEDS) {
    const uint32_t n = d.numSeeds;
    for (uint32_t base = tid; base < n; base += U * blockDim.x) {
      // (clamped indices: the loads issue back to back, with no branches)
      uint32_t x[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t i = base + u * blockDim.x;
        x[u] = sv[min(i, n - 1)];
        x[u] = seeds ? x[u] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t i = base + u * blockDim.x;
        if (i < n) *ldsAt(4 * (2 * P + 1 + i)) = x[u];
      }
    }
  }
  // split chunks: LDS from sB (16-byte aligned) holds the snapshot of the
  // bottom pair's even layer (top chunks, maps, parts) or a resumed chunk's
  // entering-state chain: vectors A and B | live positions of its top
  // boundary | lookup table | maps and parts of chunks chunk+1 .. K-2
  const uint32_t sB = (voteB + 8 + 15) & ~15u;
  const uint32_t injSlot = mode == kSplitNone || mode == kSplitTop ? 0u : static_cast<uint32_t>((start + 1) & 1) * P;
  uint32_t entB = 0, lvB = 0;  // a resumed chunk's entering state, its positions
  if (mode == kSplitResume) {
    const uint32_t K = sd.K(), P4 = quad(P);
    const uint32_t aB = sB, bB = aB + 4 * P4, miscB = bB + 8 * P4, tabB = miscB + 16;
    const uint32_t mB = tabB + 4 * 128 * ((P + 31) / 32);
    lvB = bB + 4 * P4;
    const uint32_t s0 = sd.stateOff(), mapWords = sd.stateEnd() - s0;
    uint32_t *sync = sp.sync + 2 * si;
    // the live positions it injects at (they do not change from run to run)
    {
      const u32x4 *lv4 = reinterpret_cast<const u32x4 *>(sp.live + sd.liveOff(chunk + 1));
      for (uint32_t q = tid; q < quad(sd.nLive(chunk + 1)) / 4; q += blockDim.x) *lds4At(lvB + 16 * q) = lv4[q];
    }
    // wait until the column's producers are done (thread 0 polls; the
    // producers wrote the state area through to memory before counting
    // themselves done, and it is read here past this XCD's L2)
    if (tid == 0) {
      *ldsAt(miscB) = 0u;  // misc: a zero word
      const uint32_t need = sd.producers();
      uint32_t spins = 0;
      while (__hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins >= kSplitSpinLimit) {
          atomicOr(&flags[kFlagSplitWait], 1u);
          break;
        }
      }
    }
    __syncthreads();
    stageChain<US>(__builtin_amdgcn_make_buffer_rsrc(sp.state + s0, 0, 4 * mapWords, 0x00020000), mapWords / 4, mB);
    // the chain, from the top chunk's state out (staged first)
    uint32_t cur = mB, nxt = aB, n = sd.nLive(K - 1);
    for (uint32_t j = K - 2; j > chunk; --j) {
      const uint32_t nOut = sd.nLive(j);
      const uint32_t TB = mB + 4 * (sd.tOff(j) - s0), pB = TB + 4 * splitTWords(n, nOut);
      const uint32_t fB = TB + 4 * splitListOff(n, nOut);
      if (*ldsAt(fB)) splitMatvec(n, nOut, cur, TB, pB, nxt, tabB);
      else listMatvec(nOut, cur, fB + 16, fB + 16 + 4 * quad(nOut), pB, nxt, miscB);
      cur = nxt;
      nxt = cur == aB ? bB : aB;
      n = nOut;
    }
    entB = cur;
  }
  // ring and zero word; then the injected state at the live positions of
  // the slot the chunk's first pair reads
  for (uint32_t i = tid; i <= 2 * P; i += blockDim.x) *ldsAt(4 * i) = 0u;
  if (mode == kSplitResume) {
    __syncthreads();
    const uint32_t n = sd.nLive(chunk + 1);
    for (uint32_t j = tid; j < n; j += blockDim.x) *ldsAt(4 * (injSlot + *ldsAt(lvB + 4 * j))) = *ldsAt(entB + 4 * j);
  } else if (mode == kSplitBasis) {
    const uint32_t *lv = sp.live + sd.liveOff(chunk + 1);
    const uint32_t n = sd.nLive(chunk + 1), j = 32 * group + tid;
    const uint32_t pos = tid < 32 && j < n ? lv[j] : 0u;
    __syncthreads();
    if (tid < 32 && j < n) *ldsAt(4 * (injSlot + pos)) = 1u << tid;
  }
  if (tid < 2) *ldsAt(voteB + 4 * tid) = 0u;
  __syncthreads();

  if (active) {
    st.ac = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4 *>(act + d.actOff), 0,
                                              16 * (P + 1) * (kSweepLeadPad + d.numPairs), 0x00020000);
    const uint4 reg = colRegion[d.colOff + c];
    const uint32_t rows = __builtin_amdgcn_readfirstlane(reg.w) + 2;
    // (no records: every store is out of range and dropped)
    st.sc = __builtin_amdgcn_make_buffer_rsrc(
        S + ((static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(reg.y)) << 32) | __builtin_amdgcn_readfirstlane(reg.x)),
        0, stores ? 4 * rows * P : 0u, 0x00020000);
    st.gseeds = sv;
    st.seeds = seeds;
    st.blockBytes = 16 * (P + 1);
    st.headOff = 16 * P;
    st.zeroB = 8 * P;
    st.seedB = 8 * P + 4;
    st.rowBytes = 4 * P;
    st.base = static_cast<int32_t>(__builtin_amdgcn_readfirstlane(reg.z));
    st.trash = rows - 2;
    st.snapPair = mode == kSplitTop || mode == kSplitBasis || mode == kSplitPart ? stop : INT32_MIN;
    st.snapB = sB;
    // ring slot of each trip position (pair mt - u; D is even)
#pragma unroll
    for (int u = 0; u < D; ++u) slotB[u] = static_cast<uint32_t>((mt - u) & 1) * P * 4;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      // circuits with fewer records per thread repeat their last record
      // (same value to the same addresses)
      const uint32_t j = static_cast<uint32_t>(k) < d.sweepRpt ? static_cast<uint32_t>(k) : d.sweepRpt - 1;
      const uint32_t pos = j * T + tid;
      st.off[k] = pos * 16;
      st.pos4[k] = pos * 4;
    }
    // register sets for pairs mt, mt-1, ..., mt-D+1 (block kSweepLeadPad +
    // pair), loaded right before the loop (with other loops after them the
    // compiler loses count of what is outstanding at the loop head and
    // drains the prefetches on every trip).  Each set but the last is
    // followed by the stores of one step (to the rows of the lowest padding
    // layers, which nothing reads), as each refill is followed by a step's
    // stores in the loop: the compiler's count of memory operations
    // outstanding at the loop head is then the same on entry and on the
    // back edge, and its waits keep the full prefetch distance.
    loadOff = static_cast<uint32_t>(mt + static_cast<int32_t>(kSweepLeadPad)) * st.blockBytes;
    const uint32_t padB = st.trash * st.rowBytes;
#pragma unroll
    for (int u = 0; u < D; ++u) {
      loadBlock<RPT>(rs[u], head[u], st.ac, loadOff, st.off, st.headOff);
      __builtin_amdgcn_sched_barrier(0);
      if (u + 1 < D) {
        loadOff -= st.blockBytes;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          __builtin_amdgcn_raw_buffer_store_b32(0u, st.sc, st.pos4[k], padB, 0);
          __builtin_amdgcn_raw_buffer_store_b32(0u, st.sc, st.pos4[k], padB + st.rowBytes, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  if (active) {
    for (;;) {
      uint32_t any = 0;
      const uint32_t myWord = voteB + 4 * (trip & 1), prevWord = voteB + 4 * ((trip + 1) & 1);
      // first step: also read the previous trip's vote
      sweepStep<RPT, GSEEDS>(st, rs[0], head[0], slotB[0], mt, loadOff, any);
      const uint32_t prev = trip ? *ldsAt(prevWord) : 1u;
      if (prev == 0u && 2 * mt + 2 <= lowSeed && lo < 0) lo = 2 * mt + 2;
      ldsBarrier();
      if (tid == 0) *ldsAt(prevWord) = 0u;
#pragma unroll
      for (int u = 1; u < D; ++u) {
        sweepStep<RPT, GSEEDS>(st, rs[u], head[u], slotB[u], mt - u, loadOff, any);
        if (u == D - 1 && __ballot(any != 0) != 0ull && (tid & 63) == 0) *ldsAt(myWord) = 1u;
        ldsBarrier();
      }
      mt -= D;
      ++trip;
      if (lo >= 0 || mt < stop) break;
    }
  } else {
    for (;;) {
      const uint32_t prevWord = voteB + 4 * ((trip + 1) & 1);
      const uint32_t prev = trip ? *ldsAt(prevWord) : 1u;
      if (prev == 0u && 2 * mt + 2 <= lowSeed && lo < 0) lo = 2 * mt + 2;
#pragma unroll
      for (int u = 0; u < D; ++u) ldsBarrier();
      mt -= D;
      ++trip;
      if (lo >= 0 || mt < stop) break;
    }
  }
  if (lo < 0) lo = 0;
  if (mode == kSplitResume) {
    // the last resumed chunk of the column to finish clears its row counts,
    // dense flags and sync words for the next run (every resumed chunk has
    // read the state area by then)
    uint32_t *sync = sp.sync + 2 * si;
    const uint32_t K = sd.K();
    __syncthreads();
    if (tid == 0) {
      const uint32_t fin = __hip_atomic_fetch_add(sync + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *ldsAt(sB) = fin + 2 == K ? 1u : 0u;
    }
    __syncthreads();
    if (*ldsAt(sB)) {
      const auto rsState = __builtin_amdgcn_make_buffer_rsrc(sp.state, 0, 4 * sd.stateEnd(), 0x00020000);
      for (uint32_t j = 1; j + 2 <= K; ++j) {
        const uint32_t n = sd.nLive(j), lz = sd.tOff(j) + splitListOff(sd.nLive(j + 1), n);
        for (uint32_t i = tid; i < 4 + quad(n); i += blockDim.x)
          __builtin_amdgcn_raw_buffer_store_b32(0u, rsState, 4 * (lz + i), 0, kStoreThrough);
      }
      if (tid < 2) sync[tid] = 0u;
    }
  }
  if (producer) {
    // state out: the snapshot of pair `stop`'s even layer (taken in its
    // step, behind that step's barrier) at the live positions of boundary
    // `stop` (lvOut)
    const uint32_t kb = mode == kSplitTop ? sd.K() - 1 : chunk;
    const uint32_t n = sd.nLive(kb);
    uint32_t o = sd.stateOff();
    if (mode != kSplitTop) {
      o = sd.tOff(chunk);
      o += mode == kSplitPart ? splitTWords(sd.nLive(chunk + 1), n) : group * n;
    }
    // written through to memory: the resumed chunks read them from there
    const auto rs = __builtin_amdgcn_make_buffer_rsrc(sp.state, 0, 4 * sd.stateEnd(), 0x00020000);
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const uint32_t i = tid + k * blockDim.x;
      if (i < n) __builtin_amdgcn_raw_buffer_store_b32(*ldsAt(sB + 4 * lvOut[k]), rs, 4 * (o + i), 0, kStoreThrough);
    }
    if (mode == kSplitBasis) {
      // the map's row lists: each set bit appended to its output's list
      const uint32_t lo = sd.tOff(chunk) + splitListOff(sd.nLive(chunk + 1), n);
      uint32_t *cnt = sp.state + lo + 4;
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const uint32_t i = tid + k * blockDim.x;
        uint32_t w = i < n ? *ldsAt(sB + 4 * lvOut[k]) : 0u;
        while (w) {
          const uint32_t in = 32 * group + __builtin_ctz(w);
          w &= w - 1;
          const uint32_t at = atomicAdd(&cnt[i], 1u);
          if (at < kRowBits)
            __builtin_amdgcn_raw_buffer_store_b16(static_cast<uint16_t>(in), rs,
                                                  4 * (lo + 4 + quad(n)) + 2 * (kRowBits * i + at), 0, kStoreThrough);
          else
            __builtin_amdgcn_raw_buffer_store_b32(1u, rs, 4 * lo, 0, kStoreThrough);  // a row too dense
        }
      }
    }
    // counted done once every thread's stores are acknowledged (no release
    // fence: it would write back the whole L2 of this XCD)
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(sp.sync + 2 * si, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (tid == 0 && (mode == kSplitNone || mode == kSplitTop)) {
    // a band below the stored rows is cut at them (readers stay inside the
    // region) and reported: the host stores more rows and runs again
    const int32_t base = static_cast<int32_t>(colRegion[d.colOff + c].z);
    if (lo < base) atomicOr(&flags[kFlagBandOverflow], 1u);
    band[d.colOff + c] = make_int2(lo < base ? base : lo, hi);
  }
}

// One node's frame rows (its sensitivity is their XOR), decoded from its
// compact code: up to four codes layer << kRowPosBits | position, in
// ascending layer order, kNone-padded at the end.  idx = the row's word in a
// column of S (layer * P + position); the rows span layers [lo, lo + span].
struct FrameRows {
  uint32_t lay[4], idx[4];
  bool ok[4];
  uint32_t lo, span;
};

__device__ __forceinline__ FrameRows decodeRows(const uint4 code, uint32_t P) {
  constexpr uint32_t kPosMask = (1u << kRowPosBits) - 1;
  const uint32_t cd[4] = {code.x, code.y, code.z, code.w};
  FrameRows r;
  uint32_t hi = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    r.ok[q] = cd[q] != NONE;
    r.lay[q] = r.ok[q] ? cd[q] >> kRowPosBits : 0u;
    r.idx[q] = r.ok[q] ? r.lay[q] * P + (cd[q] & kPosMask) : 0u;
    if (r.ok[q]) hi = r.lay[q];
  }
  r.lo = r.lay[0];
  r.span = hi - r.lo;
  return r;
}

// Word of column c for one node: XOR of its rows' S words over the rows
// whose layer lies in the column's band (outside the band a frame
// sensitivity is zero and its S word is not written).
__device__ __forceinline__ uint32_t bandXor(const FrameRows &r, const int2 bd, const uint32_t (&x)[4]) {
  uint32_t v = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (r.ok[q] && bd.x <= static_cast<int32_t>(r.lay[q]) && static_cast<int32_t>(r.lay[q]) <= bd.y) v ^= x[q];
  return v;
}

// Column masks of the layer windows a workgroup's nodes' rows lie in, in
// LDS.  Bit c of word w of window (s, lo) is set when column 32w + c's band
// meets layers [lo, lo + s]; a node whose rows span [lo, lo + s] has its
// targets in those columns only (the exact words then come from bandXor).
// Windows s = 0..span for lo in [l0, l0 + nl), s-major.  Each wave turns 64
// columns' bands into two mask words per window with one ballot.
constexpr uint32_t kMaskTableWords = 1024;
constexpr uint32_t kBandTableCols = 1024;  // column bands kept in LDS by classify

__host__ __device__ constexpr uint32_t maskTableWords(uint32_t nl, uint32_t span, uint32_t maskWords) {
  return (span + 1) * nl * maskWords;
}

__device__ __forceinline__ void buildMaskTable(uint32_t *tab, uint32_t l0, uint32_t nl, uint32_t span,
                                               uint32_t numCols, uint32_t maskWords,
                                               const int2 *__restrict__ band) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (uint32_t ch = wave; 2 * ch < maskWords; ch += waves) {
    const uint32_t c = ch * 64 + lane;
    const int2 bd = c < numCols ? band[c] : make_int2(1, 0);
    const uint32_t w = 2 * ch + (lane & 1);
    for (uint32_t s = 0; s <= span; ++s)
      for (uint32_t li = 0; li < nl; ++li) {
        const int32_t lo = static_cast<int32_t>(l0 + li), hi = lo + static_cast<int32_t>(s);
        const uint64_t bits = __ballot(bd.x <= hi && lo <= bd.y);
        if (lane < 2 && w < maskWords)
          tab[(s * nl + li) * maskWords + w] = static_cast<uint32_t>(lane ? bits >> 32 : bits);
      }
  }
  __syncthreads();
}

// Calls f(target) for every set bit of one node's sensitivity row, in
// ascending target order, visiting the columns set in the node's window
// mask (`mrow`, maskWords words, read two at a time: circuits with at most
// 64 columns never reload it).  Columns are taken kBatch at a time: their
// bands are read together, then the row words that lie in each column's
// band are loaded before any is used, so a node waits for one memory round
// trip per kBatch columns (16: one for nearly every node of a circuit with
// ~15 columns per layer; 4 where nodes see few columns).
template <int kBatch, bool UNIFORM, typename F>
__device__ __forceinline__ void forEachTargetMasked(const uint32_t *mrow, uint32_t maskWords, const int2 *band,
                                                    const int32_t *adj, uint32_t P, uint32_t colWords,
                                                    const FrameRows &r, const uint32_t *__restrict__ S0, F &&f) {
  if (!r.ok[0] || !maskWords) return;
  uint32_t w = 0;  // the window's low word
  uint64_t m = mrow[0] | (maskWords > 1 ? static_cast<uint64_t>(mrow[1]) << 32 : 0ull);
  for (;;) {
    uint32_t col[kBatch], x[kBatch][4];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      while (m == 0ull && w + 2 < maskWords) {
        w += 2;
        m = mrow[w] | (w + 1 < maskWords ? static_cast<uint64_t>(mrow[w + 1]) << 32 : 0ull);
      }
      col[j] = m ? w * 32 + static_cast<uint32_t>(__builtin_ctzll(m)) : NONE;
      m &= m - 1;
    }
    int2 bd[kBatch];
    uint32_t aj[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      bd[j] = col[j] != NONE ? band[col[j]] : make_int2(1, 0);
      if constexpr (UNIFORM)
        aj[j] = col[j] != NONE ? col[j] * colWords : 0u;  // regions of one size, every layer stored
      else
        aj[j] = col[j] != NONE ? static_cast<uint32_t>(adj[col[j]]) * P : 0u;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int32_t l = static_cast<int32_t>(r.lay[q]);
        x[j][q] = (r.ok[q] && bd[j].x <= l && l <= bd[j].y) ? S0[aj[j] + r.idx[q]] : 0u;
      }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      if (col[j] == NONE) continue;
      uint32_t v = x[j][0] ^ x[j][1] ^ x[j][2] ^ x[j][3];
      while (v) {
        f(col[j] * 32 + static_cast<uint32_t>(__builtin_ctz(v)));
        v &= v - 1;
      }
    }
    if (col[kBatch - 1] == NONE) break;
  }
}

// The same walk without a mask table: every column's band is tested (one
// memory round trip per column holding a row).  Used by the emit stage for
// the few nodes without a stashed target row, and by classify when a
// workgroup's mask table would not fit in LDS.
template <typename F>
__device__ __forceinline__ void forEachTargetScan(uint32_t numCols, const int2 *band, const int32_t *adj, uint32_t P,
                                                  const FrameRows &r, const uint32_t *__restrict__ S0, F &&f) {
  if (!r.ok[0]) return;
  for (uint32_t c = 0; c < numCols; ++c) {
    const int2 bd = band[c];
    if (bd.y < static_cast<int32_t>(r.lo) || bd.x > static_cast<int32_t>(r.lo + r.span)) continue;
    const uint32_t a = static_cast<uint32_t>(adj[c]) * P;
    uint32_t x[4];
    // rows outside the band are not read (they may lie outside the region)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t l = static_cast<int32_t>(r.lay[q]);
      x[q] = r.ok[q] && bd.x <= l && l <= bd.y ? S0[a + r.idx[q]] : 0u;
    }
    uint32_t v = bandXor(r, bd, x);
    while (v) {
      f(c * 32 + static_cast<uint32_t>(__builtin_ctz(v)));
      v &= v - 1;
    }
  }
}

// Class extraction + hashing of every p>0 node: the ascending list of set
// bits, FNV-1a-64 over it padded with 0xFFFFFFFF to 32 slots (lists longer
// than 32 hash all entries, unpadded), a second hash and the class size.
// The node then takes a slot of its sort bucket (first radix digit = top
// bucketBits hash bits): slot pos of bucket b is pslot[b * cap + pos], one
// 16-byte {hash, rank} record (one write per node), pos from an atomic
// counter per bucket; nodes beyond a bucket's cap slots go to the overflow
// list.  In circuits whose targets fit 16 bits (at most kStashMaxCols
// columns), the first 2 W targets of every node are also stored as one
// row of W words per node (`stash`, 16-bit entries; W = 1 << stashShift,
// at most kStashWords: 32 targets cover every class of the reference's
// circuits, 16 would leave 1.1 % of bb12 L2's to the much slower gather
// from S; a load's rows take the narrowest width that holds its largest
// class once a run has shown it), so the emit stage
// copies a class's targets from its
// representative's row instead of gathering them from S one column at a
// time: each wave assembles its 64 rows in LDS and writes them out
// contiguously.
constexpr uint32_t kStashWords = 16;                  // 32-bit words per row, at most
constexpr uint32_t kStashTargets = 2 * kStashWords;   // 16-bit targets per row, at most
constexpr uint32_t kStashMaxCols = 65536 / 32;
constexpr uint32_t kBucketCap = kBucketItemsPerWave;

struct ClassifyOut {
  uint4 *info;            // per node: nodeInfo(probability, second hash, class size)
  uint32_t *bucketCount;  // bucket b's counter at b * countStride
  uint32_t countStride;
  uint4 *pslot;           // bucket slots
  uint4 *ovSlot;          // overflow list (count in flags[3]) and each entry's bucket
  uint32_t *ovBucket;
  uint32_t *flags;
  uint32_t *stash;
  uint32_t stashMaxCols;  // circuits with more columns get no rows (0: none do)
  uint32_t stashShift;    // a row holds 1 << stashShift words (2..4)
  uint32_t packedKey;     // 0: slots carry kLocalRankNone (the bucket stage sorts exactly)
  uint32_t bucketCap;     // slots per bucket
  uint32_t ovCap;         // overflow list entries held (past them the count goes on, nothing is written)
};

// Ranks' inputs: compact frame rows, spill entries, probabilities.
struct RankIn {
  const uint2 *code;
  const uint4 *spill;
  const double *probVal;
  const uint16_t *probIdx;
};

template <bool TABLE, int BATCH>
__global__ __launch_bounds__(kClassBlock) void classifyKernel(
    const VDesc *__restrict__ desc, const uint4 *__restrict__ tiles, const int2 *__restrict__ band,
    const int32_t *__restrict__ colAdj, const uint32_t *__restrict__ S, RankIn rin, ClassifyOut o,
    uint32_t oneCircuit) {
  // per wave: 64 rows of kStashWords words, padded to kStashWords + 1 so
  // that a lane's writes and the row-major read-out avoid bank conflicts
  __shared__ uint32_t rows[kClassBlock / 64][64 * (kStashWords + 1)];
  __shared__ uint32_t maskTab[TABLE ? kMaskTableWords : 1];
  __shared__ int2 bandTab[kBandTableCols];
  __shared__ int32_t adjTab[kBandTableCols];
  // tile: {circuit, first rank, lowest layer of its ranks' rows, layers
  // spanned by them}; the host picks TABLE only when every workgroup's
  // table fits kMaskTableWords
  // (a load of one circuit: its descriptor requested with the tile)
  const VDesc d0 = desc[0];
  const uint4 t = tiles[blockIdx.x];
  VDesc d = d0;
  if (!oneCircuit) d = desc[t.x];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = t.y + threadIdx.x;
  // this node's frame rows and probability, loaded before the tables are built
  const uint64_t gi = d.rankOff + (i < d.numRanks ? i : t.y);
  const uint2 ccode = rin.code[gi];
  const double prob = rankProbability(d, rin.probVal, rin.probIdx, gi);
  const uint32_t l0 = t.z, nl = t.w;
  const int2 *bandC = band + d.colOff;
  const int32_t *adjC = colAdj + d.colOff;
  if (d.numCols <= kBandTableCols) {
    for (uint32_t c = threadIdx.x; c < d.numCols; c += blockDim.x) {
      bandTab[c] = bandC[c];
      adjTab[c] = adjC[c];  // (read by the scan path in every layout)
    }
    bandC = bandTab;  // filled by the barrier below
    adjC = adjTab;
  }
  if constexpr (TABLE)
    buildMaskTable(maskTab, l0, nl, d.rowSpan, d.numCols, d.maskWords, band + d.colOff);
  else
    __syncthreads();
  const uint32_t W = 1u << o.stashShift;  // words per row (the LDS rows are padded to W + 1)
  // 16-bit entries of the 32-bit rows: a type that may alias them
  typedef uint16_t __attribute__((may_alias)) u16_alias;
  u16_alias *row = reinterpret_cast<u16_alias *>(rows[wave] + lane * (W + 1));
  const bool stashed = d.numCols <= o.stashMaxCols;
  // this node's class hash, its bucket and the slot the bucket's counter
  // gave it (used after the wave's target rows are written: the atomic's
  // round trip overlaps those stores)
  uint64_t h = kFnvBasis;
  uint32_t cnt = 0, bkt = 0, pos = 0;
  if (i < d.numRanks) {
    uint32_t h2 = 0x2545F491u;
    const FrameRows fr = decodeRows(expandRankCode(ccode, rin.spill + d.spillOff), d.layerStride);
    auto take = [&](uint32_t tg) {
      h = (h ^ tg) * kFnvPrime;
      h2 = mix2(h2, tg);
      // targets past the row's end overwrite its last entry: the emit stage
      // reads rows of classes with at most 2 W targets only
      row[cnt < 2 * W ? cnt : 2 * W - 1] = static_cast<uint16_t>(tg);
      cnt++;
    };
    if constexpr (TABLE)
    {
      const uint32_t *mrow = maskTab + (fr.span * nl + (fr.lo - l0)) * d.maskWords;
      if (d.colRows)
        forEachTargetMasked<BATCH, true>(mrow, d.maskWords, bandC, adjC, d.layerStride, d.colRows * d.layerStride,
                                         fr, S + d.sOff, take);
      else
        forEachTargetMasked<BATCH, false>(mrow, d.maskWords, bandC, adjC, d.layerStride, 0u, fr, S + d.sOff, take);
    }
    else
      forEachTargetScan(d.numCols, bandC, adjC, d.layerStride, fr, S + d.sOff, take);
    if (cnt && cnt <= 32)
      for (uint32_t k = cnt; k < 32; ++k) h = (h ^ 0xFFFFFFFFull) * kFnvPrime;
    h2 = mix2(h2, cnt);
    o.info[gi] = nodeInfo(prob, h2, cnt);
    if (cnt) {
      bkt = d.bucketBase + (d.bucketBits ? static_cast<uint32_t>(h >> (64 - d.bucketBits)) : 0u);
      pos = atomicAdd(&o.bucketCount[static_cast<uint64_t>(bkt) * o.countStride], 1u);
    }
  }
  // the wave's rows, row-major and contiguous from its first rank: every
  // lane's row writes are made visible to the wave (LDS accesses of a wave
  // complete in order; the fences keep the compiler from moving the reads
  // above the writes)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  const uint32_t waveFirst = t.y + wave * 64;
  if (stashed && waveFirst < d.numRanks) {
    const uint32_t nRows = d.numRanks - waveFirst < 64 ? d.numRanks - waveFirst : 64;
    uint32_t *dst = o.stash + (d.rankOff + waveFirst) * W;
#pragma unroll
    for (uint32_t k = 0; k < kStashWords; ++k) {
      if (k >= W) break;
      const uint32_t e = k * 64 + lane;  // element of the 64 x W block
      const uint32_t r = e >> o.stashShift, c = e & (W - 1);
      if (r < nRows) dst[e] = rows[wave][r * (W + 1) + c];
    }
  }
  if (i < d.numRanks && cnt) {
    const uint32_t local = o.packedKey && d.numRanks < kLocalRankNone ? i : kLocalRankNone;
    const uint4 slot = make_uint4(static_cast<uint32_t>(h), static_cast<uint32_t>(h >> 32),
                                  static_cast<uint32_t>(gi), d.bucketBits << kLocalRankBits | local);
    if (pos < o.bucketCap) {
      o.pslot[static_cast<uint64_t>(bkt) * o.bucketCap + pos] = slot;
    } else {
      const uint32_t k = atomicAdd(&o.flags[3], 1u);
      if (k < o.ovCap) {
        o.ovSlot[k] = slot;
        o.ovBucket[k] = bkt;
      }
    }
  }
}

// A wave's LDS writes made visible to its other lanes (a wave's LDS
// accesses complete in order; the fences keep the compiler from moving the
// reads above the writes).
__device__ __forceinline__ void waveLdsSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Element e - 1's value of a register array held 64 per slot (lane 0 of
// slot s takes lane 63 of slot s - 1; e = 0 gets `first`).
template <int R, typename T>
__device__ __forceinline__ T prevElement(const T (&v)[R], int s, uint32_t lane, T first) {
  T p = __shfl_up(v[s], 1);
  const T prevSlotLast = __shfl(v[s > 0 ? s - 1 : 0], 63);
  if (lane == 0) p = s > 0 ? prevSlotLast : first;
  return p;
}

// One bucket of at most 64 R items, by one wave.  The sort runs on one
// 64-bit key per item: the 32 hash bits right below the bucket digit, the
// circuit-local rank and the slot position.  Where every pair of items with
// equal key bits has equal hashes (checked after the sort; two distinct
// hashes sharing those 32 bits within a bucket are rare), that order is the
// (hash, rank) order; otherwise, and for circuits whose ranks do not fit the
// key, the bucket is sorted again on the full (hash, rank).  The items' node
// records are requested as soon as their slots arrive, so that gather's
// latency passes during the sort; slots and records then meet in the wave's
// LDS by slot position, where the merge reads them in sorted order.
// Sorts a wave's 64*R packed 64-bit keys (key[s] is entry s * 64 + lane;
// the bucket stage's common case, where the key packs the hash bits below
// the bucket digit, the circuit-local rank and the slot index, so the cnt
// real keys are distinct and entries past them hold ~0, above all of them):
// each key's sorted position is the number of real keys below it, counted
// against every real key read back from LDS by broadcast -- independent
// reads and two VALU operations per comparison, no exchange chain -- then
// the keys are written to their positions and read back in order.  Uses
// the .b halves of the wave's SlotRec entries (rewritten by the caller).
template <int R>
__device__ __forceinline__ void waveRankSort64(uint64_t (&key)[R], uint32_t lane, uint32_t cnt, SlotRec *pay) {
  typedef uint64_t __attribute__((may_alias)) u64_alias;
#pragma unroll
  for (int s = 0; s < R; ++s)
    *reinterpret_cast<u64_alias *>(&pay[static_cast<uint32_t>(s) * 64 + lane].b.x) = key[s];
  waveLdsSync();
  uint32_t rank[R];
#pragma unroll
  for (int s = 0; s < R; ++s) rank[s] = 0;
#pragma unroll 8
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint64_t kj = *reinterpret_cast<const u64_alias *>(&pay[j].b.x);
#pragma unroll
    for (int s = 0; s < R; ++s) rank[s] += kj < key[s] ? 1u : 0u;
  }
#pragma unroll
  for (int s = 0; s < R; ++s)
    if (static_cast<uint32_t>(s) * 64 + lane < cnt) *reinterpret_cast<u64_alias *>(&pay[rank[s]].b.z) = key[s];
  waveLdsSync();
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    key[s] = e < cnt ? *reinterpret_cast<const u64_alias *>(&pay[e].b.z) : ~0ull;
  }
}

template <int R>
__device__ void bucketWave(uint32_t b, uint32_t bs, uint32_t ob, uint32_t cnt, uint32_t lane, const uint4 first,
                           const uint4 *__restrict__ pslot, const uint4 *__restrict__ info, SlotRec *pay,
                           const RunOut &o, uint32_t *__restrict__ runCount, uint32_t *__restrict__ runNnz,
                           uint32_t *tot) {
  uint64_t key[R], h[R];
  uint32_t r[R], li[R];
  uint32_t meta = 0;
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    key[s] = ~0ull;
    h[s] = ~0ull;
    r[s] = NONE;
    li[s] = e;
    if (e < cnt) {
      const uint4 x = s == 0 ? first : pslot[bs + e];
      pay[e].a = x;
      h[s] = (static_cast<uint64_t>(x.y) << 32) | x.x;
      r[s] = x.z;
      const uint32_t bits = x.w >> kLocalRankBits;
      const uint64_t sub = (h[s] << bits) >> 32;
      key[s] = (sub << 32) | (static_cast<uint64_t>(x.w & kLocalRankNone) << 8) | e;
      if (s == 0) meta = x.w;
    }
  }
  // node records in flight during the sort
  uint4 rec[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    rec[s] = e < cnt ? info[r[s]] : make_uint4(0u, 0u, 0u, 0u);
  }
  // every item of a bucket comes from one circuit (lane 0's item exists)
  const bool packed = (__builtin_amdgcn_readfirstlane(meta) & kLocalRankNone) != kLocalRankNone;
  if (packed) {
    waveRankSort64<R>(key, lane, cnt, pay);
    waveLdsSync();  // (the sorted keys are read before the records overwrite them)
  }
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    if (e < cnt) pay[e].b = rec[s];
  }
  waveLdsSync();
  bool ordered = false;
  double pv[R];
  uint32_t h2v[R], nzv[R];
  if (packed) {
    bool clash = false;
    uint32_t hi[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
      li[s] = static_cast<uint32_t>(key[s]) & 0xFFu;
      hi[s] = static_cast<uint32_t>(key[s] >> 32);
      h[s] = ~0ull;
      r[s] = NONE;
      if (e < cnt) {
        const SlotRec x = pay[li[s]];
        h[s] = (static_cast<uint64_t>(x.a.y) << 32) | x.a.x;
        r[s] = x.a.z;
        pv[s] = infoProb(x.b);
        h2v[s] = x.b.z;
        nzv[s] = x.b.w;
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
      const uint32_t phi = prevElement<R>(hi, s, lane, 0u);
      const uint64_t ph = prevElement<R>(h, s, lane, static_cast<uint64_t>(0));
      if (e < cnt && e > 0 && phi == hi[s] && ph != h[s]) clash = true;
    }
    ordered = __ballot(clash) == 0ull;
  }
  if (!ordered) {
    waveBitonicSort<R>(h, r, li, lane);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
      if (e < cnt) {
        const uint4 x = pay[li[s]].b;
        pv[s] = infoProb(x);
        h2v[s] = x.z;
        nzv[s] = x.w;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    if (e >= cnt) {
      pv[s] = 0.0;
      h2v[s] = 0u;
      nzv[s] = 0u;
    }
  }
  // the members' values by sorted position over the wave's LDS records
  // (read for runs across rows of 64; every record read is done), then
  // hsRunMerge's slots
  double *vals = reinterpret_cast<double *>(pay);
  if constexpr (R > 1) {
    waveLdsSync();
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
      if (e < cnt) vals[e] = pv[s];
    }
    waveLdsSync();
  }
  RunCarry cy{0.0, 0ull, 0u, 0u, 0u, 0u, 0u};
  scanRuns<R>(h, r, pv, h2v, nzv, cnt, 0u, ob, cy, lane, o, LdsRunValues{vals, vals + 64 * R});
  if (lane == 0) {
    runCount[b] = cy.runs;
    runNnz[b] = cy.nnzSum;
    tot[0] = cy.runs;
    tot[1] = cy.nnzSum;
  }
}

struct BucketIo {
  uint32_t *bucketCount;  // read and reset to zero for the next run
  uint32_t countStride;
  const uint4 *pslot;     // slots written by classify
  const uint4 *info;      // node records written by classify
  const uint4 *ovSlot;    // overflow list (count in flags[3]) and each entry's bucket
  const uint32_t *ovBucket;
  uint32_t *runCount, *runNnz, *itemCount;  // per bucket
  uint32_t *slotBase;     // per bucket: where its sorted items and runs start (compact: set by the
                          // scan of the bucket counts before this stage)
  uint32_t *wgSum;        // per workgroup: its buckets' runs, targets, items
  uint64_t *scratchKey;   // sort workspace for buckets too large for LDS
  uint32_t *scratchRank;
  uint32_t arenaBase;     // numBuckets * cap: oversized buckets' items and runs follow
  uint32_t cap;           // slots per bucket (at most kBucketCap)
  uint32_t compact;       // items and runs packed by bucket at slotBase (no arena)
  uint32_t ovCap;         // overflow list entries held
  uint64_t scratchCap;    // workspace entries held
};

// An oversized bucket (more than its slots' worth of items), by the whole
// workgroup: its items (slots, then its entries of the overflow list) are
// gathered in LDS (up to kBigLdsItems) or a global workspace range taken
// with one atomic, bitonic-sorted there on (hash, rank), and merged chunk by
// chunk by wave 0 with the items' node records; the sorted items and runs
// go to the bucket's packed range or an arena range taken with one atomic.
// Bitonic sort of P (a power of two) (hash, rank) pairs by the workgroup.
template <typename KP, typename RP>
__device__ __forceinline__ void bitonicWG(KP K, RP Rk, uint32_t P) {
  for (uint32_t k = 2; k <= P; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool asc = (i & k) == 0;
          const uint64_t ka = K[i], kb = K[ixj];
          const uint32_t ra = Rk[i], rb = Rk[ixj];
          const bool sw = asc ? keyLess(kb, rb, ka, ra) : keyLess(ka, ra, kb, rb);
          if (sw) {
            K[i] = kb;
            K[ixj] = ka;
            Rk[i] = rb;
            Rk[ixj] = ra;
          }
        }
      }
      __syncthreads();
    }
}

// Items of an oversized bucket sorted in LDS when they fit this many
// (`lds`: 12 bytes each), else in the global workspace.
constexpr uint32_t kBigLdsItems = 2048;

__device__ void bigBucket(uint32_t b, uint32_t cnt, const BucketIo &io, const RunOut &o, uint32_t *tot,
                          uint64_t *ldsKeys, uint32_t *ldsRanks) {
  __shared__ uint32_t shOff, shFill, shWs;
  __shared__ double hsSlots[32];
  const uint32_t lane = threadIdx.x & 63;
  uint32_t P = 1;
  while (P < cnt) P <<= 1;
  // sorted in LDS when the items fit, else in the global workspace
  const bool inLds = P <= kBigLdsItems;
  if (threadIdx.x == 0) {
    // its range: packed by bucket, or taken from the arena (total items
    // never exceed it); its workspace range, if it needs one
    shOff = io.compact ? io.slotBase[b] : atomicAdd(&o.flags[5], cnt);
    shFill = io.cap;
    shWs = inLds ? 0u : atomicAdd(&o.flags[kFlagScratchNeed], P);
  }
  __syncthreads();
  const uint32_t off = shOff, ws = shWs;
  if (!inLds && static_cast<uint64_t>(ws) + P > io.scratchCap) {
    // the workspace is full: the bucket is left empty, and the host grows
    // the workspace to what the run asked for and runs again
    if (threadIdx.x == 0) {
      io.runCount[b] = 0;
      io.runNnz[b] = 0;
      io.slotBase[b] = io.compact ? off : io.arenaBase + off;
      tot[0] = 0;
      tot[1] = 0;
    }
    __syncthreads();
    return;
  }
  uint64_t *K = inLds ? ldsKeys : io.scratchKey + ws;
  uint32_t *Rk = inLds ? ldsRanks : io.scratchRank + ws;
  // the items: its slots, then its entries of the overflow list, padded
  // to P
  const uint32_t slot0 = b * io.cap;
  for (uint32_t i = threadIdx.x; i < io.cap; i += blockDim.x) {
    const uint4 x = io.pslot[slot0 + i];
    K[i] = (static_cast<uint64_t>(x.y) << 32) | x.x;
    Rk[i] = x.z;
  }
  // (a list cut at its capacity holds fewer: the run is repeated)
  const uint32_t nov = min(o.flags[3], io.ovCap);
  for (uint32_t k = threadIdx.x; k < nov; k += blockDim.x)
    if (io.ovBucket[k] == b) {
      const uint32_t at = atomicAdd(&shFill, 1u);
      const uint4 x = io.ovSlot[k];
      K[at] = (static_cast<uint64_t>(x.y) << 32) | x.x;
      Rk[at] = x.z;
    }
  __syncthreads();
  const uint32_t got = shFill;  // == cnt unless the list was cut
  for (uint32_t i = got + threadIdx.x; i < P; i += blockDim.x) {
    K[i] = ~0ull;
    Rk[i] = NONE;
  }
  __syncthreads();
  if (inLds) bitonicWG(ldsKeys, ldsRanks, P);
  else bitonicWG(K, Rk, P);
  const uint32_t base = io.compact ? off : io.arenaBase + off;
  if (threadIdx.x < 64) {
    RunCarry cy{0.0, 0ull, 0u, 0u, 0u, 0u, 0u};
    // members' values for runs across chunks: from the sorted ranks' records
    const NodeRunValues members{Rk, io.info, hsSlots};
    for (uint32_t c0 = 0; c0 < got; c0 += 64) {
      uint64_t h[1];
      uint32_t r[1], h2v[1] = {0u}, nzv[1] = {0u};
      double pv[1] = {0.0};
      const uint32_t e = c0 + lane;
      h[0] = e < got ? K[e] : ~0ull;
      r[0] = e < got ? Rk[e] : NONE;
      if (e < got) {
        const uint4 x = io.info[r[0]];
        pv[0] = infoProb(x);
        h2v[0] = x.z;
        nzv[0] = x.w;
      }
      const uint32_t n = got - c0 < 64 ? got - c0 : 64;
      scanRuns<1>(h, r, pv, h2v, nzv, n, c0, base, cy, lane, o, members);
    }
    if (lane == 0) {
      io.runCount[b] = cy.runs;
      io.runNnz[b] = cy.nnzSum;
      io.slotBase[b] = base;
      tot[0] = cy.runs;
      tot[1] = cy.nnzSum;
    }
  }
  __syncthreads();
}

// One wave64 per bucket: sort in registers, merge runs, reading the
// bucket's slots.  Buckets above kBucketCap items are then taken by their
// workgroup (bigBucket).  Each bucket's counter is reset for the next run.
__global__ __launch_bounds__(256) void bucketKernel(uint32_t numBuckets, BucketIo io, RunOut o) {
  __shared__ uint32_t bigB[4], bigCnt[4], nBig;
  __shared__ uint32_t tot[4][3];          // each wave's bucket: runs, targets, items
  __shared__ SlotRec pay[4][kBucketCap];  // each wave's bucket records, by slot position
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t b = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  if (threadIdx.x == 0) nBig = 0;
  if (lane < 3) tot[wave][lane] = 0u;
  __syncthreads();
  if (b < numBuckets) {
    const uint32_t bs = b * io.cap;
    // the bucket's first 64 slots travel with its counter (slots past the
    // count hold stale records and are never used), and its packed range
    const uint4 first = io.pslot[bs + lane];
    uint32_t *ctr = io.bucketCount + static_cast<uint64_t>(b) * io.countStride;
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(*ctr);
    const uint32_t ob = io.compact ? __builtin_amdgcn_readfirstlane(io.slotBase[b]) : bs;
    if (lane == 0) {
      *ctr = 0u;
      io.itemCount[b] = cnt;
      tot[wave][2] = cnt;
    }
    if (cnt <= io.cap) {
      if (lane == 0 && !io.compact) io.slotBase[b] = bs;
      if (cnt == 0) {
        if (lane == 0) {
          io.runCount[b] = 0;
          io.runNnz[b] = 0;
        }
      } else if (cnt <= 64) {
        bucketWave<1>(b, bs, ob, cnt, lane, first, io.pslot, io.info, pay[wave], o, io.runCount, io.runNnz, tot[wave]);
      } else if (cnt <= 128) {
        bucketWave<2>(b, bs, ob, cnt, lane, first, io.pslot, io.info, pay[wave], o, io.runCount, io.runNnz, tot[wave]);
      } else {
        bucketWave<4>(b, bs, ob, cnt, lane, first, io.pslot, io.info, pay[wave], o, io.runCount, io.runNnz, tot[wave]);
      }
    } else if (lane == 0) {
      const uint32_t k = atomicAdd(&nBig, 1u);
      bigB[k] = b;
      bigCnt[k] = cnt;
    }
  }
  __syncthreads();
  const uint32_t n = nBig;
  // (the waves' LDS records are done with: the oversized buckets sort there)
  static_assert(sizeof(pay) >= kBigLdsItems * 12, "LDS for oversized buckets");
  uint64_t *ldsKeys = reinterpret_cast<uint64_t *>(&pay[0][0]);
  uint32_t *ldsRanks = reinterpret_cast<uint32_t *>(ldsKeys + kBigLdsItems);
  for (uint32_t k = 0; k < n; ++k)
    bigBucket(bigB[k], bigCnt[k], io, o, tot[bigB[k] - blockIdx.x * 4], ldsKeys, ldsRanks);
  // the workgroup's totals, from which the emit stage forms the offsets of
  // its buckets
  __syncthreads();
  if (threadIdx.x < 3)
    io.wgSum[3 * blockIdx.x + threadIdx.x] =
        tot[0][threadIdx.x] + tot[1][threadIdx.x] + tot[2][threadIdx.x] + tot[3][threadIdx.x];
}

// Writes the compact output: per class (run) its probability, hash, target
// offset and first member, and its targets (from the representative's
// stashed row, or from S for classes above kStashTargets targets).  One wave
// per bucket, one lane per run, four buckets per workgroup (the bucket
// kernel's grouping).  The buckets' class, target and member offsets come
// either from the scans (classStart, tgtStart, memberStart) or, when wgSum
// is set, from the bucket kernel's per-workgroup totals: workgroup g adds up
// those of workgroups 0..g-1 (fine for up to kEmitPrefixGroups workgroups)
// and writes memberStart for its buckets.  Each circuit's class and target
// bounds go to vBounds from the workgroup holding its first bucket.
constexpr uint32_t kEmitPrefixGroups = 2048;

struct EmitIn {
  uint32_t numVariants, numBuckets;
  const uint32_t *wgSum;  // or nullptr: offsets from the scan arrays
  const uint32_t *slotBase, *runCount, *itemCount, *classStart, *tgtStart;
  uint32_t *memberStart;
  const uint32_t *runRank, *runTgtPre, *runMember;
  const double *runP;
  const uint64_t *runHash;
  const uint32_t *runNnz, *S, *stash;
  const uint2 *rankCode;
  const uint4 *spill;
  uint32_t stashMaxCols;
  uint32_t stashShift;  // row words = 1 << stashShift
  const int2 *band;
  const int32_t *colAdj;
  uint32_t bucketCap;  // slots per bucket (0: runs packed by bucket, read ahead from entry 0)
};
struct EmitOut {
  double *p;
  uint32_t *off;
  uint64_t *hash;
  uint32_t *member;
  uint32_t *targets;
  uint16_t *targets16;  // non-null: targets are written here, 16 bits each
  uint64_t capTargets;
  uint32_t *flags, *vBounds;
};

constexpr uint32_t kEmitStageWords = 1024;  // per wave: a bucket's targets staged in LDS

__global__ __launch_bounds__(256) void emitKernel(const VDesc *__restrict__ desc, EmitIn in, EmitOut out) {
  __shared__ uint32_t red[3][4];
  // a bucket's targets are assembled in LDS and written out as one
  // contiguous, coalesced range (rows of scattered 2-byte stores otherwise)
  __shared__ uint32_t stage[4][kEmitStageWords];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g = blockIdx.x;
  const uint32_t b = __builtin_amdgcn_readfirstlane(g * 4 + wave);
  const bool live = b < in.numBuckets;
  // Everything that depends on no other load is requested first, so the
  // wave waits for one round trip before its target rows: the four buckets'
  // counts (lane q: bucket 4g + q), this bucket's first 64 runs as they lie
  // when its runs start at its own slots (every bucket but an oversized one),
  // the circuit owning the bucket, and the preceding workgroups' totals.
  uint32_t qRuns = 0, qNnz = 0, qItems = 0, qBase = 0;
  if (lane < 4 && g * 4 + lane < in.numBuckets) {
    qRuns = in.runCount[g * 4 + lane];
    qNnz = in.runNnz[g * 4 + lane];
    qItems = in.itemCount[g * 4 + lane];
    qBase = in.slotBase[g * 4 + lane];
  }
  const uint32_t bs0 = b * in.bucketCap;
  uint32_t sRank = 0, sPre = 0, sNext = 0, sMember = 0;
  double sP = 0.0;
  uint64_t sHash = 0;
  if (live) {
    sRank = in.runRank[bs0 + lane];
    sPre = in.runTgtPre[bs0 + lane];
    sNext = in.runTgtPre[bs0 + lane + 1];
    sP = in.runP[bs0 + lane];
    sHash = in.runHash[bs0 + lane];
    sMember = in.runMember[bs0 + lane];
  }
  // circuit owning bucket b
  uint32_t lo = 0;
  if (live) {
    uint32_t hi = in.numVariants - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (desc[mid].bucketBase <= b) lo = mid;
      else hi = mid - 1;
    }
  }
  // offsets (classes, targets, members) of this workgroup's first bucket
  uint32_t base[3] = {0u, 0u, 0u};
  if (in.wgSum) {
    uint32_t part[3] = {0u, 0u, 0u};
    // eight totals per thread per pass, all loads in flight together (one
    // round trip for up to 2,048 preceding workgroups)
    for (uint32_t i0 = threadIdx.x; i0 < g; i0 += 8 * blockDim.x) {
      uint32_t x[8][3];
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t i = i0 + u * blockDim.x;
#pragma unroll
        for (int q = 0; q < 3; ++q) x[u][q] = i < g ? in.wgSum[3 * i + q] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u)
#pragma unroll
        for (int q = 0; q < 3; ++q) part[q] += x[u][q];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint32_t t = waveInclusiveSum(part[q], lane);
      if (lane == 63) red[q][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) base[q] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
  }
  if (in.numBuckets == 0) {
    if (g == 0 && threadIdx.x == 0) {
      for (uint32_t v = 0; v <= in.numVariants; ++v) out.vBounds[2 * v] = out.vBounds[2 * v + 1] = 0u;
      out.off[0] = 0u;
      out.member[0] = 0u;
    }
    return;
  }
  if (!live) return;
  // this bucket's counts, and the workgroup's buckets before it
  const uint32_t nR = __shfl(qRuns, static_cast<int>(wave)), Tb = __shfl(qNnz, static_cast<int>(wave));
  const uint32_t nItems = __shfl(qItems, static_cast<int>(wave)), bs = __shfl(qBase, static_cast<int>(wave));
  uint32_t cb, tb, mb;
  if (in.wgSum) {
    cb = base[0];
    tb = base[1];
    mb = base[2];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      const uint32_t r = __shfl(qRuns, w), t = __shfl(qNnz, w), it = __shfl(qItems, w);
      if (static_cast<uint32_t>(w) < wave) {
        cb += r;
        tb += t;
        mb += it;
      }
    }
    if (lane == 0) in.memberStart[b] = mb;
  } else {
    cb = in.classStart[b];
    tb = in.tgtStart[b];
    mb = in.memberStart[b];
  }
  const VDesc d = desc[lo];
  if (lane == 0) {
    if (d.bucketBase == b) {  // the circuit's first bucket: its bounds
      out.vBounds[2 * lo] = cb;
      out.vBounds[2 * lo + 1] = tb;
    }
    if (b + 1 == in.numBuckets) {  // totals
      const uint32_t C = cb + nR, T = tb + Tb, M = mb + nItems;
      out.vBounds[2 * in.numVariants] = C;
      out.vBounds[2 * in.numVariants + 1] = T;
      out.off[C] = T;
      out.member[C] = M;
    }
  }
  // targets of this bucket: staged when they fit the wave's LDS share
  const bool staged = Tb <= (out.targets16 ? 2 * kEmitStageWords : kEmitStageWords);
  typedef uint16_t __attribute__((may_alias)) u16_alias;
  u16_alias *st16 = reinterpret_cast<u16_alias *>(stage[wave]);
  uint32_t *st32 = stage[wave];
  auto put = [&](uint64_t at, uint32_t t) {
    if (staged) {
      const uint32_t i = static_cast<uint32_t>(at - tb);
      if (out.targets16) st16[i] = static_cast<uint16_t>(t);
      else st32[i] = t;
    } else if (at < out.capTargets) {
      if (out.targets16) out.targets16[at] = static_cast<uint16_t>(t);
      else out.targets[at] = t;
    }
  };
  for (uint32_t r = lane; r < nR; r += 64) {
    const uint32_t c = cb + r;
    uint32_t rk, pre, next, mem;
    double pr;
    uint64_t hh;
    if (r < 64 && bs == bs0) {
      rk = sRank;
      pre = sPre;
      next = sNext;
      pr = sP;
      hh = sHash;
      mem = sMember;
    } else {
      rk = in.runRank[bs + r];
      pre = in.runTgtPre[bs + r];
      next = r + 1 < nR ? in.runTgtPre[bs + r + 1] : 0u;
      pr = in.runP[bs + r];
      hh = in.runHash[bs + r];
      mem = in.runMember[bs + r];
    }
    if (r + 1 == nR) next = Tb;
    uint64_t pos = tb + pre;
    out.p[c] = pr;
    out.hash[c] = hh;
    out.off[c] = static_cast<uint32_t>(pos);
    out.member[c] = mb + mem;
    // class size: the gap to the next run's target offset
    const uint32_t n = next - pre;
    const uint32_t W = 1u << in.stashShift;
    if (n <= 2 * W && d.numCols <= in.stashMaxCols) {
      // the representative's stashed row (4 W bytes, 16-bit targets)
      const uint4 *src = reinterpret_cast<const uint4 *>(in.stash + static_cast<uint64_t>(rk) * W);
      uint32_t w[kStashWords] = {};
#pragma unroll
      for (uint32_t q = 0; q < kStashWords / 4; ++q) {
        if (4 * q >= W) break;
        const uint4 x = src[q];
        w[4 * q] = x.x;
        w[4 * q + 1] = x.y;
        w[4 * q + 2] = x.z;
        w[4 * q + 3] = x.w;
      }
#pragma unroll
      for (uint32_t j = 0; j < kStashTargets; ++j)
        if (j < n) put(pos + j, (w[j / 2] >> (16 * (j % 2))) & 0xFFFFu);
      pos += n;
    } else {
      const FrameRows fr = decodeRows(expandRankCode(in.rankCode[rk], in.spill + d.spillOff), d.layerStride);
      forEachTargetScan(d.numCols, in.band + d.colOff, in.colAdj + d.colOff, d.layerStride, fr, in.S + d.sOff,
                        [&](uint32_t tg) {
        put(pos, tg);
        pos++;
      });
    }
    if (pos > out.capTargets) atomicOr(&out.flags[1], 1u);
  }
  if (staged) {
    // every lane's staged targets visible to the wave before the read-out
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    for (uint32_t i = lane; i < Tb; i += 64) {
      const uint64_t at = static_cast<uint64_t>(tb) + i;
      if (at < out.capTargets) {
        if (out.targets16) out.targets16[at] = st16[i];
        else out.targets[at] = st32[i];
      }
    }
  }
}

// Compact copy of the bucket-sorted members (for the differentiable path):
// bucket b's itemCount[b] sorted ranks, from its slots, to memberStart[b].
__global__ __launch_bounds__(256) void compactMembersKernel(uint32_t numBuckets, const uint32_t *__restrict__ slotBase,
                                                            const uint32_t *__restrict__ itemCount,
                                                            const uint32_t *__restrict__ memberStart,
                                                            const uint32_t *__restrict__ sortedRank,
                                                            uint32_t *__restrict__ dst) {
  const uint32_t b = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (b >= numBuckets) return;
  const uint32_t lane = threadIdx.x & 63, n = itemCount[b], s = slotBase[b], m = memberStart[b];
  for (uint32_t k = lane; k < n; k += 64) dst[m + k] = sortedRank[s + k];
}

// Exclusive scan of three u32 arrays (scanned independently), tile level:
// 4096 items per 1024-thread block.  With one tile it also writes the totals
// to out[n]; otherwise the tile totals go to tileSums (interleaved triples).
constexpr int kScanArrays = 3;
struct ScanArrays {
  const uint32_t *in[kScanArrays];
  uint32_t *out[kScanArrays];
  uint32_t stride = 1;              // words between consecutive entries of the inputs
  uint32_t count = kScanArrays;     // arrays in use (the first `count`)
};

__global__ __launch_bounds__(kScanBlock) void scanTileKernel(ScanArrays a, uint32_t n,
                                                             uint32_t *__restrict__ tileSums) {
  __shared__ uint32_t waveSums[kScanArrays][kScanBlock / 64];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t v[kScanArrays][4], t[kScanArrays], incl[kScanArrays];
  // every load first (one memory round trip), then the arithmetic
#pragma unroll
  for (int q = 0; q < kScanArrays; ++q)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[q][k] = q < a.count ? a.in[q][(base + k < n ? base + k : 0u) * a.stride] : 0u;
#pragma unroll
  for (int q = 0; q < kScanArrays; ++q) {
    t[q] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (base + k >= n) v[q][k] = 0u;
      t[q] += v[q][k];
    }
    incl[q] = waveInclusiveSum(t[q], lane);
    if (lane == 63) waveSums[q][wave] = incl[q];
  }
  __syncthreads();
  if (wave == 0) {
    constexpr uint32_t W = kScanBlock / 64;
#pragma unroll
    for (int q = 0; q < kScanArrays; ++q) {
      const uint32_t x = lane < W ? waveSums[q][lane] : 0u;
      const uint32_t y = waveInclusiveSum(x, lane);
      if (lane < W) waveSums[q][lane] = y - x;
      if (lane == W - 1) {
        if (gridDim.x > 1) tileSums[kScanArrays * blockIdx.x + q] = y;
        else if (q < a.count) a.out[q][n] = y;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kScanArrays; ++q) {
    if (q >= a.count) break;
    uint32_t r = waveSums[q][wave] + incl[q] - t[q];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (base + k < n) a.out[q][base + k] = r;
      r += v[q][k];
    }
  }
}

// Second level: scans the tile totals in one block; sums[3 * numTiles + q]
// gets the grand totals.
__global__ __launch_bounds__(kScanBlock) void scanTileSumsKernel(uint32_t *__restrict__ sums,
                                                                 uint32_t numTiles) {
  __shared__ uint32_t carry[kScanArrays];
  __shared__ uint32_t ws[kScanArrays][kScanBlock / 64];
  if (threadIdx.x < kScanArrays) carry[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t base = 0; base < numTiles; base += kScanBlock) {
    const uint32_t t = base + threadIdx.x;
    uint32_t x[kScanArrays], incl[kScanArrays];
#pragma unroll
    for (int q = 0; q < kScanArrays; ++q) {
      x[q] = t < numTiles ? sums[kScanArrays * t + q] : 0u;
      incl[q] = waveInclusiveSum(x[q], lane);
      if (lane == 63) ws[q][wave] = incl[q];
    }
    __syncthreads();
    if (wave == 0) {
      constexpr uint32_t W = kScanBlock / 64;
#pragma unroll
      for (int q = 0; q < kScanArrays; ++q) {
        const uint32_t y0 = lane < W ? ws[q][lane] : 0u;
        const uint32_t y = waveInclusiveSum(y0, lane);
        if (lane < W) ws[q][lane] = y - y0;
      }
    }
    __syncthreads();
    uint32_t c[kScanArrays];
#pragma unroll
    for (int q = 0; q < kScanArrays; ++q) {
      c[q] = carry[q];
      if (t < numTiles) sums[kScanArrays * t + q] = c[q] + ws[q][wave] + incl[q] - x[q];
    }
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) {
#pragma unroll
      for (int q = 0; q < kScanArrays; ++q) carry[q] = c[q] + ws[q][wave] + incl[q];
    }
    __syncthreads();
  }
  if (threadIdx.x < kScanArrays) sums[kScanArrays * numTiles + threadIdx.x] = carry[threadIdx.x];
}

__global__ __launch_bounds__(kScanBlock) void scanAddKernel(ScanArrays a, uint32_t n,
                                                            const uint32_t *__restrict__ tileOff,
                                                            uint32_t numTiles) {
  const uint32_t tile = blockIdx.x;
  const uint32_t base = tile * kScanTile;
#pragma unroll
  for (int q = 0; q < kScanArrays; ++q) {
    if (q >= a.count) break;
    const uint32_t add = tileOff[kScanArrays * tile + q];
    for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanBlock)
      if (base + k < n) a.out[q][base + k] += add;
    if (tile == 0 && threadIdx.x == 0) a.out[q][n] = tileOff[kScanArrays * numTiles + q];
  }
}

// Differentiable path: segmented XOR-merge and its adjoint.
__global__ void xorMergeFwdKernel(const double *__restrict__ values, const uint32_t *__restrict__ index,
                                  const uint32_t *__restrict__ segStart, uint32_t numSegments,
                                  double *__restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= numSegments) return;
  double p = 0.0;
  for (uint32_t k = segStart[s]; k < segStart[s + 1]; ++k) p = xmerge(p, values[index[k]]);
  out[s] = p;
}

__global__ void xorMergeBwdKernel(const double *__restrict__ values, const uint32_t *__restrict__ index,
                                  const uint32_t *__restrict__ segStart, uint32_t numSegments,
                                  const double *__restrict__ gradOut, double *__restrict__ gradValues) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= numSegments) return;
  const uint32_t a = segStart[s], e = segStart[s + 1];
  const double g = gradOut[s];
  // d out / d v_k = prod_{j != k} (1 - 2 v_j): prefix products, then suffix
  double pre = 1.0;
  for (uint32_t k = a; k < e; ++k) {
    gradValues[index[k]] = pre;
    pre *= 1.0 - 2.0 * values[index[k]];
  }
  double suf = 1.0;
  for (uint32_t k = e; k > a; --k) {
    const uint32_t i = index[k - 1];
    gradValues[i] = g * gradValues[i] * suf;
    suf *= 1.0 - 2.0 * values[i];
  }
}

// Rate of one fault component from its location's rate t, and d rate / d t
// (Mixer.hpp:164, 184): kind 0 the rate itself; 1 one of DEPOLARIZE1's
// three Paulis, 1/2 - sqrt(1 - 4t/3)/2; 2 one of DEPOLARIZE2's fifteen,
// 1/2 - (1 - 16t/15)^(1/8)/2.  The root's argument is clamped at 1e-300
// (zero derivative there), as the plain tensor formula does.  The formula
// is the reference's (its 1/2 - x/2 cancels for small rates, ~1e-11
// relative at t = 1e-4) and is evaluated operation by operation, without
// fused multiply-adds, so the rates round as the compile path's do.
__device__ __forceinline__ double componentRate(double t, uint32_t kind, double &dr) {
  if (kind == 1) {
    const double u = __dsub_rn(1.0, __ddiv_rn(__dmul_rn(4.0, t), 3.0));
    const double s = sqrt(u > 1e-300 ? u : 1e-300);
    dr = u > 1e-300 ? 1.0 / (3.0 * s) : 0.0;
    return __dsub_rn(0.5, __dmul_rn(0.5, s));
  }
  if (kind == 2) {
    const double u = __dsub_rn(1.0, __ddiv_rn(__dmul_rn(16.0, t), 15.0));
    const double uc = u > 1e-300 ? u : 1e-300;
    const double s = pow(uc, 0.125);
    dr = u > 1e-300 ? s / (15.0 * uc) : 0.0;
    return __dsub_rn(0.5, __dmul_rn(0.5, s));
  }
  dr = 1.0;
  return t;
}

// q[n] = XOR-merge over node n's components c (nodeComp[nodeStart[n] ..
// nodeStart[n+1])) of componentRate(theta[compLoc[c]], compKind[c]).
__global__ void nodeRatesFwdKernel(const double *__restrict__ theta, const uint32_t *__restrict__ compLoc,
                                   const uint8_t *__restrict__ compKind, const uint32_t *__restrict__ nodeStart,
                                   const uint32_t *__restrict__ nodeComp, uint32_t numNodes, double *__restrict__ q) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= numNodes) return;
  double p = 0.0, dr;
  for (uint32_t k = nodeStart[n]; k < nodeStart[n + 1]; ++k) {
    const uint32_t c = nodeComp[k];
    p = xmerge(p, componentRate(theta[compLoc[c]], compKind[c], dr));
  }
  q[n] = p;
}

// gradComp[c] = dL/dq[n] * d q[n] / d rate_c * d rate_c / d t for every
// component c of node n (each component belongs to one node): d q / d
// rate_c = prod over the node's other components of (1 - 2 rate).
__global__ void nodeRatesBwdKernel(const double *__restrict__ theta, const uint32_t *__restrict__ compLoc,
                                   const uint8_t *__restrict__ compKind, const uint32_t *__restrict__ nodeStart,
                                   const uint32_t *__restrict__ nodeComp, uint32_t numNodes,
                                   const double *__restrict__ gradQ, double *__restrict__ gradComp) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= numNodes) return;
  const uint32_t a = nodeStart[n], e = nodeStart[n + 1];
  const double g = gradQ[n];
  double pre = 1.0, dr;
  for (uint32_t k = a; k < e; ++k) {
    const uint32_t c = nodeComp[k];
    const double r = componentRate(theta[compLoc[c]], compKind[c], dr);
    gradComp[c] = pre * dr;
    pre *= 1.0 - 2.0 * r;
  }
  double suf = 1.0;
  for (uint32_t k = e; k > a; --k) {
    const uint32_t c = nodeComp[k - 1];
    const double r = componentRate(theta[compLoc[c]], compKind[c], dr);
    gradComp[c] = g * gradComp[c] * suf;
    suf *= 1.0 - 2.0 * r;
  }
}

// gradTheta[l] = sum of gradComp over location l's components (locComp[
// locStart[l] .. locStart[l+1]), ascending): one deterministic sum per
// location, no atomics.
__global__ void locationGradKernel(const uint32_t *__restrict__ locStart, const uint32_t *__restrict__ locComp,
                                   const double *__restrict__ gradComp, uint32_t numLocations,
                                   double *__restrict__ gradTheta) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= numLocations) return;
  double s = 0.0;
  for (uint32_t k = locStart[l]; k < locStart[l + 1]; ++k) s += gradComp[locComp[k]];
  gradTheta[l] = s;
}

// --------------------------------------------------------------------------
// host-side buffers
// --------------------------------------------------------------------------

// Device allocations are made in whole 2 MiB units, so every buffer starts
// on its own 2 MiB boundary.  (Measured: left to the runtime, buffers below a
// few MiB are packed into shared small-allocation blocks, and the class
// stage's 246k bucket-counter atomics on bb12 ran 7 us slower when the
// counters landed in such a block.)
constexpr size_t kAllocGrain = size_t{2} << 20;
size_t roundGrain(size_t x) { return (x + kAllocGrain - 1) & ~(kAllocGrain - 1); }

// A device buffer: its own allocation (`owned`), or a range of a pipeline's
// work arena (see WorkArena).
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  bool owned = false;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    owned = false;
  }
  // Its own allocation of at least `bytes` (kept when large enough).
  void ensure(size_t bytes, bool frozen, const char *what) {
    if (bytes <= cap) return;
    if (frozen)
      throw RuntimeError(std::string("circuit exceeds the driver's capacity (") + what +
                         "); create the driver from a larger circuit");
    release();
    const size_t sz = roundGrain(std::max<size_t>(bytes, 256));
    GPA_HIP(hipMalloc(&p, sz));
    cap = sz;
    owned = true;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// The work buffers of one load laid out in a single device allocation: each
// buffer of at least 1 MiB starts on a 2 MiB boundary, smaller ones are
// packed at 4 KiB alignment.  A load requests every buffer's size; the
// layout is kept while every request fits its buffer's capacity, and is
// otherwise redone with each capacity raised to its request (a frozen
// pipeline throws instead).
struct WorkArena {
  struct Req {
    DevBuf *buf;
    size_t bytes;
    const char *what;
  };
  DevBuf mem;
  std::vector<Req> reqs;
  void request(DevBuf &b, size_t bytes, const char *what) { reqs.push_back(Req{&b, std::max<size_t>(bytes, 256), what}); }
  // Arena bytes for buffers of these capacities (the layout commit makes).
  static size_t layoutBytes(const std::vector<size_t> &caps) {
    size_t off = 0;
    for (size_t c : caps) {
      const size_t align = c >= (size_t{1} << 20) ? kAllocGrain : 4096;
      off = (off + align - 1) & ~(align - 1);
      off += c;
    }
    return roundGrain(std::max<size_t>(off, 256));
  }
  // headroom: a buffer that must grow takes this much more than asked
  // (not on the arena's first allocation)
  void commit(bool frozen, double headroom = 1.0) {
    // this load's requests (none stay behind if a capacity check throws)
    std::vector<Req> rq;
    rq.swap(reqs);
    bool fits = mem.p != nullptr;
    for (const Req &r : rq)
      if (r.bytes > r.buf->cap) {
        if (frozen)
          throw RuntimeError(std::string("circuit exceeds the driver's capacity (") + r.what +
                             "); create the driver from a larger circuit");
        fits = false;
      }
    if (!fits) {
      size_t off = 0;
      std::vector<size_t> offs(rq.size()), caps(rq.size());
      for (size_t k = 0; k < rq.size(); ++k) {
        const size_t grown =
            mem.p && rq[k].bytes > rq[k].buf->cap ? static_cast<size_t>(rq[k].bytes * headroom) : 0;
        const size_t c = std::max({rq[k].bytes, rq[k].buf->cap, grown});
        const size_t align = c >= (size_t{1} << 20) ? kAllocGrain : 4096;
        off = (off + align - 1) & ~(align - 1);
        offs[k] = off;
        caps[k] = c;
        off += c;
      }
      mem.release();
      mem.ensure(off, false, "work buffers");
      for (size_t k = 0; k < rq.size(); ++k) {
        rq[k].buf->p = mem.as<char>() + offs[k];
        rq[k].buf->cap = caps[k];
        rq[k].buf->owned = false;
      }
    }
  }
};

struct HostBuf {
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  void *p = nullptr;
  size_t cap = 0;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) GPA_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    const size_t sz = std::max<size_t>(bytes + bytes / 4, 4096);
    GPA_HIP(hipHostMalloc(&p, sz, hipHostMallocDefault));
    cap = sz;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace

struct ResultBlock {
  HostBuf p, off, hash, targets;
};

namespace {

// Result blocks whose holders let go, kept for the next downloads.
struct ResultPool {
  std::mutex mu;
  std::vector<std::unique_ptr<ResultBlock>> free;
};

// A block from the pool (or a new one); when its last handle goes, it
// returns to the pool, or is freed once the pool is gone.
std::shared_ptr<ResultBlock> takeBlock(const std::shared_ptr<ResultPool> &pool) {
  std::unique_ptr<ResultBlock> b;
  {
    std::lock_guard<std::mutex> lk(pool->mu);
    if (!pool->free.empty()) {
      b = std::move(pool->free.back());
      pool->free.pop_back();
    }
  }
  if (!b) b = std::make_unique<ResultBlock>();
  std::weak_ptr<ResultPool> wp = pool;
  return std::shared_ptr<ResultBlock>(b.release(), [wp](ResultBlock *r) {
    if (auto p = wp.lock()) {
      std::lock_guard<std::mutex> lk(p->mu);
      p->free.emplace_back(r);
    } else {
      delete r;
    }
  });
}

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

}  // namespace

// Raises the sweep kernels' dynamic LDS limit to `bytes` (never lowers it:
// other pipelines may launch with more).
static void setSweepLds(size_t bytes) {
  static std::mutex mu;
  static size_t done = 64 * 1024;
  std::lock_guard<std::mutex> lock(mu);
  if (bytes <= done) return;
  done = bytes;
  const void *fns[] = {reinterpret_cast<const void *>(sweepKernel<1, false>),
                       reinterpret_cast<const void *>(sweepKernel<2, false>),
                       reinterpret_cast<const void *>(sweepKernel<4, false>),
                       reinterpret_cast<const void *>(sweepKernel<8, false>),
                       reinterpret_cast<const void *>(sweepKernel<1, true>),
                       reinterpret_cast<const void *>(sweepKernel<2, true>),
                       reinterpret_cast<const void *>(sweepKernel<4, true>),
                       reinterpret_cast<const void *>(sweepKernel<8, true>)};
  for (const void *f : fns)
    GPA_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

struct Pipeline::Impl {
  int device;
  // the stream the pipeline's work is queued on: its own, or a caller's
  // (setStream)
  hipStream_t stream = nullptr, ownStream = nullptr;
  hipEvent_t ev[8] = {};  // [0, 5): stage marks of a timed run; 6, 7: benchDevice
  // upload: the sweep's inputs on `stream`, the error nodes' inputs on
  // `side` (overlapping the sweep; classify waits for evSide)
  hipStream_t side = nullptr;
  hipEvent_t evUp0 = nullptr, evUp1 = nullptr, evPartA = nullptr, evSide = nullptr;
  bool uploadPending = false;  // upload events recorded, not yet read
  bool sidePending = false;    // the main stream has not yet waited for evSide
  // the main stream waits for the side upload (before classify, or before
  // a graph launch: the graph holds no cross-stream wait)
  void joinSide() {
    if (!sidePending) return;
    GPA_HIP(hipStreamWaitEvent(stream, evSide, 0));
    sidePending = false;
  }
  bool frozen = false;
  bool timedRun = false;
  hipEvent_t *evOverride = nullptr;  // per-iteration events of benchDevice
  std::vector<hipEvent_t> evPool;
  // how benchDevice issues back-to-back compiles: -1 not yet measured for
  // this load, 0 direct launches, 1 graph replays
  int launchMode = -1;
  // the whole pipeline captured once per load and replayed
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;

  void dropGraph() {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
  }

  // inputs (one device arena, one pinned staging arena with the same layout)
  DevBuf in;
  HostBuf inHost;
  size_t oDesc = 0, oItems = 0, oTiles = 0, oAct = 0, oSeed = 0, oRankCode = 0, oSpill = 0, oProbIdx = 0,
         oProbVal = 0, inBytes = 0;
  std::vector<uint4> hostItems;  // the sweep items of the load, in launch order
  uint32_t V = 0, nItems = 0, nTiles = 0, B = 0, R = 0, Ccols = 0, rpt = 1;
  bool classTable = true;  // every classify workgroup's mask table fits in LDS
  uint64_t Swords = 0;
  size_t sweepLdsBytes = 0;
  int seedsInLds = 1;
  uint32_t sweepThreads = 64;
  // bucket counters sit countStride words apart (16 bytes when there are
  // few buckets): classify issues one atomic per node; measured on bb12,
  // 16-byte spacing beats both packed counters (same-line atomics
  // serialize) and one counter per 128-byte line (four times the lines)
  uint32_t countStride = 1;

  // per node: info (probability, second hash, class size); stash (target
  // rows, its own allocation: its width grows with a load's largest class);
  // the overflow list (ovSlot/ovBucket) and the oversized buckets'
  // workspace (scratchKey/scratchRank), own allocations sized for the worst
  // case in a driver and grown on demand in a growable load (ovCap and
  // scratchCap entries);
  // per bucket slot (numBuckets * bucketCap): pslot (classify's records);
  // per sorted item, at the bucket's slots plus an arena of numRanks for
  // oversized buckets or, in a compact load, packed by bucket (numRanks in
  // all): prank (the sorted ranks) and the run arrays;
  // per bucket: counters, slotBase, item/run/target counts and their scans
  DevBuf band, stash, info, counters, bucketCnt, ovSlot, ovBucket, pslot, prank,
      runRank, runP, runHash, runTgtPre, runMember, runCount, runNnz, itemCount, slotBase, wgSum, classStart,
      tgtStart, memberStart, outP, outOff, outHash, outMember, outTargets, scratchKey, scratchRank,
      scanTmp, vBounds, members;
  WorkArena work;  // holds the buffers above except stash, the overflow list, the workspace, in,
                   // outTargets, members and the split buffers
  uint64_t ovCap = 0, scratchCap = 0;
  void ensureOverflow(uint64_t n, bool fz) {
    ovSlot.ensure(16 * n, fz, "nodes");
    ovBucket.ensure(4 * n, fz, "nodes");
    ovCap = std::min<uint64_t>({ovSlot.cap / 16, ovBucket.cap / 4, 0xFFFFFFFFull});
  }
  void ensureWorkspace(uint64_t n, bool fz) {
    scratchKey.ensure(8 * n, fz, "nodes");
    scratchRank.ensure(4 * n, fz, "nodes");
    scratchCap = std::min(scratchKey.cap / 8, scratchRank.cap / 4);
  }
  // A compact load places each bucket's sorted items and runs at the
  // exclusive prefix sum of the bucket counts (a scan before the bucket
  // stage) instead of at its slots, and its outputs (p, hash, offsets,
  // members) in pslot's space, which is dead once the bucket stage has run.
  bool compact = false;
  uint64_t aliasHash = 0, aliasOff = 0, aliasMember = 0;
  double *oP() const { return compact ? pslot.as<double>() : outP.as<double>(); }
  uint64_t *oHash() const {
    return compact ? reinterpret_cast<uint64_t *>(pslot.as<char>() + aliasHash) : outHash.as<uint64_t>();
  }
  uint32_t *oOff() const {
    return compact ? reinterpret_cast<uint32_t *>(pslot.as<char>() + aliasOff) : outOff.as<uint32_t>();
  }
  uint32_t *oMember() const {
    return compact ? reinterpret_cast<uint32_t *>(pslot.as<char>() + aliasMember) : outMember.as<uint32_t>();
  }
  uint64_t capTargets = 0;
  // S (its own allocation: it grows when a band needs more rows) and its
  // per-column layout (see colRegion)
  DevBuf S;
  struct ColLayout {
    uint32_t P;     // layer stride
    uint32_t top1;  // 1 + the top layer the sweep stores (0: no seeds)
    uint32_t base;  // first layer stored
  };
  std::vector<ColLayout> cols;  // every column of the load
  size_t oColRegion = 0, oColAdj = 0;
  uint32_t bandRelayouts = 0;  // loads' columns given more rows after a run
  uint32_t bucketCap = kBucketCap;  // slots per sort bucket of the load
  bool uniformCols = false;  // every column of a circuit stores all its layers in regions of one size
  // the arena's buffers, in the order workBufferSizes() sizes them
  std::array<DevBuf *, 25> workBufs() {
    return {&band,      &info,       &counters,    &bucketCnt,  &pslot,       &prank,     &runRank,
            &runP,      &runHash,    &runTgtPre,   &runMember,  &runCount,    &runNnz,    &itemCount,
            &slotBase,  &wgSum,      &classStart,  &tgtStart,   &memberStart, &outP,      &outOff,
            &outHash,   &outMember,  &scanTmp,     &vBounds};
  }

  // host results: the block of the last download, blocks released by
  // their holders for reuse
  std::shared_ptr<ResultBlock> res;
  bool wantHashes = true;  // download the class hashes with the rest
  bool narrowTargets = false;  // 16-bit targets allowed (setNarrowTargets)
  bool targets16 = false;      // this load's targets are 16-bit
  bool sizesKnown = false;  // lastC / lastT hold the DEM size of this load's circuits
  // target rows hold 1 << stashShift words: the widest until a run of this
  // load has shown its largest class, then the narrowest that holds it
  uint32_t stashShift = 4;
  uint64_t lastC = 0, lastT = 0;
  std::shared_ptr<ResultPool> pool = std::make_shared<ResultPool>();
  HostBuf hVB, hFlags;
  uint32_t totalClasses = 0;
  uint64_t totalTargets = 0;
  std::vector<uint32_t> vClass, vTgt;
  std::vector<VDesc> descs;

  // time-chunked sweep of long columns, planned after a load's first run
  // (splitBuf: descriptors | launch-1 items | launch-2 items | per-column
  // split index; splitState: the split columns' state areas)
  bool splitPlanned = false;
  // planned from the first run's bands at the start of the load's second
  // run (outside any graph capture): a load compiled once never pays for it
  bool planPending = false;
  bool capturing = false;
  // columns the class stage loads per round trip (4; 16 once a run has shown
  // nodes see more than 12 columns on average, in loads of few workgroups)
  uint32_t classBatch = 4;
  uint32_t nSplit = 0, nSplit1 = 0, nSplit3 = 0;
  size_t splitLdsBytes = 0, oSplitItems = 0, oSplitLive = 0, oSplitSync = 0;
  DevBuf splitBuf, splitState;
  SplitIo splitIo() const {
    char *b = splitBuf.as<char>();
    if (!nSplit) return SplitIo{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    return SplitIo{reinterpret_cast<const uint4 *>(b + oSplitItems), reinterpret_cast<const SplitDesc *>(b),
                   splitState.as<uint32_t>(), reinterpret_cast<const uint32_t *>(b + oSplitLive),
                   reinterpret_cast<uint32_t *>(b + oSplitSync), nSplit1 + nSplit3};
  }

  uint32_t *flags() const { return counters.as<uint32_t>(); }
  uint32_t *bucketCount() const { return bucketCnt.as<uint32_t>(); }

  // exclusive scans of n entries of three arrays, n + 1 outputs each
  void scan(const ScanArrays &a, uint32_t n) {
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles <= 1) {
      scanTileKernel<<<1, kScanBlock, 0, stream>>>(a, n, nullptr);
      return;
    }
    uint32_t *ts = scanTmp.as<uint32_t>();
    scanTileKernel<<<tiles, kScanBlock, 0, stream>>>(a, n, ts);
    scanTileSumsKernel<<<1, kScanBlock, 0, stream>>>(ts, tiles);
    scanAddKernel<<<tiles, kScanBlock, 0, stream>>>(a, n, ts, tiles);
  }
};

Pipeline::Pipeline(int device, bool growable) : device_(device), growable_(growable), impl_(new Impl) {
  impl_->device = device;
  GPA_HIP(hipSetDevice(device));
  GPA_HIP(hipStreamCreateWithFlags(&impl_->ownStream, hipStreamNonBlocking));
  impl_->stream = impl_->ownStream;
  GPA_HIP(hipStreamCreateWithFlags(&impl_->side, hipStreamNonBlocking));
  for (auto &e : impl_->ev) GPA_HIP(hipEventCreate(&e));
  GPA_HIP(hipEventCreate(&impl_->evUp0));
  GPA_HIP(hipEventCreate(&impl_->evUp1));
  GPA_HIP(hipEventCreateWithFlags(&impl_->evPartA, hipEventDisableTiming));
  GPA_HIP(hipEventCreateWithFlags(&impl_->evSide, hipEventDisableTiming));
}

Pipeline::~Pipeline() {
  (void)hipSetDevice(device_);
  if (impl_->stream) (void)hipStreamSynchronize(impl_->stream);
  if (impl_->side) (void)hipStreamSynchronize(impl_->side);
  for (auto &e : impl_->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {impl_->evUp0, impl_->evUp1, impl_->evPartA, impl_->evSide})
    if (e) (void)hipEventDestroy(e);
  if (impl_->side) (void)hipStreamDestroy(impl_->side);
  for (auto &e : impl_->evPool) (void)hipEventDestroy(e);
  impl_->dropGraph();
  if (impl_->ownStream) (void)hipStreamDestroy(impl_->ownStream);
  delete impl_;
}

void *Pipeline::stream() const { return impl_->stream; }

void Pipeline::setStream(void *stream) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  m.joinSide();
  // work queued so far completes first; later work goes to the new stream
  GPA_HIP(hipStreamSynchronize(m.stream));
  GPA_HIP(hipStreamSynchronize(m.side));
  m.stream = stream ? static_cast<hipStream_t>(stream) : m.ownStream;
}

// Default items per bucket for a load of `loadRanks` error nodes: 64 while
// the bucket stage is latency-bound (a few thousand buckets, one wave each:
// shorter per-wave sorts win; bb12: 79.4 us per compile against 101.7 at
// 128), 192 once there are millions (throughput-bound: fewer, fuller
// buckets halve the emit stage and need fewer slots; the 256-variant
// batch: 7.24 ms and 9.34 GB against 7.34 ms and 10.13 GB at 128, 7.90 ms
// at 64; at 256 buckets overflow their 256 slots and it takes 13.3 ms).
constexpr uint64_t kLargeLoadRanks = uint64_t{1} << 22;
// Whether a load packs its sorted items and runs by bucket (Impl::compact):
// growable loads above kLargeLoadRanks error nodes (the packing scan costs a
// few microseconds; the slot arrays it replaces hold about 3.9 entries per
// node there), or every growable load when compact_items is 2.
static bool compactLoad(bool growable, uint64_t R) {
  const uint32_t mode = gCompactItems.load(std::memory_order_relaxed);
  return growable && (mode == 2 || (mode == 1 && R > kLargeLoadRanks));
}

static uint32_t defaultFill(uint64_t loadRanks) {
  const uint32_t knob = gBucketFill.load(std::memory_order_relaxed);
  return knob ? knob : loadRanks > kLargeLoadRanks ? 192u : 64u;
}

// First radix digit width of a circuit: the smallest b with fill * 2^b >=
// its error nodes (fill: its size class's tuned value, else `fallback`).
static uint32_t bucketBitsFor(const StepgImage &g, uint32_t fallback) {
  uint32_t b = 0;
  const uint64_t nodes = static_cast<uint64_t>(g.numLayers) * g.npl;
  const uint64_t fill = classTuning("bucket_fill", g.level, nodes, fallback);
  while (b < 24 && (fill << b) < g.numRanks) b++;
  return b;
}

// Slots per sort bucket of a load: the largest tuned value among its
// circuits' size classes (bucket_cap where a class has none).  (Measured on
// the 256-variant batch: 128 slots at a mean fill of <= 96 leave the
// buckets of big classes -- many nodes with one hash -- to the oversized
// path, bucket stage 1.45 -> 6.5 ms; every load keeps kBucketCap.)
static uint32_t bucketCapFor(const std::vector<const StepgImage *> &imgs) {
  uint32_t cap = 0;
  for (const StepgImage *g : imgs)
    cap = std::max(cap, classTuning("bucket_cap", g->level, static_cast<uint64_t>(g->numLayers) * g->npl,
                                    gBucketCap.load(std::memory_order_relaxed)));
  return cap ? cap : kBucketCap;
}

void Pipeline::reserve(const std::vector<const StepgImage *> &imgs) {
  impl_->frozen = false;
  load(imgs);
  impl_->frozen = !growable_;
}

static constexpr size_t kLdsBudget = 150 * 1024;  // of gfx950's 160 KB per workgroup

// Growable pipelines (batches, serving slots) that must grow a buffer they
// hold take this much more than asked, so a stream of batches of varying
// size settles after a few loads instead of reallocating on each larger one.
constexpr double kGrowHeadroom = 1.25;

// bucket counters sit this many words apart (16 bytes while there are few
// buckets: see Impl::countStride)
static uint32_t bucketCountStride(uint64_t B) {
  uint32_t cs = 4;
  while (cs > 1 && B * cs > (1u << 17)) cs >>= 1;
  return cs;
}

// Every work buffer of a load, in the arena's order, by its totals: R error
// nodes, B sort buckets of `cap` slots, V circuits, Ccols detector columns.
// req(index, bytes, what) is called once per buffer.
constexpr uint32_t kWorkBuffers = 25;
// `compact`: the load's sorted items and runs packed by bucket, its outputs
// in pslot's space (see Impl::compact).
struct OutAlias {
  uint64_t hash, off, member, bytes;  // byte offsets in pslot; bytes in all
};
static OutAlias outAlias(uint64_t R) {
  OutAlias a;
  a.hash = align256(8 * R);
  a.off = a.hash + align256(8 * R);
  a.member = a.off + align256(4 * (R + 1));
  a.bytes = a.member + align256(4 * (R + 1));
  return a;
}
template <typename F>
static void workBufferSizes(uint64_t R, uint64_t B, uint64_t V, uint64_t Ccols, uint32_t countStride, uint32_t cap,
                            bool compact, F &&req) {
  R = std::max<uint64_t>(R, 1);
  const uint64_t bucketSlots = B * static_cast<uint64_t>(cap);
  if (bucketSlots + R >= 0xFFFFFFFFull) throw RuntimeError("too many sort buckets in one load");
  // sorted items and runs: at the bucket slots, then an arena of R entries
  // for oversized buckets; or packed by bucket (R, and the 65 entries the
  // emit stage reads ahead from the start)
  const uint64_t slots = compact ? R + 64 : bucketSlots + R;
  const uint64_t outs = compact ? 0 : R;  // outputs in their own buffers (else in pslot's space)
  uint32_t k = 0;
  req(k++, 8 * std::max<uint64_t>(Ccols, 1), "columns");                // band
  req(k++, 16 * R, "nodes");                                             // info
  req(k++, 4 * kFlagWords, "buckets");                                   // counters
  // (B + 1) * countStride <= 2^17 + 32 words whatever B is, so a reserved
  // pipeline fits every later load
  req(k++, 4 * std::max<uint64_t>((B + 1) * countStride, (1u << 17) + 32), "buckets");  // bucketCnt
  req(k++, std::max<uint64_t>(16 * bucketSlots, compact ? outAlias(R).bytes : 0), "nodes");  // pslot
  req(k++, 4 * slots, "nodes");                                          // prank
  req(k++, 4 * slots, "nodes");                                          // runRank
  req(k++, 8 * slots, "nodes");                                          // runP
  req(k++, 8 * slots, "nodes");                                          // runHash
  req(k++, 4 * slots, "nodes");                                          // runTgtPre
  req(k++, 4 * slots, "nodes");                                          // runMember
  for (int q = 0; q < 4; ++q) req(k++, 4 * (B + 1), "buckets");          // runCount, runNnz, itemCount, slotBase
  req(k++, 4 * 3 * ((B + 3) / 4 + 1), "buckets");                        // wgSum
  for (int q = 0; q < 3; ++q) req(k++, 4 * (B + 1), "buckets");          // classStart, tgtStart, memberStart
  req(k++, 8 * outs, "nodes");                                           // outP
  req(k++, 4 * (outs + 1), "nodes");                                     // outOff
  req(k++, 8 * outs, "nodes");                                           // outHash
  req(k++, 4 * (outs + 1), "nodes");                                     // outMember
  req(k++, 4 * kScanArrays * ((B + kScanTile - 1) / kScanTile + 2), "buckets");  // scanTmp
  req(k++, 4 * (2 * V + 8), "circuits");                                 // vBounds
}

// The layers of S a column stores: base .. top, top = the odd layer of the
// sweep's first pair.  Banded (growable pipelines): base = the column's
// lowest seed layer less twice the circuit's seed gap and 4 layers (made
// even), enough for the bands of memory experiments, whose detectors
// compare one round's measurements with the round before; columns holding
// observables, whose bands reach back to the start, and unbanded (frozen)
// pipelines store every layer (base 0).  A band that reaches below base
// anyway is reported by the sweep, and the load stores more rows.
static Pipeline::Impl::ColLayout columnLayout(const StepgImage &g, uint32_t c, bool banded) {
  const int32_t hi = g.colHi[c];
  if (hi < 0) return {g.layerStride, 0, 0};
  const uint32_t top1 = 2 * static_cast<uint32_t>(hi / 2) + 2;
  uint32_t base = 0;
  const uint32_t mode = gSBands.load();
  if (banded && mode && (c + 1) * 32 <= g.numDetectors) {
    const int64_t b = static_cast<int64_t>(g.colLow[c]) - (mode == 2 ? 0 : 2 * static_cast<int64_t>(g.seedGap) + 4);
    base = b > 0 ? static_cast<uint32_t>(b) & ~1u : 0u;
  }
  return {g.layerStride, top1, base};
}

// Words of S a column's region takes: its rows and the two rows below-base
// stores go to.
static uint64_t regionWords(const Pipeline::Impl::ColLayout &c) {
  return c.top1 ? static_cast<uint64_t>(c.top1 - c.base + 2) * c.P : 0u;
}

// Rows of a circuit's regions when they are all of one size (frozen
// pipelines): its highest column's stored layers and the two more.
static uint32_t uniformColRows(const StepgImage &g) {
  uint32_t top1 = 0;
  for (uint32_t c = 0; c < g.numCols; ++c) top1 = std::max(top1, columnLayout(g, c, false).top1);
  return top1 + 2;
}

// Bytes DevBuf::ensure allocates for `bytes` (with `head` more when it grows).
static size_t grownBytes(size_t bytes, double head) {
  return roundGrain(std::max<size_t>(static_cast<size_t>(static_cast<double>(bytes) * head), 256));
}

// Lays S out from m.cols: the regions (colRegion, colAdj) in the staged
// inputs, and on the device when `device`; S's capacity (grown with `head`;
// a frozen pipeline throws instead).
static void layoutS(Pipeline::Impl &m, bool frozen, double head, bool device) {
  uint4 *reg = reinterpret_cast<uint4 *>(m.inHost.as<char>() + m.oColRegion);
  int32_t *adj = reinterpret_cast<int32_t *>(m.inHost.as<char>() + m.oColAdj);
  VDesc *hd = reinterpret_cast<VDesc *>(m.inHost.as<char>() + m.oDesc);
  uint64_t off = 0;
  for (size_t v = 0; v < m.descs.size(); ++v) {
    VDesc &d = m.descs[v];
    d.sOff = hd[v].sOff = off;
    // frozen pipelines: one region size per circuit (classify then needs no
    // per-column offsets)
    uint32_t rows = 0;
    if (m.uniformCols)
      for (uint32_t c = 0; c < d.numCols; ++c) rows = std::max(rows, m.cols[d.colOff + c].top1 + 2);
    d.colRows = hd[v].colRows = rows;
    for (uint32_t c = 0; c < d.numCols; ++c) {
      const uint64_t k = d.colOff + c;
      const Pipeline::Impl::ColLayout &cl = m.cols[k];
      reg[k] = make_uint4(static_cast<uint32_t>(off), static_cast<uint32_t>(off >> 32), cl.base,
                          rows ? rows - 2 : cl.top1 ? cl.top1 - cl.base : 0u);
      adj[k] = static_cast<int32_t>((off - d.sOff) / cl.P) - static_cast<int32_t>(cl.base);
      off += rows ? static_cast<uint64_t>(rows) * cl.P : regionWords(cl);
    }
    if (off - d.sOff >= (uint64_t{1} << 31))
      throw RuntimeError("circuit too large: its sensitivity matrix exceeds 2^31 words");
  }
  m.Swords = off;
  const size_t bytes = 4 * std::max<uint64_t>(off, 1);
  if (bytes > m.S.cap) {
    if (frozen) m.S.ensure(bytes, true, "sensitivity matrix");  // throws
    const double grow = m.S.cap ? head : 1.0;  // (no headroom on the first allocation)
    m.S.release();
    m.S.ensure(static_cast<size_t>(static_cast<double>(bytes) * grow), false, "sensitivity matrix");
  }
  if (device && !m.cols.empty()) {
    GPA_HIP(hipMemcpy(m.in.as<char>() + m.oColRegion, reg, 16 * m.cols.size(), hipMemcpyHostToDevice));
    GPA_HIP(hipMemcpy(m.in.as<char>() + m.oColAdj, adj, 4 * m.cols.size(), hipMemcpyHostToDevice));
    GPA_HIP(hipMemcpy(m.in.as<char>() + m.oDesc, hd, sizeof(VDesc) * m.descs.size(), hipMemcpyHostToDevice));
  }
}

void Pipeline::load(const std::vector<const StepgImage *> &imgs) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  m.joinSide();
  GPA_HIP(hipStreamSynchronize(m.stream));
  GPA_HIP(hipStreamSynchronize(m.side));
  m.sizesKnown = false;
  m.launchMode = -1;
  m.splitPlanned = false;
  m.planPending = false;
  m.classBatch = 4;
  m.nSplit = m.nSplit1 = m.nSplit3 = 0;
  const bool frozen = m.frozen;
  const uint32_t V = static_cast<uint32_t>(imgs.size());
  if (V == 0) throw InvalidArgument("no circuits to compile");
  m.dropGraph();

  // layout
  m.descs.assign(V, VDesc{});
  variants_.assign(V, VariantInfo{});
  uint64_t actOff = 0, seedOff = 0, colOff = 0, rankOff = 0, spillOff = 0, probOff = 0;
  uint32_t bucketBase = 0, maxRpt = 1, maxThreads = 64;
  size_t ldsRing = 0, ldsAll = 0;  // largest ring (+ zero word), largest ring + seeds
  std::vector<uint4> items;
  std::vector<uint4> tiles;
  bool classTable = true;
  std::vector<std::pair<int32_t, uint32_t>> order;  // (band estimate, item index)
  m.cols.clear();
  m.uniformCols = !growable_;
  uint64_t loadRanks = 0;
  for (const StepgImage *g : imgs) loadRanks += g->numRanks;
  const uint32_t fill = defaultFill(loadRanks);
  for (uint32_t v = 0; v < V; ++v) {
    const StepgImage &g = *imgs[v];
    VDesc &d = m.descs[v];
    d.numLayers = g.numLayers;
    d.npl = g.npl;
    d.numCols = g.numCols;
    d.numDetectors = g.numDetectors;
    d.numObservables = g.numObservables;
    d.numRanks = g.numRanks;
    d.numSeeds = g.numSeeds;
    d.bucketBits = bucketBitsFor(g, fill);
    d.numBuckets = 1u << d.bucketBits;
    d.bucketBase = bucketBase;
    d.maskWords = (g.numCols + 31) / 32;
    d.layerStride = g.layerStride;
    d.sweepThreads = g.sweepThreads;
    d.sweepRpt = g.sweepRpt;
    d.rowSpan = g.rowSpan;
    d.actOff = actOff;
    d.seedOff = seedOff;
    d.colOff = colOff;
    d.rankOff = rankOff;
    d.numPairs = g.numPairs;
    for (uint32_t c = 0; c < g.numCols; ++c) m.cols.push_back(columnLayout(g, c, growable_));
    d.spillOff = spillOff;
    d.probOff = probOff;
    d.probDict = g.probTab.empty() ? 0u : 1u;
    variants_[v] = VariantInfo{g.numRanks, g.numCols, g.numDetectors, g.numObservables, bucketBase,
                               d.numBuckets};
    for (uint32_t c = 0; c < g.numCols; ++c) {
      const int32_t hiL = g.colHi[c];
      const int32_t est = hiL < 0 ? -1 : hiL - g.colLow[c] + 1;
      order.push_back({est, static_cast<uint32_t>(items.size())});
      items.push_back(make_uint4(v, c, static_cast<uint32_t>(hiL), static_cast<uint32_t>(g.colLow[c])));
    }
    for (uint32_t t = 0; t < g.numRanks; t += kClassBlock) {
      // the rows of a rank at layer l lie in layers l .. l + rowSpan
      const uint32_t last = std::min(t + kClassBlock, g.numRanks) - 1;
      const uint32_t nl = std::min(g.rankLayer[last] + g.rowSpan, g.numLayers - 1) - g.rankLayer[t] + 1;
      tiles.push_back(make_uint4(v, t, g.rankLayer[t], nl));
      if (maskTableWords(nl, g.rowSpan, (g.numCols + 31) / 32) > kMaskTableWords) classTable = false;
    }
    actOff += static_cast<uint64_t>(kSweepLeadPad + g.numPairs) * (g.layerStride + 1);
    seedOff += static_cast<uint64_t>(g.numCols) * g.numSeeds;
    colOff += g.numCols;
    rankOff += g.numRanks;
    spillOff += g.rankSpill.size() / 4;
    probOff += d.probDict ? g.probTab.size() : g.numRanks;
    bucketBase += d.numBuckets;
    maxRpt = std::max(maxRpt, g.sweepRpt);
    maxThreads = std::max(maxThreads, g.sweepThreads);
    ldsRing = std::max<size_t>(ldsRing, 4ull * (2ull * g.layerStride + 1 + 2));
    ldsAll = std::max<size_t>(ldsAll, 4ull * (2ull * g.layerStride + 1 + g.numSeeds + 2));
    if (16ull * (g.layerStride + 1) * (kSweepLeadPad + g.numPairs) >= (1ull << 32))
      throw RuntimeError("circuit too large for the sweep: sweep records exceed 4 GB");
  }
  if (rankOff >= 0x7FFFFFFFull) throw RuntimeError("too many error nodes in one load");
  // sweep launch shape (the widest circuit's) and LDS: the ring and zero
  // word, plus the column's seeds if they fit
  m.rpt = maxRpt;
  m.sweepThreads = maxThreads;
  if (ldsRing > kLdsBudget)
    throw RuntimeError("circuit too wide for the LDS sweep (" + std::to_string(ldsRing) +
                       " bytes of LDS per layer ring)");
  m.seedsInLds = ldsAll <= kLdsBudget;
  m.sweepLdsBytes = m.seedsInLds ? ldsAll : ldsRing;

  // longest bands first (they bound the sweep's critical path)
  std::stable_sort(order.begin(), order.end(),
                   [](const auto &a, const auto &b) { return a.first > b.first; });
  std::vector<uint4> &launchItems = m.hostItems;
  launchItems.resize(items.size());
  for (size_t k = 0; k < order.size(); ++k) launchItems[k] = items[order[k].second];

  m.V = V;
  m.nItems = static_cast<uint32_t>(launchItems.size());
  m.nTiles = static_cast<uint32_t>(tiles.size());
  m.classTable = classTable;
  m.targets16 = m.narrowTargets;
  for (const StepgImage *g : imgs)
    if (g->numDetectors + static_cast<uint64_t>(g->numObservables) > 65536) m.targets16 = false;
  m.B = bucketBase;
  m.R = static_cast<uint32_t>(rankOff);
  m.Ccols = static_cast<uint32_t>(colOff);

  // input arena
  size_t off = 0;
  auto place = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  m.oDesc = place(sizeof(VDesc) * V);
  m.oItems = place(sizeof(uint4) * m.nItems);
  m.oTiles = place(sizeof(uint4) * m.nTiles);
  m.oColRegion = place(16 * m.Ccols);
  m.oColAdj = place(4 * m.Ccols);
  m.oAct = place(16 * actOff);
  m.oSeed = place(4 * seedOff);
  m.oRankCode = place(8 * rankOff);
  m.oSpill = place(16 * spillOff);
  m.oProbIdx = place(2 * rankOff);
  m.oProbVal = place(8 * probOff);
  m.inBytes = off;
  m.in.ensure(m.inBytes, frozen, "inputs");
  m.inHost.ensure(m.inBytes);
  char *hb = m.inHost.as<char>();
  std::memcpy(hb + m.oDesc, m.descs.data(), sizeof(VDesc) * V);
  std::memcpy(hb + m.oItems, launchItems.data(), sizeof(uint4) * m.nItems);
  std::memcpy(hb + m.oTiles, tiles.data(), sizeof(uint4) * m.nTiles);
  layoutS(m, frozen, growable_ ? kGrowHeadroom : 1.0, false);
  for (uint32_t v = 0; v < V; ++v) {
    const StepgImage &g = *imgs[v];
    const VDesc &d = m.descs[v];
    std::memcpy(hb + m.oAct + 16 * d.actOff, g.act.data(), 4 * g.act.size());
    std::memcpy(hb + m.oSeed + 4 * d.seedOff, g.seedVals.data(), 4 * g.seedVals.size());
    std::memcpy(hb + m.oRankCode + 8 * d.rankOff, g.rankCode.data(), 4 * g.rankCode.size());
    std::memcpy(hb + m.oSpill + 16 * d.spillOff, g.rankSpill.data(), 4 * g.rankSpill.size());
    if (d.probDict) {
      std::memcpy(hb + m.oProbIdx + 2 * d.rankOff, g.probIdx.data(), 2 * g.probIdx.size());
      std::memcpy(hb + m.oProbVal + 8 * d.probOff, g.probTab.data(), 8 * g.probTab.size());
    } else {
      std::memset(hb + m.oProbIdx + 2 * d.rankOff, 0, 2ull * g.numRanks);
      std::memcpy(hb + m.oProbVal + 8 * d.probOff, g.rankProb.data(), 8 * g.rankProb.size());
    }
  }

  // work buffers: one arena (see WorkArena), the targets' own allocation
  const uint64_t R = std::max<uint64_t>(m.R, 1), B = m.B;
  WorkArena &wa = m.work;
  m.countStride = bucketCountStride(B);
  const auto bufs = m.workBufs();
  static_assert(std::tuple_size<decltype(bufs)>::value == kWorkBuffers, "one DevBuf per work buffer");
  // (a frozen pipeline keeps the slots per bucket it was sized with)
  if (!frozen) m.bucketCap = bucketCapFor(imgs);
  m.compact = compactLoad(growable_, R);
  if (m.compact) {
    const OutAlias a = outAlias(R);
    m.aliasHash = a.hash;
    m.aliasOff = a.off;
    m.aliasMember = a.member;
  }
  workBufferSizes(R, B, V, m.Ccols, m.countStride, m.bucketCap, m.compact,
                  [&](uint32_t k, uint64_t bytes, const char *what) { wa.request(*bufs[k], bytes, what); });
  wa.commit(frozen, growable_ ? kGrowHeadroom : 1.0);
  // target rows: a driver's 16 words per node (its capacity is fixed); a
  // growable load starts at the widest rows its stash already holds (at
  // least 4 words) and widens them, growing the stash, when a run shows a
  // larger class (rows that do not fit are read from S meanwhile)
  if (!growable_) {
    m.stashShift = 4;
  } else {
    m.stashShift = 2;
    while (m.stashShift < 4 && (4ull << (m.stashShift + 1)) * R <= m.stash.cap) ++m.stashShift;
  }
  m.stash.ensure((4ull << m.stashShift) * R, frozen, "nodes");
  // the overflow list and the oversized buckets' workspace: a driver's for
  // the worst case (every node past its bucket's slots; twice the nodes), a
  // growable load's from a start size (what it holds, at least R / 16 and
  // overflow_start entries), grown when a run asks for more
  if (!growable_) {
    m.ensureOverflow(R, frozen);
    m.ensureWorkspace(2 * R, frozen);
  } else {
    const uint64_t start = gOverflowStart.load();
    m.ensureOverflow(std::min<uint64_t>(R, std::max<uint64_t>(R / 16, start)), false);
    m.ensureWorkspace(std::min<uint64_t>(2 * R, start), false);
  }
  // targets: 16 per error node for a driver (its capacity is fixed), a
  // growable load starts with target_fill per node and grows (running again)
  // if its DEM holds more
  const uint64_t tgtFill = frozen ? 16u : gTargetFill.load();
  if (m.capTargets == 0 || (!frozen && m.capTargets < tgtFill * R)) {
    m.outTargets.ensure(4 * tgtFill * R, frozen, "targets");
    m.capTargets = m.outTargets.cap / 4;
  }
  m.hVB.ensure(4 * (2 * V + 8));
  m.hFlags.ensure(4 * kFlagWords);
  setSweepLds(m.sweepLdsBytes);

  // bucket counters start at zero; every run's bucket stage leaves them
  // zeroed; so does the band-overflow flag
  GPA_HIP(hipMemsetAsync(m.bucketCnt.p, 0, m.bucketCnt.cap, m.stream));
  GPA_HIP(hipMemsetAsync(m.counters.as<uint32_t>() + kFlagBandOverflow, 0, 4, m.stream));
  upload();
}

namespace {

// The totals a load forms from its circuits (see load()).
struct LoadTotals {
  uint64_t R = 0, B = 0, V = 0, Ccols = 0, Swords = 0, inBytes = 0;
  uint32_t cap = kBucketCap;  // slots per bucket
  uint32_t maxP = 64;  // widest layer stride
};

LoadTotals loadTotals(const std::vector<const StepgImage *> &imgs, bool banded) {
  LoadTotals t;
  uint64_t actOff = 0, seedOff = 0, spill = 0, probOff = 0, nTiles = 0, loadRanks = 0;
  for (const StepgImage *g : imgs) loadRanks += g->numRanks;
  const uint32_t fill = defaultFill(loadRanks);
  for (const StepgImage *g : imgs) {
    t.B += uint64_t{1} << bucketBitsFor(*g, fill);
    nTiles += (g->numRanks + kClassBlock - 1) / kClassBlock;
    actOff += static_cast<uint64_t>(kSweepLeadPad + g->numPairs) * (g->layerStride + 1);
    seedOff += static_cast<uint64_t>(g->numCols) * g->numSeeds;
    t.Ccols += g->numCols;
    t.R += g->numRanks;
    spill += g->rankSpill.size() / 4;
    probOff += g->probTab.empty() ? g->numRanks : g->probTab.size();
    if (banded)
      for (uint32_t c = 0; c < g->numCols; ++c) t.Swords += regionWords(columnLayout(*g, c, true));
    else
      t.Swords += static_cast<uint64_t>(g->numCols) * uniformColRows(*g) * g->layerStride;
    t.maxP = std::max(t.maxP, g->layerStride);
  }
  t.V = imgs.size();
  t.cap = bucketCapFor(imgs);
  // the input arena: each array at a 256-byte boundary
  t.inBytes = align256(sizeof(VDesc) * t.V) + align256(16 * t.Ccols) + align256(16 * nTiles) +
              align256(16 * t.Ccols) + align256(4 * t.Ccols) + align256(16 * actOff) +
              align256(4 * seedOff) + align256(8 * t.R) + align256(16 * spill) + align256(2 * t.R) +
              align256(8 * probOff);
  return t;
}

// Bytes of a buffer after DevBuf::ensure(bytes) on one of capacity `cap`.
uint64_t afterEnsure(size_t cap, uint64_t bytes) { return bytes <= cap ? cap : roundGrain(std::max<uint64_t>(bytes, 256)); }

// Capacities a pipeline holds (all zero: a new pipeline).
struct HeldCaps {
  std::array<size_t, kWorkBuffers> work{};
  bool arena = false;  // the arena is allocated
  size_t arenaCap = 0, in = 0, S = 0, targets = 0, members = 0, splitBuf = 0, splitState = 0, stash = 0;
  size_t ovSlot = 0, ovBucket = 0, scratchKey = 0, scratchRank = 0;
  uint64_t capTargets = 0;
};

// Device bytes a growable pipeline holding `h` holds after loading a load
// with totals `t` and compiling it: the work arena as WorkArena::commit
// lays it out, the input arena and targets as load() sizes them, the
// target rows at their widest, members as held, and the time-chunked
// sweep's buffers at their largest (at most kSplitMaxColumns columns of
// kSplitMaxChunks chunks of the widest circuit).  Every buffer is freed
// before its replacement is allocated, so this also bounds the peak during
// the load and its compiles -- unless the DEM has more than 16 targets per
// error node, when the target buffer grows.
uint64_t footprint(const LoadTotals &t, const HeldCaps &h, double headroom, bool growable) {
  const uint64_t R = std::max<uint64_t>(t.R, 1);
  std::vector<size_t> caps;
  bool fits = h.arena;
  workBufferSizes(R, t.B, t.V, t.Ccols, bucketCountStride(t.B), t.cap, compactLoad(growable, R),
                  [&](uint32_t k, uint64_t bytes, const char *) {
    const size_t b = std::max<size_t>(bytes, 256), cap = h.work[k];
    const size_t grown = h.arena && b > cap ? static_cast<size_t>(b * headroom) : 0;
    if (b > cap) fits = false;
    caps.push_back(std::max({b, cap, grown}));
  });
  uint64_t total = fits ? h.arenaCap : WorkArena::layoutBytes(caps);
  total += afterEnsure(h.in, t.inBytes);
  const uint64_t sBytes = 4 * std::max<uint64_t>(t.Swords, 1);
  total += sBytes <= h.S ? h.S : grownBytes(sBytes, h.S ? headroom : 1.0);
  total += h.capTargets == 0 || h.capTargets < 16 * R ? afterEnsure(h.targets, 4 * 16 * R) : h.targets;
  total += afterEnsure(h.stash, 4ull * kStashWords * R);
  // the overflow list and workspace at their largest (they grow on demand)
  total += afterEnsure(h.ovSlot, 16 * R) + afterEnsure(h.ovBucket, 4 * R);
  total += afterEnsure(h.scratchKey, 16 * R) + afterEnsure(h.scratchRank, 8 * R);
  total += h.members;
  const uint64_t G = (t.maxP + 31) / 32, P4 = quad(static_cast<uint32_t>(t.maxP));
  // descriptors, items and live lists; state areas (every boundary with
  // all positions live)
  const uint64_t splitItems = 4096 + sizeof(SplitDesc) * kSplitMaxColumns +
                              16ull * kSplitMaxColumns * (kSplitMaxChunks * (G + 1) + kSplitMaxChunks) +
                              4ull * kSplitMaxColumns * kSplitMaxChunks * P4 + 8ull * kSplitMaxColumns;
  const uint64_t splitState =
      4ull * kSplitMaxColumns * (P4 + (kSplitMaxChunks - 2) * splitMapWords(P4, P4)) + 16;
  total += afterEnsure(h.splitBuf, splitItems);
  total += afterEnsure(h.splitState, splitState);
  return total;
}

}  // namespace

uint64_t Pipeline::deviceBytesFor(const std::vector<const StepgImage *> &imgs, bool growable) {
  return footprint(loadTotals(imgs, growable), HeldCaps{}, growable ? kGrowHeadroom : 1.0, growable);
}

uint64_t Pipeline::deviceBytesAfter(const std::vector<const StepgImage *> &imgs) const {
  Impl &m = *impl_;
  if (!growable_) return deviceBytes();  // a frozen pipeline never grows (an oversized load throws)
  HeldCaps h;
  const auto bufs = m.workBufs();
  for (uint32_t k = 0; k < kWorkBuffers; ++k) h.work[k] = bufs[k]->cap;
  h.arena = m.work.mem.p != nullptr;
  h.arenaCap = m.work.mem.cap;
  h.S = m.S.cap;
  h.in = m.in.cap;
  h.targets = m.outTargets.cap;
  h.capTargets = m.capTargets;
  h.members = m.members.cap;
  h.stash = m.stash.cap;
  h.ovSlot = m.ovSlot.cap;
  h.ovBucket = m.ovBucket.cap;
  h.scratchKey = m.scratchKey.cap;
  h.scratchRank = m.scratchRank.cap;
  h.splitBuf = m.splitBuf.cap;
  h.splitState = m.splitState.cap;
  return footprint(loadTotals(imgs, true), h, kGrowHeadroom, true);
}

void Pipeline::upload() {
  Impl &m = *impl_;
  if (!m.inBytes) throw RuntimeError("nothing loaded");
  GPA_HIP(hipSetDevice(device_));
  // the arena's head (descriptors, sweep records, seeds, bands: what the
  // sweep reads) on the pipeline's stream; then its tail (the error nodes'
  // frame rows and probabilities, read from classify on) on the side
  // stream, overlapping the sweep.  Stream-ordered before the next run; the
  // duration is read once the streams have drained, in download().
  m.joinSide();
  const size_t headBytes = m.oRankCode;
  GPA_HIP(hipEventRecord(m.evUp0, m.stream));
  GPA_HIP(hipMemcpyAsync(m.in.p, m.inHost.p, headBytes, hipMemcpyHostToDevice, m.stream));
  if (m.inBytes > headBytes) {
    GPA_HIP(hipEventRecord(m.evPartA, m.stream));
    GPA_HIP(hipStreamWaitEvent(m.side, m.evPartA, 0));
    GPA_HIP(hipMemcpyAsync(m.in.as<char>() + headBytes, m.inHost.as<char>() + headBytes, m.inBytes - headBytes,
                           hipMemcpyHostToDevice, m.side));
    GPA_HIP(hipEventRecord(m.evUp1, m.side));
    GPA_HIP(hipEventRecord(m.evSide, m.side));
    m.sidePending = true;
  } else {
    GPA_HIP(hipEventRecord(m.evUp1, m.stream));
  }
  m.uploadPending = true;
}

static void planSweep(Pipeline::Impl &m);

void Pipeline::run(bool timed) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  if (m.planPending && !m.capturing) {
    m.planPending = false;
    planSweep(m);
  }
  char *ib = m.in.as<char>();
  const VDesc *desc = reinterpret_cast<const VDesc *>(ib + m.oDesc);
  const uint4 *items = reinterpret_cast<const uint4 *>(ib + m.oItems);
  const uint4 *tiles = reinterpret_cast<const uint4 *>(ib + m.oTiles);
  const uint4 *act = reinterpret_cast<const uint4 *>(ib + m.oAct);
  const uint32_t *seed = reinterpret_cast<const uint32_t *>(ib + m.oSeed);
  const uint4 *colRegion = reinterpret_cast<const uint4 *>(ib + m.oColRegion);
  const int32_t *colAdj = reinterpret_cast<const int32_t *>(ib + m.oColAdj);
  const RankIn rin{reinterpret_cast<const uint2 *>(ib + m.oRankCode), reinterpret_cast<const uint4 *>(ib + m.oSpill),
                   reinterpret_cast<const double *>(ib + m.oProbVal),
                   reinterpret_cast<const uint16_t *>(ib + m.oProbIdx)};
  hipStream_t st = m.stream;
  hipEvent_t *evs = m.evOverride ? m.evOverride : m.ev;
  auto mark = [&](int i) {
    if (timed) GPA_HIP(hipEventRecord(evs[i], st));
  };
  const uint32_t B = m.B;
  m.timedRun = timed;
  const uint32_t stashMaxCols = gStash.load() ? kStashMaxCols : 0u;

  // Five launches: sweep (block 0 also zeroes the flags), classify (hash,
  // stash, bucket slots), buckets (sort + merge; counters reset for the next
  // run), the three scans, emit.  No memset inside, so the sequence can be
  // captured in a graph.
  mark(0);
  {
    // one launch: every column (split columns: their top chunk), then the
    // middle chunks' maps and parts, then the lower chunks of split columns
    const SplitIo sp = m.splitIo();
    const size_t lds = m.nSplit ? m.splitLdsBytes : m.sweepLdsBytes;
#define GPA_SWEEP(RPT, G)                                                                                  \
  sweepKernel<RPT, G><<<std::max<uint32_t>(m.nItems + sp.numItems, 1), m.sweepThreads, lds, st>>>(          \
      desc, items, act, seed, colRegion, m.band.as<int2>(), m.S.as<uint32_t>(), m.nItems, m.flags(), sp, \
      m.V == 1 ? 1u : 0u)
    if (m.seedsInLds) {
      switch (m.rpt) {
      case 1: GPA_SWEEP(1, false); break;
      case 2: GPA_SWEEP(2, false); break;
      case 4: GPA_SWEEP(4, false); break;
      default: GPA_SWEEP(8, false); break;
      }
    } else {
      switch (m.rpt) {
      case 1: GPA_SWEEP(1, true); break;
      case 2: GPA_SWEEP(2, true); break;
      case 4: GPA_SWEEP(4, true); break;
      default: GPA_SWEEP(8, true); break;
      }
    }
#undef GPA_SWEEP
  }
  mark(1);
  m.joinSide();  // classify reads the error nodes' inputs
  if (m.nTiles) {
    ClassifyOut co{m.info.as<uint4>(),     m.bucketCount(),          m.countStride,
                   m.pslot.as<uint4>(),    m.ovSlot.as<uint4>(),     m.ovBucket.as<uint32_t>(),
                   m.flags(),              m.stash.as<uint32_t>(),   stashMaxCols,
                   m.stashShift,           gPackedSort.load(),       m.bucketCap,
                   static_cast<uint32_t>(m.ovCap)};
    if (m.classTable && gMaskTable.load() && m.classBatch == 16)
      classifyKernel<true, 16><<<m.nTiles, kClassBlock, 0, st>>>(desc, tiles, m.band.as<int2>(), colAdj,
                                                                 m.S.as<uint32_t>(), rin, co, m.V == 1 ? 1u : 0u);
    else if (m.classTable && gMaskTable.load())
      classifyKernel<true, 4><<<m.nTiles, kClassBlock, 0, st>>>(desc, tiles, m.band.as<int2>(), colAdj,
                                                                m.S.as<uint32_t>(), rin, co, m.V == 1 ? 1u : 0u);
    else
      classifyKernel<false, 4><<<m.nTiles, kClassBlock, 0, st>>>(desc, tiles, m.band.as<int2>(), colAdj,
                                                                 m.S.as<uint32_t>(), rin, co, m.V == 1 ? 1u : 0u);
  }
  mark(2);
  RunOut o{m.prank.as<uint32_t>(),
           m.runRank.as<uint32_t>(),
           m.runP.as<double>(),
           m.runHash.as<uint64_t>(),
           m.runTgtPre.as<uint32_t>(),
           m.runMember.as<uint32_t>(),
           m.flags()};
  BucketIo bio{m.bucketCount(),           m.countStride,
               m.pslot.as<uint4>(),         m.info.as<uint4>(),
               m.ovSlot.as<uint4>(),        m.ovBucket.as<uint32_t>(),
               m.runCount.as<uint32_t>(),
               m.runNnz.as<uint32_t>(),     m.itemCount.as<uint32_t>(),
               m.slotBase.as<uint32_t>(),   m.wgSum.as<uint32_t>(),
               m.scratchKey.as<uint64_t>(), m.scratchRank.as<uint32_t>(),
               B * m.bucketCap, m.bucketCap, m.compact ? 1u : 0u,
               static_cast<uint32_t>(m.ovCap), m.scratchCap};
  if (B && m.compact) {
    // each bucket's packed range: the exclusive prefix sum of the counts
    ScanArrays sa{{m.bucketCount(), nullptr, nullptr}, {m.slotBase.as<uint32_t>(), nullptr, nullptr}};
    sa.stride = m.countStride;
    sa.count = 1;
    m.scan(sa, B);
  }
  if (B) bucketKernel<<<(B + 3) / 4, 256, 0, st>>>(B, bio, o);
  mark(3);
  // bucket offsets: inside emit from the workgroup totals, or by the scans
  // when there are too many workgroups for that
  const bool prefixInEmit = (B + 3) / 4 <= kEmitPrefixGroups;
  if (!prefixInEmit) {
    ScanArrays sa{{m.runCount.as<uint32_t>(), m.runNnz.as<uint32_t>(), m.itemCount.as<uint32_t>()},
                  {m.classStart.as<uint32_t>(), m.tgtStart.as<uint32_t>(), m.memberStart.as<uint32_t>()}};
    m.scan(sa, B);
  }
  EmitIn ein{m.V, B, prefixInEmit ? m.wgSum.as<uint32_t>() : nullptr, m.slotBase.as<uint32_t>(),
             m.runCount.as<uint32_t>(), m.itemCount.as<uint32_t>(), m.classStart.as<uint32_t>(),
             m.tgtStart.as<uint32_t>(), m.memberStart.as<uint32_t>(), m.runRank.as<uint32_t>(),
             m.runTgtPre.as<uint32_t>(), m.runMember.as<uint32_t>(), m.runP.as<double>(),
             m.runHash.as<uint64_t>(), m.runNnz.as<uint32_t>(), m.S.as<uint32_t>(),
             m.stash.as<uint32_t>(), rin.code, rin.spill, stashMaxCols, m.stashShift, m.band.as<int2>(), colAdj,
             m.compact ? 0u : m.bucketCap};
  EmitOut eout{m.oP(), m.oOff(), m.oHash(), m.oMember(),
               m.outTargets.as<uint32_t>(), m.targets16 ? m.outTargets.as<uint16_t>() : nullptr, m.capTargets,
               m.flags(), m.vBounds.as<uint32_t>()};
  emitKernel<<<std::max<uint32_t>((B + 3) / 4, 1), 256, 0, st>>>(desc, ein, eout);
  mark(4);
  GPA_HIP(hipGetLastError());
}

// Critical-path model of the sweep, in steps (pairs) of one column,
// measured on MI355X (bb12, clock stamps per trip and per chain step): a
// step costs ~0.29 us at every height of a column, whole or in chunks; a
// chunk's first trip costs kSplitTrip steps more; a resumed chunk starts
// kSplitWake steps after the last producer of its column ends (the done
// count, its wait, the staging of its inputs), and each map of its chain
// costs kSplitMvFixed plus kSplitMvPer288 per 288 x 288 inputs x outputs.
constexpr double kSplitTrip = 1.0, kSplitWake = 18.0, kSplitMvFixed = 1.7, kSplitMvPer288 = 0.2;
constexpr uint32_t kClassWideMaxTiles = 2048;  // class workgroups up to which the 16-column walk is used
// boundaries a split plan chooses among (evenly thinned to this many)
constexpr size_t kSplitMaxCandidates = 16;

// The live positions of boundary q of circuit d (the ring words that pair
// q - 1 reads: the slot of layer 2q), ascending, from its staged sweep
// records.
static std::vector<uint32_t> livePositions(const Pipeline::Impl &m, const VDesc &d, int32_t q) {
  const uint32_t P = d.layerStride, slot = static_cast<uint32_t>(q & 1) * P;
  const uint4 *blk = reinterpret_cast<const uint4 *>(m.inHost.as<char>() + m.oAct) + d.actOff +
                     static_cast<uint64_t>(static_cast<int64_t>(kSweepLeadPad) + q - 1) * (P + 1);
  std::vector<uint8_t> seen(P, 0);
  for (uint32_t j = 0; j < P; ++j) {
    const uint4 r = blk[j];
    for (const uint32_t w : {r.x, r.y, r.z})
      for (const uint32_t x : {w & 0xFFFFu, w >> 16})
        if (x >= slot && x < slot + P) seen[x - slot] = 1;
  }
  std::vector<uint32_t> out;
  for (uint32_t j = 0; j < P; ++j)
    if (seen[j]) out.push_back(j);
  return out;
}

// Runs at the start of a load's second run (outside graph capture), from
// the column bands its first run found.  The
// plain items are put in the order of their measured bands, longest first
// (the launch order of the sweep's workgroups; the first run orders them by
// seed range, which misses the observables' long band).  Then the
// time-chunked sweep: the longest columns, while their band reaches the
// bottom (layer <= 3) and a plan is estimated to beat sweeping the column
// whole by 10 %.  A plan's boundaries are chosen among the pairs whose
// boundary has the fewest live positions (the kSplitMaxCandidates / 2
// fewest and every boundary with as few), every combination of K - 1 of
// them for K = 3 .. kSplitMaxChunks; its estimate is the longest of the
// producers (the top chunk, a middle chunk's maps), then the longest of the
// resumed chunks with their wake-up and chain, against the next longest
// band.  Writes the items and descriptors to the device; the results do not
// depend on the plan.
static void planSweep(Pipeline::Impl &m) {
  m.nSplit = m.nSplit1 = m.nSplit3 = 0;
  if (!m.Ccols) return;
  std::vector<int2> bands(m.Ccols);
  GPA_HIP(hipMemcpy(bands.data(), m.band.p, 8ull * m.Ccols, hipMemcpyDeviceToHost));
  auto bandLen = [&](const uint4 &it) {
    const int2 b = bands[m.descs[it.x].colOff + (it.y & ~kItemSplit)];
    return b.y < 0 ? -1 : b.y / 2 - b.x / 2 + 1;
  };
  std::stable_sort(m.hostItems.begin(), m.hostItems.end(),
                   [&](const uint4 &a, const uint4 &b) { return bandLen(a) > bandLen(b); });
  // columns a node's row window meets, on average over the error nodes: a
  // column's band (plus the rows' span) over the circuit's layers
  double seen = 0, nodes = 0;
  for (uint32_t v = 0; v < m.V; ++v) {
    const VDesc &d = m.descs[v];
    double layers = 0;
    for (uint32_t c = 0; c < d.numCols; ++c) {
      const int2 b = bands[d.colOff + c];
      if (b.y >= b.x) layers += std::min<double>(b.y - b.x + 1 + d.rowSpan, d.numLayers);
    }
    seen += d.numLayers ? layers / d.numLayers * d.numRanks : 0.0;
    nodes += d.numRanks;
  }
  // (only while the class stage is latency-bound: a few workgroups per
  // compute unit; large batches need the occupancy of the narrow walk,
  // measured: 256 variants 3.78 ms with 4, 4.04 ms with 16)
  const uint32_t batch = nodes > 0 && seen / nodes > 12.0 && m.nTiles <= kClassWideMaxTiles ? 16u : 4u;
  if (batch != m.classBatch) {
    m.classBatch = batch;
    m.dropGraph();
  }
  auto writeItems = [&]() {
    std::memcpy(m.inHost.as<char>() + m.oItems, m.hostItems.data(), 16 * m.hostItems.size());
    GPA_HIP(hipMemcpy(m.in.as<char>() + m.oItems, m.hostItems.data(), 16 * m.hostItems.size(),
                      hipMemcpyHostToDevice));
  };
  const uint32_t knob = gSweepSplit.load();
  const bool force = knob == 2;  // every column reaching the bottom, most chunks that fit
  // (loads with several long columns of one length, e.g. batches, are left
  // whole by the estimate: the next longest band bounds the sweep anyway)
  if (!knob) {
    writeItems();
    return;
  }
  struct Col {
    uint32_t v, c;
    int32_t lo, hi, len;
  };
  std::vector<Col> cols;
  for (uint32_t v = 0; v < m.V; ++v) {
    const VDesc &d = m.descs[v];
    for (uint32_t c = 0; c < d.numCols; ++c) {
      const int2 b = bands[d.colOff + c];
      if (b.y >= 0) cols.push_back({v, c, b.x, b.y, b.y / 2 - b.x / 2 + 1});
    }
  }
  std::stable_sort(cols.begin(), cols.end(), [](const Col &a, const Col &b) { return a.len > b.len; });
  struct Plan {
    SplitDesc sd{};
    std::vector<std::vector<uint32_t>> live;  // [k] for boundaries 1 .. K-1
    double est = 0;
    size_t lds = 0;
  };
  std::vector<Plan> plans;
  for (size_t i = 0; i < cols.size() && plans.size() < kSplitMaxColumns; ++i) {
    const Col &col = cols[i];
    if (col.lo > 3) {
      if (force) continue;
      break;  // the longest column left bounds the sweep anyway
    }
    const VDesc &d = m.descs[col.v];
    const uint32_t P = d.layerStride;
    const double rest = i + 1 < cols.size() ? cols[i + 1].len : 0;
    const int32_t top = col.hi / 2, total = top + 1;
    if (total < 3) continue;
    // candidate boundaries: the pairs with the fewest live positions
    std::vector<uint32_t> nl(total + 1, 0);
    for (int32_t q = 1; q <= top; ++q) nl[q] = static_cast<uint32_t>(livePositions(m, d, q).size());
    // (the boundaries with the fewest live positions, at least
    // kSplitMaxCandidates / 2 of them)
    std::vector<uint32_t> sorted(nl.begin() + 1, nl.begin() + top + 1);
    std::sort(sorted.begin(), sorted.end());
    const uint32_t limit = sorted[std::min<size_t>(sorted.size(), kSplitMaxCandidates / 2) - 1];
    std::vector<int32_t> Q;
    for (int32_t q = 1; q <= top; ++q)
      if (nl[q] <= limit) Q.push_back(q);
    if (Q.size() > kSplitMaxCandidates) {
      std::vector<int32_t> t;
      for (size_t k = 0; k < kSplitMaxCandidates; ++k) t.push_back(Q[k * (Q.size() - 1) / (kSplitMaxCandidates - 1)]);
      Q = t;
    }
    const double whole = std::max<double>(total, rest) + kSplitTrip;
    const size_t ring = 4ull * (2 * P + 1), seedB = 4ull * d.numSeeds;
    auto al16 = [](size_t x) { return (x + 15) & ~static_cast<size_t>(15); };
    Plan best;
    best.est = force ? -1.0 : whole * 0.9;
    int bestK = 0;
    for (int K = 3; K <= kSplitMaxChunks; ++K) {
      if (Q.size() < static_cast<size_t>(K - 1)) break;
      std::vector<size_t> ix(K - 1);
      for (int k = 0; k < K - 1; ++k) ix[k] = k;
      for (;;) {
        int32_t bd[kSplitMaxChunks + 1];
        bd[0] = 0;
        bd[K] = top + 1;
        for (int k = 1; k < K; ++k) bd[k] = Q[ix[k - 1]];
        double prod = bd[K] - bd[K - 1], res = 0;
        for (int k = 1; k <= K - 2; ++k) prod = std::max<double>(prod, bd[k + 1] - bd[k]);
        prod += kSplitTrip;
        double chain = 0;
        uint32_t mapWords = quad(nl[bd[K - 1]]);
        for (int k = K - 2; k >= 0; --k) {
          res = std::max(res, chain + (bd[k + 1] - bd[k]) + kSplitTrip);
          if (k >= 1) {
            chain += kSplitMvFixed + kSplitMvPer288 * nl[bd[k + 1]] * nl[bd[k]] / (288.0 * 288.0);
            mapWords += splitMapWords(nl[bd[k + 1]], nl[bd[k]]);
          }
        }
        // (one launch: a resumed chunk holds the ring, the seeds when the
        // launch keeps them in LDS, and the chain)
        const size_t lds = al16(ring + (m.seedsInLds ? seedB : 0) + 8) + 4ull * splitChainWords(P, mapWords);
        const bool ok = lds <= kLdsBudget;
        double est = std::max(prod + kSplitWake + res, rest + kSplitTrip);
        if (force) {
          // most chunks, then the shortest longest chunk
          int32_t longest = 0;
          for (int k = 0; k < K; ++k) longest = std::max(longest, bd[k + 1] - bd[k]);
          est = K * 1e6 - longest;
        }
        const bool better = force ? est > best.est : est < best.est;
        if (ok && better) {
          best.est = est;
          best.sd = SplitDesc{};
          best.sd.v = col.v;
          best.sd.c = col.c;
          best.sd.K = static_cast<uint32_t>(K);
          for (int k = 0; k <= K; ++k) best.sd.bound[k] = bd[k];
          best.lds = lds;
          bestK = K;
        }
        // next combination
        int k = K - 2;
        while (k >= 0 && ix[k] == Q.size() - (K - 1) + k) --k;
        if (k < 0) break;
        ++ix[k];
        for (int j = k + 1; j < K - 1; ++j) ix[j] = ix[j - 1] + 1;
      }
    }
    if (!bestK) {
      if (force) continue;
      break;
    }
    best.live.resize(bestK);
    for (int k = 1; k < bestK; ++k) {
      best.live[k] = livePositions(m, d, best.sd.bound[k]);
      best.sd.nLive[k] = static_cast<uint32_t>(best.live[k].size());
    }
    plans.push_back(std::move(best));
  }
  if (plans.empty()) {
    writeItems();
    return;
  }
  // state areas, live lists and items
  std::vector<SplitDesc> descs;
  std::vector<uint4> items1, items3;
  std::vector<uint32_t> live;
  uint64_t stateWords = 0;
  size_t lds = m.sweepLdsBytes;
  for (Plan &pl : plans) {
    SplitDesc &sd = pl.sd;
    const uint32_t K = sd.K, si = static_cast<uint32_t>(descs.size());
    for (uint32_t k = 1; k < K; ++k) {
      sd.liveOff[k] = static_cast<uint32_t>(live.size());
      live.insert(live.end(), pl.live[k].begin(), pl.live[k].end());
      live.resize(quad(static_cast<uint32_t>(live.size())), 0u);
    }
    sd.stateOff = static_cast<uint32_t>(stateWords);
    stateWords += quad(sd.nLive[K - 1]);
    sd.producers = 1;
    for (uint32_t k = 1; k + 2 <= K; ++k) {
      sd.tOff[k] = static_cast<uint32_t>(stateWords);
      stateWords += splitMapWords(sd.nLive[k + 1], sd.nLive[k]);
      sd.producers += (sd.nLive[k + 1] + 31) / 32 + 1;
    }
    sd.tOff[K - 1] = sd.stateEnd = static_cast<uint32_t>(stateWords);
    if (stateWords >= 0xFFFFFFF0ull) break;
    for (uint32_t k = 1; k + 2 <= K; ++k) {
      for (uint32_t g = 0; g < (sd.nLive[k + 1] + 31) / 32; ++g)
        items1.push_back(make_uint4(si, kSplitBasis | k << 8 | g << 16, sd.v, sd.c));
      items1.push_back(make_uint4(si, kSplitPart | k << 8, sd.v, sd.c));
    }
    for (uint32_t k = 0; k + 2 <= K; ++k) items3.push_back(make_uint4(si, kSplitResume | k << 8, sd.v, sd.c));
    lds = std::max(lds, pl.lds);
    descs.push_back(sd);
  }
  auto al = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
  // descriptors | maps' and parts' items, then the resumed chunks' | live
  // positions | sync words (zero: the runs leave them zero)
  m.oSplitItems = al(sizeof(SplitDesc) * descs.size());
  m.oSplitLive = al(m.oSplitItems + 16 * (items1.size() + items3.size()));
  m.oSplitSync = al(m.oSplitLive + 4 * live.size());
  const size_t bytes = m.oSplitSync + 8 * kSplitMaxColumns;
  std::vector<char> h(bytes, 0);
  std::memcpy(h.data(), descs.data(), sizeof(SplitDesc) * descs.size());
  std::memcpy(h.data() + m.oSplitItems, items1.data(), 16 * items1.size());
  std::memcpy(h.data() + m.oSplitItems + 16 * items1.size(), items3.data(), 16 * items3.size());
  std::memcpy(h.data() + m.oSplitLive, live.data(), 4 * live.size());
  m.dropGraph();  // before the buffers it may name are replaced
  // a split column's band is reported from layer 0 (its top chunk does not
  // reach the bottom): it stores every layer
  bool relayout = false;
  for (const SplitDesc &sd : descs) {
    Pipeline::Impl::ColLayout &cl = m.cols[m.descs[sd.v].colOff + sd.c];
    if (cl.base) {
      cl.base = 0;
      relayout = true;
    }
  }
  if (relayout) layoutS(m, m.frozen, kGrowHeadroom, true);
  m.splitBuf.ensure(bytes, false, "split sweep");
  m.splitState.ensure(4 * stateWords + 16, false, "split sweep");
  GPA_HIP(hipMemcpy(m.splitBuf.p, h.data(), bytes, hipMemcpyHostToDevice));
  // the maps' dense flags and row counts start at zero (each run's last
  // resumed chunk clears them again)
  GPA_HIP(hipMemset(m.splitState.p, 0, 4 * stateWords + 16));
  // the split columns' own items become their top chunks (in the staged
  // inputs, which later compiles upload, and on the device now), launched
  // first
  for (uint32_t si = 0; si < descs.size(); ++si)
    for (uint4 &it : m.hostItems)
      if (it.x == descs[si].v && it.y == descs[si].c) {
        it.y |= kItemSplit;
        it.w = si;
      }
  std::stable_partition(m.hostItems.begin(), m.hostItems.end(),
                        [](const uint4 &it) { return (it.y & kItemSplit) != 0; });
  writeItems();
  m.splitLdsBytes = lds;
  setSweepLds(lds);
  m.nSplit = static_cast<uint32_t>(descs.size());
  m.nSplit1 = static_cast<uint32_t>(items1.size());
  m.nSplit3 = static_cast<uint32_t>(items3.size());
}

void Pipeline::download() {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  const uint32_t V = m.V;
  const auto t0 = std::chrono::steady_clock::now();
  // a block still held by a caller keeps its contents; take another
  if (!m.res || m.res.use_count() > 1) m.res = takeBlock(m.pool);
  ResultBlock &rb = *m.res;
  auto copyArrays = [&](uint64_t C, uint64_t T) {
    rb.p.ensure(8 * C + 8);
    rb.off.ensure(4 * (C + 1));
    rb.hash.ensure(8 * C + 8);
    rb.targets.ensure(4 * T + 4);
    const uint64_t tb = m.targets16 ? 2 : 4;
    if (C) {
      GPA_HIP(hipMemcpyAsync(rb.p.p, m.oP(), 8 * C, hipMemcpyDeviceToHost, m.stream));
      if (m.wantHashes) GPA_HIP(hipMemcpyAsync(rb.hash.p, m.oHash(), 8 * C, hipMemcpyDeviceToHost, m.stream));
    }
    GPA_HIP(hipMemcpyAsync(rb.off.p, m.oOff(), 4 * (C + 1), hipMemcpyDeviceToHost, m.stream));
    if (T) GPA_HIP(hipMemcpyAsync(rb.targets.p, m.outTargets.p, tb * T, hipMemcpyDeviceToHost, m.stream));
  };
  GPA_HIP(hipMemcpyAsync(m.hVB.p, m.vBounds.p, 4 * (2 * V + 2), hipMemcpyDeviceToHost, m.stream));
  GPA_HIP(hipMemcpyAsync(m.hFlags.p, m.counters.p, 4 * kFlagWords, hipMemcpyDeviceToHost, m.stream));
  // The DEM of a loaded circuit has the same size on every run: once one
  // run has told it, the arrays are copied with the sizes, behind one wait
  // (and again if the sizes differ).
  const bool guessed = m.sizesKnown;
  const uint64_t gC = m.lastC, gT = m.lastT;
  if (guessed) copyArrays(gC, gT);
  GPA_HIP(hipStreamSynchronize(m.stream));
  if (m.uploadPending) {
    float ms = 0;
    GPA_HIP(hipEventSynchronize(m.evUp1));
    GPA_HIP(hipEventElapsedTime(&ms, m.evUp0, m.evUp1));
    times_.upload = ms * 1e3;
    m.uploadPending = false;
  }
  if (m.timedRun) {
    float ms[4] = {};
    for (int i = 0; i < 4; ++i) GPA_HIP(hipEventElapsedTime(&ms[i], m.ev[i], m.ev[i + 1]));
    times_.sweep = ms[0] * 1e3;
    times_.classify = ms[1] * 1e3;
    times_.partition = 0;  // the bucket slots are filled by classify
    times_.buckets = ms[2] * 1e3;
    times_.emit = ms[3] * 1e3;
    m.timedRun = false;
  }
  const uint32_t *vb = m.hVB.as<uint32_t>();
  const uint32_t *fl = m.hFlags.as<uint32_t>();
  if (fl[kFlagSplitWait]) throw RuntimeError("time-chunked sweep: a chunk's wait for its inputs timed out");
  if (fl[kFlagBandOverflow]) {
    // a column's band reached below its stored rows (cut there, so this
    // run's classes are wrong): the columns cut at their base store every
    // layer from now on, and the load runs again
    std::vector<int2> bands(m.Ccols);
    GPA_HIP(hipMemcpy(bands.data(), m.band.p, 8ull * m.Ccols, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < m.Ccols; ++k)
      if (m.cols[k].base && bands[k].y >= 0 && bands[k].x <= static_cast<int32_t>(m.cols[k].base)) m.cols[k].base = 0;
    ++m.bandRelayouts;
    m.dropGraph();
    layoutS(m, m.frozen, kGrowHeadroom, true);
    GPA_HIP(hipMemsetAsync(m.counters.as<uint32_t>() + kFlagBandOverflow, 0, 4, m.stream));
    m.sizesKnown = false;
    run(false);
    download();
    return;
  }
  if (fl[3] > m.ovCap || fl[kFlagScratchNeed] > m.scratchCap) {
    // the overflow list or the oversized buckets' workspace was too small
    // (cut there, so this run's classes are wrong): grow them to what the
    // run asked for, and run again.  A driver's hold the worst case.
    if (m.frozen) throw RuntimeError("sort overflow exceeded the driver's capacity");
    if (fl[3] > m.ovCap) m.ensureOverflow(std::min<uint64_t>(m.R, fl[3] + fl[3] / 4), false);
    if (fl[kFlagScratchNeed] > m.scratchCap)
      m.ensureWorkspace(fl[kFlagScratchNeed] + fl[kFlagScratchNeed] / 4, false);
    m.dropGraph();
    m.sizesKnown = false;
    run(false);
    download();
    return;
  }
  if (fl[0]) throw RuntimeError("64-bit class hash collision between different error classes");
  m.totalClasses = vb[2 * V];
  m.totalTargets = vb[2 * V + 1];
  if (fl[1]) {
    // target buffer too small: grow it and re-run everything
    if (m.frozen && 4 * m.totalTargets > m.outTargets.cap)
      throw RuntimeError("circuit exceeds the driver's capacity (targets)");
    m.outTargets.ensure(4 * (m.totalTargets + m.totalTargets / 4 + 64), false, "targets");
    m.capTargets = m.outTargets.cap / 4;
    m.dropGraph();
    m.sizesKnown = false;
    run(false);
    download();
    return;
  }
  m.vClass.resize(V + 1);
  m.vTgt.resize(V + 1);
  for (uint32_t v = 0; v <= V; ++v) {
    m.vClass[v] = vb[2 * v];
    m.vTgt[v] = vb[2 * v + 1];
  }
  const uint64_t C = m.totalClasses, T = m.totalTargets;
  if (!guessed || C != gC || T != gT) {
    copyArrays(C, T);
    GPA_HIP(hipStreamSynchronize(m.stream));
  }
  if (!m.sizesKnown) {
    // the load's largest class (the same on every run) sets the row width
    const uint32_t *off = m.res->off.as<uint32_t>();
    uint32_t most = 0;
    for (uint64_t c = 0; c < C; ++c) most = std::max(most, off[c + 1] - off[c]);
    uint32_t shift = 2;
    while (shift < 4 && (2u << shift) < most) ++shift;
    if (shift != m.stashShift) {
      m.stashShift = shift;
      m.stash.ensure((4ull << shift) * std::max<uint64_t>(m.R, 1), m.frozen, "nodes");
      m.dropGraph();
    }
  }
  if (!m.splitPlanned) {
    m.splitPlanned = true;
    m.planPending = true;
  }
  m.lastC = C;
  m.lastT = T;
  m.sizesKnown = true;
  times_.download =
      std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

void Pipeline::result(uint32_t v, VariantResult &out) const {
  const Impl &m = *impl_;
  if (v >= m.V) throw InvalidArgument("circuit index out of range");
  const uint32_t c0 = m.vClass[v], c1 = m.vClass[v + 1];
  const uint32_t t0 = m.vTgt[v];
  out.numErrors = c1 - c0;
  out.numTargets = m.vTgt[v + 1] - t0;
  out.numDetectors = variants_[v].numDetectors;
  out.numObservables = variants_[v].numObservables;
  if (!m.res) throw RuntimeError("no results downloaded");
  const ResultBlock &rb = *m.res;
  out.probs = rb.p.as<double>() + c0;
  out.hashes = m.wantHashes ? rb.hash.as<uint64_t>() + c0 : nullptr;
  out.targets = m.targets16 ? nullptr : rb.targets.as<uint32_t>() + t0;
  out.targets16 = m.targets16 ? rb.targets.as<uint16_t>() + t0 : nullptr;
  const uint32_t *off = rb.off.as<uint32_t>();
  if (t0 == 0) {
    // the first circuit's offsets are its own
    out.offsetsRel.clear();
    out.offsets = off + c0;
  } else {
    out.offsetsRel.resize(out.numErrors + 1);
    for (uint32_t i = 0; i <= out.numErrors; ++i) out.offsetsRel[i] = off[c0 + i] - t0;
    out.offsets = out.offsetsRel.data();
  }
}

ResultHandle Pipeline::resultBlock() const { return impl_->res; }

void Pipeline::setDownloadHashes(bool on) { impl_->wantHashes = on; }
void Pipeline::setNarrowTargets(bool on) { impl_->narrowTargets = on; }

void Pipeline::copyHashes(uint32_t v, uint64_t *dst) const {
  const Impl &m = *impl_;
  if (v >= m.V || m.vClass.size() != m.V + 1) throw RuntimeError("no results downloaded");
  const uint32_t c0 = m.vClass[v], n = m.vClass[v + 1] - c0;
  GPA_HIP(hipSetDevice(device_));
  GPA_HIP(hipStreamSynchronize(m.stream));
  if (n) GPA_HIP(hipMemcpy(dst, m.oHash() + c0, 8ull * n, hipMemcpyDeviceToHost));
}

void Pipeline::runGraph() {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  m.joinSide();
  if (m.planPending) {
    m.planPending = false;
    planSweep(m);
  }
  if (!m.graphExec) {
    GPA_HIP(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
    m.capturing = true;
    try {
      run(false);
      m.capturing = false;
    } catch (...) {
      m.capturing = false;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(m.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    GPA_HIP(hipStreamEndCapture(m.stream, &m.graph));
    GPA_HIP(hipGraphInstantiate(&m.graphExec, m.graph, nullptr, nullptr, 0));
  }
  GPA_HIP(hipGraphLaunch(m.graphExec, m.stream));
}

double Pipeline::benchDevice(uint32_t iters) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  iters = std::max<uint32_t>(iters, 1);
  runGraph();  // warm (and capture)
  // back-to-back compiles go out either as graph replays or as direct
  // launches: a graph saves host launch work, but each replay's completion
  // leaves the device idle before the next (measured on bb12: 82.4 us per
  // compile replayed, 77.5 direct).  The first call after a load times a few
  // of each and keeps the faster.
  auto timeIters = [&](uint32_t n, bool graph) {
    GPA_HIP(hipEventRecord(m.ev[6], m.stream));
    for (uint32_t i = 0; i < n; ++i) {
      if (graph) runGraph();
      else run(false);
    }
    GPA_HIP(hipEventRecord(m.ev[7], m.stream));
    GPA_HIP(hipEventSynchronize(m.ev[7]));
    float t = 0;
    GPA_HIP(hipEventElapsedTime(&t, m.ev[6], m.ev[7]));
    return t;
  };
  if (m.launchMode < 0) {
    const uint32_t n = 8;
    timeIters(n, false);  // warm the direct path
    const float g = timeIters(n, true), d = timeIters(n, false);
    m.launchMode = d < g ? 0 : 1;
  }
  // throughput: back-to-back compiles between two events
  const float ms = timeIters(iters, m.launchMode == 1);
  // per-stage device time: the same iterations launched directly with five
  // events each (before the sweep and after the sweep, classify, bucket and
  // scan + emit stages), read after the last one completes
  constexpr int kEv = 5;
  const size_t need = static_cast<size_t>(iters) * kEv;
  while (m.evPool.size() < need) {
    hipEvent_t e;
    GPA_HIP(hipEventCreate(&e));
    m.evPool.push_back(e);
  }
  for (uint32_t i = 0; i < iters; ++i) {
    m.evOverride = m.evPool.data() + static_cast<size_t>(i) * kEv;
    run(true);
  }
  m.evOverride = nullptr;
  m.timedRun = false;
  GPA_HIP(hipStreamSynchronize(m.stream));
  double st[kEv - 1] = {};
  for (uint32_t i = 0; i < iters; ++i)
    for (int k = 0; k + 1 < kEv; ++k) {
      float t = 0;
      GPA_HIP(hipEventElapsedTime(&t, m.evPool[i * kEv + k], m.evPool[i * kEv + k + 1]));
      st[k] += t;
    }
  times_.sweep = st[0] * 1e3 / iters;
  times_.classify = st[1] * 1e3 / iters;
  times_.partition = 0;
  times_.buckets = st[2] * 1e3 / iters;
  times_.emit = st[3] * 1e3 / iters;
  return ms * 1e3 / iters;
}

void setBucketFill(uint32_t fill) { gBucketFill.store(fill); }  // 0: by load size
void setStash(bool on) { gStash.store(on ? 1u : 0u); }
bool stash() { return gStash.load() != 0; }
void setSweepSplit(uint32_t mode) { gSweepSplit.store(mode); }
uint32_t sweepSplit() { return gSweepSplit.load(); }

void setMaskTable(bool on) { gMaskTable.store(on ? 1u : 0u); }
void setPackedSort(bool on) { gPackedSort.store(on ? 1u : 0u); }
void setSBands(uint32_t mode) {
  if (mode > 2) throw InvalidArgument("s_bands must be 0, 1 or 2");
  gSBands.store(mode);
}
uint32_t sBands() { return gSBands.load(); }
uint32_t Pipeline::bandRelayouts() const { return impl_->bandRelayouts; }
uint32_t Pipeline::bucketCap() const { return impl_->bucketCap; }
uint32_t Pipeline::numBuckets() const { return impl_->B; }
void setBucketCap(uint32_t cap) {
  if (cap != 64 && cap != 128 && cap != 256) throw InvalidArgument("bucket_cap must be 64, 128 or 256");
  gBucketCap.store(cap);
}
uint32_t bucketCap() { return gBucketCap.load(); }
void setTargetFill(uint32_t n) {
  if (n < 1 || n > 16) throw InvalidArgument("target_fill must be between 1 and 16");
  gTargetFill.store(n);
}
uint32_t targetFill() { return gTargetFill.load(); }
void setCompactItems(uint32_t mode) {
  if (mode > 2) throw InvalidArgument("compact_items must be 0, 1 or 2");
  gCompactItems.store(mode);
}
uint32_t compactItems() { return gCompactItems.load(); }
void setOverflowStart(uint32_t n) {
  if (n < 1) throw InvalidArgument("overflow_start must be at least 1");
  gOverflowStart.store(n);
}
uint32_t overflowStart() { return gOverflowStart.load(); }
uint64_t Pipeline::sWords() const { return impl_->Swords; }
bool packedSort() { return gPackedSort.load() != 0; }
bool maskTable() { return gMaskTable.load() != 0; }
uint32_t bucketFill() { return gBucketFill.load(); }

uint64_t Pipeline::deviceBytes() const {
  const Impl &m = *impl_;
  // the work arena and the buffers with their own allocations
  uint64_t b = m.work.mem.cap;
  for (const DevBuf *x : {&m.in, &m.S, &m.stash, &m.ovSlot, &m.ovBucket, &m.scratchKey, &m.scratchRank,
                          &m.outTargets, &m.members, &m.splitBuf, &m.splitState})
    b += x->cap;
  return b;
}

void Pipeline::sweepPlan(uint32_t out[3]) const {
  out[0] = impl_->nSplit;
  out[1] = impl_->nSplit1;
  out[2] = impl_->nSplit3;
}

uint64_t Pipeline::sortedItems() const {
  const Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  GPA_HIP(hipStreamSynchronize(m.stream));
  std::vector<uint32_t> cnt(m.B);
  if (m.B) GPA_HIP(hipMemcpy(cnt.data(), m.itemCount.as<uint32_t>(), 4ull * m.B, hipMemcpyDeviceToHost));
  uint64_t t = 0;
  for (uint32_t c : cnt) t += c;
  return t;
}

void Pipeline::debugSensitivities(uint32_t v, std::vector<uint32_t> &cols32) const {
  const Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  GPA_HIP(hipStreamSynchronize(m.stream));
  const VDesc &d = m.descs.at(v);
  std::vector<uint32_t> S(m.Swords);
  std::vector<int2> band(d.numCols);
  if (!S.empty()) GPA_HIP(hipMemcpy(S.data(), m.S.p, 4 * S.size(), hipMemcpyDeviceToHost));
  const int32_t *adj = reinterpret_cast<const int32_t *>(m.inHost.as<char>() + m.oColAdj) + d.colOff;
  if (d.numCols)
    GPA_HIP(hipMemcpy(band.data(), m.band.as<int2>() + d.colOff, 8 * d.numCols, hipMemcpyDeviceToHost));
  // each rank's word = XOR of its in-band frame rows
  std::vector<uint32_t> code(4ull * d.numRanks);
  const uint2 *cc = reinterpret_cast<const uint2 *>(m.inHost.as<char>() + m.oRankCode) + d.rankOff;
  const uint4 *spill = reinterpret_cast<const uint4 *>(m.inHost.as<char>() + m.oSpill) + d.spillOff;
  for (uint32_t i = 0; i < d.numRanks; ++i) {
    const uint4 q = expandRankCode(cc[i], spill);
    code[4ull * i] = q.x;
    code[4ull * i + 1] = q.y;
    code[4ull * i + 2] = q.z;
    code[4ull * i + 3] = q.w;
  }
  cols32.assign(static_cast<size_t>(d.numRanks) * d.numCols, 0u);
  for (uint32_t c = 0; c < d.numCols; ++c)
    for (uint32_t i = 0; i < d.numRanks; ++i) {
      uint32_t w = 0;
      for (int q = 0; q < 4; ++q) {
        const uint32_t cd = code[4ull * i + q];
        if (cd == NONE) continue;
        const int32_t l = static_cast<int32_t>(cd >> kRowPosBits);
        if (l < band[c].x || l > band[c].y) continue;
        const uint64_t row = static_cast<uint64_t>(l) * d.layerStride + (cd & ((1u << kRowPosBits) - 1));
        w ^= S[d.sOff + static_cast<uint32_t>(static_cast<uint32_t>(adj[c]) * d.layerStride + row)];
      }
      cols32[static_cast<size_t>(i) * d.numCols + c] = w;
    }
}

const uint32_t *Pipeline::devSortedRanks() {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  m.members.ensure(4 * std::max<uint64_t>(m.R, 1), false, "nodes");
  if (m.B)
    compactMembersKernel<<<(m.B + 3) / 4, 256, 0, m.stream>>>(m.B, m.slotBase.as<uint32_t>(),
                                                              m.itemCount.as<uint32_t>(),
                                                              m.memberStart.as<uint32_t>(),
                                                              m.prank.as<uint32_t>(), m.members.as<uint32_t>());
  GPA_HIP(hipGetLastError());
  GPA_HIP(hipStreamSynchronize(m.stream));
  return m.members.as<uint32_t>();
}
const uint32_t *Pipeline::devClassMember() const { return impl_->oMember(); }

void xorMergeForward(const double *values, const uint32_t *index, const uint32_t *segStart,
                     uint32_t numSegments, double *out, void *stream) {
  if (!numSegments) return;
  xorMergeFwdKernel<<<(numSegments + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      values, index, segStart, numSegments, out);
  GPA_HIP(hipGetLastError());
}

void nodeRatesForward(const double *theta, const uint32_t *compLoc, const uint8_t *compKind,
                      const uint32_t *nodeStart, const uint32_t *nodeComp, uint32_t numNodes, double *q,
                      void *stream) {
  if (!numNodes) return;
  nodeRatesFwdKernel<<<(numNodes + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      theta, compLoc, compKind, nodeStart, nodeComp, numNodes, q);
  GPA_HIP(hipGetLastError());
}

void nodeRatesBackward(const double *theta, const uint32_t *compLoc, const uint8_t *compKind,
                       const uint32_t *nodeStart, const uint32_t *nodeComp, uint32_t numNodes, const double *gradQ,
                       double *gradComp, const uint32_t *locStart, const uint32_t *locComp, uint32_t numLocations,
                       double *gradTheta, void *stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (numNodes)
    nodeRatesBwdKernel<<<(numNodes + 255) / 256, 256, 0, st>>>(theta, compLoc, compKind, nodeStart, nodeComp,
                                                               numNodes, gradQ, gradComp);
  if (numLocations)
    locationGradKernel<<<(numLocations + 255) / 256, 256, 0, st>>>(locStart, locComp, gradComp, numLocations,
                                                                   gradTheta);
  GPA_HIP(hipGetLastError());
}

void xorMergeBackward(const double *values, const uint32_t *index, const uint32_t *segStart,
                      uint32_t numSegments, const double *gradOut, double *gradValues, void *stream) {
  if (!numSegments) return;
  xorMergeBwdKernel<<<(numSegments + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      values, index, segStart, numSegments, gradOut, gradValues);
  GPA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Allocation-free backward of the node rates: the per-component gradients
// live in a fixed device array instead of a buffer the caller passes in.
// Each GPU holds its own copy for the life of the process; its address on a
// device is looked up once, with that device current, and used from then on
// (the array's name in host code is only a stand-in for the symbol).
// ---------------------------------------------------------------------------

static __device__ __align__(256) double scratch[kNoAllocMaxComponents];

namespace {

constexpr int kScratchDevices = 64;
std::mutex gScratchMu;
double *gScratchPtr[kScratchDevices] = {};

double *scratchOn(int device) {
  if (device < 0 || device >= kScratchDevices)
    throw RuntimeError("device ordinal " + std::to_string(device) + " out of range");
  std::lock_guard<std::mutex> lk(gScratchMu);
  if (!gScratchPtr[device]) {
    int prev = 0;
    GPA_HIP(hipGetDevice(&prev));
    GPA_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t r = hipSetDevice(prev);
    if (e != hipSuccess)
      throw RuntimeError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    if (r != hipSuccess) throw RuntimeError(std::string("hipSetDevice failed: ") + hipGetErrorString(r));
    if (!p) throw RuntimeError("hipGetSymbolAddress(scratch) returned NULL");
    gScratchPtr[device] = static_cast<double *>(p);
  }
  return gScratchPtr[device];
}

}  // namespace

void nodeRatesBackwardNoAlloc(const double *theta, const uint32_t *compLoc, const uint8_t *compKind,
                              const uint32_t *nodeStart, const uint32_t *nodeComp, uint32_t numNodes,
                              uint32_t numComponents, const double *gradQ, const uint32_t *locStart,
                              const uint32_t *locComp, uint32_t numLocations, double *gradTheta, void *stream) {
  if (numComponents > kNoAllocMaxComponents)
    throw InvalidArgument("num_components = " + std::to_string(numComponents) +
                          " exceeds the allocation-free maximum of " + std::to_string(kNoAllocMaxComponents) +
                          "; use gp_node_rates_backward with a caller buffer");
  if (!numNodes && !numLocations) return;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  int device = 0;
  if (st) GPA_HIP(hipStreamGetDevice(st, &device));
  else GPA_HIP(hipGetDevice(&device));
  // the same kernels as nodeRatesBackward, with the fixed array as their
  // per-component buffer (every entry the location sums read is written by
  // the node kernel of the same call, so nothing needs clearing)
  nodeRatesBackward(theta, compLoc, compKind, nodeStart, nodeComp, numNodes, gradQ, scratchOn(device), locStart,
                    locComp, numLocations, gradTheta, stream);
}

}  // namespace gpa
