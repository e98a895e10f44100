// engine.hip -- gfx950 kernels and host orchestration of the DEM compile
// pipeline (see engine.hpp for the stage list).
//
// Conventions: one wavefront = 64 lanes; "rank" = index of a p>0 STEPG node
// (ascending node order) across all circuits of a load; "column" = one
// 32-bit word of detector/observable bits; bit t of the concatenated words is
// detector t (t < D) or observable t - D.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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
constexpr int kSweepBlock = 1024;
constexpr int kClassBlock = 256;
constexpr int kScanBlock = 1024;
constexpr int kScanTile = 4 * kScanBlock;
constexpr int kBucketItemsPerWave = 256;  // wave-register sort capacity (4 per lane)
constexpr int kTargetBucketFill = 64;     // mean items per bucket the digit width aims at

// Per-circuit descriptor (device and host).
struct VDesc {
  uint32_t numLayers, npl, numCols, numDetectors;
  uint32_t numObservables, numRanks, numSeeds, bucketBits;
  uint32_t bucketBase, numBuckets, pad0, pad1;
  uint64_t layerActOff, actOff, seedOff, colOff, rankOff, sOff;
};

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

// Second, independent 64-bit hash of a class: used only to detect FNV
// collisions between different classes inside one equal-hash run.
__device__ __forceinline__ uint64_t mix2(uint64_t h, uint32_t t) {
  h ^= (static_cast<uint64_t>(t) + 0x9E3779B97F4A7C15ull) + (h << 6) + (h >> 2);
  return h * 0xD6E8FEB86659FD93ull;
}

// Bitonic sort of 64*R (hash, rank) keys held R per lane; element index
// e = s * 64 + lane.  Distances below 64 exchange through lane shuffles,
// larger ones inside a lane's registers.
template <int R>
__device__ __forceinline__ void waveBitonicSort(uint64_t (&h)[R], uint32_t (&r)[R], uint32_t lane) {
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
            const bool sw = asc ? keyLess(h[sp], r[sp], h[s], r[s]) : keyLess(h[s], r[s], h[sp], r[sp]);
            if (sw) {
              uint64_t th = h[s];
              h[s] = h[sp];
              h[sp] = th;
              uint32_t tr = r[s];
              r[s] = r[sp];
              r[sp] = tr;
            }
          }
        }
      } else {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
          const uint64_t ho = __shfl_xor(h[s], static_cast<int>(j));
          const uint32_t ro = __shfl_xor(r[s], static_cast<int>(j));
          const bool lower = (lane & j) == 0;
          const bool asc = (e & k) == 0;
          const bool otherLess = keyLess(ho, ro, h[s], r[s]);
          const bool keepMin = lower == asc;
          const bool take = keepMin ? otherLess : keyLess(h[s], r[s], ho, ro);
          if (take) {
            h[s] = ho;
            r[s] = ro;
          }
        }
      }
    }
  }
}

struct RunCarry {
  double p;
  uint64_t h2;
  uint64_t prevHash;
  uint32_t rep;
  uint32_t runs;
  uint32_t nnzSum;
  uint32_t havePrev;
};

struct RunOut {
  const double *prob;      // rank -> node probability
  const uint64_t *h2;      // rank -> second hash
  const uint32_t *nnz;     // rank -> class size
  uint32_t *sortedRank;    // bucket-sorted ranks (in place over the partition)
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
// the XOR-merge of its members, its representative the first member.
template <int R>
__device__ __forceinline__ void scanRuns(const uint64_t (&h)[R], const uint32_t (&rk)[R], uint32_t cnt,
                                         uint32_t chunkBase, uint32_t bs, RunCarry &cy, uint32_t lane,
                                         const RunOut &o) {
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
    const double own = valid ? o.prob[r] : 0.0;
    const uint64_t ownH2 = valid ? o.h2[r] : 0ull;
    double v = own;
    uint32_t rep = r;
    uint64_t hf = ownH2;
    int fl = f[s] ? 1 : 0;
    const uint32_t nz = f[s] ? o.nnz[r] : 0u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double pv = __shfl_up(v, d);
      const uint32_t pr = __shfl_up(rep, d);
      const uint64_t ph2 = __shfl_up(hf, d);
      const int pf = __shfl_up(fl, d);
      if (lane >= static_cast<uint32_t>(d)) {
        if (!fl) {
          v = xmerge(pv, v);
          rep = pr;
          hf = ph2;
        }
        fl |= pf;
      }
    }
    if (!fl) {  // the run began before this slot
      v = xmerge(cy.p, v);
      rep = cy.rep;
      hf = cy.h2;
    }
    const uint64_t bal = __ballot(f[s]);
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
      if (e + 1 == cnt || nf) o.runP[bs + runIdx] = v;
    }
    cy.p = __shfl(v, 63);
    cy.rep = __shfl(rep, 63);
    cy.h2 = __shfl(hf, 63);
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

// Backward sensitivity sweep of one 32-bit column of one circuit.  The
// sensitivities of layers l, l+1, l+2 live in an LDS ring; each layer is one
// pass over the active-node records followed by one workgroup barrier, whose
// OR-reduction also detects when the column has become identically zero so
// the sweep can stop early (no seeds below and two zero layers above).
__global__ __launch_bounds__(kSweepBlock) void sweepKernel(
    const VDesc *__restrict__ desc, const uint2 *__restrict__ items,
    const uint32_t *__restrict__ layerAct, const uint4 *__restrict__ act,
    const uint32_t *__restrict__ actRank, const uint32_t *__restrict__ seedVals,
    const int32_t *__restrict__ colHi, const int32_t *__restrict__ colLow,
    int32_t *__restrict__ bandLo, uint32_t *__restrict__ S) {
  extern __shared__ uint32_t ring[];
  const uint2 it = items[blockIdx.x];
  const VDesc d = desc[it.x];
  const uint32_t c = it.y;
  const int32_t hi = colHi[d.colOff + c];
  if (hi < 0) {
    if (threadIdx.x == 0) bandLo[d.colOff + c] = 0x7FFFFFFF;
    return;
  }
  const int32_t lowSeed = colLow[d.colOff + c];
  const uint32_t npl = d.npl;
  for (uint32_t i = threadIdx.x; i < 3 * npl; i += kSweepBlock) ring[i] = 0u;
  __syncthreads();

  const uint32_t *la = layerAct + d.layerActOff;
  const uint4 *ac = act + d.actOff;
  const uint32_t *ar = actRank + d.actOff;
  const uint32_t *sv = seedVals + d.seedOff + static_cast<uint64_t>(c) * d.numSeeds;
  uint32_t *Sc = S + d.sOff + static_cast<uint64_t>(c) * d.numRanks;

  uint32_t cur = static_cast<uint32_t>(hi % 3) * npl;
  uint32_t nb1 = static_cast<uint32_t>((hi + 1) % 3) * npl;
  uint32_t nb2 = static_cast<uint32_t>((hi + 2) % 3) * npl;
  bool prevZero = true;
  int32_t lo = 0;

  // prefetch of this thread's first record of the next layer
  uint32_t a0 = la[hi], a1 = la[hi + 1];
  uint32_t a = a0 + threadIdx.x;
  uint4 rec = make_uint4(0, NONE, NONE, NONE);
  uint32_t rrank = NONE, seed = 0;
  if (a < a1) {
    rec = ac[a];
    rrank = ar[a];
    seed = rec.w != NONE ? sv[rec.w] : 0u;
  }
  for (int32_t l = hi; l >= 0; --l) {
    uint32_t any = 0;
    if (a < a1) {
      uint32_t s = seed;
      if (rec.y != NONE) s ^= ring[((rec.y >> 30) == 1 ? nb1 : nb2) + (rec.y & 0x3FFFFFFFu)];
      if (rec.z != NONE) s ^= ring[((rec.z >> 30) == 1 ? nb1 : nb2) + (rec.z & 0x3FFFFFFFu)];
      ring[cur + rec.x] = s;
      if (rrank != NONE) Sc[rrank] = s;
      any |= s;
      for (uint32_t b = a + kSweepBlock; b < a1; b += kSweepBlock) {  // wide layers
        const uint4 rb = ac[b];
        uint32_t sb = rb.w != NONE ? sv[rb.w] : 0u;
        if (rb.y != NONE) sb ^= ring[((rb.y >> 30) == 1 ? nb1 : nb2) + (rb.y & 0x3FFFFFFFu)];
        if (rb.z != NONE) sb ^= ring[((rb.z >> 30) == 1 ? nb1 : nb2) + (rb.z & 0x3FFFFFFFu)];
        ring[cur + rb.x] = sb;
        const uint32_t rr = ar[b];
        if (rr != NONE) Sc[rr] = sb;
        any |= sb;
      }
    }
    // prefetch the next layer's first record before the barrier
    if (l > 0) {
      a1 = a0;
      a0 = la[l - 1];
      a = a0 + threadIdx.x;
      if (a < a1) {
        rec = ac[a];
        rrank = ar[a];
        seed = rec.w != NONE ? sv[rec.w] : 0u;
      }
    }
    const int anyL = __syncthreads_or(any != 0);
    if (!anyL && prevZero && l <= lowSeed) {
      lo = l;
      break;
    }
    prevZero = !anyL;
    const uint32_t t = nb2;
    nb2 = nb1;
    nb1 = cur;
    cur = t;
  }
  if (threadIdx.x == 0) bandLo[d.colOff + c] = lo;
}

// Class extraction + hashing of every p>0 node: the ascending list of set
// bits, FNV-1a-64 over it padded with 0xFFFFFFFF to 32 slots (lists longer
// than 32 hash all entries, unpadded), a second hash, the class size, and
// the bucket histogram.
__global__ __launch_bounds__(kClassBlock) void classifyKernel(
    const VDesc *__restrict__ desc, const uint2 *__restrict__ tiles,
    const int32_t *__restrict__ colHi, const int32_t *__restrict__ bandLo,
    const uint32_t *__restrict__ rankLayer, const uint32_t *__restrict__ S,
    uint64_t *__restrict__ hashOut, uint64_t *__restrict__ h2Out, uint32_t *__restrict__ nnzOut,
    uint32_t *__restrict__ bucketCount) {
  const uint2 t = tiles[blockIdx.x];
  const VDesc d = desc[t.x];
  const uint32_t i = t.y + threadIdx.x;
  if (i >= d.numRanks) return;
  const uint64_t gi = d.rankOff + i;
  const int32_t l = static_cast<int32_t>(rankLayer[gi]);
  const int32_t *hiC = colHi + d.colOff;
  const int32_t *loC = bandLo + d.colOff;
  const uint32_t *Sv = S + d.sOff + i;
  uint64_t h = kFnvBasis, h2 = 0x84222325cbf29ce4ull;
  uint32_t cnt = 0;
  for (uint32_t c = 0; c < d.numCols; ++c) {
    if (l < loC[c] || l > hiC[c]) continue;
    uint32_t w = Sv[static_cast<uint64_t>(c) * d.numRanks];
    while (w) {
      const uint32_t tg = c * 32 + static_cast<uint32_t>(__builtin_ctz(w));
      h = (h ^ tg) * kFnvPrime;
      h2 = mix2(h2, tg);
      cnt++;
      w &= w - 1;
    }
  }
  if (cnt && cnt <= 32)
    for (uint32_t k = cnt; k < 32; ++k) h = (h ^ 0xFFFFFFFFull) * kFnvPrime;
  hashOut[gi] = h;
  h2Out[gi] = mix2(h2, cnt);
  nnzOut[gi] = cnt;
  if (cnt) {
    const uint32_t b = d.bucketBase + (d.bucketBits ? static_cast<uint32_t>(h >> (64 - d.bucketBits)) : 0u);
    atomicAdd(&bucketCount[b], 1u);
  }
}

// MSD partition: scatter (hash, rank) into the buckets.  Order inside a
// bucket is arbitrary; the bucket sort orders by (hash, rank).
__global__ __launch_bounds__(kClassBlock) void partitionKernel(
    const VDesc *__restrict__ desc, const uint2 *__restrict__ tiles,
    const uint64_t *__restrict__ hash, const uint32_t *__restrict__ nnz,
    uint32_t *__restrict__ cursor, uint64_t *__restrict__ pkey, uint32_t *__restrict__ prank) {
  const uint2 t = tiles[blockIdx.x];
  const VDesc d = desc[t.x];
  const uint32_t i = t.y + threadIdx.x;
  if (i >= d.numRanks) return;
  const uint64_t gi = d.rankOff + i;
  if (!nnz[gi]) return;
  const uint64_t h = hash[gi];
  const uint32_t b = d.bucketBase + (d.bucketBits ? static_cast<uint32_t>(h >> (64 - d.bucketBits)) : 0u);
  const uint32_t pos = atomicAdd(&cursor[b], 1u);
  pkey[pos] = h;
  prank[pos] = static_cast<uint32_t>(gi);
}

template <int R>
__device__ void bucketWave(uint32_t b, uint32_t bs, uint32_t cnt, uint32_t lane,
                           const uint64_t *__restrict__ pkey, const RunOut &o,
                           uint32_t *__restrict__ runCount, uint32_t *__restrict__ runNnz) {
  uint64_t h[R];
  uint32_t r[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const uint32_t e = static_cast<uint32_t>(s) * 64 + lane;
    h[s] = e < cnt ? pkey[bs + e] : ~0ull;
    r[s] = e < cnt ? o.sortedRank[bs + e] : NONE;
  }
  waveBitonicSort<R>(h, r, lane);
  RunCarry cy{0.0, 0ull, 0ull, 0u, 0u, 0u, 0u};
  scanRuns<R>(h, r, cnt, 0u, bs, cy, lane, o);
  if (lane == 0) {
    runCount[b] = cy.runs;
    runNnz[b] = cy.nnzSum;
  }
}

// One wave64 per bucket: sort in registers, merge runs.  Buckets above the
// wave capacity are queued for bucketFallbackKernel.
__global__ __launch_bounds__(256) void bucketKernel(
    uint32_t numBuckets, const uint32_t *__restrict__ bucketStart, const uint64_t *__restrict__ pkey,
    RunOut o, uint32_t *__restrict__ runCount, uint32_t *__restrict__ runNnz,
    uint32_t *__restrict__ overflowList) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= numBuckets) return;
  const uint32_t bs = bucketStart[b];
  const uint32_t cnt = bucketStart[b + 1] - bs;
  if (cnt == 0) {
    if (lane == 0) {
      runCount[b] = 0;
      runNnz[b] = 0;
    }
  } else if (cnt <= 64) {
    bucketWave<1>(b, bs, cnt, lane, pkey, o, runCount, runNnz);
  } else if (cnt <= 128) {
    bucketWave<2>(b, bs, cnt, lane, pkey, o, runCount, runNnz);
  } else if (cnt <= kBucketItemsPerWave) {
    bucketWave<4>(b, bs, cnt, lane, pkey, o, runCount, runNnz);
  } else if (lane == 0) {
    const uint32_t idx = atomicAdd(&o.flags[2], 1u);
    overflowList[idx] = b;
  }
}

// Oversized buckets: workgroup bitonic sort in global scratch, then one
// wave merges the runs chunk by chunk.
__global__ __launch_bounds__(256) void bucketFallbackKernel(
    const uint32_t *__restrict__ bucketStart, const uint64_t *__restrict__ pkey, RunOut o,
    uint32_t *__restrict__ runCount, uint32_t *__restrict__ runNnz,
    const uint32_t *__restrict__ overflowList, uint64_t *__restrict__ scratchKey,
    uint32_t *__restrict__ scratchRank) {
  const uint32_t numOver = o.flags[2];
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t idx = blockIdx.x; idx < numOver; idx += gridDim.x) {
    const uint32_t b = overflowList[idx];
    const uint32_t bs = bucketStart[b];
    const uint32_t cnt = bucketStart[b + 1] - bs;
    uint32_t P = 1;
    while (P < cnt) P <<= 1;
    uint64_t *K = scratchKey + 2ull * bs;
    uint32_t *Rk = scratchRank + 2ull * bs;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
      K[i] = i < cnt ? pkey[bs + i] : ~0ull;
      Rk[i] = i < cnt ? o.sortedRank[bs + i] : NONE;
    }
    __syncthreads();
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
    if (threadIdx.x < 64) {
      RunCarry cy{0.0, 0ull, 0ull, 0u, 0u, 0u, 0u};
      for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {
        uint64_t h[1];
        uint32_t r[1];
        const uint32_t e = c0 + lane;
        h[0] = e < cnt ? K[e] : ~0ull;
        r[0] = e < cnt ? Rk[e] : NONE;
        const uint32_t n = cnt - c0 < 64 ? cnt - c0 : 64;
        scanRuns<1>(h, r, n, c0, bs, cy, lane, o);
      }
      if (lane == 0) {
        runCount[b] = cy.runs;
        runNnz[b] = cy.nnzSum;
      }
    }
    __syncthreads();
  }
}

// Compact output: one wave per bucket, one lane per run.
__global__ __launch_bounds__(256) void emitKernel(
    const VDesc *__restrict__ desc, uint32_t numVariants, uint32_t numBuckets,
    const uint32_t *__restrict__ bucketStart, const uint32_t *__restrict__ runCount,
    const uint32_t *__restrict__ classStart, const uint32_t *__restrict__ tgtStart,
    const uint32_t *__restrict__ runRank, const double *__restrict__ runP,
    const uint64_t *__restrict__ runHash, const uint32_t *__restrict__ runTgtPre,
    const uint32_t *__restrict__ runMember, const uint32_t *__restrict__ rankLayer,
    const int32_t *__restrict__ colHi, const int32_t *__restrict__ bandLo,
    const uint32_t *__restrict__ S, double *__restrict__ outP, uint32_t *__restrict__ outOff,
    uint64_t *__restrict__ outHash, uint32_t *__restrict__ outMember,
    uint32_t *__restrict__ outTargets, uint64_t capTargets, uint32_t *__restrict__ flags) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= numBuckets) return;
  // circuit owning bucket b
  uint32_t lo = 0, hi = numVariants - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (desc[mid].bucketBase <= b) lo = mid;
    else hi = mid - 1;
  }
  const VDesc d = desc[lo];
  const uint32_t nR = runCount[b], bs = bucketStart[b], cb = classStart[b], tb = tgtStart[b];
  const int32_t *hiC = colHi + d.colOff;
  const int32_t *loC = bandLo + d.colOff;
  for (uint32_t r = lane; r < nR; r += 64) {
    const uint32_t c = cb + r;
    const uint32_t rk = runRank[bs + r];
    uint64_t pos = tb + runTgtPre[bs + r];
    outP[c] = runP[bs + r];
    outHash[c] = runHash[bs + r];
    outOff[c] = static_cast<uint32_t>(pos);
    outMember[c] = bs + runMember[bs + r];
    const uint32_t i = static_cast<uint32_t>(rk - d.rankOff);
    const int32_t l = static_cast<int32_t>(rankLayer[rk]);
    const uint32_t *Sv = S + d.sOff + i;
    for (uint32_t col = 0; col < d.numCols; ++col) {
      if (l < loC[col] || l > hiC[col]) continue;
      uint32_t w = Sv[static_cast<uint64_t>(col) * d.numRanks];
      while (w) {
        if (pos < capTargets) outTargets[pos] = col * 32 + static_cast<uint32_t>(__builtin_ctz(w));
        pos++;
        w &= w - 1;
      }
    }
    if (pos > capTargets) atomicOr(&flags[1], 1u);
  }
  if (b == numBuckets - 1 && lane == 0) {
    outOff[classStart[numBuckets]] = tgtStart[numBuckets];
    outMember[classStart[numBuckets]] = bucketStart[numBuckets];
  }
}

// Exclusive scan, tile level: 4096 items per 1024-thread block.
__global__ __launch_bounds__(kScanBlock) void scanTileKernel(const uint32_t *__restrict__ in, uint32_t n,
                                                             uint32_t *__restrict__ out,
                                                             uint32_t *__restrict__ out2,
                                                             uint32_t *__restrict__ tileSums) {
  __shared__ uint32_t waveSums[kScanBlock / 64];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t tsum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    tsum += v[k];
  }
  const uint32_t incl = waveInclusiveSum(tsum, lane);
  if (lane == 63) waveSums[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint32_t ws = lane < kScanBlock / 64 ? waveSums[lane] : 0u;
    const uint32_t wi = waveInclusiveSum(ws, lane);
    if (lane < kScanBlock / 64) waveSums[lane] = wi - ws;
    if (lane == kScanBlock / 64 - 1) {
      if (gridDim.x == 1) {
        out[n] = wi;
        if (out2) out2[n] = wi;
      } else {
        tileSums[blockIdx.x] = wi;
      }
    }
  }
  __syncthreads();
  uint32_t run = waveSums[wave] + incl - tsum;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) {
      out[base + k] = run;
      if (out2) out2[base + k] = run;
    }
    run += v[k];
  }
}

__global__ __launch_bounds__(kScanBlock) void scanAddKernel(uint32_t n, const uint32_t *__restrict__ tileOff,
                                                            uint32_t numTiles, uint32_t *__restrict__ out,
                                                            uint32_t *__restrict__ out2) {
  const uint32_t tile = blockIdx.x;
  const uint32_t add = tileOff[tile];
  const uint32_t base = tile * kScanTile;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanBlock) {
    const uint32_t i = base + k;
    if (i < n) {
      out[i] += add;
      if (out2) out2[i] += add;
    }
  }
  if (tile == 0 && threadIdx.x == 0) {
    out[n] = tileOff[numTiles];
    if (out2) out2[n] = tileOff[numTiles];
  }
}

__global__ void zeroKernel(uint32_t *__restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0u;
}

__global__ void variantBoundsKernel(const VDesc *__restrict__ desc, uint32_t numVariants,
                                    uint32_t numBuckets, const uint32_t *__restrict__ classStart,
                                    const uint32_t *__restrict__ tgtStart, uint32_t *__restrict__ out,
                                    const uint32_t *__restrict__ flags) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > numVariants) return;
  const uint32_t bb = v < numVariants ? desc[v].bucketBase : numBuckets;
  out[2 * v] = classStart[bb];
  out[2 * v + 1] = tgtStart[bb];
  if (v == 0) {
    out[2 * numVariants + 2] = flags[0];
    out[2 * numVariants + 3] = flags[1];
    out[2 * numVariants + 4] = flags[2];
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

// --------------------------------------------------------------------------
// host-side buffers
// --------------------------------------------------------------------------

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t bytes, bool frozen, const char *what) {
    if (bytes <= cap) return;
    if (frozen)
      throw RuntimeError(std::string("circuit exceeds the driver's capacity (") + what +
                         "); create the driver from a larger circuit");
    if (p) GPA_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    const size_t sz = std::max<size_t>(bytes, 256);
    GPA_HIP(hipMalloc(&p, sz));
    cap = sz;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

struct HostBuf {
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

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

}  // namespace

struct Pipeline::Impl {
  int device;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  bool frozen = false;
  bool timedRun = false;
  hipEvent_t *evOverride = nullptr;   // per-iteration events of benchDevice
  std::vector<hipEvent_t> evPool;

  // inputs (one device arena, one pinned staging arena with the same layout)
  DevBuf in;
  HostBuf inHost;
  size_t oDesc = 0, oItems = 0, oTiles = 0, oLayerAct = 0, oAct = 0, oActRank = 0, oSeed = 0,
         oColHi = 0, oColLow = 0, oRankLayer = 0, oRankProb = 0, inBytes = 0;
  uint32_t V = 0, nItems = 0, nTiles = 0, B = 0, R = 0, Ccols = 0, maxNpl = 0;
  uint64_t Swords = 0;

  DevBuf bandLo, S, hash, h2, nnz, bucketCount, bucketStart, cursor, pkey, prank, runRank, runP,
      runHash, runTgtPre, runMember, runCount, runNnz, classStart, tgtStart, outP, outOff, outHash,
      outMember, outTargets, scratchKey, scratchRank, flags, overflowList, scanTmp, scanTmp2, vBounds;
  uint64_t capTargets = 0;

  // host results
  HostBuf hP, hOff, hHash, hTargets, hVB;
  uint32_t totalClasses = 0;
  uint64_t totalTargets = 0;
  std::vector<uint32_t> vClass, vTgt;
  std::vector<VDesc> descs;

  size_t sweepLds() const { return static_cast<size_t>(3) * maxNpl * 4; }

  void scan(const uint32_t *inp, uint32_t n, uint32_t *out, uint32_t *out2) {
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles <= 1) {
      scanTileKernel<<<1, kScanBlock, 0, stream>>>(inp, n, out, out2, nullptr);
      return;
    }
    if (tiles > kScanTile) throw RuntimeError("scan too large");
    uint32_t *ts = scanTmp.as<uint32_t>();
    uint32_t *to = scanTmp2.as<uint32_t>();
    scanTileKernel<<<tiles, kScanBlock, 0, stream>>>(inp, n, out, out2, ts);
    scanTileKernel<<<1, kScanBlock, 0, stream>>>(ts, tiles, to, nullptr, nullptr);
    scanAddKernel<<<tiles, kScanBlock, 0, stream>>>(n, to, tiles, out, out2);
  }
};

Pipeline::Pipeline(int device, bool growable) : device_(device), growable_(growable), impl_(new Impl) {
  impl_->device = device;
  GPA_HIP(hipSetDevice(device));
  GPA_HIP(hipStreamCreateWithFlags(&impl_->stream, hipStreamNonBlocking));
  for (auto &e : impl_->ev) GPA_HIP(hipEventCreate(&e));
}

Pipeline::~Pipeline() {
  (void)hipSetDevice(device_);
  if (impl_->stream) (void)hipStreamSynchronize(impl_->stream);
  for (auto &e : impl_->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : impl_->evPool) (void)hipEventDestroy(e);
  if (impl_->stream) (void)hipStreamDestroy(impl_->stream);
  delete impl_;
}

void *Pipeline::stream() const { return impl_->stream; }

static uint32_t bucketBitsFor(uint32_t numRanks) {
  uint32_t b = 0;
  while (b < 24 && (static_cast<uint64_t>(kTargetBucketFill) << b) < numRanks) b++;
  return b;
}

void Pipeline::reserve(const std::vector<const StepgImage *> &imgs) {
  impl_->frozen = false;
  load(imgs);
  impl_->frozen = !growable_;
}

void Pipeline::load(const std::vector<const StepgImage *> &imgs) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  GPA_HIP(hipStreamSynchronize(m.stream));
  const bool frozen = m.frozen;
  const uint32_t V = static_cast<uint32_t>(imgs.size());
  if (V == 0) throw InvalidArgument("no circuits to compile");

  // layout
  m.descs.assign(V, VDesc{});
  variants_.assign(V, VariantInfo{});
  uint64_t layerActOff = 0, actOff = 0, seedOff = 0, colOff = 0, rankOff = 0, sOff = 0;
  uint32_t bucketBase = 0, maxNpl = 0;
  std::vector<uint2> items, tiles;
  std::vector<std::pair<int32_t, uint32_t>> order;  // (band estimate, item index)
  for (uint32_t v = 0; v < V; ++v) {
    const StepgImage &g = *imgs[v];
    if (3ull * g.npl * 4 > 160 * 1024)
      throw RuntimeError("circuit has too many STEPG nodes per layer for the LDS sweep (" +
                         std::to_string(g.npl) + ")");
    VDesc &d = m.descs[v];
    d.numLayers = g.numLayers;
    d.npl = g.npl;
    d.numCols = g.numCols;
    d.numDetectors = g.numDetectors;
    d.numObservables = g.numObservables;
    d.numRanks = g.numRanks;
    d.numSeeds = g.numSeeds;
    d.bucketBits = bucketBitsFor(g.numRanks);
    d.numBuckets = 1u << d.bucketBits;
    d.bucketBase = bucketBase;
    d.layerActOff = layerActOff;
    d.actOff = actOff;
    d.seedOff = seedOff;
    d.colOff = colOff;
    d.rankOff = rankOff;
    d.sOff = sOff;
    variants_[v] = VariantInfo{g.numRanks, g.numCols, g.numDetectors, g.numObservables, bucketBase,
                               d.numBuckets};
    for (uint32_t c = 0; c < g.numCols; ++c) {
      const int32_t hiL = g.colHi[c];
      const int32_t est = hiL < 0 ? -1 : hiL - g.colLow[c] + 1;
      order.push_back({est, static_cast<uint32_t>(items.size())});
      items.push_back(make_uint2(v, c));
    }
    for (uint32_t t = 0; t < g.numRanks; t += kClassBlock) tiles.push_back(make_uint2(v, t));
    layerActOff += g.numLayers + 1;
    actOff += g.numActive;
    seedOff += static_cast<uint64_t>(g.numCols) * g.numSeeds;
    colOff += g.numCols;
    rankOff += g.numRanks;
    sOff += static_cast<uint64_t>(g.numRanks) * g.numCols;
    bucketBase += d.numBuckets;
    maxNpl = std::max(maxNpl, g.npl);
  }
  if (rankOff >= 0x7FFFFFFFull) throw RuntimeError("too many error nodes in one load");
  // longest bands first (they bound the sweep's critical path)
  std::stable_sort(order.begin(), order.end(),
                   [](const auto &a, const auto &b) { return a.first > b.first; });
  std::vector<uint2> sortedItems(items.size());
  for (size_t k = 0; k < order.size(); ++k) sortedItems[k] = items[order[k].second];

  m.V = V;
  m.nItems = static_cast<uint32_t>(sortedItems.size());
  m.nTiles = static_cast<uint32_t>(tiles.size());
  m.B = bucketBase;
  m.R = static_cast<uint32_t>(rankOff);
  m.Ccols = static_cast<uint32_t>(colOff);
  m.Swords = sOff;
  m.maxNpl = maxNpl;

  // input arena
  size_t off = 0;
  auto place = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  m.oDesc = place(sizeof(VDesc) * V);
  m.oItems = place(sizeof(uint2) * m.nItems);
  m.oTiles = place(sizeof(uint2) * m.nTiles);
  m.oLayerAct = place(4 * layerActOff);
  m.oAct = place(16 * actOff);
  m.oActRank = place(4 * actOff);
  m.oSeed = place(4 * seedOff);
  m.oColHi = place(4 * colOff);
  m.oColLow = place(4 * colOff);
  m.oRankLayer = place(4 * rankOff);
  m.oRankProb = place(8 * rankOff);
  m.inBytes = off;
  m.in.ensure(m.inBytes, frozen, "inputs");
  m.inHost.ensure(m.inBytes);
  char *hb = m.inHost.as<char>();
  std::memcpy(hb + m.oDesc, m.descs.data(), sizeof(VDesc) * V);
  std::memcpy(hb + m.oItems, sortedItems.data(), sizeof(uint2) * m.nItems);
  std::memcpy(hb + m.oTiles, tiles.data(), sizeof(uint2) * m.nTiles);
  for (uint32_t v = 0; v < V; ++v) {
    const StepgImage &g = *imgs[v];
    const VDesc &d = m.descs[v];
    std::memcpy(hb + m.oLayerAct + 4 * d.layerActOff, g.layerAct.data(), 4 * g.layerAct.size());
    std::memcpy(hb + m.oAct + 16 * d.actOff, g.act.data(), 4 * g.act.size());
    std::memcpy(hb + m.oActRank + 4 * d.actOff, g.actRank.data(), 4 * g.actRank.size());
    std::memcpy(hb + m.oSeed + 4 * d.seedOff, g.seedVals.data(), 4 * g.seedVals.size());
    std::memcpy(hb + m.oColHi + 4 * d.colOff, g.colHi.data(), 4 * g.colHi.size());
    std::memcpy(hb + m.oColLow + 4 * d.colOff, g.colLow.data(), 4 * g.colLow.size());
    std::memcpy(hb + m.oRankLayer + 4 * d.rankOff, g.rankLayer.data(), 4 * g.rankLayer.size());
    std::memcpy(hb + m.oRankProb + 8 * d.rankOff, g.rankProb.data(), 8 * g.rankProb.size());
  }

  // work buffers
  const uint64_t R = std::max<uint64_t>(m.R, 1), B = m.B;
  m.bandLo.ensure(4 * std::max<uint32_t>(m.Ccols, 1), frozen, "columns");
  m.S.ensure(4 * std::max<uint64_t>(m.Swords, 1), frozen, "sensitivity matrix");
  m.hash.ensure(8 * R, frozen, "nodes");
  m.h2.ensure(8 * R, frozen, "nodes");
  m.nnz.ensure(4 * R, frozen, "nodes");
  m.bucketCount.ensure(4 * (B + 1), frozen, "buckets");
  m.bucketStart.ensure(4 * (B + 1), frozen, "buckets");
  m.cursor.ensure(4 * (B + 1), frozen, "buckets");
  m.pkey.ensure(8 * R, frozen, "nodes");
  m.prank.ensure(4 * R, frozen, "nodes");
  m.runRank.ensure(4 * R, frozen, "nodes");
  m.runP.ensure(8 * R, frozen, "nodes");
  m.runHash.ensure(8 * R, frozen, "nodes");
  m.runTgtPre.ensure(4 * R, frozen, "nodes");
  m.runMember.ensure(4 * R, frozen, "nodes");
  m.runCount.ensure(4 * (B + 1), frozen, "buckets");
  m.runNnz.ensure(4 * (B + 1), frozen, "buckets");
  m.classStart.ensure(4 * (B + 1), frozen, "buckets");
  m.tgtStart.ensure(4 * (B + 1), frozen, "buckets");
  m.outP.ensure(8 * R, frozen, "nodes");
  m.outOff.ensure(4 * (R + 1), frozen, "nodes");
  m.outHash.ensure(8 * R, frozen, "nodes");
  m.outMember.ensure(4 * (R + 1), frozen, "nodes");
  if (m.capTargets == 0 || (!frozen && m.capTargets < 16 * R)) {
    m.outTargets.ensure(4 * 16 * R, frozen, "targets");
    m.capTargets = m.outTargets.cap / 4;
  }
  m.scratchKey.ensure(16 * R, frozen, "nodes");
  m.scratchRank.ensure(8 * R, frozen, "nodes");
  m.flags.ensure(64, frozen, "flags");
  m.overflowList.ensure(4 * (B + 1), frozen, "buckets");
  m.scanTmp.ensure(4 * (kScanTile + 1), frozen, "scan");
  m.scanTmp2.ensure(4 * (kScanTile + 1), frozen, "scan");
  m.vBounds.ensure(4 * (2 * V + 8), frozen, "circuits");
  m.hVB.ensure(4 * (2 * V + 8));

  if (m.sweepLds() > 64 * 1024)
    GPA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(sweepKernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(m.sweepLds())));

  // upload (one copy)
  GPA_HIP(hipEventRecord(m.ev[0], m.stream));
  GPA_HIP(hipMemcpyAsync(m.in.p, m.inHost.p, m.inBytes, hipMemcpyHostToDevice, m.stream));
  GPA_HIP(hipEventRecord(m.ev[1], m.stream));
  GPA_HIP(hipStreamSynchronize(m.stream));
  float ms = 0;
  GPA_HIP(hipEventElapsedTime(&ms, m.ev[0], m.ev[1]));
  times_.upload = ms * 1e3;
}

void Pipeline::run(bool timed) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  char *ib = m.in.as<char>();
  const VDesc *desc = reinterpret_cast<const VDesc *>(ib + m.oDesc);
  const uint2 *items = reinterpret_cast<const uint2 *>(ib + m.oItems);
  const uint2 *tiles = reinterpret_cast<const uint2 *>(ib + m.oTiles);
  const uint32_t *layerAct = reinterpret_cast<const uint32_t *>(ib + m.oLayerAct);
  const uint4 *act = reinterpret_cast<const uint4 *>(ib + m.oAct);
  const uint32_t *actRank = reinterpret_cast<const uint32_t *>(ib + m.oActRank);
  const uint32_t *seed = reinterpret_cast<const uint32_t *>(ib + m.oSeed);
  const int32_t *colHi = reinterpret_cast<const int32_t *>(ib + m.oColHi);
  const int32_t *colLow = reinterpret_cast<const int32_t *>(ib + m.oColLow);
  const uint32_t *rankLayer = reinterpret_cast<const uint32_t *>(ib + m.oRankLayer);
  const double *rankProb = reinterpret_cast<const double *>(ib + m.oRankProb);
  hipStream_t st = m.stream;
  hipEvent_t *evs = m.evOverride ? m.evOverride : m.ev;
  auto mark = [&](int i) {
    if (timed) GPA_HIP(hipEventRecord(evs[i], st));
  };
  const uint32_t B = m.B;
  m.timedRun = timed;

  // counters that must start at zero (fill kernels, safe under graph capture)
  zeroKernel<<<std::max<uint32_t>(1, std::min<uint32_t>((B + 256) / 256, 1024)), 256, 0, st>>>(
      m.bucketCount.as<uint32_t>(), B + 1);
  zeroKernel<<<1, 64, 0, st>>>(m.flags.as<uint32_t>(), 16);
  mark(0);
  if (m.nItems)
    sweepKernel<<<m.nItems, kSweepBlock, m.sweepLds(), st>>>(desc, items, layerAct, act, actRank, seed,
                                                             colHi, colLow, m.bandLo.as<int32_t>(),
                                                             m.S.as<uint32_t>());
  mark(1);
  if (m.nTiles)
    classifyKernel<<<m.nTiles, kClassBlock, 0, st>>>(desc, tiles, colHi, m.bandLo.as<int32_t>(),
                                                     rankLayer, m.S.as<uint32_t>(),
                                                     m.hash.as<uint64_t>(), m.h2.as<uint64_t>(),
                                                     m.nnz.as<uint32_t>(), m.bucketCount.as<uint32_t>());
  mark(2);
  m.scan(m.bucketCount.as<uint32_t>(), B, m.bucketStart.as<uint32_t>(), m.cursor.as<uint32_t>());
  if (m.nTiles)
    partitionKernel<<<m.nTiles, kClassBlock, 0, st>>>(desc, tiles, m.hash.as<uint64_t>(),
                                                      m.nnz.as<uint32_t>(), m.cursor.as<uint32_t>(),
                                                      m.pkey.as<uint64_t>(), m.prank.as<uint32_t>());
  mark(3);
  RunOut o{rankProb,
           m.h2.as<uint64_t>(),
           m.nnz.as<uint32_t>(),
           m.prank.as<uint32_t>(),
           m.runRank.as<uint32_t>(),
           m.runP.as<double>(),
           m.runHash.as<uint64_t>(),
           m.runTgtPre.as<uint32_t>(),
           m.runMember.as<uint32_t>(),
           m.flags.as<uint32_t>()};
  bucketKernel<<<(B + 3) / 4, 256, 0, st>>>(B, m.bucketStart.as<uint32_t>(), m.pkey.as<uint64_t>(), o,
                                            m.runCount.as<uint32_t>(), m.runNnz.as<uint32_t>(),
                                            m.overflowList.as<uint32_t>());
  bucketFallbackKernel<<<64, 256, 0, st>>>(m.bucketStart.as<uint32_t>(), m.pkey.as<uint64_t>(), o,
                                           m.runCount.as<uint32_t>(), m.runNnz.as<uint32_t>(),
                                           m.overflowList.as<uint32_t>(), m.scratchKey.as<uint64_t>(),
                                           m.scratchRank.as<uint32_t>());
  mark(4);
  m.scan(m.runCount.as<uint32_t>(), B, m.classStart.as<uint32_t>(), nullptr);
  m.scan(m.runNnz.as<uint32_t>(), B, m.tgtStart.as<uint32_t>(), nullptr);
  emitKernel<<<(B + 3) / 4, 256, 0, st>>>(
      desc, m.V, B, m.bucketStart.as<uint32_t>(), m.runCount.as<uint32_t>(),
      m.classStart.as<uint32_t>(), m.tgtStart.as<uint32_t>(), m.runRank.as<uint32_t>(),
      m.runP.as<double>(), m.runHash.as<uint64_t>(), m.runTgtPre.as<uint32_t>(),
      m.runMember.as<uint32_t>(), rankLayer, colHi, m.bandLo.as<int32_t>(), m.S.as<uint32_t>(),
      m.outP.as<double>(), m.outOff.as<uint32_t>(), m.outHash.as<uint64_t>(),
      m.outMember.as<uint32_t>(), m.outTargets.as<uint32_t>(), m.capTargets, m.flags.as<uint32_t>());
  variantBoundsKernel<<<(m.V + 1 + 63) / 64, 64, 0, st>>>(desc, m.V, B, m.classStart.as<uint32_t>(),
                                                          m.tgtStart.as<uint32_t>(),
                                                          m.vBounds.as<uint32_t>(),
                                                          m.flags.as<uint32_t>());
  mark(5);
  GPA_HIP(hipGetLastError());
}

void Pipeline::download() {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  const uint32_t V = m.V;
  const auto t0 = std::chrono::steady_clock::now();
  GPA_HIP(hipMemcpyAsync(m.hVB.p, m.vBounds.p, 4 * (2 * V + 5), hipMemcpyDeviceToHost, m.stream));
  GPA_HIP(hipStreamSynchronize(m.stream));
  if (m.timedRun) {
    float ms[5] = {};
    for (int i = 0; i < 5; ++i) GPA_HIP(hipEventElapsedTime(&ms[i], m.ev[i], m.ev[i + 1]));
    times_.sweep = ms[0] * 1e3;
    times_.classify = ms[1] * 1e3;
    times_.partition = ms[2] * 1e3;
    times_.buckets = ms[3] * 1e3;
    times_.emit = ms[4] * 1e3;
    m.timedRun = false;
  }
  const uint32_t *vb = m.hVB.as<uint32_t>();
  if (vb[2 * V + 2]) throw RuntimeError("64-bit class hash collision between different error classes");
  m.totalClasses = vb[2 * V];
  m.totalTargets = vb[2 * V + 1];
  if (vb[2 * V + 3]) {
    // target buffer too small: grow it and re-emit everything
    if (m.frozen && 4 * m.totalTargets > m.outTargets.cap)
      throw RuntimeError("circuit exceeds the driver's capacity (targets)");
    const bool fr = m.frozen;
    m.frozen = false;
    m.outTargets.ensure(4 * (m.totalTargets + m.totalTargets / 4 + 64), false, "targets");
    m.frozen = fr;
    m.capTargets = m.outTargets.cap / 4;
    run(false);
    download();
    return;
  }
  m.vTgt.resize(V + 1);
  std::vector<uint32_t> vc(V + 1);
  for (uint32_t v = 0; v <= V; ++v) {
    vc[v] = vb[2 * v];
    m.vTgt[v] = vb[2 * v + 1];
  }
  m.vClass = vc;
  const uint64_t C = m.totalClasses, T = m.totalTargets;
  m.hP.ensure(8 * C + 8);
  m.hOff.ensure(4 * (C + 1));
  m.hHash.ensure(8 * C + 8);
  m.hTargets.ensure(4 * T + 4);
  if (C) {
    GPA_HIP(hipMemcpyAsync(m.hP.p, m.outP.p, 8 * C, hipMemcpyDeviceToHost, m.stream));
    GPA_HIP(hipMemcpyAsync(m.hHash.p, m.outHash.p, 8 * C, hipMemcpyDeviceToHost, m.stream));
  }
  GPA_HIP(hipMemcpyAsync(m.hOff.p, m.outOff.p, 4 * (C + 1), hipMemcpyDeviceToHost, m.stream));
  if (T) GPA_HIP(hipMemcpyAsync(m.hTargets.p, m.outTargets.p, 4 * T, hipMemcpyDeviceToHost, m.stream));
  GPA_HIP(hipStreamSynchronize(m.stream));
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
  out.probs = m.hP.as<double>() + c0;
  out.hashes = m.hHash.as<uint64_t>() + c0;
  out.targets = m.hTargets.as<uint32_t>() + t0;
  out.offsetsRel.resize(out.numErrors + 1);
  const uint32_t *off = m.hOff.as<uint32_t>();
  for (uint32_t i = 0; i <= out.numErrors; ++i) out.offsetsRel[i] = off[c0 + i] - t0;
  out.offsets = out.offsetsRel.data();
}

double Pipeline::benchDevice(uint32_t iters) {
  Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  iters = std::max<uint32_t>(iters, 1);
  run(false);  // warm
  // six stage events per iteration, read after the last one completes
  const size_t need = static_cast<size_t>(iters) * 6;
  while (m.evPool.size() < need) {
    hipEvent_t e;
    GPA_HIP(hipEventCreate(&e));
    m.evPool.push_back(e);
  }
  GPA_HIP(hipEventRecord(m.ev[6], m.stream));
  for (uint32_t i = 0; i < iters; ++i) {
    m.evOverride = m.evPool.data() + static_cast<size_t>(i) * 6;
    run(true);
  }
  m.evOverride = nullptr;
  m.timedRun = false;
  GPA_HIP(hipEventRecord(m.ev[7], m.stream));
  GPA_HIP(hipEventSynchronize(m.ev[7]));
  float ms = 0;
  GPA_HIP(hipEventElapsedTime(&ms, m.ev[6], m.ev[7]));
  double st[5] = {0, 0, 0, 0, 0};
  for (uint32_t i = 0; i < iters; ++i)
    for (int k = 0; k < 5; ++k) {
      float t = 0;
      GPA_HIP(hipEventElapsedTime(&t, m.evPool[i * 6 + k], m.evPool[i * 6 + k + 1]));
      st[k] += t;
    }
  times_.sweep = st[0] * 1e3 / iters;
  times_.classify = st[1] * 1e3 / iters;
  times_.partition = st[2] * 1e3 / iters;
  times_.buckets = st[3] * 1e3 / iters;
  times_.emit = st[4] * 1e3 / iters;
  return ms * 1e3 / iters;
}

uint64_t Pipeline::deviceBytes() const {
  const Impl &m = *impl_;
  const DevBuf *bufs[] = {&m.in, &m.bandLo, &m.S, &m.hash, &m.h2, &m.nnz, &m.bucketCount,
                          &m.bucketStart, &m.cursor, &m.pkey, &m.prank, &m.runRank, &m.runP,
                          &m.runHash, &m.runTgtPre, &m.runMember, &m.runCount, &m.runNnz,
                          &m.classStart, &m.tgtStart, &m.outP, &m.outOff, &m.outHash, &m.outMember,
                          &m.outTargets, &m.scratchKey, &m.scratchRank, &m.flags, &m.overflowList,
                          &m.scanTmp, &m.scanTmp2, &m.vBounds};
  uint64_t s = 0;
  for (const DevBuf *b : bufs) s += b->cap;
  return s;
}

void Pipeline::debugSensitivities(uint32_t v, std::vector<uint32_t> &cols32) const {
  const Impl &m = *impl_;
  GPA_HIP(hipSetDevice(device_));
  GPA_HIP(hipStreamSynchronize(m.stream));
  const VDesc &d = m.descs.at(v);
  std::vector<uint32_t> S(static_cast<size_t>(d.numRanks) * d.numCols);
  std::vector<int32_t> lo(d.numCols), hi(d.numCols);
  if (!S.empty())
    GPA_HIP(hipMemcpy(S.data(), m.S.as<uint32_t>() + d.sOff, 4 * S.size(), hipMemcpyDeviceToHost));
  GPA_HIP(hipMemcpy(lo.data(), m.bandLo.as<int32_t>() + d.colOff, 4 * d.numCols, hipMemcpyDeviceToHost));
  std::memcpy(hi.data(), m.inHost.as<char>() + m.oColHi + 4 * d.colOff, 4 * d.numCols);
  std::vector<uint32_t> layer(d.numRanks);
  std::memcpy(layer.data(), m.inHost.as<char>() + m.oRankLayer + 4 * d.rankOff, 4 * d.numRanks);
  cols32.assign(static_cast<size_t>(d.numRanks) * d.numCols, 0u);
  for (uint32_t c = 0; c < d.numCols; ++c)
    for (uint32_t i = 0; i < d.numRanks; ++i) {
      const int32_t l = static_cast<int32_t>(layer[i]);
      if (l >= lo[c] && l <= hi[c]) cols32[static_cast<size_t>(i) * d.numCols + c] = S[static_cast<size_t>(c) * d.numRanks + i];
    }
}

const uint32_t *Pipeline::devSortedRanks() const { return impl_->prank.as<uint32_t>(); }
const uint32_t *Pipeline::devClassMember() const { return impl_->outMember.as<uint32_t>(); }

void xorMergeForward(const double *values, const uint32_t *index, const uint32_t *segStart,
                     uint32_t numSegments, double *out, void *stream) {
  if (!numSegments) return;
  xorMergeFwdKernel<<<(numSegments + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      values, index, segStart, numSegments, out);
  GPA_HIP(hipGetLastError());
}

void xorMergeBackward(const double *values, const uint32_t *index, const uint32_t *segStart,
                      uint32_t numSegments, const double *gradOut, double *gradValues, void *stream) {
  if (!numSegments) return;
  xorMergeBwdKernel<<<(numSegments + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      values, index, segStart, numSegments, gradOut, gradValues);
  GPA_HIP(hipGetLastError());
}

}  // namespace gpa
