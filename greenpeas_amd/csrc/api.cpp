// api.cpp -- driver, batch and gradient-plan objects behind the C-ABI
// declared in include/greenpeas_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/greenpeas_amd.h"
#include "circuit.hpp"
#include "engine.hpp"
#include "stepg.hpp"

using gpa::InvalidArgument;
using gpa::RuntimeError;

namespace {

thread_local std::string g_lastError;

template <typename F>
int guarded(F &&f) {
  try {
    g_lastError.clear();
    f();
    return GP_OK;
  } catch (const std::invalid_argument &e) {
    g_lastError = e.what();
    return GP_INVALID_ARGUMENT;
  } catch (const std::exception &e) {
    g_lastError = e.what();
    return GP_RUNTIME_ERROR;
  } catch (...) {
    g_lastError = "unknown error";
    return GP_RUNTIME_ERROR;
  }
}

void checkLevel(int level) {
  if (level < 0 || level > 2) throw InvalidArgument("Unsupported correlation level.");
}

void requireText(const char *text) {
  if (!text) throw InvalidArgument("circuit text is NULL");
}

// DEM text in stim's format: one "error(p) D.. L.." line per error.
size_t formatDem(const gpa::VariantResult &r, std::string &out) {
  out.clear();
  out.reserve(static_cast<size_t>(r.numErrors) * 40 + r.numTargets * 6);
  char buf[64];
  for (uint32_t e = 0; e < r.numErrors; ++e) {
    int n = std::snprintf(buf, sizeof buf, "error(%.17g)", r.probs[e]);
    out.append(buf, n);
    for (uint32_t k = r.offsets[e]; k < r.offsets[e + 1]; ++k) {
      const uint32_t t = r.target(k);
      if (t < r.numDetectors) n = std::snprintf(buf, sizeof buf, " D%u", t);
      else n = std::snprintf(buf, sizeof buf, " L%u", t - r.numDetectors);
      out.append(buf, n);
    }
    out.push_back('\n');
  }
  return out.size();
}

struct Parsed {
  gpa::Circuit circuit;
  gpa::Stepg stepg;
  gpa::StepgImage image;
};

// `parallel`: large circuits are built on the worker pool (callers that
// build many circuits on their own threads pass false).
void parseInto(const char *text, int level, Parsed &p, bool keepComponents = false,
               const double *theta = nullptr, bool parallel = true) {
  p.circuit = gpa::Circuit::parse(text);
  gpa::buildStepg(p.circuit, level, p.stepg, theta, keepComponents, parallel);
  gpa::buildImage(p.stepg, p.image, parallel);
}

}  // namespace

struct gp_driver {
  int level = 2;
  int device = 0;
  Parsed parsed;
  std::unique_ptr<gpa::Pipeline> pipe;
  bool haveCircuit = false, loaded = false, haveResult = false;
  gpa::VariantResult result;
  double wallUs = 0;
};

struct gp_batch {
  int level = 2;
  int device = 0;
  std::vector<std::shared_ptr<const gpa::StepgImage>> images;  // the circuits' device images
  std::unique_ptr<gpa::Pipeline> pipe;
  std::unique_ptr<gpa::Pipeline> single;  // one-by-one baseline
  std::vector<gpa::VariantResult> results;
  std::vector<std::vector<double>> ownP;
  std::vector<std::vector<uint32_t>> ownOff, ownTg;
  std::vector<std::vector<uint64_t>> ownHash;
  double deviceUs = 0;
  bool loaded = false;
};

// A circuit parsed and built up to its device image (the host half of a
// compile), shared by the batches it is loaded into.
struct gp_prepared {
  int level = 2;
  std::shared_ptr<const gpa::StepgImage> image;
};

struct gp_grad_plan {
  int device = 0;
  Parsed parsed;
  std::unique_ptr<gpa::Pipeline> pipe;
  gpa::VariantResult result;
  uint32_t numLoc = 0, numComp = 0, numNodes = 0, numErrors = 0;
  uint32_t *dCompLoc = nullptr, *dNodeStart = nullptr, *dNodeComp = nullptr, *dClassStart = nullptr,
           *dClassNode = nullptr;
  uint8_t *dCompKind = nullptr;
  ~gp_grad_plan() {
    for (void *p : {static_cast<void *>(dCompLoc), static_cast<void *>(dNodeStart),
                    static_cast<void *>(dNodeComp), static_cast<void *>(dClassStart),
                    static_cast<void *>(dClassNode), static_cast<void *>(dCompKind)})
      if (p) (void)hipFree(p);
  }
};

extern "C" {

const char *gp_last_error(void) { return g_lastError.c_str(); }

const char *gp_version(void) { return "greenpeas_amd 0.1.0 (gfx950)"; }

int gp_get_driver_on_stream(const char *circuit_text, int correlation_level, int device, void *stream,
                            gp_driver **out) {
  return guarded([&] {
    if (!out) throw InvalidArgument("out is NULL");
    requireText(circuit_text);
    checkLevel(correlation_level);
    auto d = std::make_unique<gp_driver>();
    d->level = correlation_level;
    d->device = device;
    parseInto(circuit_text, correlation_level, d->parsed);
    d->pipe = std::make_unique<gpa::Pipeline>(device, false);
    if (stream) d->pipe->setStream(stream);
    // the DEM is the reference's (no hashes; they are read on request),
    // with 16-bit targets where they fit
    d->pipe->setDownloadHashes(false);
    d->pipe->setNarrowTargets(true);
    d->pipe->reserve({&d->parsed.image});
    d->haveCircuit = true;
    d->loaded = true;
    *out = d.release();
  });
}

int gp_get_driver(const char *circuit_text, int correlation_level, int device, gp_driver **out) {
  return gp_get_driver_on_stream(circuit_text, correlation_level, device, nullptr, out);
}

int gp_driver_set_stream(gp_driver *driver, void *stream) {
  return guarded([&] {
    if (!driver) throw InvalidArgument("driver is NULL");
    driver->pipe->setStream(stream);
  });
}

int gp_driver_load_stepg(gp_driver *driver, const gp_stepg *st) {
  return guarded([&] {
    if (!driver || !st) throw InvalidArgument("NULL argument");
    checkLevel(st->correlation_level);
    if (st->correlation_level != driver->level)
      throw InvalidArgument("the STEPG's correlation level differs from the driver's");
    gpa::Stepg &g = driver->parsed.stepg;
    g.level = st->correlation_level;
    g.k = gpa::nodesPerQubit(g.level);
    g.numLayers = st->num_layers;
    g.numQubits = st->num_qubits;
    const uint64_t npl = static_cast<uint64_t>(g.numQubits) * g.k, nn = npl * g.numLayers;
    if (nn >= (1ull << 31) || npl >= (1u << 30)) throw RuntimeError("STEPG too large: more than 2^31 nodes");
    g.npl = static_cast<uint32_t>(npl);
    g.numNodes = static_cast<uint32_t>(nn);
    g.numMeasurements = st->num_measurements;
    g.numDetectors = st->num_detectors;
    g.numObservables = st->num_observables;
    g.numLocations = 0;  // not part of a STEPG
    g.numCols = static_cast<uint32_t>((static_cast<uint64_t>(g.numDetectors) + g.numObservables + 31) / 32);
    const uint32_t N = g.numNodes, M = g.numMeasurements;
    const uint32_t W = (g.numDetectors + g.numObservables + 63) / 64;
    if (N && (!st->adjacency || !st->probs)) throw InvalidArgument("STEPG arrays are NULL");
    if (M && (!st->keys || (W && !st->sensitivity))) throw InvalidArgument("STEPG sensitivity map is NULL");
    // flows run forward by one or two layers (Mixer.hpp rules), keys name
    // nodes, probabilities lie in [0, 1]
    for (uint32_t n = 0; n < N; ++n) {
      const uint64_t a = st->adjacency[n];
      const uint32_t l = n / g.npl;
      for (uint32_t c : {static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32)}) {
        if (c == gpa::kNone) continue;
        const uint32_t cl = c < N ? c / g.npl : 0u;
        if (c >= N || cl <= l || cl > l + 2)
          throw InvalidArgument("STEPG node " + std::to_string(n) + " flows to " + std::to_string(c) +
                                ", not one or two layers later");
      }
      if (!(st->probs[n] >= 0.0 && st->probs[n] <= 1.0))
        throw InvalidArgument("STEPG node " + std::to_string(n) + " has probability outside [0, 1]");
    }
    g.adjacency.assign(st->adjacency, st->adjacency + N);
    g.probs.assign(st->probs, st->probs + N);
    g.keys.assign(st->keys, st->keys + M);
    for (uint32_t m = 0; m < M; ++m)
      if (g.keys[m] >= N) throw InvalidArgument("STEPG key of measurement " + std::to_string(m) + " names no node");
    // 64-bit words, column-major -> 32-bit columns per measurement
    g.measBits.assign(static_cast<size_t>(M) * g.numCols, 0u);
    for (uint32_t c = 0; c < g.numCols; ++c)
      for (uint32_t m = 0; m < M; ++m) {
        const uint64_t w = st->sensitivity[static_cast<size_t>(c / 2) * M + m];
        g.measBits[static_cast<size_t>(m) * g.numCols + c] = static_cast<uint32_t>(c % 2 ? w >> 32 : w);
      }
    // bits past the last observable are not targets
    const uint32_t T = g.numDetectors + g.numObservables;
    if (T % 32)
      for (uint32_t m = 0; m < M; ++m) g.measBits[static_cast<size_t>(m) * g.numCols + g.numCols - 1] &= (1u << (T % 32)) - 1;
    g.compNode.clear();
    g.compLoc.clear();
    g.compKind.clear();
    gpa::buildImage(g, driver->parsed.image, true);
    gpa::Circuit &c = driver->parsed.circuit;
    c = gpa::Circuit{};
    c.numQubits = g.numQubits;
    c.numTicks = g.numLayers ? g.numLayers - 1 : 0;
    c.numMeasurements = M;
    c.numDetectors = g.numDetectors;
    c.numObservables = g.numObservables;
    driver->haveCircuit = true;
    driver->loaded = false;
  });
}

int gp_driver_destroy(gp_driver *driver) {
  return guarded([&] { delete driver; });
}

int gp_driver_parse(gp_driver *driver, const char *circuit_text) {
  return guarded([&] {
    if (!driver) throw InvalidArgument("driver is NULL");
    requireText(circuit_text);
    parseInto(circuit_text, driver->level, driver->parsed);
    driver->haveCircuit = true;
    driver->loaded = false;
  });
}

int gp_compile_detector_error_model(gp_driver *driver, const char *circuit_text) {
  return guarded([&] {
    if (!driver) throw InvalidArgument("driver is NULL");
    const auto t0 = std::chrono::steady_clock::now();
    if (circuit_text) {
      parseInto(circuit_text, driver->level, driver->parsed);
      driver->haveCircuit = true;
      driver->loaded = false;
    }
    if (!driver->haveCircuit) throw RuntimeError("no circuit parsed");
    driver->haveResult = false;
    // upload the STEPG (Driver.hpp:145-152), run the stages, copy back; the
    // packed input arena of an unchanged circuit is sent again as it is
    if (driver->loaded) {
      driver->pipe->upload();
    } else {
      driver->pipe->load({&driver->parsed.image});
      driver->loaded = true;
    }
    driver->pipe->run(true);
    driver->pipe->download();
    driver->pipe->result(0, driver->result);
    driver->haveResult = true;
    driver->wallUs =
        std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  });
}

int gp_dem_shape(const gp_driver *driver, uint32_t *num_errors, uint64_t *num_targets,
                 uint32_t *num_detectors, uint32_t *num_observables) {
  return guarded([&] {
    if (!driver || !driver->haveResult) throw RuntimeError("no compiled detector error model");
    const auto &r = driver->result;
    if (num_errors) *num_errors = r.numErrors;
    if (num_targets) *num_targets = r.numTargets;
    if (num_detectors) *num_detectors = r.numDetectors;
    if (num_observables) *num_observables = r.numObservables;
  });
}

int gp_dem_copy(const gp_driver *driver, double *probs, uint32_t *offsets, uint32_t *targets,
                uint64_t *hashes) {
  return guarded([&] {
    if (!driver || !driver->haveResult) throw RuntimeError("no compiled detector error model");
    const auto &r = driver->result;
    if (probs) std::memcpy(probs, r.probs, 8ull * r.numErrors);
    if (offsets) std::memcpy(offsets, r.offsets, 4ull * (r.numErrors + 1));
    if (targets) r.copyTargets(targets);
    if (hashes) {
      if (r.hashes) std::memcpy(hashes, r.hashes, 8ull * r.numErrors);
      else driver->pipe->copyHashes(0, hashes);
    }
  });
}

int gp_dem_acquire(const gp_driver *driver, gp_dem_view *view) {
  return guarded([&] {
    if (!driver || !view) throw InvalidArgument("NULL argument");
    if (!driver->haveResult) throw RuntimeError("no compiled detector error model");
    const auto &r = driver->result;
    // a driver compiles one circuit, whose offsets are the block's own
    if (!r.offsetsRel.empty()) throw RuntimeError("result offsets are not in the result block");
    auto *h = new gpa::ResultHandle(driver->pipe->resultBlock());
    view->probs = r.probs;
    view->offsets = r.offsets;
    view->targets = r.targets16 ? static_cast<const void *>(r.targets16) : static_cast<const void *>(r.targets);
    view->target_bytes = r.targets16 ? 2u : 4u;
    view->hashes = r.hashes;
    view->num_errors = r.numErrors;
    view->num_detectors = r.numDetectors;
    view->num_observables = r.numObservables;
    view->num_targets = r.numTargets;
    view->token = h;
  });
}

int gp_dem_release(void *token) {
  return guarded([&] {
    if (!token) throw InvalidArgument("token is NULL");
    delete static_cast<gpa::ResultHandle *>(token);
  });
}

int gp_dem_text(const gp_driver *driver, char *buf, size_t cap, size_t *len) {
  return guarded([&] {
    if (!driver || !driver->haveResult) throw RuntimeError("no compiled detector error model");
    std::string s;
    formatDem(driver->result, s);
    if (len) *len = s.size();
    if (buf && cap > s.size()) std::memcpy(buf, s.c_str(), s.size() + 1);
  });
}

int gp_driver_params(const gp_driver *driver, uint32_t params[7]) {
  return guarded([&] {
    if (!driver || !params) throw InvalidArgument("NULL argument");
    const auto &c = driver->parsed.circuit;
    params[0] = c.numQubits;
    params[1] = c.numTicks + 1;
    params[2] = c.numMeasurements;
    params[3] = c.numDetectors;
    params[4] = c.numObservables;
    params[5] = driver->parsed.stepg.numNodes;
    params[6] = c.numLocations;
  });
}

int gp_driver_timings(const gp_driver *driver, double t[8]) {
  return guarded([&] {
    if (!driver || !t) throw InvalidArgument("NULL argument");
    const auto &s = driver->pipe->times();
    t[0] = s.upload;
    t[1] = s.sweep;
    t[2] = s.classify;
    t[3] = s.partition;
    t[4] = s.buckets;
    t[5] = s.emit;
    t[6] = s.download;
    t[7] = driver->wallUs;
  });
}

int gp_driver_bench_device(gp_driver *driver, uint32_t iterations, double *mean_us) {
  return guarded([&] {
    if (!driver || !mean_us) throw InvalidArgument("NULL argument");
    if (!driver->haveCircuit) throw RuntimeError("no circuit parsed");
    if (!driver->loaded) {
      driver->pipe->load({&driver->parsed.image});
      driver->loaded = true;
    }
    *mean_us = driver->pipe->benchDevice(iterations);
  });
}

int gp_debug_pnodes(const gp_driver *driver, uint32_t *num_pnodes, uint32_t *words64) {
  return guarded([&] {
    if (!driver) throw InvalidArgument("driver is NULL");
    const auto &img = driver->parsed.image;
    if (num_pnodes) *num_pnodes = img.numRanks;
    if (words64) *words64 = (img.numDetectors + img.numObservables + 63) / 64;
  });
}

int gp_debug_sensitivities(const gp_driver *driver, uint32_t *pnode_index, uint64_t *sens) {
  return guarded([&] {
    if (!driver || !driver->loaded) throw RuntimeError("no compiled circuit");
    const auto &img = driver->parsed.image;
    if (pnode_index) std::memcpy(pnode_index, img.rankNode.data(), 4ull * img.numRanks);
    if (sens) {
      std::vector<uint32_t> cols;
      driver->pipe->debugSensitivities(0, cols);
      const uint32_t C = img.numCols, W = (img.numDetectors + img.numObservables + 63) / 64;
      for (uint32_t i = 0; i < img.numRanks; ++i)
        for (uint32_t w = 0; w < W; ++w) {
          const uint64_t lo = 2 * w < C ? cols[static_cast<size_t>(i) * C + 2 * w] : 0u;
          const uint64_t hi = 2 * w + 1 < C ? cols[static_cast<size_t>(i) * C + 2 * w + 1] : 0u;
          sens[static_cast<size_t>(i) * W + w] = lo | (hi << 32);
        }
    }
  });
}

int gp_debug_stepg(const gp_driver *driver, uint64_t *adjacency, double *probs) {
  return guarded([&] {
    if (!driver) throw InvalidArgument("driver is NULL");
    const auto &g = driver->parsed.stepg;
    if (adjacency) std::memcpy(adjacency, g.adjacency.data(), 8ull * g.numNodes);
    if (probs) std::memcpy(probs, g.probs.data(), 8ull * g.numNodes);
  });
}

int gp_debug_sweep_plan(const gp_driver *driver, uint32_t plan[3]) {
  return guarded([&] {
    if (!driver || !plan) throw InvalidArgument("NULL argument");
    if (!driver->pipe) throw RuntimeError("driver has not compiled a circuit");
    driver->pipe->sweepPlan(plan);
  });
}

int gp_build_stepg_host(const char *circuit_text, int correlation_level, uint32_t dims[4],
                        uint64_t *adjacency, double *probs) {
  return guarded([&] {
    requireText(circuit_text);
    checkLevel(correlation_level);
    if (!dims) throw InvalidArgument("dims is NULL");
    gpa::Circuit c = gpa::Circuit::parse(circuit_text);
    gpa::Stepg g;
    gpa::buildStepg(c, correlation_level, g, nullptr, false, true);
    dims[0] = g.numLayers;
    dims[1] = g.npl;
    dims[2] = g.numNodes;
    dims[3] = g.numLocations;
    if (adjacency) std::memcpy(adjacency, g.adjacency.data(), 8ull * g.numNodes);
    if (probs) std::memcpy(probs, g.probs.data(), 8ull * g.numNodes);
  });
}

int gp_debug_image_digest(const char *circuit_text, int correlation_level, int parallel, uint64_t digest[2]) {
  return guarded([&] {
    requireText(circuit_text);
    checkLevel(correlation_level);
    if (!digest) throw InvalidArgument("digest is NULL");
    gpa::Circuit c = gpa::Circuit::parse(circuit_text);
    gpa::Stepg g;
    gpa::StepgImage img;
    gpa::buildStepg(c, correlation_level, g, nullptr, false, parallel != 0);
    gpa::buildImage(g, img, parallel != 0);
    // FNV-1a-64 over the STEPG's and the image's arrays and sizes
    uint64_t h = 0xcbf29ce484222325ull, n = 0;
    auto mix = [&](const void *p, size_t bytes) {
      const unsigned char *b = static_cast<const unsigned char *>(p);
      for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
      n += bytes;
    };
    auto vec = [&](const auto &v) {
      const uint64_t sz = v.size();
      mix(&sz, 8);
      if (sz) mix(v.data(), sz * sizeof(v[0]));
    };
    vec(g.adjacency), vec(g.probs), vec(g.keys), vec(g.measBits);
    const uint32_t dims[] = {img.numLayers, img.npl, img.numCols, img.numRanks, img.numSeeds, img.numActive,
                             img.maxActive, img.sweepThreads, img.sweepRpt, img.layerStride, img.rowSpan,
                             img.numPairs};
    mix(dims, sizeof dims);
    vec(img.layerAct), vec(img.act), vec(img.seedVals), vec(img.colHi), vec(img.colLow), vec(img.rankLayer);
    vec(img.rankProb), vec(img.rankNode), vec(img.rankCode), vec(img.rankSpill), vec(img.probTab), vec(img.probIdx);
    digest[0] = h;
    digest[1] = n;
  });
}

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------

int gp_batch_create(int correlation_level, int device, gp_batch **out) {
  return guarded([&] {
    if (!out) throw InvalidArgument("out is NULL");
    checkLevel(correlation_level);
    auto b = std::make_unique<gp_batch>();
    b->level = correlation_level;
    b->device = device;
    b->pipe = std::make_unique<gpa::Pipeline>(device, true);
    *out = b.release();
  });
}

int gp_batch_destroy(gp_batch *batch) {
  return guarded([&] { delete batch; });
}

int gp_batch_set_circuits(gp_batch *batch, const char *const *texts, uint32_t count, uint32_t threads) {
  return guarded([&] {
    if (!batch || (!texts && count)) throw InvalidArgument("NULL argument");
    for (uint32_t i = 0; i < count; ++i) requireText(texts[i]);
    std::vector<std::shared_ptr<const gpa::StepgImage>> imgs(count);
    uint32_t nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<uint32_t>(nt, std::max<uint32_t>(count, 1));
    std::vector<std::string> errs(count);
    std::vector<int> kinds(count, 0);
    std::atomic<uint32_t> next{0};
    auto work = [&] {
      for (uint32_t i; (i = next.fetch_add(1)) < count;) {
        try {
          Parsed p;
          parseInto(texts[i], batch->level, p, false, nullptr, nt == 1);
          imgs[i] = std::make_shared<const gpa::StepgImage>(std::move(p.image));
        } catch (const std::invalid_argument &e) {
          errs[i] = e.what();
          kinds[i] = 1;
        } catch (const std::exception &e) {
          errs[i] = e.what();
          kinds[i] = 2;
        }
      }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    for (uint32_t i = 0; i < count; ++i) {
      if (kinds[i] == 1) throw InvalidArgument("circuit " + std::to_string(i) + ": " + errs[i]);
      if (kinds[i] == 2) throw RuntimeError("circuit " + std::to_string(i) + ": " + errs[i]);
    }
    batch->images = std::move(imgs);
    batch->results.clear();
    batch->loaded = false;
  });
}

int gp_prepare_circuit(const char *circuit_text, int correlation_level, gp_prepared **out) {
  return guarded([&] {
    requireText(circuit_text);
    checkLevel(correlation_level);
    if (!out) throw InvalidArgument("out is NULL");
    Parsed p;
    parseInto(circuit_text, correlation_level, p);
    auto h = std::make_unique<gp_prepared>();
    h->level = correlation_level;
    h->image = std::make_shared<const gpa::StepgImage>(std::move(p.image));
    *out = h.release();
  });
}

int gp_prepared_destroy(gp_prepared *prepared) {
  return guarded([&] { delete prepared; });
}

int gp_prepared_device_bytes(const gp_prepared *prepared, uint64_t *bytes) {
  return guarded([&] {
    if (!prepared || !bytes) throw InvalidArgument("NULL argument");
    *bytes = gpa::Pipeline::deviceBytesFor({prepared->image.get()}, true);
  });
}

// The images of `count` prepared circuits of the batch's level.
static std::vector<const gpa::StepgImage *> preparedImages(const gp_batch *batch, const gp_prepared *const *prepared,
                                                           uint32_t count) {
  if (!prepared && count) throw InvalidArgument("prepared is NULL");
  std::vector<const gpa::StepgImage *> imgs(count);
  for (uint32_t i = 0; i < count; ++i) {
    if (!prepared[i]) throw InvalidArgument("prepared circuit " + std::to_string(i) + " is NULL");
    if (prepared[i]->level != batch->level)
      throw InvalidArgument("prepared circuit " + std::to_string(i) + " was built at correlation level " +
                            std::to_string(prepared[i]->level) + ", the batch compiles at level " +
                            std::to_string(batch->level));
    imgs[i] = prepared[i]->image.get();
  }
  return imgs;
}

int gp_batch_set_prepared(gp_batch *batch, const gp_prepared *const *prepared, uint32_t count) {
  return guarded([&] {
    if (!batch) throw InvalidArgument("batch is NULL");
    preparedImages(batch, prepared, count);
    std::vector<std::shared_ptr<const gpa::StepgImage>> imgs(count);
    for (uint32_t i = 0; i < count; ++i) imgs[i] = prepared[i]->image;
    batch->images = std::move(imgs);
    batch->results.clear();
    batch->loaded = false;
  });
}

int gp_batch_bytes_after(const gp_batch *batch, const gp_prepared *const *prepared, uint32_t count,
                         uint64_t *bytes) {
  return guarded([&] {
    if (!batch || !bytes) throw InvalidArgument("NULL argument");
    if (!count) throw InvalidArgument("no circuits");
    *bytes = batch->pipe->deviceBytesAfter(preparedImages(batch, prepared, count)) +
             (batch->single ? batch->single->deviceBytes() : 0);
  });
}

int gp_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
  return guarded([&] {
    if (!free_bytes || !total_bytes) throw InvalidArgument("NULL argument");
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) throw RuntimeError("no HIP device");
    if (hipSetDevice(device) != hipSuccess) throw InvalidArgument("invalid device " + std::to_string(device));
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) throw RuntimeError(std::string("hipMemGetInfo: ") + hipGetErrorString(e));
    *free_bytes = f;
    *total_bytes = t;
  });
}

// Device microseconds of one replay of the loaded pipeline's graph (the
// first call after a load captures it; that capture is not timed).
static double timeGraph(gpa::Pipeline &p) {
  hipStream_t st = static_cast<hipStream_t>(p.stream());
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    throw RuntimeError("hipEventCreate failed");
  p.runGraph();  // capture + warm
  (void)hipEventRecord(a, st);
  p.runGraph();
  (void)hipEventRecord(b, st);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms * 1e3;
}

static void keepResult(gp_batch *b, uint32_t i, const gpa::VariantResult &r) {
  b->ownP[i].assign(r.probs, r.probs + r.numErrors);
  b->ownOff[i].assign(r.offsets, r.offsets + r.numErrors + 1);
  b->ownTg[i].resize(r.numTargets);
  r.copyTargets(b->ownTg[i].data());
  b->ownHash[i].assign(r.hashes, r.hashes + r.numErrors);
  gpa::VariantResult &k = b->results[i];
  k.numErrors = r.numErrors;
  k.numTargets = r.numTargets;
  k.numDetectors = r.numDetectors;
  k.numObservables = r.numObservables;
  k.probs = b->ownP[i].data();
  k.offsets = b->ownOff[i].data();
  k.targets = b->ownTg[i].data();
  k.hashes = b->ownHash[i].data();
}

int gp_batch_compile(gp_batch *batch, int mode) {
  return guarded([&] {
    if (!batch) throw InvalidArgument("batch is NULL");
    const uint32_t n = static_cast<uint32_t>(batch->images.size());
    if (!n) throw RuntimeError("batch has no circuits");
    batch->results.assign(n, gpa::VariantResult{});
    batch->ownP.assign(n, {});
    batch->ownOff.assign(n, {});
    batch->ownTg.assign(n, {});
    batch->ownHash.assign(n, {});
    if (mode == 0) {
      if (!batch->loaded) {
        // a new batch (the serving case): one direct run, timed by its
        // stage events -- capturing a graph would cost more than it saves
        std::vector<const gpa::StepgImage *> imgs;
        for (auto &c : batch->images) imgs.push_back(c.get());
        batch->pipe->load(imgs);
        batch->loaded = true;
        batch->pipe->run(true);
        batch->pipe->download();
        const gpa::StageTimes &t = batch->pipe->times();
        batch->deviceUs = t.sweep + t.classify + t.buckets + t.emit;
      } else {
        // the same batch again: device time of one replay of the captured
        // pipeline (captured on this first repeat)
        batch->deviceUs = timeGraph(*batch->pipe);
        batch->pipe->download();
      }
      for (uint32_t i = 0; i < n; ++i) {
        gpa::VariantResult r;
        batch->pipe->result(i, r);
        keepResult(batch, i, r);
      }
    } else if (mode == 1) {
      if (!batch->single) batch->single = std::make_unique<gpa::Pipeline>(batch->device, true);
      double total = 0;
      for (uint32_t i = 0; i < n; ++i) {
        batch->single->load({batch->images[i].get()});
        // one untimed compile first, as the batched timing follows earlier
        // runs of its load (a load's first run sizes the target rows)
        batch->single->run(false);
        batch->single->download();
        total += timeGraph(*batch->single);
        batch->single->download();
        gpa::VariantResult r;
        batch->single->result(0, r);
        keepResult(batch, i, r);
      }
      batch->deviceUs = total;
    } else {
      throw InvalidArgument("mode must be 0 (batched) or 1 (one by one)");
    }
  });
}

int gp_batch_stage_us(gp_batch *batch, double stage_us[4], uint64_t *sorted_items) {
  return guarded([&] {
    if (!batch || !stage_us) throw InvalidArgument("NULL argument");
    if (!batch->loaded) throw RuntimeError("no batched compile to time (gp_batch_compile mode 0 first)");
    batch->pipe->run(true);
    batch->pipe->download();
    const gpa::StageTimes &t = batch->pipe->times();
    stage_us[0] = t.sweep;
    stage_us[1] = t.classify;
    stage_us[2] = t.buckets;
    stage_us[3] = t.emit;
    if (sorted_items) *sorted_items = batch->pipe->sortedItems();
  });
}

int gp_batch_device_us(const gp_batch *batch, double *us) {
  return guarded([&] {
    if (!batch || !us) throw InvalidArgument("NULL argument");
    *us = batch->deviceUs;
  });
}

int gp_batch_dem_shape(const gp_batch *batch, uint32_t index, uint32_t *num_errors,
                       uint64_t *num_targets, uint32_t *num_detectors, uint32_t *num_observables) {
  return guarded([&] {
    if (!batch || index >= batch->results.size()) throw InvalidArgument("index out of range");
    const auto &r = batch->results[index];
    if (num_errors) *num_errors = r.numErrors;
    if (num_targets) *num_targets = r.numTargets;
    if (num_detectors) *num_detectors = r.numDetectors;
    if (num_observables) *num_observables = r.numObservables;
  });
}

int gp_batch_dem_copy(const gp_batch *batch, uint32_t index, double *probs, uint32_t *offsets,
                      uint32_t *targets, uint64_t *hashes) {
  return guarded([&] {
    if (!batch || index >= batch->results.size()) throw InvalidArgument("index out of range");
    const auto &r = batch->results[index];
    if (probs) std::memcpy(probs, r.probs, 8ull * r.numErrors);
    if (offsets) std::memcpy(offsets, r.offsets, 4ull * (r.numErrors + 1));
    if (targets) r.copyTargets(targets);
    if (hashes) std::memcpy(hashes, r.hashes, 8ull * r.numErrors);
  });
}

int gp_batch_device_bytes(const gp_batch *batch, uint64_t *bytes) {
  return guarded([&] {
    if (!batch || !bytes) throw InvalidArgument("NULL argument");
    *bytes = batch->pipe->deviceBytes() + (batch->single ? batch->single->deviceBytes() : 0);
  });
}

int gp_batch_layout_stats(const gp_batch *batch, uint64_t stats[4]) {
  return guarded([&] {
    if (!batch || !stats) throw InvalidArgument("NULL argument");
    stats[0] = batch->pipe->sWords();
    stats[1] = batch->pipe->bandRelayouts();
    stats[2] = batch->pipe->numBuckets();
    stats[3] = batch->pipe->bucketCap();
  });
}

int gp_set_tuning_class(int correlation_level, uint32_t size_class, const char *name, uint32_t value) {
  return guarded([&] {
    checkLevel(correlation_level);
    if (!name) throw InvalidArgument("name is NULL");
    gpa::setClassTuning(correlation_level, size_class, name, value);
  });
}

int gp_get_tuning_class(int correlation_level, uint32_t size_class, const char *name, uint32_t *value) {
  return guarded([&] {
    checkLevel(correlation_level);
    if (!name || !value) throw InvalidArgument("NULL argument");
    if (size_class > 63) throw InvalidArgument("size class must be at most 63");
    *value = gpa::classTuning(name, correlation_level, uint64_t{1} << size_class, 0u);
  });
}

int gp_clear_tuning_classes(void) {
  return guarded([&] { gpa::clearClassTuning(); });
}

int gp_circuit_summary(const char *circuit_text, uint64_t summary[4]) {
  return guarded([&] {
    requireText(circuit_text);
    if (!summary) throw InvalidArgument("summary is NULL");
    const gpa::CircuitSummary s = gpa::summarize(circuit_text);
    summary[0] = s.layers;
    summary[1] = s.qubits;
    summary[2] = s.detectors;
    summary[3] = s.observables;
  });
}

int gp_estimate_device_bytes(const char *circuit_text, int correlation_level, uint64_t *bytes) {
  return guarded([&] {
    requireText(circuit_text);
    checkLevel(correlation_level);
    if (!bytes) throw InvalidArgument("bytes is NULL");
    Parsed p;
    parseInto(circuit_text, correlation_level, p);
    *bytes = gpa::Pipeline::deviceBytesFor({&p.image}, true);
  });
}

// ---------------------------------------------------------------------------
// differentiable plan
// ---------------------------------------------------------------------------

int gp_grad_plan_create(const char *circuit_text, int correlation_level, int device,
                        gp_grad_plan **out) {
  return guarded([&] {
    if (!out) throw InvalidArgument("out is NULL");
    requireText(circuit_text);
    checkLevel(correlation_level);
    auto g = std::make_unique<gp_grad_plan>();
    g->device = device;
    // Structure only: every location gets the same positive rate so each
    // node that carries a noise component is a p>0 node of the plan.
    gpa::Circuit c = gpa::Circuit::parse(circuit_text);
    std::vector<double> unit(c.numLocations, 1e-3);
    g->parsed.circuit = std::move(c);
    gpa::buildStepg(g->parsed.circuit, correlation_level, g->parsed.stepg, unit.data(), true);
    gpa::buildImage(g->parsed.stepg, g->parsed.image);
    g->pipe = std::make_unique<gpa::Pipeline>(device, true);
    g->pipe->load({&g->parsed.image});
    g->pipe->run(false);
    g->pipe->download();
    g->pipe->result(0, g->result);

    const auto &st = g->parsed.stepg;
    const auto &img = g->parsed.image;
    g->numLoc = st.numLocations;
    g->numComp = static_cast<uint32_t>(st.compNode.size());
    g->numNodes = img.numRanks;
    g->numErrors = g->result.numErrors;
    // node -> rank, components grouped by rank (stable in circuit order)
    std::vector<uint32_t> rankOf(st.numNodes, gpa::kNone);
    for (uint32_t i = 0; i < img.numRanks; ++i) rankOf[img.rankNode[i]] = i;
    std::vector<uint32_t> nodeStart(img.numRanks + 1, 0), nodeComp(g->numComp);
    for (uint32_t k = 0; k < g->numComp; ++k) {
      const uint32_t r = rankOf[st.compNode[k]];
      if (r == gpa::kNone) throw RuntimeError("noise component on a node outside the plan");
      nodeStart[r + 1]++;
    }
    for (uint32_t i = 0; i < img.numRanks; ++i) nodeStart[i + 1] += nodeStart[i];
    std::vector<uint32_t> cur(nodeStart.begin(), nodeStart.end() - 1);
    for (uint32_t k = 0; k < g->numComp; ++k) nodeComp[cur[rankOf[st.compNode[k]]]++] = k;

    if (hipSetDevice(device) != hipSuccess) throw RuntimeError("hipSetDevice failed");
    auto up = [&](auto **dst, const auto *src, size_t count) {
      using T = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
      if (hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
        throw RuntimeError("hipMalloc failed");
      if (count && hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        throw RuntimeError("hipMemcpy failed");
    };
    up(&g->dCompLoc, st.compLoc.data(), g->numComp);
    up(&g->dCompKind, st.compKind.data(), g->numComp);
    up(&g->dNodeStart, nodeStart.data(), nodeStart.size());
    up(&g->dNodeComp, nodeComp.data(), nodeComp.size());
    // class members: the pipeline's bucket-sorted ranks, grouped by class
    if (hipMalloc(reinterpret_cast<void **>(&g->dClassStart), 4ull * (g->numErrors + 1)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&g->dClassNode), 4ull * std::max<uint32_t>(img.numRanks, 1)) !=
            hipSuccess)
      throw RuntimeError("hipMalloc failed");
    if (hipMemcpy(g->dClassStart, g->pipe->devClassMember(), 4ull * (g->numErrors + 1),
                  hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(g->dClassNode, g->pipe->devSortedRanks(), 4ull * img.numRanks,
                  hipMemcpyDeviceToDevice) != hipSuccess)
      throw RuntimeError("hipMemcpy failed");
    *out = g.release();
  });
}

int gp_grad_plan_destroy(gp_grad_plan *plan) {
  return guarded([&] { delete plan; });
}

int gp_grad_plan_sizes(const gp_grad_plan *plan, uint32_t sizes[4]) {
  return guarded([&] {
    if (!plan || !sizes) throw InvalidArgument("NULL argument");
    sizes[0] = plan->numLoc;
    sizes[1] = plan->numComp;
    sizes[2] = plan->numNodes;
    sizes[3] = plan->numErrors;
  });
}

int gp_grad_plan_dem_structure(const gp_grad_plan *plan, uint32_t *offsets, uint32_t *targets) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("plan is NULL");
    const auto &r = plan->result;
    if (offsets) std::memcpy(offsets, r.offsets, 4ull * (r.numErrors + 1));
    if (targets) r.copyTargets(targets);
  });
}

int gp_grad_plan_device_arrays(const gp_grad_plan *plan, const uint32_t **comp_loc,
                               const uint8_t **comp_kind, const uint32_t **node_start,
                               const uint32_t **node_comp, const uint32_t **class_start,
                               const uint32_t **class_node) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("plan is NULL");
    if (comp_loc) *comp_loc = plan->dCompLoc;
    if (comp_kind) *comp_kind = plan->dCompKind;
    if (node_start) *node_start = plan->dNodeStart;
    if (node_comp) *node_comp = plan->dNodeComp;
    if (class_start) *class_start = plan->dClassStart;
    if (class_node) *class_node = plan->dClassNode;
  });
}

int gp_grad_plan_copy_arrays(const gp_grad_plan *plan, uint32_t *comp_loc, uint8_t *comp_kind,
                             uint32_t *node_start, uint32_t *node_comp, uint32_t *class_start,
                             uint32_t *class_node) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("plan is NULL");
    if (hipSetDevice(plan->device) != hipSuccess) throw RuntimeError("hipSetDevice failed");
    auto down = [](void *dst, const void *src, size_t bytes) {
      if (dst && bytes && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        throw RuntimeError("hipMemcpy failed");
    };
    down(comp_loc, plan->dCompLoc, 4ull * plan->numComp);
    down(comp_kind, plan->dCompKind, plan->numComp);
    down(node_start, plan->dNodeStart, 4ull * (plan->numNodes + 1));
    down(node_comp, plan->dNodeComp, 4ull * plan->numComp);
    down(class_start, plan->dClassStart, 4ull * (plan->numErrors + 1));
    down(class_node, plan->dClassNode, 4ull * plan->numNodes);
  });
}

int gp_set_tuning(const char *name, uint32_t value) {
  return guarded([&] {
    if (!name) throw InvalidArgument("name is NULL");
    if (std::string(name) == "bucket_fill") gpa::setBucketFill(value);
    else if (std::string(name) == "sweep_rpt") gpa::setSweepMinRpt(value);
    else if (std::string(name) == "stash") gpa::setStash(value != 0);
    else if (std::string(name) == "mask_table") gpa::setMaskTable(value != 0);
    else if (std::string(name) == "prob_table") gpa::setProbTable(value != 0);
    else if (std::string(name) == "packed_sort") gpa::setPackedSort(value != 0);
    else if (std::string(name) == "s_bands") gpa::setSBands(value);
    else if (std::string(name) == "bucket_cap") gpa::setBucketCap(value);
    else if (std::string(name) == "target_fill") gpa::setTargetFill(value);
    else if (std::string(name) == "compact_items") gpa::setCompactItems(value);
    else if (std::string(name) == "overflow_start") gpa::setOverflowStart(value);
    else if (std::string(name) == "sweep_split") {
      if (value > 2) throw InvalidArgument("sweep_split must be 0, 1 or 2");
      gpa::setSweepSplit(value);
    }
    else throw InvalidArgument(std::string("unknown tuning parameter: ") + name);
  });
}

int gp_get_tuning(const char *name, uint32_t *value) {
  return guarded([&] {
    if (!name || !value) throw InvalidArgument("NULL argument");
    if (std::string(name) == "bucket_fill") *value = gpa::bucketFill();
    else if (std::string(name) == "sweep_rpt") *value = gpa::sweepMinRpt();
    else if (std::string(name) == "stash") *value = gpa::stash() ? 1u : 0u;
    else if (std::string(name) == "mask_table") *value = gpa::maskTable() ? 1u : 0u;
    else if (std::string(name) == "prob_table") *value = gpa::probTable() ? 1u : 0u;
    else if (std::string(name) == "packed_sort") *value = gpa::packedSort() ? 1u : 0u;
    else if (std::string(name) == "s_bands") *value = gpa::sBands();
    else if (std::string(name) == "bucket_cap") *value = gpa::bucketCap();
    else if (std::string(name) == "target_fill") *value = gpa::targetFill();
    else if (std::string(name) == "compact_items") *value = gpa::compactItems();
    else if (std::string(name) == "overflow_start") *value = gpa::overflowStart();
    else if (std::string(name) == "sweep_split") *value = gpa::sweepSplit();
    else throw InvalidArgument(std::string("unknown tuning parameter: ") + name);
  });
}

int gp_xor_merge_forward(const double *values, const uint32_t *index, const uint32_t *seg_start,
                         uint32_t num_segments, double *out, void *stream) {
  return guarded([&] {
    if (num_segments && (!values || !index || !seg_start || !out)) throw InvalidArgument("NULL argument");
    gpa::xorMergeForward(values, index, seg_start, num_segments, out, stream);
  });
}

int gp_node_rates_forward(const double *theta, const uint32_t *comp_loc, const uint8_t *comp_kind,
                          const uint32_t *node_start, const uint32_t *node_comp, uint32_t num_nodes, double *q,
                          void *stream) {
  return guarded([&] {
    if (num_nodes && (!theta || !comp_loc || !comp_kind || !node_start || !node_comp || !q))
      throw InvalidArgument("NULL argument");
    gpa::nodeRatesForward(theta, comp_loc, comp_kind, node_start, node_comp, num_nodes, q, stream);
  });
}

int gp_node_rates_backward(const double *theta, const uint32_t *comp_loc, const uint8_t *comp_kind,
                           const uint32_t *node_start, const uint32_t *node_comp, uint32_t num_nodes,
                           const double *grad_q, double *grad_comp, const uint32_t *loc_start,
                           const uint32_t *loc_comp, uint32_t num_locations, double *grad_theta, void *stream) {
  return guarded([&] {
    if ((num_nodes && (!theta || !comp_loc || !comp_kind || !node_start || !node_comp || !grad_q || !grad_comp)) ||
        (num_locations && (!loc_start || !loc_comp || !grad_comp || !grad_theta)))
      throw InvalidArgument("NULL argument");
    gpa::nodeRatesBackward(theta, comp_loc, comp_kind, node_start, node_comp, num_nodes, grad_q, grad_comp,
                           loc_start, loc_comp, num_locations, grad_theta, stream);
  });
}

static_assert(gpa::kNoAllocMaxComponents == GP_NOALLOC_MAX_COMPONENTS, "header and engine disagree");

int gp_node_rates_backward_noalloc(const double *theta, const uint32_t *comp_loc, const uint8_t *comp_kind,
                                   const uint32_t *node_start, const uint32_t *node_comp, uint32_t num_nodes,
                                   uint32_t num_components, const double *grad_q, const uint32_t *loc_start,
                                   const uint32_t *loc_comp, uint32_t num_locations, double *grad_theta,
                                   void *stream) {
  return guarded([&] {
    if ((num_nodes && (!theta || !comp_loc || !comp_kind || !node_start || !node_comp || !grad_q)) ||
        (num_locations && (!loc_start || !loc_comp || !grad_theta)))
      throw InvalidArgument("NULL argument");
    gpa::nodeRatesBackwardNoAlloc(theta, comp_loc, comp_kind, node_start, node_comp, num_nodes, num_components,
                                  grad_q, loc_start, loc_comp, num_locations, grad_theta, stream);
  });
}

int gp_xor_merge_backward(const double *values, const uint32_t *index, const uint32_t *seg_start,
                          uint32_t num_segments, const double *grad_out, double *grad_values,
                          void *stream) {
  return guarded([&] {
    if (num_segments && (!values || !index || !seg_start || !grad_out || !grad_values))
      throw InvalidArgument("NULL argument");
    gpa::xorMergeBackward(values, index, seg_start, num_segments, grad_out, grad_values, stream);
  });
}

}  // extern "C"
