/*
 * greenpeas_amd.h -- C-ABI of the MI355X detector-error-model (DEM) compiler.
 *
 * Plain C types only (no torch, no C++ types).  Every entry point returns a
 * status code: GP_OK, GP_INVALID_ARGUMENT (the reference's
 * std::invalid_argument -> Python ValueError) or GP_RUNTIME_ERROR (the
 * reference's std::runtime_error -> Python RuntimeError); gp_last_error()
 * then holds the message for the calling thread.
 *
 * The drop-in surface mirrors the reference's pybind module
 * (source/QEC/ErrorAnalysis/ErrorAnalysis.pybind.cu:47-79) and its C++
 * Driver (include/GreenPeas/QEC/ErrorAnalysis/Driver.hpp:37-219): a driver is
 * sized once from a "max" circuit and then compiles any circuit that fits,
 * reusing its device buffers.
 */
#ifndef GREENPEAS_AMD_H
#define GREENPEAS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_OK 0
#define GP_INVALID_ARGUMENT 1
#define GP_RUNTIME_ERROR 2

/* CorrelationLevel (Mixer.hpp:13-20; pybind enum ErrorAnalysis.pybind.cu:48-52) */
#define GP_L0 0
#define GP_L1 1
#define GP_L2 2

typedef struct gp_driver gp_driver;
typedef struct gp_batch gp_batch;
typedef struct gp_grad_plan gp_grad_plan;

/* Message of the last failed call on this thread ("" if none). */
const char *gp_last_error(void);

/* Library version string. */
const char *gp_version(void);

/* ------------------------------------------------------------------------
 * Drop-in driver
 * --------------------------------------------------------------------- */

/* get_driver(circuit, correlation_level=L2)
 * Replaces ErrorAnalysis.pybind.cu:72-78 / getDriver (:26-41) and
 * Driver::fromStimCircuit (Driver.hpp:216-218): parses `circuit_text` (stim
 * circuit text), builds its STEPG and allocates device buffers sized for it
 * on HIP device `device`.  Later compiles must fit within these sizes. */
int gp_get_driver(const char *circuit_text, int correlation_level, int device,
                  gp_driver **out);

/* get_driver with the driver's work queued on a caller's stream (a
 * hipStream_t of `device`; NULL: a stream of the driver's own), as
 * gp_get_driver otherwise.  No counterpart in the reference, which works on
 * the legacy default stream (CUDA.hpp:161-163). */
int gp_get_driver_on_stream(const char *circuit_text, int correlation_level, int device,
                            void *stream, gp_driver **out);

/* Queues the driver's later work on `stream` (NULL: its own stream again),
 * after the work queued so far has completed.  A driver is used by one
 * thread at a time. */
int gp_driver_set_stream(gp_driver *driver, void *stream);

/* A space-time error-propagation graph built by the caller, in the
 * reference's layout (Circuit::stepg and Circuit::sensitivityMap,
 * Circuit.hpp:200-231; STEPG.hpp:40-132): node = layer * num_qubits * k +
 * qubit * k + slot with k = 3, 5, 8 nodes per qubit at L0, L1, L2
 * (Mixer.hpp:59-84). */
typedef struct gp_stepg {
  int correlation_level;
  uint32_t num_qubits, num_layers, num_measurements, num_detectors, num_observables;
  /* per node: child0 | child1 << 32, 0xFFFFFFFF = none; children lie one or
   * two layers later */
  const uint64_t *adjacency;
  const double *probs;           /* per node, in [0, 1] */
  const uint32_t *keys;          /* per measurement: its node */
  /* per measurement, ceil((num_detectors + num_observables) / 64) 64-bit
   * words, column-major (word w of measurement m at
   * sensitivity[w * num_measurements + m]): bit t = detector t, or
   * observable t - num_detectors */
  const uint64_t *sensitivity;
} gp_stepg;

/* Loads a caller's STEPG into the driver in place of parsing circuit text
 * (the reference's Driver state before compileDetectorErrorModel(),
 * Driver.hpp:208-211: the span it times, Cli.hpp:293-302).  The next
 * gp_compile_detector_error_model(driver, NULL) compiles it; capacity rules
 * are a parsed circuit's.  GP_INVALID_ARGUMENT for a level other than the
 * driver's, flows that do not run one or two layers forward, keys that name
 * no node or probabilities outside [0, 1]; GP_RUNTIME_ERROR where the graph
 * breaks the structure the sweep relies on (as for a parsed circuit). */
int gp_driver_load_stepg(gp_driver *driver, const gp_stepg *stepg);

/* Releases the driver and all of its host and device memory. */
int gp_driver_destroy(gp_driver *driver);

/* Driver.compile_detector_error_model(circuit)
 * Replaces ErrorAnalysis.pybind.cu:56-66 and
 * Driver::compileDetectorErrorModel(const stim::Circuit&) (Driver.hpp:199-204):
 * parses the circuit into the driver (throws GP_RUNTIME_ERROR if it exceeds
 * the driver's capacity, like fitto at Vector.hpp:118-123), compiles on the
 * GPU and keeps the resulting DEM in the driver until the next compile.
 * With circuit_text == NULL it recompiles the circuit parsed last
 * (ErrorAnalysis.pybind.cu:67-70, Driver.hpp:208-211). */
int gp_compile_detector_error_model(gp_driver *driver, const char *circuit_text);

/* Parses `circuit_text` into the driver without compiling
 * (Circuit::parseFromStimCircuit, Circuit.hpp:471-478). */
int gp_driver_parse(gp_driver *driver, const char *circuit_text);

/* Shape of the last compiled DEM (Driver::getDetectorErrorModel,
 * Driver.hpp:166-194): errors in the reference's order (ascending FNV-1a-64
 * class hash), total number of targets, detector and observable counts. */
int gp_dem_shape(const gp_driver *driver, uint32_t *num_errors,
                 uint64_t *num_targets, uint32_t *num_detectors,
                 uint32_t *num_observables);

/* Copies the last DEM as CSR: probs[num_errors], offsets[num_errors + 1],
 * targets[num_targets] where a target t < num_detectors is detector t and
 * otherwise observable t - num_detectors.  hashes (optional, may be NULL)
 * receives each error's 64-bit class hash (read from the device: a
 * compile downloads only the DEM itself). */
int gp_dem_copy(const gp_driver *driver, double *probs, uint32_t *offsets,
                uint32_t *targets, uint64_t *hashes);

/* The last DEM's arrays in place, without a copy (the pinned host memory
 * the driver downloaded them into; no counterpart in the reference, whose
 * DEM is returned by value).  The arrays stay valid until
 * gp_dem_release(view->token), also across later compiles and after the
 * driver is destroyed; the driver downloads later DEMs elsewhere while a
 * view is held.  Layout as gp_dem_copy, except that targets are 16-bit
 * (target_bytes = 2) when a circuit has at most 65536 detectors and
 * observables. */
typedef struct gp_dem_view {
  const double *probs;       /* num_errors */
  const uint32_t *offsets;   /* num_errors + 1 */
  const void *targets;       /* num_targets, target_bytes (2 or 4) each */
  const uint64_t *hashes;    /* NULL: not downloaded (gp_dem_copy reads them) */
  uint32_t num_errors, num_detectors, num_observables, target_bytes;
  uint64_t num_targets;
  void *token;
} gp_dem_view;
int gp_dem_acquire(const gp_driver *driver, gp_dem_view *view);
/* Lets go of a view's arrays (token from gp_dem_acquire; thread-safe). */
int gp_dem_release(void *token);

/* Writes the last DEM as stim DEM text ("error(p) D.. L..\n" per error).
 * If buf is NULL or too small, only *len (bytes, excluding the NUL) is set
 * and GP_OK is returned; call again with a buffer of *len + 1 bytes. */
int gp_dem_text(const gp_driver *driver, char *buf, size_t cap, size_t *len);

/* Circuit parameters of the parsed circuit (CircuitParameters,
 * Circuit.hpp:20-80): qubits, layers, measurements, detectors, observables,
 * nodes (at the driver's level), fault locations. */
int gp_driver_params(const gp_driver *driver, uint32_t params[7]);

/* Per-stage device timings of the last compile, in microseconds, measured
 * with HIP events on the driver's stream: [0] upload, [1] sensitivity sweep,
 * [2] classes + hash + scatter into sort buckets, [3] unused (0; the
 * partition is part of [2]), [4] bucket sort + segmented reduce,
 * [5] offset scans + emit, [6] download, [7] whole compile (host wall
 * clock). */
int gp_driver_timings(const gp_driver *driver, double timings_us[8]);

/* Compiles the already-uploaded circuit `iterations` times back to back on
 * the device (no host transfer, no DEM download) and returns the mean
 * device time per compile in microseconds; the driver keeps the result of the
 * last iteration on the device.  Used to time the device-resident path.  The
 * compiles go out as HIP graph replays or as direct launches, whichever the
 * first call after a load measured faster over a few of each. */
int gp_driver_bench_device(gp_driver *driver, uint32_t iterations,
                           double *mean_us);

/* Stage outputs for parity tests (small circuits): the dense sensitivity
 * matrix of the p>0 nodes as computed on the GPU, sens[num_pnodes * words64]
 * row-major (64-bit words, bit t = detector t or observable t - D), and the
 * node index of each row.  Call after a compile. */
int gp_debug_pnodes(const gp_driver *driver, uint32_t *num_pnodes,
                    uint32_t *words64);
int gp_debug_sensitivities(const gp_driver *driver, uint32_t *pnode_index,
                           uint64_t *sens);
/* The driver's host STEPG: adjacency[nodes] (child0 low 32 bits, child1 high
 * 32 bits, 0xFFFFFFFF = none) and probs[nodes] (STEPG.hpp:40-132). */
int gp_debug_stepg(const gp_driver *driver, uint64_t *adjacency, double *probs);

/* The driver's time-chunked sweep plan (made at the first compile of a
 * circuit, see the "sweep_split" knob): plan[0] = columns swept in chunks,
 * plan[1] = sweep blocks the first launch adds for their middle chunks'
 * transfer maps, plan[2] = blocks of the second launch (their lower
 * chunks).  All zero: every column is swept whole. */
int gp_debug_sweep_plan(const gp_driver *driver, uint32_t plan[3]);

/* Host-only STEPG build (no device is touched): dims[0..3] = layers, nodes
 * per layer, nodes, fault locations.  Pass NULL arrays to query dims only,
 * then adjacency[nodes] and probs[nodes] as gp_debug_stepg. */
int gp_build_stepg_host(const char *circuit_text, int correlation_level,
                        uint32_t dims[4], uint64_t *adjacency, double *probs);

/* Host STEPG and device image of `circuit_text` built on the calling thread
 * alone (parallel = 0) or on the library's worker pool (parallel = 1),
 * reduced to a digest: digest[0] = FNV-1a-64 over all of their arrays,
 * digest[1] = bytes hashed.  Both builds must agree (a parity check of the
 * pooled build; no device is touched). */
int gp_debug_image_digest(const char *circuit_text, int correlation_level, int parallel, uint64_t digest[2]);

/* ------------------------------------------------------------------------
 * Batched compile: many circuits in one set of launches (no counterpart in
 * the reference, which compiles one circuit per Driver::compile call).
 * --------------------------------------------------------------------- */

/* Creates a batch context on `device`; buffers grow on demand. */
int gp_batch_create(int correlation_level, int device, gp_batch **out);
int gp_batch_destroy(gp_batch *batch);
/* Parses and builds `count` circuits (host threads: `threads`, 0 = auto). */
int gp_batch_set_circuits(gp_batch *batch, const char *const *circuit_texts,
                          uint32_t count, uint32_t threads);
/* Compiles all circuits of the batch on the GPU.  mode 0: one batched
 * pipeline for all circuits (the first compile of a new set of circuits
 * runs it directly; compiling the same set again captures it in a HIP graph
 * and replays it); mode 1: the circuits one after another through the
 * same kernels (the one-by-one baseline). */
int gp_batch_compile(gp_batch *batch, int mode);
/* Runs the loaded batch once more with HIP events between the stages and
 * writes their device times in microseconds: [0] sensitivity sweep,
 * [1] classes + scatter into sort buckets, [2] bucket sort + segmented
 * reduce, [3] emit; *sorted_items (may be NULL) = the items that run sorted.
 * Needs a gp_batch_compile in mode 0 before; the DEMs are unchanged. */
int gp_batch_stage_us(gp_batch *batch, double stage_us[4], uint64_t *sorted_items);
/* Device time of the last gp_batch_compile in microseconds. */
int gp_batch_device_us(const gp_batch *batch, double *us);
/* Shape / copy of the DEM of circuit `index` (as gp_dem_shape/gp_dem_copy). */
int gp_batch_dem_shape(const gp_batch *batch, uint32_t index,
                       uint32_t *num_errors, uint64_t *num_targets,
                       uint32_t *num_detectors, uint32_t *num_observables);
int gp_batch_dem_copy(const gp_batch *batch, uint32_t index, double *probs,
                      uint32_t *offsets, uint32_t *targets, uint64_t *hashes);
/* Bytes of device memory the batch holds now. */
int gp_batch_device_bytes(const gp_batch *batch, uint64_t *bytes);
/* Device layout of the batch's load: stats[0] = 32-bit words of its
 * sensitivity matrix (each detector column stores the layers its time band
 * can reach; see "s_bands"), stats[1] = times a compile found a band
 * reaching below a column's stored layers and compiled again with more,
 * stats[2] = sort buckets, stats[3] = slots per bucket. */
int gp_batch_layout_stats(const gp_batch *batch, uint64_t stats[4]);
/* Sizes read off circuit text in one pass, without parsing or building
 * anything (no counterpart in the reference; for memory admission of
 * queued requests): summary = {layers (TICKs + 1, REPEAT bodies counted per
 * repetition), qubits (highest qubit target + 1), detectors, observables}. */
int gp_circuit_summary(const char *circuit_text, uint64_t summary[4]);
/* Device bytes a new batch context allocates to compile circuit text at a
 * level alone.  An upper bound of the peak during the compile unless the
 * DEM has more than 16 targets per error node or a detector column's time
 * band reaches further back than the rows the batch stores for it (see
 * "s_bands"). */
int gp_estimate_device_bytes(const char *circuit_text, int correlation_level,
                             uint64_t *bytes);

/* ------------------------------------------------------------------------
 * Prepared circuits and memory admission (no counterpart in the reference;
 * for servers that build circuits on their own threads and admit batches
 * by device memory).
 * --------------------------------------------------------------------- */

typedef struct gp_prepared gp_prepared;
/* Parses circuit text and builds its STEPG and device image (the host half
 * of a compile) on the calling thread (and the library's worker pool when
 * it is free).  Any number of threads may prepare circuits at once. */
int gp_prepare_circuit(const char *circuit_text, int correlation_level, gp_prepared **out);
int gp_prepared_destroy(gp_prepared *prepared);
/* gp_estimate_device_bytes of the prepared circuit (no parse). */
int gp_prepared_device_bytes(const gp_prepared *prepared, uint64_t *bytes);
/* Makes `count` prepared circuits (of the batch's correlation level) the
 * batch's circuits, as gp_batch_set_circuits would; the batch shares their
 * images, so the caller may destroy its handles right after. */
int gp_batch_set_prepared(gp_batch *batch, const gp_prepared *const *prepared, uint32_t count);
/* Device bytes the batch holds after compiling these prepared circuits in
 * mode 0: its buffers where they fit, grown with headroom where not.  Every
 * buffer is freed before it is replaced, so this bounds the batch's peak
 * during that compile too (unless the DEM has more than 16 targets per
 * error node).  Nothing is loaded. */
int gp_batch_bytes_after(const gp_batch *batch, const gp_prepared *const *prepared, uint32_t count,
                         uint64_t *bytes);
/* Free and total device memory of `device` (hipMemGetInfo). */
int gp_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* Autotuning knobs (process-wide):
 *   "bucket_fill": mean number of (hash, node) items per sort bucket the
 *   first radix digit aims at; the digit width per circuit is the smallest b
 *   with bucket_fill * 2^b >= its p>0 node count (default 0: 64, or 192
 *   in loads of more than 2^22 error nodes; read at the
 *   next load).
 *   "sweep_rpt": fewest STEPG records per thread of the sensitivity sweep
 *   (1, 2, 4 or 8; default 1, raised as needed to fit a layer in 1024
 *   threads; read when a circuit is parsed).
 *   "stash": 1 (default) stores each node's first 32 targets as a 16-bit
 *   row for the emit stage; 0 makes emit gather every class's targets from
 *   the sensitivity matrix (what circuits above 65536 targets always do);
 *   read at each compile.
 *   "mask_table": 1 (default) lets the class stage keep per-workgroup column
 *   masks in LDS; 0 makes every node test every column's time band (what
 *   circuits whose masks do not fit in LDS always do); read at each compile.
 *   "prob_table": 1 (default) sends a circuit's probabilities as 16-bit
 *   indices into its table of distinct values when it has at most 65536 of
 *   them; 0 always sends one double per error node; read when a circuit is
 *   parsed.
 *   "packed_sort": 1 (default) lets the bucket stage sort on one packed
 *   64-bit key per node (the 32 hash bits below the bucket digit, the
 *   node's rank within its circuit, its slot), falling back to the exact
 *   (hash, rank) sort where two distinct hashes share those bits or a
 *   circuit has 2^24 - 1 or more error nodes; 0 always sorts exactly (same
 *   results); read at each compile.
 *   "s_bands": how batches lay out the sensitivity matrix: 1 (default)
 *   each detector column stores only the layers its time band can reach,
 *   from its lowest seed layer less twice the circuit's largest gap between
 *   seeded layers (columns holding observables store every layer); 0 every
 *   column stores every layer; 2 from its lowest seed layer only, so that
 *   bands reach below and compiles run again with more rows (a test
 *   setting); same results; read at each load.  Drivers always store every
 *   layer (their capacity is fixed at creation).
 *   "bucket_cap": slots per sort bucket, 64, 128 or 256 (default 256):
 *   classify writes a bucket's first bucket_cap nodes to
 *   its slots and the rest to an overflow list that the bucket's workgroup
 *   sorts; slot storage is buckets x bucket_cap; with bucket_fill it
 *   sets the shape of the hash table; same results; read at each load (a
 *   driver keeps the value it was created with).
 *   "target_fill": target slots per error node that a batch's (growable)
 *   output buffer starts with, 1..16 (default 4); a DEM with more targets
 *   grows the buffer and compiles again; same results.  Drivers hold 16.
 *   "compact_items": how a batch stores the bucket stage's sorted nodes and
 *   runs: 1 (default) packed by bucket (at the prefix sums of the bucket
 *   counts: one entry per error node, the outputs in the space of the
 *   bucket slots) for loads of more than 2^22 error nodes, else at each
 *   bucket's slots (buckets x bucket_cap entries plus one per error node);
 *   2 packed for every batch (a test setting); 0 never; same results; read
 *   at each load.  Drivers always use the slots.
 *   "overflow_start": entries a batch's sort overflow list (nodes past
 *   their bucket's slots; at least one in 16 error nodes) and its workspace
 *   for buckets too large for LDS start with, at least 1 (default 2^20); a
 *   compile that needs more grows them and runs again; same results; read
 *   at each load.  Drivers hold the worst case.
 *   "sweep_split": 1 (default) lets a load's second compile plan, from the
 *   bands its first compile found, a time-chunked sweep of a column much
 *   longer than the rest (cut where few nodes are live; the middle chunks'
 *   transfer maps and the top chunk first, the lower chunks waiting for
 *   them in the same launch; identical results); 2 splits every column whose
 *   band reaches the first layers into as many chunks as fit (a test
 *   setting); 0 sweeps every column whole; read when the plan is made. */
int gp_set_tuning(const char *name, uint32_t value);
int gp_get_tuning(const char *name, uint32_t *value);
/* Tuned knobs per size class (what an autotuner installs): circuits at
 * `correlation_level` whose STEPG has n nodes with floor(log2 n) ==
 * size_class take `value` for "bucket_fill", "bucket_cap" or "sweep_rpt"
 * instead of the process-wide setting, when they are parsed (sweep_rpt) or
 * loaded (the others; a batch takes the largest bucket_cap among its
 * circuits).  value 0 removes the entry; gp_get_tuning_class reads 0 for
 * none. */
int gp_set_tuning_class(int correlation_level, uint32_t size_class, const char *name, uint32_t value);
int gp_get_tuning_class(int correlation_level, uint32_t size_class, const char *name, uint32_t *value);
int gp_clear_tuning_classes(void);

/* ------------------------------------------------------------------------
 * Differentiable probabilities (no counterpart in the reference).
 *
 * A plan fixes the structure of a circuit (STEPG and error classes are
 * independent of the noise rates); probabilities then follow from the
 * per-location rates theta[num_locations] (one per DEPOLARIZE1 / X_ERROR /
 * M target and one per DEPOLARIZE2 pair, circuit order) by two segmented
 * XOR-merges: components -> STEPG nodes -> error classes.  All pointers
 * below are device pointers; work is queued on `stream` (a hipStream_t,
 * NULL = default stream).
 * --------------------------------------------------------------------- */

int gp_grad_plan_create(const char *circuit_text, int correlation_level,
                        int device, gp_grad_plan **out);
int gp_grad_plan_destroy(gp_grad_plan *plan);
/* sizes[0] = locations, [1] = components, [2] = p>0 nodes, [3] = errors */
int gp_grad_plan_sizes(const gp_grad_plan *plan, uint32_t sizes[4]);
/* Copies the plan's DEM structure (offsets/targets as gp_dem_copy) to host. */
int gp_grad_plan_dem_structure(const gp_grad_plan *plan, uint32_t *offsets,
                               uint32_t *targets);
/* Device arrays describing the plan, each of the length given by sizes:
 * comp_loc[components], comp_kind[components] (0 rate, 1 DEPOLARIZE1 split,
 * 2 DEPOLARIZE2 split), node_start[nodes+1] + node_comp[components]
 * (components grouped by node), class_start[errors+1] + class_node[nodes]
 * (nodes grouped by error class). */
int gp_grad_plan_device_arrays(const gp_grad_plan *plan,
                               const uint32_t **comp_loc,
                               const uint8_t **comp_kind,
                               const uint32_t **node_start,
                               const uint32_t **node_comp,
                               const uint32_t **class_start,
                               const uint32_t **class_node);

/* Same arrays copied to host memory (any pointer may be NULL). */
int gp_grad_plan_copy_arrays(const gp_grad_plan *plan, uint32_t *comp_loc,
                             uint8_t *comp_kind, uint32_t *node_start,
                             uint32_t *node_comp, uint32_t *class_start,
                             uint32_t *class_node);

/* Segmented XOR-merge, forward: out[s] = XOR-merge of values[index[k]] for k
 * in [seg_start[s], seg_start[s+1]); the merge p (+) q = p + q - 2pq is the
 * identity 1 - 2 out = prod(1 - 2 values), evaluated without cancellation.
 * Replaces the reference's reduce-by-key reducer (CUDAHandlers.hpp:133-138,
 * CUDA.hpp:315-343) for differentiable use. */
int gp_xor_merge_forward(const double *values, const uint32_t *index,
                         const uint32_t *seg_start, uint32_t num_segments,
                         double *out, void *stream);
/* Backward: grad_values[index[k]] = grad_out[s] * prod_{j != k in s}
 * (1 - 2 values[index[j]]).  index must not name a value twice (within or
 * across segments: a partial permutation, as a DEM plan's class members
 * are); entries in no segment are left untouched.  Repeated values are
 * merged from a gathered copy with an identity index instead, their
 * gradients summed back by the caller (greenpeas_amd.autograd.xor_merge does
 * this). */
int gp_xor_merge_backward(const double *values, const uint32_t *index,
                          const uint32_t *seg_start, uint32_t num_segments,
                          const double *grad_out, double *grad_values,
                          void *stream);

/* Node probabilities from fault-location rates, fused: q[n] = XOR-merge over
 * node n's components c = node_comp[node_start[n] .. node_start[n+1]) of the
 * component's split rate of theta[comp_loc[c]] (comp_kind 0: the rate; 1:
 * 1/2 - sqrt(1 - 4t/3)/2, a DEPOLARIZE1 Pauli; 2: 1/2 - (1 - 16t/15)^(1/8)/2,
 * a DEPOLARIZE2 pair -- Mixer.hpp:164, 184).  The arrays are a grad plan's
 * (gp_grad_plan_copy_arrays).  Replaces the reference's host Mixer splits and
 * STEPG::mergeProbabilities (STEPG.hpp:120-131) for differentiable use. */
int gp_node_rates_forward(const double *theta, const uint32_t *comp_loc,
                          const uint8_t *comp_kind, const uint32_t *node_start,
                          const uint32_t *node_comp, uint32_t num_nodes,
                          double *q, void *stream);
/* Backward: grad_comp[c] (scratch, one per component) = grad_q[n] * dq[n] /
 * d rate_c * d rate_c / d t, then grad_theta[l] = sum of grad_comp over
 * location l's components loc_comp[loc_start[l] .. loc_start[l+1]) (a CSR
 * the caller builds from comp_loc; one ordered sum per location). */
int gp_node_rates_backward(const double *theta, const uint32_t *comp_loc,
                           const uint8_t *comp_kind, const uint32_t *node_start,
                           const uint32_t *node_comp, uint32_t num_nodes,
                           const double *grad_q, double *grad_comp,
                           const uint32_t *loc_start, const uint32_t *loc_comp,
                           uint32_t num_locations, double *grad_theta,
                           void *stream);

/* Allocation-free variant of gp_node_rates_backward.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving or fitting loops).  The
 * per-component gradients live in a fixed device array of
 * GP_NOALLOC_MAX_COMPONENTS doubles (8 MiB) that each GPU the library runs
 * on holds until the process ends, so there is no grad_comp argument;
 * num_components (greater than every index in node_comp and loc_comp) above
 * that maximum is rejected with GP_INVALID_ARGUMENT.  Calls no hipMalloc or
 * hipFree.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race; calls queued on one
 * stream are ordered.  Every component loc_comp names must be one node_comp
 * names (as in a grad plan), since each call writes the entries it reads.
 * Results are identical to gp_node_rates_backward's. */
#define GP_NOALLOC_MAX_COMPONENTS 1048576u
int gp_node_rates_backward_noalloc(const double *theta, const uint32_t *comp_loc,
                                   const uint8_t *comp_kind, const uint32_t *node_start,
                                   const uint32_t *node_comp, uint32_t num_nodes,
                                   uint32_t num_components, const double *grad_q,
                                   const uint32_t *loc_start, const uint32_t *loc_comp,
                                   uint32_t num_locations, double *grad_theta, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GREENPEAS_AMD_H */
