/*
 * gp_oracle.h -- CPU oracle for the detector-error-model (DEM) compile path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (greenpeas_amd/) may
 * include, link or call this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * Two independent CPU compilers live here:
 *
 *  1. gpo_compile_stepg(): a plain-C restatement of the reference's STEPG
 *     algorithm -- parse (Circuit.hpp:290-478), Mixer flow rules and noise
 *     splits (Mixer.hpp:97-533), XOR-merge of node probabilities
 *     (STEPG.hpp:128-131), backward sensitivity sweep (CUDA.hpp:37-51 with
 *     CUDAHandlers.hpp:46-61), class extraction + FNV-1a-64 over 32 padded
 *     slots (CUDAHandlers.hpp:86-121, Array.hpp:52-65), stable sort by hash
 *     (CUDA.hpp:284-296) and a sequential fold with the reference reducer
 *     (CUDAHandlers.hpp:133-138).  Output order is the reference's order
 *     (ascending FNV hash).
 *
 *  2. gpo_compile_frames(): an independent forward Pauli-frame compiler with
 *     stim's semantics (each independent Pauli component of every noise
 *     channel is propagated forward through the Clifford gates; measurement
 *     flips are read off; identical symptom sets are XOR-merged).  It shares
 *     no code with (1) beyond the text parser and stands in for stim's
 *     ErrorAnalyzer, which is absent from this environment.
 *
 * Semantics chosen where the reference has quirks (all harmless on the
 * reference's circuits): detector/observable records are XORed (the
 * reference ORs, Circuit.hpp:299); OBSERVABLE_INCLUDE(k) sets bit D+k (the
 * reference takes the next detector counter, Circuit.hpp:305-307); the sweep
 * XORs a node's seed with its children (the reference overwrites,
 * CUDAHandlers.hpp:57-60); classes are not truncated at 32 targets
 * (CUDAHandlers.hpp:94) -- longer classes hash all targets with no padding;
 * nodes with probability 0 are dropped before grouping (the reference keeps
 * them and could print error(0) lines, Driver.hpp:185); R/CX/H in the last
 * layer add or remove no flow (the reference throws, Graph.hpp:163-165).
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpo_dem gpo_dem;
typedef struct gpo_stepg gpo_stepg;

/* Last error message of this thread (empty string if none). */
const char *gpo_last_error(void);

/* ---- whole-path compilers -------------------------------------------- */
/* Returns 0 on success, 1 for invalid input (bad text / unsupported gate),
 * 2 for a runtime failure (flow rule violation, bad probability). */
int gpo_compile_stepg(const char *text, int level, gpo_dem **out);
/* Same, with theta[num_theta] overriding the rate of each fault location
 * (one per DEPOLARIZE1 / X_ERROR / M target, one per DEPOLARIZE2 pair, in
 * circuit order). */
int gpo_compile_stepg_theta(const char *text, int level, const double *theta,
                            uint32_t num_theta, gpo_dem **out);
int gpo_compile_frames(const char *text, gpo_dem **out);
void gpo_dem_free(gpo_dem *dem);

uint32_t gpo_dem_num_errors(const gpo_dem *dem);
uint64_t gpo_dem_num_targets(const gpo_dem *dem);
uint32_t gpo_dem_num_detectors(const gpo_dem *dem);
uint32_t gpo_dem_num_observables(const gpo_dem *dem);
/* probs[num_errors], offsets[num_errors+1], targets[num_targets]; a target
 * t < num_detectors is detector t, otherwise observable t - num_detectors. */
void gpo_dem_copy(const gpo_dem *dem, double *probs, uint32_t *offsets,
                  uint32_t *targets);
/* FNV hash of each error, in output order (STEPG compiler only). */
void gpo_dem_hashes(const gpo_dem *dem, uint64_t *hashes);

/* ---- intermediate stages (for stage-by-stage parity) ------------------- */
int gpo_build_stepg(const char *text, int level, gpo_stepg **out);
void gpo_stepg_free(gpo_stepg *g);
/* dims[0..7] = layers, qubits, nodes_per_qubit, nodes, measurements,
 *              detectors, observables, words_per_node */
void gpo_stepg_dims(const gpo_stepg *g, uint32_t *dims);
void gpo_stepg_copy(const gpo_stepg *g, uint64_t *adjacency, double *probs,
                    uint32_t *keys, uint64_t *vals_row_major);
/* Dense sensitivity matrix S[node][word] (row-major), nodes*words u64. */
void gpo_stepg_sweep(const gpo_stepg *g, uint64_t *sens_row_major);

/* ---- primitives -------------------------------------------------------- */
uint64_t gpo_fnv1a64_slots(const uint32_t *vals, uint32_t n, uint32_t pad_to);
double gpo_split_dep1(double p);
double gpo_split_dep2(double p);
double gpo_xor_merge(double p_old, double p_new);

/* Host-side gradient of L = sum_e w_e * logit(p_e) with respect to the
 * per-location noise rates theta of a circuit whose noise instructions
 * carry one target (or target pair) each, in circuit order.  Uses the STEPG
 * compiler's grouping; grad has one entry per fault location. */
int gpo_location_count(const char *text, uint32_t *count);
int gpo_logit_loss_grad(const char *text, int level, const double *theta,
                        uint32_t num_theta, const double *w, uint32_t num_w,
                        double *loss, double *grad, uint32_t num_locations);

#ifdef __cplusplus
}
#endif

#endif /* GP_ORACLE_H */
