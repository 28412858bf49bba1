/*
 * pp_hip.h -- C ABI of the MI355X rotation-layer engine (libpp_hip.so).
 *
 * The drop-in boundary for the reference's device backend of Pauli propagation
 * (lkdvos/paulistrings-rs, crates/paulistrings/src/engine/gpu/): a sum resident
 * on one GPU, stepped through Pauli-rotation channels under a built-in
 * truncation policy, read back in the reference's canonical order.  Plain
 * pointers and sizes only; a Rust, ctypes or cgo host binds it as-is.
 *
 * Citations are into crates/paulistrings/src/ of the reference.
 *
 * Conventions:
 *   - keys are the symplectic encoding of pauli_string.rs:15-20: term i's x
 *     words are x[i*W .. i*W+W), its z words z[i*W .. i*W+W) (term-major, as
 *     the reference's x_array()/z_array() hand them out); W is 1 (<= 64 qubits)
 *     or 2 (<= 128 qubits);
 *   - coefficients are float64 or float32 (coefficient.rs:199-200);
 *   - calls on one context are serialized by the caller; work inside a call is
 *     asynchronous on the context's stream and pp_sync() waits for it;
 *   - every entry point returns a pp_status; pp_last_error() explains the last
 *     failure on a context.
 */
#ifndef PP_HIP_H
#define PP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapping the reference's GpuError (engine/gpu/error.rs:9-75). */
typedef enum {
    PP_OK = 0,
    PP_ERR_NO_DEVICE = 1,        /* GpuError::NoDevice */
    PP_ERR_OUT_OF_MEMORY = 2,    /* GpuError::OutOfMemory { device, bytes } */
    PP_ERR_UNSUPPORTED = 3,      /* GpuError::Unsupported(reason) */
    PP_ERR_DRIVER = 4,           /* GpuError::Driver(hipError_t) */
    PP_ERR_INVALID_ARGUMENT = 5
} pp_status;

typedef enum { PP_COEFF_F64 = 0, PP_COEFF_F32 = 1 } pp_coeff_kind;

typedef enum { PP_FORWARD = 0, PP_HEISENBERG = 1 } pp_direction; /* engine/mod.rs Direction */

/* A Pauli rotation exp(-i theta P / 2) about P = (gx, gz): PauliRotation::new
 * (channel/rotation.rs:15-25).  Words beyond W are ignored. */
typedef struct {
    uint64_t gx[4];
    uint64_t gz[4];
    double theta;
} pp_channel;

/* A built-in truncation policy (truncation/builtin.rs, truncation/tree.rs):
 * the per-term filter as a postfix program of at most 15 nodes (the reference's
 * device KeepProgram, engine/gpu/truncation.rs:8-102), and an optional layer
 * pass after every layer: TopN::finalize_layer (builtin.rs:73-129) or ApproxTopN's (builtin.rs:241-468). */
typedef enum {
    PP_KEEP_TRUE = 0,
    PP_KEEP_COEFF = 1,  /* CoefficientThreshold(eps): eps < 0 or |c|^2 > eps^2 */
    PP_KEEP_WEIGHT = 2, /* WeightCutoff(k): Pauli weight <= k */
    PP_KEEP_AND = 3,
    PP_KEEP_OR = 4
} pp_keep_op;

typedef struct {
    int32_t op;
    uint32_t k;
    double eps;
} pp_keep_node;

typedef struct {
    int32_t n_nodes; /* 0: keep every nonzero term */
    pp_keep_node nodes[15];
    int32_t has_topn; /* the layer pass after every layer: 0 none, 1 TopN(topn), 2 ApproxTopN(topn) */
    uint64_t topn;
} pp_policy;

/* PropagateOptions (engine/mod.rs:62-106): the reference's bucket-count knobs,
 * which decide its canonical order and TopN's tie order. */
typedef struct {
    uint32_t target_bucket_len; /* default 1024 */
    uint32_t min_buckets;       /* default 128 */
} pp_options;

typedef struct pp_context pp_context;
typedef struct pp_sum pp_sum;

/* ---- context ------------------------------------------------------------------------------ */

/* One device and one stream (a lone GpuSum's own context, engine/gpu/sum.rs:71-80). */
pp_status pp_context_create(int device, pp_context** out);
void pp_context_destroy(pp_context* ctx);
const char* pp_last_error(const pp_context* ctx);
pp_status pp_sync(pp_context* ctx);
/* The hipStream_t every call of this context is queued on. */
pp_status pp_context_stream(pp_context* ctx, void** stream);

/* ---- resident sums (GpuSum, engine/gpu/sum.rs:47-253) ----------------------------------------- */

/* Upload n terms with distinct keys (a PauliSum's terms), hash seeded by `hash_seed`
 * (DEFAULT_HASH_SEED = 0x9E3779B97F4A7C15, pauli_sum/storage.rs:12).  Exact zeros are dropped.
 * The sum starts at the reference's bucket count desired_bits(n, 1024, 128) (pauli_sum/accumulator.rs:49-52). */
pp_status pp_sum_upload(pp_context* ctx, uint32_t num_qubits, uint32_t width, pp_coeff_kind kind, uint64_t n,
                        const uint64_t* x, const uint64_t* z, const void* coeff, uint64_t hash_seed,
                        pp_sum** out);
/* A device-side copy (same context), e.g. to reset a benchmark input without a host round trip. */
pp_status pp_sum_clone(const pp_sum* src, pp_sum** out);
/* src copied into dst's own memory (same context, width, coefficient type, qubit count and hash seed): dst's terms are
 * discarded and it takes src's layout; its columns grow only when too small.  To reset a working sum from a resident
 * input many times (an in-place layer per reset) with no allocation after the first.  Asynchronous. */
pp_status pp_sum_copy(pp_sum* dst, pp_sum* src);
pp_status pp_sum_free(pp_sum* sum);
pp_status pp_sum_len(pp_sum* sum, uint64_t* len);
/* Download in canonical order: the reference's bucket index ascending, then lexicographic (x, z)
 * (pauli_sum/storage.rs:285-295, GpuSum::to_host engine/gpu/sum.rs:216-253).  Arrays hold len terms. */
pp_status pp_sum_download(pp_sum* sum, uint64_t* x, uint64_t* z, void* coeff);
/* The reference's bucket bits for this sum, and the device's own. */
pp_status pp_sum_bucket_bits(pp_sum* sum, int32_t* reference_bits, int32_t* device_bits);
/* The resident layout's invariants, counted (the reference's debug-build k_check_invariants,
 * engine/gpu/kernels/kern/invariants.cu:5, run from engine/gpu/sum.rs:559-565): out[0] rows whose key hashes to
 * another device bucket, [1] zero coefficients, [2] buckets whose segment is misaligned, outside the columns or
 * longer than its room, [3] keys held twice in one bucket, [4] buckets too long for the duplicate scan (over 1024
 * rows), [5] overlapping segments, [6] 1 when the bucket lengths do not add up to the sum's length.  All zero for a
 * sound layout.  With PP_CHECK_INVARIANTS=1 in the environment every layer entry point runs it after its layer and
 * fails with PP_ERR_DRIVER on any violation. */
pp_status pp_sum_check(pp_sum* sum, uint64_t* out);

/* ---- layers -------------------------------------------------------------------------------- */

/* propagate (engine/mod.rs:121-187 -> engine/driver.rs:319-592): every channel in circuit order
 * (reversed, each adjoint, for PP_HEISENBERG), each layer's outputs kept iff nonzero and the policy's
 * per-term filter passes, the TopN pass after every run of layers (Runs::take).  A run of one-to-one steps
 * (engine/bucketed/permutation_run.rs) is walked in one pass with the same keys, coefficients and per-layer
 * lengths as layer by layer (PP_PERMUTATION_RUNS=layers in the environment selects the latter).  layer_len
 * (optional, n_channels entries) receives the length after every layer.  options may be NULL for the
 * defaults. */
pp_status pp_propagate(pp_sum* sum, const pp_channel* channels, uint64_t n_channels, pp_direction direction,
                       const pp_policy* policy, const pp_options* options, uint64_t* layer_len);

/* One rotation layer on the device's current bucketing, with no rebucket and no layer pass: the
 * kernel the benchmark times.  It is PauliRotation::prepare_for_layer (channel/rotation.rs:94-123) and
 * DevicePartition::apply_layer (engine/gpu/partition.rs:315-374) in one call, for a host that holds angles rather
 * than prepared tables; pp_prepare_channel + pp_apply_prepared below are the two halves. */
pp_status pp_apply_rotation(pp_sum* sum, const pp_channel* channel, int adjoint, const pp_policy* policy);

/* The same layer out of place: dst receives the layer applied to src, which is left as it was (dst's previous
 * terms are discarded).  Asynchronous: no host synchronization; pp_sum_len(dst) waits for the counts.  For
 * stepping many layers from one resident input, e.g. every channel of a dense layer on the same saturated sum.
 * Both sums belong to one context and share width, coefficient type, qubit count and hash seed. */
pp_status pp_apply_rotation_to(pp_sum* src, pp_sum* dst, const pp_channel* channel, int adjoint,
                               const pp_policy* policy);

/* pp_apply_rotation_to with no device allocation, for callers that cannot allocate (tight serving loops, code that
 * must not call hipMalloc); no counterpart in the reference.  The join's scratch (offset-scan tile sums, counters,
 * redo queue) lives in a fixed per-GPU array of the library, for inputs of at most 2^18 device buckets (~2^26
 * terms); dst's columns and buckets must already have room (one pp_apply_rotation_to of the same input sizes them)
 * and src's device bucketing must be settled; otherwise it returns PP_ERR_INVALID_ARGUMENT / PP_ERR_UNSUPPORTED
 * before queuing anything.  Calls on one GPU share that array: calls that overlap on one GPU (from two contexts'
 * streams) race.  Not capturable into a graph. */
pp_status pp_apply_rotation_to_noalloc(pp_sum* src, pp_sum* dst, const pp_channel* channel, int adjoint,
                                       const pp_policy* policy);

/* TopN(n) as a standalone pass (truncation/builtin.rs:73-129): ties at the n-th magnitude are kept
 * in the reference's canonical order. */
pp_status pp_truncate_topn(pp_sum* sum, uint64_t n);

/* ApproxTopN(n) as a standalone pass (truncation/builtin.rs:241-468, approx_top_n_solo): octaves of |c|^2
 * kept whole from the top while they fit in n, the boundary octave's first terms in the reference's
 * canonical order up to n in all; NaN coefficients are never kept once a cut happens. */
pp_status pp_truncate_approx_topn(pp_sum* sum, uint64_t n);

/* expectation_product (readout/product_state.rs:84-101) in the product state whose qubit q is the
 * +-1 eigenstate of the axis (axis_x, axis_z) bits, -1 where `negative` is set.  W words each. */
pp_status pp_expectation_product(pp_sum* sum, const uint64_t* axis_x, const uint64_t* axis_z,
                                 const uint64_t* negative, double* out);

/* ---- the layer seam: PartitionBackend / PartitionStorage (engine/backend.rs:26-108) ---------------- */

/* What a host that keeps its own run_layers (engine/driver.rs:319-592), Channel and TruncationPolicy calls per
 * layer, in place of pp_propagate.  The host prepares each channel (Channel::prepare_for_layer), decides the
 * reference's bucket count and the runs (Runs::take, engine/driver.rs:202-258), and calls the layer pass after
 * every run; the device keeps its own bucketing underneath. */

/* DeltaEntry (channel/prepared.rs:35-46): one key delta of a local table.  Pattern s packs the support qubits'
 * bits, bit 2j the x bit of support qubit j and bit 2j+1 its z bit (prepared.rs:141-151); amplitude[s] takes
 * pattern s to s ^ local_delta, and an exact 0 emits nothing.  A float32 sum's host passes its float32 amplitudes
 * widened to double; the device narrows them back exactly. */
typedef struct {
    uint32_t local_delta;
    double amplitude[16];
} pp_delta_entry;

typedef enum {
    PP_PREPARED_LOCAL = 0,   /* Prepared::Local(LocalPtm): support <= 2 qubits, tabulated (prepared.rs:79-86) */
    PP_PREPARED_ROTATION = 1 /* Prepared::Rotation(PreparedRotation): a wider Pauli rotation (prepared.rs:303-317) */
} pp_prepared_kind;

/* Prepared (channel/prepared.rs:343-350).  LOCAL: k support qubits, ascending, and n_deltas entries ascending by
 * local_delta (the identity first when present).  ROTATION: the generator and its cos and sin.  Bucket deltas
 * are the device's own and not passed. */
typedef struct {
    int32_t kind;
    uint32_t k;
    uint32_t qubits[2];
    uint32_t n_deltas;
    pp_delta_entry deltas[16];
    uint64_t gx[4];
    uint64_t gz[4];
    double cos;
    double sin;
} pp_prepared;

/* PauliRotation::prepare_for_layer (channel/rotation.rs:94-123, Prepared::derive_local prepared.rs:401-426): the
 * rotation's table on its support (weight <= 2) or its PreparedRotation, with sin and cos snapped at quarter turns
 * (rotation.rs:70-78) and negated angle for the adjoint, amplitudes rounded to the coefficient type.  For a host
 * that holds a PauliRotation; host-only (no device needed). */
pp_status pp_prepare_channel(uint32_t num_qubits, uint32_t width, pp_coeff_kind kind, const pp_channel* channel,
                             int adjoint, pp_prepared* out);

/* PartitionBackend::apply_layer (engine/backend.rs:80-89; the device one engine/gpu/partition.rs:315-374): one
 * prepared channel under the policy's per-term filter, no layer pass.  Supported: a table shaped like a rotation
 * (RotationLayer::of, engine/bucketed/rotation_layer.rs:101-140) whose amplitudes are a rotation's (the pair join,
 * its Clifford arm at quarter turns); a key-preserving table (one entry, local_delta 0: an in-place rescale by
 * amplitude[s]); a wider rotation; any other local table of at most 16 entries (the general table layer).  The
 * sum's length is exact on return. */
pp_status pp_apply_prepared(pp_sum* sum, const pp_prepared* prepared, const pp_policy* policy);

/* PartitionBackend::apply_run (engine/backend.rs:99-104): the layers of one run (Run::Rotations, at most 8
 * rotation tables whose bucket deltas span at most 3 dimensions; or Run::Permutations, one-to-one steps) with the
 * same keys, coefficients and per-layer lengths as layer by layer; a permutation run is walked in one pass.
 * layer_len (optional, n entries) receives the length after each layer.  The caller keeps the run within the
 * policy's run_length_limit (truncation/builtin.rs:147-149), as run_layers does. */
typedef enum { PP_RUN_ROTATIONS = 0, PP_RUN_PERMUTATIONS = 1 } pp_run_kind;
pp_status pp_apply_run(pp_sum* sum, pp_run_kind kind, const pp_prepared* layers, uint32_t n, const pp_policy* policy,
                       uint64_t* layer_len);

/* PartitionBackend::finalize_layer (engine/backend.rs:90; engine/gpu/partition.rs:439-447): the policy's layer
 * pass (TopN or ApproxTopN; nothing for a per-term policy).  An exact TopN's retain is left to the next layer on
 * the device, which reads its input through the selection; pp_settle (or any read of the sum) applies it. */
pp_status pp_finalize_layer(pp_sum* sum, const pp_policy* policy);

/* PartitionStorage::settle (engine/backend.rs:60-62; engine/gpu/partition.rs:270): apply what the last layer pass deferred. */
pp_status pp_settle(pp_sum* sum);

/* PartitionStorage::refine / coarsen_to (engine/backend.rs:35-37; engine/gpu/partition.rs:230-239): the reference's bucket count, which fixes the
 * canonical order and TopN's choice among ties.  The device bucketing follows on its own. */
pp_status pp_sum_refine(pp_sum* sum);
pp_status pp_sum_coarsen_to(pp_sum* sum, int32_t bits);

/* ---- measurement --------------------------------------------------------------------------- */

/* While enabled, the rotation-join kernel of every layer is bracketed by hipEvents on the context's
 * stream; pp_profile_read returns the launches and their total milliseconds since it was enabled
 * (and the rows read and written by those launches), after a stream synchronization. */
pp_status pp_profile_enable(pp_context* ctx, int enable);
pp_status pp_profile_read(pp_context* ctx, uint64_t* launches, double* total_ms, uint64_t* rows_in,
                          uint64_t* rows_out);
/* The same for the learned pass's score kernel (pp_truncate_learned, pp_propagate_learned): its launches, their
 * total milliseconds and the terms they scored since profiling was enabled.  No counterpart in the reference. */
pp_status pp_profile_read_score(pp_context* ctx, uint64_t* launches, double* total_ms, uint64_t* terms);

/* An HBM ceiling probe beside the measurements (the reference's membench, crates/membench/kernels/membench.cu,
 * crates/membench/src/gpu.rs:17-60): STREAM read, write, copy and triad passes over `bytes`-byte f64 arrays on the
 * context's stream, each timed with hipEvents, best of `reps`.  gbs[0..4) receives read, write, copy and triad in
 * GB/s, counting 1, 1, 2 and 3 arrays moved per pass.  It allocates its three arrays and frees them. */
pp_status pp_stream_probe(pp_context* ctx, uint64_t bytes, uint32_t reps, double* gbs);

/* ---- batches ------------------------------------------------------------------------------------ */

/* Many independent propagations stepped together (no counterpart in the reference: each of its instances is one
 * propagate call, engine/mod.rs:121-132).  A batch holds sums of one context, one width and one coefficient type;
 * freeing it leaves the sums.  pp_batch_propagate steps every sum i through its own circuit
 * channels[offsets[i] .. offsets[i + 1]) in directions[i] under policies[i] (NULL: keep every nonzero term),
 * each with exactly pp_propagate's result; per step, one launch of each batched kernel serves every sum that
 * takes a rotation join (a key-preserving layer, a refinement or a layer pass runs on its sum alone). */
typedef struct pp_batch pp_batch;
pp_status pp_batch_create(pp_context* ctx, pp_sum* const* sums, uint32_t count, pp_batch** out);
pp_status pp_batch_free(pp_batch* batch);
pp_status pp_batch_propagate(pp_batch* batch, const pp_channel* channels, const uint64_t* offsets,
                             const pp_direction* directions, const pp_policy* policies, const pp_options* options);

/* ---- the learned truncation policy --------------------------------------------------------------- */

/* No counterpart in the reference: a TruncationPolicy (truncation/mod.rs:30-113) whose layer pass keeps the n
 * terms of largest predicted contribution to the final expectation value, with plain |c| TopN
 * (truncation/builtin.rs:61-150) as its fallback and judge.  The predictor is a 35-128-128-1 ReLU MLP over a
 * one-word term's features: X and Z occupancy counts of 16 qubit groups (a lattice's rows, then its columns;
 * unused groups 0), its Pauli weight, ln|c| and the layers still to apply (divided by depth_scale, capped at 1).
 * It is trained to output ln|<psi| P propagated through the remaining layers |psi>|, the term's read-out factor;
 * the pass ranks by ln|c| plus that output.  Its two hidden GEMMs run in bf16 on MFMA with fp32 accumulation;
 * parameters and Adam state are fp32 on the device.
 * Parameters: pp_model_param_count() floats, W1 [64][128] (input-major, rows 35..63 zero), b1 [128],
 * W2 [128][128], b2 [128], w3 [128], b3.  Samples are one-word terms: x, z, f64 coefficient, remaining layers. */
typedef struct pp_model pp_model;
uint32_t pp_model_param_count(void);
pp_status pp_model_create(pp_context* ctx, const uint64_t* groups, float depth_scale, uint64_t seed,
                          pp_model** out);
pp_status pp_model_free(pp_model* model);
pp_status pp_model_set_params(pp_model* model, const float* params);  /* also resets the optimizer */
pp_status pp_model_get_params(pp_model* model, float* params);
pp_status pp_model_predict(pp_model* model, const uint64_t* x, const uint64_t* z, const double* coeff,
                           const float* remaining, uint64_t n, float* out);
/* Gradient of the mean squared error of the predictions against labels, and that error.  Both are summed in a
 * fixed order (per 64-sample tile, then over tiles): the same inputs give the same bits on every call. */
pp_status pp_model_gradient(pp_model* model, const uint64_t* x, const uint64_t* z, const double* coeff,
                            const float* remaining, const float* labels, uint64_t n, float* grads, double* loss);
/* Adam (0.9, 0.999, 1e-8) over `epochs` shuffled passes in batches of `batch`; loss = the last epoch's mean.
 * Deterministic for a given seed (the shuffle is a splitmix64 stream, the gradients as pp_model_gradient's). */
pp_status pp_model_train(pp_model* model, const uint64_t* x, const uint64_t* z, const double* coeff,
                         const float* remaining, const float* labels, uint64_t n, uint32_t epochs, uint32_t batch,
                         float lr, uint64_t seed, double* loss);
/* The learned layer pass on a one-word sum: keep the n terms of largest score ln|c| + prediction (ties in canonical
 * order), the prediction being the network's output for the term's features with `remaining` layers left. */
pp_status pp_truncate_learned(pp_sum* sum, pp_model* model, uint64_t n, float remaining);
/* pp_propagate with the learned pass (keep n) after every run; `filter` is a per-term policy without a pass. */
pp_status pp_propagate_learned(pp_sum* sum, const pp_channel* channels, uint64_t n_channels, pp_direction direction,
                               const pp_policy* filter, pp_model* model, uint64_t keep, const pp_options* options,
                               uint64_t* layer_len);

/* ---- tuning ------------------------------------------------------------------------------------- */

/* The device path's sizes, per context.  No counterpart in the reference, which fixes its device sizes
 * (records_per_block = 4096, target_bucket_len = 1024: engine/gpu/layer/options.rs:11, pauli_sum/storage.rs:54);
 * results do not depend on them.  Defaults come from PP_DEVICE_BUCKET_LEN, PP_JOIN_STAGE_ROWS and
 * PP_JOIN_GROUPS_PER_CU, PP_JOIN_BLOCK and PP_JOIN_IN_PLACE when set. */
typedef struct {
    uint32_t bucket_target; /* rows per device bucket, 16..512 (default 256) */
    uint32_t stage_rows;    /* anticommuting rows a join workgroup stages in its first launch, 64..1024 (a pair
                               with more is redone by a second launch that stages the largest pair);
                               0: the largest pair, no second launch (default) */
    uint32_t groups_per_cu; /* the join's grid: 0xFFFFFFFF one workgroup per pair (default); 0 a persistent
                               grid of as many as fit; 1..32 persistent with at most that many per CU */
    uint32_t join_block;    /* threads per join workgroup, the rows of one pass: 256 (default) or 512 */
    uint32_t join_in_place; /* 1 (default): a layer of pp_propagate / pp_apply_rotation / pp_apply_prepared runs in
                               place when the sum's segments have room for it (no pending TopN selection, not a
                               Clifford arm), writing only the rows it changes; 0: always into the spare columns */
} pp_tuning;
pp_status pp_get_tuning(pp_context* ctx, pp_tuning* out);
pp_status pp_set_tuning(pp_context* ctx, const pp_tuning* tuning);
/* Time candidate tunings (bucket_target 128/192/256/384 x stage_rows auto/128/256 x one-per-pair/persistent x
 * join_block 256/512) on
 * `sample`, each channel applied out of place from it as pp_apply_rotation_to does, keep the fastest in the
 * context and return it (out may be NULL).  The sample is left as it was. */
pp_status pp_autotune(pp_sum* sample, const pp_channel* channels, uint64_t n_channels, const pp_policy* policy,
                      pp_tuning* out);

/* ---- layer sequences captured into a HIP graph ---------------------------------------------- */

/* No counterpart in the reference (its device engine does not capture: research/FINDINGS.md:489, :663).
 * Between pp_graph_capture_begin and pp_graph_capture_end, the context's stream is captured instead of run:
 * only pp_apply_rotation_to is accepted, from inputs whose counts are settled (no pending layer) and into
 * outputs already sized by an earlier identical call; any entry point that would synchronize or allocate
 * returns PP_ERR_UNSUPPORTED (the capture is then unusable: end it and free the graph).  pp_graph_launch
 * replays the captured layers asynchronously on the context's stream with the captured buffers; the outputs'
 * counts are read, as after the eager call, once the replay has run (pp_sum_len waits for it). */
typedef struct pp_graph pp_graph;
pp_status pp_graph_capture_begin(pp_context* ctx);
pp_status pp_graph_capture_end(pp_context* ctx, pp_graph** out);
pp_status pp_graph_launch(pp_graph* graph);
pp_status pp_graph_free(pp_graph* graph);

/* The reference's hash rows (Gf2Hash::new, pauli_sum/hash.rs:212-258), for checks: count*W words each. */
int32_t pp_hash_rows(uint32_t num_qubits, uint64_t seed, uint32_t width, uint64_t* rows_x, uint64_t* rows_z);

#ifdef __cplusplus
}
#endif

#endif
