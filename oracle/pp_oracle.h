/*
 * pp_oracle.h -- CPU oracle for the rotation-layer hot path of Pauli propagation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product (the HIP library and its host driver) never links, loads or calls it.
 *
 * It restates, in plain C, the host engine of lkdvos/paulistrings-rs for
 * Pauli-rotation channels (citations are into crates/paulistrings/src/):
 *   keys, anticommutation, product phase ........ pauli_string.rs:15-123
 *   splitmix64 / mix64 .......................... rng.rs:5-18
 *   GF(2) bucket hash with rank completion ....... pauli_sum/hash.rs:22-130, :203-361
 *   bucket-count policy, canonical storage ...... pauli_sum/storage.rs:15-58, :290-336, :456-487
 *   rotation table (sin/cos snap, a0/a1) ........ channel/rotation.rs:43-123, channel/prepared.rs:257-281, :436-493
 *   rotation layer scan / partner order / merge . engine/bucketed/rotation_layer.rs:159-229, :386-463,
 *                                                 engine/bucketed/rotation_layer/merge.rs:53-91, :209-286
 *   key-preserving rescale ...................... engine/bucketed.rs:646-662
 *   truncation policies ......................... truncation/builtin.rs:19-239, :470-526
 *   layer loop (order, rebucket, runs, pass) .... engine/driver.rs:186-258, :319-592
 *   product-state read-out ...................... readout/product_state.rs:84-101, readout/mod.rs:16-27
 * and the reference's differential oracle naive_apply_layer
 * (test_support/oracle.rs:21-54) as an independent second check.
 *
 * What pins each part (tests/test_oracle.py unless noted):
 *   ppo_mix64 / ppo_splitmix64 ... the reference's known answer, state 0 -> 0xE220A8397B1DCDAF (rng/tests.rs:3-8)
 *   ppo_hash_rows ................ at 127 qubits the rows are the attempt-0 draws (the reference's own property,
 *                                  pauli_sum/hash/tests.rs:209-221), plus full rank, width independence, linearity;
 *                                  at 1..32, 36 and 64 qubits and four seeds the exact rows, redrawn ones included,
 *                                  equal an independent restatement of row_word and accept_rows with its own GF(2)
 *                                  elimination (pauli_sum/hash.rs:22-26, :102-130)
 *   ppo_desired_bits / ppo_rebucketed_bits ... hand-computed values of the storage.rs:15-41 formulas
 *   ppo_sin_cos .................. quarter-turn snaps (theta = -pi/2 gives exactly -YZ); the ulp-level values
 *                                  are glibc's sincos, the library boundary the reference also crosses:
 *                                  parity unpinned below the 17-digit golden read-outs
 *   ppo_propagate (layer) ........ statevector brute force on 4 qubits; the reference's golden kicked-Ising
 *                                  records (final and peak term counts, <Z62> to 17 digits); the naive layer
 *   keep programs ................ golden records (threshold, weight cap); an independent predicate over the
 *                                  unfiltered layer (And, Or, nested, disabled threshold)
 *   ppo_topn ..................... first_n_by_rank, a stable-sort transcription of test_support/compare.rs:33-55
 *   ppo_approx_topn .............. an independent histogram walk over the canonical export (tests), and the
 *                                  reference's own cases (truncation/tests.rs ApproxTopN tests, restated)
 *   ppo_expectation_product ...... statevector read-out on every axis with sign masks; golden <Z62>
 *   ppo_naive_layer .............. the bucketed engine (two restatements of one semantics, different code)
 */
#ifndef PP_ORACLE_H
#define PP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PPO_F64 = 0, PPO_F32 = 1 };

/* Postfix keep program; node kinds. */
enum {
    PPO_NODE_TRUE = 0,   /* keeps everything */
    PPO_NODE_COEFF = 1,  /* CoefficientThreshold(eps): eps < 0 || |c|^2 > eps^2 */
    PPO_NODE_WEIGHT = 2, /* WeightCutoff(k): weight <= k */
    PPO_NODE_AND = 3,
    PPO_NODE_OR = 4
};

typedef struct {
    int32_t op;
    uint32_t k;
    double eps;
} ppo_node;

#define PPO_MAX_NODES 15

typedef struct {
    int32_t n_nodes;          /* 0: keep everything */
    ppo_node nodes[PPO_MAX_NODES];
    int32_t has_topn;         /* a layer pass after every layer: 1 TopN(topn), 2 ApproxTopN(topn) */
    uint64_t topn;
} ppo_policy;

/* A Pauli rotation exp(-i theta P / 2) about the generator P = (gx, gz). */
typedef struct {
    uint64_t gx[4];
    uint64_t gz[4];
    double theta;
} ppo_channel;

typedef struct ppo_sum ppo_sum;

/* Build a sum like BuildAccumulator::finalize (pauli_sum/accumulator.rs:35-55):
 * repeated keys summed in input order, exact zeros dropped, sorted, bucketed at
 * desired_bits(n, 1024, 128) under the hash drawn from `seed`.
 * x, z: n*W words each, term-major (term i's words at [i*W, i*W+W)). */
ppo_sum* ppo_sum_new(int W, int kind, int num_qubits, size_t n, const uint64_t* x, const uint64_t* z,
                     const void* coeff, uint64_t seed);
void ppo_sum_free(ppo_sum* s);
ppo_sum* ppo_sum_clone(const ppo_sum* s);
size_t ppo_sum_len(const ppo_sum* s);
int ppo_sum_bits(const ppo_sum* s);
/* Canonical order: bucket ascending, lex (x, z) within a bucket.  `bucket` may be NULL. */
void ppo_sum_export(const ppo_sum* s, uint64_t* x, uint64_t* z, void* coeff, uint32_t* bucket);

/* propagate (engine/mod.rs:121-132): direction 0 = forward, 1 = Heisenberg.
 * layer_len (optional, n_channels entries) receives the length after every layer.
 * threads <= 0 uses the OpenMP default.  Returns 0 or a negative error code. */
int ppo_propagate(ppo_sum* s, const ppo_channel* channels, size_t n_channels, int direction,
                  const ppo_policy* policy, int target_bucket_len, int min_buckets, int threads,
                  uint64_t* layer_len);

/* One rotation layer exactly as one step of propagate would apply it (rebucket first). */
int ppo_apply_rotation(ppo_sum* s, const ppo_channel* channel, int adjoint, const ppo_policy* policy,
                       int threads);

/* TopN::finalize_layer (truncation/builtin.rs:73-129). */
void ppo_topn(ppo_sum* s, uint64_t n);
/* ApproxTopN(n)'s layer pass (truncation/builtin.rs:241-468): octaves of |c|^2 kept whole from the top, the
 * boundary octave's first terms in bucket order up to n. */
void ppo_approx_topn(ppo_sum* s, uint64_t n);

/* expectation_product (readout/product_state.rs:84-101); buckets summed in ascending order. */
double ppo_expectation_product(const ppo_sum* s, const uint64_t* axis_x, const uint64_t* axis_z,
                               const uint64_t* negative);

/* The reference's differential oracle (test_support/oracle.rs:21-54): one layer, term by term,
 * into a sorted accumulator, keep program applied, no bucketing. Inputs/outputs term-major, any order in,
 * lex order out. Returns the output count; out arrays must hold 2*n terms. */
size_t ppo_naive_layer(int W, int kind, size_t n, const uint64_t* x, const uint64_t* z, const void* coeff,
                       const ppo_channel* channel, int adjoint, const ppo_policy* policy, uint64_t* out_x,
                       uint64_t* out_z, void* out_coeff);

/* Building blocks exposed for the known-answer tests. */
uint64_t ppo_mix64(uint64_t z);
uint64_t ppo_splitmix64(uint64_t* state);
/* All max_hash_bits(num_qubits) rows of Gf2Hash::new(num_qubits, _, seed); returns the row count.
 * rows_x/rows_z receive count*W words. */
int ppo_hash_rows(int num_qubits, uint64_t seed, int W, uint64_t* rows_x, uint64_t* rows_z);
int ppo_desired_bits(size_t len, size_t target, size_t min_buckets);
int ppo_rebucketed_bits(size_t len, int bits, size_t target, size_t min_buckets);
/* sin_cos with quarter-turn snapping (channel/rotation.rs:70-78). */
void ppo_sin_cos(double theta, double* s, double* c);

#ifdef __cplusplus
}
#endif

#endif
