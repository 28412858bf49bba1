/*
 * pp_oracle.c -- CPU oracle for the rotation-layer hot path (see pp_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP product path and the CPU
 * baseline of the benchmark, never part of the product.
 *
 * This file holds the width-independent parts: the splitmix64 streams, the
 * GF(2) bucket hash, the bucket-count policy, the rotation table, and the layer
 * loop with the reference's run grouping (which decides the bucket count every
 * layer sees, and so the tie order of TopN and the canonical output order).
 */
#define _GNU_SOURCE /* sincos */
#include "pp_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define PPO_INDEX_BITS 31 /* pauli_sum/hash.rs:7 */
#define PPO_MIN_TERMS_PER_TASK 64 /* pauli_sum/storage.rs:48 */
#define PPO_COARSEN_BELOW 8 /* pauli_sum/storage.rs:44 */
#define PPO_REDRAW_ATTEMPTS 4096 /* pauli_sum/hash.rs:18 */
#define PPO_SPLITMIX_GAMMA 0x9E3779B97F4A7C15ull /* rng.rs:5 */
#define PPO_MAX_ROTATION_RUN_LAYERS 8 /* engine/bucketed/rotation_layer.rs:508 */
#define PPO_MAX_ROTATION_RUN_DIMENSION 3 /* engine/bucketed/rotation_layer.rs:511 */

struct ppo_sum {
    int W;
    int kind;
    int num_qubits;
    int bits;
    int max_bits;
    uint64_t seed;
    uint64_t rows_x[PPO_INDEX_BITS][4];
    uint64_t rows_z[PPO_INDEX_BITS][4];
    size_t len;
    void* buckets;
    int known; /* every term passed keep_term in an earlier layer of this call, supports valid */
};

/* ---- splitmix64 (rng.rs:8-18) --------------------------------------------------------------- */

uint64_t ppo_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t ppo_splitmix64(uint64_t* state) {
    *state += PPO_SPLITMIX_GAMMA;
    return ppo_mix64(*state);
}

/* ---- the GF(2) bucket hash (pauli_sum/hash.rs:22-130, :203-258) ------------------------------ */

static uint64_t word_mask(int num_qubits, int word) {
    int first = 64 * word;
    if (num_qubits >= first + 64) return ~0ull;
    if (num_qubits <= first) return 0;
    return (1ull << (num_qubits - first)) - 1ull;
}

/* row_word: word `word` of the `half` (0 = x, 1 = z) of row `row`, draw `attempt`. */
static uint64_t row_word(uint64_t seed, int row, uint32_t attempt, int word, int half) {
    uint64_t position = ((uint64_t)attempt << 32) | ((uint64_t)row << 16) | ((uint64_t)word << 1) | (uint64_t)half;
    return ppo_mix64(seed + PPO_SPLITMIX_GAMMA * (position + 1ull));
}

/* An incrementally grown echelon basis over key-space rows of up to 4+4 words.
 * Independence does not depend on which column a pivot is named by, so columns
 * are numbered x-words first, most significant bit first. */
typedef struct {
    int n;
    int pivot[2 * PPO_INDEX_BITS + 2];
    uint64_t row[2 * PPO_INDEX_BITS + 2][8];
} gf2_basis;

static int leading_column(const uint64_t* r, int words) {
    for (int w = 0; w < words; ++w)
        if (r[w]) return 64 * w + __builtin_clzll(r[w]);
    return -1;
}

static int basis_insert(gf2_basis* b, const uint64_t* row, int words) {
    uint64_t red[8];
    memcpy(red, row, (size_t)words * sizeof(uint64_t));
    for (;;) {
        int lead = leading_column(red, words);
        if (lead < 0) return 0;
        int found = -1;
        for (int i = 0; i < b->n; ++i)
            if (b->pivot[i] == lead) {
                found = i;
                break;
            }
        if (found < 0) {
            b->pivot[b->n] = lead;
            memcpy(b->row[b->n], red, (size_t)words * sizeof(uint64_t));
            b->n++;
            return 1;
        }
        for (int w = 0; w < words; ++w) red[w] ^= b->row[found][w];
    }
}

int ppo_hash_rows(int num_qubits, uint64_t seed, int W, uint64_t* rows_x, uint64_t* rows_z) {
    int rows = 2 * num_qubits < PPO_INDEX_BITS ? 2 * num_qubits : PPO_INDEX_BITS;
    int independent = rows; /* min(rows, 2 * num_qubits) */
    gf2_basis basis;
    basis.n = 0;
    for (int r = 0; r < rows; ++r) {
        uint64_t key[8];
        for (uint32_t attempt = 0;; ++attempt) {
            int zero = 1;
            for (int w = 0; w < W; ++w) {
                uint64_t m = word_mask(num_qubits, w);
                key[w] = row_word(seed, r, attempt, w, 0) & m;
                key[W + w] = row_word(seed, r, attempt, w, 1) & m;
                zero &= key[w] == 0 && key[W + w] == 0;
            }
            int accepted = basis.n < independent ? basis_insert(&basis, key, 2 * W) : !zero;
            if (accepted || attempt == PPO_REDRAW_ATTEMPTS) break;
        }
        for (int w = 0; w < W; ++w) {
            rows_x[r * W + w] = key[w];
            rows_z[r * W + w] = key[W + w];
        }
    }
    return rows;
}

/* ---- bucket-count policy (pauli_sum/storage.rs:15-41) --------------------------------------- */

int ppo_desired_bits(size_t len, size_t target, size_t min_buckets) {
    int floor_bits = 0;
    if (len >= min_buckets * PPO_MIN_TERMS_PER_TASK) {
        while (((size_t)1 << floor_bits) < min_buckets && floor_bits < PPO_INDEX_BITS) floor_bits++;
    }
    int b = floor_bits;
    while (b < PPO_INDEX_BITS && len > target * ((size_t)1 << b)) b++;
    return b;
}

int ppo_rebucketed_bits(size_t len, int bits, size_t target, size_t min_buckets) {
    int want = ppo_desired_bits(len, target, min_buckets);
    int sparse = len * PPO_COARSEN_BELOW < target * ((size_t)1 << bits);
    return (want > bits || sparse) ? want : bits;
}

/* ---- the rotation table (channel/rotation.rs:43-123) ---------------------------------------- */

/* Rust's f64::sin_cos is sin and cos of one argument as LLVM intrinsics, which LLVM lowers to one call of
 * glibc's sincos on x86_64 Linux; glibc's sincos and its separate sin and cos differ in the last bit on rare
 * arguments, so the call is spelled out here rather than left to the compiler's own fusion. */
void ppo_sin_cos(double theta, double* s, double* c) {
    const double snap = 1e-15;
    double sv, cv;
    sincos(theta, &sv, &cv);
    if (fabs(sv) < snap) {
        *s = 0.0;
        *c = copysign(1.0, cv);
    } else if (fabs(cv) < snap) {
        *s = copysign(1.0, sv);
        *c = 0.0;
    } else {
        *s = sv;
        *c = cv;
    }
}

/* Descending |c|^2 with NaN after every number (truncation/builtin.rs:101-104). */
static int ppo_cmp_magnitude_desc(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    int xn = x != x, yn = y != y;
    if (xn || yn) return xn - yn;
    return x > y ? -1 : (x < y ? 1 : 0);
}

/* ---- instantiations ------------------------------------------------------------------------- */

#define W 1
#define REAL double
#define SFX _w1_f64
#include "pp_oracle_tmpl.h"
#undef W
#undef REAL
#undef SFX

#define W 1
#define REAL float
#define SFX _w1_f32
#include "pp_oracle_tmpl.h"
#undef W
#undef REAL
#undef SFX

#define W 2
#define REAL double
#define SFX _w2_f64
#include "pp_oracle_tmpl.h"
#undef W
#undef REAL
#undef SFX

#define W 2
#define REAL float
#define SFX _w2_f32
#include "pp_oracle_tmpl.h"
#undef W
#undef REAL
#undef SFX

#define DISPATCH(s, call_f64_1, call_f32_1, call_f64_2, call_f32_2) \
    do {                                                            \
        if ((s)->W == 1 && (s)->kind == PPO_F64) { call_f64_1; }    \
        else if ((s)->W == 1) { call_f32_1; }                      \
        else if ((s)->kind == PPO_F64) { call_f64_2; }              \
        else { call_f32_2; }                                        \
    } while (0)

/* ---- sums ------------------------------------------------------------------------------------ */

ppo_sum* ppo_sum_new(int W_, int kind, int num_qubits, size_t n, const uint64_t* x, const uint64_t* z,
                     const void* coeff, uint64_t seed) {
    if ((W_ != 1 && W_ != 2) || (kind != PPO_F64 && kind != PPO_F32) || num_qubits < 0 || num_qubits > 64 * W_)
        return NULL;
    ppo_sum* s = (ppo_sum*)calloc(1, sizeof(ppo_sum));
    s->W = W_;
    s->kind = kind;
    s->num_qubits = num_qubits;
    s->seed = seed;
    uint64_t rx[PPO_INDEX_BITS * 4], rz[PPO_INDEX_BITS * 4];
    s->max_bits = ppo_hash_rows(num_qubits, seed, W_, rx, rz);
    for (int r = 0; r < s->max_bits; ++r)
        for (int w = 0; w < W_; ++w) {
            s->rows_x[r][w] = rx[r * W_ + w];
            s->rows_z[r][w] = rz[r * W_ + w];
        }
    DISPATCH(s, build_w1_f64(s, n, x, z, (const double*)coeff), build_w1_f32(s, n, x, z, (const float*)coeff),
             build_w2_f64(s, n, x, z, (const double*)coeff), build_w2_f32(s, n, x, z, (const float*)coeff));
    return s;
}

void ppo_sum_free(ppo_sum* s) {
    if (!s) return;
    DISPATCH(s, free_buckets_w1_f64(s), free_buckets_w1_f32(s), free_buckets_w2_f64(s), free_buckets_w2_f32(s));
    free(s);
}

ppo_sum* ppo_sum_clone(const ppo_sum* src) {
    ppo_sum* s = (ppo_sum*)malloc(sizeof(ppo_sum));
    *s = *src;
    s->buckets = NULL;
    DISPATCH(src, clone_buckets_w1_f64(src, s), clone_buckets_w1_f32(src, s), clone_buckets_w2_f64(src, s),
             clone_buckets_w2_f32(src, s));
    return s;
}

size_t ppo_sum_len(const ppo_sum* s) { return s->len; }
int ppo_sum_bits(const ppo_sum* s) { return s->bits; }

void ppo_sum_export(const ppo_sum* s, uint64_t* x, uint64_t* z, void* coeff, uint32_t* bucket) {
    DISPATCH(s, export_w1_f64(s, x, z, (double*)coeff, bucket), export_w1_f32(s, x, z, (float*)coeff, bucket),
             export_w2_f64(s, x, z, (double*)coeff, bucket), export_w2_f32(s, x, z, (float*)coeff, bucket));
}

double ppo_expectation_product(const ppo_sum* s, const uint64_t* ax, const uint64_t* az, const uint64_t* neg) {
    double v = 0.0;
    DISPATCH(s, v = expectation_w1_f64(s, ax, az, neg), v = expectation_w1_f32(s, ax, az, neg),
             v = expectation_w2_f64(s, ax, az, neg), v = expectation_w2_f32(s, ax, az, neg));
    return v;
}

void ppo_topn(ppo_sum* s, uint64_t n) {
    DISPATCH(s, topn_w1_f64(s, n), topn_w1_f32(s, n), topn_w2_f64(s, n), topn_w2_f32(s, n));
}

void ppo_approx_topn(ppo_sum* s, uint64_t n) {
    DISPATCH(s, approx_topn_w1_f64(s, n), approx_topn_w1_f32(s, n), approx_topn_w2_f64(s, n),
             approx_topn_w2_f32(s, n));
}

static void rebucket_to(ppo_sum* s, int want) {
    DISPATCH(s, rebucket_to_w1_f64(s, want), rebucket_to_w1_f32(s, want), rebucket_to_w2_f64(s, want),
             rebucket_to_w2_f32(s, want));
    s->known = 0;
}

/* bucket_of under `bits` active rows, for the width-independent driver. */
static uint32_t bucket_of_mask(const ppo_sum* s, const uint64_t* gx, const uint64_t* gz, int bits) {
    uint32_t b = 0;
    for (int r = 0; r < bits; ++r) {
        uint64_t f = 0;
        for (int w = 0; w < s->W; ++w) f ^= (gx[w] & s->rows_x[r][w]) ^ (gz[w] & s->rows_z[r][w]);
        b |= (uint32_t)(__builtin_popcountll(f) & 1) << r;
    }
    return b;
}

/* ---- one prepared channel ------------------------------------------------------------------- */

typedef struct {
    double sin, cos;
    int weight;        /* generator weight */
    int key_preserving; /* only the identity delta (sine snapped to zero, or identity generator) */
    int clifford;      /* a rotation table whose identity residue is dropped */
    int local;         /* Prepared::Local (weight <= 2) rather than Prepared::Rotation */
} prepared_t;

static prepared_t prepare(const ppo_sum* s, const ppo_channel* ch, int adjoint) {
    prepared_t p;
    double theta = adjoint ? -ch->theta : ch->theta;
    ppo_sin_cos(theta, &p.sin, &p.cos);
    int wt = 0;
    for (int w = 0; w < s->W; ++w) wt += __builtin_popcountll(ch->gx[w] | ch->gz[w]);
    p.weight = wt;
    p.local = wt <= 2;
    /* channel/prepared.rs:257-281: the generator entry exists iff some amplitude is nonzero */
    p.key_preserving = p.local && (wt == 0 || p.sin == 0.0);
    /* rotation_layer.rs:120 and :277: |a0|^2 <= eps^2 |a1|^2 on every anticommuting pattern */
    double a0 = s->kind == PPO_F64 ? p.cos : (double)(float)p.cos;
    double a1 = s->kind == PPO_F64 ? p.sin : (double)(float)p.sin;
    const double eps = 2.220446049250313e-16;
    p.clifford = !p.key_preserving && (a0 * a0 <= eps * eps * (a1 * a1));
    return p;
}

/* max_fanout (engine/bucketed.rs:628-641) */
static size_t fanout(const prepared_t* p) {
    if (!p->local) return 2;
    if (p->key_preserving || p->clifford) return 1;
    return 2;
}

/* PermutationStep::of (engine/bucketed/permutation_run.rs:53-121): a Clifford rotation table,
 * or a key-preserving table on a nonempty support. */
static int is_permutation_step(const prepared_t* p) {
    if (!p->local || p->weight == 0) return 0;
    return p->key_preserving || p->clifford;
}

/* RotationLayer::of is Some for a Local table with an identity and one generator entry. */
static int is_rotation_table(const prepared_t* p) { return p->local && !p->key_preserving; }

/* Rank over GF(2) of bucket deltas (engine/coset.rs:95-160). */
static int span_dimension(const uint32_t* deltas, int n) {
    uint32_t basis[32];
    int nb = 0;
    for (int i = 0; i < n; ++i) {
        uint32_t r = deltas[i];
        for (int j = 0; j < nb; ++j) {
            uint32_t hb = 1u << (31 - __builtin_clz(basis[j]));
            if (r & hb) r ^= basis[j];
        }
        if (!r) continue;
        /* keep the basis with distinct leading bits */
        basis[nb++] = r;
        for (int a = 0; a < nb; ++a)
            for (int b = a + 1; b < nb; ++b)
                if (basis[b] > basis[a]) {
                    uint32_t t = basis[a];
                    basis[a] = basis[b];
                    basis[b] = t;
                }
        for (int j = 0; j < nb; ++j)
            for (int k = 0; k < nb; ++k) {
                if (j == k) continue;
                uint32_t hb = 1u << (31 - __builtin_clz(basis[j]));
                if (basis[k] & hb) basis[k] ^= basis[j];
            }
    }
    return nb;
}

static void apply_prepared(ppo_sum* s, const ppo_channel* ch, const prepared_t* p, const ppo_policy* pol,
                           int threads) {
    if (p->key_preserving) {
        DISPATCH(s, rescale_layer_w1_f64(s, ch->gx, ch->gz, p->cos, pol),
                 rescale_layer_w1_f32(s, ch->gx, ch->gz, p->cos, pol),
                 rescale_layer_w2_f64(s, ch->gx, ch->gz, p->cos, pol),
                 rescale_layer_w2_f32(s, ch->gx, ch->gz, p->cos, pol));
        return;
    }
    DISPATCH(s, rotation_layer_w1_f64(s, ch->gx, ch->gz, p->cos, p->sin, p->clifford, pol, threads),
             rotation_layer_w1_f32(s, ch->gx, ch->gz, p->cos, p->sin, p->clifford, pol, threads),
             rotation_layer_w2_f64(s, ch->gx, ch->gz, p->cos, p->sin, p->clifford, pol, threads),
             rotation_layer_w2_f32(s, ch->gx, ch->gz, p->cos, p->sin, p->clifford, pol, threads));
}

/* ---- the layer loop (engine/driver.rs:319-592 at one partition) ------------------------------ */

static size_t run_length_limit(const ppo_policy* pol) {
    /* TopN::run_length_limit and ApproxTopN's (builtin.rs:147-149, :462-464) */
    if (pol && pol->has_topn) return (size_t)pol->topn;
    return (size_t)-1;                                    /* no layer pass: usize::MAX */
}

int ppo_propagate(ppo_sum* s, const ppo_channel* channels, size_t n_channels, int direction,
                  const ppo_policy* pol, int target, int min_buckets, int threads, uint64_t* layer_len) {
    if (target < 1 || min_buckets < 1) return -1;
    int adjoint = direction == 1;
    size_t limit = run_length_limit(pol);
    int finalizes = pol && pol->has_topn;
    s->known = 0; /* begin_call */
    size_t i = 0;
    while (i < n_channels) {
        const ppo_channel* ch = &channels[adjoint ? n_channels - 1 - i : i];
        /* refine_unprepared (engine/backend.rs:155-165) */
        int want = ppo_rebucketed_bits(s->len, s->bits, (size_t)target, (size_t)min_buckets);
        if (want > s->max_bits) want = s->max_bits;
        if (want != s->bits) rebucket_to(s, want);
        prepared_t p = prepare(s, ch, adjoint);
        /* Runs::take (engine/driver.rs:202-258): a run applies its layers under this bucket count,
         * with no rebucket between them and the layer pass once at its end. */
        size_t end = i + 1;
        int run = 0;
        if (limit > 0) {
            if (s->len <= limit && is_permutation_step(&p)) {
                size_t j = i + 1;
                while (j < n_channels) {
                    prepared_t q = prepare(s, &channels[adjoint ? n_channels - 1 - j : j], adjoint);
                    if (!is_permutation_step(&q)) break;
                    ++j;
                }
                if (j > i + 1) {
                    end = j;
                    run = 1;
                }
            }
            if (!run) {
                size_t worst = s->len * fanout(&p);
                if (s->len != 0 && worst / s->len != fanout(&p)) worst = (size_t)-1;
                uint32_t deltas[2 * PPO_MAX_ROTATION_RUN_LAYERS];
                int nd = 0;
                if (worst <= limit && is_rotation_table(&p)) {
                    deltas[nd++] = 0;
                    deltas[nd++] = bucket_of_mask(s, ch->gx, ch->gz, s->bits);
                    if (span_dimension(deltas, nd) <= PPO_MAX_ROTATION_RUN_DIMENSION) {
                        size_t j = i + 1;
                        while (j < n_channels && j < i + PPO_MAX_ROTATION_RUN_LAYERS) {
                            const ppo_channel* cj = &channels[adjoint ? n_channels - 1 - j : j];
                            prepared_t q = prepare(s, cj, adjoint);
                            size_t f = fanout(&q);
                            size_t grown = worst * f;
                            if (worst != 0 && grown / worst != f) grown = (size_t)-1;
                            if (grown > limit || !is_rotation_table(&q)) break;
                            deltas[nd] = 0;
                            deltas[nd + 1] = bucket_of_mask(s, cj->gx, cj->gz, s->bits);
                            if (span_dimension(deltas, nd + 2) > PPO_MAX_ROTATION_RUN_DIMENSION) break;
                            nd += 2;
                            worst = grown;
                            ++j;
                        }
                        if (j > i + 1) {
                            end = j;
                            run = 1;
                        }
                    }
                }
            }
        }
        for (size_t j = i; j < end; ++j) {
            const ppo_channel* cj = &channels[adjoint ? n_channels - 1 - j : j];
            prepared_t q = j == i ? p : prepare(s, cj, adjoint);
            apply_prepared(s, cj, &q, pol, threads);
            if (layer_len) layer_len[j] = s->len;
        }
        if (finalizes) {
            if (pol->has_topn == 2) ppo_approx_topn(s, pol->topn);
            else ppo_topn(s, pol->topn);
            if (layer_len) layer_len[end - 1] = s->len;
        }
        i = end;
    }
    s->known = 0;
    return 0;
}

int ppo_apply_rotation(ppo_sum* s, const ppo_channel* ch, int adjoint, const ppo_policy* pol, int threads) {
    return ppo_propagate(s, ch, 1, adjoint ? 1 : 0, pol, 1024, 128, threads, NULL);
}

size_t ppo_naive_layer(int W_, int kind, size_t n, const uint64_t* x, const uint64_t* z, const void* coeff,
                       const ppo_channel* ch, int adjoint, const ppo_policy* pol, uint64_t* ox, uint64_t* oz,
                       void* oc) {
    double sv, cv;
    ppo_sin_cos(adjoint ? -ch->theta : ch->theta, &sv, &cv);
    if (W_ == 1 && kind == PPO_F64)
        return naive_w1_f64(n, x, z, (const double*)coeff, ch->gx, ch->gz, sv, cv, pol, ox, oz, (double*)oc);
    if (W_ == 1)
        return naive_w1_f32(n, x, z, (const float*)coeff, ch->gx, ch->gz, sv, cv, pol, ox, oz, (float*)oc);
    if (kind == PPO_F64)
        return naive_w2_f64(n, x, z, (const double*)coeff, ch->gx, ch->gz, sv, cv, pol, ox, oz, (double*)oc);
    return naive_w2_f32(n, x, z, (const float*)coeff, ch->gx, ch->gz, sv, cv, pol, ox, oz, (float*)oc);
}
