// pp_topn.h -- TopN as a radix select on the bits of |c|^2 (truncation/builtin.rs:73-129; the reference's device
// form is engine/gpu/finalize.rs:238-409 with kern/truncate.cu:31-141).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_device.h"

namespace pp {

// Device-resident state of one selection, advanced by k_topn_select after every histogram pass.
struct TopnState {
    unsigned long long prefix;   // the digits of the n-th largest |c|^2 found so far
    unsigned long long mask;     // which bits of prefix are decided
    unsigned long long k;        // rank still sought among the terms matching prefix (1 = the largest)
    unsigned long long above;    // terms strictly above the boundary found so far
    unsigned long long finite;   // terms whose |c|^2 is not NaN
    unsigned long long ties;     // terms at the final boundary
    unsigned long long keep_all_finite;  // at most n finite terms: keep every one of them
    unsigned long long gathered;  // tie terms written by k_topn_gather_ties
    unsigned long long done;      // digits decided so far (a pass whose digit is decided does nothing)
    unsigned long long seed;      // a seeded selection: the previous selection's prefix (its first two digits)
    unsigned long long seeded;
};

__device__ __forceinline__ bool is_nan_bits(uint64_t u) { return (u & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull; }

template <typename T>
__device__ __forceinline__ uint64_t magnitude_bits(T c) {
    const double v = (double)c;  // norm_sqr widens to f64 first (coefficient.rs:121-125)
    return (uint64_t)__double_as_longlong(v * v);
}

// A float score as a key in the high half of a u64 that orders as the score (a NaN score sorts lowest): the learned
// TopN pass ranks terms by their score through the same selection as TopN.
__device__ __forceinline__ uint64_t score_bits(float s) {
    if (s != s) return 0ull;
    const uint32_t u = __float_as_uint(s);
    const uint32_t k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return (uint64_t)k << 32;
}

// Whether a term whose ranking key is u (at column position i) survives a finished selection: above the boundary
// bin on the decided bits, or in it and flagged (without flags, every term in it); with at most n finite terms every
// finite one; a NaN |c|^2 never (a score key is never NaN-coded: score_nan marks the |c| case).
__device__ __forceinline__ bool topn_keeps(const TopnState& st, uint64_t u, bool score_nan,
                                           const uint8_t* tie_keep, uint64_t i) {
    if (st.keep_all_finite) return !score_nan;
    if (score_nan) return false;
    const uint64_t um = u & st.mask;
    if (um > st.prefix) return true;
    if (um == st.prefix) return tie_keep == nullptr || tie_keep[i] != 0;
    return false;
}

constexpr int kTopnDigitBits = 11;
constexpr int kTopnBins = 1 << kTopnDigitBits;
constexpr int kTopnPasses = 6;  // digits at shifts 53, 42, 31, 20, 9 (11 bits) and 0 (9 bits)
constexpr uint32_t kHistBucketsMax = 256;  // buckets per histogram workgroup (one length per thread)
// Copies of the global histogram: workgroup w adds its counts to copy w % kHistCopies, so that the ~2048 workgroups'
// same-address atomics on a few occupied bins spread over kHistCopies addresses; the digit select sums the copies.
constexpr uint32_t kHistCopies = 4;
constexpr uint32_t kHistWords = kHistCopies * kTopnBins;

__host__ __device__ inline int topn_shift(int pass) { return pass < 5 ? 53 - 11 * pass : 0; }
__host__ __device__ inline int topn_bits(int pass) { return pass < 5 ? 11 : 9; }

template <typename T>
struct TopnHistArgs {
    const T* c;
    const float* score;  // when set, the ranking key is this score (largest kept) instead of |c|^2
    const uint64_t* start;
    const uint32_t* len;
    uint32_t num_buckets;
    uint32_t buckets_per_block;
    int32_t pass;
    const TopnState* state;
    unsigned long long* hist;  // kHistWords (kHistCopies copies of kTopnBins), zero on entry (k_topn_select clears
                               // it after each pass)
};

template <int W, typename T>
struct TopnTieArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    const float* score;  // when set, the ranking key is this score (largest kept) instead of |c|^2
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    const uint64_t* rows;
    int32_t ref_bits;
    TopnState* state;
    uint64_t* tie_x;  // term-major [ties][W]
    uint64_t* tie_z;
    uint32_t* tie_bucket;
    uint64_t* tie_pos;
    uint64_t* tie_rank;  // the rank key of each gathered term (the boundary bin's terms need not be equal)
    uint64_t tie_cap;
    uint32_t num_buckets;
    uint32_t buckets_per_block;  // at most kHistBucketsMax
};

template <int W, typename T>
struct TopnRetainArgs {
    uint64_t* x;
    uint64_t* z;
    T* c;
    const float* score;  // when set, the ranking key is this score (largest kept) instead of |c|^2
    uint64_t cap;
    const uint64_t* start;
    uint32_t* len;
    const TopnState* state;
    const uint8_t* tie_keep;  // per column position; null keeps every tie
    unsigned long long* stats;
};

// Many selections at once (a batch of sums, each with its own state and histogram): an item per sum.
struct TopnSelectItem {
    TopnState* state;
    unsigned long long* hist;  // kHistWords
    uint64_t n;                // terms to keep
    TopnState* out;            // the state's copy after the last pass
};
// The boundary flags of one selection: its boundary bin's positions in rank order, the first `kept` of them kept.
struct TopnFlagItem {
    const uint64_t* pos;
    uint64_t ties;
    uint64_t kept;
    uint8_t* flags;
};
hipError_t launch_topn_init_batch(hipStream_t st, const TopnSelectItem* items, uint32_t num_items);
// block_start: exclusive prefix of the items' histogram workgroups (num_items + 1 entries, blocks in all).
template <typename T>
hipError_t launch_topn_hist_batch(hipStream_t st, const TopnHistArgs<T>* items, const uint32_t* block_start,
                                  uint32_t num_items, uint32_t blocks, int pass);
hipError_t launch_topn_select_batch(hipStream_t st, const TopnSelectItem* items, uint32_t num_items, int pass,
                                    bool last);
template <int W, typename T>
hipError_t launch_topn_gather_batch(hipStream_t st, const TopnTieArgs<W, T>* items, const uint32_t* block_start,
                                    uint32_t num_items, uint32_t blocks);
hipError_t launch_topn_flags_batch(hipStream_t st, const TopnFlagItem* items, uint32_t num_items);

template <typename T>
hipError_t launch_topn_hist(hipStream_t st, const TopnHistArgs<T>& a, uint32_t blocks);
hipError_t launch_topn_select(hipStream_t st, TopnState* state, unsigned long long* hist, int pass, uint64_t n);
// A new selection's state, its histogram (hist_words words) cleared.
hipError_t launch_topn_init(hipStream_t st, TopnState* state, uint64_t n, unsigned long long* hist,
                            uint32_t hist_words, uint64_t seed = 0, bool seeded = false);
// The seeded first pass: besides the first digit's histogram, the second digit's over the terms whose first digit is
// the seed's; the select then decides one or two digits (the next pass does nothing for a digit decided).  hist holds
// kSeededHistWords.
constexpr uint32_t kSeededHistWords = 2 * kHistWords;
template <typename T>
hipError_t launch_topn_hist_seeded(hipStream_t st, const TopnHistArgs<T>& a, uint32_t blocks);
hipError_t launch_topn_select_seeded(hipStream_t st, TopnState* state, unsigned long long* hist, uint64_t n);
template <int W, typename T>
hipError_t launch_topn_gather_ties(hipStream_t st, const TopnTieArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
hipError_t launch_topn_retain(hipStream_t st, const TopnRetainArgs<W, T>& a, uint32_t num_buckets);
hipError_t launch_mark_positions(hipStream_t st, const uint64_t* pos, uint64_t n, uint8_t* flags);
// flags[pos[j]] = j < kept for j < ties: a boundary bin's positions in rank order.  Only positions in the boundary
// bin are read (topn_keeps), so no other flag needs a value.
hipError_t launch_topn_flag_ties(hipStream_t st, const uint64_t* pos, uint64_t ties, uint64_t kept, uint8_t* flags);

// ---- ApproxTopN (truncation/builtin.rs:241-468): the cut at an octave of |c|^2 ----------------------------------

constexpr int kOctaveBins = 2048;  // APPROX_BINS, one per f64 exponent (builtin.rs:241-242)

template <typename T>
struct OctaveHistArgs {
    const T* c;
    const uint64_t* start;
    const uint32_t* len;
    uint32_t num_buckets;
    uint32_t buckets_per_block;
    unsigned long long* hist;  // kOctaveBins
};

// The boundary octave's terms: counted per reference bucket, gathered from one reference bucket, and the in-place
// retain that keeps every term above the octave and the boundary terms the quota reaches.
template <int W, typename T>
struct OctaveCutArgs {
    uint64_t* x;
    uint64_t* z;
    T* c;
    uint64_t cap;
    const uint64_t* start;
    uint32_t* len;
    const uint64_t* rows;
    int32_t ref_bits;
    uint32_t bin;           // the boundary octave
    uint32_t split_bucket;  // boundary terms of reference buckets below it are kept; 1 << ref_bits keeps all
    uint32_t* counts;       // k_octave_counts: boundary terms per reference bucket
    const uint8_t* keep;    // k_octave_retain: per column position, which of split_bucket's boundary terms stay
    uint64_t* out_x;        // k_octave_gather: split_bucket's boundary terms, term-major [n][W], and positions
    uint64_t* out_z;
    uint64_t* out_pos;
    unsigned long long* gathered;
    uint64_t out_cap;
};

template <typename T>
hipError_t launch_octave_hist(hipStream_t st, const OctaveHistArgs<T>& a, uint32_t blocks);
template <int W, typename T>
hipError_t launch_octave_counts(hipStream_t st, const OctaveCutArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
hipError_t launch_octave_gather(hipStream_t st, const OctaveCutArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
hipError_t launch_octave_retain(hipStream_t st, const OctaveCutArgs<W, T>& a, uint32_t num_buckets);

}  // namespace pp
