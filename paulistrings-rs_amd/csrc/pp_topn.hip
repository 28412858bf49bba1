// pp_topn.hip -- TopN kernels: a radix select on the IEEE bits of |c|^2 (non-negative doubles order as their bits;
// three 11-bit passes, six for a crowded boundary bin), the gather of the boundary bin's terms, and the in-place
// retain.
//
// Reference semantics (truncation/builtin.rs:73-129, :152-239): with t the n-th largest |c|^2, keep every term above
// t and the first n - above of the terms at t in canonical order (the reference's bucket ascending, then the
// lexicographic key); a NaN is never kept, and with at most n other terms every non-NaN term is kept.
#include "pp_topn.h"

namespace pp {

template <typename T>
__device__ __forceinline__ uint64_t rank_bits(const T* c, const float* score, uint64_t i) {
    return score ? score_bits(score[i]) : magnitude_bits<T>(c[i]);
}

// The state of a new selection, and (one workgroup) its histogram cleared: ApproxTopN's octave pass, which shares
// it, leaves its counts behind.
__global__ void __launch_bounds__(kBlock) k_topn_init(TopnState* s, uint64_t n, uint64_t seed, uint32_t seeded,
                                                      unsigned long long* hist, uint32_t hist_words) {
    for (uint32_t i = threadIdx.x; i < hist_words; i += kBlock) hist[i] = 0ull;
    if (threadIdx.x != 0) return;
    s->prefix = 0;
    s->mask = 0;
    s->k = n;
    s->above = 0;
    s->finite = 0;
    s->ties = 0;
    s->keep_all_finite = 0;
    s->gathered = 0;
    s->done = 0;
    s->seed = seed;
    s->seeded = seeded;
}

// The digit holding the k-th largest after pass `pass`, from the complete histogram, run by one workgroup; then the
// histogram is cleared for the next pass.  The digit is the largest d whose suffix count sum_{e >= d} hist[e]
// reaches k: each thread owns kPerThread consecutive bins, a suffix scan of the per-thread sums (in LDS) finds the
// one thread whose bins hold d, and that thread walks them.
__device__ __forceinline__ void select_digit(TopnState* s, unsigned long long* hist, int pass, uint64_t n) {
    constexpr int kPerThread = kTopnBins / kBlock;
    static_assert(kPerThread * kBlock == kTopnBins, "bins split evenly over the workgroup");
    __shared__ unsigned long long suffix[kBlock];
    __shared__ int keep_all;
    __shared__ unsigned long long k_sh;  // k as every thread must see it, read before the holder updates it
    const int t = threadIdx.x;
    const int bins = 1 << topn_bits(pass);
    unsigned long long mine[kPerThread];
    unsigned long long sum = 0;
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const int d = t * kPerThread + j;
        mine[j] = 0ull;
        if (d < bins) {
#pragma unroll
            for (uint32_t c = 0; c < kHistCopies; ++c) mine[j] += hist[c * kTopnBins + d];
        }
        sum += mine[j];
    }
    suffix[t] = sum;
    __syncthreads();
    // inclusive suffix scan: suffix[t] = sum over threads u >= t
    for (int off = 1; off < kBlock; off <<= 1) {
        const unsigned long long v = t + off < kBlock ? suffix[t + off] : 0ull;
        __syncthreads();
        suffix[t] += v;
        __syncthreads();
    }
    if (t == 0) {
        int ka = s->keep_all_finite;
        if (!ka && pass == 0) {
            s->finite = suffix[0];
            if (suffix[0] <= n) {
                s->keep_all_finite = 1;
                ka = 1;
            }
        }
        keep_all = ka;
        k_sh = s->k;
    }
    __syncthreads();
    if (!keep_all) {
        const unsigned long long k = k_sh;
        const unsigned long long above_mine = t + 1 < kBlock ? suffix[t + 1] : 0ull;  // bins of higher threads
        // the thread holding the digit; with no digit reaching k (not reached in a selection) thread 0 takes bin 0
        const bool holder = (above_mine < k && above_mine + sum >= k) || (t == 0 && suffix[0] < k);
        if (holder) {
            unsigned long long acc = above_mine, at = mine[0];
            int chosen = t * kPerThread;
            bool found = false;
#pragma unroll
            for (int j = kPerThread - 1; j >= 0; --j) {
                if (!found && acc + mine[j] >= k) {
                    chosen = t * kPerThread + j;
                    at = mine[j];
                    found = true;
                }
                if (!found) acc += mine[j];
            }
            const int shift = topn_shift(pass);
            s->above += acc;
            s->k -= acc;
            s->prefix |= (unsigned long long)chosen << shift;
            s->mask |= ((1ull << topn_bits(pass)) - 1ull) << shift;
            s->ties = at;
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < kHistWords; i += kBlock) hist[i] = 0;
}

// One radix pass: each workgroup histograms the digit of its buckets' terms that match the prefix found so far and
// adds its counts to the global histogram; k_topn_select picks the digit.  The workgroup walks its buckets' rows
// as one list, so that its threads' loads do not wait bucket by bucket.
// The rows of buckets [b_lo, b_lo + nb) (nb <= kHistBucketsMax) as one list: first[j] rows precede bucket j, whose
// segment starts at seg[j]; returns the rows in all.  Every thread of the workgroup calls it (two barriers).
struct RowList {
    uint32_t first[kHistBucketsMax + 1];
    uint64_t seg[kHistBucketsMax];
    uint32_t wave_rows[kWaves];
};

__device__ __forceinline__ uint32_t list_rows(const uint64_t* start, const uint32_t* len, uint32_t b_lo, uint32_t nb,
                                              RowList& l) {
    const uint32_t t = threadIdx.x;
    uint32_t v = t < nb ? len[b_lo + t] : 0u;
    if (t < nb) l.seg[t] = start[b_lo + t];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane_id() >= (uint32_t)d) v += o;
    }
    if (lane_id() == 63) l.wave_rows[wave_id()] = v;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        before += (uint32_t)w < wave_id() ? l.wave_rows[w] : 0u;
        total += l.wave_rows[w];
    }
    if (t < nb) l.first[t + 1] = before + v;
    if (t == 0) l.first[0] = 0;
    __syncthreads();
    return total;
}

// Column positions of list rows g0 + q * kBlock (q < 4) that exist; j is the caller's bucket cursor (rows only move
// forward, so it does too).
__device__ __forceinline__ void list_positions(const RowList& l, uint32_t total, uint32_t g0, uint32_t& j,
                                               uint64_t (&pos)[4], bool (&in)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t g = g0 + (uint32_t)q * kBlock;
        in[q] = g < total;
        if (in[q]) {
            while (l.first[j + 1] <= g) ++j;
            pos[q] = l.seg[j] + (g - l.first[j]);
        }
    }
}

// One histogram workgroup's share (workgroup `block` of the selection's grid).
template <typename T>
__device__ __forceinline__ void topn_hist_body(const TopnHistArgs<T>& a, uint32_t block) {
    __shared__ uint32_t h[kTopnBins];
    __shared__ RowList list;
    const TopnState st = *a.state;
    if (st.done > (unsigned long long)a.pass) return;  // decided by a seeded first pass: the whole workgroup
    for (int i = threadIdx.x; i < kTopnBins; i += kBlock) h[i] = 0;
    const uint32_t b_lo = block * a.buckets_per_block;
    const uint32_t nb = min(a.num_buckets, b_lo + a.buckets_per_block) - b_lo;  // <= kHistBucketsMax
    const uint32_t t = threadIdx.x;
    const uint32_t total = list_rows(a.start, a.len, b_lo, nb, list);
    if (!st.keep_all_finite) {
        const int shift = topn_shift(a.pass);
        const uint64_t dmask = (1ull << topn_bits(a.pass)) - 1ull;
        // Row g of the list is row g - first[j] of bucket j; a thread's rows only move forward, so j does too.
        // Four rows per thread per step, their loads issued together.
        uint32_t j = 0;
        for (uint32_t g0 = t; g0 < total; g0 += 4 * kBlock) {
            uint64_t pos[4];
            bool in[4];
            list_positions(list, total, g0, j, pos, in);
            uint64_t u[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (in[q]) u[q] = rank_bits<T>(a.c, a.score, pos[q]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (!in[q]) continue;
                const bool match = (a.score || !is_nan_bits(u[q])) && (u[q] & st.mask) == st.prefix;
                if (match) atomicAdd(&h[(uint32_t)((u[q] >> shift) & dmask)], 1u);
            }
        }
    }
    __syncthreads();
    unsigned long long* const copy = a.hist + (block % kHistCopies) * kTopnBins;
    for (int i = threadIdx.x; i < kTopnBins; i += kBlock)
        if (h[i]) atomicAdd(&copy[i], (unsigned long long)h[i]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_topn_hist(const TopnHistArgs<T> a) {
    topn_hist_body<T>(a, blockIdx.x);
}

// The item of a batched launch whose range [block_start[i], block_start[i + 1]) of workgroups holds blockIdx.x.
__device__ __forceinline__ uint32_t batch_item(const uint32_t* block_start, uint32_t num_items) {
    uint32_t lo = 0, hi = num_items;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (block_start[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Pass `pass` of many selections in one launch (items: each sum's histogram arguments).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_topn_hist_batch(const TopnHistArgs<T>* __restrict__ items,
                                                             const uint32_t* __restrict__ block_start,
                                                             uint32_t num_items, int32_t pass) {
    const uint32_t i = batch_item(block_start, num_items);
    TopnHistArgs<T> a = items[i];
    a.pass = pass;
    topn_hist_body<T>(a, blockIdx.x - block_start[i]);
}

// One workgroup after each histogram pass (a kernel boundary, not a device-scope fence per histogram workgroup: on
// a multi-XCD device such a fence writes back the XCD's L2).
__global__ void __launch_bounds__(kBlock) k_topn_select(TopnState* s, unsigned long long* hist, int pass, uint64_t n) {
    if (s->done > (unsigned long long)pass) return;  // uniform: one load, the same for every thread
    select_digit(s, hist, pass, n);
    if (threadIdx.x == 0) s->done = (unsigned long long)pass + 1ull;
}

// Many selections' states: one workgroup each clears its histogram's first region and starts the state (unseeded).
__global__ void __launch_bounds__(kBlock) k_topn_init_batch(const TopnSelectItem* __restrict__ items) {
    const TopnSelectItem it = items[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < kHistWords; i += kBlock) it.hist[i] = 0ull;
    if (threadIdx.x == 0) {
        TopnState* s = it.state;
        s->prefix = 0;
        s->mask = 0;
        s->k = it.n;
        s->above = 0;
        s->finite = 0;
        s->ties = 0;
        s->keep_all_finite = 0;
        s->gathered = 0;
        s->done = 0;
        s->seed = 0;
        s->seeded = 0;
    }
}

// The digit select of pass `pass` for many selections, one workgroup each; with `last`, each state is also copied
// to its item's slot of the batch's contiguous state array (one read-back for all of them).
__global__ void __launch_bounds__(kBlock) k_topn_select_batch(const TopnSelectItem* __restrict__ items, int pass,
                                                              int last) {
    const TopnSelectItem it = items[blockIdx.x];
    TopnState* s = it.state;
    if (s->done <= (unsigned long long)pass) {  // uniform: one load, the same for every thread
        select_digit(s, it.hist, pass, it.n);
        if (threadIdx.x == 0) s->done = (unsigned long long)pass + 1ull;
    }
    if (last) {
        __syncthreads();  // the holder's updates of the state
        if (threadIdx.x == 0) *it.out = *s;
    }
}

// The seeded first pass's histograms: region 0 of hist counts the first digit of every term, region 1 the second
// digit of the terms whose first digit is the seed's.  (A third level, over the terms whose first two digits are the
// seed's, measured slower: at a TopN cap the second digit of the boundary moves from layer to layer, and the third
// histogram's LDS cost the pass more than it saved.)  Walks its buckets as k_topn_hist does.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_topn_hist_seeded(const TopnHistArgs<T> a) {
    __shared__ uint32_t h[2][kTopnBins];
    __shared__ RowList list;
    for (int i = threadIdx.x; i < 2 * kTopnBins; i += kBlock) (&h[0][0])[i] = 0;
    const TopnState st = *a.state;
    const uint32_t spec0 = (uint32_t)(st.seed >> topn_shift(0)) & (kTopnBins - 1);
    const uint32_t b_lo = blockIdx.x * a.buckets_per_block;
    const uint32_t nb = min(a.num_buckets, b_lo + a.buckets_per_block) - b_lo;  // <= kHistBucketsMax
    const uint32_t total = list_rows(a.start, a.len, b_lo, nb, list);
    uint32_t j = 0;
    for (uint32_t g0 = threadIdx.x; g0 < total; g0 += 4 * kBlock) {
        uint64_t pos[4];
        bool in[4];
        list_positions(list, total, g0, j, pos, in);
        uint64_t u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (in[q]) u[q] = rank_bits<T>(a.c, a.score, pos[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!in[q] || (!a.score && is_nan_bits(u[q]))) continue;
            const uint32_t d0 = (uint32_t)(u[q] >> topn_shift(0)) & (kTopnBins - 1);
            atomicAdd(&h[0][d0], 1u);
            if (d0 != spec0) continue;
            atomicAdd(&h[1][(uint32_t)(u[q] >> topn_shift(1)) & (kTopnBins - 1)], 1u);
        }
    }
    __syncthreads();
    const uint32_t copy = (blockIdx.x % kHistCopies) * kTopnBins;
    for (int i = threadIdx.x; i < 2 * kTopnBins; i += kBlock) {
        const uint32_t v = (&h[0][0])[i];
        if (v) atomicAdd(&a.hist[(i / kTopnBins) * kHistWords + copy + (i % kTopnBins)], (unsigned long long)v);
    }
}

// The first digit from region 0, the second from region 1 when the first is the seed's; region 1 is cleared for the
// next selection when unused.
__global__ void __launch_bounds__(kBlock) k_topn_select_seeded(TopnState* s, unsigned long long* hist, uint64_t n) {
    __shared__ int go;
    select_digit(s, hist, 0, n);
    int done = 1;
    for (int r = 1; r < 2; ++r) {
        __syncthreads();  // the holder's update of the state is visible
        if (threadIdx.x == 0) {
            const uint64_t d = (s->prefix >> topn_shift(r - 1)) & (kTopnBins - 1);
            go = done == r && !s->keep_all_finite && d == ((s->seed >> topn_shift(r - 1)) & (kTopnBins - 1));
        }
        __syncthreads();
        if (go) {
            select_digit(s, hist + (size_t)r * kHistWords, r, n);
            done = r + 1;
        } else {
            for (uint32_t i = threadIdx.x; i < kHistWords; i += kBlock) hist[(size_t)r * kHistWords + i] = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) s->done = s->keep_all_finite ? 3ull : (unsigned long long)done;
}

// The terms of the boundary bin: rank key equal to the prefix on the decided bits (a NaN never matches a finite
// prefix: its exponent bits are all ones).  Buckets are walked as in k_topn_hist.
template <int W, typename T>
__device__ __forceinline__ void topn_gather_body(const TopnTieArgs<W, T>& a, uint32_t block) {
    __shared__ RowList list;
    const uint64_t t = a.state->prefix, mask = a.state->mask;
    const uint32_t b_lo = block * a.buckets_per_block;
    const uint32_t nb = min(a.num_buckets, b_lo + a.buckets_per_block) - b_lo;  // <= kHistBucketsMax
    const uint32_t total = list_rows(a.start, a.len, b_lo, nb, list);
    uint32_t j = 0;
    for (uint32_t g0 = threadIdx.x; g0 < total; g0 += 4 * kBlock) {
        uint64_t pos[4];
        bool in[4];
        list_positions(list, total, g0, j, pos, in);
        uint64_t u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (in[q]) u[q] = rank_bits<T>(a.c, a.score, pos[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!in[q] || (u[q] & mask) != t || (!a.score && is_nan_bits(u[q]))) continue;
            const uint64_t slot = atomicAdd(&a.state->gathered, 1ull);
            if (slot >= a.tie_cap) continue;
            const Key<W> k = load_key<W>(a.x, a.z, a.cap, pos[q]);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                a.tie_x[slot * W + w] = k.x[w];
                a.tie_z[slot * W + w] = k.z[w];
            }
            a.tie_bucket[slot] = bucket_of<W>(k, a.rows, a.ref_bits);
            a.tie_pos[slot] = pos[q];
            a.tie_rank[slot] = u[q];
        }
    }
}

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_topn_gather_ties(const TopnTieArgs<W, T> a) {
    topn_gather_body<W, T>(a, blockIdx.x);
}

// The boundary bins of many selections in one launch (each item's ties into its own range of the batch's arrays).
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_topn_gather_batch(const TopnTieArgs<W, T>* __restrict__ items,
                                                              const uint32_t* __restrict__ block_start,
                                                              uint32_t num_items) {
    const uint32_t i = batch_item(block_start, num_items);
    const TopnTieArgs<W, T> a = items[i];
    topn_gather_body<W, T>(a, blockIdx.x - block_start[i]);
}

// Whether the term at column position i (coefficient c) survives the selection: above the boundary bin on the
// decided bits, or in it and flagged (without flags, every term in it).
template <int W, typename T>
__device__ __forceinline__ bool retain_keeps(const TopnRetainArgs<W, T>& a, const TopnState& st, T c, uint64_t i) {
    const uint64_t u = a.score ? score_bits(a.score[i]) : magnitude_bits<T>(c);
    return topn_keeps(st, u, !a.score && is_nan_bits(u), a.tie_keep, i);
}

// In-place retain, 2 kBlock rows per step: both rows' loads are issued before the step's one scan, and a dropped
// row's key is never read.  Every write lands at or before its row's own position, and every row of the step was
// read before the scan's barrier.
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_topn_retain(const TopnRetainArgs<W, T> a) {
    __shared__ uint32_t scan_scratch[2 * kWaves];
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    const TopnState st = *a.state;
    uint32_t cur = 0;
    for (uint32_t base = 0; base < n; base += 2 * kBlock) {
        const uint32_t r0 = base + threadIdx.x, r1 = r0 + kBlock;
        T c0 = T(0), c1 = T(0);
        if (r0 < n) c0 = a.c[s + r0];
        if (r1 < n) c1 = a.c[s + r1];
        const bool keep0 = r0 < n && retain_keeps<W, T>(a, st, c0, s + r0);
        const bool keep1 = r1 < n && retain_keeps<W, T>(a, st, c1, s + r1);
        Key<W> k0, k1;
        if (keep0) k0 = load_key<W>(a.x, a.z, a.cap, s + r0);
        if (keep1) k1 = load_key<W>(a.x, a.z, a.cap, s + r1);
        bool f[2] = {keep0, keep1};
        uint32_t off[2], tot[2];
        block_flag_scan<2>(f, off, tot, scan_scratch);
        if (keep0 && cur + off[0] != r0) {
            const uint64_t pos = s + cur + off[0];
            store_key<W>(a.x, a.z, a.cap, pos, k0);
            store_coeff<T>(a.c, pos, c0);
        }
        if (keep1 && cur + tot[0] + off[1] != r1) {
            const uint64_t pos = s + cur + tot[0] + off[1];
            store_key<W>(a.x, a.z, a.cap, pos, k1);
            store_coeff<T>(a.c, pos, c1);
        }
        cur += tot[0] + tot[1];
    }
    if (threadIdx.x == 0) a.len[b] = cur;
}

__global__ void k_mark_positions(const uint64_t* pos, uint64_t n, uint8_t* flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[pos[i]] = 1;
}

// The boundary flags of one selection: pos holds its boundary bin's positions in rank order, the first `kept` kept.
__global__ void k_topn_flag_ties(const uint64_t* pos, uint64_t ties, uint64_t kept, uint8_t* flags) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ties) flags[pos[j]] = j < kept ? 1 : 0;
}

// The boundary flags of many selections, one workgroup each: every position of the item's boundary bin, in rank
// order, is flagged kept (its first `kept`) or dropped.  Only positions in the boundary bin are ever read, so no
// other flag needs a value.
__global__ void __launch_bounds__(kBlock) k_topn_flags_batch(const TopnFlagItem* __restrict__ items) {
    const TopnFlagItem it = items[blockIdx.x];
    for (uint64_t j = threadIdx.x; j < it.ties; j += kBlock) it.flags[it.pos[j]] = j < it.kept ? 1 : 0;
}

// ---- ApproxTopN ---------------------------------------------------------------------------------------------

// octave(norm_sqr) = norm_sqr.to_bits() >> 52 (builtin.rs:276-279).  A NaN of either sign counts in octave 2047,
// as a positive NaN does in the reference; classify_octave (:281-288) puts it below every octave.
__device__ __forceinline__ uint32_t octave_of(uint64_t u) { return (uint32_t)((u >> 52) & 0x7FFu); }

template <typename T>
__global__ void __launch_bounds__(kBlock) k_octave_hist(const OctaveHistArgs<T> a) {
    __shared__ uint32_t h[kOctaveBins];
    for (int i = threadIdx.x; i < kOctaveBins; i += kBlock) h[i] = 0;
    __syncthreads();
    const uint32_t b_lo = blockIdx.x * a.buckets_per_block;
    const uint32_t b_hi = min(a.num_buckets, b_lo + a.buckets_per_block);
    for (uint32_t b = b_lo; b < b_hi; ++b) {
        const uint32_t n = a.len[b];
        const uint64_t s = a.start[b];
        for (uint32_t r = threadIdx.x; r < n; r += kBlock) atomicAdd(&h[octave_of(magnitude_bits<T>(a.c[s + r]))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kOctaveBins; i += kBlock)
        if (h[i]) atomicAdd(&a.hist[i], (unsigned long long)h[i]);
}

// The boundary side of a term: 1 above the octave, 0 at it, -1 below (or NaN).
__device__ __forceinline__ int octave_side(uint64_t u, uint32_t bin) {
    if (is_nan_bits(u)) return -1;
    const uint32_t o = octave_of(u);
    return o > bin ? 1 : (o == bin ? 0 : -1);
}

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_octave_counts(const OctaveCutArgs<W, T> a) {
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
        if (octave_side(magnitude_bits<T>(a.c[s + r]), a.bin) != 0) continue;
        const Key<W> k = load_key<W>(a.x, a.z, a.cap, s + r);
        atomicAdd(&a.counts[bucket_of<W>(k, a.rows, a.ref_bits)], 1u);
    }
}

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_octave_gather(const OctaveCutArgs<W, T> a) {
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
        if (octave_side(magnitude_bits<T>(a.c[s + r]), a.bin) != 0) continue;
        const Key<W> k = load_key<W>(a.x, a.z, a.cap, s + r);
        if (bucket_of<W>(k, a.rows, a.ref_bits) != a.split_bucket) continue;
        const uint64_t slot = atomicAdd(a.gathered, 1ull);
        if (slot >= a.out_cap) continue;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            a.out_x[slot * W + w] = k.x[w];
            a.out_z[slot * W + w] = k.z[w];
        }
        a.out_pos[slot] = s + r;
    }
}

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_octave_retain(const OctaveCutArgs<W, T> a) {
    __shared__ uint32_t scan_scratch[kWaves];
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    uint32_t cur = 0;
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t r = base + threadIdx.x;
        Key<W> k;
        T c = T(0);
        bool keep = false;
        if (r < n) {
            k = load_key<W>(a.x, a.z, a.cap, s + r);
            c = a.c[s + r];
            const int side = octave_side(magnitude_bits<T>(c), a.bin);
            if (side > 0) {
                keep = true;
            } else if (side == 0) {
                const uint32_t rb = bucket_of<W>(k, a.rows, a.ref_bits);
                keep = rb < a.split_bucket || (rb == a.split_bucket && a.keep != nullptr && a.keep[s + r] != 0);
            }
        }
        bool f[1] = {keep};
        uint32_t off[1], tot[1];
        block_flag_scan<1>(f, off, tot, scan_scratch);  // every read of the chunk precedes its writes
        if (keep && cur + off[0] != r) {
            const uint64_t pos = s + cur + off[0];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                a.x[w * a.cap + pos] = k.x[w];
                a.z[w * a.cap + pos] = k.z[w];
            }
            a.c[pos] = c;
        }
        cur += tot[0];
    }
    if (threadIdx.x == 0) a.len[b] = cur;
}

template <typename T>
hipError_t launch_octave_hist(hipStream_t st, const OctaveHistArgs<T>& a, uint32_t blocks) {
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_octave_hist<T>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_octave_counts(hipStream_t st, const OctaveCutArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_octave_counts<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_octave_gather(hipStream_t st, const OctaveCutArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_octave_gather<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_octave_retain(hipStream_t st, const OctaveCutArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_octave_retain<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_topn_init_batch(hipStream_t st, const TopnSelectItem* items, uint32_t num_items) {
    if (num_items == 0) return hipSuccess;
    hipLaunchKernelGGL(k_topn_init_batch, dim3(num_items), dim3(kBlock), 0, st, items);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_topn_hist_batch(hipStream_t st, const TopnHistArgs<T>* items, const uint32_t* block_start,
                                  uint32_t num_items, uint32_t blocks, int pass) {
    if (num_items == 0 || blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_topn_hist_batch<T>), dim3(blocks), dim3(kBlock), 0, st, items, block_start, num_items,
                       (int32_t)pass);
    return hipGetLastError();
}

hipError_t launch_topn_select_batch(hipStream_t st, const TopnSelectItem* items, uint32_t num_items, int pass,
                                    bool last) {
    if (num_items == 0) return hipSuccess;
    hipLaunchKernelGGL(k_topn_select_batch, dim3(num_items), dim3(kBlock), 0, st, items, pass, last ? 1 : 0);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_topn_gather_batch(hipStream_t st, const TopnTieArgs<W, T>* items, const uint32_t* block_start,
                                    uint32_t num_items, uint32_t blocks) {
    if (num_items == 0 || blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_topn_gather_batch<W, T>), dim3(blocks), dim3(kBlock), 0, st, items, block_start,
                       num_items);
    return hipGetLastError();
}

hipError_t launch_topn_flag_ties(hipStream_t st, const uint64_t* pos, uint64_t ties, uint64_t kept, uint8_t* flags) {
    if (ties == 0) return hipSuccess;
    hipLaunchKernelGGL(k_topn_flag_ties, dim3((uint32_t)((ties + 255) / 256)), dim3(256), 0, st, pos, ties, kept,
                       flags);
    return hipGetLastError();
}

hipError_t launch_topn_flags_batch(hipStream_t st, const TopnFlagItem* items, uint32_t num_items) {
    if (num_items == 0) return hipSuccess;
    hipLaunchKernelGGL(k_topn_flags_batch, dim3(num_items), dim3(kBlock), 0, st, items);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_topn_hist(hipStream_t st, const TopnHistArgs<T>& a, uint32_t blocks) {
    hipLaunchKernelGGL((k_topn_hist<T>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_topn_select(hipStream_t st, TopnState* state, unsigned long long* hist, int pass, uint64_t n) {
    hipLaunchKernelGGL(k_topn_select, dim3(1), dim3(kBlock), 0, st, state, hist, pass, n);
    return hipGetLastError();
}

hipError_t launch_topn_init(hipStream_t st, TopnState* state, uint64_t n, unsigned long long* hist,
                            uint32_t hist_words, uint64_t seed, bool seeded) {
    hipLaunchKernelGGL(k_topn_init, dim3(1), dim3(kBlock), 0, st, state, n, seed, seeded ? 1u : 0u, hist,
                       hist_words);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_topn_hist_seeded(hipStream_t st, const TopnHistArgs<T>& a, uint32_t blocks) {
    hipLaunchKernelGGL((k_topn_hist_seeded<T>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_topn_select_seeded(hipStream_t st, TopnState* state, unsigned long long* hist, uint64_t n) {
    hipLaunchKernelGGL(k_topn_select_seeded, dim3(1), dim3(kBlock), 0, st, state, hist, n);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_topn_gather_ties(hipStream_t st, const TopnTieArgs<W, T>& a, uint32_t num_buckets) {
    (void)num_buckets;
    if (a.buckets_per_block == 0 || a.buckets_per_block > kHistBucketsMax) return hipErrorInvalidValue;
    const uint32_t blocks = (a.num_buckets + a.buckets_per_block - 1) / a.buckets_per_block;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_topn_gather_ties<W, T>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_topn_retain(hipStream_t st, const TopnRetainArgs<W, T>& a, uint32_t num_buckets) {
    hipLaunchKernelGGL((k_topn_retain<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mark_positions(hipStream_t st, const uint64_t* pos, uint64_t n, uint8_t* flags) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mark_positions, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, pos, n, flags);
    return hipGetLastError();
}

template hipError_t launch_topn_hist_batch<double>(hipStream_t, const TopnHistArgs<double>*, const uint32_t*, uint32_t,
                                                   uint32_t, int);
template hipError_t launch_topn_hist_batch<float>(hipStream_t, const TopnHistArgs<float>*, const uint32_t*, uint32_t,
                                                  uint32_t, int);
template hipError_t launch_topn_hist<double>(hipStream_t, const TopnHistArgs<double>&, uint32_t);
template hipError_t launch_topn_hist<float>(hipStream_t, const TopnHistArgs<float>&, uint32_t);
template hipError_t launch_topn_hist_seeded<double>(hipStream_t, const TopnHistArgs<double>&, uint32_t);
template hipError_t launch_topn_hist_seeded<float>(hipStream_t, const TopnHistArgs<float>&, uint32_t);
template hipError_t launch_octave_hist<double>(hipStream_t, const OctaveHistArgs<double>&, uint32_t);
template hipError_t launch_octave_hist<float>(hipStream_t, const OctaveHistArgs<float>&, uint32_t);
#define PP_TOPN_INSTANTIATE(W, T)                                                                       \
    template hipError_t launch_topn_gather_ties<W, T>(hipStream_t, const TopnTieArgs<W, T>&, uint32_t); \
    template hipError_t launch_topn_gather_batch<W, T>(hipStream_t, const TopnTieArgs<W, T>*, const uint32_t*, \
                                                       uint32_t, uint32_t);                               \
    template hipError_t launch_topn_retain<W, T>(hipStream_t, const TopnRetainArgs<W, T>&, uint32_t);   \
    template hipError_t launch_octave_counts<W, T>(hipStream_t, const OctaveCutArgs<W, T>&, uint32_t);  \
    template hipError_t launch_octave_gather<W, T>(hipStream_t, const OctaveCutArgs<W, T>&, uint32_t);  \
    template hipError_t launch_octave_retain<W, T>(hipStream_t, const OctaveCutArgs<W, T>&, uint32_t);
PP_TOPN_INSTANTIATE(1, double)
PP_TOPN_INSTANTIATE(1, float)
PP_TOPN_INSTANTIATE(2, double)
PP_TOPN_INSTANTIATE(2, float)

}  // namespace pp
