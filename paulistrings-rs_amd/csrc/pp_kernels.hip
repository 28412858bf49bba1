// pp_kernels.hip -- the HIP kernels of the rotation-layer engine, for gfx950 (MI355X, wave64).
//
// The sum is resident in HBM as word-major columns grouped into buckets of the reference's GF(2) hash
// (crates/paulistrings/src/pauli_sum/hash.rs:198-361) at a bucket count the device picks for itself.  A bucket
// is a CSR segment (start, len) with room to spare; rows within a bucket are in no particular order.  Because the
// hash is linear, a rotation about g moves a key of bucket b only to bucket b ^ h(g), so one workgroup owns the
// pair {b, b ^ h(g)} and writes both output buckets: every input row is read once and every output row written
// once, which is the whole HBM traffic of the layer.
#include "pp_device.h"
#include "pp_kernels.h"

namespace pp {

// ---------------------------------------------------------------------------------------------------------------
// Bucket offsets of the next layout: a three-kernel exclusive scan over per-bucket sizes (tiles of
// kScanTile buckets: tile sums, one workgroup scans the tile sums, then each tile scans itself).
// mode 0: size = len[b]; mode 1: rotation output room len[b] + len[b ^ delta] (2 len[b] when delta = 0);
// mode 2: coarsened bucket b < coarse_buckets gathers b + e * coarse_buckets for e < fold; mode 3: len[b] as a
// resident segment (a permutation run's counts).
// ---------------------------------------------------------------------------------------------------------------

constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kBlock * kScanItems;

struct OffsetsArgs {
    const uint32_t* len;
    uint32_t out_n;
    uint32_t mode;
    uint32_t delta;
    uint32_t coarse_buckets;
    uint32_t fold;
    uint64_t* out_start;
    uint32_t* out_cap;  // when set (a resident layout), each segment's rows
    uint64_t* tile_sums;
    unsigned long long* total;
    unsigned long long* clear;  // when set, clear[0 .. clear_n) is zeroed before *total is written
    uint32_t clear_n;
};

// Mode 0 packs the buckets densely (the export layout); modes 1 (a join's outputs: at most both buckets of the
// pair), 2 (coarsening) and 3 (a permutation run's destinations) allocate resident segments, rounded up to
// kSegAlign rows.
__device__ __forceinline__ uint64_t bucket_size(const OffsetsArgs& a, uint32_t b) {
    if (a.mode == 0) return a.len[b];
    uint64_t s = 0;
    if (a.mode == 1) {
        s = a.delta ? (uint64_t)a.len[b] + a.len[b ^ a.delta] : 2ull * a.len[b];
    } else if (a.mode == 3) {
        s = a.len[b];
    } else {
        for (uint32_t e = 0; e < a.fold; ++e) s += a.len[b + e * a.coarse_buckets];
    }
    return (s + (kSegAlign - 1)) & ~(uint64_t)(kSegAlign - 1);
}

// Exclusive scan of one u64 per thread across the workgroup; returns the block total in *total.
__device__ __forceinline__ uint64_t block_exclusive_scan_u64(uint64_t v, uint64_t* wave_sums, uint64_t* total) {
    const uint32_t lane = lane_id(), wave = wave_id();
    uint64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t o = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += o;
    }
    if (lane == 63) wave_sums[wave] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint64_t s = wave_sums[w];
        before += (uint32_t)w < wave ? s : 0ull;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

__global__ void __launch_bounds__(kBlock) k_offsets_tiles(const OffsetsArgs a) {
    __shared__ uint64_t wave_sums[kWaves];
    const uint32_t lo = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i)
        if (lo + i < a.out_n) mine += bucket_size(a, lo + i);
    uint64_t total;
    block_exclusive_scan_u64(mine, wave_sums, &total);
    if (threadIdx.x == 0) a.tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_offsets_top(const OffsetsArgs a, uint32_t tiles) {
    __shared__ uint64_t wave_sums[kWaves];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < tiles; base += kBlock) {
        const uint32_t t = base + threadIdx.x;
        const uint64_t v = t < tiles ? a.tile_sums[t] : 0ull;
        uint64_t total;
        const uint64_t ex = block_exclusive_scan_u64(v, wave_sums, &total);
        if (t < tiles) a.tile_sums[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        for (uint32_t i = 0; i < a.clear_n; ++i) a.clear[i] = 0ull;
        if (a.total) *a.total = carry;
    }
}

__global__ void __launch_bounds__(kBlock) k_offsets_final(const OffsetsArgs a) {
    __shared__ uint64_t wave_sums[kWaves];
    const uint32_t lo = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint64_t sz[kScanItems];
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        sz[i] = lo + i < a.out_n ? bucket_size(a, lo + i) : 0ull;
        mine += sz[i];
    }
    uint64_t total;
    uint64_t base = a.tile_sums[blockIdx.x] + block_exclusive_scan_u64(mine, wave_sums, &total);
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i)
        if (lo + i < a.out_n) {
            a.out_start[lo + i] = base;
            if (a.out_cap) a.out_cap[lo + i] = (uint32_t)sz[i];
            base += sz[i];
        }
}

// The whole scan in one workgroup when the layout fits one tile (kScanTile buckets): the small sums' layers, whose
// time is mostly launches, take one launch here instead of three.
__global__ void __launch_bounds__(kBlock) k_offsets_one(const OffsetsArgs a) {
    __shared__ uint64_t wave_sums[kWaves];
    const uint32_t lo = threadIdx.x * kScanItems;
    uint64_t sz[kScanItems];
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        sz[i] = lo + i < a.out_n ? bucket_size(a, lo + i) : 0ull;
        mine += sz[i];
    }
    uint64_t total;
    uint64_t base = block_exclusive_scan_u64(mine, wave_sums, &total);
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i)
        if (lo + i < a.out_n) {
            a.out_start[lo + i] = base;
            if (a.out_cap) a.out_cap[lo + i] = (uint32_t)sz[i];
            base += sz[i];
        }
    if (threadIdx.x == 0) {
        for (uint32_t i = 0; i < a.clear_n; ++i) a.clear[i] = 0ull;
        if (a.total) *a.total = total;
    }
}

// Total and longest bucket of a layout, with a handful of atomics per launch (one per workgroup), and the
// total added to a measurement counter when one is given.
__global__ void __launch_bounds__(kBlock) k_reduce_lens(const uint32_t* __restrict__ len, uint32_t num_buckets,
                                                        unsigned long long* __restrict__ out,
                                                        unsigned long long* __restrict__ prof) {
    __shared__ uint64_t wave_sums[kWaves];
    __shared__ uint32_t wave_max[kWaves];
    const uint32_t lo = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint64_t s = 0;
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i)
        if (lo + i < num_buckets) {
            const uint32_t v = len[lo + i];
            s += v;
            m = max(m, v);
        }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        s += __shfl_down(s, d, 64);
        m = max(m, (uint32_t)__shfl_down((int)m, d, 64));
    }
    if (lane_id() == 0) {
        wave_sums[wave_id()] = s;
        wave_max[wave_id()] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ts = 0;
        uint32_t tm = 0;
        for (int w = 0; w < kWaves; ++w) {
            ts += wave_sums[w];
            tm = max(tm, wave_max[w]);
        }
        if (ts) atomicAdd(&out[0], (unsigned long long)ts);
        if (tm) atomicMax(&out[1], (unsigned long long)tm);
        if (prof && ts) atomicAdd(&prof[0], (unsigned long long)ts);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// The rotation layer (reference semantics: engine/bucketed/rotation_layer.rs:159-229, :386-463;
// rotation_layer/merge.rs:53-91, :209-286; the device join it replaces is kern/join.cu:42-212).
//
// Per output key there are at most two contributions, own * cos and partner * sin, so each sum is one rounded
// add and the result is bitwise the reference's whatever the order.
//
// A workgroup joins the bucket pairs {b, b ^ delta} blockIdx.x, blockIdx.x + G, ... (G = gridDim.x), streaming each
// pair in chunks of kBlock rows.  Phase A (per chunk): a row commuting with g is final (kept iff nonzero and
// keep_term) and written at once; an anticommuting row is staged in LDS and entered in an open-addressing table.
// Phase B (after a pair's last chunk) walks the staged rows: row k's output is own*cos (0 on the Clifford arm)
// plus, if k ^ g is staged, that row's c * a1; if k ^ g is not staged it is a new key k ^ g with c * a1 in the
// other bucket.
//
// The loads of the next chunk (of this pair, or the first of the next pair) are issued before the current chunk
// is processed, so they stay in flight through its scan, its writes and the pair's phase B; the next pair's
// metadata is loaded one pair ahead.  Nothing between a load and its use waits on younger loads: phase A and B
// read LDS only.
// ---------------------------------------------------------------------------------------------------------------

// Where a pair lives: its two buckets (b1 = b0 and n1 = 0 for a single-bucket coset), their input segments and
// their output segments.
struct PairMeta {
    uint32_t b0, b1, n0, n1;
    uint64_t s0, s1, o0, o1;
};

template <int W, typename T>
__device__ __forceinline__ void pair_buckets(const JoinArgs<W, T>& a, uint32_t p, uint32_t& b0, uint32_t& b1) {
    if (a.delta) {
        const uint32_t lo = p & ((1u << a.hb) - 1u);
        b0 = ((p >> a.hb) << (a.hb + 1u)) | lo;
        b1 = b0 ^ a.delta;
    } else {
        b0 = b1 = p;
    }
}

// A pair's metadata in flight: every lane loads len, start and ostart of b0 (even lanes) or b1 (odd lanes), three
// unconditional vector loads read out of lanes 0 and 1 with v_readlane once they land.  Nothing is computed from
// the loaded words before then: a select on them, or a branch merging two of the loads, would make the loads
// issued after them wait.
struct MetaWords {
    uint32_t len;
    uint64_t start, ostart;
};

template <int W, typename T>
__device__ __forceinline__ MetaWords load_meta(const JoinArgs<W, T>& a, uint32_t p) {
    uint32_t b0, b1;
    pair_buckets(a, p, b0, b1);
    const uint32_t b = (threadIdx.x & 1u) ? b1 : b0;
    MetaWords w;
    w.len = a.len[b];
    w.start = a.start[b];
    w.ostart = a.ostart[b];
    return w;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

template <int W, typename T>
__device__ __forceinline__ PairMeta read_meta(const JoinArgs<W, T>& a, uint32_t p, const MetaWords& w) {
    PairMeta m;
    pair_buckets(a, p, m.b0, m.b1);
    m.n0 = (uint32_t)__builtin_amdgcn_readlane((int)w.len, 0);
    m.n1 = a.delta ? (uint32_t)__builtin_amdgcn_readlane((int)w.len, 1) : 0u;
    m.s0 = readlane_u64(w.start, 0);
    m.s1 = readlane_u64(w.start, 1);
    m.o0 = readlane_u64(w.ostart, 0);
    m.o1 = readlane_u64(w.ostart, 1);
    return m;
}

template <int W, typename T>
struct ChunkRow {
    Key<W> k;
    T c;
};

// Row r of a pair (bucket b0's rows, then b1's); rows at or past n are not loaded.
template <int W, typename T>
__device__ __forceinline__ void load_chunk_row(const JoinArgs<W, T>& a, const PairMeta& m, uint32_t r,
                                               ChunkRow<W, T>& row) {
    if (r < m.n0 + m.n1) {
        const uint64_t i = r < m.n0 ? m.s0 + r : m.s1 + (r - m.n0);
        row.k = load_key<W>(a.x, a.z, a.cap, i);
        row.c = a.c[i];
    }
}

__device__ __forceinline__ uint32_t table_bits(uint32_t n) {
    uint32_t t = 1;
    while ((1u << t) < 2u * n) ++t;
    return t;
}

// Phase B over staged rows lb .. lb + BLOCK - 1 of a pair (see k_rotation_join).
template <int W, typename T>
struct Staged {
    uint64_t* x;  // [W][stride]
    uint64_t* z;  // [W][stride]
    T* c;
    uint32_t* fold;
    uint8_t* side;  // 0: the row came from b0, 1: from b1
    uint32_t* table;
    uint32_t stride;  // staged rows held (the word planes' stride)
};

template <int W, typename T, int BLOCK>
__device__ __forceinline__ void join_staged_pass(const JoinArgs<W, T>& a, const PairMeta& m, const Key<W>& g,
                                                 uint32_t supp, uint32_t fold_g, uint32_t tbits, uint32_t nl,
                                                 uint32_t lb, const Staged<W, T>& st, uint32_t& cur0,
                                                 uint32_t& cur1, uint32_t* scan_scratch) {
    const bool pair = a.delta != 0;
    const uint32_t stride = st.stride;
    const uint64_t* Lx = st.x;
    const uint64_t* Lz = st.z;
    const T* Lc = st.c;
    const uint32_t* Lf = st.fold;
    const uint8_t* Lm = st.side;
    const uint32_t* table = st.table;
    const uint32_t tmask = (1u << tbits) - 1u;
    const uint32_t i = lb + threadIdx.x;
    const bool valid = i < nl;
    Key<W> k, q;
    T own = T(0), fresh = T(0);
    uint32_t j = 0;
    bool keep_own = false, keep_new = false;
    if (valid) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            k.x[w] = Lx[w * stride + i];
            k.z[w] = Lz[w * stride + i];
        }
        const T c = Lc[i];
        j = Lm[i];
        q = key_xor<W>(k, g);
        const uint32_t h = fold_mix(Lf[i] ^ fold_g);
        const uint32_t tag = h << 16;
        uint32_t s = h >> (32u - tbits);
        int found = -1;
        for (;;) {
            const uint32_t e = table[s];
            if (e == 0u) break;
            if ((e & 0xFFFF0000u) == tag) {
                const uint32_t idx = (e & 0xFFFFu) - 1u;
                bool eq = true;
#pragma unroll
                for (int w = 0; w < W; ++w)
                    eq = eq && (Lx[w * stride + idx] == q.x[w]) && (Lz[w * stride + idx] == q.z[w]);
                if (eq) {
                    found = (int)idx;
                    break;
                }
            }
            s = (s + 1u) & tmask;
        }
        // one phase evaluation: of the partner q when it is staged (its a1 lands on k), else of k
        // (whose a1 makes the new key q)
        const T a1 = partner_amplitude_on<W, T>(found >= 0 ? q : k, g, supp, a.gphase, a.sin_t);
        own = a.clifford ? T(0) : c * a.cos_t;
        if (found >= 0) own = own + Lc[found] * a1;
        keep_own = survives<W, T>(a.keep, k, own);
        if (found < 0) {
            fresh = c * a1;
            keep_new = survives<W, T>(a.keep, q, fresh);
        }
    }
    const uint32_t jn = pair ? (j ^ 1u) : j;
    bool f[4] = {keep_own && j == 0, keep_own && j == 1, keep_new && jn == 0, keep_new && jn == 1};
    uint32_t off[4], tot[4];
    block_flag_scan<4, BLOCK / 64>(f, off, tot, scan_scratch);
    // one store sequence per output kind, for either bucket
    if (keep_own) {
        const uint64_t pos = j == 0 ? m.o0 + cur0 + off[0] : m.o1 + cur1 + off[1];
        store_key<W>(a.ox, a.oz, a.ocap, pos, k);
        store_coeff<T>(a.oc, pos, own);
    }
    if (keep_new) {
        const uint64_t pos = jn == 0 ? m.o0 + cur0 + tot[0] + off[2] : m.o1 + cur1 + tot[1] + off[3];
        store_key<W>(a.ox, a.oz, a.ocap, pos, q);
        store_coeff<T>(a.oc, pos, fresh);
    }
    cur0 += tot[0] + tot[2];
    cur1 += tot[1] + tot[3];
}

// REDO: the launch over the queued pairs (a compile-time switch, so that the first launch's pair index needs no
// load: a lookup there, merged with the prefetch's path, would make the compiler wait for the prefetch).
// The join of work items t, t + G, t + G + G, ... below count (see k_rotation_join); lds and scan_scratch are the
// workgroup's.  FILTER: the input may carry a pending TopN selection (a compile-time switch: the arguments it reads
// stay out of the scalar registers of the common launch, which has none).
// The output segment sizes of a small layer (at most kSmallJoinBuckets buckets) below a pair's two buckets and in
// all, from the input lengths: each thread's share of the buckets is loaded first (small_lens_load) and summed
// later (small_lens_sum), so that the loads overlap the pair's own first loads.
constexpr uint32_t kSmallLensPerThread = kSmallJoinBuckets / kBlock;
struct SmallLens {
    uint32_t n[kSmallLensPerThread];  // rows of bucket i and its partner i ^ delta
};

template <int W, typename T>
__device__ __forceinline__ SmallLens small_lens_load(const JoinArgs<W, T>& a, uint32_t num_buckets) {
    SmallLens l;
#pragma unroll
    for (uint32_t k = 0; k < kSmallLensPerThread; ++k) {
        const uint32_t i = threadIdx.x + k * kBlock;
        l.n[k] = 0;
        if (i < num_buckets) l.n[k] = a.delta ? a.len[i] + a.len[i ^ a.delta] : 2u * a.len[i];
    }
    return l;
}

// Sizes below the pair's b0, below its b1 and in all (a workgroup reduction through part[3][kWaves]) into m.o0,
// m.o1 and all; thread 0 also writes the starts and rooms of b0 and b1.
template <int W, typename T>
__device__ __forceinline__ void small_lens_sum(const JoinArgs<W, T>& a, const SmallLens& l, PairMeta& m,
                                               uint32_t* orows, uint64_t* part, uint64_t& all) {
    const uint32_t b0 = m.b0, b1 = m.b1;
    uint64_t below0 = 0, below1 = 0, tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSmallLensPerThread; ++k) {
        const uint32_t i = threadIdx.x + k * kBlock;
        const uint64_t sz = ((uint64_t)l.n[k] + (kSegAlign - 1)) & ~(uint64_t)(kSegAlign - 1);
        tot += sz;
        below0 += i < b0 ? sz : 0ull;
        below1 += i < b1 ? sz : 0ull;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        below0 += __shfl_down(below0, d, 64);
        below1 += __shfl_down(below1, d, 64);
        tot += __shfl_down(tot, d, 64);
    }
    if (lane_id() == 0) {
        part[wave_id()] = below0;
        part[kWaves + wave_id()] = below1;
        part[2 * kWaves + wave_id()] = tot;
    }
    __syncthreads();
    uint64_t o0 = 0, o1 = 0;
    all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        o0 += part[w];
        o1 += part[kWaves + w];
        all += part[2 * kWaves + w];
    }
    m.o0 = o0;
    m.o1 = o1;
    if (threadIdx.x == 0) {
        const uint64_t n = a.delta ? (uint64_t)m.n0 + m.n1 : 2ull * m.n0;  // the pair's rows (n1 = 0 alone)
        const uint32_t room = (uint32_t)((n + (kSegAlign - 1)) & ~(uint64_t)(kSegAlign - 1));
        const_cast<uint64_t*>(a.ostart)[b0] = o0;
        orows[b0] = room;
        if (a.delta) {
            const_cast<uint64_t*>(a.ostart)[b1] = o1;
            orows[b1] = room;
        }
    }
}

// SMALL: one pair per workgroup, which lays out its own output segments (small_lens_*: small_buckets buckets,
// small_rows receiving the rooms, small_part the reduction's LDS, *small_all the layout's end).
template <int W, typename T, bool REDO, int BLOCK, bool FILTER, bool SMALL = false>
__device__ __forceinline__ void join_items(const JoinArgs<W, T>& a, uint32_t t, uint32_t G, uint32_t count,
                                           unsigned char* lds, uint32_t* scan_scratch, uint32_t small_buckets = 0,
                                           uint32_t* small_rows = nullptr, uint64_t* small_part = nullptr,
                                           uint64_t* small_all = nullptr) {
    const uint32_t scap = a.scap;  // staged rows this launch holds
    uint64_t* const Lx = reinterpret_cast<uint64_t*>(lds);   // [W][scap]
    uint64_t* const Lz = Lx + (size_t)W * scap;              // [W][scap]
    T* const Lc = reinterpret_cast<T*>(Lz + (size_t)W * scap);
    uint32_t* const Lf = reinterpret_cast<uint32_t*>(Lc + scap);  // key_fold of each staged row
    uint32_t* const table = Lf + scap;
    uint8_t* const Lm = reinterpret_cast<uint8_t*>(table + (1u << a.tcap_bits));
    const Staged<W, T> st{Lx, Lz, Lc, Lf, Lm, table, scap};

    // Work item t is pair t, or in the redo launch the t-th pair the first launch queued.
    constexpr bool redo = REDO;
    if (t >= count) return;  // uniform
    const uint32_t last = count - 1u;
    auto pair_at = [&](uint32_t i) { return redo ? a.redo[i] : i; };
    const bool pair = a.delta != 0;
    Key<W> g;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        g.x[w] = a.gx[w];
        g.z[w] = a.gz[w];
    }
    const uint32_t supp = support_words<W>(g);
    const uint32_t fold_g = key_fold<W>(g);
    TopnState fs{};
    if constexpr (FILTER) {
        if (a.filter) fs = *a.filter;
    }

    uint32_t p = pair_at(t);
    const MetaWords w0 = load_meta(a, p);
    SmallLens sl;
    if constexpr (SMALL) sl = small_lens_load(a, small_buckets);
    PairMeta m = read_meta(a, p, w0);
    ChunkRow<W, T> cb;
    load_chunk_row(a, m, threadIdx.x, cb);
    MetaWords mw = load_meta(a, pair_at(min(t + G, last)));
    if constexpr (SMALL) small_lens_sum(a, sl, m, small_rows, small_part, *small_all);
    // Land the first chunk here, so that inside the loop the current chunk is always one copied from the
    // previous iteration's prefetch (which the copy waited for) and its first use waits on nothing younger.
    __builtin_amdgcn_s_waitcnt(0);
    uint32_t base = 0, cur0 = 0, cur1 = 0, nl = 0;
    bool overflow = false;  // more anticommuting rows than scap: the pair is left to the redo launch
    uint32_t tbits = min(table_bits(min(m.n0 + m.n1, scap)), a.tcap_bits);
    for (uint32_t i = threadIdx.x; i < (1u << tbits); i += BLOCK) table[i] = 0u;
    for (;;) {
        const uint32_t n = m.n0 + m.n1;
        // ---- the next chunk: this pair's, or the first of the next pair ----
        const bool last_chunk = base + BLOCK >= n;
        bool have_next = true;
        PairMeta mn = m;
        uint32_t nbase = base + BLOCK;
        uint32_t pn = p;
        if (last_chunk) {
            if (t + G <= last) {
                pn = pair_at(t + G);
                mn = read_meta(a, pn, mw);
                nbase = 0;
            } else {
                have_next = false;
            }
        }
        ChunkRow<W, T> nb;
        if (have_next) load_chunk_row(a, mn, nbase + threadIdx.x, nb);
        if (last_chunk && have_next) mw = load_meta(a, pair_at(min(t + 2u * G, last)));

        // ---- phase A over the current chunk ----
        if (!overflow) {  // uniform
            const uint32_t r = base + threadIdx.x;
            bool valid = r < n;
            const uint32_t j = r >= m.n0 ? 1u : 0u;
            if (FILTER && valid && a.filter) {  // uniform test: a pending TopN retain
                const uint64_t i = j ? m.s1 + (r - m.n0) : m.s0 + r;
                const uint64_t u = a.filter_score ? score_bits(a.filter_score[i]) : magnitude_bits<T>(cb.c);
                valid = topn_keeps(fs, u, !a.filter_score && is_nan_bits(u), a.filter_flags, i);
            }
            bool anti = false, keep = false;
            if (valid) {
                anti = anticommutes_on<W>(cb.k, g, supp);
                keep = !anti && survives<W, T>(a.keep, cb.k, cb.c);
            }
            bool f[3] = {keep && j == 0, keep && j == 1, anti};
            uint32_t off[3], tot[3];
            block_flag_scan<3, BLOCK / 64>(f, off, tot, scan_scratch);
            overflow = nl + tot[2] > scap;
            if (keep) {  // one store sequence for either bucket (a redo rewrites the same positions)
                const uint64_t pos = j == 0 ? m.o0 + cur0 + off[0] : m.o1 + cur1 + off[1];
                store_key<W>(a.ox, a.oz, a.ocap, pos, cb.k);
                store_coeff<T>(a.oc, pos, cb.c);
            }
            if (anti && !overflow) {
                const uint32_t li = nl + off[2];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    Lx[w * scap + li] = cb.k.x[w];
                    Lz[w * scap + li] = cb.k.z[w];
                }
                Lc[li] = cb.c;
                Lm[li] = (uint8_t)j;
                const uint32_t fk = key_fold<W>(cb.k);
                Lf[li] = fk;
                // the table: slot = (tag << 16) | (index + 1), load at most 1/2
                const uint32_t h = fold_mix(fk);
                const uint32_t entry = (h << 16) | (li + 1u);
                const uint32_t tmask = (1u << tbits) - 1u;
                uint32_t s = h >> (32u - tbits);
                while (atomicCAS(&table[s], 0u, entry) != 0u) s = (s + 1u) & tmask;
            }
            cur0 += tot[0];
            cur1 += tot[1];
            nl += tot[2];
        }

        if (last_chunk) {
            if (overflow) {
                // Queue the pair for the redo launch, which stages up to the largest pair; in the redo launch
                // itself an overflow means the host's bound was wrong.
                if (threadIdx.x == 0) {
                    if (redo || a.redo == nullptr) {
                        atomicOr(&a.stats[2], 1ull);
                    } else {
                        a.redo[atomicAdd(a.redo_count, 1ull)] = p;
                    }
                }
            } else {
                // ---- phase B over the pair's staged rows ----
                // One pass in the common case nl <= BLOCK: a loop holding stores and no loads would make the
                // compiler wait for the prefetch before entering it.
                if (nl != 0) {
                    __syncthreads();  // the staged rows and the table are complete
                    if (nl <= BLOCK) {
                        join_staged_pass<W, T, BLOCK>(a, m, g, supp, fold_g, tbits, nl, 0, st, cur0, cur1, scan_scratch);
                    } else {
                        for (uint32_t lb = 0; lb < nl; lb += BLOCK)
                            join_staged_pass<W, T, BLOCK>(a, m, g, supp, fold_g, tbits, nl, lb, st, cur0, cur1,
                                                   scan_scratch);
                    }
                }
                // The counts stay per bucket: a total through one global atomic per workgroup would serialize
                // ~10^5 same-address atomics per launch, so k_reduce_lens sums them afterwards.
                if (threadIdx.x == 0) {
                    a.olen[m.b0] = cur0;
                    if (pair) a.olen[m.b1] = cur1;
                    if (a.count_totals) {  // few pairs: these atomics cost less than a reduce launch
                        atomicAdd(&a.stats[0], (unsigned long long)(cur0 + cur1));
                        atomicMax(&a.stats[1], (unsigned long long)max(cur0, cur1));
                    }
                }
            }
            if (!have_next) break;
            // The next pair's table: every lookup above precedes phase B's scan barriers, every insert below
            // follows the next scan's barrier.
            t += G;
            p = pn;
            m = mn;
            cur0 = cur1 = nl = 0;
            overflow = false;
            tbits = min(table_bits(min(m.n0 + m.n1, scap)), a.tcap_bits);
            for (uint32_t i = threadIdx.x; i < (1u << tbits); i += BLOCK) table[i] = 0u;
        }
        base = nbase;
        cb = nb;
    }
}

template <int W, typename T, bool REDO, int BLOCK, bool FILTER>
__global__ void __launch_bounds__(BLOCK) k_rotation_join(const JoinArgs<W, T> a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t scan_scratch[4 * (BLOCK / 64)];
    const uint32_t count = REDO ? (uint32_t)*a.redo_count : a.num_cosets;
    join_items<W, T, REDO, BLOCK, FILTER>(a, blockIdx.x, gridDim.x, count, lds, scan_scratch);
}

// A small layer's counters to the host: the four words, then (once they have landed) the sequence number the host
// waits on; then the counters are cleared for the next launch that adds into them.  Thread 0 of one workgroup.
__device__ __forceinline__ void publish_counters(unsigned long long* counters, unsigned long long* host,
                                                 unsigned long long seq) {
    volatile unsigned long long* c = counters;
    volatile unsigned long long* h = host;  // fine-grained host memory: the stores bypass the L2
    const unsigned long long rows = c[0], longest = c[1], flags = c[2], end = c[3];
    h[0] = rows;
    h[1] = longest;
    h[2] = flags;
    h[3] = end;
    c[0] = 0;
    c[1] = 0;
    c[2] = 0;
    c[3] = 0;
    __builtin_amdgcn_s_waitcnt(0);  // the slot's words land before its sequence number
    h[4] = seq;
}

// A small layer's join in one launch (see SmallJoinArgs).  The workgroup of pair p sums the output segment sizes
// (each pair's rows rounded up to kSegAlign, as the offset scan lays them out) below each of its two buckets and in
// all, over every bucket of the layout (at most kSmallJoinBuckets, read by every workgroup while its pair's first
// rows load), writes its buckets' starts and rooms, and joins the pair.  The workgroup after the last pair
// publishes the launch before's counters, which that launch's end made complete: a layer's counters reach the host
// while the next layer runs, and no workgroup waits on another's.
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_rotation_join_small(const SmallJoinArgs<W, T> a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t scan_scratch[4 * kWaves];
    __shared__ uint64_t part[3][kWaves];
    const JoinArgs<W, T>& j = a.j;
    if (blockIdx.x == j.num_cosets) {  // uniform: the publishing workgroup
        if (threadIdx.x == 0 && a.prev_counters) publish_counters(a.prev_counters, a.prev_host, a.prev_seq);
        return;
    }
    uint64_t all = 0;
    join_items<W, T, false, kBlock, false, true>(j, blockIdx.x, j.num_cosets, j.num_cosets, lds, scan_scratch,
                                                 a.num_buckets, a.orows, &part[0][0], &all);
    if (blockIdx.x == 0 && threadIdx.x == 0) j.stats[3] = all;
}

__global__ void k_publish_counters(unsigned long long* counters, unsigned long long* host, unsigned long long seq) {
    if (threadIdx.x == 0) publish_counters(counters, host, seq);
}

// ---------------------------------------------------------------------------------------------------------------
// The rotation layer in place.  A workgroup per pair {b0, b1} reads every row of the pair, as k_rotation_join does,
// but writes only what the layer changes: the coefficient of an anticommuting row that stays, and the new partner
// rows, which go first into the rows the layer drops on their side (holes) and then after the side's last row.  A
// commuting row that stays is not written.  Every decision is taken before the first write.  The first launch
// leaves a pair untouched and queues it when it stages more anticommuting rows than the workgroup holds or does not
// fit its segments (more rows than a segment's room), adding the rows its move may take to counter 7.  A side that
// loses more rows than it gains closes its remaining holes in place: the new rows take its lowest holes, and the rows
// that stay above its final length move down into the others (their order does not matter in a bucket).  With
// FILTER the input carries a pending TopN selection: a row the selection drops is a hole.  The redo launch (REDO) stages the largest pair: a queued pair that fits is done in
// place, any other is written whole to fresh segments after the columns' last used row (tail_base plus counter 5,
// which the host has made room for), its old segments left as dead space.
// ---------------------------------------------------------------------------------------------------------------

constexpr uint32_t kInplaceHoles = 1024;  // holes per side a workgroup records (a bucket holds at most 512 rows)

// The rows a moved bucket's fresh segment holds: its rows, an eighth more for the next layers, rounded to kSegAlign.
__host__ __device__ constexpr uint32_t moved_segment_rows(uint32_t rows) {
    return (rows + rows / 8 + (kSegAlign - 1)) & ~(kSegAlign - 1);
}

template <int W, typename T>
size_t inplace_lds_bytes(uint32_t scap, uint32_t tcap_bits) {
    return (size_t)scap * (2 * W * sizeof(uint64_t) + 3 * sizeof(T) + 4 + 2 + 2 + 1 + 1) + ((size_t)4 << tcap_bits) +
           2 * (size_t)kInplaceHoles * 2;
}

// A side's holes as a bitmap of its rows (kInplaceHoles bits) with the holes before each word (pre[w], pre[32] all):
// the k-th hole in ascending order, and the holes below row p, each in a few steps.
constexpr uint32_t kHoleWords = kInplaceHoles / 32;
__device__ __forceinline__ uint32_t holes_below(const uint32_t* bits, const uint16_t* pre, uint32_t p) {
    const uint32_t w = p >> 5, b = p & 31u;
    return pre[w] + (b ? (uint32_t)__popc(bits[w] & ((1u << b) - 1u)) : 0u);
}

__device__ __forceinline__ uint32_t kth_hole(const uint32_t* bits, const uint16_t* pre, uint32_t k) {
    uint32_t lo = 0, hi = kHoleWords;  // the word w with pre[w] <= k < pre[w + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= k) lo = mid;
        else hi = mid;
    }
    uint32_t v = bits[lo], pos = 0;
    k -= pre[lo];
#pragma unroll
    for (uint32_t half = 16; half >= 1; half >>= 1) {  // select the k-th set bit by halves
        const uint32_t c = (uint32_t)__popc(v & ((1u << half) - 1u));
        if (k >= c) {
            k -= c;
            v >>= half;
            pos += half;
        }
    }
    return lo * 32u + pos;
}

template <int W, typename T, bool REDO, bool FILTER>
__global__ void __launch_bounds__(kBlock) k_rotation_inplace(const InplaceArgs<W, T> a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t scan_scratch[4 * kWaves];
    __shared__ uint64_t reloc_base;
    __shared__ uint32_t hbits[2][kHoleWords];
    __shared__ uint16_t hpre[2][kHoleWords + 1];
    const uint32_t S = a.scap;
    uint64_t* const Lx = reinterpret_cast<uint64_t*>(lds);
    uint64_t* const Lz = Lx + (size_t)W * S;
    T* const Lc = reinterpret_cast<T*>(Lz + (size_t)W * S);
    T* const Lown = Lc + S;
    T* const Lfresh = Lown + S;
    uint32_t* const Lf = reinterpret_cast<uint32_t*>(Lfresh + S);
    uint32_t* const table = Lf + S;
    uint16_t* const Li = reinterpret_cast<uint16_t*>(table + (1u << a.tcap_bits));
    uint16_t* const Lrank = Li + S;
    uint16_t* const holes = Lrank + S;  // [2][kInplaceHoles]
    uint8_t* const Lm = reinterpret_cast<uint8_t*>(holes + 2 * kInplaceHoles);
    uint8_t* const Lflag = Lm + S;

    const uint32_t count = REDO ? (uint32_t)*a.redo_count : a.num_cosets;
    const bool pair = a.delta != 0;
    Key<W> g;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        g.x[w] = a.gx[w];
        g.z[w] = a.gz[w];
    }
    const uint32_t supp = support_words<W>(g);
    const uint32_t fold_g = key_fold<W>(g);
    TopnState fs{};
    if constexpr (FILTER) fs = *a.filter;
    // whether the TopN selection keeps the row at column position i (coefficient c)
    auto selected = [&](uint64_t i, T c) -> bool {
        if constexpr (!FILTER) {
            return true;
        } else {
            const uint64_t u = a.filter_score ? score_bits(a.filter_score[i]) : magnitude_bits<T>(c);
            return topn_keeps(fs, u, !a.filter_score && is_nan_bits(u), a.filter_flags, i);
        }
    };
    for (uint32_t t = blockIdx.x; t < count; t += gridDim.x) {  // uniform
        const uint32_t p = REDO ? a.redo[t] : t;
        uint32_t b0, b1;
        if (pair) {
            const uint32_t lo = p & ((1u << a.hb) - 1u);
            b0 = ((p >> a.hb) << (a.hb + 1u)) | lo;
            b1 = b0 ^ a.delta;
        } else {
            b0 = b1 = p;
        }
        const uint32_t n0 = a.len[b0], n1 = pair ? a.len[b1] : 0u;
        const uint64_t s0 = a.start[b0], s1 = a.start[b1];
        const uint32_t cap0 = a.bcap[b0], cap1 = a.bcap[b1];
        const uint32_t n = n0 + n1;
        const uint32_t tbits = min(table_bits(min(max(n, 1u), S)), a.tcap_bits);
        const uint32_t tmask = (1u << tbits) - 1u;
        __syncthreads();  // the previous pair's LDS is no longer read
        for (uint32_t i = threadIdx.x; i < (1u << tbits); i += kBlock) table[i] = 0u;
        if (threadIdx.x < 2 * kHoleWords) (&hbits[0][0])[threadIdx.x] = 0u;
        // a side of more rows than the bitmap covers cannot close its holes in place
        const bool compactable = a.compact && n0 <= kInplaceHoles && n1 <= kInplaceHoles;
        __syncthreads();
        // ---- phase A: holes and staging; no global write ----
        uint32_t hc0 = 0, hc1 = 0, nl = 0;
        bool overflow = false;
        // each thread's row of a chunk, loaded one chunk ahead: the next chunk's loads are in flight while this
        // chunk's scan and staging run
        auto fetch = [&](uint32_t base, Key<W>& k, T& c) {
            const uint32_t r = base + threadIdx.x;
            if (r < n) {
                const uint64_t row = r >= n0 ? s1 + (r - n0) : s0 + r;
                k = load_key<W>(a.x, a.z, a.cap, row);
                c = a.c[row];
            }
        };
        Key<W> kn{};
        T cn = T(0);
        fetch(0, kn, cn);
        for (uint32_t base = 0; base < n; base += kBlock) {
            const uint32_t r = base + threadIdx.x;
            const bool in = r < n;
            const uint32_t j = r >= n0 ? 1u : 0u;
            const uint32_t idx = j ? r - n0 : r;
            const uint64_t row = j ? s1 + idx : s0 + idx;
            const Key<W> k = kn;
            const T c = cn;
            if (base + kBlock < n) fetch(base + kBlock, kn, cn);  // uniform
            bool valid = in;
            if (in) valid = selected(row, c);
            const bool anti = valid && anticommutes_on<W>(k, g, supp);
            const bool hole = in && !anti && !(valid && survives<W, T>(a.keep, k, c));
            bool f[4] = {hole && j == 0, hole && j == 1, anti && j == 0, anti && j == 1};
            uint32_t off[4], tot[4];
            block_flag_scan<4>(f, off, tot, scan_scratch);
            if (hole) {
                holes[j * kInplaceHoles + (j ? hc1 + off[1] : hc0 + off[0])] = (uint16_t)idx;
                if (compactable) atomicOr(&hbits[j][idx >> 5], 1u << (idx & 31u));
            }
            overflow = overflow || nl + tot[2] + tot[3] > S;
            if (anti && !overflow) {
                const uint32_t li = nl + off[2] + off[3];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    Lx[w * S + li] = k.x[w];
                    Lz[w * S + li] = k.z[w];
                }
                Lc[li] = c;
                Lm[li] = (uint8_t)j;
                Li[li] = (uint16_t)idx;
                const uint32_t fk = key_fold<W>(k);
                Lf[li] = fk;
                const uint32_t h = fold_mix(fk);
                const uint32_t entry = (h << 16) | (li + 1u);
                uint32_t sl = h >> (32u - tbits);
                while (atomicCAS(&table[sl], 0u, entry) != 0u) sl = (sl + 1u) & tmask;
            }
            hc0 += tot[0];
            hc1 += tot[1];
            nl += tot[2] + tot[3];
        }
        if (overflow) {
            if (threadIdx.x == 0) {
                if (REDO) {
                    atomicOr(&a.stats[2], 1ull);  // the redo launch stages the largest pair: never expected
                } else {
                    a.redo[atomicAdd(a.redo_count, 1ull)] = p;
                    // a side's rows after the layer are at most the pair's
                    atomicAdd(&a.stats[7], (unsigned long long)(2u * moved_segment_rows(n)));
                }
            }
            continue;
        }
        __syncthreads();  // staging, table and phase A's holes complete
        // ---- phase B: every staged row's decision, into LDS; own drops extend the holes ----
        uint32_t nc0 = 0, nc1 = 0;
        for (uint32_t lb = 0; lb < nl; lb += kBlock) {
            const uint32_t i = lb + threadIdx.x;
            const bool valid = i < nl;
            bool keep_own = false, keep_new = false;
            uint32_t jj = 0;
            if (valid) {
                Key<W> kk;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    kk.x[w] = Lx[w * S + i];
                    kk.z[w] = Lz[w * S + i];
                }
                const T cc = Lc[i];
                jj = Lm[i];
                const Key<W> q = key_xor<W>(kk, g);
                const uint32_t h = fold_mix(Lf[i] ^ fold_g);
                const uint32_t tag = h << 16;
                uint32_t sl = h >> (32u - tbits);
                int found = -1;
                for (;;) {
                    const uint32_t e = table[sl];
                    if (e == 0u) break;
                    if ((e & 0xFFFF0000u) == tag) {
                        const uint32_t ix = (e & 0xFFFFu) - 1u;
                        bool eq = true;
#pragma unroll
                        for (int w = 0; w < W; ++w)
                            eq = eq && (Lx[w * S + ix] == q.x[w]) && (Lz[w * S + ix] == q.z[w]);
                        if (eq) {
                            found = (int)ix;
                            break;
                        }
                    }
                    sl = (sl + 1u) & tmask;
                }
                const T a1 = partner_amplitude_on<W, T>(found >= 0 ? q : kk, g, supp, a.gphase, a.sin_t);
                T own = a.clifford ? T(0) : cc * a.cos_t;
                if (found >= 0) own = own + Lc[found] * a1;
                keep_own = survives<W, T>(a.keep, kk, own);
                Lown[i] = own;
                if (found < 0) {
                    const T fresh = cc * a1;
                    keep_new = survives<W, T>(a.keep, q, fresh);
                    Lfresh[i] = fresh;
                }
            }
            const uint32_t jn = pair ? (jj ^ 1u) : jj;
            const bool drop = valid && !keep_own;
            bool f[4] = {drop && jj == 0, drop && jj == 1, keep_new && jn == 0, keep_new && jn == 1};
            uint32_t off[4], tot[4];
            block_flag_scan<4>(f, off, tot, scan_scratch);
            if (drop) {
                holes[jj * kInplaceHoles + (jj ? hc1 + off[1] : hc0 + off[0])] = Li[i];
                if (compactable) atomicOr(&hbits[jj][Li[i] >> 5], 1u << (Li[i] & 31u));
            }
            if (valid) {
                Lflag[i] = (uint8_t)((keep_own ? 1u : 0u) | (keep_new ? 2u : 0u));
                Lrank[i] = (uint16_t)(jn ? nc1 + off[3] : nc0 + off[2]);
            }
            hc0 += tot[0];
            hc1 += tot[1];
            nc0 += tot[2];
            nc1 += tot[3];
        }
        // rows that stay, per side: every row of the side but its holes
        const uint32_t fin0 = n0 - hc0 + nc0, fin1 = n1 - hc1 + nc1;
        const bool fits = fin0 <= cap0 && (!pair || fin1 <= cap1) &&
                          (compactable || (hc0 <= nc0 && (!pair || hc1 <= nc1)));
        __syncthreads();  // every decision and hole in LDS
        if (!fits && !REDO) {
            if (threadIdx.x == 0) {
                a.redo[atomicAdd(a.redo_count, 1ull)] = p;
                atomicAdd(&a.stats[7], (unsigned long long)(moved_segment_rows(fin0) + (pair ? moved_segment_rows(fin1) : 0u)));
            }
            continue;
        }
        if (fits) {
            // ---- phase C in place: kept own coefficients, new rows into the holes and then after the rows; a
            // side left with holes (more holes than new rows) takes its new rows into its lowest holes and moves the
            // rows that stay above its final length into the holes left below it ----
            const bool shrink0 = hc0 > nc0, shrink1 = pair && hc1 > nc1;
            if (shrink0 || shrink1) {  // uniform: the holes before each word of the bitmap
                if (threadIdx.x < 2) {
                    uint32_t acc = 0;
                    for (uint32_t w = 0; w < kHoleWords; ++w) {
                        hpre[threadIdx.x][w] = (uint16_t)acc;
                        acc += (uint32_t)__popc(hbits[threadIdx.x][w]);
                    }
                    hpre[threadIdx.x][kHoleWords] = (uint16_t)acc;
                }
                __syncthreads();
            }
            for (uint32_t i = threadIdx.x; i < nl; i += kBlock) {
                const uint32_t fl = Lflag[i];
                const uint32_t jj = Lm[i];
                if (fl & 1u) {
                    const uint32_t at = Li[i];
                    const uint32_t fin = jj ? fin1 : fin0;
                    if ((jj ? shrink1 : shrink0) && at >= fin) {  // a staying row above the final length moves down
                        // its rank among the rows that stay above the final length, after the new rows
                        const uint32_t above = (at - fin) - (holes_below(hbits[jj], hpre[jj], at) -
                                                             holes_below(hbits[jj], hpre[jj], fin));
                        const uint32_t to = kth_hole(hbits[jj], hpre[jj], (jj ? nc1 : nc0) + above);
                        const uint64_t pos = (jj ? s1 : s0) + to;
                        Key<W> kk;
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            kk.x[w] = Lx[w * S + i];
                            kk.z[w] = Lz[w * S + i];
                        }
                        store_key<W>(a.x, a.z, a.cap, pos, kk);
                        store_coeff<T>(a.c, pos, Lown[i]);
                    } else {
                        a.c[(jj ? s1 : s0) + at] = Lown[i];
                    }
                }
                if (fl & 2u) {
                    const uint32_t jn = pair ? (jj ^ 1u) : jj;
                    const uint32_t rk = Lrank[i];
                    const uint32_t hcn = jn ? hc1 : hc0;
                    const uint32_t slot = (jn ? shrink1 : shrink0) ? kth_hole(hbits[jn], hpre[jn], rk)
                                          : rk < hcn ? holes[jn * kInplaceHoles + rk]
                                                     : (jn ? n1 : n0) + (rk - hcn);
                    const uint64_t pos = (jn ? s1 : s0) + slot;
                    Key<W> q;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        q.x[w] = Lx[w * S + i] ^ g.x[w];
                        q.z[w] = Lz[w * S + i] ^ g.z[w];
                    }
                    store_key<W>(a.x, a.z, a.cap, pos, q);
                    store_coeff<T>(a.c, pos, Lfresh[i]);
                }
            }
            // commuting rows that stay above a shrinking side's final length: read here (nothing writes at or above
            // the final length) and moved into the holes after the new rows'
            for (int jj = 0; jj < (pair ? 2 : 1); ++jj) {
                if (!(jj ? shrink1 : shrink0)) continue;  // uniform
                const uint32_t fin = jj ? fin1 : fin0, nn = jj ? n1 : n0;
                const uint64_t sj = jj ? s1 : s0;
                for (uint32_t at = fin + threadIdx.x; at < nn; at += kBlock) {
                    if ((hbits[jj][at >> 5] >> (at & 31u)) & 1u) continue;
                    const Key<W> k = load_key<W>(a.x, a.z, a.cap, sj + at);
                    if (anticommutes_on<W>(k, g, supp)) continue;  // a staged row: moved above
                    const T c = a.c[sj + at];
                    const uint32_t above = (at - fin) - (holes_below(hbits[jj], hpre[jj], at) -
                                                         holes_below(hbits[jj], hpre[jj], fin));
                    const uint32_t to = kth_hole(hbits[jj], hpre[jj], (jj ? nc1 : nc0) + above);
                    store_key<W>(a.x, a.z, a.cap, sj + to, k);
                    store_coeff<T>(a.c, sj + to, c);
                }
            }
            if (threadIdx.x == 0) {
                a.len[b0] = fin0;
                if (pair) a.len[b1] = fin1;
                if (a.count_totals) {
                    atomicAdd(&a.stats[0], (unsigned long long)(fin0 + fin1));
                    atomicMax(&a.stats[1], (unsigned long long)max(fin0, fin1));
                }
            }
            continue;
        }
        // ---- phase C moved: the pair's rows that stay, written whole to fresh segments ----
        const uint32_t seg0 = moved_segment_rows(fin0);
        const uint32_t seg1 = pair ? moved_segment_rows(fin1) : 0u;
        if (threadIdx.x == 0) {
            const uint64_t base = a.tail_base + atomicAdd(&a.stats[5], (unsigned long long)(seg0 + seg1));
            if (base + seg0 + seg1 > a.tail_limit) {
                atomicOr(&a.stats[2], 2ull);  // the host's bound on moved rows was wrong: nothing is written
                reloc_base = ~0ull;
            } else {
                reloc_base = base;
            }
        }
        __syncthreads();
        const uint64_t nb0 = reloc_base;
        if (nb0 == ~0ull) continue;
        const uint64_t nb1 = nb0 + seg0;
        // commuting rows that stay, in the order of the pair's rows (the same reads as phase A: nothing was written)
        uint32_t cur0 = 0, cur1 = 0;
        for (uint32_t base = 0; base < n; base += kBlock) {
            const uint32_t r = base + threadIdx.x;
            const bool in = r < n;
            const uint32_t j = r >= n0 ? 1u : 0u;
            const uint64_t row = j ? s1 + (r - n0) : s0 + r;
            Key<W> k;
            T c = T(0);
            bool valid = in;
            if (in) {
                k = load_key<W>(a.x, a.z, a.cap, row);
                c = a.c[row];
                valid = selected(row, c);
            }
            const bool stay = valid && !anticommutes_on<W>(k, g, supp) && survives<W, T>(a.keep, k, c);
            bool f[2] = {stay && j == 0, stay && j == 1};
            uint32_t off[2], tot[2];
            block_flag_scan<2>(f, off, tot, scan_scratch);
            if (stay) {
                const uint64_t pos = j ? nb1 + cur1 + off[1] : nb0 + cur0 + off[0];
                store_key<W>(a.x, a.z, a.cap, pos, k);
                store_coeff<T>(a.c, pos, c);
            }
            cur0 += tot[0];
            cur1 += tot[1];
        }
        // then the kept own rows, then the new rows (their ranks from phase B)
        for (uint32_t lb = 0; lb < nl; lb += kBlock) {
            const uint32_t i = lb + threadIdx.x;
            const bool valid = i < nl;
            const uint32_t fl = valid ? Lflag[i] : 0u;
            const uint32_t jj = valid ? Lm[i] : 0u;
            bool f[2] = {(fl & 1u) && jj == 0, (fl & 1u) && jj == 1};
            uint32_t off[2], tot[2];
            block_flag_scan<2>(f, off, tot, scan_scratch);
            if (fl & 1u) {
                const uint64_t pos = jj ? nb1 + cur1 + off[1] : nb0 + cur0 + off[0];
                Key<W> kk;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    kk.x[w] = Lx[w * S + i];
                    kk.z[w] = Lz[w * S + i];
                }
                store_key<W>(a.x, a.z, a.cap, pos, kk);
                store_coeff<T>(a.c, pos, Lown[i]);
            }
            cur0 += tot[0];
            cur1 += tot[1];
        }
        for (uint32_t i = threadIdx.x; i < nl; i += kBlock) {
            if (!(Lflag[i] & 2u)) continue;
            const uint32_t jn = pair ? (Lm[i] ^ 1u) : Lm[i];
            const uint64_t pos = jn ? nb1 + cur1 + Lrank[i] : nb0 + cur0 + Lrank[i];
            Key<W> q;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                q.x[w] = Lx[w * S + i] ^ g.x[w];
                q.z[w] = Lz[w * S + i] ^ g.z[w];
            }
            store_key<W>(a.x, a.z, a.cap, pos, q);
            store_coeff<T>(a.c, pos, Lfresh[i]);
        }
        if (threadIdx.x == 0) {
            a.start[b0] = nb0;
            a.len[b0] = fin0;
            a.bcap[b0] = seg0;
            if (pair) {
                a.start[b1] = nb1;
                a.len[b1] = fin1;
                a.bcap[b1] = seg1;
            }
            atomicAdd(&a.stats[6], 1ull);
            if (a.count_totals) {
                atomicAdd(&a.stats[0], (unsigned long long)(fin0 + fin1));
                atomicMax(&a.stats[1], (unsigned long long)max(fin0, fin1));
            }
        }
    }
}

// Many sums' layers in one launch: workgroup w joins pair w - pair_start[i] of item i, where item i is the one
// whose range [pair_start[i], pair_start[i + 1]) holds w.  Every item stages its largest pair (no redo).
template <int W, typename T, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_rotation_join_batch(const JoinArgs<W, T>* __restrict__ items,
                                                              const uint32_t* __restrict__ pair_start,
                                                              uint32_t num_items) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t scan_scratch[4 * (BLOCK / 64)];
    uint32_t lo = 0, hi = num_items;  // the last item whose start is at most blockIdx.x
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (pair_start[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const JoinArgs<W, T>& a = items[lo];
    const uint32_t t = blockIdx.x - pair_start[lo];
    join_items<W, T, false, BLOCK, true>(a, t, a.num_cosets, a.num_cosets, lds, scan_scratch);
}

// The batch's bucket offsets and counters, one workgroup per item: the output segments of a join (as
// k_offsets_* mode 1: len[b] + len[b ^ delta], or 2 len[b], rounded to kSegAlign) and, after the join, the total
// and longest bucket of its output.  Small sums are the batch's use, so one workgroup scans an item's buckets.
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_offsets_batch(const JoinArgs<W, T>* __restrict__ items) {
    __shared__ uint64_t wave_sums[kWaves];
    const JoinArgs<W, T>& a = items[blockIdx.x];
    const uint32_t B = a.delta ? 2u * a.num_cosets : a.num_cosets;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < B; base += kBlock) {
        const uint32_t b = base + threadIdx.x;
        uint64_t v = 0;
        if (b < B) {
            v = a.delta ? (uint64_t)a.len[b] + a.len[b ^ a.delta] : 2ull * a.len[b];
            v = (v + (kSegAlign - 1)) & ~(uint64_t)(kSegAlign - 1);
        }
        uint64_t total;
        const uint64_t ex = block_exclusive_scan_u64(v, wave_sums, &total);
        if (b < B) const_cast<uint64_t*>(a.ostart)[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x < 8) a.stats[threadIdx.x] = 0;
}

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_reduce_batch(const JoinArgs<W, T>* __restrict__ items,
                                                        unsigned long long* __restrict__ out) {
    __shared__ uint64_t wave_sums[kWaves];
    __shared__ uint32_t wave_max[kWaves];
    const JoinArgs<W, T>& a = items[blockIdx.x];
    const uint32_t B = a.delta ? 2u * a.num_cosets : a.num_cosets;
    uint64_t s = 0;
    uint32_t m = 0;
    for (uint32_t b = threadIdx.x; b < B; b += kBlock) {
        const uint32_t v = a.olen[b];
        s += v;
        m = max(m, v);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, 64);
        m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
    }
    if (lane_id() == 0) {
        wave_sums[wave_id()] = s;
        wave_max[wave_id()] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ts = 0;
        uint32_t tm = 0;
        for (int w = 0; w < kWaves; ++w) {
            ts += wave_sums[w];
            tm = max(tm, wave_max[w]);
        }
        a.stats[0] = ts;
        a.stats[1] = tm;
        out[4 * blockIdx.x + 0] = ts;
        out[4 * blockIdx.x + 1] = tm;
        out[4 * blockIdx.x + 2] = a.stats[2];
    }
}

// ---------------------------------------------------------------------------------------------------------------
// A key-preserving layer (a table whose only entry is the identity delta, e.g. a rotation whose sine snapped to zero):
// rescale in place by the pattern's amplitude, with the zero drop and keep_term (engine/bucketed.rs:646-662).  One
// workgroup per bucket compacts its segment in order.
// ---------------------------------------------------------------------------------------------------------------

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_rescale(const RescaleArgs<W, T> a) {
    __shared__ uint32_t scan_scratch[kWaves];
    __shared__ T amp[16];
    if (threadIdx.x < 16) amp[threadIdx.x] = (T)a.step.amp[threadIdx.x];
    __syncthreads();
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    uint32_t cur = 0;
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t r = base + threadIdx.x;
        Key<W> k;
        T v = T(0);
        bool keep = false;
        if (r < n) {
            k = load_key<W>(a.x, a.z, a.cap, s + r);
            const T c = a.c[s + r];
            v = c * amp[pattern_of<W>(k, a.step.word, a.step.bit, a.step.mask)];
            keep = survives<W, T>(a.keep, k, v);
        }
        bool f[1] = {keep};
        uint32_t off[1], tot[1];
        block_flag_scan<1>(f, off, tot, scan_scratch);  // every read of the chunk precedes its writes
        if (keep) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                a.x[w * a.cap + s + cur + off[0]] = k.x[w];
                a.z[w * a.cap + s + cur + off[0]] = k.z[w];
            }
            a.c[s + cur + off[0]] = v;
        }
        cur += tot[0];
    }
    if (threadIdx.x == 0) a.len[b] = cur;
}

// ---------------------------------------------------------------------------------------------------------------
// A permutation run (engine/bucketed/permutation_run.rs:1-4, :216-293, :297-...): steps that map distinct keys to
// distinct keys, so no two rows ever meet and each row walks the whole run alone, in registers.  Per step the row's
// pattern on the step's support selects one output: the key moved by that pattern's local delta with the
// coefficient times its amplitude (a Clifford rotation's arm that drops the identity residue, a key-preserving
// rescale, or any other one-to-one local table); an inert pattern leaves the row as it is.  A row is dropped where
// the per-layer path drops it: after a step that changed it, when it no longer survives, and when it never changed
// or first changes after step 0, where the first layer would have refused it unchanged.  The walk counts every
// row's output bucket and every step's drops; k_perm_scatter then moves the rows into their buckets.  The steps
// pass through LDS kWalkChunk at a time, each row's state staying in registers across the chunks.
// ---------------------------------------------------------------------------------------------------------------

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_perm_walk(const PermWalkArgs<W, T> a,
                                                      const TableStep* __restrict__ steps) {
    __shared__ TableStep lsteps[kWalkChunk];
    __shared__ uint32_t died[kMaxPermSteps];
    for (uint32_t j = threadIdx.x; j < a.num_steps; j += kBlock) died[j] = 0;
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s0 = a.start[b];
    constexpr uint32_t kStepWords = sizeof(TableStep) / 8;
    static_assert(sizeof(TableStep) % 8 == 0, "TableStep is copied as 8-byte words");
    for (uint32_t base = 0; base < n; base += kBlock) {  // uniform: every thread reaches the barriers below
        const uint32_t r = base + threadIdx.x;
        const bool live = r < n;
        const uint64_t i = s0 + r;
        Key<W> k;
        T c = T(0);
        if (live) {
            k = load_key<W>(a.x, a.z, a.cap, i);
            c = a.c[i];
        }
        uint32_t dest = b;
        bool tested = false;  // whether the row, as it now is, has been asked survives
        int dead_at = live ? -1 : -2;
        for (uint32_t j0 = 0; j0 < a.num_steps; j0 += kWalkChunk) {
            const uint32_t m = min(kWalkChunk, a.num_steps - j0);
            __syncthreads();  // the previous chunk is no longer read
            {
                const uint64_t* src = reinterpret_cast<const uint64_t*>(steps + j0);
                uint64_t* dst = reinterpret_cast<uint64_t*>(lsteps);
                for (uint32_t w = threadIdx.x; w < m * kStepWords; w += kBlock) dst[w] = src[w];
            }
            __syncthreads();
            for (uint32_t jj = 0; jj < m; ++jj) {
                if (dead_at != -1) continue;
                const TableStep& st = lsteps[jj];
                const uint32_t p = pattern_of<W>(k, st.word, st.bit, st.mask);
                if ((st.inert >> p) & 1u) continue;
                const uint32_t j = j0 + jj;
                if (!tested && j > 0 && !survives<W, T>(a.keep, k, c)) {
                    dead_at = 0;
                    continue;
                }
                c = c * (T)st.amp[p];
                xor_local_delta<W>(k, st.delta[p], st.word, st.bit);
                dest ^= st.bucket[p];
                tested = true;
                if (!survives<W, T>(a.keep, k, c)) dead_at = (int)j;
            }
        }
        if (dead_at == -1 && !tested && !survives<W, T>(a.keep, k, c)) dead_at = 0;
        if (dead_at >= 0) {
            atomicAdd(&died[dead_at], 1u);
            a.dest[i] = kDeadRow;
        } else if (dead_at == -1) {
            store_key<W>(a.x, a.z, a.cap, i, k);
            store_coeff<T>(a.c, i, c);
            a.dest[i] = dest;
            atomicAdd(&a.counts[dest], 1u);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < a.num_steps; j += kBlock)
        if (died[j]) atomicAdd(&a.deaths[j], (unsigned long long)died[j]);
}

// Each surviving row to the next free slot of its output bucket (rows within a bucket are in no particular order).
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_perm_scatter(const PermScatterArgs<W, T> a) {
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s0 = a.start[b];
    for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
        const uint64_t i = s0 + r;
        const uint32_t d = a.dest[i];
        if (d == kDeadRow) continue;
        const Key<W> k = load_key<W>(a.x, a.z, a.cap, i);
        const T c = a.c[i];
        const uint64_t pos = a.ostart[d] + atomicAdd(&a.olen[d], 1u);
        store_key<W>(a.ox, a.oz, a.ocap, pos, k);
        store_coeff<T>(a.oc, pos, c);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Refine by one hash row: bucket b splits into b (row parity 0) and b + B (parity 1) in one pass; the parity-0
// rows fill the segment from the front, the parity-1 rows from the back.
// ---------------------------------------------------------------------------------------------------------------

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_refine(const RefineArgs<W, T> a) {
    __shared__ uint32_t scan_scratch[2 * kWaves];
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    uint32_t cur0 = 0, cur1 = 0;
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t r = base + threadIdx.x;
        Key<W> k;
        T c = T(0);
        uint32_t par = 0;
        if (r < n) {
            k = load_key<W>(a.x, a.z, a.cap, s + r);
            c = a.c[s + r];
            par = row_parity<W>(k, a.rows, a.row);
        }
        bool f[2] = {r < n && par == 0, r < n && par == 1};
        uint32_t off[2], tot[2];
        block_flag_scan<2>(f, off, tot, scan_scratch);
        if (r < n) {
            const uint64_t pos = par ? s + n - 1u - (cur1 + off[1]) : s + cur0 + off[0];
            store_key<W>(a.ox, a.oz, a.ocap, pos, k);
            store_coeff<T>(a.oc, pos, c);
        }
        cur0 += tot[0];
        cur1 += tot[1];
    }
    if (threadIdx.x == 0) {
        a.ostart[b] = s;
        a.olen[b] = cur0;
        a.ostart[b + a.num_buckets] = s + cur0;
        a.olen[b + a.num_buckets] = cur1;
        if (a.obcap) {  // the segment's room after its rows goes with the second half
            a.obcap[b] = cur0;
            a.obcap[b + a.num_buckets] = a.bcap[b] - cur0;
        }
    }
}

// Coarsen: bucket b of the old layout is copied into its coarse bucket b & (coarse - 1) after the old buckets of
// that coarse bucket with a smaller index.
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_coarsen(const CoarsenArgs<W, T> a) {
    const uint32_t b = blockIdx.x;
    const uint32_t cb = b & (a.coarse_buckets - 1u);
    uint64_t dst = a.ostart[cb];
    for (uint32_t e = cb; e < b; e += a.coarse_buckets) dst += a.len[e];
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
        store_key<W>(a.ox, a.oz, a.ocap, dst + r, load_key<W>(a.x, a.z, a.cap, s + r));
        store_coeff<T>(a.oc, dst + r, a.c[s + r]);
    }
    if (threadIdx.x == 0 && b < a.coarse_buckets) {
        uint32_t total = 0;
        for (uint32_t e = cb; e < a.num_buckets; e += a.coarse_buckets) total += a.len[e];
        a.olen[cb] = total;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Read-out and export.
// ---------------------------------------------------------------------------------------------------------------

// <psi|O|psi> for a product state (readout/product_state.rs:84-101): a term counts iff it matches the axis on its
// support, with the sign of the negative mask's parity.  Each workgroup sums a fixed run of buckets in a fixed
// order, so the result is deterministic.
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_expectation(const ExpectationArgs<W, T> a) {
    __shared__ double part[kBlock];
    double acc = 0.0;
    const uint32_t b_lo = blockIdx.x * a.buckets_per_block;
    const uint32_t b_hi = min(a.num_buckets, b_lo + a.buckets_per_block);
    for (uint32_t b = b_lo; b < b_hi; ++b) {
        const uint32_t n = a.len[b];
        const uint64_t s = a.start[b];
        for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
            const Key<W> k = load_key<W>(a.x, a.z, a.cap, s + r);
            uint64_t mismatch = 0;
            uint32_t sign = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint64_t sup = k.x[w] | k.z[w];
                mismatch |= (k.x[w] ^ (sup & a.axis_x[w])) | (k.z[w] ^ (sup & a.axis_z[w]));
                sign += __popcll(sup & a.negative[w]);
            }
            if (mismatch == 0) {
                const double v = (double)a.c[s + r];
                acc += (sign & 1u) ? -v : v;
            }
        }
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[blockIdx.x] = part[0];
}

// Dense, term-major copy of the sum with each term's bucket under `ref_bits` rows of the reference hash, for the
// canonical download (bucket ascending, then lexicographic key) done on the host.
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_export(const ExportArgs<W, T> a) {
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    const uint64_t d = a.dense_start[b];
    for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
        const Key<W> k = load_key<W>(a.x, a.z, a.cap, s + r);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            a.out_x[(d + r) * W + w] = k.x[w];
            a.out_z[(d + r) * W + w] = k.z[w];
        }
        a.out_c[d + r] = a.c[s + r];
        a.out_bucket[d + r] = bucket_of<W>(k, a.rows, a.ref_bits);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------------------------

uint32_t offsets_tiles(uint32_t out_n) { return (out_n + kScanTile - 1) / kScanTile; }

hipError_t launch_bucket_offsets(hipStream_t st, const uint32_t* len, uint32_t num_buckets, uint32_t mode,
                                 uint32_t delta, uint32_t coarse_buckets, uint32_t fold, uint64_t* out_start,
                                 unsigned long long* total, uint64_t* tile_sums, unsigned long long* clear,
                                 uint32_t clear_n, uint32_t* out_cap) {
    OffsetsArgs a{};
    a.out_cap = mode == 0 ? nullptr : out_cap;
    a.clear = clear;
    a.clear_n = clear ? clear_n : 0u;
    a.len = len;
    a.out_n = mode == 2 ? coarse_buckets : num_buckets;
    a.mode = mode;
    a.delta = delta;
    a.coarse_buckets = coarse_buckets;
    a.fold = fold;
    a.out_start = out_start;
    a.tile_sums = tile_sums;
    a.total = total;
    const uint32_t tiles = offsets_tiles(a.out_n);
    if (tiles == 0) return hipSuccess;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_offsets_one, dim3(1), dim3(kBlock), 0, st, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_offsets_tiles, dim3(tiles), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(k_offsets_top, dim3(1), dim3(kBlock), 0, st, a, tiles);
    hipLaunchKernelGGL(k_offsets_final, dim3(tiles), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

// A fill kernel for the counters: a memset node recorded into a captured graph is not replayed reliably.
__global__ void k_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(hipStream_t st, unsigned long long* p, uint32_t n, unsigned long long v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 63) / 64), dim3(64), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_reduce_lens(hipStream_t st, const uint32_t* len, uint32_t num_buckets, unsigned long long* out,
                              unsigned long long* prof) {
    const uint32_t blocks = offsets_tiles(num_buckets);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_lens, dim3(blocks), dim3(kBlock), 0, st, len, num_buckets, out, prof);
    return hipGetLastError();
}

template <int W, typename T>
size_t join_lds_bytes(uint32_t scap, uint32_t tcap_bits) {
    return (size_t)scap * (2 * W * sizeof(uint64_t) + sizeof(T) + sizeof(uint32_t) + 1) + ((size_t)4 << tcap_bits);
}

// A persistent grid: as many workgroups as stay resident at once (at most groups_per_cu per CU when that is
// nonzero), never more than there are pairs; kJoinOnePairPerGroup launches one workgroup per pair instead.
template <int W, typename T>
hipError_t launch_rotation_join(hipStream_t st, const JoinArgs<W, T>& a, uint32_t groups_per_cu, uint32_t block) {
    if (a.num_cosets == 0) return hipSuccess;
    const size_t lds = join_lds_bytes<W, T>(a.scap, a.tcap_bits);
    static int cus = 0;
    static int resident[2][2][2][17] = {};  // by block, launch kind, filter and scap / 64 (scap <= 1024)
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    const bool wide = block == 512, filter = a.filter != nullptr;
    using Kernel = void (*)(const JoinArgs<W, T>);
    static constexpr Kernel kernels[2][2][2] = {
        {{k_rotation_join<W, T, false, 256, false>, k_rotation_join<W, T, false, 256, true>},
         {k_rotation_join<W, T, false, 512, false>, k_rotation_join<W, T, false, 512, true>}},
        {{k_rotation_join<W, T, true, 256, false>, k_rotation_join<W, T, true, 256, true>},
         {k_rotation_join<W, T, true, 512, false>, k_rotation_join<W, T, true, 512, true>}}};
    const Kernel kernel = kernels[a.redo_mode ? 1 : 0][wide ? 1 : 0][filter ? 1 : 0];
    int& per = resident[wide ? 1 : 0][a.redo_mode ? 1 : 0][filter ? 1 : 0][std::min<uint32_t>(a.scap >> 6, 16u)];
    if (per == 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, lds) != hipSuccess || per <= 0))
        per = 1;
    const int per_cu = groups_per_cu ? std::min<int>(per, (int)std::min<uint32_t>(groups_per_cu, 1024u)) : per;
    // the redo launch does not know its pair count: a persistent grid that reads it on the device
    const uint32_t grid = a.redo_mode ? (uint32_t)(cus * per_cu)
                          : groups_per_cu == kJoinOnePairPerGroup
                              ? a.num_cosets
                              : std::min<uint32_t>(a.num_cosets, (uint32_t)(cus * per_cu));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_rotation_join_small(hipStream_t st, const SmallJoinArgs<W, T>& a) {
    if (a.j.num_cosets == 0 || a.num_buckets > kSmallJoinBuckets || a.j.redo_mode || a.j.filter || !a.j.count_totals)
        return hipErrorInvalidValue;
    const size_t lds = join_lds_bytes<W, T>(a.j.scap, a.j.tcap_bits);
    hipLaunchKernelGGL((k_rotation_join_small<W, T>), dim3(a.j.num_cosets + 1), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_publish_counters(hipStream_t st, unsigned long long* counters, unsigned long long* host,
                                   unsigned long long seq) {
    hipLaunchKernelGGL(k_publish_counters, dim3(1), dim3(64), 0, st, counters, host, seq);
    return hipGetLastError();
}

// One workgroup per pair; the redo launch, over the queued pairs, is a grid of as many as stay resident.
template <int W, typename T>
hipError_t launch_rotation_inplace(hipStream_t st, const InplaceArgs<W, T>& a) {
    if (a.num_cosets == 0) return hipSuccess;
    if (a.scap > kPairRowsMax || a.tcap_bits > 11) return hipErrorInvalidValue;
    const size_t lds = inplace_lds_bytes<W, T>(a.scap, a.tcap_bits);
    const bool filter = a.filter != nullptr;
    if (!a.redo_mode) {
        if (filter)
            hipLaunchKernelGGL((k_rotation_inplace<W, T, false, true>), dim3(a.num_cosets), dim3(kBlock), lds, st, a);
        else
            hipLaunchKernelGGL((k_rotation_inplace<W, T, false, false>), dim3(a.num_cosets), dim3(kBlock), lds, st, a);
        return hipGetLastError();
    }
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per = 0;
    const hipError_t oe = filter ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                       &per, k_rotation_inplace<W, T, true, true>, kBlock, lds)
                                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                       &per, k_rotation_inplace<W, T, true, false>, kBlock, lds);
    if (oe != hipSuccess || per <= 0) per = 1;
    if (filter)
        hipLaunchKernelGGL((k_rotation_inplace<W, T, true, true>), dim3(cus * per), dim3(kBlock), lds, st, a);
    else
        hipLaunchKernelGGL((k_rotation_inplace<W, T, true, false>), dim3(cus * per), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_join_batch(hipStream_t st, const JoinArgs<W, T>* items, const uint32_t* pair_start,
                             uint32_t num_items, uint32_t total_pairs, uint32_t scap_max, uint32_t tcap_bits_max,
                             unsigned long long* out) {
    if (num_items == 0) return hipSuccess;
    hipLaunchKernelGGL((k_offsets_batch<W, T>), dim3(num_items), dim3(kBlock), 0, st, items);
    if (total_pairs) {
        const size_t lds = join_lds_bytes<W, T>(scap_max, tcap_bits_max);
        hipLaunchKernelGGL((k_rotation_join_batch<W, T, 256>), dim3(total_pairs), dim3(256), lds, st, items,
                           pair_start, num_items);
    }
    hipLaunchKernelGGL((k_reduce_batch<W, T>), dim3(num_items), dim3(kBlock), 0, st, items, out);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_rescale(hipStream_t st, const RescaleArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_rescale<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_perm_walk(hipStream_t st, const PermWalkArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0 || a.num_steps == 0) return hipSuccess;
    if (a.num_steps > kMaxPermSteps) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_perm_walk<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a, a.steps);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_perm_scatter(hipStream_t st, const PermScatterArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_perm_scatter<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_refine(hipStream_t st, const RefineArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_refine<W, T>), dim3(a.num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_coarsen(hipStream_t st, const CoarsenArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_coarsen<W, T>), dim3(a.num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_expectation(hipStream_t st, const ExpectationArgs<W, T>& a, uint32_t blocks) {
    hipLaunchKernelGGL((k_expectation<W, T>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_export(hipStream_t st, const ExportArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_export<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

// One workgroup per bucket: the segment's bounds, then every row's bucket, coefficient and key against the others
// (an LDS table of row indices, keys compared in the columns).
template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_check_layout(const CheckArgs<W, T> a) {
    constexpr uint32_t kRows = 1024, kSlots = 2048;
    __shared__ uint32_t table[kSlots];
    __shared__ uint32_t counts[kCheckCounts];
    const uint32_t b = blockIdx.x;
    const uint64_t s0 = a.start[b];
    const uint32_t n = a.len[b];
    const uint64_t extent = a.room ? a.room[b] : n;
    if (threadIdx.x < kCheckCounts) counts[threadIdx.x] = 0u;
    for (uint32_t i = threadIdx.x; i < kSlots; i += kBlock) table[i] = 0u;
    __syncthreads();
    const bool seg_ok = (n == 0 || s0 % kSegAlign == 0) && n <= extent && (extent == 0 || s0 + extent <= a.end) &&
                        a.end <= a.cap;
    if (!seg_ok) {
        if (threadIdx.x == 0) atomicAdd(&a.out[kCheckSegments], 1ull);
        return;  // the whole workgroup: its rows are not read
    }
    const bool scan = n <= kRows;
    for (uint32_t r = threadIdx.x; r < n; r += kBlock) {
        const Key<W> k = load_key<W>(a.x, a.z, a.cap, s0 + r);
        if (bucket_of<W>(k, a.rows, a.bits) != b) atomicAdd(&counts[kCheckMisplaced], 1u);
        if (a.c[s0 + r] == T(0)) atomicAdd(&counts[kCheckZero], 1u);
        if (!scan) continue;
        uint32_t slot = fold_mix(key_fold<W>(k)) >> (32u - 11u);
        for (;;) {
            const uint32_t prev = atomicCAS(&table[slot], 0u, r + 1u);
            if (prev == 0u) break;
            const Key<W> q = load_key<W>(a.x, a.z, a.cap, s0 + prev - 1u);
            if (key_eq<W>(k, q)) {
                atomicAdd(&counts[kCheckDuplicates], 1u);
                break;
            }
            slot = (slot + 1u) & (kSlots - 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && !scan) atomicAdd(&a.out[kCheckUnscanned], 1ull);
    if (threadIdx.x < kCheckCounts && counts[threadIdx.x]) atomicAdd(&a.out[threadIdx.x], (unsigned long long)counts[threadIdx.x]);
}

template <int W, typename T>
hipError_t launch_check_layout(hipStream_t st, const CheckArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_check_layout<W, T>), dim3(a.num_buckets), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

#define PP_INSTANTIATE(W, T)                                                                           \
    template size_t join_lds_bytes<W, T>(uint32_t, uint32_t);                                          \
    template hipError_t launch_rotation_join<W, T>(hipStream_t, const JoinArgs<W, T>&, uint32_t, uint32_t); \
    template hipError_t launch_rotation_join_small<W, T>(hipStream_t, const SmallJoinArgs<W, T>&);      \
    template hipError_t launch_join_batch<W, T>(hipStream_t, const JoinArgs<W, T>*, const uint32_t*, uint32_t,     \
                                                uint32_t, uint32_t, uint32_t, unsigned long long*);             \
    template hipError_t launch_rescale<W, T>(hipStream_t, const RescaleArgs<W, T>&, uint32_t);          \
    template size_t inplace_lds_bytes<W, T>(uint32_t, uint32_t);                                        \
    template hipError_t launch_rotation_inplace<W, T>(hipStream_t, const InplaceArgs<W, T>&);           \
    template hipError_t launch_perm_walk<W, T>(hipStream_t, const PermWalkArgs<W, T>&, uint32_t);       \
    template hipError_t launch_perm_scatter<W, T>(hipStream_t, const PermScatterArgs<W, T>&, uint32_t); \
    template hipError_t launch_refine<W, T>(hipStream_t, const RefineArgs<W, T>&);                      \
    template hipError_t launch_coarsen<W, T>(hipStream_t, const CoarsenArgs<W, T>&);                    \
    template hipError_t launch_expectation<W, T>(hipStream_t, const ExpectationArgs<W, T>&, uint32_t);  \
    template hipError_t launch_export<W, T>(hipStream_t, const ExportArgs<W, T>&, uint32_t);                 \
    template hipError_t launch_check_layout<W, T>(hipStream_t, const CheckArgs<W, T>&);

PP_INSTANTIATE(1, double)
PP_INSTANTIATE(1, float)
PP_INSTANTIATE(2, double)
PP_INSTANTIATE(2, float)

}  // namespace pp
