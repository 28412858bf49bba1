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
// mode 2: coarsened bucket b < coarse_buckets gathers b + e * coarse_buckets for e < fold.
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
    uint64_t* tile_sums;
    unsigned long long* total;
};

__device__ __forceinline__ uint64_t bucket_size(const OffsetsArgs& a, uint32_t b) {
    if (a.mode == 0) return a.len[b];
    if (a.mode == 1) return a.delta ? (uint64_t)a.len[b] + a.len[b ^ a.delta] : 2ull * a.len[b];
    uint64_t s = 0;
    for (uint32_t e = 0; e < a.fold; ++e) s += a.len[b + e * a.coarse_buckets];
    return s;
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
    if (threadIdx.x == 0 && a.total) *a.total = carry;
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
            base += sz[i];
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
// Phase A streams both buckets of the pair: a row commuting with g is final (kept iff nonzero and keep_term);
// an anticommuting row is staged in LDS.  Phase B builds an open-addressing LDS table over the staged keys and
// walks them: row k's output is own*cos (0 on the Clifford arm) plus, if k ^ g is staged, that row's c * a1;
// if k ^ g is not staged it is a new key k ^ g with c * a1 in the other bucket.
// ---------------------------------------------------------------------------------------------------------------

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_rotation_join(const JoinArgs<W, T> a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t scan_scratch[4 * kWaves];
    const uint32_t lcap = a.lcap;
    uint64_t* const Lx = reinterpret_cast<uint64_t*>(lds);   // [W][lcap]
    uint64_t* const Lz = Lx + (size_t)W * lcap;              // [W][lcap]
    T* const Lc = reinterpret_cast<T*>(Lz + (size_t)W * lcap);
    uint32_t* const table = reinterpret_cast<uint32_t*>(Lc + lcap);
    uint8_t* const Lm = reinterpret_cast<uint8_t*>(table + (1u << a.tcap_bits));

    const uint32_t p = blockIdx.x;
    uint32_t b0, b1;
    if (a.delta) {
        const uint32_t lo = p & ((1u << a.hb) - 1u);
        b0 = ((p >> a.hb) << (a.hb + 1u)) | lo;
        b1 = b0 ^ a.delta;
    } else {
        b0 = b1 = p;
    }
    const bool pair = a.delta != 0;
    const uint32_t n0 = a.len[b0];
    const uint32_t n1 = pair ? a.len[b1] : 0u;
    const uint32_t n = n0 + n1;
    const uint64_t s0 = a.start[b0], s1 = a.start[b1];
    const uint64_t o0 = a.ostart[b0], o1 = a.ostart[b1];
    if (n > lcap) {  // the host sizes lcap for the largest pair; never taken, guards the LDS bounds
        if (threadIdx.x == 0) atomicOr(&a.stats[2], 1ull);
        return;
    }
    Key<W> g;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        g.x[w] = a.gx[w];
        g.z[w] = a.gz[w];
    }

    uint32_t cur0 = 0, cur1 = 0, nl = 0;
    // ---- phase A ----
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t r = base + threadIdx.x;
        const bool valid = r < n;
        const uint32_t j = r >= n0 ? 1u : 0u;
        Key<W> k;
        T c = T(0);
        bool anti = false, keep = false;
        if (valid) {
            const uint64_t row = j ? s1 + (r - n0) : s0 + r;
            k = load_key<W>(a.x, a.z, a.cap, row);
            c = a.c[row];
            anti = anticommutes<W>(k, g);
            keep = !anti && survives<W, T>(a.keep, k, c);
        }
        bool f[3] = {keep && j == 0, keep && j == 1, anti};
        uint32_t off[3], tot[3];
        block_flag_scan<3>(f, off, tot, scan_scratch);
        if (f[0]) {
            store_key<W>(a.ox, a.oz, a.ocap, o0 + cur0 + off[0], k);
            store_coeff<T>(a.oc, o0 + cur0 + off[0], c);
        }
        if (f[1]) {
            store_key<W>(a.ox, a.oz, a.ocap, o1 + cur1 + off[1], k);
            store_coeff<T>(a.oc, o1 + cur1 + off[1], c);
        }
        if (anti) {
            const uint32_t li = nl + off[2];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                Lx[w * lcap + li] = k.x[w];
                Lz[w * lcap + li] = k.z[w];
            }
            Lc[li] = c;
            Lm[li] = (uint8_t)j;
        }
        cur0 += tot[0];
        cur1 += tot[1];
        nl += tot[2];
    }

    if (nl != 0) {
        // ---- the LDS table over the staged keys: slot = (tag << 16) | (index + 1), load at most 1/2 ----
        uint32_t tbits = 1;
        while ((1u << tbits) < 2u * nl) ++tbits;
        const uint32_t tsize = 1u << tbits, tmask = tsize - 1u;
        for (uint32_t i = threadIdx.x; i < tsize; i += kBlock) table[i] = 0u;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nl; i += kBlock) {
            Key<W> k;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                k.x[w] = Lx[w * lcap + i];
                k.z[w] = Lz[w * lcap + i];
            }
            const uint32_t h = key_hash<W>(k);
            const uint32_t entry = (h << 16) | (i + 1u);
            uint32_t s = h >> (32u - tbits);
            while (atomicCAS(&table[s], 0u, entry) != 0u) s = (s + 1u) & tmask;
        }
        __syncthreads();

        // ---- phase B ----
        for (uint32_t base = 0; base < nl; base += kBlock) {
            const uint32_t i = base + threadIdx.x;
            const bool valid = i < nl;
            Key<W> k, q;
            T own = T(0), fresh = T(0);
            uint32_t j = 0;
            bool keep_own = false, keep_new = false;
            if (valid) {
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    k.x[w] = Lx[w * lcap + i];
                    k.z[w] = Lz[w * lcap + i];
                }
                const T c = Lc[i];
                j = Lm[i];
                q = key_xor<W>(k, g);
                const uint32_t h = key_hash<W>(q);
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
                            eq = eq && (Lx[w * lcap + idx] == q.x[w]) && (Lz[w * lcap + idx] == q.z[w]);
                        if (eq) {
                            found = (int)idx;
                            break;
                        }
                    }
                    s = (s + 1u) & tmask;
                }
                own = a.clifford ? T(0) : c * a.cos_t;
                if (found >= 0) own = own + Lc[found] * partner_amplitude<W, T>(q, g, a.sin_t);
                keep_own = survives<W, T>(a.keep, k, own);
                if (found < 0) {
                    fresh = c * partner_amplitude<W, T>(k, g, a.sin_t);
                    keep_new = survives<W, T>(a.keep, q, fresh);
                }
            }
            const uint32_t jn = pair ? (j ^ 1u) : j;
            bool f[4] = {keep_own && j == 0, keep_own && j == 1, keep_new && jn == 0, keep_new && jn == 1};
            uint32_t off[4], tot[4];
            block_flag_scan<4>(f, off, tot, scan_scratch);
            if (f[0]) {
                store_key<W>(a.ox, a.oz, a.ocap, o0 + cur0 + off[0], k);
                store_coeff<T>(a.oc, o0 + cur0 + off[0], own);
            }
            if (f[1]) {
                store_key<W>(a.ox, a.oz, a.ocap, o1 + cur1 + off[1], k);
                store_coeff<T>(a.oc, o1 + cur1 + off[1], own);
            }
            if (f[2]) {
                const uint64_t pos = o0 + cur0 + tot[0] + off[2];
                store_key<W>(a.ox, a.oz, a.ocap, pos, q);
                store_coeff<T>(a.oc, pos, fresh);
            }
            if (f[3]) {
                const uint64_t pos = o1 + cur1 + tot[1] + off[3];
                store_key<W>(a.ox, a.oz, a.ocap, pos, q);
                store_coeff<T>(a.oc, pos, fresh);
            }
            cur0 += tot[0] + tot[2];
            cur1 += tot[1] + tot[3];
        }
    }
    // The counts stay per bucket: a total through one global atomic per workgroup would serialize ~10^5
    // same-address atomics per launch, so k_reduce_lens sums them afterwards.
    if (threadIdx.x == 0) {
        a.olen[b0] = cur0;
        if (pair) a.olen[b1] = cur1;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// A key-preserving layer (a rotation whose sine snapped to zero): rescale in place with the zero drop and
// keep_term (engine/bucketed.rs:646-662).  One workgroup per bucket compacts its segment in order.
// ---------------------------------------------------------------------------------------------------------------

template <int W, typename T>
__global__ void __launch_bounds__(kBlock) k_rescale(const RescaleArgs<W, T> a) {
    __shared__ uint32_t scan_scratch[kWaves];
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t s = a.start[b];
    Key<W> g;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        g.x[w] = a.gx[w];
        g.z[w] = a.gz[w];
    }
    uint32_t cur = 0;
    for (uint32_t base = 0; base < n; base += kBlock) {
        const uint32_t r = base + threadIdx.x;
        Key<W> k;
        T v = T(0);
        bool keep = false;
        if (r < n) {
            k = load_key<W>(a.x, a.z, a.cap, s + r);
            const T c = a.c[s + r];
            v = anticommutes<W>(k, g) ? c * a.a0 : c;
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
                                 unsigned long long* total, uint64_t* tile_sums) {
    OffsetsArgs a{};
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
    hipLaunchKernelGGL(k_offsets_tiles, dim3(tiles), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(k_offsets_top, dim3(1), dim3(kBlock), 0, st, a, tiles);
    hipLaunchKernelGGL(k_offsets_final, dim3(tiles), dim3(kBlock), 0, st, a);
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
size_t join_lds_bytes(uint32_t lcap, uint32_t tcap_bits) {
    return (size_t)lcap * (2 * W * sizeof(uint64_t) + sizeof(T) + 1) + ((size_t)4 << tcap_bits);
}

template <int W, typename T>
hipError_t launch_rotation_join(hipStream_t st, const JoinArgs<W, T>& a, uint32_t num_cosets) {
    const size_t lds = join_lds_bytes<W, T>(a.lcap, a.tcap_bits);
    if (num_cosets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_rotation_join<W, T>), dim3(num_cosets), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_rescale(hipStream_t st, const RescaleArgs<W, T>& a, uint32_t num_buckets) {
    if (num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_rescale<W, T>), dim3(num_buckets), dim3(kBlock), 0, st, a);
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

#define PP_INSTANTIATE(W, T)                                                                           \
    template size_t join_lds_bytes<W, T>(uint32_t, uint32_t);                                          \
    template hipError_t launch_rotation_join<W, T>(hipStream_t, const JoinArgs<W, T>&, uint32_t);       \
    template hipError_t launch_rescale<W, T>(hipStream_t, const RescaleArgs<W, T>&, uint32_t);          \
    template hipError_t launch_refine<W, T>(hipStream_t, const RefineArgs<W, T>&);                      \
    template hipError_t launch_coarsen<W, T>(hipStream_t, const CoarsenArgs<W, T>&);                    \
    template hipError_t launch_expectation<W, T>(hipStream_t, const ExpectationArgs<W, T>&, uint32_t);  \
    template hipError_t launch_export<W, T>(hipStream_t, const ExportArgs<W, T>&, uint32_t);

PP_INSTANTIATE(1, double)
PP_INSTANTIATE(1, float)
PP_INSTANTIATE(2, double)
PP_INSTANTIATE(2, float)

}  // namespace pp
