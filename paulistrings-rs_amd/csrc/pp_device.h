// pp_device.h -- device-side building blocks of the MI355X rotation-layer engine.
//
// Keys are the symplectic Pauli encoding of the reference (crates/paulistrings/src/pauli_string.rs:15-20):
// W 64-bit words of x bits and W of z bits.  Columns live in HBM word-major ("SoA per word"):
// x[w * cap + i], z[w * cap + i], c[i], so every wave loads 512 contiguous bytes per word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

constexpr int kBlock = 256;  // threads per workgroup: four wave64s
constexpr int kWaves = kBlock / 64;
constexpr int kMaxKeepNodes = 15;
// Bucket segments of a resident layout start on a multiple of kSegAlign rows (128 bytes of every 8-byte column),
// so that no cache line holds the end of one bucket and the start of the next: two workgroups, possibly on
// different XCDs with their own L2, would each move that line.
constexpr uint32_t kSegAlign = 16;
constexpr int kMaxHashRows = 31;
constexpr uint32_t kPairRowsMax = 1024;  // the largest bucket pair a join stages (buckets are refined below it)

enum KeepOp : int32_t { KEEP_TRUE = 0, KEEP_COEFF = 1, KEEP_WEIGHT = 2, KEEP_AND = 3, KEEP_OR = 4 };

// A built-in policy's per-term filter as a postfix program (truncation/builtin.rs:19-59, :470-526; the
// reference's device form is src/engine/gpu/truncation.rs:22-102).  eps2 = eps * eps is formed on the host
// in f64, exactly as the reference forms it per term.
struct KeepProgram {
    int32_t n;
    int32_t op[kMaxKeepNodes];
    uint32_t k[kMaxKeepNodes];
    uint32_t disabled[kMaxKeepNodes];  // CoefficientThreshold with eps < 0 keeps everything
    double eps2[kMaxKeepNodes];
};

template <int W>
struct Key {
    uint64_t x[W];
    uint64_t z[W];
};

template <int W>
__device__ __forceinline__ Key<W> load_key(const uint64_t* __restrict__ x, const uint64_t* __restrict__ z,
                                           uint64_t cap, uint64_t i) {
    Key<W> k;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        k.x[w] = x[w * cap + i];
        k.z[w] = z[w * cap + i];
    }
    return k;
}

// Plain stores: with the non-temporal hint the saturated join layer ran 2% slower (4.28 against 4.37 TB/s).
template <int W>
__device__ __forceinline__ void store_key(uint64_t* __restrict__ x, uint64_t* __restrict__ z, uint64_t cap,
                                          uint64_t i, const Key<W>& k) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        x[w * cap + i] = k.x[w];
        z[w * cap + i] = k.z[w];
    }
}

template <typename T>
__device__ __forceinline__ void store_coeff(T* __restrict__ c, uint64_t i, T v) {
    c[i] = v;
}

// commutes_with (pauli_string.rs:116-123): anticommute iff popcount(x.gz ^ z.gx) is odd.
template <int W>
__device__ __forceinline__ bool anticommutes(const Key<W>& k, const Key<W>& g) {
    uint64_t f = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) f ^= (k.x[w] & g.z[w]) ^ (k.z[w] & g.x[w]);
    return __popcll(f) & 1;
}

// The exponent of the phase i^d of k * g (pauli_string.rs:73-93), mod 4.
template <int W>
__device__ __forceinline__ uint32_t product_phase(const Key<W>& k, const Key<W>& g) {
    uint32_t d = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t a = k.x[w], b = k.z[w], c = g.x[w], e = g.z[w];
        d += 2u * __popcll(b & c) + __popcll(a & b) + __popcll(c & e) - __popcll((a ^ c) & (b ^ e));
    }
    return d & 3u;
}

// The generator's support: bit w set when word w of g has an x or z bit.  Outside it c = e = 0, where the
// |a & b| and |(a ^ c) & (b ^ e)| terms of the phase cancel and x.gz ^ z.gx vanishes, so only the support's
// words are read.  The mask is uniform (a kernel argument), so the skipped words cost scalar branches only.
template <int W>
__device__ __forceinline__ uint32_t support_words(const Key<W>& g) {
    uint32_t m = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) m |= ((g.x[w] | g.z[w]) != 0 ? 1u : 0u) << w;
    return m;
}

template <int W>
__device__ __forceinline__ bool anticommutes_on(const Key<W>& k, const Key<W>& g, uint32_t supp) {
    uint32_t f = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (supp & (1u << w)) f += __popcll((k.x[w] & g.z[w]) ^ (k.z[w] & g.x[w]));
    return f & 1u;
}

// product_phase over the support words; gphase = |gx & gz| summed over every word (a per-launch constant).
template <int W>
__device__ __forceinline__ uint32_t product_phase_on(const Key<W>& k, const Key<W>& g, uint32_t supp,
                                                     uint32_t gphase) {
    uint32_t d = gphase;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if (supp & (1u << w)) {
            const uint64_t a = k.x[w], b = k.z[w], c = g.x[w], e = g.z[w];
            d += 2u * __popcll(b & c) + __popcll(a & b) - __popcll((a ^ c) & (b ^ e));
        }
    }
    return d & 3u;
}

// The partner amplitude of an anticommuting k: i * sin * (k g) carries the real phase i^(1 + d),
// a sign (channel/rotation.rs:57-65).
template <int W, typename T>
__device__ __forceinline__ T partner_amplitude(const Key<W>& k, const Key<W>& g, T sin_t) {
    const uint32_t total = (1u + product_phase<W>(k, g)) & 3u;
    return (total & 2u) ? -sin_t : sin_t;
}

template <int W, typename T>
__device__ __forceinline__ T partner_amplitude_on(const Key<W>& k, const Key<W>& g, uint32_t supp,
                                                  uint32_t gphase, T sin_t) {
    const uint32_t total = (1u + product_phase_on<W>(k, g, supp, gphase)) & 3u;
    return (total & 2u) ? -sin_t : sin_t;
}

template <int W>
__device__ __forceinline__ uint32_t weight(const Key<W>& k) {
    uint32_t n = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) n += __popcll(k.x[w] | k.z[w]);
    return n;
}

template <int W>
__device__ __forceinline__ bool key_eq(const Key<W>& a, const Key<W>& b) {
    uint64_t d = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) d |= (a.x[w] ^ b.x[w]) | (a.z[w] ^ b.z[w]);
    return d == 0;
}

template <int W>
__device__ __forceinline__ Key<W> key_xor(const Key<W>& a, const Key<W>& b) {
    Key<W> r;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        r.x[w] = a.x[w] ^ b.x[w];
        r.z[w] = a.z[w] ^ b.z[w];
    }
    return r;
}

// keep: nonzero and the policy's keep_term (rotation_layer/merge.rs:232-239).  The program is uniform
// across the wave (a kernel argument), so the interpreter loop costs scalar branches only.
template <int W, typename T>
__device__ __forceinline__ bool survives(const KeepProgram& p, const Key<W>& k, T c) {
    if (!(c != T(0))) return false;
    if (p.n == 0) return true;
    const double v = (double)c;
    if (p.n == 1 && p.op[0] == KEEP_COEFF) return p.disabled[0] || v * v > p.eps2[0];  // the common policy
    // The operand stack is a bit string (top = bit 0): a private array indexed by a loop counter would
    // live in scratch memory.
    uint32_t stack = 0;
    for (int i = 0; i < p.n; ++i) {
        const int32_t op = p.op[i];
        if (op == KEEP_COEFF) {
            stack = (stack << 1) | (uint32_t)(p.disabled[i] || (v * v > p.eps2[i]));
        } else if (op == KEEP_WEIGHT) {
            stack = (stack << 1) | (uint32_t)(weight<W>(k) <= p.k[i]);
        } else if (op == KEEP_AND) {
            const uint32_t r = stack & (stack >> 1) & 1u;
            stack = ((stack >> 2) << 1) | r;
        } else if (op == KEEP_OR) {
            const uint32_t r = (stack | (stack >> 1)) & 1u;
            stack = ((stack >> 2) << 1) | r;
        } else {
            stack = (stack << 1) | 1u;
        }
    }
    return stack & 1u;
}

// A key's pattern on a local table's two support slots (LocalPtm::support_bits, channel/prepared.rs:141-151): bit 2j
// the x bit of slot j, bit 2j+1 its z bit; `mask` keeps the support's live bits (3 for one qubit, 0 for none).  The
// slots' words and bits are uniform, so selecting a word costs a scalar branch or a v_cndmask.
template <int W>
__device__ __forceinline__ uint32_t pattern_of(const Key<W>& k, const uint32_t (&word)[2], const uint32_t (&bit)[2],
                                               uint32_t mask) {
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        uint64_t xw = k.x[0], zw = k.z[0];
        if (W > 1 && word[j] == 1) {
            xw = k.x[W > 1 ? 1 : 0];
            zw = k.z[W > 1 ? 1 : 0];
        }
        s |= (uint32_t)(((xw >> bit[j]) & 1ull) | (((zw >> bit[j]) & 1ull) << 1)) << (2 * j);
    }
    return s & mask;
}

// k ^ lift(d) (LocalPtm::lift, channel/prepared.rs:284-299) for a local delta d on the same two slots.
template <int W>
__device__ __forceinline__ void xor_local_delta(Key<W>& k, uint32_t d, const uint32_t (&word)[2],
                                                const uint32_t (&bit)[2]) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint64_t mx = (uint64_t)((d >> (2 * j)) & 1u) << bit[j];
        const uint64_t mz = (uint64_t)((d >> (2 * j + 1)) & 1u) << bit[j];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (word[j] == (uint32_t)w) {
                k.x[w] ^= mx;
                k.z[w] ^= mz;
            }
        }
    }
}

// Row parity of a GF(2) hash row: bit r of H.v (pauli_sum/hash.rs:170-179).
template <int W>
__device__ __forceinline__ uint32_t row_parity(const Key<W>& k, const uint64_t* __restrict__ rows, int r) {
    uint64_t f = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) f ^= (k.x[w] & rows[(2 * r) * W + w]) ^ (k.z[w] & rows[(2 * r + 1) * W + w]);
    return __popcll(f) & 1u;
}

template <int W>
__device__ __forceinline__ uint32_t bucket_of(const Key<W>& k, const uint64_t* __restrict__ rows, int bits) {
    uint32_t b = 0;
    for (int r = 0; r < bits; ++r) b |= row_parity<W>(k, rows, r) << r;
    return b;
}

// A 32-bit mix of the key for the LDS join table; any function works, a good one keeps probes short.
// The LDS table's hash: a GF(2)-linear fold of the key's 32-bit halves, each rotated by its own amount (distinct
// for the x and z halves of one qubit range and for neighbouring ranges, so that one- and two-qubit differences
// do not cancel), then a multiplicative finish.  Linearity gives fold(k ^ g) = fold(k) ^ fold(g): a partner's fold
// costs one XOR.
template <int W>
__device__ __forceinline__ uint32_t key_fold(const Key<W>& k) {
    uint32_t f = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint32_t r = 4u * (uint32_t)w;
        f ^= __builtin_rotateleft32((uint32_t)k.x[w], r);
        f ^= __builtin_rotateleft32((uint32_t)(k.x[w] >> 32), r + 8u);
        f ^= __builtin_rotateleft32((uint32_t)k.z[w], r + 16u);
        f ^= __builtin_rotateleft32((uint32_t)(k.z[w] >> 32), r + 24u);
    }
    return f;
}

// Slot mix of a fold: the table index is its top bits, the 16-bit tag its low half.
__device__ __forceinline__ uint32_t fold_mix(uint32_t f) { return f * 0x9E3779B1u; }

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x >> 6; }
__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Ordered compaction of NF one-bit flags across the workgroup: ballot + popcount within a wave64, the wave
// totals through LDS.  `off[i]` is the thread's rank among set flags i, `tot[i]` the block's count.  The wave
// totals travel packed, two 16-bit fields per word (a block count is at most 64 * WAVES), so the cross-wave sums
// cost one add per word rather than per flag.  `scratch` holds at least NF * WAVES words.  Two barriers;
// every thread of the block (WAVES wave64s) must call it.
template <int NF, int WAVES = kWaves>
__device__ __forceinline__ void block_flag_scan(const bool (&f)[NF], uint32_t (&off)[NF], uint32_t (&tot)[NF],
                                                uint32_t* scratch) {
    constexpr int NW = (NF + 1) / 2;
    const uint32_t lane = lane_id();
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_id());
    const uint64_t lt = lanemask_lt();
    uint32_t packed[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) packed[k] = 0;
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const uint64_t m = __ballot(f[i]);
        off[i] = __popcll(m & lt);
        packed[i >> 1] |= (uint32_t)__popcll(m) << (16 * (i & 1));
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NW; ++k) scratch[k * WAVES + wave] = packed[k];
    }
    __syncthreads();
    uint32_t before[NW], all[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        before[k] = 0;
        all[k] = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = scratch[k * WAVES + w];
            if ((uint32_t)w < wave) before[k] += c;
            all[k] += c;
        }
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const int sh = 16 * (i & 1);
        off[i] += (before[i >> 1] >> sh) & 0xFFFFu;
        tot[i] = (all[i >> 1] >> sh) & 0xFFFFu;
    }
    __syncthreads();
}

}  // namespace pp
