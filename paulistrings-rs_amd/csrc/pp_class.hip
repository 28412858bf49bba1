// pp_class.hip -- the class pass (pp_class.h): a program of up to eight local tables applied in one pass, each class
// bucket's strings held in LDS through every layer.
//
// Layout of one apply workgroup's LDS: the program's amplitudes; two row sets of lcap rows (word-major keys, then
// coefficients) that trade places every layer; per layer an open-addressing table of the live set's classes under
// that layer (the key with the layer's support cleared) and one record per class: the row of each of its (at most
// 16) patterns.  A layer gathers every output from its sources: an output K with pattern t has, for each entry e,
// the source of pattern t ^ delta[e] in its class, with amplitude amp[e][t ^ delta[e]]; the source that comes first
// in entry order among those present owns K, sums its contributions in entry order (one rounded multiply and one
// rounded add each, as a rotation's pair update or the host engine's merge), asks keep and appends it to the other
// set.  A rotation table (identity entry first, then the generator) gives own = c cos + c_partner a1 exactly as the
// pair join does.

#include "pp_class.h"

namespace pp {

namespace {

template <int W, typename T>
__device__ __forceinline__ bool class_touched(const ClassArgs<W, T>& a, const Key<W>& k) {
    bool t = false;
    for (uint32_t l = 0; l < a.n_layers; ++l) {
        const ClassLayer& L = a.layers[l];
        t = t || ((L.active >> pattern_of<W>(k, L.word, L.bit, L.mask)) & 1u);
    }
    return t;
}

template <int W>
__device__ __forceinline__ Key<W> class_key(const Key<W>& k, const uint64_t (&qx)[W], const uint64_t (&qz)[W]) {
    Key<W> r;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        r.x[w] = k.x[w] & ~(qx[w] | qz[w]);
        r.z[w] = k.z[w] & ~(qx[w] | qz[w]);
    }
    return r;
}

__device__ __forceinline__ uint32_t lds_slot(uint32_t fold, uint32_t tbits) { return fold_mix(fold) >> (32u - tbits); }

// One wave per input bucket.  Route (GATHER false): count every touched string into its class bucket and every
// untouched one that keep passes into ukeep[b].  Gather (GATHER true): place the touched strings in class-bucket order
// (each class bucket's count counts down to zero as its rows are placed).
template <int W, typename T, bool GATHER>
__global__ void __launch_bounds__(kClassBlock) k_class_route(const ClassArgs<W, T> a) {
    const uint32_t b = blockIdx.x * (kClassBlock / 64) + wave_id();
    if (b >= a.num_buckets) return;  // the whole wave
    const uint32_t lane = lane_id();
    const uint64_t s0 = a.start[b];
    const uint32_t n = a.len[b];
    uint32_t kept = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += 64) {
        const uint32_t r = r0 + lane;
        if (r >= n) continue;
        const Key<W> k = load_key<W>(a.x, a.z, a.cap, s0 + r);
        const T c = a.c[s0 + r];
        if (class_touched<W, T>(a, k)) {
            const uint32_t cb = bucket_of<W>(class_key<W>(k, a.qx, a.qz), a.rows, a.cbits);
            if constexpr (GATHER) {
                const uint64_t pos = a.cstart[cb] + (uint64_t)(atomicSub(&a.ccount[cb], 1u) - 1u);
                store_key<W>(a.sx, a.sz, a.scap, pos, k);
                a.sc[pos] = c;
            } else {
                atomicAdd(&a.ccount[cb], 1u);
            }
        } else if (!GATHER) {
            kept += survives<W, T>(a.keep, k, c) ? 1u : 0u;
        }
    }
    if constexpr (!GATHER) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) kept += (uint32_t)__shfl_xor((int)kept, d, 64);
        if (lane == 0) {
            a.ukeep[b] = kept;
            if (kept) atomicAdd(&a.stats[kClassUntouched], (unsigned long long)kept);
        }
    }
}

// A persistent grid over the class buckets; each workgroup runs the whole program on one class bucket at a time.
// The program's tables are copied to LDS once per workgroup.  Per layer three barriers: the table cleared; the live
// set inserted; the outputs written.  The output counters alternate between two slots, so the next layer resets its
// slot while the last readers of this one are still behind the third barrier.
template <int W, typename T>
__global__ void __launch_bounds__(kClassBlock) k_class_apply(const ClassArgs<W, T> a) {
    extern __shared__ __align__(16) unsigned char class_lds[];
    const uint32_t cap = a.lcap;
    double* const P = reinterpret_cast<double*>(class_lds);  // every layer's amplitudes, entry-major
    uint64_t* const X0 = reinterpret_cast<uint64_t*>(P + a.prog_words);
    uint64_t* const Z0 = X0 + (size_t)W * cap;
    uint64_t* const X1 = Z0 + (size_t)W * cap;
    uint64_t* const Z1 = X1 + (size_t)W * cap;
    T* const C0 = reinterpret_cast<T*>(Z1 + (size_t)W * cap);
    T* const C1 = C0 + cap;
    uint32_t* const table = reinterpret_cast<uint32_t*>(C1 + cap);
    // per layer, one record per class of the live set (the key with the layer's support cleared): a row holding it
    // and the row of each of its patterns (kNoRow where the class has none)
    uint32_t* const rec_rep = table + (1u << a.tbits);
    uint16_t* const rec_mem = reinterpret_cast<uint16_t*>(rec_rep + cap);  // [cap][16]
    uint16_t* const row_rec = rec_mem + (size_t)cap * 16;                  // [cap]
    __shared__ uint32_t s_word[kClassMaxLayers][2];
    __shared__ uint32_t s_bit[kClassMaxLayers][2];
    __shared__ uint32_t s_mask[kClassMaxLayers];
    __shared__ uint32_t s_ne[kClassMaxLayers];
    __shared__ uint32_t s_off[kClassMaxLayers];
    __shared__ uint32_t s_delta[kClassMaxLayers][kClassMaxEntries];
    __shared__ uint32_t s_count[2];
    __shared__ uint32_t s_bad[2];
    __shared__ unsigned long long s_base;
    const uint32_t tid = threadIdx.x;
    const uint32_t nl = a.n_layers;

    if (tid == 0) {
        uint32_t off = 0;
        for (uint32_t l = 0; l < nl; ++l) {
            s_off[l] = off;
            off += a.layers[l].n_entries * 16u;
        }
    }
    if (tid < nl) {
        const ClassLayer& L = a.layers[tid];
        s_word[tid][0] = L.word[0];
        s_word[tid][1] = L.word[1];
        s_bit[tid][0] = L.bit[0];
        s_bit[tid][1] = L.bit[1];
        s_mask[tid] = L.mask;
        s_ne[tid] = L.n_entries;
        for (uint32_t e = 0; e < kClassMaxEntries; ++e) s_delta[tid][e] = L.delta[e];
    }
    __syncthreads();
    for (uint32_t l = 0; l < nl; ++l)
        for (uint32_t i = tid; i < s_ne[l] * 16u; i += kClassBlock) P[s_off[l] + i] = a.layers[l].amp[i >> 4][i & 15u];
    __syncthreads();

    // Chunk q holds the class buckets whose rows start in [q R, (q + 1) R) of the staging (R = chunk_rows): whole
    // classes, independent of one another, share the LDS set and its table.
    const uint64_t staged = a.stats[kClassStaged];
    const uint64_t R = a.chunk_rows;
    const uint64_t chunks = (staged + R - 1) / R;
    __shared__ unsigned long long s_range[2];
    for (uint64_t q = blockIdx.x; q < chunks; q += gridDim.x) {
        if (tid < 2) {  // the first class bucket starting at or after q R (or (q + 1) R), as a row of the staging
            const uint64_t at = (q + tid) * R;
            uint32_t lo = 0, hi = a.num_classes;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.cstart[mid] < at) lo = mid + 1;
                else hi = mid;
            }
            s_range[tid] = lo < a.num_classes ? a.cstart[lo] : staged;
        }
        __syncthreads();
        const uint64_t base = s_range[0];
        const uint64_t end = s_range[1];
        __syncthreads();  // read before the next chunk's search writes them
        const uint32_t n0 = (uint32_t)(end - base);
        if (n0 == 0u) continue;  // uniform
        if (n0 > cap) {
            if (tid == 0) atomicOr(&a.stats[kClassFlags], 1ull);
            continue;
        }
        for (uint32_t i = tid; i < n0; i += kClassBlock) {
            store_key<W>(X0, Z0, cap, i, load_key<W>(a.sx, a.sz, a.scap, base + i));
            C0[i] = a.sc[base + i];
        }
        uint32_t n = n0;
        bool in0 = true, bad = false;
        for (uint32_t l = 0; l < nl; ++l) {
            const uint32_t slot = l & 1u;
            const uint64_t* IX = in0 ? X0 : X1;
            const uint64_t* IZ = in0 ? Z0 : Z1;
            const T* IC = in0 ? C0 : C1;
            uint64_t* OX = in0 ? X1 : X0;
            uint64_t* OZ = in0 ? Z1 : Z0;
            T* OC = in0 ? C1 : C0;
            const double* A = P + s_off[l];  // A[e * 16 + s]
            const uint32_t* D = s_delta[l];
            const uint32_t ne = s_ne[l];
            // a table of at least twice the live rows (a probe then ends), at most the launch's
            uint32_t tb = 4;
            while (tb < a.tbits && (1u << tb) < 2u * n) ++tb;
            for (uint32_t i = tid; i < (1u << tb); i += kClassBlock) table[i] = 0u;
            if (tid == 0) {
                s_count[slot] = 0u;
                s_bad[slot] = 0u;
            }
            // the layer's support, cleared in a class key
            Key<W> sup{};
            xor_local_delta<W>(sup, s_mask[l], s_word[l], s_bit[l]);
            __syncthreads();
            // every row into its class's record (a record is written whole before its slot is published)
            for (uint32_t i = tid; i < n; i += kClassBlock) {
                const Key<W> k = load_key<W>(IX, IZ, cap, i);
                Key<W> ck;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    ck.x[w] = k.x[w] & ~sup.x[w];
                    ck.z[w] = k.z[w] & ~sup.z[w];
                }
                const uint32_t s = pattern_of<W>(k, s_word[l], s_bit[l], s_mask[l]);
                const uint32_t tmask = (1u << tb) - 1u;
                uint32_t h = lds_slot(key_fold<W>(ck), tb);
                uint32_t r = 0xFFFFFFFFu;
                while (r == 0xFFFFFFFFu) {
                    uint32_t e = table[h];
                    if (e == 0u) {  // publish a record of this row's own (record i, unused if another row wins)
                        rec_rep[i] = i;
#pragma unroll
                        for (int q = 0; q < 8; ++q) reinterpret_cast<uint32_t*>(rec_mem + (size_t)i * 16)[q] = 0xFFFFFFFFu;
                        __threadfence_block();
                        e = atomicCAS(&table[h], 0u, i + 1u);
                        if (e == 0u) {
                            r = i;
                            break;
                        }
                    }
                    const Key<W> rk = load_key<W>(IX, IZ, cap, rec_rep[e - 1u]);
                    bool same = true;
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        same = same && (rk.x[w] & ~sup.x[w]) == ck.x[w] && (rk.z[w] & ~sup.z[w]) == ck.z[w];
                    if (same) r = e - 1u;
                    else h = (h + 1u) & tmask;
                }
                rec_mem[(size_t)r * 16 + s] = (uint16_t)i;
                row_rec[i] = (uint16_t)r;
            }
            __syncthreads();
            for (uint32_t i = tid; i < n; i += kClassBlock) {
                const Key<W> k = load_key<W>(IX, IZ, cap, i);
                const T c = IC[i];
                const uint32_t s = pattern_of<W>(k, s_word[l], s_bit[l], s_mask[l]);
                const uint16_t* const mem = rec_mem + (size_t)row_rec[i] * 16;
                // a row alone in its class is every output's only source: one product each, no gather
                uint32_t members = 0;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const uint32_t v = reinterpret_cast<const uint32_t*>(mem)[q];
                    members += ((v & 0xFFFFu) != 0xFFFFu ? 1u : 0u) + ((v >> 16) != 0xFFFFu ? 1u : 0u);
                }
                const bool alone = members == 1u;
                for (uint32_t e = 0; e < ne; ++e) {
                    if (A[e * 16u + s] == 0.0) continue;
                    const uint32_t t = s ^ D[e];
                    Key<W> K = k;
                    xor_local_delta<W>(K, D[e], s_word[l], s_bit[l]);
                    T w = T(0);
                    bool any = false, owner = true;
                    if (alone) {
                        w = c * (T)A[e * 16u + s];
                        any = true;
                    }
                    for (uint32_t f = 0; f < ne && !alone; ++f) {
                        const double af = A[f * 16u + (t ^ D[f])];
                        if (af == 0.0) continue;
                        T cf = c;
                        if (f != e) {
                            const uint32_t j = mem[t ^ D[f]];  // the class's row of the source pattern
                            if (j == 0xFFFFu) continue;
                            if (!any) {  // an earlier entry's source is present: it owns K
                                owner = false;
                                break;
                            }
                            cf = IC[j];
                        }
                        const T v = cf * (T)af;
                        w = any ? w + v : v;
                        any = true;
                    }
                    if (!owner || !survives<W, T>(a.keep, K, w)) continue;
                    const uint32_t j = atomicAdd(&s_count[slot], 1u);
                    if (j < cap) {
                        store_key<W>(OX, OZ, cap, j, K);
                        OC[j] = w;
                    } else {
                        s_bad[slot] = 1u;
                    }
                }
            }
            __syncthreads();
            n = s_count[slot];
            bad = s_bad[slot] != 0u;
            if (bad) break;  // uniform
            in0 = !in0;
            if (tid == 0) atomicAdd(&a.stats[kClassLayerRows + l], (unsigned long long)n);
        }
        if (bad) {
            if (tid == 0) atomicOr(&a.stats[kClassFlags], 2ull);
            __syncthreads();  // the counters are read before the next class bucket resets them
            continue;
        }
        const uint64_t* FX = in0 ? X0 : X1;
        const uint64_t* FZ = in0 ? Z0 : Z1;
        const T* FC = in0 ? C0 : C1;
        if (tid == 0) s_base = atomicAdd(&a.stats[kClassOutRows], (unsigned long long)n);
        __syncthreads();
        const uint64_t o0 = s_base;
        if (o0 + n <= a.ocap) {
            for (uint32_t i = tid; i < n; i += kClassBlock) {
                const Key<W> K = load_key<W>(FX, FZ, cap, i);
                const uint32_t ob = bucket_of<W>(K, a.rows, a.bits);
                store_key<W>(a.ox, a.oz, a.ocap, o0 + i, K);
                a.oc[o0 + i] = FC[i];
                a.obucket[o0 + i] = ob;
                atomicAdd(&a.ocount[ob], 1u);
            }
        } else if (tid == 0) {
            atomicOr(&a.stats[kClassFlags], 4ull);
        }
        __syncthreads();  // the sets are read before the next class bucket loads over them
    }
}

__global__ void __launch_bounds__(256) k_class_lens(const uint32_t* __restrict__ ukeep, const uint32_t* __restrict__ ocount,
                                                    uint32_t* __restrict__ plen, uint32_t n,
                                                    unsigned long long* __restrict__ longest) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v = 0;
    if (i < n) {
        v = ukeep[i] + ocount[i];
        plen[i] = v;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
    if (lane_id() == 0 && v) atomicMax(longest, (unsigned long long)v);
}

// The output layout in one launch: the first `place_blocks` workgroups put each bucket's untouched strings that keep
// passes, in their order, at the front of its segment (a wave per bucket); the others put one class output each
// behind its bucket's untouched strings (each bucket's count counts down to zero).
template <int W, typename T>
__global__ void __launch_bounds__(kClassBlock) k_class_place(const ClassArgs<W, T> a, uint32_t place_blocks,
                                                             uint64_t out_rows) {
    if (blockIdx.x >= place_blocks) {
        const uint64_t i = (uint64_t)(blockIdx.x - place_blocks) * kClassBlock + threadIdx.x;
        if (i >= out_rows) return;
        const uint32_t ob = a.obucket[i];
        const uint64_t pos = a.pstart[ob] + a.ukeep[ob] + (uint64_t)(atomicSub(&a.ocount[ob], 1u) - 1u);
        store_key<W>(a.px, a.pz, a.pcap, pos, load_key<W>(a.ox, a.oz, a.ocap, i));
        a.pc[pos] = a.oc[i];
        return;
    }
    const uint32_t b = blockIdx.x * (kClassBlock / 64) + wave_id();
    if (b >= a.num_buckets) return;  // the whole wave
    const uint32_t lane = lane_id();
    const uint64_t s0 = a.start[b], d0 = a.pstart[b];
    const uint32_t n = a.len[b];
    uint32_t placed = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += 64) {
        const uint32_t r = r0 + lane;
        bool f = false;
        Key<W> k{};
        T c = T(0);
        if (r < n) {
            k = load_key<W>(a.x, a.z, a.cap, s0 + r);
            c = a.c[s0 + r];
            f = !class_touched<W, T>(a, k) && survives<W, T>(a.keep, k, c);
        }
        const uint64_t m = __ballot(f);
        if (f) {
            const uint64_t pos = d0 + placed + (uint32_t)__popcll(m & lanemask_lt());
            store_key<W>(a.px, a.pz, a.pcap, pos, k);
            a.pc[pos] = c;
        }
        placed += (uint32_t)__popcll(m);
    }
}

inline uint32_t wave_grid(uint32_t buckets) { return (buckets + kClassBlock / 64 - 1) / (kClassBlock / 64); }

}  // namespace

template <int W, typename T>
size_t class_lds_bytes(uint32_t lcap, uint32_t tbits, uint32_t prog_words) {
    return (size_t)prog_words * 8 + (size_t)2 * lcap * (2 * W * sizeof(uint64_t) + sizeof(T)) + ((size_t)4 << tbits) +
           (size_t)lcap * (4 + 32 + 2);
}

template <int W, typename T>
hipError_t launch_class_route(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    if (!a.layers || !a.rows || !a.ccount || !a.ukeep || !a.stats) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_class_route<W, T, false>), dim3(wave_grid(a.num_buckets)), dim3(kClassBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_gather(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    if (!a.layers || !a.rows || !a.ccount || !a.cstart || !a.sx || !a.sz || !a.sc) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_class_route<W, T, true>), dim3(wave_grid(a.num_buckets)), dim3(kClassBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_apply(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_classes == 0) return hipSuccess;
    if (!a.layers || !a.rows || !a.cstart || !a.sx || !a.ox || !a.obucket || !a.ocount || !a.stats ||
        a.chunk_rows == 0)
        return hipErrorInvalidValue;
    if (a.tbits < 2 || a.tbits > 14 || ((size_t)1 << a.tbits) < 2 * (size_t)a.lcap || a.lcap > 0xFFFFu)
        return hipErrorInvalidValue;
    if (a.prog_words > kClassMaxLayers * kClassMaxEntries * 16 || a.n_layers > kClassMaxLayers)
        return hipErrorInvalidValue;
    const size_t lds = class_lds_bytes<W, T>(a.lcap, a.tbits, a.prog_words);
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    static int resident[15][5] = {};  // by tbits (which fixes lcap) and the program's size in 4 KiB
    int& per = resident[a.tbits][std::min<uint32_t>(a.prog_words / 512u, 4u)];
    if (per == 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_class_apply<W, T>, kClassBlock, lds) != hipSuccess ||
         per <= 0))
        per = 1;
    const uint32_t grid = std::min<uint32_t>(a.num_classes, (uint32_t)(cus * per));
    hipLaunchKernelGGL((k_class_apply<W, T>), dim3(grid), dim3(kClassBlock), lds, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_lens(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    if (!a.stats || !a.ukeep || !a.ocount || !a.plen) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_class_lens, dim3((a.num_buckets + 255) / 256), dim3(256), 0, st, a.ukeep, a.ocount, a.plen,
                       a.num_buckets, a.stats + kClassLongest);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_place(hipStream_t st, const ClassArgs<W, T>& a, uint64_t out_rows) {
    if (!a.ukeep || !a.ocount || !a.pstart || !a.px || (out_rows && !a.obucket)) return hipErrorInvalidValue;
    const uint32_t place_blocks = wave_grid(a.num_buckets);
    const uint64_t blocks = place_blocks + (out_rows + kClassBlock - 1) / kClassBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_class_place<W, T>), dim3((uint32_t)blocks), dim3(kClassBlock), 0, st, a, place_blocks,
                       out_rows);
    return hipGetLastError();
}

#define PP_CLASS_INSTANTIATE(W, T)                                                              \
    template size_t class_lds_bytes<W, T>(uint32_t, uint32_t, uint32_t);                                  \
    template hipError_t launch_class_route<W, T>(hipStream_t, const ClassArgs<W, T>&);          \
    template hipError_t launch_class_gather<W, T>(hipStream_t, const ClassArgs<W, T>&);         \
    template hipError_t launch_class_apply<W, T>(hipStream_t, const ClassArgs<W, T>&);          \
    template hipError_t launch_class_lens<W, T>(hipStream_t, const ClassArgs<W, T>&);           \
    template hipError_t launch_class_place<W, T>(hipStream_t, const ClassArgs<W, T>&, uint64_t);

PP_CLASS_INSTANTIATE(1, double)
PP_CLASS_INSTANTIATE(1, float)
PP_CLASS_INSTANTIATE(2, double)
PP_CLASS_INSTANTIATE(2, float)

}  // namespace pp
