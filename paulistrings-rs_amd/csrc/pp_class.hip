// pp_class.hip -- the class pass (pp_class.h): a program of up to eight local tables applied in one pass, each class
// bucket's strings held in LDS through every layer.
//
// Layout of one apply workgroup's LDS: two row sets of lcap rows (word-major keys, then coefficients) that trade
// places every layer, and an open-addressing table over the live set (slot = row + 1, 0 = empty).  A layer gathers
// every output from its sources: an output K with pattern t has, for each entry e, the source K ^ lift(delta[e])
// with pattern t ^ delta[e] and amplitude amp[e][t ^ delta[e]]; the source that comes first in entry order among
// those present owns K, sums its contributions in entry order (one rounded multiply and one rounded add each, as a
// rotation's pair update or the host engine's merge), asks keep and appends it to the other set.  A rotation table
// (identity entry first, then the generator) gives own = c cos + c_partner a1 exactly as the pair join does.

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

template <int W>
__device__ __forceinline__ void lds_insert(uint32_t* table, uint32_t tbits, const Key<W>& k, uint32_t i) {
    const uint32_t tmask = (1u << tbits) - 1u;
    uint32_t s = lds_slot(key_fold<W>(k), tbits);
    while (atomicCAS(&table[s], 0u, i + 1u) != 0u) s = (s + 1u) & tmask;
}

// The row of the live set holding q, or -1.  The table is at most half full, so a probe ends.
template <int W>
__device__ __forceinline__ int lds_find(const uint64_t* X, const uint64_t* Z, uint32_t cap, const uint32_t* table,
                                        uint32_t tbits, const Key<W>& q) {
    const uint32_t tmask = (1u << tbits) - 1u;
    uint32_t s = lds_slot(key_fold<W>(q), tbits);
    for (;;) {
        const uint32_t e = table[s];
        if (e == 0u) return -1;
        const uint32_t i = e - 1u;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < W; ++w) eq = eq && X[w * cap + i] == q.x[w] && Z[w * cap + i] == q.z[w];
        if (eq) return (int)i;
        s = (s + 1u) & tmask;
    }
}

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
// Per layer three barriers: the table cleared and the layer's entries copied to LDS; the live set inserted; the
// outputs written.  The output counters alternate between two slots, so the next layer resets its slot while the
// last readers of this one are still behind the third barrier.
template <int W, typename T>
__global__ void __launch_bounds__(kClassBlock) k_class_apply(const ClassArgs<W, T> a) {
    extern __shared__ __align__(16) unsigned char class_lds[];
    const uint32_t cap = a.lcap;
    uint64_t* const X0 = reinterpret_cast<uint64_t*>(class_lds);
    uint64_t* const Z0 = X0 + (size_t)W * cap;
    uint64_t* const X1 = Z0 + (size_t)W * cap;
    uint64_t* const Z1 = X1 + (size_t)W * cap;
    T* const C0 = reinterpret_cast<T*>(Z1 + (size_t)W * cap);
    T* const C1 = C0 + cap;
    uint32_t* const table = reinterpret_cast<uint32_t*>(C1 + cap);
    __shared__ double s_amp[kClassMaxEntries][16];
    __shared__ uint32_t s_delta[kClassMaxEntries];
    __shared__ uint32_t s_count[2];
    __shared__ uint32_t s_bad[2];
    __shared__ unsigned long long s_base;
    const uint32_t tid = threadIdx.x;

    for (uint32_t cb = blockIdx.x; cb < a.num_classes; cb += gridDim.x) {
        const uint64_t base = a.cstart[cb];
        const uint64_t end = cb + 1u < a.num_classes ? a.cstart[cb + 1u] : (uint64_t)a.stats[kClassStaged];
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
        for (uint32_t l = 0; l < a.n_layers; ++l) {
            const ClassLayer& L = a.layers[l];
            const uint32_t slot = l & 1u;
            const uint64_t* IX = in0 ? X0 : X1;
            const uint64_t* IZ = in0 ? Z0 : Z1;
            const T* IC = in0 ? C0 : C1;
            uint64_t* OX = in0 ? X1 : X0;
            uint64_t* OZ = in0 ? Z1 : Z0;
            T* OC = in0 ? C1 : C0;
            // a table of at least twice the live rows (a probe then ends), at most the launch's
            uint32_t tb = 4;
            while (tb < a.tbits && (1u << tb) < 2u * n) ++tb;
            for (uint32_t i = tid; i < (1u << tb); i += kClassBlock) table[i] = 0u;
            const uint32_t ne = L.n_entries;
            for (uint32_t i = tid; i < ne * 16u; i += kClassBlock) s_amp[i >> 4][i & 15u] = L.amp[i >> 4][i & 15u];
            if (tid < ne) s_delta[tid] = L.delta[tid];
            if (tid == 0) {
                s_count[slot] = 0u;
                s_bad[slot] = 0u;
            }
            __syncthreads();
            for (uint32_t i = tid; i < n; i += kClassBlock) lds_insert<W>(table, tb, load_key<W>(IX, IZ, cap, i), i);
            __syncthreads();
            for (uint32_t i = tid; i < n; i += kClassBlock) {
                const Key<W> k = load_key<W>(IX, IZ, cap, i);
                const T c = IC[i];
                const uint32_t s = pattern_of<W>(k, L.word, L.bit, L.mask);
                for (uint32_t e = 0; e < ne; ++e) {
                    if (s_amp[e][s] == 0.0) continue;
                    const uint32_t t = s ^ s_delta[e];
                    Key<W> K = k;
                    xor_local_delta<W>(K, s_delta[e], L.word, L.bit);
                    T w = T(0);
                    bool any = false, owner = true;
                    for (uint32_t f = 0; f < ne; ++f) {
                        const double af = s_amp[f][t ^ s_delta[f]];
                        if (af == 0.0) continue;
                        T cf = c;
                        if (f != e) {
                            Key<W> src = K;
                            xor_local_delta<W>(src, s_delta[f], L.word, L.bit);
                            const int j = lds_find<W>(IX, IZ, cap, table, tb, src);
                            if (j < 0) continue;
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
size_t class_lds_bytes(uint32_t lcap, uint32_t tbits) {
    return (size_t)2 * lcap * (2 * W * sizeof(uint64_t) + sizeof(T)) + ((size_t)4 << tbits);
}

template <int W, typename T>
hipError_t launch_class_route(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_class_route<W, T, false>), dim3(wave_grid(a.num_buckets)), dim3(kClassBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_gather(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_class_route<W, T, true>), dim3(wave_grid(a.num_buckets)), dim3(kClassBlock), 0, st, a);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_apply(hipStream_t st, const ClassArgs<W, T>& a) {
    if (a.num_classes == 0) return hipSuccess;
    if (a.tbits < 2 || a.tbits > 14 || ((size_t)1 << a.tbits) < 2 * (size_t)a.lcap) return hipErrorInvalidValue;
    const size_t lds = class_lds_bytes<W, T>(a.lcap, a.tbits);
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    static int resident[15] = {};  // by tbits (which fixes lcap)
    int& per = resident[a.tbits];
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
    hipLaunchKernelGGL(k_class_lens, dim3((a.num_buckets + 255) / 256), dim3(256), 0, st, a.ukeep, a.ocount, a.plen,
                       a.num_buckets, a.stats + kClassLongest);
    return hipGetLastError();
}

template <int W, typename T>
hipError_t launch_class_place(hipStream_t st, const ClassArgs<W, T>& a, uint64_t out_rows) {
    const uint32_t place_blocks = wave_grid(a.num_buckets);
    const uint64_t blocks = place_blocks + (out_rows + kClassBlock - 1) / kClassBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_class_place<W, T>), dim3((uint32_t)blocks), dim3(kClassBlock), 0, st, a, place_blocks,
                       out_rows);
    return hipGetLastError();
}

#define PP_CLASS_INSTANTIATE(W, T)                                                              \
    template size_t class_lds_bytes<W, T>(uint32_t, uint32_t);                                  \
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
