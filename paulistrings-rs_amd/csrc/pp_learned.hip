// pp_learned.hip -- the learned truncation policy's network: features, a 35-128-128-1 MLP whose two hidden GEMMs
// run on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with fp32 accumulation, its backward pass, Adam, and the per-term score.
//
// A workgroup of four waves takes a tile of 64 terms; wave w owns rows 16w .. 16w + 15 of every activation matrix,
// so a layer's output tile feeds the next layer within the wave.  MFMA operand layout (16x16x32, bf16): lane l
// supplies row (l & 15) of A and column (l & 15) of B, each at k = 8 (l >> 4) .. 8 (l >> 4) + 7; it receives
// D[4 (l >> 4) + q][l & 15] for q = 0..3.  Both operands are read with k contiguous: A row-major, B as its transpose.
#include "pp_learned.h"

namespace pp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// acc += A[16 x 32] * B[32 x 16]: A row-major at a (lda elements per row), B given transposed at bt (ldb per row).
__device__ __forceinline__ f32x4 mfma_16x16x32(const uint16_t* a, int lda, const uint16_t* bt, int ldb, f32x4 acc) {
    const int l = threadIdx.x & 63;
    const int r = l & 15, kq = (l >> 4) * 8;
    const uint4 av = *reinterpret_cast<const uint4*>(a + r * lda + kq);
    const uint4 bv = *reinterpret_cast<const uint4*>(bt + r * ldb + kq);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                    acc, 0, 0, 0);
}

// The 64 bf16 features of one term into row[0 .. 64): four threads per term, `part` = which 16.
__device__ __forceinline__ void term_features(const MlpDev& m, uint64_t x, uint64_t z, double c, float remaining,
                                              bool valid, int part, uint16_t* row) {
    float f[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) f[i] = 0.f;
    if (valid) {
        if (part == 0) {
#pragma unroll
            for (int g = 0; g < 16; ++g) f[g] = (float)__popcll(x & m.groups[g]) / 8.f;
        } else if (part == 1) {
#pragma unroll
            for (int g = 0; g < 16; ++g) f[g] = (float)__popcll(z & m.groups[g]) / 8.f;
        } else if (part == 2) {
            f[0] = (float)__popcll(x | z) / 16.f;
            f[1] = (float)(log(fmax(fabs(c), 1e-300)) / 8.0);
            f[2] = fminf(remaining / m.depth_scale, 1.f);
        }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) row[16 * part + i] = bf16_bits(f[i]);
}

// Shared memory of a forward tile.
struct FwdLds {
    uint16_t a0[kMlpTile * kMlpInPad];  // features, bf16
    uint16_t a1[kMlpTile * kMlpHid];    // bf16 relu(z1)
    float y[kMlpTile];
};

// Layers 1 and 2 and the output for the tile in s.a0; h2 (fp32 relu(z2)) is written when h2 is not null.
__device__ __forceinline__ void mlp_forward_tile(const MlpDev& m, FwdLds& s, float* h2) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const float* b1 = m.params + kOffB1;
    const float* b2 = m.params + kOffB2;
    const float* w3 = m.params + kOffW3;
    // layer 1: rows 16w .. 16w + 15, all 128 outputs, K = 64
    for (int ct = 0; ct < kMlpHid / 16; ++ct) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kMlpInPad / 32; ++ks)
            acc = mfma_16x16x32(s.a0 + (16 * w) * kMlpInPad + 32 * ks, kMlpInPad, m.w1t + (16 * ct) * kMlpInPad + 32 * ks,
                                kMlpInPad, acc);
        const int col = 16 * ct + (l & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * w + 4 * (l >> 4) + q;
            s.a1[row * kMlpHid + col] = bf16_bits(fmaxf(acc[q] + b1[col], 0.f));
        }
    }
    __syncthreads();
    // layer 2 and the output: K = 128
    float part[4] = {0.f, 0.f, 0.f, 0.f};
    for (int ct = 0; ct < kMlpHid / 16; ++ct) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kMlpHid / 32; ++ks)
            acc = mfma_16x16x32(s.a1 + (16 * w) * kMlpHid + 32 * ks, kMlpHid, m.w2t + (16 * ct) * kMlpHid + 32 * ks,
                                kMlpHid, acc);
        const int col = 16 * ct + (l & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float h = fmaxf(acc[q] + b2[col], 0.f);
            if (h2) h2[(16 * w + 4 * (l >> 4) + q) * kMlpHid + col] = h;
            part[q] += h * w3[col];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float v = part[q];
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) v += __shfl_xor(v, d, 64);
        if ((l & 15) == 0) s.y[16 * w + 4 * (l >> 4) + q] = v + m.params[kOffB3];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_mlp_refresh(const float* __restrict__ p, uint16_t* __restrict__ w1t,
                                                     uint16_t* __restrict__ w2t, uint16_t* __restrict__ w2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kMlpInPad * kMlpHid) {  // W1[k][n] -> w1t[n][k]
        const int k = i / kMlpHid, n = i % kMlpHid;
        w1t[n * kMlpInPad + k] = bf16_bits(p[kOffW1 + i]);
    }
    if (i < kMlpHid * kMlpHid) {  // W2[k][n] -> w2t[n][k] and w2[k][n]
        const int k = i / kMlpHid, n = i % kMlpHid;
        const uint16_t b = bf16_bits(p[kOffW2 + i]);
        w2t[n * kMlpHid + k] = b;
        w2[i] = b;
    }
}

__global__ void __launch_bounds__(256) k_mlp_predict(const MlpDev m, const MlpBatch b, float* __restrict__ out) {
    __shared__ FwdLds s;
    const uint64_t base = (uint64_t)blockIdx.x * kMlpTile;
    const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
    const uint64_t i = base + row;
    const bool valid = i < b.count;
    term_features(m, valid ? b.x[i] : 0, valid ? b.z[i] : 0, valid ? b.c[i] : 1.0, valid ? b.remaining[i] : 0.f,
                  valid, part, s.a0 + row * kMlpInPad);
    __syncthreads();
    mlp_forward_tile(m, s, nullptr);
    if (threadIdx.x < kMlpTile && base + threadIdx.x < b.count) out[base + threadIdx.x] = s.y[threadIdx.x];
}

struct GradLds {
    FwdLds f;
    uint16_t a0t[kMlpInPad * kMlpTile];  // [64 inputs][64 rows]
    uint16_t a1t[kMlpHid * kMlpTile];    // [128 hidden][64 rows]
    uint16_t dz2[kMlpTile * kMlpHid];    // bf16 dz2 [rows][hidden]
    uint16_t dzt[kMlpHid * kMlpTile];    // bf16 dz2, later dz1, transposed [hidden][rows]
    float h[kMlpTile * kMlpHid];         // fp32 h2, then dz2, then dz1
    float dy[kMlpTile];
};

// Column sums over the tile's 64 rows of an fp32 [64][128] matrix, added to g[0 .. 128).
__device__ __forceinline__ void add_column_sums(const float* h, float* g) {
    if (threadIdx.x < kMlpHid) {
        float s = 0.f;
        for (int r = 0; r < kMlpTile; ++r) s += h[r * kMlpHid + threadIdx.x];
        atomicAdd(&g[threadIdx.x], s);
    }
}

__global__ void __launch_bounds__(256) k_mlp_grad(const MlpDev m, const MlpBatch b, float* __restrict__ grads,
                                                  double* __restrict__ loss_sum) {
    __shared__ GradLds s;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kMlpTile;
    {
        const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
        const uint64_t i = base + row;
        const bool valid = i < b.count;
        term_features(m, valid ? b.x[i] : 0, valid ? b.z[i] : 0, valid ? b.c[i] : 1.0, valid ? b.remaining[i] : 0.f,
                      valid, part, s.f.a0 + row * kMlpInPad);
    }
    __syncthreads();
    mlp_forward_tile(m, s.f, s.h);
    // dy and the loss
    if (threadIdx.x < kMlpTile) {
        const uint64_t i = base + threadIdx.x;
        float d = 0.f;
        if (i < b.count) {
            const float e = s.f.y[threadIdx.x] - b.label[i];
            d = 2.f * e / (float)b.count;
            atomicAdd(loss_sum, (double)e * (double)e);
        }
        s.dy[threadIdx.x] = d;
    }
    // transposes of a0 and a1 for the weight-gradient GEMMs
    for (int i = threadIdx.x; i < kMlpTile * kMlpInPad; i += 256) {
        const int r = i / kMlpInPad, k = i % kMlpInPad;
        s.a0t[k * kMlpTile + r] = s.f.a0[i];
    }
    for (int i = threadIdx.x; i < kMlpTile * kMlpHid; i += 256) {
        const int r = i / kMlpHid, k = i % kMlpHid;
        s.a1t[k * kMlpTile + r] = s.f.a1[i];
    }
    __syncthreads();
    // w3, b3 gradients; dz2 = dy w3 (h2 > 0)
    const float* w3 = m.params + kOffW3;
    if (threadIdx.x < kMlpHid) {
        float g = 0.f;
        for (int r = 0; r < kMlpTile; ++r) g += s.h[r * kMlpHid + threadIdx.x] * s.dy[r];
        atomicAdd(&grads[kOffW3 + threadIdx.x], g);
    } else if (threadIdx.x == kMlpHid) {
        float g = 0.f;
        for (int r = 0; r < kMlpTile; ++r) g += s.dy[r];
        atomicAdd(&grads[kOffB3], g);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMlpTile * kMlpHid; i += 256) {
        const int r = i / kMlpHid, j = i % kMlpHid;
        const float d = s.h[i] > 0.f ? s.dy[r] * w3[j] : 0.f;
        s.h[i] = d;
        const uint16_t bd = bf16_bits(d);
        s.dz2[i] = bd;
        s.dzt[j * kMlpTile + r] = bd;
    }
    __syncthreads();
    add_column_sums(s.h, grads + kOffB2);
    // dW2[i][o] = sum_m a1[m][i] dz2[m][o]: D[o][i] = dz2t[o][.] . a1t[i][.], 64 tiles of 16x16, K = 64 rows
    for (int t = w; t < 64; t += 4) {
        const int ot = t >> 3, it = t & 7;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kMlpTile / 32; ++ks)
            acc = mfma_16x16x32(s.dzt + (16 * ot) * kMlpTile + 32 * ks, kMlpTile, s.a1t + (16 * it) * kMlpTile + 32 * ks,
                                kMlpTile, acc);
        const int i = 16 * it + (l & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int o = 16 * ot + 4 * (l >> 4) + q;
            atomicAdd(&grads[kOffW2 + i * kMlpHid + o], acc[q]);
        }
    }
    __syncthreads();
    // dh1[m][i] = sum_o dz2[m][o] W2[i][o]; dz1 = dh1 (a1 > 0)
    for (int it = 0; it < kMlpHid / 16; ++it) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kMlpHid / 32; ++ks)
            acc = mfma_16x16x32(s.dz2 + (16 * w) * kMlpHid + 32 * ks, kMlpHid, m.w2 + (16 * it) * kMlpHid + 32 * ks,
                                kMlpHid, acc);
        const int i = 16 * it + (l & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * w + 4 * (l >> 4) + q;
            const float d = bf16_float(s.f.a1[r * kMlpHid + i]) > 0.f ? acc[q] : 0.f;
            s.h[r * kMlpHid + i] = d;
            s.dzt[i * kMlpTile + r] = bf16_bits(d);
        }
    }
    __syncthreads();
    add_column_sums(s.h, grads + kOffB1);
    // dW1[k][o] = sum_m a0[m][k] dz1[m][o]: D[o][k] = dz1t[o][.] . a0t[k][.], 8 x 4 tiles (inputs >= 35 skipped)
    for (int t = w; t < 32; t += 4) {
        const int ot = t >> 2, kt = t & 3;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kMlpTile / 32; ++ks)
            acc = mfma_16x16x32(s.dzt + (16 * ot) * kMlpTile + 32 * ks, kMlpTile, s.a0t + (16 * kt) * kMlpTile + 32 * ks,
                                kMlpTile, acc);
        const int k = 16 * kt + (l & 15);
        if (k < kMlpIn) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int o = 16 * ot + 4 * (l >> 4) + q;
                atomicAdd(&grads[kOffW1 + k * kMlpHid + o], acc[q]);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_mlp_adam(float* __restrict__ p, float* __restrict__ am,
                                                  float* __restrict__ av, const float* __restrict__ g, float lr,
                                                  float beta1, float beta2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kMlpParams) return;
    if (i >= kOffW1 + kMlpIn * kMlpHid && i < kOffB1) return;  // W1's padding rows stay zero
    const float gi = g[i];
    const float mi = beta1 * am[i] + (1.f - beta1) * gi;
    const float vi = beta2 * av[i] + (1.f - beta2) * gi * gi;
    am[i] = mi;
    av[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mlp_score(const MlpDev m, const ScoreArgs<T> a) {
    __shared__ FwdLds s;
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.len[b];
    const uint64_t start = a.start[b];
    const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
    for (uint32_t t0 = 0; t0 < n; t0 += kMlpTile) {
        const uint32_t r = t0 + row;
        const bool valid = r < n;
        const uint64_t i = start + r;
        term_features(m, valid ? a.x[i] : 0, valid ? a.z[i] : 0, valid ? (double)a.c[i] : 1.0, a.remaining, valid,
                      part, s.a0 + row * kMlpInPad);
        __syncthreads();
        mlp_forward_tile(m, s, nullptr);
        // score = ln|c| + the predicted ln|read-out factor| = the predicted ln|contribution|
        if (threadIdx.x < kMlpTile && t0 + threadIdx.x < n) {
            const uint64_t j = start + t0 + threadIdx.x;
            a.score[j] = s.y[threadIdx.x] + (float)log(fabs((double)a.c[j]));
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_mlp_refresh(hipStream_t st, const float* params, uint16_t* w1t, uint16_t* w2t, uint16_t* w2) {
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpHid * kMlpHid + 255) / 256), dim3(256), 0, st, params, w1t, w2t, w2);
    return hipGetLastError();
}

hipError_t launch_mlp_predict(hipStream_t st, const MlpDev& m, const MlpBatch& b, float* out) {
    if (b.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_predict, dim3((uint32_t)((b.count + kMlpTile - 1) / kMlpTile)), dim3(256), 0, st, m, b,
                       out);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(hipStream_t st, const MlpDev& m, const MlpBatch& b, float* grads, double* loss_sum) {
    if (b.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_grad, dim3((uint32_t)((b.count + kMlpTile - 1) / kMlpTile)), dim3(256), 0, st, m, b,
                       grads, loss_sum);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(hipStream_t st, float* params, float* adam_m, float* adam_v, const float* grads, float lr,
                           float beta1, float beta2, float eps, int step) {
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, params, adam_m, adam_v, grads,
                       lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mlp_score(hipStream_t st, const MlpDev& m, const ScoreArgs<T>& a) {
    if (a.num_buckets == 0) return hipSuccess;
    hipLaunchKernelGGL((k_mlp_score<T>), dim3(a.num_buckets), dim3(256), 0, st, m, a);
    return hipGetLastError();
}

template hipError_t launch_mlp_score<double>(hipStream_t, const MlpDev&, const ScoreArgs<double>&);
template hipError_t launch_mlp_score<float>(hipStream_t, const MlpDev&, const ScoreArgs<float>&);

}  // namespace pp
