// pp_learned.hip -- the learned truncation policy's network: features, a 35-128-128-1 MLP whose two hidden GEMMs
// run on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with fp32 accumulation, its backward pass, Adam, and the per-term score.
//
// A workgroup of four waves takes a tile of 64 terms.  In the forward pass wave w owns hidden columns 32w .. 32w + 31
// of both layers, with those columns' weight fragments in registers for the whole kernel, and reads the activation
// rows from LDS; the backward's GEMMs read both operands from LDS or global memory.  MFMA operand layout (16x16x32, bf16): lane l
// supplies row (l & 15) of A and column (l & 15) of B, each at k = 8 (l >> 4) .. 8 (l >> 4) + 7; it receives
// D[4 (l >> 4) + q][l & 15] for q = 0..3.  Both operands are read with k contiguous: A row-major, B as its transpose.
#include "pp_learned.h"

#include <algorithm>

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

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Set bits of a 64-bit word as a 32-bit count (two v_bcnt; the 64-bit popcount's result would be converted to
// float as a 64-bit integer).
__device__ __forceinline__ uint32_t popc64(uint64_t v) { return __popc((uint32_t)v) + __popc((uint32_t)(v >> 32)); }

// max(v, 0) as one integer max on the bits: a float with the sign bit set is a negative int.  Equal to fmaxf(v, 0)
// except that -0 and negative NaNs give +0 (a NaN with the sign clear stays NaN).
__device__ __forceinline__ float relu(float v) { return __int_as_float(max(__float_as_int(v), 0)); }

// Four floats to bf16 (round to nearest even, v_cvt_pk_bf16_f32), packed in 8 bytes.
__device__ __forceinline__ uint2 pack_bf16x4(float a, float b, float c, float d) {
    bf16x4 v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    v[2] = (__bf16)c;
    v[3] = (__bf16)d;
    return __builtin_bit_cast(uint2, v);
}

// The 64 bf16 features of one term into row[0 .. 64), one wave per part: the term is lane (threadIdx & 63) and
// part = threadIdx >> 6 picks which 16 features the thread writes -- 0: X occupancies, 1: Z occupancies, 2: the
// scalar features, 3: padding -- so every branch below is uniform across a wave.  Part 2 also leaves
// ln max(|c|, 1e-300) in *lnc when lnc is not null.
__device__ __forceinline__ void term_features(const MlpDev& m, uint64_t x, uint64_t z, double c, float remaining,
                                              bool valid, int part, uint16_t* row, float* lnc) {
    float f[16];
    if (part < 2) {
        const uint64_t v = part == 0 ? x : z;
#pragma unroll
        for (int g = 0; g < 16; ++g) f[g] = valid ? (float)popc64(v & m.groups[g]) * 0.125f : 0.f;
    } else {
#pragma unroll
        for (int g = 0; g < 16; ++g) f[g] = 0.f;
        if (part == 2) {
            const float L = (float)log(fmax(fabs(c), 1e-300));
            if (lnc) *lnc = L;
            if (valid) {
                f[0] = (float)popc64(x | z) / 16.f;
                f[1] = L / 8.f;  // exact: a power-of-two scale of a value far from the subnormal range
                f[2] = fminf(remaining / m.depth_scale, 1.f);
            }
        }
    }
    uint4 lo, hi;
    const uint2 p0 = pack_bf16x4(f[0], f[1], f[2], f[3]), p1 = pack_bf16x4(f[4], f[5], f[6], f[7]);
    const uint2 p2 = pack_bf16x4(f[8], f[9], f[10], f[11]), p3 = pack_bf16x4(f[12], f[13], f[14], f[15]);
    lo = make_uint4(p0.x, p0.y, p1.x, p1.y);
    hi = make_uint4(p2.x, p2.y, p3.x, p3.y);
    reinterpret_cast<uint4*>(row + 16 * part)[0] = lo;
    reinterpret_cast<uint4*>(row + 16 * part)[1] = hi;
}

// Row strides (bf16 elements) of the activation tiles in LDS, padded by 16 bytes so that the 16 rows one MFMA operand
// read touches start in different banks.
constexpr int kLd0 = kMlpInPad + 8, kLd1 = kMlpHid + 8;

// Shared memory of a forward tile.
struct alignas(16) FwdLds {
    uint16_t a0[kMlpTile * kLd0];  // features, bf16 [row][input]
    uint16_t a1[kMlpTile * kLd1];  // bf16 relu(z1) [row][hidden]
    float ypart[2][kWaves][kMlpTile];  // each wave's share of the output dot product, by tile parity
    float y[kMlpTile];
    float b1[kMlpHid], b2[kMlpHid], w3[kMlpHid];  // read as float4 per use: registers hold only the weights
};

// The weights a wave multiplies by, held in registers for the whole kernel.  The GEMMs are computed transposed,
// Z^T = W^T A^T, with the weights as the MFMA's A operand and the terms as its B operand: wave w owns hidden units
// 32w .. 32w + 31 (two 16-row tiles j), lane l supplies weight row 32w + 16j + (l & 15) and receives, for term
// (l & 15) of the term tile, the four consecutive hidden units 32w + 16j + 4 (l >> 4) + q, which are contiguous in
// the next layer's [term][hidden] operand.
struct FwdWeights {
    bf16x8 w1[2][kMlpInPad / 32];  // A operand fragments of W1^T, per hidden tile and k step
    bf16x8 w2[2][kMlpHid / 32];
    float b3;
};

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// The weight fragments into registers, and b1, b2, w3 into s; the caller's next barrier publishes s.
__device__ __forceinline__ FwdWeights load_fwd_weights(const MlpDev& m, FwdLds& s) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = l & 15, lk = (l >> 4) * 8;
    FwdWeights f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int row = 32 * w + 16 * j + lr;
#pragma unroll
        for (int ks = 0; ks < kMlpInPad / 32; ++ks) f.w1[j][ks] = load_frag(m.w1t + row * kMlpInPad + 32 * ks + lk);
#pragma unroll
        for (int ks = 0; ks < kMlpHid / 32; ++ks) f.w2[j][ks] = load_frag(m.w2t + row * kMlpHid + 32 * ks + lk);
    }
    if (threadIdx.x < kMlpHid) {
        s.b1[threadIdx.x] = m.params[kOffB1 + threadIdx.x];
        s.b2[threadIdx.x] = m.params[kOffB2 + threadIdx.x];
        s.w3[threadIdx.x] = m.params[kOffW3 + threadIdx.x];
    }
    f.b3 = m.params[kOffB3];
    return f;
}

// The four consecutive hidden units lane l receives in hidden tile j, from a [128] array in LDS.
__device__ __forceinline__ float4 lane_units(const float* v, int j) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    return *reinterpret_cast<const float4*>(v + 32 * w + 16 * j + 4 * (l >> 4));
}

// Layer 1 of the tile in s.a0 into s.a1 (K = 64).  The eight (term tile, hidden tile) accumulators are independent
// MFMA chains, stepped through K together.
__device__ __forceinline__ void mlp_layer1(const FwdWeights& fw, FwdLds& s) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = l & 15, lk = (l >> 4) * 8, lq = 4 * (l >> 4);
    f32x4 acc[kMlpTile / 16][2];
#pragma unroll
    for (int i = 0; i < kMlpTile / 16; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float4 b = lane_units(s.b1, j);
            acc[i][j] = f32x4{b.x, b.y, b.z, b.w};
        }
#pragma unroll
    for (int ks = 0; ks < kMlpInPad / 32; ++ks) {
#pragma unroll
        for (int i = 0; i < kMlpTile / 16; ++i) {
            const bf16x8 a = load_frag(s.a0 + (16 * i + lr) * kLd0 + 32 * ks + lk);
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw.w1[j][ks], a, acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < kMlpTile / 16; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            *reinterpret_cast<uint2*>(s.a1 + (16 * i + lr) * kLd1 + 32 * w + 16 * j + lq) =
                pack_bf16x4(relu(acc[i][j][0]), relu(acc[i][j][1]), relu(acc[i][j][2]), relu(acc[i][j][3]));
}

// Layer 2 of the tile in s.a1 (K = 128) and this wave's share of the output into ypart[w]; h2 (fp32 relu(z2),
// [row][hidden]) is written when not null.
__device__ __forceinline__ void mlp_layer2(const FwdWeights& fw, const FwdLds& s, float* h2,
                                           float (*ypart)[kMlpTile]) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = l & 15, lk = (l >> 4) * 8, lq = 4 * (l >> 4);
    f32x4 acc[kMlpTile / 16][2];
#pragma unroll
    for (int i = 0; i < kMlpTile / 16; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float4 b = lane_units(s.b2, j);
            acc[i][j] = f32x4{b.x, b.y, b.z, b.w};
        }
#pragma unroll
    for (int ks = 0; ks < kMlpHid / 32; ++ks) {
#pragma unroll
        for (int i = 0; i < kMlpTile / 16; ++i) {
            const bf16x8 a = load_frag(s.a1 + (16 * i + lr) * kLd1 + 32 * ks + lk);
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw.w2[j][ks], a, acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < kMlpTile / 16; ++i) {
        float part = 0.f;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float4 w3v = lane_units(s.w3, j);
            const float w3q[4] = {w3v.x, w3v.y, w3v.z, w3v.w};
            float h[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                h[q] = relu(acc[i][j][q]);
                part = __builtin_fmaf(h[q], w3q[q], part);
            }
            if (h2)
                *reinterpret_cast<float4*>(h2 + (16 * i + lr) * kMlpHid + 32 * w + 16 * j + lq) =
                    make_float4(h[0], h[1], h[2], h[3]);
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        ypart[w][16 * i + lr] = part;  // the four lanes of a term hold the same sum
    }
}

// The output of term t: the waves' shares in order, then b3.
__device__ __forceinline__ float mlp_output(const FwdWeights& fw, const float (*ypart)[kMlpTile], int t) {
    float v = ypart[0][t];
#pragma unroll
    for (int ww = 1; ww < kWaves; ++ww) v += ypart[ww][t];
    return v + fw.b3;
}

// Layers 1 and 2 and the output for the tile in s.a0, into s.y; h2 as in mlp_layer2.
__device__ __forceinline__ void mlp_forward_tile(const FwdWeights& fw, FwdLds& s, float* h2) {
    mlp_layer1(fw, s);
    __syncthreads();
    mlp_layer2(fw, s, h2, s.ypart[0]);
    __syncthreads();
    if (threadIdx.x < kMlpTile) s.y[threadIdx.x] = mlp_output(fw, s.ypart[0], threadIdx.x);
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
    const FwdWeights fw = load_fwd_weights(m, s);
    const uint64_t base = (uint64_t)blockIdx.x * kMlpTile;
    const int row = threadIdx.x & 63, part = threadIdx.x >> 6;
    const uint64_t i = base + row;
    const bool valid = i < b.count;
    term_features(m, valid ? b.x[i] : 0, valid ? b.z[i] : 0, valid ? b.c[i] : 1.0, valid ? b.remaining[i] : 0.f,
                  valid, part, s.a0 + row * kLd0, nullptr);
    __syncthreads();
    mlp_forward_tile(fw, s, nullptr);
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
    double e2[kMlpTile];                 // each row's squared error
};

// Column sums over the tile's 64 rows of an fp32 [64][128] matrix, stored to g[0 .. 128).
__device__ __forceinline__ void store_column_sums(const float* h, float* g) {
    if (threadIdx.x < kMlpHid) {
        float s = 0.f;
        for (int r = 0; r < kMlpTile; ++r) s += h[r * kMlpHid + threadIdx.x];
        g[threadIdx.x] = s;
    }
}

// Workgroup t writes the gradient of its 64 rows to partial[t * kMlpParams ...] (every parameter but W1's padding
// rows, each by one thread) and their squared error to loss_part[t]; k_mlp_reduce sums the tiles in tile order, so
// that a gradient, and the training built on it, is the same on every run.
__global__ void __launch_bounds__(256) k_mlp_grad(const MlpDev m, const MlpBatch b, float* __restrict__ partial,
                                                  double* __restrict__ loss_part) {
    __shared__ GradLds s;
    float* const grads = partial + (uint64_t)blockIdx.x * kMlpParams;
    const FwdWeights fw = load_fwd_weights(m, s.f);
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kMlpTile;
    {
        const int row = threadIdx.x & 63, part = threadIdx.x >> 6;
        const uint64_t i = base + row;
        const bool valid = i < b.count;
        term_features(m, valid ? b.x[i] : 0, valid ? b.z[i] : 0, valid ? b.c[i] : 1.0, valid ? b.remaining[i] : 0.f,
                      valid, part, s.f.a0 + row * kLd0, nullptr);
    }
    __syncthreads();
    mlp_forward_tile(fw, s.f, s.h);
    // dy and the loss
    if (threadIdx.x < kMlpTile) {
        const uint64_t i = base + threadIdx.x;
        float d = 0.f;
        double e2 = 0.0;
        if (i < b.count) {
            const float e = s.f.y[threadIdx.x] - b.label[i];
            d = 2.f * e / (float)b.count;
            e2 = (double)e * (double)e;
        }
        s.dy[threadIdx.x] = d;
        s.e2[threadIdx.x] = e2;
    }
    // transposes of a0 and a1 for the weight-gradient GEMMs
    for (int i = threadIdx.x; i < kMlpTile * kMlpInPad; i += 256) {
        const int r = i / kMlpInPad, k = i % kMlpInPad;
        s.a0t[k * kMlpTile + r] = s.f.a0[r * kLd0 + k];
    }
    for (int i = threadIdx.x; i < kMlpTile * kMlpHid; i += 256) {
        const int r = i / kMlpHid, k = i % kMlpHid;
        s.a1t[k * kMlpTile + r] = s.f.a1[r * kLd1 + k];
    }
    __syncthreads();
    // w3, b3 gradients; dz2 = dy w3 (h2 > 0)
    const float* w3 = m.params + kOffW3;
    if (threadIdx.x < kMlpHid) {
        float g = 0.f;
        for (int r = 0; r < kMlpTile; ++r) g += s.h[r * kMlpHid + threadIdx.x] * s.dy[r];
        grads[kOffW3 + threadIdx.x] = g;
    } else if (threadIdx.x == kMlpHid) {
        float g = 0.f;
        for (int r = 0; r < kMlpTile; ++r) g += s.dy[r];
        grads[kOffB3] = g;
    } else if (threadIdx.x == kMlpHid + 1) {
        double l = 0.0;
        for (int r = 0; r < kMlpTile; ++r) l += s.e2[r];
        loss_part[blockIdx.x] = l;
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
    store_column_sums(s.h, grads + kOffB2);
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
            grads[kOffW2 + i * kMlpHid + o] = acc[q];
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
            const float d = bf16_float(s.f.a1[r * kLd1 + i]) > 0.f ? acc[q] : 0.f;
            s.h[r * kMlpHid + i] = d;
            s.dzt[i * kMlpTile + r] = bf16_bits(d);
        }
    }
    __syncthreads();
    store_column_sums(s.h, grads + kOffB1);
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
                grads[kOffW1 + k * kMlpHid + o] = acc[q];
            }
        }
    }
}

// grads[i] = the tiles' partial gradients summed in tile order (W1's padding rows zero); loss[0] likewise.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ partial, uint32_t tiles,
                                                    float* __restrict__ grads, const double* __restrict__ loss_part,
                                                    double* __restrict__ loss) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        double l = 0.0;
        for (uint32_t t = 0; t < tiles; ++t) l += loss_part[t];
        loss[0] = l;
    }
    if (i >= kMlpParams) return;
    float g = 0.f;
    if (!(i >= kOffW1 + kMlpIn * kMlpHid && i < kOffB1))
        for (uint32_t t = 0; t < tiles; ++t) g += partial[(uint64_t)t * kMlpParams + i];
    grads[i] = g;
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

// The first nonempty bucket at or after b among this workgroup's (b, b + gridDim.x, ...); num_buckets when none.
template <typename T>
__device__ __forceinline__ uint32_t next_bucket(const ScoreArgs<T>& a, uint32_t b, uint32_t& n, uint64_t& start) {
    for (; b < a.num_buckets; b += gridDim.x) {
        n = a.len[b];
        if (n) {
            start = a.start[b];
            return b;
        }
    }
    return b;
}

// The words a thread's part of the features reads (x for parts 0 and 2, z for 1 and 2, c for 2).
template <typename T>
__device__ __forceinline__ void load_term(const ScoreArgs<T>& a, uint64_t i, bool valid, int part, uint64_t& x,
                                          uint64_t& z, double& c) {
    x = 0;
    z = 0;
    c = 1.0;
    if (!valid) return;
    if (part == 0 || part == 2) x = a.x[i];
    if (part == 1 || part == 2) z = a.z[i];
    if (part == 2) c = (double)a.c[i];
}

// Workgroups stride over the buckets, each loading its weight fragments once and walking its buckets' 64-term tiles
// with the next tile's words loaded while the current tile is scored.  Two barriers per tile: a tile's output sum
// and score are written after the next tile's first barrier, from the parity-indexed ypart and lnc.
// Four waves per SIMD: measured 312 us on 2^22 terms against 340 us at the compiler's own three (it spills five
// VGPRs to reach 128; the biases live in LDS so that the weights fit).
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_mlp_score(const MlpDev m, const ScoreArgs<T> a) {
    __shared__ FwdLds s;
    __shared__ float lnc[2][kMlpTile];  // ln max(|c|, 1e-300) by tile parity
    const FwdWeights fw = load_fwd_weights(m, s);
    const int r = threadIdx.x & 63, part = threadIdx.x >> 6;
    uint32_t n = 0;
    uint64_t start = 0;
    uint32_t b = next_bucket(a, blockIdx.x, n, start);
    if (b >= a.num_buckets) return;  // uniform: depends on blockIdx only
    uint32_t t0 = 0;
    uint64_t cx, cz;
    double cc;
    load_term(a, start + r, r < n, part, cx, cz, cc);
    uint64_t prev = 0;  // column position of the previous tile's term r, when prev_valid
    bool prev_valid = false;
    int parity = 0;
    for (;; parity ^= 1) {
        uint32_t nb = b, nt = t0 + kMlpTile, nn = n;
        uint64_t ns = start;
        if (nt >= n) {
            nt = 0;
            nb = next_bucket(a, b + gridDim.x, nn, ns);
        }
        const bool more = nb < a.num_buckets;
        uint64_t px = 0, pz = 0;
        double pc = 1.0;
        if (more) load_term(a, ns + nt + r, nt + r < nn, part, px, pz, pc);
        const bool valid = t0 + r < n;
        term_features(m, cx, cz, cc, a.remaining, valid, part, s.a0 + r * kLd0, part == 2 ? &lnc[parity][r] : nullptr);
        __syncthreads();
        // score = ln|c| + the predicted ln|read-out factor| = the predicted ln|contribution|
        if (part == 0 && prev_valid)
            a.score[prev] = mlp_output(fw, s.ypart[parity ^ 1], r) + lnc[parity ^ 1][r];
        mlp_layer1(fw, s);
        __syncthreads();
        mlp_layer2(fw, s, nullptr, s.ypart[parity]);
        prev = start + t0 + r;
        prev_valid = valid;
        if (!more) break;
        b = nb;
        t0 = nt;
        n = nn;
        start = ns;
        cx = px;
        cz = pz;
        cc = pc;
    }
    __syncthreads();
    if (part == 0 && prev_valid) a.score[prev] = mlp_output(fw, s.ypart[parity], r) + lnc[parity][r];
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

hipError_t launch_mlp_grad(hipStream_t st, const MlpDev& m, const MlpBatch& b, float* partial, double* loss_part,
                           uint64_t partial_tiles, float* grads, double* loss) {
    const uint64_t tiles = (b.count + kMlpTile - 1) / kMlpTile;
    if (tiles > partial_tiles) return hipErrorInvalidValue;
    if (tiles)
        hipLaunchKernelGGL(k_mlp_grad, dim3((uint32_t)tiles), dim3(256), 0, st, m, b, partial, loss_part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, partial, (uint32_t)tiles,
                       grads, loss_part, loss);
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
    // one resident grid: compute units x resident workgroups per unit, read once per device and type
    static int slots_of[64][2] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    int& slots = slots_of[dev][sizeof(T) == 8 ? 0 : 1];
    if (!slots) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mlp_score<T>, 256, 0) != hipSuccess || per_cu <= 0)
            per_cu = 2;
        slots = cus * per_cu;
    }
    const uint32_t grid = std::min<uint32_t>(a.num_buckets, (uint32_t)slots);
    hipLaunchKernelGGL((k_mlp_score<T>), dim3(grid), dim3(256), 0, st, m, a);
    return hipGetLastError();
}

template hipError_t launch_mlp_score<double>(hipStream_t, const MlpDev&, const ScoreArgs<double>&);
template hipError_t launch_mlp_score<float>(hipStream_t, const MlpDev&, const ScoreArgs<float>&);

}  // namespace pp
