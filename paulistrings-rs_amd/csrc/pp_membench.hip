// pp_membench.hip -- an HBM ceiling probe for the measurement calls: STREAM read, write, copy and triad over f64
// arrays (the semantics of the reference's crates/membench/kernels/membench.cu), written for gfx950: 16-byte loads
// and stores, four in flight per thread, one workgroup per 4096 elements so the grid covers the chip many times over.
#include "pp_membench.h"

namespace pp {

namespace {

constexpr int kStreamBlock = 256;
constexpr int kStreamUnroll = 4;
constexpr uint64_t kStreamTile = (uint64_t)kStreamBlock * kStreamUnroll;  // double2 vectors per workgroup

// Each workgroup's partial sum lands in out[blockIdx.x]; nothing reads it, it only keeps the loads alive.
__global__ void __launch_bounds__(kStreamBlock) k_stream_read(const double2* __restrict__ a, double* __restrict__ out,
                                                              uint64_t n2) {
    __shared__ double part[kStreamBlock / 64];
    const uint64_t base = blockIdx.x * kStreamTile + threadIdx.x;
    double2 v[kStreamUnroll];
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamBlock;
        v[u] = i < n2 ? a[i] : make_double2(0.0, 0.0);
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) acc += v[u].x + v[u].y;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kStreamBlock / 64; ++w) s += part[w];
        out[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kStreamBlock) k_stream_write(double2* __restrict__ a, double v, uint64_t n2) {
    const uint64_t base = blockIdx.x * kStreamTile + threadIdx.x;
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamBlock;
        if (i < n2) a[i] = make_double2(v, v);
    }
}

__global__ void __launch_bounds__(kStreamBlock) k_stream_copy(double2* __restrict__ a, const double2* __restrict__ b,
                                                              uint64_t n2) {
    const uint64_t base = blockIdx.x * kStreamTile + threadIdx.x;
    double2 v[kStreamUnroll];
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamBlock;
        if (i < n2) v[u] = b[i];
    }
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamBlock;
        if (i < n2) a[i] = v[u];
    }
}

__global__ void __launch_bounds__(kStreamBlock) k_stream_triad(double2* __restrict__ a, const double2* __restrict__ b,
                                                               const double2* __restrict__ c, double s, uint64_t n2) {
    const uint64_t base = blockIdx.x * kStreamTile + threadIdx.x;
    double2 vb[kStreamUnroll], vc[kStreamUnroll];
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamBlock;
        if (i < n2) {
            vb[u] = b[i];
            vc[u] = c[i];
        }
    }
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamBlock;
        if (i < n2) a[i] = make_double2(vb[u].x + s * vc[u].x, vb[u].y + s * vc[u].y);
    }
}

}  // namespace

uint64_t stream_blocks(uint64_t n2) { return (n2 + kStreamTile - 1) / kStreamTile; }

hipError_t launch_stream_pass(hipStream_t st, int pass, double* a, double* b, double* c, double* partial,
                              uint64_t n2) {
    const uint64_t blocks = stream_blocks(n2);
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    const dim3 grid((uint32_t)blocks), block(kStreamBlock);
    switch (pass) {
        case 0:
            hipLaunchKernelGGL(k_stream_read, grid, block, 0, st, (const double2*)a, partial, n2);
            break;
        case 1:
            hipLaunchKernelGGL(k_stream_write, grid, block, 0, st, (double2*)a, 1.5, n2);
            break;
        case 2:
            hipLaunchKernelGGL(k_stream_copy, grid, block, 0, st, (double2*)a, (const double2*)b, n2);
            break;
        default:
            hipLaunchKernelGGL(k_stream_triad, grid, block, 0, st, (double2*)a, (const double2*)b, (const double2*)c,
                               3.0, n2);
            break;
    }
    return hipGetLastError();
}

}  // namespace pp
