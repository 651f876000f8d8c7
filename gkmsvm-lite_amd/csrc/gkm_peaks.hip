// Peak microbenchmarks of the rooflines the hot kernels are measured against:
// HBM read bandwidth (a streaming read of a large buffer), random 128-byte
// line reads (the l-mer index's access pattern) and the int32 VALU rate
// (independent chains of XOR, add and popcount, written as single
// instructions so that each one is one lane operation).  Measurement only:
// nothing on the scoring path calls them.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gkm_internal.h"

constexpr int PEAK_THREADS = 256;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Streaming read: every thread takes 4 16-byte vectors per round (loads issued
// together), XORs them into a register; the result is stored only if it hits
// a value the data never produces, so the loads are kept and nothing is written.
__global__ __launch_bounds__(PEAK_THREADS) void k_peak_read(const u32x4 *__restrict__ src,
                                                            int64_t n, uint32_t *sink)
{
    const int64_t stride = (int64_t)gridDim.x * PEAK_THREADS;
    uint32_t acc = 0;
    int64_t i = (int64_t)blockIdx.x * PEAK_THREADS + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i);
        const u32x4 b = __builtin_nontemporal_load(src + i + stride);
        const u32x4 c = __builtin_nontemporal_load(src + i + 2 * stride);
        const u32x4 d = __builtin_nontemporal_load(src + i + 3 * stride);
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w;
        acc ^= c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; i < n; i += stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i);
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    if (acc == 0x9E3779B9u)
        sink[0] = acc;
}

// int32 VALU: 8 independent accumulators, per round each takes v_xor_b32,
// v_add_u32 and v_bcnt_u32_b32 (3 lane operations); the asm keeps the
// compiler from fusing or removing them.
__global__ __launch_bounds__(PEAK_THREADS) void k_peak_valu(int rounds, uint32_t seed, uint32_t *sink)
{
    uint32_t a[8];
    const uint32_t t = blockIdx.x * PEAK_THREADS + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++)
        a[k] = t * 2654435761u + seed + (uint32_t)k;
    const uint32_t m = seed | 1u;
    for (int r = 0; r < rounds; r++) {
        // in passes over the chains, so that consecutive instructions are independent
#pragma unroll
        for (int k = 0; k < 8; k++)
            asm volatile("v_xor_b32 %0, %1, %0" : "+v"(a[k]) : "v"(m));
#pragma unroll
        for (int k = 0; k < 8; k++)
            asm volatile("v_add_u32 %0, %1, %0" : "+v"(a[k]) : "v"(m));
#pragma unroll
        for (int k = 0; k < 8; k++)
            asm volatile("v_bcnt_u32_b32 %0, %0, %0" : "+v"(a[k]));
    }
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        x ^= a[k];
    if (x == 0x9E3779B9u)
        sink[0] = x;
}

// Random 128-byte lines: each pair of lanes reads two 16-byte pieces of one
// line (32 bytes per lane, the index query's record pattern), `depth` lines
// per lane pair in flight per round; the line of round r is a hash of the lane
// pair and r, below n_lines.  Blocks of 1,024 threads.
constexpr int GATHER_THREADS = 1024;

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <int DEPTH>
__global__ __launch_bounds__(GATHER_THREADS) void k_peak_gather(const u32x4 *__restrict__ src,
                                                                uint64_t n_lines, int rounds,
                                                                uint32_t *sink)
{
    const uint32_t gid = blockIdx.x * GATHER_THREADS + threadIdx.x;
    const uint32_t pair = gid >> 1, half = gid & 1u;
    uint32_t acc = 0;
    for (int r = 0; r < rounds; r += DEPTH) {
        u32x4 v[DEPTH][2];
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            const uint64_t line =
                (uint64_t)mix32(pair * 7919u + (uint32_t)(r + d) * 104729u) % n_lines;
            const u32x4 *q = src + line * 8 + half * 4;  // 8 vectors per 128-byte line
            v[d][0] = q[0];
            v[d][1] = q[1];
        }
#pragma unroll
        for (int d = 0; d < DEPTH; ++d)
            acc ^= v[d][0].x ^ v[d][0].y ^ v[d][0].z ^ v[d][0].w ^ v[d][1].x ^ v[d][1].y ^
                   v[d][1].z ^ v[d][1].w;
    }
    if (acc == 0x9E3779B9u)
        sink[0] = acc;
}

GKM_API int gkm_peak_gather(const void *src, int64_t bytes, void *sink, int blocks, int rounds,
                            void *stream)
{
    if (bytes < 128 || bytes % 128 != 0 || blocks < 1 || rounds < 2 || rounds % 2 != 0)
        return fail(-1, "gkm_peak_gather: bytes must be a positive multiple of 128, blocks >= 1, "
                        "rounds even and >= 2");
    hipLaunchKernelGGL(k_peak_gather<2>, dim3((unsigned)blocks), dim3(GATHER_THREADS), 0,
                       S_(stream), (const u32x4 *)src, (uint64_t)bytes / 128, rounds,
                       (uint32_t *)sink);
    GKM_CHECK_LAUNCH("gkm_peak_gather");
    return 0;
}

GKM_API int gkm_peak_read(const void *src, int64_t bytes, void *sink, int blocks, void *stream)
{
    if (bytes < 16 || bytes % 16 != 0 || blocks < 1)
        return fail(-1, "gkm_peak_read: bytes must be a positive multiple of 16, blocks >= 1");
    hipLaunchKernelGGL(k_peak_read, dim3((unsigned)blocks), dim3(PEAK_THREADS), 0, S_(stream),
                       (const u32x4 *)src, bytes / 16, (uint32_t *)sink);
    GKM_CHECK_LAUNCH("gkm_peak_read");
    return 0;
}

GKM_API int gkm_peak_valu(int rounds, void *sink, int blocks, void *stream)
{
    if (rounds < 1 || blocks < 1)
        return fail(-1, "gkm_peak_valu: rounds and blocks must be >= 1");
    hipLaunchKernelGGL(k_peak_valu, dim3((unsigned)blocks), dim3(PEAK_THREADS), 0, S_(stream),
                       rounds, 0x5bd1e995u, (uint32_t *)sink);
    GKM_CHECK_LAUNCH("gkm_peak_valu");
    return 0;
}
