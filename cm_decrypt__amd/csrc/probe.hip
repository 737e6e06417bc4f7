// Measured VALU issue rate of the device: the peak against which the decoder's integer
// work is compared.  Each lane runs 8 independent chains of v_xor_b32 (written as single
// instructions so the compiler cannot merge two XORs into one three-input v_bitop3), long
// enough that launch overhead vanishes; 256 threads per block, 8 blocks per CU.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "engine.h"

using namespace cmd;

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 32;

__global__ __launch_bounds__(256) void k_valu_probe(uint32_t *out, uint32_t seed, int iters)
{
    uint32_t a[kChains];
    const uint32_t b = seed * 0x9E3779B9u + threadIdx.x;
#pragma unroll
    for (int c = 0; c < kChains; c++) a[c] = seed ^ (threadIdx.x * 0x85EBCA6Bu) ^ (c * 0xC2B2AE35u);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
#pragma unroll
            for (int c = 0; c < kChains; c++) asm volatile("v_xor_b32 %0, %1, %0" : "+v"(a[c]) : "v"(b));
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int c = 0; c < kChains; c++) x ^= a[c];
    // never true for the seeds used; keeps the chains live without a store per thread
    if (x == 0x7FFFFFFFu && seed == 0u) out[blockIdx.x] = x;
}

// The same chains with a compile-time count CH of independent chains per lane (1 = every
// instruction waits for the previous one's result).
template <int CH>
__global__ __launch_bounds__(256) void k_valu_chains(uint32_t *out, uint32_t seed, int iters)
{
    uint32_t a[CH];
    const uint32_t b = seed * 0x9E3779B9u + threadIdx.x;
#pragma unroll
    for (int c = 0; c < CH; c++) a[c] = seed ^ (threadIdx.x * 0x85EBCA6Bu) ^ (c * 0xC2B2AE35u);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
#pragma unroll
            for (int c = 0; c < CH; c++) asm volatile("v_xor_b32 %0, %1, %0" : "+v"(a[c]) : "v"(b));
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int c = 0; c < CH; c++) x ^= a[c];
    if (x == 0x7FFFFFFFu && seed == 0u) out[blockIdx.x] = x;
}

// Dependent-chain latencies of one wavefront in shader-clock ticks (s_memtime) per step:
// [0] v_xor_b32 -> v_xor_b32, [1] DPP row shift -> v_xor_b32, [2] v_readlane -> scalar ->
// v_xor_b32.  Written by lane 0.
__global__ void k_chain_latency(unsigned long long *out, uint32_t seed, int iters)
{
    uint32_t v = seed ^ threadIdx.x;
    const uint32_t b = seed * 7u + threadIdx.x;
    const unsigned long long t0 = __builtin_amdgcn_s_memtime();
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) asm volatile("v_xor_b32 %0, %1, %0" : "+v"(v) : "v"(b));
    }
    const unsigned long long t1 = __builtin_amdgcn_s_memtime();
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
            v = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111 /* row_shr:1 */, 0xF, 0xF, true) ^ b;
    }
    const unsigned long long t2 = __builtin_amdgcn_s_memtime();
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v ^= (uint32_t)__builtin_amdgcn_readlane((int)v, 5) + 1u;
    }
    const unsigned long long t3 = __builtin_amdgcn_s_memtime();
    if (threadIdx.x == 0) {
        out[0] = t1 - t0;
        out[1] = t2 - t1;
        out[2] = t3 - t2;
        out[3] = v;
    }
}

template <int CH>
hipError_t launch_chains(int blocks, uint32_t *out, uint32_t seed, int iters, hipStream_t st)
{
    hipLaunchKernelGGL(k_valu_chains<CH>, dim3(blocks), dim3(256), 0, st, out, seed, iters);
    return hipGetLastError();
}

}  // namespace

extern "C" int cmd_valu_issue_rate(cmd_ctx *ctx, int32_t waves_per_simd, int32_t chains, double *wave_insts_per_clk)
{
    if (!ctx || !wave_insts_per_clk || waves_per_simd < 1 || waves_per_simd > 8) return CMD_ERR_ARG;
    if (chains != 1 && chains != 2 && chains != 4 && chains != 8) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    int clock_khz = 0;
    CMD_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, ctx->device));
    // 256-thread blocks hold one wavefront per SIMD; waves_per_simd blocks per CU
    const int blocks = ctx->n_cu * waves_per_simd;
    DevBuf out;
    int rc = dev_alloc(ctx, out, (size_t)blocks * 4);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int iters = 2048;
    double best = 0.0;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    for (int rep = 0; rep < 3 && err == hipSuccess; rep++) {
        err = hipEventRecord(e0, ctx->stream);
        uint32_t *o = (uint32_t *)out.ptr;
        if (err == hipSuccess) {
            switch (chains) {
            case 1: err = launch_chains<1>(blocks, o, 0x1234u + rep, iters, ctx->stream); break;
            case 2: err = launch_chains<2>(blocks, o, 0x1234u + rep, iters, ctx->stream); break;
            case 4: err = launch_chains<4>(blocks, o, 0x1234u + rep, iters, ctx->stream); break;
            default: err = launch_chains<8>(blocks, o, 0x1234u + rep, iters, ctx->stream); break;
            }
        }
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && rep > 0 && ms > 0.f) {
            // wave instructions per SIMD per clock at the device's peak clock
            const double insts = (double)blocks * 4.0 * iters * kUnroll * chains;
            const double per = insts / ((double)ctx->n_cu * 4.0) / (ms * 1e-3 * clock_khz * 1e3);
            best = std::max(best, per);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(ctx, out);
    if (err != hipSuccess) return hip_fail(err, "cmd_valu_issue_rate");
    *wave_insts_per_clk = best;
    return CMD_OK;
}

extern "C" int cmd_chain_latency(cmd_ctx *ctx, double ticks[3])
{
    if (!ctx || !ticks) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    DevBuf out;
    int rc = dev_alloc(ctx, out, 4 * sizeof(unsigned long long));
    if (rc) return rc;
    const int iters = 256;
    hipLaunchKernelGGL(k_chain_latency, dim3(1), dim3(64), 0, ctx->stream, (unsigned long long *)out.ptr, 7u, iters);
    hipError_t err = hipGetLastError();
    unsigned long long h[4] = {};
    if (err == hipSuccess) err = hipMemcpyAsync(h, out.ptr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, out);
    if (err != hipSuccess) return hip_fail(err, "cmd_chain_latency");
    for (int i = 0; i < 3; i++) ticks[i] = (double)h[i] / ((double)iters * kUnroll);
    return CMD_OK;
}

extern "C" int cmd_valu_peak(cmd_ctx *ctx, double *lane_ops_per_s)
{
    if (!ctx || !lane_ops_per_s) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    DevBuf out;
    const int blocks = ctx->n_cu * 8;
    int rc = dev_alloc(ctx, out, (size_t)blocks * 4);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best = 0.0;
    const int iters = 4096;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    for (int rep = 0; rep < 4 && err == hipSuccess; rep++) {
        err = hipEventRecord(e0, ctx->stream);
        if (err == hipSuccess) {
            hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, ctx->stream, (uint32_t *)out.ptr,
                               0x1234u + rep, iters);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && rep > 0 && ms > 0.f) {
            const double ops = (double)blocks * 256.0 * iters * kUnroll * kChains;
            best = std::max(best, ops / (ms * 1e-3));
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(ctx, out);
    if (err != hipSuccess) return hip_fail(err, "cmd_valu_peak");
    *lane_ops_per_s = best;
    return CMD_OK;
}
