// On-device key expansion: the per-key arrays of a resident key (KeyDev) built on the GPU
// from the secret key's Goppa coefficients and Benes control bits, instead of on the host.
// It computes exactly what expand_key (keys.cpp) computes:
//   L        support_gen (CPU/src/common/root_13.c:181-211): the Benes network applied to
//            bitrev_m(i), i < 2^m, keeping the first n values;
//   table    compute_inverses (CPU/param/param6960119/src/util.c:86-100):
//            table[i][j] = L_i^j / g(L_i)^2 for i < sb, j < 2t;
//   lbits    bit planes of L, 32 support positions per word;
//   benes    per layer and 32-bit word, the positions whose switch is set (for strides of 32
//            and more both words of a pair carry the pair's bit);
//   ypl      bit planes of 1/g(a)^2 at every field element a in FFT order (word w, bit b holds
//            a = bitrev_m(32w + b)), the weights of the decapsulation check's syndrome.
#include <hip/hip_runtime.h>

#include "dev_gf.h"
#include "kernels.h"
#include "params.h"

namespace cmd {

namespace {

constexpr int kSupportThreads = 1024;

// One workgroup: the 2^m values in LDS through the 2m-1 Benes layers, then L and its planes.
template <int IDX>
__global__ __launch_bounds__(kSupportThreads) void k_key_support(const uint8_t *__restrict__ cond, uint16_t *L,
                                                                 uint32_t *lbits, int lstride)
{
    using Pm = P<IDX>;
    constexpr int M = Pm::m, N = Pm::n, COUNT = 1 << M;
    constexpr int LAYER_BYTES = COUNT / 16;
    __shared__ uint16_t v[COUNT];
    const int tid = threadIdx.x;
    for (int i = tid; i < COUNT; i += kSupportThreads)
        v[i] = (uint16_t)(__builtin_bitreverse32((uint32_t)i) >> (32 - M));
    __syncthreads();
    for (int layer = 0; layer < 2 * M - 1; layer++) {
        const int lg = layer < M ? layer : 2 * M - 2 - layer;
        const int s = 1 << lg;
        const uint8_t *bits = cond + (size_t)layer * LAYER_BYTES;
        for (int j = tid; j < COUNT / 2; j += kSupportThreads) {
            const int p = ((j >> lg) << (lg + 1)) | (j & (s - 1));
            if ((bits[j >> 3] >> (j & 7)) & 1) {
                const uint16_t a = v[p], b = v[p + s];
                v[p] = b;
                v[p + s] = a;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < N; i += kSupportThreads) L[i] = v[i];
    // bit planes: wave w covers positions 64c .. 64c+63 for its chunks c
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int CHUNKS = (N + 63) / 64;
    for (int c = wave; c < CHUNKS; c += kSupportThreads / 64) {
        const int i = 64 * c + lane;
        const uint32_t val = i < N ? v[i] : 0u;
#pragma unroll
        for (int k = 0; k < M; k++) {
            const uint64_t ball = __ballot((val >> k) & 1u);
            if (lane == 0) lbits[(size_t)k * lstride + 2 * c] = (uint32_t)ball;
            if (lane == 1 && 2 * c + 1 < lstride) lbits[(size_t)k * lstride + 2 * c + 1] = (uint32_t)(ball >> 32);
        }
    }
    // the rest of each plane row is zero
    for (int idx = tid; idx < M * lstride; idx += kSupportThreads) {
        const int w = idx % lstride;
        if (w >= 2 * CHUNKS) lbits[idx] = 0;
    }
}

// One wavefront per table row i < sb: g(L_i) by Horner, 1/g(L_i)^2, then the 2t powers.
template <int IDX>
__global__ __launch_bounds__(256) void k_key_table(const uint8_t *__restrict__ gbytes, const uint16_t *__restrict__ L,
                                                   uint16_t *table)
{
    using Pm = P<IDX>;
    constexpr int M = Pm::m, T = Pm::t, SB = Pm::sb;
    constexpr uint32_t MASK = (1u << M) - 1u;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= SB) return;
    const uint32_t a = L[row];
    // Horner with g_t = 1 leading: g(a) = (((1*a + g_{t-1})*a + ...)*a + g_0
    uint32_t ev = 1;
    for (int i = T - 1; i >= 0; i--) {
        const uint32_t gi = ((uint32_t)gbytes[2 * i] | ((uint32_t)gbytes[2 * i + 1] << 8)) & MASK;
        ev = gf_mul<M>(ev, a) ^ gi;
    }
    const uint32_t inv = gf_inv<M>(gf_mul<M>(ev, ev));
    // a^lane by square and multiply over the 6 bits of lane, a^64 by six squarings
    uint32_t pw = 1, sq = a;
#pragma unroll
    for (int b = 0; b < 6; b++) {
        if ((lane >> b) & 1) pw = gf_mul<M>(pw, sq);
        sq = gf_mul<M>(sq, sq);
    }
    uint32_t val = gf_mul<M>(inv, pw);  // inv * a^lane
    uint16_t *out = table + (size_t)row * 2 * T;
    for (int j = lane; j < 2 * T; j += 64) {
        out[j] = (uint16_t)val;
        val = gf_mul<M>(val, sq);  // times a^64
    }
}

// One thread per field element c (FFT order): y = 1/g(bitrev_m(c))^2, stored as m bit planes.
template <int IDX>
__global__ __launch_bounds__(256) void k_key_ypl(const uint8_t *__restrict__ gbytes, uint32_t *ypl)
{
    using Pm = P<IDX>;
    constexpr int M = Pm::m, T = Pm::t, WORDS = 1 << (M - 5);
    constexpr uint32_t MASK = (1u << M) - 1u;
    const int c = blockIdx.x * 256 + threadIdx.x;  // the grid covers 2^m exactly
    const int lane = threadIdx.x & 63;
    const uint32_t a = __builtin_bitreverse32((uint32_t)c) >> (32 - M);
    uint32_t ev = 1;
    for (int i = T - 1; i >= 0; i--) {
        const uint32_t gi = ((uint32_t)gbytes[2 * i] | ((uint32_t)gbytes[2 * i + 1] << 8)) & MASK;
        ev = gf_mul<M>(ev, a) ^ gi;
    }
    const uint32_t y = gf_inv<M>(gf_mul<M>(ev, ev));
    const int w = 2 * (c >> 6);
#pragma unroll
    for (int k = 0; k < M; k++) {
        const uint64_t ball = __ballot((y >> k) & 1u);
        if (lane == 0) ypl[k * WORDS + w] = (uint32_t)ball;
        if (lane == 1) ypl[k * WORDS + w + 1] = (uint32_t)(ball >> 32);
    }
}

// One thread per (layer, 32-bit word) of the swap masks.
template <int IDX>
__global__ __launch_bounds__(256) void k_key_benes(const uint8_t *__restrict__ cond, uint32_t *benes)
{
    using Pm = P<IDX>;
    constexpr int M = Pm::m, WORDS = 1 << (M - 5), LAYER_BYTES = (1 << M) / 16;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= (2 * M - 1) * WORDS) return;
    const int layer = idx / WORDS, w = idx % WORDS;
    const int sft = layer < M ? layer : 2 * M - 2 - layer;
    const uint8_t *bits = cond + (size_t)layer * LAYER_BYTES;
    uint32_t mask = 0;
    for (int b = 0; b < 32; b++) {
        int c = 32 * w + b;
        if ((c >> sft) & 1) {
            if (sft < 5) continue;  // the lower position of an in-word pair carries the bit
            c -= 1 << sft;          // the upper word carries its pair's bit too
        }
        const int j = ((c >> (sft + 1)) << sft) | (c & ((1 << sft) - 1));
        if ((bits[j >> 3] >> (j & 7)) & 1) mask |= 1u << b;
    }
    benes[idx] = mask;
}

template <int IDX>
hipError_t launch_expand(const uint8_t *d_key, const KeyDev &k, hipStream_t st)
{
    using Pm = P<IDX>;
    const uint8_t *gbytes = d_key;               // t coefficients, 2 bytes each (little endian)
    const uint8_t *cond = d_key + 2 * Pm::t;     // control bits
    hipLaunchKernelGGL(k_key_support<IDX>, dim3(1), dim3(kSupportThreads), 0, st, cond, const_cast<uint16_t *>(k.L),
                       const_cast<uint32_t *>(k.lbits), (int)k.lstride);
    hipLaunchKernelGGL(k_key_table<IDX>, dim3((Pm::sb + 3) / 4), dim3(256), 0, st, gbytes, k.L,
                       const_cast<uint16_t *>(k.table));
    const int mask_words = (2 * Pm::m - 1) << (Pm::m - 5);
    hipLaunchKernelGGL(k_key_benes<IDX>, dim3((mask_words + 255) / 256), dim3(256), 0, st, cond,
                       const_cast<uint32_t *>(k.benes));
    static_assert(((1 << Pm::m) % 256) == 0, "whole blocks of field elements");
    hipLaunchKernelGGL(k_key_ypl<IDX>, dim3((1 << Pm::m) / 256), dim3(256), 0, st, gbytes,
                       const_cast<uint32_t *>(k.ypl));
    return hipGetLastError();
}

}  // namespace

hipError_t launch_key_expand(int pidx, const uint8_t *d_key, const KeyDev &k, hipStream_t st)
{
    switch (pidx) {
    case 0: return launch_expand<0>(d_key, k, st);
    case 1: return launch_expand<1>(d_key, k, st);
    case 2: return launch_expand<2>(d_key, k, st);
    case 3: return launch_expand<3>(d_key, k, st);
    case 4: return launch_expand<4>(d_key, k, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace cmd
