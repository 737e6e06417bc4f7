// Plain per-stage decryption pipeline: three kernels per key group (synd, bm, chien),
// scalar uint16 field elements, one launch each.  Semantics follow the reference CPU
// stages (CPU/param/param6960119/src/util.c:102-114, CPU/src/common/bm.c:12-68,
// CPU/src/common/root_13.c:233-240 with decrypt.c:43-53); the launch pattern is the
// per-stage batch loop of the reference GPU drivers
// (GPU_Optimised/param/param8192128/src/cuda/Decrypt.cu:220-314).
#include <hip/hip_runtime.h>

#include "dev_gf.h"
#include "kernels.h"
#include "params.h"

namespace cmd {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// ---------------------------------------------------------------------------------------
// synd: s_j = XOR_{i < sb, bit i of ct set} table[i][j].
// One wavefront handles CPW ciphertexts of the same key; lane l owns the column pairs
// (2c, 2c+1) for c = l + 64r, so each table row streams in coalesced once per CPW
// ciphertexts.  Ciphertext words are staged in LDS and read back as wave-uniform scalars.
template <int IDX, int CPW>
__global__ __launch_bounds__(kBlock) void k_plain_synd(const uint16_t *__restrict__ table, int count,
                                                       const int32_t *__restrict__ idx,
                                                       const uint8_t *__restrict__ ct, size_t ct_stride,
                                                       uint16_t *__restrict__ s_out)
{
    using Pm = P<IDX>;
    constexpr int PAIRS = Pm::t;             // dwords per table row (2t uint16)
    constexpr int R = (PAIRS + 63) / 64;
    constexpr int WORDS = (Pm::synd_bytes + 3) / 4;
    __shared__ uint32_t sct[kWaves][CPW][WORDS];

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int first = (blockIdx.x * kWaves + wv) * CPW;

    uint8_t *bytes = reinterpret_cast<uint8_t *>(&sct[wv][0][0]);
    for (int q = lane; q < CPW * WORDS * 4; q += 64) {
        int x = q / (WORDS * 4), b = q % (WORDS * 4), k = first + x;
        uint8_t v = 0;
        if (k < count && b < Pm::synd_bytes) v = ct[(size_t)(idx ? idx[k] : k) * ct_stride + b];
        bytes[q] = v;
    }
    __syncthreads();
    if (first >= count) return;

    uint32_t acc[CPW][R];
#pragma unroll
    for (int x = 0; x < CPW; x++)
#pragma unroll
        for (int r = 0; r < R; r++) acc[x][r] = 0;

    const uint32_t *rows = reinterpret_cast<const uint32_t *>(table);
    for (int w = 0; w < WORDS; w++) {
        uint32_t cw[CPW];
#pragma unroll
        for (int x = 0; x < CPW; x++) cw[x] = __builtin_amdgcn_readfirstlane(sct[wv][x][w]);
        const int nbits = (w * 32 + 32 <= Pm::sb) ? 32 : (Pm::sb - w * 32);
#pragma unroll 8
        for (int bit = 0; bit < nbits; bit++) {
            const uint32_t *row = rows + (size_t)(w * 32 + bit) * PAIRS;
            uint32_t v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = (lane + 64 * r < PAIRS) ? row[lane + 64 * r] : 0u;
#pragma unroll
            for (int x = 0; x < CPW; x++) {
                uint32_t msk = (uint32_t)__builtin_amdgcn_sbfe((int)cw[x], bit, 1);
#pragma unroll
                for (int r = 0; r < R; r++) acc[x][r] ^= v[r] & msk;
            }
        }
    }
#pragma unroll
    for (int x = 0; x < CPW; x++) {
        int k = first + x;
        if (k >= count) break;
        uint32_t *out = reinterpret_cast<uint32_t *>(s_out + (size_t)k * 2 * Pm::t);
#pragma unroll
        for (int r = 0; r < R; r++)
            if (lane + 64 * r < PAIRS) out[lane + 64 * r] = acc[x][r];
    }
}

// ---------------------------------------------------------------------------------------
// bm: one wavefront per ciphertext, lane l holds coefficients l + 64b of C and B.
// Inversion-free Massey update C' = b*C + d*B (B, b, L replaced under the same masks as
// bm.c:42-59); every iterate is a nonzero multiple of the reference's C, and dividing by
// the constant term at the end gives the reference's normalised locator exactly.
// Control flow and addressing do not depend on secret values.
__device__ __forceinline__ uint32_t wave_xor(uint32_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v ^= __shfl_xor(v, off, 64);
    return v;
}

template <int IDX>
__global__ __launch_bounds__(kBlock) void k_plain_bm(int count, const uint16_t *__restrict__ s_in,
                                                     uint16_t *__restrict__ loc_out)
{
    using Pm = P<IDX>;
    constexpr int t = Pm::t, M = Pm::m;
    constexpr int NB = (t + 1 + 63) / 64;
    __shared__ uint16_t sS[kWaves][3 * t];  // [0, t) zero, [t, 3t) = s_0 .. s_{2t-1}

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = blockIdx.x * kWaves + wv;
    const bool active = k < count;
    for (int j = lane; j < 3 * t; j += 64)
        sS[wv][j] = (active && j >= t) ? s_in[(size_t)k * 2 * t + (j - t)] : (uint16_t)0;
    __syncthreads();
    if (!active) return;

    uint32_t C[NB], B[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        C[b] = (b == 0 && lane == 0) ? 1u : 0u;
        B[b] = (b == 0 && lane == 1) ? 1u : 0u;
    }
    uint32_t L = 0, bb = 1;
    const uint16_t *S = &sS[wv][t];
    for (int N = 0; N < 2 * t; N++) {
        uint32_t part = 0;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int i = lane + 64 * b;
            if (i <= t) part ^= gf_mul<M>(C[b], S[N - i]);
        }
        const uint32_t d = __builtin_amdgcn_readfirstlane(wave_xor(part));
        const uint32_t mne = 0u - (uint32_t)(d != 0);
        const uint32_t mle = mne & (0u - (uint32_t)(2 * L <= (uint32_t)N));
        uint32_t cb[M], cd[M];
        gf_mul_cols<M>(bb, cb);
        gf_mul_cols<M>(d, cd);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            uint32_t cn = gf_mul_by_cols<M>(cb, C[b]) ^ gf_mul_by_cols<M>(cd, B[b]);
            B[b] = (C[b] & mle) | (B[b] & ~mle);
            C[b] = cn;
        }
        L = ((N + 1 - L) & mle) | (L & ~mle);
        bb = (d & mle) | (bb & ~mle);
        // B <- x * B
        uint32_t carry = 0;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            uint32_t up = __shfl_up(B[b], 1, 64);
            uint32_t top = __shfl(B[b], 63, 64);
            B[b] = (lane == 0) ? carry : up;
            carry = top;
        }
    }
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(C[0]);
    const uint32_t inv = gf_inv<M>(c0);
    uint16_t *out = loc_out + (size_t)k * (t + 1);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int i = lane + 64 * b;
        if (i <= t) out[t - i] = (uint16_t)gf_mul<M>(C[b], inv);
    }
}

// ---------------------------------------------------------------------------------------
// chien: one workgroup per ciphertext; thread p evaluates the reversed locator at L[p]
// by Horner (root_13.c:216-230) and a ballot turns 64 zero tests into 8 bytes of e.
template <int IDX>
__global__ __launch_bounds__(kBlock) void k_plain_chien(const uint16_t *__restrict__ Lsup, int count,
                                                        const int32_t *__restrict__ idx,
                                                        const uint16_t *__restrict__ loc_in,
                                                        uint8_t *__restrict__ e, size_t e_stride)
{
    using Pm = P<IDX>;
    constexpr int t = Pm::t, M = Pm::m, n = Pm::n;
    __shared__ uint32_t sf[t + 1];
    const int k = blockIdx.x;
    for (int j = threadIdx.x; j <= t; j += kBlock) sf[j] = loc_in[(size_t)k * (t + 1) + j];
    __syncthreads();
    uint8_t *eo = e + (size_t)(idx ? idx[k] : k) * e_stride;
    const int lane = threadIdx.x & 63;
    for (int base = 0; base < n; base += kBlock) {
        const int p = base + threadIdx.x;
        const uint32_t a = p < n ? Lsup[p] : 0u;
        uint32_t r = sf[t];
        for (int j = t - 1; j >= 0; j--) r = gf_mul<M>(r, a) ^ sf[j];
        const uint64_t z = __ballot(p < n && r == 0);
        const int p0 = p - lane;  // first position of this wavefront's 64
        if (lane < 8 && p0 + 8 * lane < n) eo[p0 / 8 + lane] = (uint8_t)(z >> (8 * lane));
    }
}

// ---------------------------------------------------------------------------------------
template <int IDX>
static hipError_t synd_for(const KeyDev &key, int count, const int32_t *idx, const uint8_t *ct,
                           size_t ct_stride, uint16_t *s, hipStream_t st)
{
    constexpr int CPW = 4;
    const int waves = (count + CPW - 1) / CPW;
    const int blocks = (waves + kWaves - 1) / kWaves;
    hipLaunchKernelGGL((k_plain_synd<IDX, CPW>), dim3(blocks), dim3(kBlock), 0, st, key.table, count, idx,
                       ct, ct_stride, s);
    return hipGetLastError();
}

hipError_t launch_plain_synd(int pidx, const KeyDev &key, int count, const int32_t *idx, const uint8_t *ct,
                             size_t ct_stride, uint16_t *s, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    switch (pidx) {
    case 0: return synd_for<0>(key, count, idx, ct, ct_stride, s, st);
    case 1: return synd_for<1>(key, count, idx, ct, ct_stride, s, st);
    case 2: return synd_for<2>(key, count, idx, ct, ct_stride, s, st);
    case 3: return synd_for<3>(key, count, idx, ct, ct_stride, s, st);
    case 4: return synd_for<4>(key, count, idx, ct, ct_stride, s, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_plain_bm(int pidx, int count, const uint16_t *s, uint16_t *loc, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    const int blocks = (count + kWaves - 1) / kWaves;
    switch (pidx) {
    case 0: hipLaunchKernelGGL(k_plain_bm<0>, dim3(blocks), dim3(kBlock), 0, st, count, s, loc); break;
    case 1: hipLaunchKernelGGL(k_plain_bm<1>, dim3(blocks), dim3(kBlock), 0, st, count, s, loc); break;
    case 2: hipLaunchKernelGGL(k_plain_bm<2>, dim3(blocks), dim3(kBlock), 0, st, count, s, loc); break;
    case 3: hipLaunchKernelGGL(k_plain_bm<3>, dim3(blocks), dim3(kBlock), 0, st, count, s, loc); break;
    case 4: hipLaunchKernelGGL(k_plain_bm<4>, dim3(blocks), dim3(kBlock), 0, st, count, s, loc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_plain_chien(int pidx, const KeyDev &key, int count, const int32_t *idx, const uint16_t *loc,
                              uint8_t *e, size_t e_stride, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    switch (pidx) {
#define CMD_CHIEN_CASE(I)                                                                                    \
    case I:                                                                                                  \
        hipLaunchKernelGGL(k_plain_chien<I>, dim3(count), dim3(kBlock), 0, st, key.L, count, idx, loc, e,    \
                           e_stride);                                                                        \
        break;
        CMD_CHIEN_CASE(0) CMD_CHIEN_CASE(1) CMD_CHIEN_CASE(2) CMD_CHIEN_CASE(3) CMD_CHIEN_CASE(4)
#undef CMD_CHIEN_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cmd
