// GPU side of the seeded ciphertext generator: the same ciphertexts as the host generator
// (gen.cpp, encryptor_init / encrypt_one), for many keys at once.
//
// The reference's encryption multiplies e by the systematic public key
// (kem/common/core/encrypt.c:87-120, pk from the Gaussian elimination of pk_gen.c:35-149).
// Equivalently, with H the mt x n parity-check matrix of (g, L) (bit i*m + k of column j =
// bit k of L_j^i / g(L_j)) and A its first mt columns, ct = A^-1 H e.  Per key:
//   k_gen_cols    the n columns of H (one thread per column: g(L_j) by Horner, inverse, powers);
//   k_gen_aug     [A | I] as bit rows (32 x 32 bit transposes of the column words);
//   k_gen_invert  Gauss-Jordan on [A | I] (one workgroup per key; pivot search, row swap and
//                 elimination by whole wavefronts per row) leaving [I | A^-1]; a key whose A
//                 is singular is flagged (with rank_only, within A alone: key-pair candidates
//                 need only that flag, the systematic-form test of pk_gen.c:35-149);
//   k_gen_ainvt   A^-1 as columns.
// Per ciphertext (one wavefront): y = H e as the XOR of the columns at its error positions
// (lane = word), then ct = A^-1 y as the XOR of the columns of A^-1 at the set bits of y.
#include <hip/hip_runtime.h>

#include "dev_gf.h"
#include "kernels.h"
#include "params.h"

namespace cmd {

namespace {

template <int IDX>
struct Gen {
    using Pm = P<IDX>;
    static constexpr int M = Pm::m, T = Pm::t, N = Pm::n, MT = M * T;
    static constexpr int CW = (MT + 31) / 32;  // 32-bit words of an mt-bit column or A row
    static constexpr int RW = 2 * CW;          // words of an [A | I] row
};

template <int IDX>
__global__ __launch_bounds__(256) void k_gen_cols(GenArgs a)
{
    using G = Gen<IDX>;
    constexpr uint32_t MASK = (1u << G::M) - 1u;
    const int j = blockIdx.x * 256 + threadIdx.x, ks = blockIdx.y;
    if (j >= G::N) return;
    uint32_t x, ev = 1;  // g(x), g_t = 1 leading
    if (a.cand_L) {
        const uint16_t *cg = a.cand_g + (size_t)ks * G::T;
        x = a.cand_L[(size_t)ks * G::N + j];
        for (int i = G::T - 1; i >= 0; i--) ev = gf_mul<G::M>(ev, x) ^ ((uint32_t)cg[i] & MASK);
    } else {
        const KeyDev &kd = a.keys[a.key_ids[ks]];
        x = kd.L[j];
        for (int i = G::T - 1; i >= 0; i--) {
            const uint32_t gi = ((uint32_t)kd.gbytes[2 * i] | ((uint32_t)kd.gbytes[2 * i + 1] << 8)) & MASK;
            ev = gf_mul<G::M>(ev, x) ^ gi;
        }
    }
    uint32_t v = gf_inv<G::M>(ev);
    uint32_t *col = a.cols + ((size_t)ks * G::N + j) * G::CW;
    uint32_t word = 0;
    int r = 0;
    for (int i = 0; i < G::T; i++) {
#pragma unroll
        for (int k = 0; k < G::M; k++, r++) {
            word |= ((v >> k) & 1u) << (r & 31);
            if ((r & 31) == 31) {
                col[r >> 5] = word;
                word = 0;
            }
        }
        v = gf_mul<G::M>(v, x);
    }
    if (r & 31) col[r >> 5] = word;
}

// 32 x 32 bit transpose: out[i] bit b = in[b] bit i.
__device__ __forceinline__ void transpose32(const uint32_t (&in)[32], uint32_t (&out)[32])
{
#pragma unroll
    for (int i = 0; i < 32; i++) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 32; b++) v |= ((in[b] >> i) & 1u) << b;
        out[i] = v;
    }
}

// Thread (rb, w): rows 32rb .. 32rb+31 of [A | I], word w of A and of I.
template <int IDX>
__global__ __launch_bounds__(256) void k_gen_aug(GenArgs a)
{
    using G = Gen<IDX>;
    const int idx = blockIdx.x * 256 + threadIdx.x, ks = blockIdx.y;
    if (idx >= G::CW * G::CW) return;
    const int rb = idx / G::CW, w = idx % G::CW;
    const uint32_t *cols = a.cols + (size_t)ks * G::N * G::CW;
    uint32_t in[32], out[32];
#pragma unroll
    for (int b = 0; b < 32; b++) {
        const int c = 32 * w + b;  // column of A
        in[b] = c < G::MT ? cols[(size_t)c * G::CW + rb] : 0u;
    }
    transpose32(in, out);
    uint32_t *aug = a.aug + (size_t)ks * G::MT * G::RW;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        const int r = 32 * rb + i;
        if (r >= G::MT) break;
        aug[(size_t)r * G::RW + w] = out[i];
        aug[(size_t)r * G::RW + G::CW + w] = (w == rb) ? (1u << i) : 0u;
    }
}

constexpr int kInvThreads = 1024;

template <int IDX>
__global__ __launch_bounds__(kInvThreads) void k_gen_invert(GenArgs a)
{
    using G = Gen<IDX>;
    __shared__ uint32_t prow[G::RW];
    __shared__ uint32_t flag[G::CW];
    __shared__ int piv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, ks = blockIdx.x;
    constexpr int WAVES = kInvThreads / 64;
    uint32_t *A = a.aug + (size_t)ks * G::MT * G::RW;
    // words of a row taking part: all of [A | I], or only A when just its rank is wanted
    const int RWU = a.rank_only ? G::CW : G::RW;
    for (int c = 0; c < G::MT; c++) {
        const int cw = c >> 5;
        const uint32_t cb = 1u << (c & 31);
        if (tid == 0) piv = G::MT;
        for (int i = tid; i < G::CW; i += kInvThreads) flag[i] = 0;
        __syncthreads();
        for (int r = c + tid; r < G::MT; r += kInvThreads)
            if (A[(size_t)r * G::RW + cw] & cb) {
                atomicMin(&piv, r);
                break;
            }
        __syncthreads();
        const int p = piv;
        if (p >= G::MT) {  // every thread sees the same value
            if (tid == 0) a.singular[ks] = 1;
            return;
        }
        if (p != c) {
            for (int w = tid; w < RWU; w += kInvThreads) {
                const uint32_t t0 = A[(size_t)p * G::RW + w];
                A[(size_t)p * G::RW + w] = A[(size_t)c * G::RW + w];
                A[(size_t)c * G::RW + w] = t0;
            }
        }
        __syncthreads();
        for (int w = tid; w < RWU; w += kInvThreads) prow[w] = A[(size_t)c * G::RW + w];
        for (int r = tid; r < G::MT; r += kInvThreads)
            if (r != c && (A[(size_t)r * G::RW + cw] & cb)) atomicOr(&flag[r >> 5], 1u << (r & 31));
        __syncthreads();
        // rows with bit c set ^= pivot row, words cw .. RW-1 (those below cw are zero in it)
        for (int r = wave; r < G::MT; r += WAVES) {
            if (!((flag[r >> 5] >> (r & 31)) & 1u)) continue;
            for (int w = cw + lane; w < RWU; w += 64) A[(size_t)r * G::RW + w] ^= prow[w];
        }
        __syncthreads();
    }
    if (tid == 0) a.singular[ks] = 0;
}

// Thread (cb, w): columns 32cb .. 32cb+31 of A^-1, row word w (rows 32w .. 32w+31).
template <int IDX>
__global__ __launch_bounds__(256) void k_gen_ainvt(GenArgs a)
{
    using G = Gen<IDX>;
    const int idx = blockIdx.x * 256 + threadIdx.x, ks = blockIdx.y;
    if (idx >= G::CW * G::CW) return;
    const int cb = idx / G::CW, w = idx % G::CW;
    const uint32_t *A = a.aug + (size_t)ks * G::MT * G::RW;
    uint32_t in[32], out[32];
#pragma unroll
    for (int b = 0; b < 32; b++) {
        const int r = 32 * w + b;  // row of A^-1
        in[b] = r < G::MT ? A[(size_t)r * G::RW + G::CW + cb] : 0u;
    }
    transpose32(in, out);
    uint32_t *dst = a.ainvt + (size_t)ks * G::MT * G::CW;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        const int c = 32 * cb + i;
        if (c >= G::MT) break;
        dst[(size_t)c * G::CW + w] = out[i];
    }
}

// One wavefront per ciphertext (key-major order); lane w < CW holds word w.
template <int IDX>
__global__ __launch_bounds__(256) void k_gen_encrypt(GenArgs a)
{
    using G = Gen<IDX>;
    static_assert(G::CW <= 64, "one word per lane");
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (q >= a.n_keys * a.per_key) return;
    const int ks = q / a.per_key;
    const uint32_t *cols = a.cols + (size_t)ks * G::N * G::CW;
    const uint16_t *pp = a.pos + (size_t)q * a.weight;
    uint32_t y = 0;
    for (int i = 0; i < a.weight; i++) {
        const int j = pp[i];
        if (lane < G::CW) y ^= cols[(size_t)j * G::CW + lane];
    }
    const uint32_t *inv = a.ainvt + (size_t)ks * G::MT * G::CW;
    uint32_t acc = 0;
    for (int w = 0; w < G::CW; w++) {
        uint32_t bits = (uint32_t)__builtin_amdgcn_readlane((int)y, w);
        while (bits) {
            const int k = 32 * w + __builtin_ctz(bits);
            bits &= bits - 1;
            if (lane < G::CW) acc ^= inv[(size_t)k * G::CW + lane];
        }
    }
    uint8_t *ct = a.ct + (size_t)q * P<IDX>::synd_bytes;
    if (lane < G::CW) {
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (4 * lane + b < P<IDX>::synd_bytes) ct[4 * lane + b] = (uint8_t)(acc >> (8 * b));
    }
}

template <int IDX>
hipError_t gen_keys(const GenArgs &a, hipStream_t st)
{
    using G = Gen<IDX>;
    hipLaunchKernelGGL(k_gen_cols<IDX>, dim3((G::N + 255) / 256, a.n_keys), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_gen_aug<IDX>, dim3((G::CW * G::CW + 255) / 256, a.n_keys), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_gen_invert<IDX>, dim3(a.n_keys), dim3(kInvThreads), 0, st, a);
    hipLaunchKernelGGL(k_gen_ainvt<IDX>, dim3((G::CW * G::CW + 255) / 256, a.n_keys), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int IDX>
hipError_t gen_check(const GenArgs &a, hipStream_t st)
{
    using G = Gen<IDX>;
    hipLaunchKernelGGL(k_gen_cols<IDX>, dim3((G::N + 255) / 256, a.n_keys), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_gen_aug<IDX>, dim3((G::CW * G::CW + 255) / 256, a.n_keys), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_gen_invert<IDX>, dim3(a.n_keys), dim3(kInvThreads), 0, st, a);
    return hipGetLastError();
}

template <int IDX>
hipError_t gen_encrypt(const GenArgs &a, hipStream_t st)
{
    const int total = a.n_keys * a.per_key;
    hipLaunchKernelGGL(k_gen_encrypt<IDX>, dim3((total + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gen_keys(const GenArgs &a, hipStream_t st)
{
    if (a.n_keys <= 0) return hipSuccess;
    switch (a.pidx) {
    case 0: return gen_keys<0>(a, st);
    case 1: return gen_keys<1>(a, st);
    case 2: return gen_keys<2>(a, st);
    case 3: return gen_keys<3>(a, st);
    case 4: return gen_keys<4>(a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gen_check(const GenArgs &a, hipStream_t st)
{
    if (a.n_keys <= 0) return hipSuccess;
    switch (a.pidx) {
    case 0: return gen_check<0>(a, st);
    case 1: return gen_check<1>(a, st);
    case 2: return gen_check<2>(a, st);
    case 3: return gen_check<3>(a, st);
    case 4: return gen_check<4>(a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gen_encrypt(const GenArgs &a, hipStream_t st)
{
    if (a.n_keys <= 0 || a.per_key <= 0) return hipSuccess;
    switch (a.pidx) {
    case 0: return gen_encrypt<0>(a, st);
    case 1: return gen_encrypt<1>(a, st);
    case 2: return gen_encrypt<2>(a, st);
    case 3: return gen_encrypt<3>(a, st);
    case 4: return gen_encrypt<4>(a, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace cmd
