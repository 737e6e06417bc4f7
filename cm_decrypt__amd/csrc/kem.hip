// KEM decapsulation around the decoder: the check of a decoded error vector and the session
// key of crypto_kem_dec (Classic McEliece).  The reference stops at the error vector; its
// encapsulation (kem/common/core/operations.c:15-34) fixes the session key of a valid
// ciphertext as SHAKE256(1 || e || c), 32 bytes.  Decapsulation accepts e when its weight is t
// and its syndrome over the whole support equals the ciphertext's, and otherwise returns
// SHAKE256(0 || s || c) with the secret key's s (implicit rejection), selecting by mask so that
// the two cases run the same instructions.
//
// k_decap_check: one workgroup per tile of a fused plan.  By linearity the check is s(v) == 0
// for v = (c || 0) ^ e, where s_j(v) = sum_{i < n} v_i L_i^j / g(L_i)^2 (j < 2t).  Per
// ciphertext, v is taken from support order to FFT order by the key's Benes network in reverse,
// multiplied by the weights 1/g(a)^2 (key planes in FFT order), and reduced to the 2t power
// sums by the transposed additive FFT: butterflies in reverse, leaves summed, then the
// coefficient phase in reverse, bit-sliced across the tile.  Cost is that of the root search's
// FFT, against n * 2t multiply-adds for the direct sum.
//
// k_kem_pack / k_keccak: the padded SHAKE256 input per item (one wavefront per item, 64-bit
// stores), then Keccak-f[1600] per item (one lane per item).
#include "fused_impl.h"

namespace cmd {

namespace {

__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v |= (uint32_t)__shfl_xor((int)v, off);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

template <int IDX>
struct DecapTile {
    static constexpr int TILE = 32;
    using Pm = P<IDX>;
    static constexpr int M = Pm::m, T = Pm::t, EB = Pm::e_bytes, CB = Pm::synd_bytes;
    static constexpr int LOG = (2 * T > 128) ? 8 : 7;  // 2^LOG >= 2t power sums
    using FP = FftPlan<M, LOG>;
    static constexpr int NP = FP::NP, NW = FP::NW, WPL = FP::WPL;
    static_assert(FP::WPLEAF == 1, "one word per leaf");
    static_assert(TILE * (1 << LOG) * 2 <= kWaveLds * 4, "leaves must fit the wave's LDS");

    // Bytes 4g .. 4g+3 of a row of `limit` bytes (missing bytes read as zero).
    static __device__ __forceinline__ uint32_t row_word(const uint8_t *row, int g, int limit, bool aligned)
    {
        if (aligned && 4 * g + 4 <= limit) return *reinterpret_cast<const uint32_t *>(row + 4 * g);
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (4 * g + b < limit) v |= (uint32_t)row[4 * g + b] << (8 * b);
        return v;
    }

    // The tile's wavefronts share the per-ciphertext part (slot x goes to wavefront x mod
    // nwaves); wavefront 0 then runs the coefficient phase over the leaves in LDS.  *wbits
    // (LDS, zero on entry) collects the weight bits.
    static __device__ __forceinline__ void run(const DecapArgs &a, const FusedTile &tl, uint32_t *lds,
                                               uint32_t *wbits, int lane, int wave, int nwaves)
    {
        const FftTab *tab = a.fft + fft_variant(M, LOG);
        uint16_t *leaf = reinterpret_cast<uint16_t *>(lds);
        const bool e_al = ((a.e_stride & 3u) == 0) && ((reinterpret_cast<uintptr_t>(a.e) & 3u) == 0);
        const bool c_al = ((a.ct_stride & 3u) == 0) && ((reinterpret_cast<uintptr_t>(a.ct) & 3u) == 0);
        uint32_t weight_ok = 0;  // bit x: e of slot x has weight t (wave-uniform)
        for (int x = wave; x < tl.count; x += nwaves) {
            const uint32_t wl_fin = FP::lane_part(FP::SCHED.fin, opaque(lane), 64);
            const int item = a.items[tl.first + x];
            const KeyDev *kd = &a.keys[a.item_key[tl.first + x]];
            const uint8_t *er = a.e + (size_t)item * a.e_stride;
            const uint8_t *cr = a.ct ? a.ct + (size_t)item * a.ct_stride : nullptr;
            uint32_t z[WPL], wsum = 0;
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                const int g = (int)FP::final_word(wl_fin, r);
                const uint32_t ev = row_word(er, g, EB, e_al);
                wsum += __builtin_popcount(ev);
                z[r] = ev ^ (cr ? row_word(cr, g, CB, c_al) : 0u);
            }
            if (wave_sum(wsum) == (uint32_t)T) weight_ok |= 1u << x;
            // support order -> FFT order, then the weights 1/g(a)^2
            FP::benes_all_rev(z, gptr(kd->benes), lane, wl_fin, std::make_integer_sequence<int, 2 * M - 1>{});
            gu32 *yp = gptr(kd->ypl);
            uint32_t W[WPL][M];
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                const int g = (int)FP::final_word(wl_fin, r);
#pragma unroll
                for (int k = 0; k < M; k++) W[r][k] = z[r] & yp[k * NW + g];
            }
            FP::template butterfly_tr<0>(W, lane, tab);
            FP::template butterfly_tr<1>(W, lane, tab);
            FP::template butterfly_tr<2>(W, lane, tab);
            FP::template butterfly_tr<3>(W, lane, tab);
            FP::template butterfly_tr<4>(W, lane, tab);
            FP::template butterfly_tr<5>(W, lane, tab);
            FP::template butterfly_tr<6>(W, lane, tab);
            if constexpr (LOG == 8) FP::template butterfly_tr<7>(W, lane, tab);
            // leaves: the sum over the 32 points of word w = lane + 64 r
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                const int w = lane + 64 * r;
                uint32_t v = 0;
#pragma unroll
                for (int k = 0; k < M; k++) v |= (uint32_t)(__builtin_popcount(W[r][k]) & 1) << k;
                leaf[x * (1 << LOG) + bitrev_bits((uint32_t)w, LOG)] = (uint16_t)v;
            }
        }
        if (lane == 0 && weight_ok) atomicOr(wbits, weight_ok);
        __syncthreads();
        if (wave != 0) return;
        weight_ok = uni(*wbits);
        // coefficient phase in reverse, bit-sliced across the tile (bit x = slot x)
        uint32_t F[NP][M];
#pragma unroll
        for (int pb = 0; pb < NP; pb++)
#pragma unroll
            for (int k = 0; k < M; k++) F[pb][k] = 0;
        for (int x = 0; x < tl.count; x++) {
#pragma unroll
            for (int pb = 0; pb < NP; pb++) {
                const uint32_t v = leaf[x * (1 << LOG) + lane + 64 * pb];
#pragma unroll
                for (int k = 0; k < M; k++) F[pb][k] |= ((v >> k) & 1u) << x;
            }
        }
        if constexpr (LOG == 8) FP::template coeff_depth_tr<7>(F, lane, tab);
        FP::template coeff_depth_tr<6>(F, lane, tab);
        FP::template coeff_depth_tr<5>(F, lane, tab);
        FP::template coeff_depth_tr<4>(F, lane, tab);
        FP::template coeff_depth_tr<3>(F, lane, tab);
        FP::template coeff_depth_tr<2>(F, lane, tab);
        FP::template coeff_depth_tr<1>(F, lane, tab);
        FP::template coeff_depth_tr<0>(F, lane, tab);
        // F[pb][k] bit x: bit k of s_p(v) of slot x, p = lane + 64 pb
        if (a.s_out) {
            for (int x = 0; x < tl.count; x++) {
                uint16_t *so = a.s_out + (size_t)a.items[tl.first + x] * (2 * T);
#pragma unroll
                for (int pb = 0; pb < NP; pb++) {
                    const int p = lane + 64 * pb;
                    if (p >= 2 * T) continue;
                    uint32_t v = 0;
#pragma unroll
                    for (int k = 0; k < M; k++) v |= ((F[pb][k] >> x) & 1u) << k;
                    so[p] = (uint16_t)v;
                }
            }
        }
        if (a.ok) {
            uint32_t nz = 0;
#pragma unroll
            for (int pb = 0; pb < NP; pb++) {
                if (lane + 64 * pb >= 2 * T) continue;
#pragma unroll
                for (int k = 0; k < M; k++) nz |= F[pb][k];
            }
            const uint32_t good = weight_ok & ~wave_or(nz);
            if (lane < tl.count) a.ok[a.items[tl.first + lane]] = (good >> lane) & 1u;
        }
    }
};

constexpr int kDecapWaves = 4;  // wavefronts per tile (workgroup)

__global__ __launch_bounds__(64 * kDecapWaves) void k_decap_check(DecapArgs a)
{
    __shared__ uint32_t lds[32 * 256 / 2];
    __shared__ uint32_t wbits;
    const FusedTile tl = a.tiles[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) wbits = 0;
    __syncthreads();
    switch (tl.pidx) {
    case 0: DecapTile<0>::run(a, tl, lds, &wbits, lane, wave, kDecapWaves); break;
    case 1: DecapTile<1>::run(a, tl, lds, &wbits, lane, wave, kDecapWaves); break;
    case 2: DecapTile<2>::run(a, tl, lds, &wbits, lane, wave, kDecapWaves); break;
    case 3: DecapTile<3>::run(a, tl, lds, &wbits, lane, wave, kDecapWaves); break;
    case 4: DecapTile<4>::run(a, tl, lds, &wbits, lane, wave, kDecapWaves); break;
    default: break;
    }
}

// ---- SHAKE256 (FIPS 202) -----------------------------------------------------------------
constexpr int kRate = 136;  // SHAKE256 rate in bytes (capacity 512 bits)

// rc(t) of FIPS 202 3.2.5: the LFSR over x^8 + x^6 + x^5 + x^4 + 1 started at R = 1.
constexpr int rc_bit(int t)
{
    unsigned R = 1;
    for (int i = 1; i <= t % 255; i++) {
        R <<= 1;
        const unsigned r8 = (R >> 8) & 1u;
        R = (R ^ (r8 | (r8 << 4) | (r8 << 5) | (r8 << 6))) & 0xFFu;
    }
    return (int)(R & 1u);
}

// iota's round constant: bit 2^j - 1 is rc(j + 7 i).
constexpr uint64_t round_constant(int i)
{
    uint64_t c = 0;
    for (int j = 0; j <= 6; j++)
        if (rc_bit(j + 7 * i)) c |= 1ull << ((1 << j) - 1);
    return c;
}

// rho offsets of FIPS 202 3.2.2, lane x + 5y.
struct RhoTable {
    int r[25];
};
constexpr RhoTable make_rho()
{
    RhoTable o{};
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
        o.r[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        const int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx;
        y = ny;
    }
    return o;
}
constexpr RhoTable kRho = make_rho();

__device__ __forceinline__ uint64_t rotl64(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

// One round of Keccak-f[1600]; RC is iota's constant, a compile-time value per round.
template <uint64_t RC>
__device__ __forceinline__ void keccak_round(uint64_t (&A)[25])
{
    uint64_t C[5], B[25];
#pragma unroll
    for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        const uint64_t D = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
#pragma unroll
        for (int y = 0; y < 5; y++) A[x + 5 * y] ^= D;
    }
    // rho and pi: lane (x, y) moves to (y, 2x + 3y)
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
        for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[x + 5 * y], kRho.r[x + 5 * y]);
    // chi
#pragma unroll
    for (int y = 0; y < 5; y++)
#pragma unroll
        for (int x = 0; x < 5; x++) A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
    A[0] ^= RC;
}

template <int... Rs>
__device__ __forceinline__ void keccak_rounds(uint64_t (&A)[25], std::integer_sequence<int, Rs...>)
{
    (keccak_round<round_constant(Rs)>(A), ...);
}

__device__ __forceinline__ void keccak_f1600(uint64_t (&A)[25])
{
    keccak_rounds(A, std::make_integer_sequence<int, 24>{});
}

struct KemSizes {
    int eb, cb, blocks;
};
__device__ __forceinline__ KemSizes kem_sizes(int pidx)
{
    int eb = 0, cb = 0;
    switch (pidx) {
    case 0: eb = P<0>::e_bytes; cb = P<0>::synd_bytes; break;
    case 1: eb = P<1>::e_bytes; cb = P<1>::synd_bytes; break;
    case 2: eb = P<2>::e_bytes; cb = P<2>::synd_bytes; break;
    case 3: eb = P<3>::e_bytes; cb = P<3>::synd_bytes; break;
    default: eb = P<4>::e_bytes; cb = P<4>::synd_bytes; break;
    }
    return KemSizes{eb, cb, (1 + eb + cb) / kRate + 1};
}
static_assert(((1 + kMaxEBytes + kMaxSyndBytes) / kRate + 1) * kRate <= kKemMsgStride, "message stride");

// One wavefront per slot: b || (b ? e : s) || c, then the SHAKE padding (0x1F ... 0x80).
__global__ __launch_bounds__(64) void k_kem_pack(KemHashArgs a)
{
    const int q = blockIdx.x, lane = threadIdx.x;
    const int item = a.items[q];
    const KeyDev &kd = a.keys[a.item_key[q]];
    const KemSizes z = kem_sizes(kd.pidx);
    const int len = 1 + z.eb + z.cb, total = z.blocks * kRate;
    const uint32_t b = a.ok[item] & 1u, keep = 0u - b;  // all ones when e is accepted
    const uint8_t *er = a.e + (size_t)item * a.e_stride;
    const uint8_t *cr = a.ct + (size_t)item * a.ct_stride;
    uint64_t *dst = reinterpret_cast<uint64_t *>(a.msg + (size_t)q * kKemMsgStride);
    for (int w = lane; w < total / 8; w += 64) {
        uint64_t word = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int o = 8 * w + j;
            uint32_t v = 0;
            if (o == 0)
                v = b;
            else if (o <= z.eb)
                v = ((uint32_t)er[o - 1] & keep) | ((uint32_t)kd.s[o - 1] & ~keep);
            else if (o < len)
                v = cr[o - 1 - z.eb];
            if (o == len) v ^= 0x1Fu;
            if (o == total - 1) v ^= 0x80u;
            word |= (uint64_t)(v & 0xFFu) << (8 * j);
        }
        dst[w] = word;
    }
}

// One lane per slot: absorb the padded blocks, squeeze 32 bytes.
__global__ __launch_bounds__(256) void k_keccak(KemHashArgs a)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= a.count) return;
    const KemSizes z = kem_sizes(a.keys[a.item_key[q]].pidx);
    const uint64_t *m = reinterpret_cast<const uint64_t *>(a.msg + (size_t)q * kKemMsgStride);
    uint64_t A[25];
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
    for (int blk = 0; blk < z.blocks; blk++) {
#pragma unroll
        for (int i = 0; i < kRate / 8; i++) A[i] ^= m[blk * (kRate / 8) + i];
        keccak_f1600(A);
    }
    // A ciphertext of 6960119 (m t = 1547 bits in 194 bytes) whose five padding bits are not
    // zero is refused outright: its session key is all zero bytes, as the Classic McEliece
    // specification's decapsulation for that set prescribes (the reference has no
    // decapsulation).  The padding bits are public, so the choice leaks nothing.
    const int item = a.items[q];
    const int pidx = a.keys[a.item_key[q]].pidx;
    constexpr int kPadShift = (P<3>::m * P<3>::t) % 8;
    uint32_t last = 0;
    if (pidx == 3) last = a.ct[(size_t)item * a.ct_stride + P<3>::synd_bytes - 1];
    const uint64_t keep = (last >> kPadShift) != 0 ? 0ull : ~0ull;
    uint8_t *out = a.key_out + (size_t)item * a.key_stride;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int b = 0; b < 8; b++) out[8 * i + b] = (uint8_t)((A[i] & keep) >> (8 * b));
}

}  // namespace

hipError_t launch_decap_check(const DecapArgs &a, hipStream_t st)
{
    if (a.n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_decap_check, dim3(a.n_tiles), dim3(64 * kDecapWaves), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_kem_hash(const KemHashArgs &a, hipStream_t st)
{
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kem_pack, dim3(a.count), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_keccak, dim3((a.count + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace cmd
