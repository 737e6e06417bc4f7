// Fused persistent decryption kernel (device code shared by the kernel variants):
// syndrome -> Berlekamp-Massey -> root search in one launch, for any mix of parameter sets
// and keys.
//
// Work unit: a tile of up to 32 ciphertexts of one parameter set (the host may plan smaller
// tiles), handled by one wavefront from start to finish.  Wavefronts pull tiles from a global counter until
// the list is exhausted (every wave exits when the counter passes n_tiles).
//
//  1. syndrome (CPU/param/param6960119/src/util.c:102-114): lane l owns table columns
//     2c, 2c+1 for c = l + 64r; each run of ciphertexts of one key streams that key's table
//     rows once for the tile, four rows at a time through per-wavefront LDS tables of their
//     16 XOR combinations, each ciphertext adding the entry its nibble selects (tiles of many
//     keys, and tiles of at most three ciphertexts, read only the rows of set bits instead).
//     The 2t syndromes are transposed into LDS bit planes: word [j][k] holds bit k of s_j
//     for every ciphertext of the tile (bit x = tile slot x).
//  2. Berlekamp-Massey (CPU/src/common/bm.c:12-68), bit-sliced across the tile: lane l holds
//     planes of coefficients l + 64b of C and B (coefficient t, when t is a multiple of 64,
//     transposed over lanes instead); discrepancies reduce over lanes with DPP, the
//     per-ciphertext discrepancy, masks mne/mle, L and b live in scalar registers.
//     Inversion-free update C' = b*C + d*B keeps every iterate a nonzero multiple of the
//     reference's C, so the zeros of the locator (all that the error vector depends on) are
//     the reference's.  Tiles of at most three ciphertexts run the recurrence per ciphertext
//     with scalar field elements.
//  3. root search (root_13.c:233-240, decrypt.c:43-53), per ciphertext: either Horner at 32
//     support points per lane (bit-sliced L planes of the key), or an additive FFT over all
//     2^m field elements followed by the key's Benes network, which moves each root indicator
//     to its support position; a position is an error iff all m planes are zero there,
//     which yields the packed error-vector word directly.
// No step indexes memory or branches on secret data (the public ciphertext bits select table
// rows and LDS entries); the control flow depends only on public sizes, the ciphertext and
// the iteration counter.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <utility>

#include "dev_gf.h"
#include "kernels.h"
#include "params.h"

namespace cmd {

// Stage-truncated builds for timing the stages of the single-wavefront kernel (Makefile target
// `stages`): 1 stops each tile after the syndrome, 2 after Berlekamp-Massey, 3 after the FFT
// coefficient phase; 0 (the library) runs the whole decryption.  Truncated builds write no
// error vectors.
#ifndef CMD_STAGE_STOP
#define CMD_STAGE_STOP 0
#endif

constexpr int kRowWords = 16;                          // LDS words per syndrome plane row (m used)
constexpr int kWaveLds = (2 * kMaxT + 1) * kRowWords;  // words of LDS per wavefront

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp0(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}

// XOR over the 64 lanes, returned as a wave-uniform value.
__device__ __forceinline__ uint32_t wave_xor_all(uint32_t v)
{
    v ^= dpp0<0xB1>(v);   // quad_perm [1,0,3,2]
    v ^= dpp0<0x4E>(v);   // quad_perm [2,3,0,1]
    v ^= dpp0<0x124>(v);  // row_ror:4
    v ^= dpp0<0x128>(v);  // row_ror:8
    return (uint32_t)(__builtin_amdgcn_readlane((int)v, 0) ^ __builtin_amdgcn_readlane((int)v, 16) ^
                      __builtin_amdgcn_readlane((int)v, 32) ^ __builtin_amdgcn_readlane((int)v, 48));
}

// m ? a : b per bit, for a lane mask m (all-ones or zero per lane) held in a vector
// register.  A select written on a lane condition can be compiled as a branch under an
// execution mask, and a DPP move inside such a branch reads its disabled source lanes as
// zero; this form always runs as one instruction on every lane.
__device__ __forceinline__ uint32_t sel_mask(uint32_t m, uint32_t a, uint32_t b)
{
    return __builtin_amdgcn_bitop3_b32(m, a, b, 0xCA);
}

// XOR over the 64 lanes of each of N <= 16 values at once.  Each step folds one lane bit of
// two registers into one: the permlane swaps exchange the halves (lane bit 5) or the odd and
// even 16-lane rows (bit 4) of a register pair, so one XOR leaves the first value's partial
// sums where that bit is 0 and the second's where it is 1; bits 2 and 3 do the same with a
// lane-dependent select and one row rotation.  After the two in-quad steps every lane of quad
// wave_xor_lane(k) / 4 holds the sum of value k.  `extra` is XORed into that register before
// the in-quad steps (a term that a value's quad should add once).  About 3.5 instructions per
// value against 11 for separate reductions.
__device__ __forceinline__ constexpr int wave_xor_lane(int k)
{
    return ((k & 1) << 5) | (((k >> 1) & 1) << 4) | (((k >> 3) & 1) << 3) | (((k >> 2) & 1) << 2);
}

template <int N>
__device__ __forceinline__ void wave_xor_many(const uint32_t (&v)[N], uint32_t extra, uint32_t m2, uint32_t m3,
                                              uint32_t (&out)[N])
{
    static_assert(N >= 1 && N <= 16, "up to 16 values");
    constexpr int N5 = (N + 1) / 2, N4 = (N5 + 1) / 2, N2 = (N4 + 1) / 2;
    uint32_t r5[N5], r4[N4], r2[N2];
#pragma unroll
    for (int i = 0; i < N5; i++) {
        const uint32_t a = v[2 * i], b = 2 * i + 1 < N ? v[2 * i + 1] : v[2 * i];
        const auto s = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        r5[i] = s[0] ^ s[1];
    }
#pragma unroll
    for (int i = 0; i < N4; i++) {
        const uint32_t a = r5[2 * i], b = 2 * i + 1 < N5 ? r5[2 * i + 1] : r5[2 * i];
        const auto s = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        r4[i] = s[0] ^ s[1];
    }
#pragma unroll
    for (int i = 0; i < N2; i++) {
        // lanes with bit 2 clear keep a and pair with a lane 4 away (bit 2 set) that sends its a
        const uint32_t a = r4[2 * i], b = 2 * i + 1 < N4 ? r4[2 * i + 1] : r4[2 * i];
        const uint32_t keep = sel_mask(m2, b, a), send = sel_mask(m2, a, b);
        r2[i] = keep ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x124 /* row_ror:4 */, 0xF, 0xF, true);
    }
    uint32_t r;
    {
        const uint32_t a = r2[0], b = N2 > 1 ? r2[N2 - 1] : r2[0];
        const uint32_t keep = sel_mask(m3, b, a), send = sel_mask(m3, a, b);
        r = keep ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x128 /* row_ror:8 */, 0xF, 0xF, true);
    }
    r ^= extra;
    r ^= dpp0<0xB1>(r);  // quad_perm [1,0,3,2]
    r ^= dpp0<0x4E>(r);  // quad_perm [2,3,0,1]
#pragma unroll
    for (int k = 0; k < N; k++) out[k] = (uint32_t)__builtin_amdgcn_readlane((int)r, wave_xor_lane(k));
}

// Lane l receives v from lane l-1; lane 0 receives `carry`.
__device__ __forceinline__ uint32_t shift_up1(uint32_t v, uint32_t carry)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)carry, (int)v, 0x138 /* wave_shr:1 */, 0xF, 0xF, false);
}

// Lane l receives v from lane l-1; lane 0 receives 0 (one DPP move).
__device__ __forceinline__ uint32_t shift_up1_zero(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138 /* wave_shr:1 */, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Key tables and masks live in device global memory; saying so lets loads use a scalar base
// plus a 32-bit lane offset instead of a 64-bit address per lane (flat).
using gu32 = const __attribute__((address_space(1))) uint32_t;
__device__ __forceinline__ gu32 *gptr(const uint32_t *p) { return (gu32 *)p; }

// The same value, hidden from the optimiser.  Values derived from the lane id inside a loop
// would otherwise be hoisted to the kernel entry, where hundreds of them no longer fit in
// registers and are spilled; deriving them from an opaque copy keeps them where they are used.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Bit-sliced "L <= K" over 9-bit values L (planes Lp), K a public constant.
__device__ __forceinline__ uint32_t bs_le_const(const uint32_t Lp[9], uint32_t K)
{
    uint32_t lt = 0, eq = ~0u;
#pragma unroll
    for (int q = 8; q >= 0; q--) {
        if ((K >> q) & 1u) {
            lt |= eq & ~Lp[q];
            eq &= Lp[q];
        } else {
            eq &= ~Lp[q];
        }
    }
    return lt | eq;
}

// Bit-sliced V - L for a public constant V (9 bits, modulo 512).
__device__ __forceinline__ void bs_const_sub(uint32_t V, const uint32_t Lp[9], uint32_t out[9])
{
    uint32_t borrow = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const uint32_t l = Lp[q];
        if ((V >> q) & 1u) {
            out[q] = ~(l ^ borrow);
            borrow = l & borrow;
        } else {
            out[q] = l ^ borrow;
            borrow = l | borrow;
        }
    }
}


// Value of lane (l ^ OFF), OFF a power of two below 64.
template <int OFF>
__device__ __forceinline__ uint32_t xor_lanes(uint32_t v)
{
    if constexpr (OFF == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    } else if constexpr (OFF == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    } else if constexpr (OFF == 4) {
        const int y = __builtin_amdgcn_mov_dpp((int)v, 0x1B, 0xF, 0xF, true);     // l ^ 3
        return (uint32_t)__builtin_amdgcn_mov_dpp(y, 0x141, 0xF, 0xF, true);     // row_half_mirror: l ^ 7
    } else if constexpr (OFF == 8) {
        const int y = __builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true);    // row_mirror: l ^ 15
        return (uint32_t)__builtin_amdgcn_mov_dpp(y, 0x141, 0xF, 0xF, true);     // then l ^ 7
    } else if constexpr (OFF == 16) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (16 << 10) | 0x1F);  // bitmask mode, xor 16
    } else {
        static_assert(OFF == 32, "lane offset");
        const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, (int)v);
    }
}

// Value of lane (l ^ OFF) for any OFF < 64 (composed from the power-of-two exchanges).
template <int OFF>
__device__ __forceinline__ uint32_t xor_lanes_any(uint32_t v)
{
    if constexpr (OFF == 0) {
        return v;
    } else if constexpr (OFF == 3) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x1B, 0xF, 0xF, true);  // quad_perm [3,2,1,0]
    } else {
        constexpr int low = OFF & -OFF;
        return xor_lanes<low>(xor_lanes_any<OFF - low>(v));
    }
}

// a <- a * c for a per-lane constant c (bit-sliced over 32 elements sharing c)
template <int M>
__device__ __forceinline__ void bs_mul_scalar(uint32_t a[M], uint32_t c)
{
    uint32_t cm[M];
#pragma unroll
    for (int k = 0; k < M; k++) cm[k] = bitmask(c, k);
    bs_mul<M>(a, cm, a);
}

__device__ __forceinline__ uint32_t bitrev_bits(uint32_t v, int bits)
{
    return __builtin_bitreverse32(v) >> (32 - bits);
}

// Index layout of the bit-sliced additive FFT over the 2^M field elements (32 evaluation
// points per word) with 2^LOG leaves, and its linear steps together with their transposes.
// The forward direction evaluates a polynomial of 2^LOG coefficients at every point; the
// transposed direction maps values at the points to the power sums sum_c v_c a_c^p, p < 2^LOG,
// by the same steps in reverse order, each replaced by its transpose (a conditional swap and a
// multiplication by a constant are their own transposes; "target ^= source" becomes
// "source ^= target").
template <int M, int LOG>
struct FftPlan {
    static constexpr int NP = (1 << LOG) / 64;         // positions per lane in the coefficient phase
    static constexpr int NW = 1 << (M - 5);            // 32-point words per ciphertext
    static constexpr int WPL = NW / 64;                // words per lane in the butterfly phase
    static constexpr int WPLEAF = 1 << (M - 5 - LOG);  // words covered by one leaf

    // ---- coefficient phase (bit-sliced across the tile, lane = coefficient position) ---------

    template <int D, int LANEX, int PBX>
    static __device__ __forceinline__ void radix_pass(uint32_t (&F)[NP][M], int lane, int s, int sel)
    {
        // XOR into targets (bits s-1, s-2 of the node-local index == sel) the values at p ^ (LANEX + 64 PBX).
        // No source is a target of the same pass, so every exchange is issued before the first
        // XOR: the lane moves through LDS (lane distances 16 and 32) are then in flight together
        // instead of one round trip each.
        uint32_t X[NP][M];
#pragma unroll
        for (int pb = 0; pb < NP; pb++) {
            if ((pb ^ PBX) >= NP) continue;
#pragma unroll
            for (int k = 0; k < M; k++) X[pb][k] = xor_lanes_any<LANEX>(F[pb ^ PBX][k]);
        }
#pragma unroll
        for (int pb = 0; pb < NP; pb++) {
            const int p = lane + 64 * pb;
            const int i = p >> D;
            const uint32_t hit = (((i >> (s - 2)) & 3) == sel) ? ~0u : 0u;
            if ((pb ^ PBX) >= NP) continue;
#pragma unroll
            for (int k = 0; k < M; k++) F[pb][k] ^= X[pb][k] & hit;
        }
    }

    // Radix sub-levels S .. 2 of depth D: pass 1 targets node-local bits (S-1, S-2) = (1,0) from
    // distance q, pass 2 targets (0,1) from distance 3q; q = 2^(S-2) node-local = 2^(S-2+D) positions.
    template <int D, int S>
    static __device__ __forceinline__ void radix_levels(uint32_t (&F)[NP][M], int lane)
    {
        if constexpr (S >= 2) {
            constexpr int Q = 1 << (S - 2 + D);
            radix_pass<D, (Q & 63), (Q >> 6)>(F, lane, S, 2);
            radix_pass<D, ((3 * Q) & 63), ((3 * Q) >> 6)>(F, lane, S, 1);
            radix_levels<D, S - 1>(F, lane);
        }
    }

    template <int D>
    static __device__ __forceinline__ void coeff_depth(uint32_t (&F)[NP][M], int lane, const FftTab *tab)
    {
        if constexpr (D > 0) {
#pragma unroll
            for (int pb = 0; pb < NP; pb++) bs_mul_scalar<M>(F[pb], tab->tw[D][lane + 64 * pb]);
        }
        radix_levels<D, LOG - D>(F, lane);
    }

    // Transpose of radix_pass: the values at the targets (node-local bits (s-1, s-2) == sel) are
    // XORed into the positions p ^ X they were read from.
    template <int D, int LANEX, int PBX>
    static __device__ __forceinline__ void radix_pass_tr(uint32_t (&F)[NP][M], int lane, int s, int sel)
    {
        constexpr int X = LANEX + 64 * PBX;
#pragma unroll
        for (int pb = 0; pb < NP; pb++) {
            const int p = lane + 64 * pb;
            const int i = (p ^ X) >> D;
            const uint32_t hit = (((i >> (s - 2)) & 3) == sel) ? ~0u : 0u;
            if ((pb ^ PBX) >= NP) continue;
#pragma unroll
            for (int k = 0; k < M; k++) F[pb][k] ^= xor_lanes_any<LANEX>(F[pb ^ PBX][k]) & hit;
        }
    }

    // Transpose of radix_levels<D, S>: sub-levels 2 .. S, each pass 2 before pass 1.
    template <int D, int S>
    static __device__ __forceinline__ void radix_levels_tr(uint32_t (&F)[NP][M], int lane)
    {
        if constexpr (S >= 2) {
            radix_levels_tr<D, S - 1>(F, lane);
            constexpr int Q = 1 << (S - 2 + D);
            radix_pass_tr<D, ((3 * Q) & 63), ((3 * Q) >> 6)>(F, lane, S, 1);
            radix_pass_tr<D, (Q & 63), (Q >> 6)>(F, lane, S, 2);
        }
    }

    template <int D>
    static __device__ __forceinline__ void coeff_depth_tr(uint32_t (&F)[NP][M], int lane, const FftTab *tab)
    {
        radix_levels_tr<D, LOG - D>(F, lane);
        if constexpr (D > 0) {
#pragma unroll
            for (int pb = 0; pb < NP; pb++) bs_mul_scalar<M>(F[pb], tab->tw[D][lane + 64 * pb]);
        }
    }


    // ---- phase B data layout --------------------------------------------------------------
    // Word w (32 evaluation points) of a ciphertext lives in lane l, register r.  Its index bits
    // start as w = l + 64 r.  Butterfly level J pairs words that differ in index bit J; when that
    // bit is a lane bit, it is first exchanged with a register bit (a 2 x 2 transpose between
    // lane pairs), so every butterfly runs inside one lane on half the words: no product is
    // computed twice.  The exchange schedule is fixed at compile time: when bit J is in a lane,
    // it replaces the register bit whose index is needed furthest in the future (or never).
    static constexpr int RBITS = WPL == 4 ? 2 : 1;      // register bits of the word index
    static constexpr int NLEV = LOG;                    // butterfly levels J = JMIN .. JMIN+LOG-1
    static constexpr int JMIN = M - 5 - LOG;
    struct Layout {
        int posr[2];   // word-index bit held by register bit q
        int posl[6];   // word-index bit held by lane bit i
    };
    struct Schedule {
        int swap_lane[8];   // lane bit exchanged with register bit swap_q[i] before level i (-1: none)
        int swap_q[8];
        int bq[8];          // register bit holding bit J at level i
        Layout at[8];       // layout during level i
        Layout fin;         // layout after the last level
    };
    static constexpr Schedule make_schedule()
    {
        Schedule sc{};
        Layout cur{};
        for (int q = 0; q < 2; q++) cur.posr[q] = q < RBITS ? 6 + q : -1;
        for (int i = 0; i < 6; i++) cur.posl[i] = i;
        for (int i = 0; i < NLEV; i++) {
            const int J = JMIN + i;
            int q_hold = -1;
            for (int q = 0; q < RBITS; q++)
                if (cur.posr[q] == J) q_hold = q;
            sc.swap_lane[i] = -1;
            sc.swap_q[i] = -1;
            if (q_hold < 0) {
                int lb = 0;
                for (int b = 0; b < 6; b++)
                    if (cur.posl[b] == J) lb = b;
                int best = 0, key = -1;
                for (int q = 0; q < RBITS; q++) {
                    const int k = cur.posr[q] < J ? 1000 : cur.posr[q];
                    if (k > key) { key = k; best = q; }
                }
                sc.swap_lane[i] = lb;
                sc.swap_q[i] = best;
                cur.posl[lb] = cur.posr[best];
                cur.posr[best] = J;
                q_hold = best;
            }
            sc.bq[i] = q_hold;
            sc.at[i] = cur;
        }
        sc.fin = cur;
        return sc;
    }
    static constexpr Schedule SCHED = make_schedule();

    // Index bits of lane l's words that come from the lane: bits of l placed where the layout
    // says, restricted to word-index bits below LIMIT.
    static __device__ __forceinline__ uint32_t lane_part(const Layout &lay, int lane, int limit)
    {
        uint32_t w = 0;
#pragma unroll
        for (int i = 0; i < 6; i++)
            if (lay.posl[i] < limit) w |= (uint32_t)((lane >> i) & 1) << lay.posl[i];
        return w;
    }
    static constexpr uint32_t reg_part(const Layout &lay, int r, int limit)
    {
        uint32_t w = 0;
        for (int q = 0; q < RBITS; q++)
            if (lay.posr[q] < limit) w |= (uint32_t)((r >> q) & 1) << lay.posr[q];
        return w;
    }

    // Exchanges lane bit LB with register bit Q: lane l keeps the words whose register bit Q
    // equals its lane bit LB and trades the others with lane l ^ 2^LB.
    // Lane bits 5 and 4 are one permlane swap per word pair (the swap exchanges exactly the
    // halves / 16-lane rows that change owner); bits 3 and 2 two row rotations that write
    // only the 4-lane banks that receive; bits 1 and 0 a quad permutation and selects.
    template <int LB, int Q>
    static __device__ __forceinline__ void swap_bits(uint32_t (&W)[WPL][M], int lane)
    {
        [[maybe_unused]] const uint32_t upper = ((opaque(lane) >> LB) & 1) ? ~0u : 0u;
#pragma unroll
        for (int r0 = 0; r0 < WPL; r0++) {
            if ((r0 >> Q) & 1) continue;
            const int r1 = r0 | (1 << Q);
#pragma unroll
            for (int k = 0; k < M; k++) {
                const uint32_t a = W[r0][k], b = W[r1][k];
                if constexpr (LB == 5) {
                    const auto s = __builtin_amdgcn_permlane32_swap(a, b, false, false);
                    W[r0][k] = s[0];
                    W[r1][k] = s[1];
                } else if constexpr (LB == 4) {
                    const auto s = __builtin_amdgcn_permlane16_swap(a, b, false, false);
                    W[r0][k] = s[0];
                    W[r1][k] = s[1];
                } else if constexpr (LB == 3) {
                    // banks 2, 3 take b from 8 lanes down; banks 0, 1 take a from 8 lanes up
                    W[r0][k] = (uint32_t)__builtin_amdgcn_update_dpp((int)a, (int)b, 0x128 /* row_ror:8 */, 0xF, 0xC, false);
                    W[r1][k] = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)a, 0x128 /* row_ror:8 */, 0xF, 0x3, false);
                } else if constexpr (LB == 2) {
                    // banks 1, 3 take b from 4 lanes down; banks 0, 2 take a from 4 lanes up
                    W[r0][k] = (uint32_t)__builtin_amdgcn_update_dpp((int)a, (int)b, 0x124 /* row_ror:4 */, 0xF, 0xA, false);
                    W[r1][k] = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)a, 0x12C /* row_ror:12 */, 0xF, 0x5, false);
                } else {
                    const uint32_t send = sel_mask(upper, a, b);
                    const uint32_t recv = xor_lanes<(1 << LB)>(send);
                    W[r0][k] = sel_mask(upper, recv, a);
                    W[r1][k] = sel_mask(upper, b, recv);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // The m bit planes of the butterfly multiplier G of level Dl for word offset i, read
    // precomputed (lowpat ^ hi) from the table in 16-byte loads.
    template <int Dl>
    static __device__ __forceinline__ void load_g(const FftTab *tab, uint32_t i, uint32_t (&G)[M])
    {
        const uint4 *gp = reinterpret_cast<const uint4 *>(tab->gpl[Dl][i]);
#pragma unroll
        for (int q = 0; q < (M + 3) / 4; q++) {
            const uint4 g4 = gp[q];
            const uint32_t gw[4] = {g4.x, g4.y, g4.z, g4.w};
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (4 * q + e < M) G[4 * q + e] = gw[e];
        }
    }

    // Butterfly level Dl (word-index bit J = M - 6 - Dl): w0 = u + G v, w1 = w0 + v, where
    // G = lowpat (by bit position) + hi (by the low J bits of the lower word's index).
    template <int Dl>
    static __device__ __forceinline__ void butterfly(uint32_t (&W)[WPL][M], int lane, const FftTab *tab)
    {
        constexpr int J = M - 6 - Dl;
        constexpr int LI = J - JMIN;
        constexpr int SL = SCHED.swap_lane[LI], SQ = SCHED.swap_q[LI];
        if constexpr (SL >= 0) swap_bits<(SL >= 0 ? SL : 0), (SQ >= 0 ? SQ : 0)>(W, lane);
        constexpr int RB = 1 << SCHED.bq[LI];
        constexpr Layout LAY = SCHED.at[LI];
        // Load this level's multipliers here: hoisted out of the per-ciphertext loop, the
        // multipliers of all levels would not fit in registers and would go to scratch.
        asm volatile("" ::: "memory");
        const uint32_t wl = lane_part(LAY, opaque(lane), J);
#pragma unroll
        for (int r = 0; r < WPL; r++) {
            if (r & RB) continue;
            uint32_t G[M], t[M];
            load_g<Dl>(tab, wl | reg_part(LAY, r, J), G);
            bs_mul<M>(G, W[r | RB], t);
#pragma unroll
            for (int k = 0; k < M; k++) {
                W[r][k] ^= t[k];
                W[r | RB][k] ^= W[r][k];
            }
            // one word pair at a time: interleaving the products would multiply the live
            // registers
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // Multiplier index of the first word pair (register 0) of level Dl.
    template <int Dl>
    static __device__ __forceinline__ uint32_t first_g_index(int lane)
    {
        constexpr int J = M - 6 - Dl;
        constexpr Layout LAY = SCHED.at[J - JMIN];
        return lane_part(LAY, opaque(lane), J);
    }

    // butterfly<Dl> with the multiplier loads one word pair ahead: G holds the multipliers of
    // this level's first pair on entry (loaded by the caller) and those of the next level's
    // first pair on return, so each product runs while the next pair's planes are in flight
    // instead of waiting for its own.
    template <int Dl>
    static __device__ __forceinline__ void butterfly_ahead(uint32_t (&W)[WPL][M], int lane, const FftTab *tab,
                                                           uint32_t (&G)[M])
    {
        constexpr int J = M - 6 - Dl;
        constexpr int LI = J - JMIN;
        constexpr int SL = SCHED.swap_lane[LI], SQ = SCHED.swap_q[LI];
        if constexpr (SL >= 0) swap_bits<(SL >= 0 ? SL : 0), (SQ >= 0 ? SQ : 0)>(W, lane);
        constexpr int RB = 1 << SCHED.bq[LI];
        constexpr Layout LAY = SCHED.at[LI];
        asm volatile("" ::: "memory");
        const uint32_t wl = lane_part(LAY, opaque(lane), J);
#pragma unroll
        for (int r = 0; r < WPL; r++) {
            if (r & RB) continue;
            int rn = -1;  // the next pair of this level
#pragma unroll
            for (int q = WPL - 1; q > r; q--)
                if (!(q & RB)) rn = q;
            uint32_t Gn[M];
            const bool ahead = rn >= 0 || Dl > 0;
            if (rn >= 0)
                load_g<Dl>(tab, wl | reg_part(LAY, rn, J), Gn);
            else if constexpr (Dl > 0)
                load_g<Dl - 1>(tab, first_g_index<Dl - 1>(lane), Gn);
            uint32_t t[M];
            bs_mul<M>(G, W[r | RB], t);
#pragma unroll
            for (int k = 0; k < M; k++) {
                W[r][k] ^= t[k];
                W[r | RB][k] ^= W[r][k];
            }
            if (ahead) {
#pragma unroll
                for (int k = 0; k < M; k++) G[k] = Gn[k];
            }
            // one word pair at a time: interleaving the products would multiply the live
            // registers
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // Full word index of lane l's register r in the final layout.
    static __device__ __forceinline__ uint32_t final_word(uint32_t wl_fin, int r)
    {
        return wl_fin | reg_part(SCHED.fin, r, 64);
    }

    template <int K>
    static __device__ __forceinline__ void benes_masks(gu32 *masks, uint32_t wl_fin, uint32_t (&mw)[WPL])
    {
        gu32 *mrow = masks + K * NW + opaque((int)wl_fin);
#pragma unroll
        for (int r = 0; r < WPL; r++) mw[r] = mrow[reg_part(SCHED.fin, r, 64)];
    }

    template <int K>
    static __device__ __forceinline__ void benes_apply(uint32_t (&z)[WPL], const uint32_t (&mw)[WPL])
    {
        constexpr int S = K < M ? K : 2 * M - 2 - K;
        if constexpr (S < 5) {
            constexpr int SH = 1 << S;
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                const uint32_t t = ((z[r] >> SH) ^ z[r]) & mw[r];
                z[r] ^= t ^ (t << SH);
            }
        } else {
            constexpr int JB = S - 5;  // word-index bit
            constexpr int QB = SCHED.fin.posr[0] == JB ? 0 : (RBITS > 1 && SCHED.fin.posr[1] == JB ? 1 : -1);
            if constexpr (QB >= 0) {
                constexpr int RB = 1 << QB;
#pragma unroll
                for (int r = 0; r < WPL; r++) {
                    if (r & RB) continue;
                    const uint32_t d = (z[r] ^ z[r | RB]) & mw[r];
                    z[r] ^= d;
                    z[r | RB] ^= d;
                }
            } else {
                constexpr int LBB = SCHED.fin.posl[0] == JB ? 0 : SCHED.fin.posl[1] == JB ? 1
                                  : SCHED.fin.posl[2] == JB ? 2 : SCHED.fin.posl[3] == JB ? 3
                                  : SCHED.fin.posl[4] == JB ? 4 : 5;
#pragma unroll
                for (int r = 0; r < WPL; r++) {
                    const uint32_t other = xor_lanes<(1 << LBB)>(z[r]);
                    z[r] ^= (z[r] ^ other) & mw[r];
                }
            }
        }
    }

    template <int K>
    static __device__ __forceinline__ void benes_layer(uint32_t (&z)[WPL], gu32 *masks, int, uint32_t wl_fin)
    {
        uint32_t mw[WPL];
        benes_masks<K>(masks, wl_fin, mw);
        benes_apply<K>(z, mw);
    }

    // Layers K .. 2M-2 in order with three layers' masks in flight: on entry m0, m1, m2 hold
    // the masks of layers K, K+1, K+2 (those that exist); layer K+3's load is issued before
    // layer K is applied, so each load has three layers' work to land in.
    template <int K>
    static __device__ __forceinline__ void benes_from3(uint32_t (&z)[WPL], gu32 *masks, uint32_t wl_fin,
                                                       const uint32_t (&m0)[WPL], const uint32_t (&m1)[WPL],
                                                       const uint32_t (&m2)[WPL])
    {
        constexpr int LAST = 2 * M - 2;
        if constexpr (K + 3 <= LAST) {
            uint32_t mn[WPL];
            benes_masks<K + 3>(masks, wl_fin, mn);
            benes_apply<K>(z, m0);
            benes_from3<K + 1>(z, masks, wl_fin, m1, m2, mn);
        } else {
            benes_apply<K>(z, m0);
            if constexpr (K + 1 <= LAST) benes_from3<K + 1>(z, masks, wl_fin, m1, m2, m2);
        }
    }

    // Transpose of butterfly<Dl>: (a, b) -> (a + b, G (a + b) + b), then the lane/register
    // exchange that preceded the forward level, which undoes itself.
    template <int Dl>
    static __device__ __forceinline__ void butterfly_tr(uint32_t (&W)[WPL][M], int lane, const FftTab *tab)
    {
        constexpr int J = M - 6 - Dl;
        constexpr int LI = J - JMIN;
        constexpr int SL = SCHED.swap_lane[LI], SQ = SCHED.swap_q[LI];
        constexpr int RB = 1 << SCHED.bq[LI];
        constexpr Layout LAY = SCHED.at[LI];
        asm volatile("" ::: "memory");
        const uint32_t wl = lane_part(LAY, opaque(lane), J);
#pragma unroll
        for (int r = 0; r < WPL; r++) {
            if (r & RB) continue;
            uint32_t G[M], t[M];
            load_g<Dl>(tab, wl | reg_part(LAY, r, J), G);
#pragma unroll
            for (int k = 0; k < M; k++) W[r][k] ^= W[r | RB][k];
            bs_mul<M>(G, W[r], t);
#pragma unroll
            for (int k = 0; k < M; k++) W[r | RB][k] ^= t[k];
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (SL >= 0) swap_bits<(SL >= 0 ? SL : 0), (SQ >= 0 ? SQ : 0)>(W, lane);
    }

    // Benes layers last to first: the transpose (and inverse) of benes_all.
    template <int... Ks>
    static __device__ __forceinline__ void benes_all_rev(uint32_t (&z)[WPL], gu32 *masks, int lane,
                                                         uint32_t wl_fin, std::integer_sequence<int, Ks...>)
    {
        (benes_layer<2 * M - 2 - Ks>(z, masks, lane, wl_fin), ...);
    }
};


template <int IDX, int ALG>
struct FusedTileKernel {
    static constexpr int TILE = 32;  // bit-sliced width (ciphertext slots per wavefront tile)
    using Pm = P<IDX>;
    static constexpr int M = Pm::m, T = Pm::t, N = Pm::n, G = Pm::groups;
    static constexpr int NBL = (T + 1 + 63) / 64;        // lane blocks of 64 coefficients
    static constexpr int PAIRS = T;                      // table dwords per row (2t uint16)
    static constexpr int R = (PAIRS + 63) / 64;
    static constexpr int WORDS = (Pm::synd_bytes + 3) / 4;
    static_assert(Pm::sb % 16 == 0, "row groups come in pairs");
    static constexpr int CH = TILE;                      // ciphertexts per syndrome chunk
    static constexpr int CST = (T + 2) & ~1;             // LDS uint16 stride of a coefficient row
    static_assert(TILE * CST * 2 <= kWaveLds * 4, "coefficient rows must fit the wave's LDS");
    static_assert(CH == TILE && CH % 16 == 0, "one syndrome chunk covers the tile");

    // A row group is the eight table rows of one ciphertext byte (sb is a multiple of 8).
    static constexpr int NGROUP = Pm::sb / 8;

    static __device__ __forceinline__ void load_group(gu32 *rows, int g, int lane, uint32_t (&v)[8][R])
    {
        gu32 *base = rows + (size_t)(8 * g) * PAIRS;
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = (lane + 64 * r < PAIRS) ? base[(size_t)q * PAIRS + lane + 64 * r] : 0u;
    }

    // The eight rows of group g into the accumulators of the CH ciphertexts of a run of one
    // key: row 8g+q is added where bit BIT0+q of the ciphertext's word is set (BIT0 = 8 (g mod
    // 4)).  Four rows at a time ("four Russians"): the 16 XOR combinations of rows 4h .. 4h+3
    // are written to LDS (table h, entry c, lane-major; Gray-code order, one XOR each), then
    // every ciphertext adds the entry its nibble selects: one LDS read per nibble and one
    // three-input XOR per word in place of four wave-uniform masks and 4R AND-XORs.  Entry 0
    // of each table is zero, written once per tile.  Ciphertext bits are public, so the
    // data-dependent LDS address leaks nothing.
    static __device__ __forceinline__ void tab_put(uint32_t *tab, int e, int lane, const uint32_t (&v)[R])
    {
        if constexpr (R == 2)
            reinterpret_cast<uint2 *>(tab)[e * 64 + lane] = make_uint2(v[0], v[1]);
        else
            tab[e * 64 + lane] = v[0];
    }
    static __device__ __forceinline__ uint2 tab_get(const uint32_t *tab, uint32_t e, int lane)
    {
        if constexpr (R == 2)
            return reinterpret_cast<const uint2 *>(tab)[e * 64 + lane];
        else
            return make_uint2(tab[e * 64 + lane], 0u);
    }

    template <int BIT0>
    static __device__ __forceinline__ void add_group(const uint32_t (&v)[8][R], const uint32_t (&cw)[CH],
                                                     uint32_t (&acc)[CH][R], uint32_t *tab, int lane)
    {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t cur[R];
#pragma unroll
            for (int r = 0; r < R; r++) cur[r] = 0;
#pragma unroll
            for (int i = 1; i < 16; i++) {
                const int q = 4 * h + __builtin_ctz(i);
#pragma unroll
                for (int r = 0; r < R; r++) cur[r] ^= v[q][r];
                tab_put(tab, 16 * h + (i ^ (i >> 1)), lane, cur);
            }
        }
        // the entries of ciphertext x + PD are read while x adds its own, so PD lookups are in
        // flight instead of one LDS round trip per ciphertext
        constexpr int PD = 4;
        uint2 lo[PD], hi[PD];
#pragma unroll
        for (int x = 0; x < PD; x++) {
            lo[x] = tab_get(tab, (cw[x] >> BIT0) & 15u, lane);
            hi[x] = tab_get(tab, 16u + ((cw[x] >> (BIT0 + 4)) & 15u), lane);
        }
#pragma unroll
        for (int x = 0; x < CH; x++) {
            const uint2 l = lo[x % PD], h = hi[x % PD];
            if (x + PD < CH) {
                lo[x % PD] = tab_get(tab, (cw[x + PD] >> BIT0) & 15u, lane);
                hi[x % PD] = tab_get(tab, 16u + ((cw[x + PD] >> (BIT0 + 4)) & 15u), lane);
            }
            acc[x][0] = xor3(acc[x][0], l.x, h.x);
            if constexpr (R == 2) acc[x][R - 1] = xor3(acc[x][R - 1], l.y, h.y);
            // pin the sum here: left free, the compiler defers these XORs and keeps the
            // loaded entries of many groups live, which overflows the register file
            asm volatile("" : "+v"(acc[x][0]), "+v"(acc[x][R - 1]));
        }
    }
    static_assert(32 * 64 * R <= kWaveLds, "four-row tables must fit the wave's LDS");

    // Bytes 4w .. 4w+3 of a ciphertext (those that exist) as a little-endian word; one dword
    // load when the rows are 4-byte aligned (al) and the word is whole.
    static __device__ __forceinline__ uint32_t ct_word(const uint8_t *c, int w, bool al)
    {
        uint32_t v = 0;
        if (c) {
            if (al && 4 * w + 4 <= Pm::synd_bytes) return *reinterpret_cast<const uint32_t *>(c + 4 * w);
#pragma unroll
            for (int b = 0; b < 4; b++)
                if (4 * w + b < Pm::synd_bytes) v |= (uint32_t)c[4 * w + b] << (8 * b);
        }
        return v;
    }
    static_assert(WORDS <= 64, "a ciphertext's words fit one per lane");

    static __device__ __forceinline__ bool ct_aligned(const FusedArgs &a)
    {
        return ((a.ct_stride & 3u) == 0) && ((reinterpret_cast<uintptr_t>(a.ct) & 3u) == 0);
    }

    // ---- stage 1: syndromes into LDS bit planes ---------------------------------------------
    // A tile's ciphertexts fall into runs of one key (the planner orders items by key).  Each
    // run costs one streamed pass over its key's whole table, shared by the run's ciphertexts;
    // a tile of many short runs (mixed serving batches) instead reads, ciphertext by
    // ciphertext, only the rows of set bits.
    static __device__ __forceinline__ void syndrome(const FusedArgs &a, const FusedTile &tl, uint32_t *lds, int lane)
    {
        const bool al = ct_aligned(a);
        const int chunks = (tl.count + CH - 1) / CH;
        for (int c = 0; c < chunks; c++) {
            const int x0 = c * CH;
            const bool live = lane < CH && x0 + lane < tl.count;
            const int my_key = live ? a.item_key[tl.first + x0 + lane] : -1;
            const uint8_t *ct_ptr = live ? a.ct + (size_t)a.items[tl.first + x0 + lane] * a.ct_stride : nullptr;
            const int n_live = tl.count - x0 < CH ? tl.count - x0 : CH;
            // distinct keys of the chunk (wave-uniform)
            int runs = 0;
            for (uint64_t rem = __ballot(live); rem; runs++) {
                const int k = __builtin_amdgcn_readlane(my_key, __builtin_ctzll(rem));
                rem &= ~__ballot(my_key == k);
            }

            uint32_t acc[CH][R];
#pragma unroll
            for (int x = 0; x < CH; x++)
#pragma unroll
                for (int r = 0; r < R; r++) acc[x][r] = 0;

            if (tl.count > SCALAR_BM && 2 * runs <= n_live) {
                // entry 0 of both four-row tables
                {
                    const uint32_t zero[R] = {};
                    tab_put(lds, 0, lane, zero);
                    tab_put(lds, 16, lane, zero);
                }
                for (uint64_t rem = __ballot(live); rem;) {
                    const int key = __builtin_amdgcn_readlane(my_key, __builtin_ctzll(rem));
                    const bool in_run = my_key == key;
                    rem &= ~__ballot(in_run);
                    // ciphertexts outside the run read as zero words and add nothing
                    const uint8_t *my_ct = in_run ? ct_ptr : nullptr;
                    // Row groups are software-pipelined through two register buffers: the loads of
                    // the next group are in flight while the current one is added, hiding the L2
                    // latency.  One iteration is one ciphertext word (four groups); 6960119's last
                    // word has two.
                    const KeyDev *kd = &a.keys[key];
                    gu32 *rows = gptr(reinterpret_cast<const uint32_t *>(kd->table));
                    uint32_t va[8][R], vb[8][R], cw[CH];
                    load_group(rows, 0, lane, va);
                    uint32_t next = ct_word(my_ct, 0, al);  // the next word's load is in flight
                    for (int w = 0; w < WORDS; w++) {
                        const uint32_t mine = next;
                        if (w + 1 < WORDS) next = ct_word(my_ct, w + 1, al);
#pragma unroll
                        for (int x = 0; x < CH; x++) cw[x] = (uint32_t)__builtin_amdgcn_readlane((int)mine, x);
                        const int g = 4 * w;
                        if (g + 1 < NGROUP) load_group(rows, g + 1, lane, vb);
                        add_group<0>(va, cw, acc, lds, lane);
                        if (g + 2 < NGROUP) load_group(rows, g + 2, lane, va);
                        if (g + 1 < NGROUP) add_group<8>(vb, cw, acc, lds, lane);
                        if (g + 2 >= NGROUP) break;
                        load_group(rows, g + 3, lane, vb);
                        add_group<16>(va, cw, acc, lds, lane);
                        if (g + 4 < NGROUP) load_group(rows, g + 4, lane, va);
                        add_group<24>(vb, cw, acc, lds, lane);
                    }
                }
            } else {
                // one key per ciphertext: each row feeds one accumulator.  The ciphertext is
                // public, so only the rows of its set bits are read (a wave-uniform bit scan),
                // GRP rows at a time to keep loads in flight.
                constexpr int GRP = 16;
#pragma unroll
                for (int x = 0; x < CH; x++) {
                    if (x0 + x >= tl.count) continue;
                    const KeyDev *kd = &a.keys[a.item_key[tl.first + x0 + x]];
                    gu32 *rows = gptr(reinterpret_cast<const uint32_t *>(kd->table));
                    const uint8_t *ctp = a.ct + (size_t)a.items[tl.first + x0 + x] * a.ct_stride;
                    // the whole ciphertext in one load, word w in lane w
                    const uint32_t ctw = lane < WORDS ? ct_word(ctp, lane, al) : 0u;
                    for (int w = 0; w < WORDS; w++) {
                        uint32_t bits = (uint32_t)__builtin_amdgcn_readlane((int)ctw, w);
                        gu32 *wrows = rows + (size_t)(w * 32) * PAIRS;
                        while (bits) {
                            uint32_t v[GRP][R];
#pragma unroll
                            for (int j = 0; j < GRP; j++) {
                                const bool any = bits != 0;
                                const int bit = any ? __builtin_ctz(bits) : 0;
                                bits &= bits - 1;
#pragma unroll
                                for (int r = 0; r < R; r++)
                                    v[j][r] = (any && lane + 64 * r < PAIRS) ? wrows[(size_t)bit * PAIRS + lane + 64 * r] : 0u;
                            }
#pragma unroll
                            for (int j = 0; j < GRP; j++)
#pragma unroll
                                for (int r = 0; r < R; r++) acc[x][r] ^= v[j][r];
                        }
                    }
                }
            }

            // row 0: the zero row read for s_{negative} (after the four-row tables are done with)
            if (lane < kRowWords) lds[lane] = 0;
            // transpose: plane (j, k) gets bit k of s_j of chunk ciphertext x at bit x0 + x
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int cp = lane + 64 * r;
                if (cp >= PAIRS) continue;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int j = 2 * cp + h;
#pragma unroll
                    for (int k = 0; k < M; k++) {
                        uint32_t plane = 0;
#pragma unroll
                        for (int x = 0; x < CH; x++) plane |= ((acc[x][r] >> (16 * h + k)) & 1u) << x;
                        lds[(j + 1) * kRowWords + k] = plane;
                    }
                }
            }
        }
    }

    // ---- stage 2: bit-sliced Berlekamp-Massey; leaves coefficient rows in LDS ----------------
    // Lane l holds coefficient i = l + 64b of C and B (b < NBV).  Before iteration n's update C
    // has degree <= n and B degree <= n+1 (B = x^(n+1) until the first nonzero discrepancy), so
    // block b joins the products once 64b <= n+1 and the shift of B once 64b <= n+2 (public
    // bounds).
    //
    // When t is a multiple of 64 the top coefficient t would occupy a lane block of its own,
    // one live lane doing a whole block's products.  It is held transposed instead: lane k
    // (k < m) holds plane k of C_t (ct) and of B_t (bt), and its products with the wave-uniform
    // field elements b, d and s_{n-t} are Horner evaluations over those planes, one shift up the
    // lanes (multiplication by x, the overflowing plane folded back by x^m = low(x)) and one
    // AND-XOR per plane of the scalar.  Lanes >= m hold values that only move upwards and are
    // never read.
    static constexpr bool TOPX = (T % 64) == 0;
    static constexpr int NBV = TOPX ? T / 64 : NBL;  // lane blocks of coefficients

    // x * v for a transposed element v (planes in lanes 0..M-1); pm has the planes of low(x).
    static __device__ __forceinline__ uint32_t tx_mulx(uint32_t v, uint32_t pm)
    {
        const uint32_t top = (uint32_t)__builtin_amdgcn_readlane((int)v, M - 1);
        return and_xor(shift_up1_zero(v), pm, top);
    }
    // The same with plane M-1 broadcast by DPP (row_newbcast: lane M-1 of each 16-lane row to
    // its row; pm is zero outside row 0), so the value never leaves the vector registers.
    static __device__ __forceinline__ uint32_t tx_mulx_v(uint32_t v, uint32_t pm)
    {
        const uint32_t top = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + (M - 1), 0xF, 0xF, true);
        return and_xor(shift_up1_zero(v), pm, top);
    }

    // Iteration n of the bit-sliced recurrence.  FULL: every lane block takes part and the
    // transposed top coefficient is live (n >= NFULL, a public bound), so the step is one
    // straight-line block the scheduler can interleave; otherwise blocks join by the bounds
    // above and the top coefficient is still zero.
    static constexpr int NFULL = TOPX ? T - 2 : 64 * (NBV - 1) - 1;
    static_assert(NFULL >= 0, "full iterations start at a valid index");

    // Block b takes part in the products of iteration n once 64b <= n+1 and in the shift once
    // 64b <= n+2; inside a phase [LO, HI) of the loop these are decided at compile time where
    // the phase bounds allow.
    template <int LO, int HI>
    static __device__ __forceinline__ bool mac_on(int b, int n)
    {
        return 64 * b <= LO + 1 ? true : (64 * b > HI ? false : 64 * b <= n + 1);
    }
    template <int LO, int HI>
    static __device__ __forceinline__ bool shift_on(int b, int n)
    {
        return 64 * b <= LO + 2 ? true : (64 * b > HI + 1 ? false : 64 * b <= n + 2);
    }

    // Bit-sliced D = n - 2L (two's complement, DBITS planes): the reference's test 2L <= n is
    // D >= 0, and its length change L' = n + 1 - L turns D into -D - 1 = ~D, while an unchanged
    // L makes it D + 1; so one sign plane and a select per plane replace the comparison and the
    // subtraction on L.  |D| <= 2t + 2 < 2^(DBITS-1).
    static constexpr int DBITS = 10;
    static_assert(2 * T + 2 < (1 << (DBITS - 1)), "D fits its planes");

    // Per-lane constants of the recurrence, computed once per tile.
    struct BmLanes {
        uint32_t keep_last;  // coefficients above T (last block) are dropped
        uint32_t pm;         // planes of low(x) in lanes 0..M-1 (transposed top coefficient)
        uint32_t top_addr;   // ds_bpermute source of plane k of the top term: lane k ...
        uint32_t top_keep;   // ... for the lane wave_xor_lane(k) that reduces plane k
        uint32_t m2, m3;     // lane bits 2 and 3 as masks (reduction selects)
    };

    // Iteration n of the bit-sliced recurrence.  The wave-uniform per-ciphertext state (b, D
    // and the selection mask) stays in vector registers, where a select is one three-input
    // operation instead of three scalar ones; the discrepancy planes come out of the reduction
    // in scalar registers.
    template <int LO, int HI>
    static __device__ __forceinline__ void bm_step(int n, const uint4 *rows4, int lane, uint32_t (&C)[NBL][M],
                                                   uint32_t (&B)[NBV][M], uint32_t (&bb)[M], uint32_t (&Dp)[DBITS],
                                                   uint32_t &ct, uint32_t &bt, const BmLanes &ln)
    {
        constexpr bool FULL = LO >= NFULL;
        constexpr int KL = Kara<M>::L, NQ0 = Kara<M>::NQ0, NQ2 = Kara<M>::NQ2;
        // discrepancy d = sum_i C_i s_{n-i} (rows hold s_{j} at j+1, row 0 is zero), its
        // products accumulated as Karatsuba half products and recombined once
        uint32_t q0[NQ0], q1[NQ0], q2[NQ2];
#pragma unroll
        for (int s2 = 0; s2 < NQ0; s2++) q0[s2] = q1[s2] = 0;
#pragma unroll
        for (int s2 = 0; s2 < NQ2; s2++) q2[s2] = 0;
#pragma unroll
        for (int b = 0; b < NBV; b++) {
            if (!mac_on<LO, HI>(b, n)) continue;
            const int i = lane + 64 * b;
            const int row = (n - i + 1) > 0 ? (n - i + 1) : 0;
            uint32_t sv[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 w = rows4[row * 4 + q];
                sv[4 * q] = w.x; sv[4 * q + 1] = w.y; sv[4 * q + 2] = w.z; sv[4 * q + 3] = w.w;
            }
            uint32_t cs[KL], ss[KL];
            kara_half_sum<M>(C[b], cs);
            kara_half_sum<M>(sv, ss);
            kara_acc<M>(C[b], sv, cs, ss, q0, q1, q2);
        }
        uint32_t p[2 * M - 1];
        kara_combine<M>(q0, q1, q2, p);
        // + C_t s_{n-t} (TOPX): s_{n-t} (row n-t+1; row 0, zero, while n < t) as wave-uniform
        // planes, Horner over them; plane k of the result is in lane k, and goes to the lane
        // that reduces plane k of the discrepancy
        uint32_t extra = 0;
        if constexpr (TOPX && FULL) {
            const int row = n - T + 1 > 0 ? n - T + 1 : 0;
            uint32_t sv[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 w = rows4[row * 4 + q];
                sv[4 * q] = w.x; sv[4 * q + 1] = w.y; sv[4 * q + 2] = w.z; sv[4 * q + 3] = w.w;
            }
            uint32_t top = and_xor(0u, ct, sv[M - 1]);
#pragma unroll
            for (int a = M - 2; a >= 0; a--) top = and_xor(tx_mulx_v(top, ln.pm), ct, sv[a]);
            extra = (uint32_t)__builtin_amdgcn_ds_bpermute((int)ln.top_addr, (int)top) & ln.top_keep;
        }
        uint32_t part[M], d[M];
        bs_fold<M>(p, part);
        wave_xor_many<M>(part, extra, ln.m2, ln.m3, d);
        uint32_t mne = 0;
#pragma unroll
        for (int k = 0; k < M; k++) mne |= d[k];
        const uint32_t mle = __builtin_amdgcn_bitop3_b32(mne, Dp[DBITS - 1], 0u, 0x30);  // mne & (D >= 0)

        // C' = b*C + d*B ; B' = mle ? C : B ; then B <- x*B
        if constexpr (TOPX && FULL) {
            // C_t' = b C_t + d B_t (B_t' comes from the shift below)
            uint32_t acc = and_xor(and_xor(0u, ct, bb[M - 1]), bt, d[M - 1]);
#pragma unroll
            for (int a = M - 2; a >= 0; a--) acc = and_xor(and_xor(tx_mulx_v(acc, ln.pm), ct, bb[a]), bt, d[a]);
            ct = acc;
        }
        // b*C + d*B per block as Karatsuba half products (the half sums of b and d are
        // wave-uniform, made once), one recombination and one fold per block
        uint32_t bbs[KL], ds[KL];
        kara_half_sum<M>(bb, bbs);
        kara_half_sum<M>(d, ds);
#pragma unroll
        for (int b = NBV - 1; b >= 0; b--) {
            if (!mac_on<LO, HI>(b, n)) continue;
            uint32_t u0[NQ0], u1[NQ0], u2[NQ2], cs[KL], bs[KL];
#pragma unroll
            for (int s2 = 0; s2 < NQ0; s2++) u0[s2] = u1[s2] = 0;
#pragma unroll
            for (int s2 = 0; s2 < NQ2; s2++) u2[s2] = 0;
            kara_half_sum<M>(C[b], cs);
            kara_half_sum<M>(B[b], bs);
            kara_acc<M>(bb, C[b], bbs, cs, u0, u1, u2);
            kara_acc<M>(d, B[b], ds, bs, u0, u1, u2);
            uint32_t cw[2 * M - 1], cn[M];
            kara_combine<M>(u0, u1, u2, cw);
            bs_fold<M>(cw, cn);
#pragma unroll
            for (int k = 0; k < M; k++) {
                B[b][k] = __builtin_amdgcn_bitop3_b32(mle, C[b][k], B[b][k], 0xCA);  // mle ? C : B
                C[b][k] = cn[k];
            }
        }
        // the shift moves coefficient t-1 of the selected B into B_t (TOPX); the old B_t
        // moves past t and is dropped
        if constexpr (TOPX && FULL) {
#pragma unroll
            for (int k = 0; k < M; k++) {
                const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)B[NBV - 1][k], 63);
                bt = lane == k ? c : bt;
            }
        }
        // shift the highest block first so each carry is read before its source moves; B has
        // degree <= n+1 here, so block b can receive a carry once 64b <= n+2.  The carry (lane
        // 63 of the block below) reaches lane 0 by a wave rotation.
#pragma unroll
        for (int b = NBV - 1; b >= 0; b--) {
            if (!shift_on<LO, HI>(b, n)) continue;
#pragma unroll
            for (int k = 0; k < M; k++) {
                if (b == 0) {
                    B[b][k] = shift_up1_zero(B[b][k]);
                } else {
                    const uint32_t carry = (uint32_t)__builtin_amdgcn_update_dpp(
                        0, (int)B[b > 0 ? b - 1 : 0][k], 0x13C /* wave_ror:1 */, 0xF, 0xF, true);
                    B[b][k] = (uint32_t)__builtin_amdgcn_update_dpp((int)carry, (int)B[b][k], 0x138 /* wave_shr:1 */,
                                                                    0xF, 0xF, false);
                }
            }
        }
        if constexpr (!TOPX) {
#pragma unroll
            for (int k = 0; k < M; k++) B[NBV - 1][k] &= ln.keep_last;
        }
        // D' = mle ? ~D : D + 1 (plane 0 flips either way); b' = mle ? d : b
        uint32_t cy = Dp[0];
        Dp[0] = ~Dp[0];
#pragma unroll
        for (int q = 1; q < DBITS; q++) {
            const uint32_t ncy = __builtin_amdgcn_bitop3_b32(Dp[q], cy, 0u, 0xC0);  // D_q & carry
            Dp[q] = __builtin_amdgcn_bitop3_b32(mle, Dp[q], cy, 0x36);              // mle ? ~D_q : D_q ^ carry
            cy = ncy;
        }
#pragma unroll
        for (int k = 0; k < M; k++) bb[k] = __builtin_amdgcn_bitop3_b32(mle, d[k], bb[k], 0xCA);
    }

    static __device__ __forceinline__ void berlekamp_massey(uint32_t *lds, int lane, uint32_t (&C)[NBL][M])
    {
        uint32_t B[NBV][M];
#pragma unroll
        for (int b = 0; b < NBV; b++)
#pragma unroll
            for (int k = 0; k < M; k++) {
                C[b][k] = (b == 0 && k == 0 && lane == 0) ? ~0u : 0u;  // C = 1
                B[b][k] = (b == 0 && k == 0 && lane == 1) ? ~0u : 0u;  // B = x
            }
        uint32_t bb[M], Dp[DBITS];
#pragma unroll
        for (int k = 0; k < M; k++) bb[k] = k == 0 ? ~0u : 0u;  // b = 1
#pragma unroll
        for (int q = 0; q < DBITS; q++) Dp[q] = 0;  // D = 0

        BmLanes ln;
        ln.keep_last = (lane + 64 * (NBV - 1) <= T) ? ~0u : 0u;
        ln.pm = (lane < M && ((Modulus<M>::low >> lane) & 1u)) ? ~0u : 0u;
        ln.top_addr = 0;
        ln.top_keep = 0;
#pragma unroll
        for (int k = 0; k < M; k++)
            if (lane == wave_xor_lane(k)) {
                ln.top_addr = 4u * k;
                ln.top_keep = ~0u;
            }
        ln.m2 = ((lane >> 2) & 1) ? ~0u : 0u;
        ln.m3 = ((lane >> 3) & 1) ? ~0u : 0u;
        uint32_t ct = 0, bt = 0;  // TOPX: planes of C_t and B_t, transposed
        const uint4 *rows4 = reinterpret_cast<const uint4 *>(lds);

        // phases between the indices where a block joins the shift (64b-2) or the products
        // (64b-1); blocks 1 and 2 at most, as t <= 128
        static_assert(NBV <= 2, "phase split covers two lane blocks");
        constexpr int P1 = NBV > 1 ? 62 : NFULL, P2 = NBV > 1 ? 63 : NFULL;
        static_assert(P1 <= P2 && P2 <= NFULL, "phases in order");
#pragma unroll 1
        for (int n = 0; n < P1; n++) bm_step<0, P1>(n, rows4, lane, C, B, bb, Dp, ct, bt, ln);
#pragma unroll 1
        for (int n = P1; n < P2; n++) bm_step<P1, P2>(n, rows4, lane, C, B, bb, Dp, ct, bt, ln);
#pragma unroll 1
        for (int n = P2; n < NFULL; n++) bm_step<P2, NFULL>(n, rows4, lane, C, B, bb, Dp, ct, bt, ln);
#pragma unroll 1
        for (int n = NFULL; n < 2 * T; n++) bm_step<NFULL, 2 * T>(n, rows4, lane, C, B, bb, Dp, ct, bt, ln);
        if constexpr (TOPX) {
            // C_t into lane 0 of the last block
#pragma unroll
            for (int k = 0; k < M; k++) {
                const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)ct, k);
                C[NBL - 1][k] = lane == 0 ? w : 0u;
            }
        }
    }

    // ---- stage 2 for tiles of at most SCALAR_BM ciphertexts --------------------------------
    // The bit-sliced loop costs the same for 1 or 32 ciphertexts.  Small tiles (low-load
    // serving) instead run the same inversion-free recurrence per ciphertext with scalar field
    // elements, lane l holding coefficient l + 64b, and pack the result into the bit-sliced
    // planes the root search reads (bit x of C[b][k] = bit k of ciphertext x's coefficient).
    static constexpr int SCALAR_BM = 3;
    static constexpr int NQ = (2 * T + 63) / 64;  // syndromes per lane

    // One iteration of the per-ciphertext recurrence, inside a phase [LO, HI) of n; the
    // selects on mle (secret) are branch-free.
    template <int LO, int HI>
    static __device__ __forceinline__ void scalar_bm_step(int n, const uint16_t *Sx, int lane, uint32_t keep_last,
                                                          uint32_t (&c)[NBL], uint32_t (&bv)[NBL], uint32_t &bb,
                                                          uint32_t &L, uint32_t (&colb)[M])
    {
        // the lanes' products are summed unreduced; the sum reduces once
        uint32_t part = 0;
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            if (!mac_on<LO, HI>(b, n)) continue;
            const int j = n - (lane + 64 * b);
            const uint32_t sv = j >= 0 ? (uint32_t)Sx[j] : 0u;
            part ^= gf_clmul<M>(c[b], sv);
        }
        const uint32_t d = gf_reduce<M>(uni(wave_xor_all(part)));
        const bool mle = d != 0 && 2 * L <= (uint32_t)n;
        uint32_t cold[M];
        gf_mul_cols<M>(d, cold);
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            if (!mac_on<LO, HI>(b, n)) continue;
            const uint32_t cn = gf_mul_by_cols<M>(colb, c[b]) ^ gf_mul_by_cols<M>(cold, bv[b]);
            bv[b] = mle ? c[b] : bv[b];
            c[b] = cn;
        }
#pragma unroll
        for (int b = NBL - 1; b >= 0; b--) {
            if (!shift_on<LO, HI>(b, n)) continue;
            const uint32_t carry = b == 0 ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)bv[b > 0 ? b - 1 : 0], 63);
            bv[b] = shift_up1(bv[b], carry);
        }
        bv[NBL - 1] &= keep_last;
        // wave-uniform selects (one s_cselect each), no branch on the secret mle
        L = mle ? (uint32_t)n + 1 - L : L;
        bb = mle ? d : bb;
#pragma unroll
        for (int k = 0; k < M; k++) colb[k] = mle ? cold[k] : colb[k];
    }

    // Scalar syndromes of the first SCALAR_BM slots out of the planes, as uint16 rows
    // (2t per slot) at the start of the wave's LDS.
    static __device__ __forceinline__ void scalar_syndrome_rows(uint32_t *lds, int lane)
    {
        uint32_t sx[SCALAR_BM][NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int j = lane + 64 * q;
            uint32_t pl[M];
#pragma unroll
            for (int k = 0; k < M; k++) pl[k] = (j < 2 * T) ? lds[(j + 1) * kRowWords + k] : 0u;
#pragma unroll
            for (int x = 0; x < SCALAR_BM; x++) {
                uint32_t v = 0;
#pragma unroll
                for (int k = 0; k < M; k++) v |= ((pl[k] >> x) & 1u) << k;
                sx[x][q] = v;
            }
        }
        wave_sync();
        uint16_t *S = reinterpret_cast<uint16_t *>(lds);
#pragma unroll
        for (int x = 0; x < SCALAR_BM; x++)
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int j = lane + 64 * q;
                if (j < 2 * T) S[x * 2 * T + j] = (uint16_t)sx[x][q];
            }
        wave_sync();
    }

    static __device__ __forceinline__ void berlekamp_massey_scalar(uint32_t *lds, int lane, int count,
                                                                   uint32_t (&C)[NBL][M])
    {
        scalar_syndrome_rows(lds, lane);
        const uint16_t *S = reinterpret_cast<const uint16_t *>(lds);

#pragma unroll
        for (int b = 0; b < NBL; b++)
#pragma unroll
            for (int k = 0; k < M; k++) C[b][k] = 0;
        const uint32_t keep_last = (lane + 64 * (NBL - 1) <= T) ? ~0u : 0u;
        for (int x = 0; x < count; x++) {
            const uint16_t *Sx = S + x * 2 * T;
            uint32_t c[NBL], bv[NBL];
#pragma unroll
            for (int b = 0; b < NBL; b++) {
                c[b] = (b == 0 && lane == 0) ? 1u : 0u;   // C = 1
                bv[b] = (b == 0 && lane == 1) ? 1u : 0u;  // B = x
            }
            uint32_t bb = 1, L = 0;
            uint32_t colb[M];  // multiplication table of bb, renewed when bb changes
            gf_mul_cols<M>(bb, colb);
            // one loop with run-time block gating: this path serves latency-bound launches of
            // one to three ciphertexts, where a smaller body matters more than branch-free steps
#pragma unroll 1
            for (int n = 0; n < 2 * T; n++) scalar_bm_step<0, 2 * T>(n, Sx, lane, keep_last, c, bv, bb, L, colb);
#pragma unroll
            for (int b = 0; b < NBL; b++)
#pragma unroll
                for (int k = 0; k < M; k++) C[b][k] |= ((c[b] >> k) & 1u) << x;
        }
    }

    // ---- the per-ciphertext recurrence on the two wavefronts of a pair -----------------------
    // Wavefront 0 holds lane block 0 of c and B, wavefront 1 the blocks above it.  Every
    // iteration each wavefront sums its blocks' products and posts the wave-uniform partial sum
    // in an exchange slot, wavefront 0 also its pre-update c and B of lane 63 (the selected B
    // of that lane is what the shift moves into block 1); after one barrier both form the same
    // discrepancy and keep the same wave-uniform state (L, b, its multiplication table).  Slots
    // alternate with the iteration's parity, so a slot is rewritten only after the other
    // wavefront has passed the barrier that follows its read.
    static constexpr int SX_SLOT = 8;  // words per slot: partial sum of wave 0 | of wave 1 | c_63 | B_63
    template <int W>
    static __device__ __forceinline__ void scalar_bm_step_pair(int n, const uint16_t *Sx, uint32_t *xch, int lane,
                                                               uint32_t keep_last, uint32_t (&c)[NBL],
                                                               uint32_t (&bv)[NBL], uint32_t &bb, uint32_t &L,
                                                               uint32_t (&colb)[M])
    {
        uint32_t part = 0;
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            if ((W == 0) != (b == 0) || !mac_on<0, 2 * T>(b, n)) continue;
            const int j = n - (lane + 64 * b);
            const uint32_t sv = j >= 0 ? (uint32_t)Sx[j] : 0u;
            part ^= gf_clmul<M>(c[b], sv);
        }
        const uint32_t pw = uni(wave_xor_all(part));
        uint32_t *slot = xch + (n & 1) * SX_SLOT;
        if (lane == 0) slot[W] = pw;
        if constexpr (W == 0) {
            const uint32_t c63 = (uint32_t)__builtin_amdgcn_readlane((int)c[0], 63);
            const uint32_t b63 = (uint32_t)__builtin_amdgcn_readlane((int)bv[0], 63);
            if (lane == 0) {
                slot[2] = c63;
                slot[3] = b63;
            }
        }
        __syncthreads();
        const uint32_t d = gf_reduce<M>(pw ^ uni(slot[1 - W]));
        [[maybe_unused]] uint32_t c63 = 0, b63 = 0;
        if constexpr (W == 1) {
            c63 = uni(slot[2]);
            b63 = uni(slot[3]);
        }
        const bool mle = d != 0 && 2 * L <= (uint32_t)n;
        uint32_t cold[M];
        gf_mul_cols<M>(d, cold);
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            if ((W == 0) != (b == 0) || !mac_on<0, 2 * T>(b, n)) continue;
            const uint32_t cn = gf_mul_by_cols<M>(colb, c[b]) ^ gf_mul_by_cols<M>(cold, bv[b]);
            bv[b] = mle ? c[b] : bv[b];
            c[b] = cn;
        }
#pragma unroll
        for (int b = NBL - 1; b >= 0; b--) {
            if ((W == 0) != (b == 0) || !shift_on<0, 2 * T>(b, n)) continue;
            uint32_t carry = 0;
            if (b == 1 && W == 1)
                carry = mle ? c63 : b63;  // wavefront 0's lane 63 after the selection
            else if (b > 1)
                carry = (uint32_t)__builtin_amdgcn_readlane((int)bv[b > 0 ? b - 1 : 0], 63);
            bv[b] = shift_up1(bv[b], carry);
        }
        if constexpr (W == 1) bv[NBL - 1] &= keep_last;
        // wave-uniform selects, no branch on the secret mle
        L = mle ? (uint32_t)n + 1 - L : L;
        bb = mle ? d : bb;
#pragma unroll
        for (int k = 0; k < M; k++) colb[k] = mle ? cold[k] : colb[k];
    }

    // berlekamp_massey_scalar on the two wavefronts of a pair: wavefront 0 extracts the scalar
    // syndromes, both run the recurrence, and each writes its own blocks' coefficient planes.
    template <int W>
    static __device__ __forceinline__ void berlekamp_massey_scalar_pair(uint32_t *lds, uint32_t *xch, int lane,
                                                                        int count)
    {
        static_assert(NBL >= 2, "wavefront 1 holds the blocks above block 0");
        if constexpr (W == 0) scalar_syndrome_rows(lds, lane);
        __syncthreads();
        const uint16_t *S = reinterpret_cast<const uint16_t *>(lds);
        uint32_t C[NBL][M];
#pragma unroll
        for (int b = 0; b < NBL; b++)
#pragma unroll
            for (int k = 0; k < M; k++) C[b][k] = 0;
        const uint32_t keep_last = (lane + 64 * (NBL - 1) <= T) ? ~0u : 0u;
        for (int x = 0; x < count; x++) {
            const uint16_t *Sx = S + x * 2 * T;
            uint32_t c[NBL], bv[NBL];
#pragma unroll
            for (int b = 0; b < NBL; b++) {
                c[b] = (W == 0 && b == 0 && lane == 0) ? 1u : 0u;   // C = 1
                bv[b] = (W == 0 && b == 0 && lane == 1) ? 1u : 0u;  // B = x
            }
            uint32_t bb = 1, L = 0;
            uint32_t colb[M];
            gf_mul_cols<M>(bb, colb);
#pragma unroll 1
            for (int n = 0; n < 2 * T; n++) scalar_bm_step_pair<W>(n, Sx, xch, lane, keep_last, c, bv, bb, L, colb);
#pragma unroll
            for (int b = 0; b < NBL; b++)
#pragma unroll
                for (int k = 0; k < M; k++) C[b][k] |= ((c[b] >> k) & 1u) << x;
        }
        __syncthreads();  // both wavefronts past their last syndrome read: the rows may go
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            if ((W == 0) != (b == 0)) continue;
            const int i = lane + 64 * b;
            if (i > T) continue;
#pragma unroll
            for (int k = 0; k < M; k++) lds[i * kRowWords + k] = C[b][k];
        }
    }

    // Coefficient rows coef[x][i] (i <= T) for ciphertext slot x, from the bit-sliced C.
    static __device__ __forceinline__ void write_coef_rows(const uint32_t (&C)[NBL][M], uint32_t *lds, int lane)
    {
        wave_sync();
        uint16_t *coef = reinterpret_cast<uint16_t *>(lds);
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            const int i = lane + 64 * b;
            if (i > T) continue;
#pragma unroll 2
            for (int x = 0; x < TILE; x++) {
                uint32_t v = 0;
#pragma unroll
                for (int k = 0; k < M; k++) v |= ((C[b][k] >> x) & 1u) << k;
                coef[x * CST + i] = (uint16_t)v;
            }
        }
        wave_sync();
    }

    // ---- stage 3: root search, 32 support points per lane per pass ----------------------------
    static __device__ __forceinline__ void roots(const FusedArgs &a, const FusedTile &tl, const uint32_t *lds, int lane,
                                 bool e_aligned)
    {
        const uint16_t *coef = reinterpret_cast<const uint16_t *>(lds);
        for (int x = 0; x < tl.count; x++) {
            const int item = a.items[tl.first + x];
            const KeyDev *kd = &a.keys[a.item_key[tl.first + x]];
            gu32 *lb = gptr(kd->lbits);
            const int ls = kd->lstride;
            uint8_t *eo = a.e + (size_t)item * a.e_stride;
            const uint16_t *cr = coef + x * CST;
            for (int g0 = 0; g0 < G; g0 += 64) {
                const int g = g0 + lane;
                const bool act = g < G;
                uint32_t av[M], r[M];
#pragma unroll
                for (int k = 0; k < M; k++) av[k] = act ? lb[k * ls + g] : 0u;
                const uint32_t c0 = uni(cr[0]);
#pragma unroll
                for (int k = 0; k < M; k++) r[k] = bitmask(c0, k);
                for (int j = 1; j <= T; j++) {
                    const uint32_t cj = uni(cr[j]);
                    bs_mul<M>(r, av, r);
#pragma unroll
                    for (int k = 0; k < M; k++) r[k] ^= bitmask(cj, k);
                }
                uint32_t nz = 0;
#pragma unroll
                for (int k = 0; k < M; k++) nz |= r[k];
                uint32_t z = ~nz;
                if (act) {
                    const int nbytes = (N / 8 - 4 * g) < 4 ? (N / 8 - 4 * g) : 4;
                    uint8_t *dst = eo + 4 * g;
                    if (nbytes == 4 && e_aligned) {
                        *reinterpret_cast<uint32_t *>(dst) = z;
                    } else {
                        for (int q = 0; q < nbytes; q++) dst[q] = (uint8_t)(z >> (8 * q));
                    }
                }
            }
        }
    }


    // ---- stage 3, additive-FFT variant --------------------------------------------------------
    // The reversed locator f_p = C_{T-p} (the polynomial root() evaluates, root_13.c:216-240) is
    // evaluated at every field element by the Gao-Mateer recursion (constants in FftTab):
    //  A. coefficient phase, bit-sliced across the tile like C: per depth a twist by public
    //     constants and the radix conversion f(x) = f0(x^2+x) + x f1(x^2+x) as XOR passes;
    //     the 2^LOG leaves (one per recursion path) go to LDS per ciphertext;
    //  B. per ciphertext, bit-sliced over 32 evaluation points per word: leaves broadcast, then
    //     LOG butterfly levels w0 = u + G v, w1 = w0 + v.  Word w, bit b is evaluation index
    //     c = 32w + b, the value at bitrev_m(c).  A zero value marks a root.
    // The key's own Benes network (the one support_gen applies to bitrev(i), root_13.c:181-211)
    // then moves the root indicator of index c to support position i, giving e directly.
    static constexpr int LOG = (T + 1 > 128) ? 8 : 7;  // 2^LOG >= t + 1 coefficients
    using FP = FftPlan<M, LOG>;
    static constexpr int NP = FP::NP, NW = FP::NW, WPL = FP::WPL, WPLEAF = FP::WPLEAF;
    static_assert(TILE * (1 << LOG) * 2 <= kWaveLds * 4, "leaf rows must fit the wave's LDS");

    // C planes to LDS rows (row i = coefficient i, word k = plane k).
    static __device__ __forceinline__ void write_coef_planes(const uint32_t (&C)[NBL][M], uint32_t *lds, int lane)
    {
#pragma unroll
        for (int b = 0; b < NBL; b++) {
            const int i = lane + 64 * b;
            if (i > T) continue;
#pragma unroll
            for (int k = 0; k < M; k++) lds[i * kRowWords + k] = C[b][k];
        }
    }

    static __device__ __forceinline__ void fft_coeff_phase(const uint32_t (&C)[NBL][M], uint32_t *lds, int lane,
                                                           const FftTab *tab)
    {
        wave_sync();
        write_coef_planes(C, lds, lane);
        wave_sync();
        fft_coeff_from_rows(lds, lane, tab);
    }

    // The coefficient phase from the coefficient rows already in LDS: gather f_p = C_{T-p},
    // then the twists and radix passes; leaves replace the rows.
    static __device__ __forceinline__ void fft_coeff_from_rows(uint32_t *lds, int lane, const FftTab *tab)
    {
        uint32_t F[NP][M];
#pragma unroll
        for (int pb = 0; pb < NP; pb++) {
            const int p = lane + 64 * pb;
            const bool live = p <= T;
            const int row = live ? T - p : 0;
#pragma unroll
            for (int k = 0; k < M; k++) F[pb][k] = live ? lds[row * kRowWords + k] : 0u;
        }
        wave_sync();
        FP::template coeff_depth<0>(F, lane, tab);
        FP::template coeff_depth<1>(F, lane, tab);
        FP::template coeff_depth<2>(F, lane, tab);
        FP::template coeff_depth<3>(F, lane, tab);
        FP::template coeff_depth<4>(F, lane, tab);
        FP::template coeff_depth<5>(F, lane, tab);
        FP::template coeff_depth<6>(F, lane, tab);
        if constexpr (LOG == 8) FP::template coeff_depth<7>(F, lane, tab);
        // leaves per ciphertext: leaf[x][p]
        uint16_t *leaf = reinterpret_cast<uint16_t *>(lds);
#pragma unroll
        for (int pb = 0; pb < NP; pb++) {
            const int p = lane + 64 * pb;
#pragma unroll 2
            for (int x = 0; x < TILE; x++) {
                uint32_t v = 0;
#pragma unroll
                for (int k = 0; k < M; k++) v |= ((F[pb][k] >> x) & 1u) << k;
                leaf[x * (1 << LOG) + p] = (uint16_t)v;
            }
        }
        wave_sync();
    }

    static __device__ __forceinline__ void roots_fft(const FusedArgs &a, const FusedTile &tl, const uint32_t *lds,
                                                     int lane, bool e_aligned, const FftTab *tab, int x0 = 0,
                                                     int dx = 1)
    {
        const uint16_t *leaf = reinterpret_cast<const uint16_t *>(lds);
        for (int x = x0; x < tl.count; x += dx) {
            const uint32_t wl_fin = FP::lane_part(FP::SCHED.fin, opaque(lane), 64);
            const int item = a.items[tl.first + x];
            const KeyDev *kd = &a.keys[a.item_key[tl.first + x]];
            uint8_t *eo = a.e + (size_t)item * a.e_stride;
            // the first level's multipliers load while the leaves are read
            uint32_t Gm[M];
            FP::template load_g<LOG - 1>(tab, FP::template first_g_index<LOG - 1>(lane), Gm);
            uint32_t W[WPL][M];
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                const int w = lane + 64 * r;  // initial layout
                const uint32_t v = leaf[x * (1 << LOG) + bitrev_bits((uint32_t)(w / WPLEAF), LOG)];
#pragma unroll
                for (int k = 0; k < M; k++) W[r][k] = bitmask(v, k);
            }
            if constexpr (LOG == 8) FP::template butterfly_ahead<7>(W, lane, tab, Gm);
            FP::template butterfly_ahead<6>(W, lane, tab, Gm);
            FP::template butterfly_ahead<5>(W, lane, tab, Gm);
            FP::template butterfly_ahead<4>(W, lane, tab, Gm);
            FP::template butterfly_ahead<3>(W, lane, tab, Gm);
            FP::template butterfly_ahead<2>(W, lane, tab, Gm);
            FP::template butterfly_ahead<1>(W, lane, tab, Gm);
            // the Benes network's first masks load during the last butterfly level
            gu32 *bmask = gptr(kd->benes);
            uint32_t mw0[WPL];
            FP::template benes_masks<0>(bmask, wl_fin, mw0);
            FP::template butterfly_ahead<0>(W, lane, tab, Gm);
            uint32_t mw1[WPL], mw2[WPL];
            FP::template benes_masks<1>(bmask, wl_fin, mw1);
            FP::template benes_masks<2>(bmask, wl_fin, mw2);
            uint32_t z[WPL];
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                uint32_t nz = 0;
#pragma unroll
                for (int k = 0; k < M; k++) nz |= W[r][k];
                z[r] = ~nz;
            }
            FP::template benes_from3<0>(z, bmask, wl_fin, mw0, mw1, mw2);
#pragma unroll
            for (int r = 0; r < WPL; r++) {
                const int g = (int)FP::final_word(wl_fin, r);
                if (g >= G) continue;
                const int nbytes = (N / 8 - 4 * g) < 4 ? (N / 8 - 4 * g) : 4;
                uint8_t *dst = eo + 4 * g;
                if (nbytes == 4 && e_aligned) {
                    *reinterpret_cast<uint32_t *>(dst) = z[r];
                } else {
                    for (int q = 0; q < nbytes; q++) dst[q] = (uint8_t)(z[r] >> (8 * q));
                }
            }
        }
    }

    static __device__ __forceinline__ void run(const FusedArgs &a, const FusedTile &tl, uint32_t *lds, int lane, bool e_aligned)
    {
        syndrome(a, tl, lds, lane);
        wave_sync();
        if constexpr (CMD_STAGE_STOP == 1) return;
        uint32_t C[NBL][M];
        if (tl.count <= SCALAR_BM)
            berlekamp_massey_scalar(lds, lane, tl.count, C);
        else
            berlekamp_massey(lds, lane, C);
        if constexpr (CMD_STAGE_STOP == 2) {
            // the locator's planes to LDS, so that the recurrence is not optimised away
            write_coef_planes(C, lds, lane);
            wave_sync();
            return;
        }
        if constexpr (ALG == 0) {
            write_coef_rows(C, lds, lane);
            roots(a, tl, lds, lane, e_aligned);
        } else {
            const FftTab *tab = a.fft + fft_variant(M, LOG);
            fft_coeff_phase(C, lds, lane, tab);
            if constexpr (CMD_STAGE_STOP == 3) return;
            roots_fft(a, tl, lds, lane, e_aligned, tab);
        }
        wave_sync();
    }

    // ---- two wavefronts per tile (waves_per_ct == 2) -----------------------------------------
    // One workgroup of two wavefronts shares a tile and its LDS, so that four wavefronts per
    // SIMD fit where the single-wavefront kernel fits two, without adding Berlekamp-Massey work
    // (a tile's recurrence is split, not repeated):
    //  1. the syndrome is split by table column: wave w accumulates the dword column
    //     lane + 64 w of every row (syndromes 2c, 2c+1 of column c) for every ciphertext of the
    //     tile, with four-row tables of one dword per lane (4 KB per wave), and stores the bit
    //     planes of its own syndromes (when t <= 64 every column belongs to wave 0);
    //  2. wave 0 holds lane block 0 of C and B, wave 1 block 1 (when there is one) and, when t
    //     is a multiple of 64, the transposed top coefficient; per iteration each wave reduces
    //     its part of the discrepancy, the parts meet in LDS (one barrier per iteration, slots
    //     alternating so a slot is rewritten only after both waves passed the next barrier),
    //     and wave 0 publishes lane 63 of its C and B, from which wave 1 takes the carry of the
    //     shift of B (and for t = 64 the new B_t);
    //  3. wave 0 runs the FFT coefficient phase; both waves run the butterflies, Benes network
    //     and stores of alternate ciphertexts.
    // Tiles of at most SCALAR_BM ciphertexts run their Berlekamp-Massey in wave 0 alone.
    static constexpr int XCH = 64;  // words per exchange slot: d of wave 0 | C_63 | B_63 | d of wave 1

    // The eight rows of group g, one dword column per lane.
    static __device__ __forceinline__ void load_group_col(gu32 *rows, int g, int col, uint32_t (&v)[8])
    {
        gu32 *base = rows + (size_t)(8 * g) * PAIRS;
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = col < PAIRS ? base[(size_t)q * PAIRS + col] : 0u;
    }

    // Group v into the accumulators: rows 4h .. 4h+3 through one 16-entry table at a time
    // (entry c = XOR of the rows set in c, lane-major); ciphertext x adds the entry its nibble
    // selects.  This wave's LDS operations execute in order, so a table is rewritten only after
    // the reads of the previous one.
    template <int BIT0>
    static __device__ __forceinline__ void add_group_col(const uint32_t (&v)[8], const uint32_t (&cw)[CH],
                                                         uint32_t (&acc)[CH], uint32_t *tab, int lane)
    {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t cur = 0;
#pragma unroll
            for (int i = 1; i < 16; i++) {
                cur ^= v[4 * h + __builtin_ctz(i)];
                tab[(i ^ (i >> 1)) * 64 + lane] = cur;
            }
#pragma unroll
            for (int x = 0; x < CH; x++) {
                acc[x] ^= tab[((cw[x] >> (BIT0 + 4 * h)) & 15u) * 64 + lane];
                asm volatile("" : "+v"(acc[x]));
            }
        }
    }

    // All row groups of one key's table for the ciphertexts of a run (others read as zero),
    // software-pipelined: the next group's loads are in flight while the current one is added.
    static __device__ __forceinline__ void stream_col(gu32 *rows, const uint8_t *my_ct, int col,
                                                      uint32_t (&acc)[CH], uint32_t *tab, int lane, bool al)
    {
        uint32_t va[8], vb[8], cw[CH];
        load_group_col(rows, 0, col, va);
        uint32_t next = ct_word(my_ct, 0, al);  // the next word's load is in flight
        for (int w = 0; w < WORDS; w++) {
            const uint32_t mine = next;
            if (w + 1 < WORDS) next = ct_word(my_ct, w + 1, al);
#pragma unroll
            for (int x = 0; x < CH; x++) cw[x] = (uint32_t)__builtin_amdgcn_readlane((int)mine, x);
            const int g = 4 * w;
            if (g + 1 < NGROUP) load_group_col(rows, g + 1, col, vb);
            add_group_col<0>(va, cw, acc, tab, lane);
            if (g + 2 < NGROUP) load_group_col(rows, g + 2, col, va);
            if (g + 1 < NGROUP) add_group_col<8>(vb, cw, acc, tab, lane);
            if (g + 2 >= NGROUP) break;
            load_group_col(rows, g + 3, col, vb);
            add_group_col<16>(va, cw, acc, tab, lane);
            if (g + 4 < NGROUP) load_group_col(rows, g + 4, col, va);
            add_group_col<24>(vb, cw, acc, tab, lane);
        }
    }

    static __device__ __forceinline__ void syndrome_pair(const FusedArgs &a, const FusedTile &tl, uint32_t *lds,
                                                         int lane, int wv)
    {
        const bool al = ct_aligned(a);
        const int col = lane + 64 * wv;                // this wave's table column
        const bool has_col = R == 2 || wv == 0;        // t <= 64: wave 0 owns every column
        const bool live = lane < CH && lane < tl.count;
        const int my_key = live ? a.item_key[tl.first + lane] : -1;
        const uint8_t *ct_ptr = live ? a.ct + (size_t)a.items[tl.first + lane] * a.ct_stride : nullptr;
        const int n_live = tl.count < CH ? tl.count : CH;
        int runs = 0;
        for (uint64_t rem = __ballot(live); rem; runs++) {
            const int k = __builtin_amdgcn_readlane(my_key, __builtin_ctzll(rem));
            rem &= ~__ballot(my_key == k);
        }
        uint32_t acc[CH];
#pragma unroll
        for (int x = 0; x < CH; x++) acc[x] = 0;
        if (has_col && tl.count > SCALAR_BM && 2 * runs <= n_live) {
            uint32_t *tab = lds + wv * 16 * 64;
            tab[lane] = 0;  // entry 0
            for (uint64_t rem = __ballot(live); rem;) {
                const int key = __builtin_amdgcn_readlane(my_key, __builtin_ctzll(rem));
                const bool in_run = my_key == key;
                rem &= ~__ballot(in_run);
                const KeyDev *kd = &a.keys[key];
                stream_col(gptr(reinterpret_cast<const uint32_t *>(kd->table)), in_run ? ct_ptr : nullptr, col, acc,
                           tab, lane, al);
            }
        } else if (has_col) {
            // one key per ciphertext: only the rows of its set bits (the ciphertext is public),
            // GRP at a time to keep loads in flight
            constexpr int GRP = 16;
#pragma unroll
            for (int x = 0; x < CH; x++) {
                if (x >= tl.count) continue;
                const KeyDev *kd = &a.keys[a.item_key[tl.first + x]];
                gu32 *rows = gptr(reinterpret_cast<const uint32_t *>(kd->table));
                const uint8_t *ctp = a.ct + (size_t)a.items[tl.first + x] * a.ct_stride;
                // the whole ciphertext in one load, word w in lane w
                const uint32_t ctw = lane < WORDS ? ct_word(ctp, lane, al) : 0u;
                for (int w = 0; w < WORDS; w++) {
                    uint32_t bits = (uint32_t)__builtin_amdgcn_readlane((int)ctw, w);
                    gu32 *wrows = rows + (size_t)(w * 32) * PAIRS;
                    while (bits) {
                        uint32_t v[GRP];
#pragma unroll
                        for (int j = 0; j < GRP; j++) {
                            const bool any = bits != 0;
                            const int bit = any ? __builtin_ctz(bits) : 0;
                            bits &= bits - 1;
                            v[j] = (any && col < PAIRS) ? wrows[(size_t)bit * PAIRS + col] : 0u;
                        }
#pragma unroll
                        for (int j = 0; j < GRP; j++) acc[x] ^= v[j];
                    }
                }
            }
        }
        __syncthreads();  // both waves' tables done with: the planes may overwrite them
        if (wv == 0 && lane < kRowWords) lds[lane] = 0;  // row 0: the zero row read for s_{negative}
        if (has_col && col < PAIRS) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int j = 2 * col + h;
#pragma unroll
                for (int k = 0; k < M; k++) {
                    uint32_t plane = 0;
#pragma unroll
                    for (int x = 0; x < CH; x++) plane |= ((acc[x] >> (16 * h + k)) & 1u) << x;
                    lds[(j + 1) * kRowWords + k] = plane;
                }
            }
        }
        __syncthreads();
    }

    template <int W, int LO, int HI>
    static __device__ __forceinline__ void bm_step_pair(int n, const uint4 *rows4, uint32_t *xch, int lane,
                                                        uint32_t (&C)[M], uint32_t (&B)[M], uint32_t (&bb)[M],
                                                        uint32_t (&Lp)[9], uint32_t &ct, uint32_t &bt,
                                                        uint32_t keep_last, uint32_t pm)
    {
        constexpr bool FULL = LO >= NFULL;
        constexpr bool BLOCK = W < NBV;
        constexpr bool TOP = TOPX && W == 1 && FULL;
        // inside every phase a block's part in the products and in the shift is fixed
        constexpr bool MAC = BLOCK && 64 * W <= LO + 1;
        constexpr bool SHIFT = BLOCK && 64 * W <= LO + 2;
        static_assert(!BLOCK || 64 * W <= LO + 1 || 64 * W > HI, "products join at a phase boundary");
        static_assert(!BLOCK || 64 * W <= LO + 2 || 64 * W > HI + 1, "shift joins at a phase boundary");
        uint32_t dp[M];
        if constexpr (MAC) {
            uint32_t p[2 * M - 1];
#pragma unroll
            for (int s2 = 0; s2 < 2 * M - 1; s2++) p[s2] = 0;
            const int i = lane + 64 * W;
            const int row = (n - i + 1) > 0 ? (n - i + 1) : 0;
            uint32_t sv[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 w = rows4[row * 4 + q];
                sv[4 * q] = w.x; sv[4 * q + 1] = w.y; sv[4 * q + 2] = w.z; sv[4 * q + 3] = w.w;
            }
            bs_mac_wide<M>(C, sv, p);
            uint32_t part[M];
            bs_fold<M>(p, part);
#pragma unroll
            for (int k = 0; k < M; k++) dp[k] = wave_xor_all(part[k]);
        } else {
#pragma unroll
            for (int k = 0; k < M; k++) dp[k] = 0;
        }
        if constexpr (TOP) {
            const int row = n - T + 1 > 0 ? n - T + 1 : 0;
            uint32_t sv[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 w = rows4[row * 4 + q];
                sv[4 * q] = w.x; sv[4 * q + 1] = w.y; sv[4 * q + 2] = w.z; sv[4 * q + 3] = w.w;
            }
            uint32_t top = and_xor(0u, ct, uni(sv[M - 1]));
#pragma unroll
            for (int q = M - 2; q >= 0; q--) top = and_xor(tx_mulx(top, pm), ct, uni(sv[q]));
#pragma unroll
            for (int k = 0; k < M; k++) dp[k] ^= (uint32_t)__builtin_amdgcn_readlane((int)top, k);
        }
        // exchange through LDS
        uint32_t *slot = xch + (n & 1) * XCH;
        if (lane == 0) {
            uint32_t w16[16];
#pragma unroll
            for (int k = 0; k < 16; k++) w16[k] = k < M ? dp[k] : 0u;
            uint4 *dst = reinterpret_cast<uint4 *>(slot + (W == 0 ? 0 : 48));
#pragma unroll
            for (int q = 0; q < 4; q++) dst[q] = make_uint4(w16[4 * q], w16[4 * q + 1], w16[4 * q + 2], w16[4 * q + 3]);
        }
        if constexpr (W == 0) {
            if (lane == 63) {
#pragma unroll
                for (int k = 0; k < M; k++) {
                    slot[16 + k] = C[k];
                    slot[32 + k] = B[k];
                }
            }
        }
        __syncthreads();
        uint32_t d[M];
        {
            const uint4 *oth = reinterpret_cast<const uint4 *>(slot + (W == 0 ? 48 : 0));
            uint32_t o[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 w = oth[q];
                o[4 * q] = w.x; o[4 * q + 1] = w.y; o[4 * q + 2] = w.z; o[4 * q + 3] = w.w;
            }
#pragma unroll
            for (int k = 0; k < M; k++) d[k] = dp[k] ^ uni(o[k]);
        }
        [[maybe_unused]] uint32_t c63[M], b63[M];
        if constexpr (W == 1) {
#pragma unroll
            for (int k = 0; k < M; k++) {
                c63[k] = slot[16 + k];
                b63[k] = slot[32 + k];
            }
        }
        uint32_t mne = 0;
#pragma unroll
        for (int k = 0; k < M; k++) mne |= d[k];
        const uint32_t mle = mne & bs_le_const(Lp, (uint32_t)n >> 1);

        if constexpr (TOP) {
            // C_t' = b C_t + d B_t
            uint32_t acc = and_xor(and_xor(0u, ct, bb[M - 1]), bt, d[M - 1]);
#pragma unroll
            for (int q = M - 2; q >= 0; q--) acc = and_xor(and_xor(tx_mulx(acc, pm), ct, bb[q]), bt, d[q]);
            ct = acc;
        }
        if constexpr (MAC) {
            uint32_t cn[M];
            bs_mul_add2<M>(bb, C, d, B, cn);
#pragma unroll
            for (int k = 0; k < M; k++) {
                B[k] = __builtin_amdgcn_bitop3_b32(mle, C[k], B[k], 0xCA);  // mle ? C : B
                C[k] = cn[k];
            }
        }
        [[maybe_unused]] uint32_t carry[M];
        if constexpr (W == 1) {
            // lane 63 of wave 0's B after the selection: the coefficient its shift moves up
#pragma unroll
            for (int k = 0; k < M; k++) carry[k] = __builtin_amdgcn_bitop3_b32(mle, c63[k], b63[k], 0xCA);
        }
        if constexpr (TOP) {
            // the shift moves coefficient t-1 of the selected B into B_t
#pragma unroll
            for (int k = 0; k < M; k++) {
                uint32_t c;
                if constexpr (NBV == 2)
                    c = (uint32_t)__builtin_amdgcn_readlane((int)B[k], 63);
                else
                    c = carry[k];
                bt = lane == k ? c : bt;
            }
        }
        if constexpr (SHIFT) {
#pragma unroll
            for (int k = 0; k < M; k++) {
                if constexpr (W == 0)
                    B[k] = shift_up1_zero(B[k]);
                else
                    B[k] = shift_up1(B[k], carry[k]);
            }
        }
        if constexpr (!TOPX && BLOCK && W == NBV - 1) {
#pragma unroll
            for (int k = 0; k < M; k++) B[k] &= keep_last;
        }
        uint32_t Ln[9];
        bs_const_sub((uint32_t)n + 1, Lp, Ln);
#pragma unroll
        for (int q = 0; q < 9; q++) Lp[q] = (Ln[q] & mle) | (Lp[q] & ~mle);
#pragma unroll
        for (int k = 0; k < M; k++) bb[k] = (d[k] & mle) | (bb[k] & ~mle);
    }

    // Wave W's part of the recurrence; leaves its coefficient rows (and the top row) in LDS.
    template <int W>
    static __device__ __forceinline__ void berlekamp_massey_pair(uint32_t *lds, uint32_t *xch, int lane)
    {
        uint32_t C[M], B[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
            C[k] = (W == 0 && k == 0 && lane == 0) ? ~0u : 0u;  // C = 1
            B[k] = (W == 0 && k == 0 && lane == 1) ? ~0u : 0u;  // B = x
        }
        uint32_t bb[M], Lp[9];
#pragma unroll
        for (int k = 0; k < M; k++) bb[k] = k == 0 ? ~0u : 0u;
#pragma unroll
        for (int q = 0; q < 9; q++) Lp[q] = 0;
        const uint32_t keep_last = (lane + 64 * (NBV - 1) <= T) ? ~0u : 0u;
        const uint32_t pm = (lane < M && ((Modulus<M>::low >> lane) & 1u)) ? ~0u : 0u;
        uint32_t ct = 0, bt = 0;
        const uint4 *rows4 = reinterpret_cast<const uint4 *>(lds);
        static_assert(NBV <= 2, "two waves cover two lane blocks");
        constexpr int P1 = NBV > 1 ? 62 : NFULL, P2 = NBV > 1 ? 63 : NFULL;
#pragma unroll 1
        for (int n = 0; n < P1; n++) bm_step_pair<W, 0, P1>(n, rows4, xch, lane, C, B, bb, Lp, ct, bt, keep_last, pm);
#pragma unroll 1
        for (int n = P1; n < P2; n++) bm_step_pair<W, P1, P2>(n, rows4, xch, lane, C, B, bb, Lp, ct, bt, keep_last, pm);
#pragma unroll 1
        for (int n = P2; n < NFULL; n++)
            bm_step_pair<W, P2, NFULL>(n, rows4, xch, lane, C, B, bb, Lp, ct, bt, keep_last, pm);
#pragma unroll 1
        for (int n = NFULL; n < 2 * T; n++)
            bm_step_pair<W, NFULL, 2 * T>(n, rows4, xch, lane, C, B, bb, Lp, ct, bt, keep_last, pm);
        // every syndrome read of both waves finished before the last exchange barrier: the
        // rows can now hold the coefficients
        if constexpr (W < NBV) {
            const int i = lane + 64 * W;
            if (i <= T) {
#pragma unroll
                for (int k = 0; k < M; k++) lds[i * kRowWords + k] = C[k];
            }
        }
        if constexpr (TOPX && W == 1) {
            if (lane < M) lds[T * kRowWords + lane] = ct;  // lane k holds plane k of C_t
        }
    }

    static __device__ __forceinline__ void run_pair(const FusedArgs &a, const FusedTile &tl, uint32_t *lds,
                                                    uint32_t *xch, int lane, int wv, bool e_aligned)
    {
        static_assert(ALG == 1, "two-wavefront tiles use the additive-FFT root search");
        syndrome_pair(a, tl, lds, lane, wv);
        if (tl.count <= SCALAR_BM) {
            if (wv == 0)
                berlekamp_massey_scalar_pair<0>(lds, xch, lane, tl.count);
            else
                berlekamp_massey_scalar_pair<1>(lds, xch, lane, tl.count);
        } else if (wv == 0) {
            berlekamp_massey_pair<0>(lds, xch, lane);
        } else {
            berlekamp_massey_pair<1>(lds, xch, lane);
        }
        __syncthreads();
        const FftTab *tab = a.fft + fft_variant(M, LOG);
        if (wv == 0) fft_coeff_from_rows(lds, lane, tab);
        __syncthreads();
        roots_fft(a, tl, lds, lane, e_aligned, tab, wv, 2);
    }
};

// LDS of a two-wavefront workgroup: the tile's planes/tables/rows/leaves, two exchange slots
// and the tile-id slot.
constexpr int kPairXchWords = 128;
constexpr int kPairLds = kWaveLds + kPairXchWords + 16;
static_assert(kWaveLds % 4 == 0, "exchange slots 16-byte aligned");


// Workgroups are dispatched to the XCDs in turn (workgroup b to XCD b % kXcds), and each XCD
// has its own L2.  The first tile of every wavefront is assigned statically so that the
// wavefronts of one XCD take one contiguous range of the tile list, which the host sorts by
// set and key: each XCD then reads the tables of few keys into its L2 instead of all of them.
// Wavefront (b, wv) takes index (prefix(b % kXcds) + b / kXcds) * waves + wv, prefix(x) being
// the number of groups on XCDs below x: a bijection of the grid's wavefronts onto [0, W).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_first_tile(uint32_t b, uint32_t groups, uint32_t waves, uint32_t wv)
{
    const uint32_t x = b % kXcds, per = groups / kXcds, extra = groups % kXcds;
    const uint32_t prefix = x * per + (x < extra ? x : extra);
    return (prefix + b / kXcds) * waves + wv;
}

// Persistent kernel: each wavefront takes its first tile by the XCD-major assignment above,
// then pulls tiles W, W+1, ... from the counter (W = wavefronts in the grid) until the list is
// exhausted.  Dynamic LDS of kWaveLds words per wavefront; the workgroup size (64..256) gives
// the wavefronts per group.  OCC is the number of wavefronts per SIMD the build is
// register-limited for: 1 (up to 512 VGPR+AGPR per lane) or 2 (256).
template <int ALG, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, OCC))) void k_fused(FusedArgs a)
{
    extern __shared__ uint32_t lds_dyn[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *lds = lds_dyn + wv * kWaveLds;
    const bool e_aligned = ((a.e_stride & 3u) == 0) && ((reinterpret_cast<uintptr_t>(a.e) & 3u) == 0);
    const uint32_t waves = blockDim.x >> 6, n_waves = gridDim.x * waves;
    uint32_t tid = (uint32_t)__builtin_amdgcn_readfirstlane(
        (int)xcd_first_tile(blockIdx.x, gridDim.x, waves, (uint32_t)wv));
    for (bool first = true;; first = false) {
        if (!first) {
            uint32_t q = 0;
            if (lane == 0) q = atomicAdd(a.tile_counter, 1u);
            tid = n_waves + (uint32_t)__builtin_amdgcn_readfirstlane((int)q);
        }
        if (tid >= (uint32_t)a.n_tiles) break;
        const FusedTile tl = a.tiles[tid];
        switch (tl.pidx) {
        case 0: FusedTileKernel<0, ALG>::run(a, tl, lds, lane, e_aligned); break;
        case 1: FusedTileKernel<1, ALG>::run(a, tl, lds, lane, e_aligned); break;
        case 2: FusedTileKernel<2, ALG>::run(a, tl, lds, lane, e_aligned); break;
        case 3: FusedTileKernel<3, ALG>::run(a, tl, lds, lane, e_aligned); break;
        case 4: FusedTileKernel<4, ALG>::run(a, tl, lds, lane, e_aligned); break;
        default: break;
        }
    }
}

// Persistent kernel, two wavefronts per tile: each workgroup is one pair, which takes tiles
// from the counter (wave 0 fetches, the id goes through LDS) until the list is exhausted.
// OCC is the number of wavefronts per SIMD the build is register-limited for.
template <int ALG, int OCC>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(OCC, OCC))) void k_fused_pair(FusedArgs a)
{
    extern __shared__ uint32_t lds_dyn[];
    // the wave index as a wave-uniform value: branches on it (which hold barriers) stay
    // scalar branches, never executed on both sides under an exec mask
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint32_t *xch = lds_dyn + kWaveLds;
    uint32_t *tid_slot = xch + kPairXchWords;
    const bool e_aligned = ((a.e_stride & 3u) == 0) && ((reinterpret_cast<uintptr_t>(a.e) & 3u) == 0);
    for (;;) {
        if (threadIdx.x == 0) *tid_slot = atomicAdd(a.tile_counter, 1u);
        __syncthreads();
        const uint32_t tid = uni(*tid_slot);
        if (tid >= (uint32_t)a.n_tiles) break;  // both waves read the same id
        const FusedTile tl = a.tiles[tid];
        switch (tl.pidx) {
        case 0: FusedTileKernel<0, ALG>::run_pair(a, tl, lds_dyn, xch, lane, wv, e_aligned); break;
        case 1: FusedTileKernel<1, ALG>::run_pair(a, tl, lds_dyn, xch, lane, wv, e_aligned); break;
        case 2: FusedTileKernel<2, ALG>::run_pair(a, tl, lds_dyn, xch, lane, wv, e_aligned); break;
        case 3: FusedTileKernel<3, ALG>::run_pair(a, tl, lds_dyn, xch, lane, wv, e_aligned); break;
        case 4: FusedTileKernel<4, ALG>::run_pair(a, tl, lds_dyn, xch, lane, wv, e_aligned); break;
        default: break;
        }
        __syncthreads();  // the tile's LDS and the id slot are free again
    }
}

template <int ALG, int OCC>
static hipError_t launch_fused_pair_variant(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st)
{
    int groups = n_cu * s.groups_per_cu;
    if (groups > a.n_tiles) groups = a.n_tiles;  // one tile per pair at most: spread small launches
    if (groups < 1) groups = 1;
    hipLaunchKernelGGL((k_fused_pair<ALG, OCC>), dim3(groups), dim3(128), (size_t)kPairLds * 4, st, a);
    return hipGetLastError();
}

// Launches k_fused<ALG, OCC> with the schedule's workgroup size and persistent grid.
template <int ALG, int OCC>
static hipError_t launch_fused_variant(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st)
{
    // the dynamic-LDS limit of the kernel is raised once per process (any thread may launch)
    static std::once_flag once;
    static hipError_t set_rc = hipSuccess;
    std::call_once(once, [] {
        set_rc = hipFuncSetAttribute((const void *)k_fused<ALG, OCC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     4 * kWaveLds * 4);
    });
    if (set_rc != hipSuccess) return set_rc;
    const size_t lds_bytes = (size_t)s.waves_per_group * kWaveLds * 4;
    int groups = n_cu * s.groups_per_cu;
    const int max_groups = (a.n_tiles + s.waves_per_group - 1) / s.waves_per_group;
    if (groups > max_groups) groups = max_groups;
    if (groups < 1) groups = 1;
    hipLaunchKernelGGL((k_fused<ALG, OCC>), dim3(groups), dim3(64 * s.waves_per_group), lds_bytes, st, a);
    return hipGetLastError();
}

}  // namespace cmd
