// Device-side GF(2^m) arithmetic for gfx950 (wave64, 32-bit VALU lanes).
//
// Two representations:
//  * scalar: one field element per lane in the low m bits of a 32-bit register
//    (the reference's uint16 gf, CPU/param/param6960119/include/common.h:34);
//  * bit-sliced: m registers, register k holding bit k of 32 independent elements
//    (one per bit position).  A product of 32 element pairs then costs m*m AND-XOR
//    steps plus the fold of the m-1 high planes, all plain bitwise VALU work with no
//    data-dependent addressing.
// Field moduli: x^13+x^4+x^3+x+1 (CPU/src/common/gf_13.c:103-108) and x^12+x^3+1
// (CPU/param/param348864/src/gf.c:116-122).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cmd {

// Low part of the modulus (x^m term removed).
template <int M>
struct Modulus;
template <>
struct Modulus<13> { static constexpr uint32_t low = 0x1Bu; };
template <>
struct Modulus<12> { static constexpr uint32_t low = 0x9u; };

// all-ones if bit k of x is set, else 0
__device__ __forceinline__ uint32_t bitmask(uint32_t x, int k)
{
    return (uint32_t)__builtin_amdgcn_sbfe((int)x, k, 1);
}

// gfx950's three-input v_bitop3_b32 evaluates any function of three operands in one
// instruction; its 8-bit table is indexed by (src0 << 2) | (src1 << 1) | src2.
// 0x6A: src2 ^ (src0 & src1)      0x96: src0 ^ src1 ^ src2
// Written as builtins so that every AND-XOR term of a product is one instruction (left to
// itself the compiler re-associates the sums into trees of separate ANDs and XORs).
__device__ __forceinline__ uint32_t and_xor(uint32_t acc, uint32_t x, uint32_t y)
{
    return __builtin_amdgcn_bitop3_b32(x, y, acc, 0x6A);
}

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Reduces a product of degree <= 2M-2 modulo the field polynomial.
template <int M>
__device__ __forceinline__ uint32_t gf_reduce(uint32_t v)
{
    constexpr uint32_t low = Modulus<M>::low;
    constexpr uint32_t mask = (1u << M) - 1u;
    uint32_t hi = v >> M;
    v &= mask;
#pragma unroll
    for (int b = 0; b < 5; b++)
        if ((low >> b) & 1u) v ^= hi << b;
    hi = v >> M;
    v &= mask;
#pragma unroll
    for (int b = 0; b < 5; b++)
        if ((low >> b) & 1u) v ^= hi << b;
    return v;
}

// Carry-less product of two reduced elements (degree <= 2M-2, not reduced): shift-and-add
// over the bits of b, one AND-XOR per bit, two partial sums.  Sums of such products reduce
// once (the reduction is linear).
template <int M>
__device__ __forceinline__ uint32_t gf_clmul(uint32_t a, uint32_t b)
{
    uint32_t acc0 = 0, acc1 = 0;
#pragma unroll
    for (int i = 0; i < M; i += 2) acc0 = and_xor(acc0, a << i, bitmask(b, i));
#pragma unroll
    for (int i = 1; i < M; i += 2) acc1 = and_xor(acc1, a << i, bitmask(b, i));
    return acc0 ^ acc1;
}

// Scalar product.
template <int M>
__device__ __forceinline__ uint32_t gf_mul(uint32_t a, uint32_t b)
{
    return gf_reduce<M>(gf_clmul<M>(a, b));
}

// x * a for a reduced element
template <int M>
__device__ __forceinline__ uint32_t gf_xtime(uint32_t a)
{
    uint32_t v = a << 1;
    return (v & ((1u << M) - 1u)) ^ (Modulus<M>::low & (0u - (v >> M)));
}

// Multiplication table of a fixed element a: col[k] = a * x^k.  Then a*b is the XOR of
// col[k] over the set bits k of b (no reduction step needed).
template <int M>
__device__ __forceinline__ void gf_mul_cols(uint32_t a, uint32_t col[M])
{
    col[0] = a;
#pragma unroll
    for (int k = 1; k < M; k++) col[k] = gf_xtime<M>(col[k - 1]);
}

template <int M>
__device__ __forceinline__ uint32_t gf_mul_by_cols(const uint32_t col[M], uint32_t b)
{
    uint32_t acc0 = 0, acc1 = 0;
#pragma unroll
    for (int k = 0; k < M; k += 2) acc0 = and_xor(acc0, col[k], bitmask(b, k));
#pragma unroll
    for (int k = 1; k < M; k += 2) acc1 = and_xor(acc1, col[k], bitmask(b, k));
    return acc0 ^ acc1;
}

// a^(2^M - 2): the inverse, with inv(0) == 0.
template <int M>
__device__ __forceinline__ uint32_t gf_inv(uint32_t a)
{
    // a^(2^(M-1)-1) by an addition chain on runs of ones, then one squaring.
    uint32_t x1 = a;                                   // 1
    uint32_t x2 = gf_mul<M>(gf_mul<M>(x1, x1), x1);     // 11
    uint32_t x4 = x2;
    x4 = gf_mul<M>(x4, x4); x4 = gf_mul<M>(x4, x4);
    x4 = gf_mul<M>(x4, x2);                            // 1111
    uint32_t x8 = x4;
#pragma unroll
    for (int i = 0; i < 4; i++) x8 = gf_mul<M>(x8, x8);
    x8 = gf_mul<M>(x8, x4);                            // 1^8
    uint32_t r = x8;
    constexpr int rest = M - 1 - 8;                    // 4 (M=13) or 3 (M=12)
    if constexpr (rest == 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) r = gf_mul<M>(r, r);
        r = gf_mul<M>(r, x4);                          // 1^12
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) r = gf_mul<M>(r, r);
        r = gf_mul<M>(r, gf_mul<M>(gf_mul<M>(x2, x2), x1));  // 1^11: append 111 = a^7
    }
    return gf_mul<M>(r, r);                            // 1^(M-1) 0
}

// ---- bit-sliced ------------------------------------------------------------------------
// Which high planes of an unreduced product land in plane k after reduction: x^s mod the
// field polynomial for s = M .. 2M-2, as a table computed at compile time.
template <int M>
struct FoldMap {
    int count[M];
    int src[M][M];
};

template <int M>
constexpr FoldMap<M> make_fold_map()
{
    FoldMap<M> f{};
    constexpr uint32_t poly = (1u << M) | Modulus<M>::low;
    for (int s = M; s <= 2 * M - 2; s++) {
        uint32_t v = 1u << s;
        for (int b = 2 * M - 2; b >= M; b--)
            if ((v >> b) & 1u) v ^= poly << (b - M);
        for (int k = 0; k < M; k++)
            if ((v >> k) & 1u) f.src[k][f.count[k]++] = s;
    }
    return f;
}

// Folds planes 2M-2 .. M of an unreduced product back into planes 0 .. M-1: each low plane
// takes its high sources two at a time with one three-input XOR (30 instructions for
// GF(2^13), 14 for GF(2^12), against 48 and 22 for the plane-by-plane fold).
template <int M>
__device__ __forceinline__ void bs_fold_planes(uint32_t p[2 * M - 1])
{
    constexpr FoldMap<M> F = make_fold_map<M>();
#pragma unroll
    for (int k = 0; k < M; k++) {
        uint32_t acc = p[k];
#pragma unroll
        for (int i = 0; i + 1 < F.count[k]; i += 2) acc = xor3(acc, p[F.src[k][i]], p[F.src[k][i + 1]]);
        if (F.count[k] & 1) acc ^= p[F.src[k][F.count[k] - 1]];
        p[k] = acc;
    }
}

// out = a * b for 32 element pairs (bit lanes).  out may alias a or b.
template <int M>
__device__ __forceinline__ void bs_mul(const uint32_t a[M], const uint32_t b[M], uint32_t out[M])
{
    uint32_t p[2 * M - 1];
#pragma unroll
    for (int s = 0; s < 2 * M - 1; s++) p[s] = 0;
#pragma unroll
    for (int i = 0; i < M; i++)
#pragma unroll
        for (int j = 0; j < M; j++) p[i + j] = and_xor(p[i + j], a[i], b[j]);
    bs_fold_planes<M>(p);
#pragma unroll
    for (int k = 0; k < M; k++) out[k] = p[k];
}

// out = a*b + c*d for 32 element pairs each, with a single fold.
template <int M>
__device__ __forceinline__ void bs_mul_add2(const uint32_t a[M], const uint32_t b[M], const uint32_t c[M],
                                            const uint32_t d[M], uint32_t out[M])
{
    uint32_t p[2 * M - 1];
#pragma unroll
    for (int s = 0; s < 2 * M - 1; s++) p[s] = 0;
#pragma unroll
    for (int i = 0; i < M; i++)
#pragma unroll
        for (int j = 0; j < M; j++) {
            p[i + j] = and_xor(p[i + j], a[i], b[j]);
            p[i + j] = and_xor(p[i + j], c[i], d[j]);
        }
    bs_fold_planes<M>(p);
#pragma unroll
    for (int k = 0; k < M; k++) out[k] = p[k];
}

// Unreduced accumulation p[0..2M-2] += a*b (carry-less, no fold).
template <int M>
__device__ __forceinline__ void bs_mac_wide(const uint32_t a[M], const uint32_t b[M], uint32_t p[2 * M - 1])
{
#pragma unroll
    for (int i = 0; i < M; i++)
#pragma unroll
        for (int j = 0; j < M; j++) p[i + j] = and_xor(p[i + j], a[i], b[j]);
}

// One Karatsuba level over the planes: an element splits into lo (planes 0..L-1) and hi
// (planes L..M-1), a = a0 + x^L a1, and a*b = q0 + x^L (q1 + q0 + q2) + x^2L q2 with
// q0 = a0 b0, q2 = a1 b1, q1 = (a0 + a1)(b0 + b1) (characteristic 2: no signs).  The three
// half products cost L^2 + H^2 + L^2 AND-XOR terms instead of M^2 (134 against 169 for
// GF(2^13)); the sums and the recombination cost about as much again, so the split pays only
// for a sum of products, where the half products of every term accumulate apart and
// recombine once.
template <int M>
struct Kara {
    static constexpr int L = (M + 1) / 2, H = M - L;
    static constexpr int NQ0 = 2 * L - 1, NQ2 = 2 * H - 1;
};

// s = a0 + a1 (L planes; plane L-1 is a0's alone when H < L).
template <int M>
__device__ __forceinline__ void kara_half_sum(const uint32_t a[M], uint32_t s[Kara<M>::L])
{
    constexpr int L = Kara<M>::L, H = Kara<M>::H;
#pragma unroll
    for (int i = 0; i < L; i++) s[i] = i < H ? a[i] ^ a[L + i] : a[i];
}

// q0 += a0 b0, q2 += a1 b1, q1 += as bs, with as, bs the half sums of a and b.
template <int M>
__device__ __forceinline__ void kara_acc(const uint32_t a[M], const uint32_t b[M], const uint32_t as[Kara<M>::L],
                                         const uint32_t bs[Kara<M>::L], uint32_t q0[Kara<M>::NQ0],
                                         uint32_t q1[Kara<M>::NQ0], uint32_t q2[Kara<M>::NQ2])
{
    constexpr int L = Kara<M>::L, H = Kara<M>::H;
#pragma unroll
    for (int i = 0; i < L; i++)
#pragma unroll
        for (int j = 0; j < L; j++) q0[i + j] = and_xor(q0[i + j], a[i], b[j]);
#pragma unroll
    for (int i = 0; i < H; i++)
#pragma unroll
        for (int j = 0; j < H; j++) q2[i + j] = and_xor(q2[i + j], a[L + i], b[L + j]);
#pragma unroll
    for (int i = 0; i < L; i++)
#pragma unroll
        for (int j = 0; j < L; j++) q1[i + j] = and_xor(q1[i + j], as[i], bs[j]);
}

// p[0..2M-2] = q0 + x^L (q1 + q0 + q2) + x^2L q2 (unreduced).
template <int M>
__device__ __forceinline__ void kara_combine(const uint32_t q0[Kara<M>::NQ0], const uint32_t q1[Kara<M>::NQ0],
                                             const uint32_t q2[Kara<M>::NQ2], uint32_t p[2 * M - 1])
{
    constexpr int L = Kara<M>::L, NQ0 = Kara<M>::NQ0, NQ2 = Kara<M>::NQ2;
    static_assert(2 * L + NQ2 == 2 * M - 1, "the high half product ends at plane 2M-2");
#pragma unroll
    for (int k = 0; k < 2 * M - 1; k++) {
        // terms landing on plane k: q0[k], q2[k-2L], and q1 + q0 + q2 at k-L
        uint32_t t[5];
        int n = 0;
        if (k < NQ0) t[n++] = q0[k];
        if (k >= 2 * L && k - 2 * L < NQ2) t[n++] = q2[k - 2 * L];
        if (k >= L && k - L < NQ0) {
            t[n++] = q1[k - L];
            t[n++] = q0[k - L];
            if (k - L < NQ2) t[n++] = q2[k - L];
        }
        uint32_t v = n > 0 ? t[0] : 0u;
        int i = 1;
#pragma unroll
        for (; i + 1 < n; i += 2) v = xor3(v, t[i], t[i + 1]);
        if (i < n) v ^= t[i];
        p[k] = v;
    }
}

// Folds an unreduced product back to m planes.
template <int M>
__device__ __forceinline__ void bs_fold(uint32_t p[2 * M - 1], uint32_t out[M])
{
    bs_fold_planes<M>(p);
#pragma unroll
    for (int k = 0; k < M; k++) out[k] = p[k];
}

// a <- a * a (squaring is linear: spread then fold).
template <int M>
__device__ __forceinline__ void bs_sq(uint32_t a[M])
{
    constexpr uint32_t low = Modulus<M>::low;
    uint32_t p[2 * M - 1];
#pragma unroll
    for (int s = 0; s < 2 * M - 1; s++) p[s] = (s & 1) ? 0u : a[s / 2];
#pragma unroll
    for (int s = 2 * M - 2; s >= M; s--)
#pragma unroll
        for (int b2 = 0; b2 < 5; b2++)
            if ((low >> b2) & 1u) p[s - M + b2] ^= p[s];
#pragma unroll
    for (int k = 0; k < M; k++) a[k] = p[k];
}

}  // namespace cmd
