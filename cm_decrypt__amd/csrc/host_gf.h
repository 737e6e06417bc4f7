// Host-side GF(2^m) arithmetic for key expansion and the vector generator.
//
// Field: GF(2^13) mod x^13+x^4+x^3+x+1 (CPU/src/common/gf_13.c:103-108) and
// GF(2^12) mod x^12+x^3+1 (CPU/param/param348864/src/gf.c:116-122).
// The product uses the CPU's carry-less multiply instruction, so its timing
// does not depend on the operands (the reference's gf_mul is table-free for
// the same reason).
#pragma once

#include <cstdint>
#include <immintrin.h>

namespace cmd {

struct HostField {
    int m;
    uint32_t low;   // modulus without x^m
    uint32_t mask;  // 2^m - 1

    explicit HostField(int m_) : m(m_), low(m_ == 13 ? 0x1Bu : 0x9u), mask((1u << m_) - 1u) {}

    static inline uint32_t clmul(uint32_t a, uint32_t b)
    {
        __m128i x = _mm_cvtsi32_si128((int)a);
        __m128i y = _mm_cvtsi32_si128((int)b);
        return (uint32_t)_mm_cvtsi128_si32(_mm_clmulepi64_si128(x, y, 0));
    }

    // Fold bits >= m back with x^m = low(x); two folds suffice for products of
    // two reduced elements for both moduli (their low parts have degree <= 4).
    inline uint16_t reduce(uint32_t v) const
    {
        uint32_t hi = v >> m;
        v = (v & mask) ^ clmul(hi, low);
        hi = v >> m;
        v = (v & mask) ^ clmul(hi, low);
        return (uint16_t)v;
    }

    inline uint16_t mul(uint16_t a, uint16_t b) const { return reduce(clmul(a, b)); }
    inline uint16_t sq(uint16_t a) const { return mul(a, a); }

    // a^(2^m - 2); inv(0) == 0 like the reference's exponent chains.
    inline uint16_t inv(uint16_t a) const
    {
        uint16_t r = 1, b = a;
        uint32_t e = (1u << m) - 2u;
        while (e) {
            if (e & 1u) r = mul(r, b);
            b = mul(b, b);
            e >>= 1;
        }
        return r;
    }

    // bitrev (gf_13.c:8-16 / param348864/src/gf.c:10-18): reverse 16 bits, keep the top m.
    inline uint16_t bitrev(uint16_t a) const
    {
        uint32_t r = 0;
        for (int i = 0; i < 16; i++) r |= ((a >> i) & 1u) << (15 - i);
        return (uint16_t)(r >> (16 - m));
    }

    // Horner with f[deg] leading (root_13.c:216-230).
    inline uint16_t eval(const uint16_t *f, int deg, uint16_t a) const
    {
        uint16_t r = f[deg];
        for (int i = deg - 1; i >= 0; i--) r = (uint16_t)(mul(r, a) ^ f[i]);
        return r;
    }
};

}  // namespace cmd
