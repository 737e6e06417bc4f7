/*
 * oracle/cmref.h -- CPU restatement of the reference Classic McEliece decoder.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests, the smoke check and
 * the CPU-baseline leg of bench.py call it to verify (and time) the product.
 * The product library never links, loads or calls anything under oracle/.
 *
 * Every function restates the algorithm of the reference CPU decoder
 * (CPU/src/common/{gf_13.c,bm.c,root_13.c}, CPU/param/param348864/src/{gf.c,root.c},
 * CPU/param/param<p>/src/util.c, CPU/param/param6960119/src/decrypt.c) for a
 * parameter set chosen at run time.  Citations sit beside each function.
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_gf_mul      exhaustive vs. a carry-less textbook product + reduction,
 *                   field axioms, and vs. the reference build (oracle/_ref).
 *   orc_gf_inv/frac a*inv(a)==1 for all nonzero a, inv(0)==0, vs. oracle/_ref.
 *   orc_bitrev      involution on m bits, vs. oracle/_ref.
 *   orc_apply_benes vs. oracle/_ref on random condition bits; inverse property.
 *   orc_support_gen vs. oracle/_ref on the five shipped secret keys; the
 *                   round trip support_gen(controlbits(pi)) == bitrev(pi).
 *   orc_key_expand  table rows vs. the closed form L^j / g(L)^2 and oracle/_ref.
 *   orc_synd        linearity, unit-vector rows, vs. oracle/_ref.
 *   orc_bm          closed form prod(1 - L_i x) for true errors, vs. oracle/_ref
 *                   on random (non-codeword) syndromes.
 *   orc_root        Horner closed form; end to end recovered e == generated e.
 *   orc_errorstream byte format vs. the reference executable's output file.
 *   orc_synd_full   equals orc_synd (the reference's table) on words inside the first
 *                   sb positions; s(e) == s(c) for every generated ciphertext c of
 *                   error e (the code's defining relation, through the public key).
 *
 * Decapsulation (tests/_oracle.py kem_dec) composes these with SHAKE256 from Python's
 * hashlib (FIPS 202).  Its success key is the reference encapsulation's formula
 * SHAKE256(1 || e || c) (kem/common/core/operations.c:15-34); the rejection key
 * SHAKE256(0 || s || c) and the acceptance test (weight t, equal syndromes) follow the
 * published Classic McEliece specification's crypto_kem_dec, which the reference does not
 * implement: that branch has no reference counterpart to pin it against.
 */
#ifndef CMREF_H
#define CMREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int id;          /* 348864, 460896, 6688128, 6960119, 8192128 */
    int m, n, t;
    int synd_bytes;  /* ceil(m*t/8) == ciphertext bytes */
    int sb;          /* 8 * synd_bytes: syndrome input bits */
    int sk_bytes;
    int cond_bytes;  /* (2m-1) * 2^(m-4) */
} orc_params;

/* Returns 0 and fills *out for a known parameter id, -1 otherwise. */
int orc_param(int id, orc_params *out);

/* GF(2^m) scalar arithmetic, m in {12, 13}. */
uint16_t orc_gf_mul(int m, uint16_t a, uint16_t b);
uint16_t orc_gf_inv(int m, uint16_t a);
uint16_t orc_gf_frac(int m, uint16_t den, uint16_t num); /* num/den */
uint16_t orc_bitrev(int m, uint16_t a);

void orc_gf_mul_batch(int m, const uint16_t *a, const uint16_t *b, uint16_t *out, size_t count);
void orc_gf_inv_batch(int m, const uint16_t *a, uint16_t *out, size_t count);

/* Benes network over 2^m bits, reference bit layout; rev!=0 applies it backwards. */
int orc_apply_benes(int m, uint8_t *r, const uint8_t *cond, int rev);
/* Support L[0..n) from the condition bits of a secret key. */
int orc_support_gen(int id, const uint8_t *cond, uint16_t *L);

/*
 * Key expansion from a raw secret key (sk_bytes): g[t+1] (monic), L[n],
 * and table[sb][2t] with table[i][j] = L_i^j / g(L_i)^2.
 * Any output pointer may be NULL.
 */
int orc_key_expand(int id, const uint8_t *sk, uint16_t *g, uint16_t *L, uint16_t *table);

/* Per-ciphertext stages; `table`, `L` come from orc_key_expand. */
int orc_synd(int id, const uint16_t *table, const uint8_t *ct, uint16_t *s);     /* s[2t] */
int orc_bm(int id, const uint16_t *s, uint16_t *locator);                        /* locator[t+1], reversed */
int orc_root(int id, const uint16_t *locator, const uint16_t *L, uint16_t *images); /* images[n] */
int orc_error_vector(int id, const uint16_t *images, uint8_t *e);                /* e[n/8] */

/*
 * Syndrome of an n-bit word over the whole support (the specification's synd over all n
 * positions): s[j] = XOR over i < n with bit i of r set of L_i^j / g(L_i)^2, j < 2t.
 * g: t+1 coefficients (g[t] = 1), L: n support elements (orc_key_expand).
 */
int orc_synd_full(int id, const uint16_t *g, const uint16_t *L, const uint8_t *r, uint16_t *s);

/* Whole decryption of `count` ciphertexts (ct_stride bytes apart) under one key. */
int orc_decrypt(int id, const uint8_t *sk, const uint8_t *ct, size_t ct_stride,
                size_t count, uint8_t *e_out);

/*
 * Text line of one error vector: " %d" per set bit, ascending, then '\n'.
 * Returns the number of bytes written (<= cap), or -1 if cap is too small.
 */
long orc_errorstream_line(int id, const uint8_t *e, char *buf, size_t cap);

/*
 * Timed single-thread decryption loop (CPU baseline).  Key expansion is
 * outside the timed region, as in the reference driver.  Writes the stage
 * totals in microseconds into stage_us[3] (synd, bm, root) and returns the
 * wall time of the loop in seconds.
 */
double orc_timed_decrypt(int id, const uint8_t *sk, const uint8_t *ct, size_t ct_stride,
                         size_t count, uint8_t *e_out, double *stage_us);

#ifdef __cplusplus
}
#endif

#endif /* CMREF_H */
