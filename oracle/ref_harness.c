/*
 * oracle/ref_harness.c -- thin C entry points over the reference CPU decoder.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
 * the reference's own sources where they lie (CPU/src/common/*.c and
 * CPU/param/param<p>/src/util.c, with CPU/param/param<p>/include/common.h),
 * once per parameter set, into oracle/_ref/libref_<p>.so.  Nothing of the
 * reference is copied here: this file only defines the globals its util.c
 * expects (CPU/param/param6960119/src/decrypt.c:93-100 defines them in the
 * reference driver) and forwards to its functions.
 */
#include <string.h>
#include <time.h>

#include "common.h"
#include "gf.h"
#include "root.h"
#include "bm.h"

gf g[SYS_T + 1];
gf L[SYS_N];
gf e_inv[SYS_N];
gf inverse_elements[sb][2 * SYS_T];

void ref_params(int *out)
{
    out[0] = GFBITS;
    out[1] = SYS_N;
    out[2] = SYS_T;
    out[3] = SYND_BYTES;
    out[4] = sb;
    out[5] = crypto_kem_SECRETKEYBYTES;
    out[6] = COND_BYTES;
}

void ref_gf_mul_batch(const gf *a, const gf *b, gf *out, long count)
{
    for (long i = 0; i < count; i++) out[i] = gf_mul(a[i], b[i]);
}

void ref_gf_inv_batch(const gf *a, gf *out, long count)
{
    for (long i = 0; i < count; i++) out[i] = gf_inv(a[i]);
}

void ref_gf_frac_batch(const gf *den, const gf *num, gf *out, long count)
{
    for (long i = 0; i < count; i++) out[i] = gf_frac(den[i], num[i]);
}

void ref_bitrev_batch(const gf *a, gf *out, long count)
{
    for (long i = 0; i < count; i++) out[i] = bitrev(a[i]);
}

void ref_apply_benes(unsigned char *r, const unsigned char *bits, int rev)
{
    apply_benes(r, bits, rev);
}

void ref_support_gen(const unsigned char *cond, gf *out)
{
    support_gen(out, cond);
}

/* Same steps as initialise_secret_key_state (CPU/param/param6960119/src/util.c:60-72)
 * followed by compute_inverses (util.c:86-100), on an in-memory key. */
void ref_set_key(const unsigned char *sk)
{
    const unsigned char *p = sk + 40;
    for (int i = 0; i < SYS_T; i++, p += 2) g[i] = load_gf(p);
    g[SYS_T] = 1;
    support_gen(L, p);
    compute_inverses();
}

void ref_get_key(gf *g_out, gf *L_out, gf *table_out)
{
    if (g_out) memcpy(g_out, g, sizeof g);
    if (L_out) memcpy(L_out, L, sizeof L);
    if (table_out) memcpy(table_out, inverse_elements, sizeof inverse_elements);
}

/* The driver's per-ciphertext copy and zero tail (decrypt.c:163-164), then synd. */
void ref_synd(const unsigned char *ct, gf *s)
{
    unsigned char r[SYS_N / 8];
    memcpy(r, ct, SYND_BYTES);
    memset(r + SYND_BYTES, 0, sizeof r - SYND_BYTES);
    synd(s, r);
}

void ref_bm(const gf *s, gf *out)
{
    gf tmp[2 * SYS_T];
    memcpy(tmp, s, sizeof tmp);
    bm(out, tmp);
}

void ref_root(const gf *f, gf *images)
{
    gf tmp[SYS_T + 1];
    memcpy(tmp, f, sizeof tmp);
    root(images, tmp, L);
}

static gf h_s[2 * SYS_T], h_loc[SYS_T + 1], h_img[SYS_N];

static void decrypt_one(const unsigned char *ct, unsigned char *e)
{
    ref_synd(ct, h_s);
    bm(h_loc, h_s);
    root(h_img, h_loc, L);
    memset(e, 0, SYS_N / 8);
    for (int i = 0; i < SYS_N; i++)
        if (h_img[i] == 0) e[i >> 3] |= (unsigned char)(1u << (i & 7));
}

/* Decrypt `count` ciphertexts under the key set by ref_set_key. */
void ref_decrypt(const unsigned char *cts, long ct_stride, long count, unsigned char *e_out)
{
    for (long k = 0; k < count; k++) decrypt_one(cts + k * ct_stride, e_out + k * (SYS_N / 8));
}

/* Wall seconds for the loop (key expansion excluded, as in the reference driver). */
double ref_timed_decrypt(const unsigned char *cts, long ct_stride, long count, unsigned char *e_out)
{
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    ref_decrypt(cts, ct_stride, count, e_out);
    clock_gettime(CLOCK_MONOTONIC, &b);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}
