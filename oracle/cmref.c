/*
 * oracle/cmref.c -- CPU restatement of the reference Classic McEliece decoder.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * bench.py's CPU-baseline leg use it.  See cmref.h for what pins each part.
 *
 * The reference fixes the parameter set at compile time; this restatement
 * takes it at run time so one library serves all five sets.  Arithmetic and
 * control flow follow the reference line by line where the result depends on
 * it (notably bm(), whose masks decide the output for non-codeword inputs).
 */
#include "cmref.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

/* CPU/param/param<p>/include/common.h: crypto_kem_* sizes, GFBITS, SYS_N, SYS_T. */
static const orc_params k_params[5] = {
    {348864, 12, 3488, 64, 96, 768, 6492, 5888},
    {460896, 13, 4608, 96, 156, 1248, 13608, 12800},
    {6688128, 13, 6688, 128, 208, 1664, 13932, 12800},
    {6960119, 13, 6960, 119, 194, 1552, 13948, 12800},
    {8192128, 13, 8192, 128, 208, 1664, 14120, 12800},
};

int orc_param(int id, orc_params *out)
{
    for (int i = 0; i < 5; i++) {
        if (k_params[i].id == id) {
            if (out) *out = k_params[i];
            return 0;
        }
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* GF(2^m)                                                                   */
/* ------------------------------------------------------------------------ */

/*
 * Product in GF(2^m): carry-less multiply, then reduce.
 * m=13: CPU/src/common/gf_13.c:86-111, modulus x^13+x^4+x^3+x+1 (:103-108).
 * m=12: CPU/param/param348864/src/gf.c:103-125, modulus x^12+x^3+1 (:116-122).
 * Both reference routines reduce completely, so the value is the field product.
 */
uint16_t orc_gf_mul(int m, uint16_t a, uint16_t b)
{
    uint32_t acc = 0;
    for (int i = 0; i < m; i++) {
        uint32_t take = 0u - (((uint32_t)b >> i) & 1u);
        acc ^= ((uint32_t)a << i) & take;
    }
    /* low part of the modulus (without x^m) */
    const uint32_t low = (m == 13) ? 0x1Bu : 0x9u;
    for (int k = 2 * m - 2; k >= m; k--) {
        uint32_t hit = 0u - ((acc >> k) & 1u);
        acc ^= hit & ((low << (k - m)) | (1u << k));
    }
    return (uint16_t)(acc & ((1u << m) - 1u));
}

/*
 * Inverse as a^(2^m - 2) (so inv(0) == 0).
 * m=13: gf_inv -> gf_frac(den, 1) (gf_13.c:136-139, :257-271) computes den^(2^13-2).
 * m=12: gf_inv (param348864/src/gf.c:150-178) computes in^(2^12-2).
 */
uint16_t orc_gf_inv(int m, uint16_t a)
{
    uint16_t result = 1, base = a;
    uint32_t e = (1u << m) - 2u;
    while (e) {
        if (e & 1u) result = orc_gf_mul(m, result, base);
        base = orc_gf_mul(m, base, base);
        e >>= 1;
    }
    return result;
}

/* num/den; gf_frac(den, num) in gf_13.c:257-271 and param348864/src/gf.c:181-184. */
uint16_t orc_gf_frac(int m, uint16_t den, uint16_t num)
{
    return orc_gf_mul(m, orc_gf_inv(m, den), num);
}

/* 16-bit reversal then drop (16-m) bits: gf_13.c:8-16, param348864/src/gf.c:10-18. */
uint16_t orc_bitrev(int m, uint16_t a)
{
    uint16_t r = 0;
    for (int i = 0; i < 16; i++)
        r |= (uint16_t)(((a >> i) & 1u) << (15 - i));
    return (uint16_t)(r >> (16 - m));
}

void orc_gf_mul_batch(int m, const uint16_t *a, const uint16_t *b, uint16_t *out, size_t count)
{
    for (size_t i = 0; i < count; i++) out[i] = orc_gf_mul(m, a[i], b[i]);
}

void orc_gf_inv_batch(int m, const uint16_t *a, uint16_t *out, size_t count)
{
    for (size_t i = 0; i < count; i++) out[i] = orc_gf_inv(m, a[i]);
}

/* ------------------------------------------------------------------------ */
/* Benes network (support generation)                                        */
/* ------------------------------------------------------------------------ */

static uint64_t le64(const uint8_t *p)
{
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}

static uint32_t le32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

static void put64(uint8_t *p, uint64_t v)
{
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}

/* 64x64 bit-matrix transpose: out[i] bit j = in[j] bit i (root_13.c:12-43). */
static void transpose64(uint64_t *out, const uint64_t *in)
{
    uint64_t tmp[64];
    for (int i = 0; i < 64; i++) {
        uint64_t w = 0;
        for (int j = 0; j < 64; j++) w |= ((in[j] >> i) & 1u) << j;
        tmp[i] = w;
    }
    memcpy(out, tmp, sizeof tmp);
}

/* Conditional swap of words w[j], w[j+s] under mask, blocks of 2s over `len` words. */
static void cswap_words(uint64_t *w, int len, int s, const uint64_t *masks)
{
    int k = 0;
    for (int base = 0; base < len; base += 2 * s)
        for (int j = base; j < base + s; j++) {
            uint64_t d = (w[j] ^ w[j + s]) & masks[k++];
            w[j] ^= d;
            w[j + s] ^= d;
        }
}

/*
 * m = 13 network: 8192 bits held as two 64x64 bit matrices.
 * Follows apply_benes in CPU/src/common/root_13.c:95-177:
 *   7 outer layers on the transposed (128-word) form, masks transposed (:119-131),
 *   6 + 5 inner layers on the row form, masks taken as loaded, the two halves
 *   interleaving mask words (layer_in, :47-69; :136-150),
 *   7 outer layers in reverse stride order (:155-167).
 * Each layer consumes 512 condition bytes; rev walks the layers backwards (:107-108).
 */
static void benes13(uint8_t *r, const uint8_t *cond, int rev)
{
    uint64_t rows[128], cols[128], mv[64], mh[64];
    const int layer_bytes = 512;
    int layer = rev ? 24 : 0;
    const int step = rev ? -1 : 1;

    for (int i = 0; i < 64; i++) {
        rows[i] = le64(r + i * 16);
        rows[64 + i] = le64(r + i * 16 + 8);
    }
    transpose64(cols, rows);
    transpose64(cols + 64, rows + 64);

    for (int lg = 0; lg <= 6; lg++, layer += step) {
        for (int i = 0; i < 64; i++) mv[i] = le64(cond + layer * layer_bytes + 8 * i);
        transpose64(mh, mv);
        cswap_words(cols, 128, 1 << lg, mh);
    }
    transpose64(rows, cols);
    transpose64(rows + 64, cols + 64);

    static const int inner_lg[11] = {0, 1, 2, 3, 4, 5, 4, 3, 2, 1, 0};
    for (int q = 0; q < 11; q++, layer += step) {
        for (int i = 0; i < 64; i++) mv[i] = le64(cond + layer * layer_bytes + 8 * i);
        /* the two 64-word halves share one layer; mask words alternate between them */
        int s = 1 << inner_lg[q], k = 0;
        for (int base = 0; base < 64; base += 2 * s)
            for (int j = base; j < base + s; j++)
                for (int h = 0; h < 2; h++) {
                    uint64_t *w = rows + 64 * h;
                    uint64_t d = (w[j] ^ w[j + s]) & mv[k++];
                    w[j] ^= d;
                    w[j + s] ^= d;
                }
    }

    transpose64(cols, rows);
    transpose64(cols + 64, rows + 64);
    for (int lg = 6; lg >= 0; lg--, layer += step) {
        for (int i = 0; i < 64; i++) mv[i] = le64(cond + layer * layer_bytes + 8 * i);
        transpose64(mh, mv);
        cswap_words(cols, 128, 1 << lg, mh);
    }
    transpose64(rows, cols);
    transpose64(rows + 64, cols + 64);

    for (int i = 0; i < 64; i++) {
        put64(r + i * 16, rows[i]);
        put64(r + i * 16 + 8, rows[64 + i]);
    }
}

/*
 * m = 12 network: 4096 bits as one 64x64 matrix.
 * Follows apply_benes in CPU/param/param348864/src/root.c:253-332:
 *   6 outer layers on the transposed form with 32-bit mask words (load4) transposed
 *   (:289-299), 6 + 5 inner layers with 64-bit mask words (:301-312),
 *   6 outer layers in reverse (:314-324).  Each layer consumes 256 bytes.
 */
static void benes12(uint8_t *r, const uint8_t *cond, int rev)
{
    uint64_t bs[64], m[64];
    const int layer_bytes = 256;
    int layer = rev ? 22 : 0;
    const int step = rev ? -1 : 1;

    for (int i = 0; i < 64; i++) bs[i] = le64(r + 8 * i);
    transpose64(bs, bs);
    for (int lg = 0; lg <= 5; lg++, layer += step) {
        for (int i = 0; i < 64; i++) m[i] = le32(cond + layer * layer_bytes + 4 * i);
        transpose64(m, m);
        cswap_words(bs, 64, 1 << lg, m);
    }
    transpose64(bs, bs);
    for (int lg = 0; lg <= 5; lg++, layer += step) {
        for (int i = 0; i < 32; i++) m[i] = le64(cond + layer * layer_bytes + 8 * i);
        cswap_words(bs, 64, 1 << lg, m);
    }
    for (int lg = 4; lg >= 0; lg--, layer += step) {
        for (int i = 0; i < 32; i++) m[i] = le64(cond + layer * layer_bytes + 8 * i);
        cswap_words(bs, 64, 1 << lg, m);
    }
    transpose64(bs, bs);
    for (int lg = 5; lg >= 0; lg--, layer += step) {
        for (int i = 0; i < 64; i++) m[i] = le32(cond + layer * layer_bytes + 4 * i);
        transpose64(m, m);
        cswap_words(bs, 64, 1 << lg, m);
    }
    transpose64(bs, bs);
    for (int i = 0; i < 64; i++) put64(r + 8 * i, bs[i]);
}

int orc_apply_benes(int m, uint8_t *r, const uint8_t *cond, int rev)
{
    if (m == 13) benes13(r, cond, rev);
    else if (m == 12) benes12(r, cond, rev);
    else return -1;
    return 0;
}

/*
 * support_gen: root_13.c:181-211 (m=13), param348864/src/root.c:336-385 (m=12).
 * Bit-plane j of bitrev(i) for all i < 2^m goes through the network; L[i]
 * gathers bit i of each plane.
 */
int orc_support_gen(int id, const uint8_t *cond, uint16_t *L)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    const int q = 1 << p.m, bytes = q / 8;
    uint8_t *planes = (uint8_t *)calloc((size_t)p.m * bytes, 1);
    if (!planes) return -2;
    for (int i = 0; i < q; i++) {
        uint16_t a = orc_bitrev(p.m, (uint16_t)i);
        for (int j = 0; j < p.m; j++)
            planes[j * bytes + i / 8] |= (uint8_t)(((a >> j) & 1u) << (i % 8));
    }
    for (int j = 0; j < p.m; j++) orc_apply_benes(p.m, planes + j * bytes, cond, 0);
    for (int i = 0; i < p.n; i++) {
        uint16_t v = 0;
        for (int j = p.m - 1; j >= 0; j--)
            v = (uint16_t)((v << 1) | ((planes[j * bytes + i / 8] >> (i % 8)) & 1u));
        L[i] = v;
    }
    free(planes);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Key expansion and decoding stages                                          */
/* ------------------------------------------------------------------------ */

/* eval: Horner with f[t] leading (root_13.c:216-230, param348864/src/root.c:400-415). */
static uint16_t eval_poly(int m, int t, const uint16_t *f, uint16_t a)
{
    uint16_t r = f[t];
    for (int i = t - 1; i >= 0; i--) r = (uint16_t)(orc_gf_mul(m, r, a) ^ f[i]);
    return r;
}

/* load_gf: little-endian 16 bits masked to m (gf_13.c:17-27). */
static uint16_t load_elem(int m, const uint8_t *src)
{
    return (uint16_t)((src[0] | (src[1] << 8)) & ((1 << m) - 1));
}

/*
 * initialise_secret_key_state (CPU/param/param6960119/src/util.c:60-72): skip the
 * 40-byte prefix, g[i] = load_gf x t, g[t] = 1, support_gen on what follows.
 * compute_inverses (util.c:86-100): table[i][0] = 1/g(L_i)^2, table[i][j] = table[i][j-1]*L_i
 * for i < sb (the table spans all sb syndrome bits, not only mt).
 */
int orc_key_expand(int id, const uint8_t *sk, uint16_t *g_out, uint16_t *L_out, uint16_t *table)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    uint16_t *g = (uint16_t *)malloc(sizeof(uint16_t) * (p.t + 1));
    uint16_t *L = (uint16_t *)malloc(sizeof(uint16_t) * p.n);
    if (!g || !L) { free(g); free(L); return -2; }
    const uint8_t *ptr = sk + 40;
    for (int i = 0; i < p.t; i++, ptr += 2) g[i] = load_elem(p.m, ptr);
    g[p.t] = 1;
    orc_support_gen(id, ptr, L);
    if (table) {
        const int w = 2 * p.t;
        for (int i = 0; i < p.sb; i++) {
            uint16_t v = eval_poly(p.m, p.t, g, L[i]);
            uint16_t inv = orc_gf_inv(p.m, orc_gf_mul(p.m, v, v));
            table[(size_t)i * w] = inv;
            for (int j = 1; j < w; j++)
                table[(size_t)i * w + j] = orc_gf_mul(p.m, table[(size_t)i * w + j - 1], L[i]);
        }
    }
    if (g_out) memcpy(g_out, g, sizeof(uint16_t) * (p.t + 1));
    if (L_out) memcpy(L_out, L, sizeof(uint16_t) * p.n);
    free(g);
    free(L);
    return 0;
}

/*
 * synd (util.c:102-114): s[j] = XOR over bits i < sb of table[i][j] * bit_i, bit i of the
 * ciphertext taken LSB first.  The driver zero-fills past SYND_BYTES (decrypt.c:163-164);
 * sb == 8*SYND_BYTES so every bit read here lies inside the ciphertext.
 */
int orc_synd(int id, const uint16_t *table, const uint8_t *ct, uint16_t *s)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    const int w = 2 * p.t;
    memset(s, 0, sizeof(uint16_t) * w);
    for (int i = 0; i < p.sb; i++) {
        uint16_t c = (uint16_t)((ct[i >> 3] >> (i & 7)) & 1u);
        for (int j = 0; j < w; j++) s[j] ^= orc_gf_mul(p.m, table[(size_t)i * w + j], c);
    }
    return 0;
}

int orc_synd_full(int id, const uint16_t *g, const uint16_t *L, const uint8_t *r, uint16_t *s)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    const int w = 2 * p.t;
    memset(s, 0, sizeof(uint16_t) * w);
    for (int i = 0; i < p.n; i++) {
        if (!((r[i >> 3] >> (i & 7)) & 1u)) continue;
        uint16_t v = eval_poly(p.m, p.t, g, L[i]);
        uint16_t c = orc_gf_inv(p.m, orc_gf_mul(p.m, v, v));
        for (int j = 0; j < w; j++) {
            s[j] ^= c;
            c = orc_gf_mul(p.m, c, L[i]);
        }
    }
    return 0;
}

/*
 * Constant-time Berlekamp-Massey, CPU/src/common/bm.c:12-68, step for step:
 * C = 1, B = x, b = 1, L = 0; for N < 2t: d = sum_{i<=min(N,t)} C_i s_{N-i};
 * mne = (d != 0), mle = (2L <= N) & mne as 16-bit all-ones masks (:42-44);
 * C ^= (d/b) B & mne; L, B, b conditionally replaced; B <- x B.  Output reversed.
 */
int orc_bm(int id, const uint16_t *s, uint16_t *out)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    const int t = p.t, m = p.m;
    uint16_t C[257], B[257], T[257];
    uint16_t L = 0, b = 1;
    for (int i = 0; i <= t; i++) C[i] = B[i] = 0;
    C[0] = 1;
    B[1] = 1;
    for (int N = 0; N < 2 * t; N++) {
        uint16_t d = 0;
        int top = N < t ? N : t;
        for (int i = 0; i <= top; i++) d ^= orc_gf_mul(m, C[i], s[N - i]);
        uint16_t mne = (uint16_t)(((uint16_t)(d - 1u) >> 15) - 1u);
        uint16_t mle = (uint16_t)(((uint16_t)(N - 2 * L) >> 15) - 1u);
        mle &= mne;
        memcpy(T, C, sizeof(uint16_t) * (t + 1));
        uint16_t f = orc_gf_frac(m, b, d);
        for (int i = 0; i <= t; i++) C[i] ^= orc_gf_mul(m, f, B[i]) & mne;
        L = (uint16_t)((L & ~mle) | ((N + 1 - L) & mle));
        for (int i = 0; i <= t; i++) B[i] = (uint16_t)((B[i] & ~mle) | (T[i] & mle));
        b = (uint16_t)((b & ~mle) | (d & mle));
        for (int i = t; i >= 1; i--) B[i] = B[i - 1];
        B[0] = 0;
    }
    for (int i = 0; i <= t; i++) out[i] = C[t - i];
    return 0;
}

/* root (root_13.c:233-240): images[i] = eval(locator, L[i]). */
int orc_root(int id, const uint16_t *locator, const uint16_t *L, uint16_t *images)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    for (int i = 0; i < p.n; i++) images[i] = eval_poly(p.m, p.t, locator, L[i]);
    return 0;
}

/* build_error_vector (CPU/param/param6960119/src/decrypt.c:43-53). */
int orc_error_vector(int id, const uint16_t *images, uint8_t *e)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    memset(e, 0, (size_t)p.n / 8);
    for (int i = 0; i < p.n; i++)
        if (images[i] == 0) e[i >> 3] |= (uint8_t)(1u << (i & 7));
    return 0;
}

/* write_error_positions (decrypt.c:55-66): " %d" per set bit, then '\n'. */
long orc_errorstream_line(int id, const uint8_t *e, char *buf, size_t cap)
{
    orc_params p;
    if (orc_param(id, &p)) return -1;
    size_t pos = 0;
    for (int i = 0; i < p.n; i++) {
        if (!(e[i >> 3] & (1u << (i & 7)))) continue;
        char tmp[16];
        int len = 0, v = i;
        char digits[8];
        int nd = 0;
        do { digits[nd++] = (char)('0' + v % 10); v /= 10; } while (v);
        tmp[len++] = ' ';
        while (nd) tmp[len++] = digits[--nd];
        if (pos + (size_t)len > cap) return -1;
        memcpy(buf + pos, tmp, (size_t)len);
        pos += (size_t)len;
    }
    if (pos + 1 > cap) return -1;
    buf[pos++] = '\n';
    return (long)pos;
}

typedef struct {
    orc_params p;
    uint16_t *L, *table, *s, *loc, *img;
} orc_ws;

static int ws_open(orc_ws *w, int id, const uint8_t *sk)
{
    memset(w, 0, sizeof *w);
    if (orc_param(id, &w->p)) return -1;
    w->L = (uint16_t *)malloc(sizeof(uint16_t) * w->p.n);
    w->table = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)w->p.sb * 2 * w->p.t);
    w->s = (uint16_t *)malloc(sizeof(uint16_t) * 2 * w->p.t);
    w->loc = (uint16_t *)malloc(sizeof(uint16_t) * (w->p.t + 1));
    w->img = (uint16_t *)malloc(sizeof(uint16_t) * w->p.n);
    if (!w->L || !w->table || !w->s || !w->loc || !w->img) return -2;
    return orc_key_expand(id, sk, NULL, w->L, w->table);
}

static void ws_close(orc_ws *w)
{
    free(w->L); free(w->table); free(w->s); free(w->loc); free(w->img);
}

int orc_decrypt(int id, const uint8_t *sk, const uint8_t *ct, size_t ct_stride,
                size_t count, uint8_t *e_out)
{
    orc_ws w;
    int rc = ws_open(&w, id, sk);
    if (rc) { ws_close(&w); return rc; }
    for (size_t k = 0; k < count; k++) {
        orc_synd(id, w.table, ct + k * ct_stride, w.s);
        orc_bm(id, w.s, w.loc);
        orc_root(id, w.loc, w.L, w.img);
        orc_error_vector(id, w.img, e_out + k * (size_t)(w.p.n / 8));
    }
    ws_close(&w);
    return 0;
}

static double usec_between(const struct timespec *a, const struct timespec *b)
{
    return 1.0e6 * (double)(b->tv_sec - a->tv_sec) + 1.0e-3 * (double)(b->tv_nsec - a->tv_nsec);
}

/* Timing follows CPU/param/param6960119/src/decrypt.c:145-201 (CLOCK_MONOTONIC per stage). */
double orc_timed_decrypt(int id, const uint8_t *sk, const uint8_t *ct, size_t ct_stride,
                         size_t count, uint8_t *e_out, double *stage_us)
{
    orc_ws w;
    if (ws_open(&w, id, sk)) { ws_close(&w); return -1.0; }
    double acc[3] = {0, 0, 0};
    struct timespec w0, w1, t0, t1, t2, t3;
    clock_gettime(CLOCK_MONOTONIC, &w0);
    for (size_t k = 0; k < count; k++) {
        clock_gettime(CLOCK_MONOTONIC, &t0);
        orc_synd(id, w.table, ct + k * ct_stride, w.s);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        orc_bm(id, w.s, w.loc);
        clock_gettime(CLOCK_MONOTONIC, &t2);
        orc_root(id, w.loc, w.L, w.img);
        clock_gettime(CLOCK_MONOTONIC, &t3);
        orc_error_vector(id, w.img, e_out + k * (size_t)(w.p.n / 8));
        acc[0] += usec_between(&t0, &t1);
        acc[1] += usec_between(&t1, &t2);
        acc[2] += usec_between(&t2, &t3);
    }
    clock_gettime(CLOCK_MONOTONIC, &w1);
    if (stage_us) memcpy(stage_us, acc, sizeof acc);
    ws_close(&w);
    return usec_between(&w0, &w1) * 1e-6;
}
