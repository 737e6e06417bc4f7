// Seeded test-vector generator: key pairs and ciphertexts in the reference's formats.
//
// Formats follow kem/common/core/operations.c:36-107 (secret-key layout) and
// encrypt.c:40-120 (ciphertext = systematic parity check times e).  The randomness is
// this library's own seeded xoshiro256** stream, not the NIST AES DRBG.
//
//  * Goppa polynomial: minimal polynomial over GF(2^m) of a random element of
//    GF(2^m)[y]/F(y) (the method of kem/common/core/sk_gen.c), with the extension
//    moduli F of the five sets (kem/variant_<p>/gf.c GF_mul reductions).
//  * Support: a seeded permutation pi of GF(2^m); L_i = bitrev(pi_i) (kem/common/core/pk_gen.c:69-70).
//  * Condition bits: Benes routing of pi by the looping algorithm; checked by re-applying
//    the network (as controlbitsfrompermutation does, kem/common/core/controlbits.c:186-218).
//  * Ciphertext: with H the m*t x n binary parity-check matrix of (g, L) and A its first
//    m*t columns, ct = A^-1 H e.  Then v = (ct, 0...0) satisfies H v = H e, which is what
//    the systematic-form syndrome (encrypt.c:87-120) gives.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/cm_decrypt.h"
#include "engine.h"
#include "host_gf.h"
#include "keys.h"
#include "params.h"

namespace cmd {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        for (int i = 0; i < 4; i++) {  // splitmix64 expansion
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    uint32_t below(uint32_t bound)  // unbiased in [0, bound)
    {
        uint64_t lim = (~0ull) - ((~0ull) % bound);
        uint64_t x;
        do { x = next(); } while (x >= lim);
        return (uint32_t)(x % bound);
    }
};

// Extension modulus F(y) = y^t + sum taps (coefficient, power), kem/variant_<p>/gf.c.
struct ExtTap { int power; uint16_t coef; };
static void ext_taps(int pidx, std::vector<ExtTap> &taps)
{
    switch (kParams[pidx].id) {
    case 348864: taps = {{3, 1}, {1, 1}, {0, 2}}; break;              // y^64 + y^3 + y + z
    case 460896: taps = {{10, 1}, {9, 1}, {6, 1}, {0, 1}}; break;      // y^96 + y^10 + y^9 + y^6 + 1
    case 6960119: taps = {{8, 1}, {0, 1}}; break;                       // y^119 + y^8 + 1
    default: taps = {{7, 1}, {2, 1}, {1, 1}, {0, 1}}; break;           // y^128 + y^7 + y^2 + y + 1
    }
}

// Carry-less products of one element with two packed in 32-bit halves (both fit: 2m-1 < 32).
static inline uint64_t clmul_pair(uint32_t a, uint64_t b2)
{
    return (uint64_t)_mm_cvtsi128_si64(_mm_clmulepi64_si128(_mm_cvtsi64_si128((long long)a),
                                                            _mm_cvtsi64_si128((long long)b2), 0));
}

// Both halves of a clmul_pair result reduced modulo the field polynomial (two folds, as
// HostField::reduce).
static inline uint64_t reduce_pair(const HostField &F, uint64_t v)
{
    const uint64_t mask2 = (uint64_t)F.mask | ((uint64_t)F.mask << 32);
    for (int r = 0; r < 2; r++) {
        const uint64_t hi = (v >> F.m) & mask2;
        v = (v & mask2) ^ clmul_pair(F.low, hi);
    }
    return v;
}

// Folds the terms y^(2t-2) .. y^t of a reduced product with F(y) = y^t + taps.
static void ext_fold(const HostField &F, int t, const std::vector<ExtTap> &taps, uint16_t *prod)
{
    for (int i = 2 * t - 2; i >= t; i--) {
        uint16_t c = prod[i];
        if (!c) continue;
        for (const ExtTap &tp : taps) prod[i - t + tp.power] ^= F.mul(c, tp.coef);
    }
}

// out = a * b mod F in GF(2^m)[y]/F(y), t coefficients each: the coefficient products are
// summed unreduced (two per carry-less multiply) and each sum reduced once.
static void ext_mul(const HostField &F, int t, const std::vector<ExtTap> &taps, const uint16_t *a,
                    const uint16_t *b, uint16_t *out)
{
    uint32_t acc[2 * kMaxT] = {};
    uint64_t b2[kMaxT / 2 + 1];
    for (int j = 0; j < t; j += 2) b2[j / 2] = b[j] | (j + 1 < t ? (uint64_t)b[j + 1] << 32 : 0ull);
    for (int i = 0; i < t; i++) {
        if (!a[i]) continue;
        for (int j = 0; j < t; j += 2) {
            const uint64_t pr = clmul_pair(a[i], b2[j / 2]);
            acc[i + j] ^= (uint32_t)pr;
            acc[i + j + 1] ^= (uint32_t)(pr >> 32);
        }
    }
    uint16_t prod[2 * kMaxT];
    for (int s = 0; s < 2 * t - 1; s++) prod[s] = F.reduce(acc[s]);
    ext_fold(F, t, taps, prod);
    std::memcpy(out, prod, sizeof(uint16_t) * t);
}

// out = a^2 mod F: squaring is linear in characteristic 2, (sum a_i y^i)^2 = sum a_i^2 y^2i.
static void ext_sq(const HostField &F, int t, const std::vector<ExtTap> &taps, const uint16_t *a, uint16_t *out)
{
    uint16_t prod[2 * kMaxT] = {};
    for (int i = 0; i < t; i++) prod[2 * i] = F.sq(a[i]);
    ext_fold(F, t, taps, prod);
    std::memcpy(out, prod, sizeof(uint16_t) * t);
}

// Minimal polynomial of f (t coefficients) over GF(2^m): g monic of degree t with
// f^t + sum_{j<t} g_j f^j = 0.  Returns false when the powers are dependent.  The columns
// f^0 .. f^t (even powers by squaring an earlier column) form the system sum_j g_j f^j = f^t,
// solved by forward elimination with the first nonzero pivot of each column and back
// substitution; its solution is unique, so it is the same g as any other elimination order.
static bool minimal_poly(int pidx, const uint16_t *f, uint16_t *g)
{
    const int t = kParams[pidx].t, W = t + 1;
    HostField F(kParams[pidx].m);
    std::vector<ExtTap> taps;
    ext_taps(pidx, taps);
    // pw[j] = f^j (t coefficients each); M[row r][col j] = (f^j)_r
    std::vector<uint16_t> pw((size_t)W * t, 0);
    pw[0] = 1;
    for (int j = 1; j <= t; j++) {
        if (j % 2 == 0)
            ext_sq(F, t, taps, &pw[(size_t)(j / 2) * t], &pw[(size_t)j * t]);
        else
            ext_mul(F, t, taps, &pw[(size_t)(j - 1) * t], f, &pw[(size_t)j * t]);
    }
    std::vector<uint16_t> M((size_t)t * W);
    for (int j = 0; j <= t; j++)
        for (int r = 0; r < t; r++) M[(size_t)r * W + j] = pw[(size_t)j * t + r];
    std::vector<uint64_t> prow2(W / 2 + 1);
    for (int c = 0; c < t; c++) {
        int piv = -1;
        for (int r = c; r < t; r++)
            if (M[(size_t)r * W + c]) { piv = r; break; }
        if (piv < 0) return false;
        uint16_t *pr = &M[(size_t)c * W];
        if (piv != c)
            for (int k = c; k < W; k++) std::swap(M[(size_t)piv * W + k], pr[k]);
        const uint16_t inv = F.inv(pr[c]);
        for (int k = c; k < W; k++) pr[k] = F.mul(pr[k], inv);
        // the pivot row from column c on, two entries per word
        const int k0 = c, nk = W - c;
        for (int q = 0; q < nk; q += 2)
            prow2[q / 2] = pr[k0 + q] | (q + 1 < nk ? (uint64_t)pr[k0 + q + 1] << 32 : 0ull);
        for (int r = c + 1; r < t; r++) {
            uint16_t *row = &M[(size_t)r * W];
            const uint16_t fac = row[c];
            if (!fac) continue;
            for (int q = 0; q < nk; q += 2) {
                const uint64_t v = reduce_pair(F, clmul_pair(fac, prow2[q / 2]));
                row[k0 + q] ^= (uint16_t)v;
                if (q + 1 < nk) row[k0 + q + 1] ^= (uint16_t)(v >> 32);
            }
        }
    }
    // back substitution (unit diagonal): g_r = M[r][t] + sum_{k>r} M[r][k] g_k
    for (int r = t - 1; r >= 0; r--) {
        uint32_t acc = 0;
        for (int k = r + 1; k < t; k++) acc ^= HostField::clmul(M[(size_t)r * W + k], g[k]);
        g[r] = (uint16_t)(M[(size_t)r * W + t] ^ F.reduce(acc));
    }
    return true;
}

// ---- Benes routing -------------------------------------------------------------------

static inline void set_cond_bit(uint8_t *cond, int m, int layer, int p, int e)
{
    // switch index of the pair (p, p + 2^e) with bit e of p clear
    const int idx = ((p >> (e + 1)) << e) | (p & ((1 << e) - 1));
    const size_t bit = (size_t)layer * ((size_t)1 << (m - 1)) + (size_t)idx;
    cond[bit >> 3] |= (uint8_t)(1u << (bit & 7));
}

// Routes pi (size 2^w, output i receives input pi[i]) through the sub-network at depth d
// whose global positions are (q << d) | c.
static void route(int m, int w, int d, int c, const std::vector<int> &pi, uint8_t *cond)
{
    const int n = 1 << w;
    const int first = d, last = 2 * m - 2 - d;
    if (w == 1) {
        if (pi[0] == 1) set_cond_bit(cond, m, first, c, d);
        return;
    }
    std::vector<int> where(n);  // where[v] = output position of value v
    for (int i = 0; i < n; i++) where[pi[i]] = i;
    std::vector<int> color(n, -1);  // sub-network of each value
    for (int start = 0; start < n; start++) {
        if (color[start] >= 0) continue;
        int v = start, col = 0;
        while (color[v] < 0) {
            color[v] = col;
            int u = pi[where[v] ^ 1];  // shares v's output pair
            color[u] = col ^ 1;
            v = u ^ 1;                 // shares u's input pair
        }
    }
    std::vector<int> sub0(n / 2), sub1(n / 2);
    for (int a = 0; a < n / 2; a++)
        if (color[2 * a] == 1) set_cond_bit(cond, m, first, ((2 * a) << d) | c, d);
    for (int b = 0; b < n / 2; b++) {
        int v0 = pi[2 * b], v1 = pi[2 * b + 1];
        if (color[v0] == 1) set_cond_bit(cond, m, last, ((2 * b) << d) | c, d);
        int in0 = color[v0] == 0 ? v0 : v1;
        int in1 = color[v0] == 0 ? v1 : v0;
        sub0[b] = in0 >> 1;
        sub1[b] = in1 >> 1;
    }
    route(m, w - 1, d + 1, c, sub0, cond);
    route(m, w - 1, d + 1, c | (1 << d), sub1, cond);
}

// Condition bits (cond_bytes) for permutation pi of 2^m; verified by re-applying them.
static bool controlbits(int m, const std::vector<int> &pi, uint8_t *cond, size_t cond_bytes)
{
    std::memset(cond, 0, cond_bytes);
    route(m, m, 0, 0, pi, cond);
    std::vector<uint16_t> check((size_t)1 << m);
    for (size_t i = 0; i < check.size(); i++) check[i] = (uint16_t)i;
    benes_apply(m, check.data(), cond);
    for (size_t i = 0; i < check.size(); i++)
        if (check[i] != pi[i]) return false;
    return true;
}

// ---- encryption ------------------------------------------------------------------------

struct Encryptor {
    int pidx = -1, mt = 0, words = 0;
    std::vector<uint64_t> cols;  // n columns of H, mt bits each
    std::vector<uint64_t> ainv;  // mt rows of A^-1, mt bits each
};

// Builds H (rows i*m + k: bit k of L_j^i / g(L_j)) and A^-1.  False when A is singular.
static bool encryptor_init(int pidx, const uint16_t *g, const uint16_t *L, Encryptor &E)
{
    const ParamDesc &p = kParams[pidx];
    HostField F(p.m);
    const int mt = p.m * p.t, words = (mt + 63) / 64;
    E.pidx = pidx;
    E.mt = mt;
    E.words = words;
    E.cols.assign((size_t)p.n * words, 0);
    for (int j = 0; j < p.n; j++) {
        uint16_t v = F.inv(F.eval(g, p.t, L[j]));
        uint64_t *col = &E.cols[(size_t)j * words];
        for (int i = 0; i < p.t; i++) {
            for (int k = 0; k < p.m; k++) {
                int r = i * p.m + k;
                col[r >> 6] |= (uint64_t)((v >> k) & 1u) << (r & 63);
            }
            v = F.mul(v, L[j]);
        }
    }
    // [A | I] as rows of 2*words words; A[r][c] = bit r of column c.
    const int w2 = 2 * words;
    std::vector<uint64_t> aug((size_t)mt * w2, 0);
    for (int c = 0; c < mt; c++) {
        const uint64_t *col = &E.cols[(size_t)c * words];
        for (int r = 0; r < mt; r++)
            if ((col[r >> 6] >> (r & 63)) & 1u) aug[(size_t)r * w2 + (c >> 6)] |= 1ull << (c & 63);
    }
    for (int r = 0; r < mt; r++) aug[(size_t)r * w2 + words + (r >> 6)] |= 1ull << (r & 63);
    for (int c = 0; c < mt; c++) {
        int piv = -1;
        for (int r = c; r < mt; r++)
            if ((aug[(size_t)r * w2 + (c >> 6)] >> (c & 63)) & 1u) { piv = r; break; }
        if (piv < 0) return false;
        if (piv != c)
            for (int k = 0; k < w2; k++) std::swap(aug[(size_t)piv * w2 + k], aug[(size_t)c * w2 + k]);
        const uint64_t *prow = &aug[(size_t)c * w2];
        for (int r = 0; r < mt; r++) {
            if (r == c) continue;
            uint64_t *row = &aug[(size_t)r * w2];
            if ((row[c >> 6] >> (c & 63)) & 1u)
                for (int k = c >> 6; k < w2; k++) row[k] ^= prow[k];
        }
    }
    E.ainv.assign((size_t)mt * words, 0);
    for (int r = 0; r < mt; r++)
        std::memcpy(&E.ainv[(size_t)r * words], &aug[(size_t)r * w2 + words], sizeof(uint64_t) * words);
    return true;
}

// Error positions of one ciphertext: the first `weight` entries of a partial Fisher-Yates
// shuffle of 0 .. n-1 (entry i is final once step i is done).  The shuffled array is virtual:
// entries moved in this draw are marked with the draw's epoch, the rest equal their index.
struct ShuffleScratch {
    std::vector<int> val;
    std::vector<uint32_t> epoch_of;
    uint32_t epoch = 0;
};

static void draw_positions(Rng &rng, int n, int weight, uint16_t *out, ShuffleScratch &sc)
{
    if ((int)sc.val.size() < n) {
        sc.val.assign(n, 0);
        sc.epoch_of.assign(n, 0);
        sc.epoch = 0;
    }
    const uint32_t ep = ++sc.epoch;
    auto get = [&](int i) { return sc.epoch_of[i] == ep ? sc.val[i] : i; };
    auto put = [&](int i, int v) {
        sc.val[i] = v;
        sc.epoch_of[i] = ep;
    };
    for (int i = 0; i < weight; i++) {
        const int j = i + (int)rng.below((uint32_t)(n - i));
        const int vi = get(i), vj = get(j);
        put(i, vj);
        put(j, vi);
        out[i] = (uint16_t)vj;
    }
}

static void encrypt_one(const Encryptor &E, Rng &rng, int weight, uint8_t *ct, uint8_t *e, ShuffleScratch &sc)
{
    const ParamDesc &p = kParams[E.pidx];
    std::vector<uint16_t> pos((size_t)weight);
    draw_positions(rng, p.n, weight, pos.data(), sc);
    std::memset(e, 0, (size_t)p.n / 8);
    std::vector<uint64_t> y(E.words, 0);
    for (int i = 0; i < weight; i++) {
        int j = pos[i];
        e[j >> 3] |= (uint8_t)(1u << (j & 7));
        const uint64_t *col = &E.cols[(size_t)j * E.words];
        for (int k = 0; k < E.words; k++) y[k] ^= col[k];
    }
    std::memset(ct, 0, (size_t)p.synd_bytes);
    for (int r = 0; r < E.mt; r++) {
        const uint64_t *row = &E.ainv[(size_t)r * E.words];
        uint64_t acc = 0;
        for (int k = 0; k < E.words; k++) acc ^= row[k] & y[k];
        if (__builtin_parityll(acc)) ct[r >> 3] |= (uint8_t)(1u << (r & 7));
    }
}

// ---- key pairs -------------------------------------------------------------------------

// The randomness of one key-pair attempt up to its support: the Goppa polynomial (minimal
// polynomial of a random element of the extension) and the support permutation, with the
// random stream positioned after the permutation for the key's remaining bytes.
struct KeyCandidate {
    Rng rng{0};
    std::vector<uint16_t> g;  // t+1 coefficients, monic
    std::vector<int> pi;      // permutation of 0 .. 2^m - 1; L_i = bitrev(pi_i)
    bool poly_ok = false;     // false: the element's powers are dependent, no key from this seed
};

static void draw_candidate(int pidx, uint64_t seed, KeyCandidate &c)
{
    const ParamDesc &p = kParams[pidx];
    HostField F(p.m);
    c.rng = Rng(seed ^ ((uint64_t)p.id << 32));
    std::vector<uint16_t> f(p.t);
    c.g.assign(p.t + 1, 0);
    for (int i = 0; i < p.t; i++) f[i] = (uint16_t)(c.rng.next() & F.mask);
    c.poly_ok = minimal_poly(pidx, f.data(), c.g.data());
    if (!c.poly_ok) return;
    c.g[p.t] = 1;
    const int q = 1 << p.m;
    c.pi.resize(q);
    for (int i = 0; i < q; i++) c.pi[i] = i;
    for (int i = q - 1; i > 0; i--) std::swap(c.pi[i], c.pi[c.rng.below((uint32_t)(i + 1))]);
}

// The secret key of an accepted candidate (operations.c:62-101 layout): 32 random bytes,
// the 8-byte pivot prefix, g, the condition bits, then n/8 random bytes of s.
static void write_sk(int pidx, KeyCandidate &c, const uint8_t *cond, uint8_t *sk)
{
    const ParamDesc &p = kParams[pidx];
    uint8_t *w = sk;
    for (int i = 0; i < 32; i++) w[i] = (uint8_t)(c.rng.next() & 0xFF);
    w += 32;
    const uint8_t pivots[8] = {0xFF, 0xFF, 0xFF, 0xFF, 0, 0, 0, 0};  // operations.c:101
    std::memcpy(w, pivots, 8);
    w += 8;
    for (int i = 0; i < p.t; i++) { w[2 * i] = (uint8_t)(c.g[i] & 0xFF); w[2 * i + 1] = (uint8_t)(c.g[i] >> 8); }
    w += 2 * p.t;
    std::memcpy(w, cond, (size_t)p.cond_bytes);
    w += p.cond_bytes;
    for (int i = 0; i < p.n / 8; i++) w[i] = (uint8_t)(c.rng.next() & 0xFF);
}

// One attempt per seed value; returns false when this seed gives no usable key.
static bool keypair_attempt(int pidx, uint64_t seed, uint8_t *sk)
{
    const ParamDesc &p = kParams[pidx];
    KeyCandidate c;
    draw_candidate(pidx, seed, c);
    if (!c.poly_ok) return false;
    std::vector<uint8_t> cond(p.cond_bytes);
    if (!controlbits(p.m, c.pi, cond.data(), cond.size())) return false;
    std::vector<uint16_t> L(p.n);
    support_gen(pidx, cond.data(), L.data());
    Encryptor E;
    if (!encryptor_init(pidx, c.g.data(), L.data(), E)) return false;
    write_sk(pidx, c, cond.data(), sk);
    return true;
}

// Seed of attempt a of a key whose base seed is s (cmd_gen_keypair's sequence).
static inline uint64_t attempt_seed(uint64_t s, int a) { return s + ((uint64_t)a << 40); }
constexpr int kMaxAttempts = 64;

}  // namespace cmd

using namespace cmd;

extern "C" int cmd_gen_keypair(int32_t param_id, uint64_t seed, uint8_t *sk)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return CMD_ERR_PARAM;
    if (!sk) return CMD_ERR_ARG;
    // Deterministic: the first seed value s, s+2^40, s+2*2^40, ... that yields a key.
    for (int a = 0; a < kMaxAttempts; a++)
        if (keypair_attempt(pidx, attempt_seed(seed, a), sk)) return CMD_OK;
    return CMD_ERR_SINGULAR;
}

extern "C" int cmd_gen_encrypt(int32_t param_id, const uint8_t *sk, size_t sk_len, uint64_t seed,
                               size_t count, int32_t weight, uint8_t *ct, uint8_t *e)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return CMD_ERR_PARAM;
    const ParamDesc &p = kParams[pidx];
    if (!sk || (!ct && count)) return CMD_ERR_ARG;
    if (weight < 0) weight = p.t;
    if (weight > p.t) return CMD_ERR_ARG;
    ExpandedKey K;
    int rc = expand_key(pidx, sk, sk_len, K);
    if (rc) return rc;
    Encryptor E;
    if (!encryptor_init(pidx, K.g.data(), K.L.data(), E)) return CMD_ERR_SINGULAR;
    Rng rng(seed ^ 0xC1F3A5E7D2B49680ull ^ ((uint64_t)p.id << 20));
    std::vector<uint8_t> tmp_e((size_t)p.n / 8);
    ShuffleScratch sc;
    for (size_t i = 0; i < count; i++)
        encrypt_one(E, rng, weight, ct + i * (size_t)p.synd_bytes, e ? e + i * (size_t)(p.n / 8) : tmp_e.data(), sc);
    return CMD_OK;
}

extern "C" int cmd_controlbits(int32_t m, const int16_t *pi, uint8_t *cond)
{
    if ((m != 12 && m != 13) || !pi || !cond) return CMD_ERR_ARG;
    const int q = 1 << m;
    std::vector<int> perm(q);
    std::vector<char> seen(q, 0);
    for (int i = 0; i < q; i++) {
        int v = pi[i];
        if (v < 0 || v >= q || seen[v]) return CMD_ERR_ARG;  // not a permutation
        seen[v] = 1;
        perm[i] = v;
    }
    const size_t bytes = (size_t)(2 * m - 1) * ((size_t)1 << (m - 4));
    return controlbits(m, perm, cond, bytes) ? CMD_OK : CMD_ERR_ARG;
}

extern "C" int cmd_gen_keys(int32_t param_id, uint64_t base_seed, size_t n_keys, uint8_t *sk_out, int threads)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return CMD_ERR_PARAM;
    if (!sk_out && n_keys) return CMD_ERR_ARG;
    const size_t skb = (size_t)kParams[pidx].sk_bytes;
    if (threads < 1) threads = 1;
    std::atomic<size_t> next{0};
    std::atomic<int> status{CMD_OK};
    auto work = [&]() {
        for (;;) {
            size_t k = next.fetch_add(1);
            if (k >= n_keys) return;
            int rc = cmd_gen_keypair(param_id, base_seed + k, sk_out + k * skb);
            if (rc) status.store(rc);
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; i++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    return status.load();
}

// ---- ciphertexts on the GPU --------------------------------------------------------------

extern "C" int cmd_gen_encrypt_device(cmd_ctx *ctx, size_t n_keys, const int32_t *key_ids, const uint64_t *seeds,
                                      size_t per_key, int32_t weight, uint8_t *ct, uint8_t *e)
{
    if (!ctx || (n_keys && (!key_ids || !seeds || (per_key && !ct)))) return CMD_ERR_ARG;
    if (!n_keys || !per_key) return CMD_OK;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    int pidx = -1;
    for (size_t k = 0; k < n_keys; k++) {
        const int32_t id = key_ids[k];
        if (id < 0 || (size_t)id >= ctx->keys.size() || !ctx->keys[id].used)
            return set_error(CMD_ERR_KEY, "key %zu: id %d is not loaded", k, (int)id);
        if (pidx < 0) pidx = ctx->keys[id].pidx;
        if (ctx->keys[id].pidx != pidx) return set_error(CMD_ERR_ARG, "keys of one parameter set per call");
    }
    const ParamDesc &p = kParams[pidx];
    if (weight < 0) weight = p.t;
    if (weight > p.t) return CMD_ERR_ARG;
    if (per_key > (size_t)INT32_MAX / 2) return CMD_ERR_ARG;
    CMD_HIP(hipSetDevice(ctx->device));

    // error positions: the host generator's random stream, drawn per key in parallel
    std::vector<uint16_t> pos(n_keys * per_key * (size_t)weight);
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            ShuffleScratch sc;
            for (size_t k; (k = next.fetch_add(1)) < n_keys;) {
                Rng rng(seeds[k] ^ 0xC1F3A5E7D2B49680ull ^ ((uint64_t)p.id << 20));
                for (size_t i = 0; i < per_key; i++)
                    draw_positions(rng, p.n, weight, &pos[(k * per_key + i) * (size_t)weight], sc);
                if (e) {
                    for (size_t i = 0; i < per_key; i++) {
                        uint8_t *row = e + (k * per_key + i) * (size_t)(p.n / 8);
                        std::memset(row, 0, (size_t)p.n / 8);
                        const uint16_t *pp = &pos[(k * per_key + i) * (size_t)weight];
                        for (int j = 0; j < weight; j++) row[pp[j] >> 3] |= (uint8_t)(1u << (pp[j] & 7));
                    }
                }
            }
        };
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < nt && i < n_keys; i++) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
    }

    const size_t mt = (size_t)p.m * p.t, cw = (mt + 31) / 32;
    const size_t kBatch = 128;  // keys per device pass (bounds the per-key matrices)
    int rc = CMD_OK;
    for (size_t k0 = 0; k0 < n_keys && !rc; k0 += kBatch) {
        const size_t nk = std::min(kBatch, n_keys - k0);
        DevBuf d_cols, d_aug, d_ainvt, d_small, d_pos, d_ct;
        const size_t ct_bytes = nk * per_key * (size_t)p.synd_bytes;
        rc = dev_alloc(ctx, d_cols, nk * (size_t)p.n * cw * 4);
        if (!rc) rc = dev_alloc(ctx, d_aug, nk * mt * 2 * cw * 4);
        if (!rc) rc = dev_alloc(ctx, d_ainvt, nk * mt * cw * 4);
        if (!rc) rc = dev_alloc(ctx, d_small, nk * 8);  // key ids, then singular flags
        if (!rc) rc = dev_alloc(ctx, d_pos, nk * per_key * (size_t)weight * 2);
        if (!rc) rc = dev_alloc(ctx, d_ct, ct_bytes);
        std::vector<int32_t> flags(nk, 1);
        if (!rc) {
            GenArgs a;
            a.pidx = pidx;
            a.n_keys = (int32_t)nk;
            a.keys = ctx->d_keys;
            a.key_ids = (const int32_t *)d_small.ptr;
            a.cols = (uint32_t *)d_cols.ptr;
            a.aug = (uint32_t *)d_aug.ptr;
            a.ainvt = (uint32_t *)d_ainvt.ptr;
            a.singular = (int32_t *)d_small.ptr + nk;
            a.per_key = (int32_t)per_key;
            a.weight = weight;
            a.pos = (const uint16_t *)d_pos.ptr;
            a.ct = (uint8_t *)d_ct.ptr;
            a.cand_g = nullptr;
            a.cand_L = nullptr;
            a.rank_only = 0;
            a.pad0 = 0;
            hipError_t he = hipMemcpyAsync(d_small.ptr, key_ids + k0, nk * 4, hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess)
                he = hipMemcpyAsync((int32_t *)d_small.ptr + nk, flags.data(), nk * 4, hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess)
                he = hipMemcpyAsync(d_pos.ptr, &pos[k0 * per_key * (size_t)weight], nk * per_key * (size_t)weight * 2,
                                    hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess) he = launch_gen_keys(a, ctx->stream);
            if (he == hipSuccess) he = launch_gen_encrypt(a, ctx->stream);
            if (he == hipSuccess)
                he = hipMemcpyAsync(ct + k0 * per_key * (size_t)p.synd_bytes, d_ct.ptr, ct_bytes, hipMemcpyDeviceToHost,
                                    ctx->stream);
            if (he == hipSuccess)
                he = hipMemcpyAsync(flags.data(), (int32_t *)d_small.ptr + nk, nk * 4, hipMemcpyDeviceToHost, ctx->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
            if (he != hipSuccess) rc = hip_fail(he, "GPU ciphertext generator");
        }
        for (size_t k = 0; k < nk && !rc; k++)
            if (flags[k]) rc = set_error(CMD_ERR_SINGULAR, "key %zu has no systematic form", k0 + k);
        for (DevBuf *b : {&d_cols, &d_aug, &d_ainvt, &d_small, &d_pos, &d_ct}) dev_free(ctx, *b);
    }
    return rc;
}

// ---- key pairs with the systematic-form test on the GPU ---------------------------------------
//
// The same keys as cmd_gen_keys (key k from base_seed + k, attempt a of a key from seed
// base_seed + k + a 2^40, the first attempt that yields a key wins).  Per round, host threads
// draw the next attempts of every unfinished key (Goppa polynomial and permutation: the host
// generator's own code), the GPU tests all their A matrices at once (H's first mt columns from
// g and L_i = bitrev(pi_i), Gauss-Jordan for the rank), and per key the first attempt, in seed
// order, with a minimal polynomial, an invertible A and a routable permutation becomes the key.

extern "C" int cmd_gen_keys_device(cmd_ctx *ctx, int32_t param_id, uint64_t base_seed, size_t n_keys, uint8_t *sk_out,
                                   int threads)
{
    const int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    if (!ctx || (n_keys && !sk_out)) return CMD_ERR_ARG;
    if (!n_keys) return CMD_OK;
    const ParamDesc &p = kParams[pidx];
    if (threads < 1) threads = 1;
    std::lock_guard<std::recursive_mutex> lock(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    const size_t mt = (size_t)p.m * p.t, cw = (mt + 31) / 32;
    const size_t skb = (size_t)p.sk_bytes;
    constexpr int kSpec = 2;                   // attempts per unfinished key and round
    const size_t per_cand = (size_t)p.n * cw * 4 + mt * 2 * cw * 4;
    const size_t kBatch = std::max<size_t>(16, std::min<size_t>(512, (size_t)(2ull << 30) / per_cand));

    std::vector<int> next_attempt(n_keys, 0);
    std::vector<char> done(n_keys, 0);
    DevBuf d_cols, d_aug, d_small, d_g, d_L;
    int rc = CMD_OK;
    for (size_t remaining = n_keys; remaining && !rc;) {
        // this round's candidates: (key, attempt)
        struct Cand { size_t key; int attempt; };
        std::vector<Cand> cands;
        for (size_t k = 0; k < n_keys; k++) {
            if (done[k]) continue;
            for (int s = 0; s < kSpec && next_attempt[k] + s < kMaxAttempts; s++) cands.push_back({k, next_attempt[k] + s});
            if (next_attempt[k] >= kMaxAttempts) {
                rc = set_error(CMD_ERR_SINGULAR, "key %zu: no key in %d attempts", k, kMaxAttempts);
                break;
            }
        }
        if (rc) break;
        std::vector<KeyCandidate> kc(cands.size());
        {
            std::atomic<size_t> next{0};
            auto work = [&]() {
                for (size_t i; (i = next.fetch_add(1)) < cands.size();)
                    draw_candidate(pidx, attempt_seed(base_seed + cands[i].key, cands[i].attempt), kc[i]);
            };
            std::vector<std::thread> pool;
            for (int i = 1; i < threads && (size_t)i < cands.size(); i++) pool.emplace_back(work);
            work();
            for (auto &th : pool) th.join();
        }
        // the GPU's rank test for every candidate with a Goppa polynomial
        std::vector<int32_t> singular(cands.size(), 1);
        std::vector<size_t> idx;
        for (size_t i = 0; i < cands.size(); i++)
            if (kc[i].poly_ok) idx.push_back(i);
        for (size_t b0 = 0; b0 < idx.size() && !rc; b0 += kBatch) {
            const size_t nb = std::min(kBatch, idx.size() - b0);
            std::vector<uint16_t> hg(nb * p.t), hL(nb * (size_t)p.n);
            const HostField F(p.m);
            for (size_t q = 0; q < nb; q++) {
                const KeyCandidate &c = kc[idx[b0 + q]];
                std::memcpy(&hg[q * p.t], c.g.data(), sizeof(uint16_t) * p.t);
                for (int i = 0; i < p.n; i++) hL[q * p.n + i] = F.bitrev((uint16_t)c.pi[i]);
            }
            rc = dev_reserve(ctx, d_cols, nb * (size_t)p.n * cw * 4);
            if (!rc) rc = dev_reserve(ctx, d_aug, nb * mt * 2 * cw * 4);
            if (!rc) rc = dev_reserve(ctx, d_small, nb * 4);
            if (!rc) rc = dev_reserve(ctx, d_g, hg.size() * 2);
            if (!rc) rc = dev_reserve(ctx, d_L, hL.size() * 2);
            if (rc) break;
            GenArgs a{};
            a.pidx = pidx;
            a.n_keys = (int32_t)nb;
            a.keys = ctx->d_keys;
            a.cols = (uint32_t *)d_cols.ptr;
            a.aug = (uint32_t *)d_aug.ptr;
            a.singular = (int32_t *)d_small.ptr;
            a.cand_g = (const uint16_t *)d_g.ptr;
            a.cand_L = (const uint16_t *)d_L.ptr;
            a.rank_only = 1;
            hipError_t he = hipMemcpyAsync(d_g.ptr, hg.data(), hg.size() * 2, hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess) he = hipMemcpyAsync(d_L.ptr, hL.data(), hL.size() * 2, hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess) he = launch_gen_check(a, ctx->stream);
            std::vector<int32_t> flags(nb, 1);
            if (he == hipSuccess)
                he = hipMemcpyAsync(flags.data(), d_small.ptr, nb * 4, hipMemcpyDeviceToHost, ctx->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
            if (he != hipSuccess) rc = hip_fail(he, "key-pair rank test");
            for (size_t q = 0; q < nb && !rc; q++) singular[idx[b0 + q]] = flags[q];
        }
        if (rc) break;
        // condition bits of every candidate still in the running (host threads), then per key
        // the first winning attempt in seed order
        std::vector<std::vector<uint8_t>> conds(cands.size());
        std::vector<char> routed(cands.size(), 0);
        {
            std::atomic<size_t> next{0};
            auto work = [&]() {
                for (size_t i; (i = next.fetch_add(1)) < cands.size();) {
                    if (!kc[i].poly_ok || singular[i]) continue;
                    conds[i].assign(p.cond_bytes, 0);
                    routed[i] = controlbits(p.m, kc[i].pi, conds[i].data(), conds[i].size());
                }
            };
            std::vector<std::thread> pool;
            for (int i = 1; i < threads && (size_t)i < cands.size(); i++) pool.emplace_back(work);
            work();
            for (auto &th : pool) th.join();
        }
        std::vector<size_t> winner(n_keys, SIZE_MAX);
        for (size_t i = 0; i < cands.size(); i++)
            if (routed[i] && winner[cands[i].key] == SIZE_MAX) winner[cands[i].key] = i;
        for (size_t k = 0; k < n_keys; k++) {
            if (done[k]) continue;
            if (winner[k] != SIZE_MAX) {
                const size_t i = winner[k];
                write_sk(pidx, kc[i], conds[i].data(), sk_out + k * skb);
                done[k] = 1;
                remaining--;
            } else {
                next_attempt[k] += kSpec;
            }
        }
    }
    for (DevBuf *b : {&d_cols, &d_aug, &d_small, &d_g, &d_L}) dev_free(ctx, *b);
    return rc;
}
