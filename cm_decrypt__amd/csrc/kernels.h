// Launch interface between the host engine and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace cmd {

// Device view of one resident key.
struct KeyDev {
    const uint16_t *table;  // sb x 2t: L_i^j / g(L_i)^2
    const uint16_t *L;      // n support elements
    const uint32_t *lbits;  // m x lbits_stride bit planes of L (32 positions per word)
    const uint32_t *benes;  // (2m-1) x 2^(m-5) swap masks of the key's Benes network, per 32-bit word
    const uint32_t *ypl;    // m x 2^(m-5) bit planes of 1/g(a)^2 in FFT order (word w, bit b: a = bitrev_m(32w+b))
    const uint8_t *s;       // n/8 bytes of the secret key's rejection value s
    const uint8_t *gbytes;  // 2t bytes: the Goppa coefficients g_0 .. g_{t-1} as stored in the key
    int32_t pidx;
    int32_t lstride;        // words per bit-plane row
};

// Constants of the additive FFT that evaluates a polynomial of 2^l coefficients at all 2^m
// field elements (evaluation index c holds the value at bitrev_m(c)).  They depend only on
// (m, l); see fft_tables.cpp for their definition.
struct FftTab {
    uint16_t tw[8][256];      // twist at depth d for position p: lambda_d^(p >> d)
    uint16_t hi[8][128];      // butterfly multiplier, high part, by word offset in the half
    uint32_t lowpat[8][16];   // butterfly multiplier, low part, bit-sliced over the 32 bits of a word
    uint32_t gpl[8][128][16]; // whole multiplier planes: lowpat ^ (bit q of hi ? ~0 : 0), 16-aligned rows
};
constexpr int kFftVariants = 3;  // (m, l) = (12, 7), (13, 7), (13, 8)
constexpr int fft_variant(int m, int l) { return m == 12 ? 0 : (l == 7 ? 1 : 2); }

// ---- on-device key expansion ---------------------------------------------------------
// d_key holds the secret key from byte 40 on: t little-endian Goppa coefficients, then the
// Benes control bits.  Fills k.L, k.lbits, k.table, k.benes and k.ypl (all of set pidx).
hipError_t launch_key_expand(int pidx, const uint8_t *d_key, const KeyDev &k, hipStream_t st);

// ---- plain per-stage pipeline (one key per launch group) ----------------------------
// Items k < count use ciphertext ct + idx[k]*ct_stride and write e + idx[k]*e_stride;
// idx == nullptr means identity.  s: count x 2t, loc: count x (t+1).
hipError_t launch_plain_synd(int pidx, const KeyDev &key, int count, const int32_t *idx,
                             const uint8_t *ct, size_t ct_stride, uint16_t *s, hipStream_t st);
hipError_t launch_plain_bm(int pidx, int count, const uint16_t *s, uint16_t *loc, hipStream_t st);
hipError_t launch_plain_chien(int pidx, const KeyDev &key, int count, const int32_t *idx,
                              const uint16_t *loc, uint8_t *e, size_t e_stride, hipStream_t st);

// ---- fused persistent pipeline -------------------------------------------------------
// One work item of the fused kernel: up to 32 ciphertexts of one parameter set.
struct FusedTile {
    int32_t pidx;   // parameter set
    int32_t first;  // offset into the item list
    int32_t count;  // items in this tile (<= tile width)
    int32_t pad;
};

struct FusedArgs {
    const FusedTile *tiles;
    int32_t n_tiles;
    int32_t pad0;
    const int32_t *items;      // item index (into ct/e) per tile slot
    const int32_t *item_key;   // key slot per tile slot
    const KeyDev *keys;        // resident key table (device)
    const uint8_t *ct;
    uint8_t *e;
    uint64_t ct_stride, e_stride;
    uint32_t *tile_counter;    // work-queue head, zeroed before each launch
    const FftTab *fft;         // kFftVariants tables
};

struct FusedSchedule {
    int tile_cts;         // 16 or 32
    int waves_per_group;  // 1, 2 or 4
    int waves_per_ct;     // 1
    int groups_per_cu;    // persistent workgroups per CU
    int root_alg;         // 0: Horner at each support point, 1: additive FFT + Benes network
};

// ---- KEM decapsulation -----------------------------------------------------------------
// Re-encryption check of decoded error vectors, over the tiles of a fused plan.  For item x,
// v = (c || 0) ^ e (or v = e alone when ct is null) and the check computes the power sums
// s_j(v) = sum_{i < n} v_i L_i^j / g(L_i)^2, j < 2t, by the transposed additive FFT.
//   ok != null:    ok[item] = 1 if weight(e) == t and s(v) == 0, else 0
//   s_out != null: s_out[item * 2t + j] = s_j(v)
struct DecapArgs {
    const FusedTile *tiles;
    int32_t n_tiles;
    int32_t pad0;
    const int32_t *items;
    const int32_t *item_key;
    const KeyDev *keys;
    const uint8_t *ct;         // ciphertexts, or null
    const uint8_t *e;          // error vectors (n/8 bytes each)
    uint64_t ct_stride, e_stride;
    uint32_t *ok;
    uint16_t *s_out;
    const FftTab *fft;
};
hipError_t launch_decap_check(const DecapArgs &a, hipStream_t st);

// Session keys: per item, SHAKE256(b || (b ? e : s) || c) truncated to 32 bytes, b = ok[item]
// (crypto_kem_dec).  Slot q < count is item items[q] of key slot item_key[q] (the arrays of a
// fused plan); msg is scratch of count * kKemMsgStride bytes.
constexpr int kKemMsgStride = 1360;  // 10 SHAKE256 blocks: the largest padded preimage (8192128)
struct KemHashArgs {
    int32_t count;
    int32_t pad0;
    const int32_t *items;
    const int32_t *item_key;
    const KeyDev *keys;
    const uint8_t *ct;
    const uint8_t *e;
    uint64_t ct_stride, e_stride;
    const uint32_t *ok;
    uint8_t *msg;
    uint8_t *key_out;
    uint64_t key_stride;
};
hipError_t launch_kem_hash(const KemHashArgs &a, hipStream_t st);

// ---- error positions -------------------------------------------------------------------
// Slot q < count: the set bits of item items[q]'s decoded vector (e + item * e_stride, a
// 4-byte-aligned row of n/8 bytes) as t ascending uint16 at out + item * out_stride, padded
// with 0xFFFF.
struct PositionsArgs {
    int32_t count;
    int32_t pad0;
    const int32_t *items;
    const int32_t *item_key;
    const KeyDev *keys;
    const uint8_t *e;
    uint64_t e_stride;
    uint8_t *out;
    uint64_t out_stride;
};
hipError_t launch_positions(const PositionsArgs &a, hipStream_t st);

// ---- seeded ciphertext generator on the GPU (same bytes as the host generator) -----------
// For n_keys resident keys of one set: H's columns (bit i*m + k of column j = bit k of
// L_j^i / g(L_j)), [A | I] for A = the first mt columns, and A^-1 by Gauss-Jordan; then per
// ciphertext ct = A^-1 (sum of the columns at its error positions).
struct GenArgs {
    int32_t pidx;
    int32_t n_keys;
    const KeyDev *keys;
    const int32_t *key_ids;   // per key slot
    uint32_t *cols;           // per key: n columns of cw words (cw = 32-bit words of mt bits)
    uint32_t *aug;            // per key: mt rows of 2 cw words ([A | I], then [I | A^-1])
    uint32_t *ainvt;          // per key: the mt columns of A^-1, cw words each
    int32_t *singular;        // per key: 1 when A has no inverse
    int32_t per_key;          // ciphertexts per key
    int32_t weight;           // error positions per ciphertext
    const uint16_t *pos;      // per ciphertext: `weight` positions (key-major order)
    uint8_t *ct;              // per ciphertext: synd_bytes
    // key-pair candidates instead of resident keys (both non-null): per candidate t Goppa
    // coefficients (g_t = 1 implied) and n support elements
    const uint16_t *cand_g;
    const uint16_t *cand_L;
    int32_t rank_only;        // 1: eliminate within A only (its invertibility), not A^-1
    int32_t pad0;
};
hipError_t launch_gen_keys(const GenArgs &a, hipStream_t st);     // columns and inverses
hipError_t launch_gen_check(const GenArgs &a, hipStream_t st);    // columns and invertibility of A
hipError_t launch_gen_encrypt(const GenArgs &a, hipStream_t st);  // ciphertexts

// Host: fills the FFT constants for (m, l).
void build_fft_tab(int m, int l, FftTab *out);

// Variant of the fused kernel compiled for a schedule, or false if not built.
bool fused_supported(const FusedSchedule &s);
hipError_t launch_fused(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st);
hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t st);

}  // namespace cmd
