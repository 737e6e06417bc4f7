// Secret-key parsing and expansion into the decoder's per-key tables (host).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "params.h"

namespace cmd {

// Applies the (2m-1)-layer Benes network of a secret key to 2^m 16-bit values in place.
// Layer k (k < m) pairs positions p, p + 2^k; layers m..2m-2 repeat strides 2^(m-2)..1.
// Within a layer, switch number j acts on the j-th pair in position order and is bit j of
// that layer's 2^(m-1) condition bits (LSB first); layers are 2^(m-4) bytes apart.
// This is the network whose bits controlbitsfrompermutation writes
// (kem/common/core/controlbits.c:186-210) and that apply_benes evaluates in bit-sliced form
// (CPU/src/common/root_13.c:95-177, CPU/param/param348864/src/root.c:253-332).
void benes_apply(int m, uint16_t *values, const uint8_t *cond);

// support_gen (root_13.c:181-211): L[i] for i < n from the condition bits.
int support_gen(int pidx, const uint8_t *cond, uint16_t *L);

struct ExpandedKey {
    int pidx = -1;
    std::vector<uint16_t> g;      // t+1 coefficients, monic
    std::vector<uint16_t> L;      // n support elements
    std::vector<uint16_t> table;  // sb x 2t: table[i][j] = L_i^j / g(L_i)^2 (util.c:86-100)
    std::vector<uint32_t> lbits;  // m x groups: bit b of word [k][gi] = bit k of L[32*gi + b]
    std::vector<uint32_t> benes;  // (2m-1) x 2^(m-5): per layer and 32-bit word, the positions whose
                                  // switch is set (both words of a pair carry the pair's mask)
};

// initialisation + compute_inverses (CPU/param/param6960119/src/util.c:60-100) for one key.
// Returns 0, or a negative cmd error code.
int expand_key(int pidx, const uint8_t *sk, size_t sk_len, ExpandedKey &out);

// Number of 32-bit words per row of the bit-plane support table (groups rounded up to 64).
inline int lbits_stride(int pidx) { return ((kParams[pidx].n + 31) / 32 + 63) / 64 * 64; }

}  // namespace cmd
