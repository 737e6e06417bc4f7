// Parameter sets of the Classic McEliece decoder (host and device).
//
// Sizes follow CPU/param/param<p>/include/common.h (e.g. param6960119/include/common.h:10-34):
// SYND_BYTES = ceil(m*t/8) is also the ciphertext size, sb = 8*SYND_BYTES is the number of
// ciphertext bits the syndrome consumes, the secret key is
// 32 B seed | 8 B prefix | t*2 B Goppa coefficients | COND_BYTES Benes bits | n/8 B s
// (kem/common/core/operations.c:62-101).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define CMD_HD __host__ __device__
#else
#define CMD_HD
#endif

namespace cmd {

struct ParamDesc {
    int32_t id;
    int32_t m, n, t;
    int32_t synd_bytes;  // ciphertext bytes
    int32_t sb;          // syndrome bits (8 * synd_bytes)
    int32_t sk_bytes;
    int32_t cond_bytes;  // (2m-1) * 2^(m-4)
    uint32_t poly_low;   // field modulus without x^m: x^13+x^4+x^3+x+1 -> 0x1B, x^12+x^3+1 -> 0x9
};

constexpr int kNumParams = 5;
constexpr int kMaxT = 128;
constexpr int kMaxN = 8192;
constexpr int kMaxSyndBytes = 208;
constexpr int kMaxEBytes = kMaxN / 8;

constexpr ParamDesc kParams[kNumParams] = {
    {348864, 12, 3488, 64, 96, 768, 6492, 5888, 0x9u},
    {460896, 13, 4608, 96, 156, 1248, 13608, 12800, 0x1Bu},
    {6688128, 13, 6688, 128, 208, 1664, 13932, 12800, 0x1Bu},
    {6960119, 13, 6960, 119, 194, 1552, 13948, 12800, 0x1Bu},
    {8192128, 13, 8192, 128, 208, 1664, 14120, 12800, 0x1Bu},
};

// Index into kParams, or -1.
CMD_HD inline int param_index(int32_t id)
{
    for (int i = 0; i < kNumParams; i++)
        if (kParams[i].id == id) return i;
    return -1;
}

// Compile-time view of one parameter set, for kernel templates.
template <int IDX>
struct P {
    static constexpr int idx = IDX;
    static constexpr int id = kParams[IDX].id;
    static constexpr int m = kParams[IDX].m;
    static constexpr int n = kParams[IDX].n;
    static constexpr int t = kParams[IDX].t;
    static constexpr int synd_bytes = kParams[IDX].synd_bytes;
    static constexpr int sb = kParams[IDX].sb;
    static constexpr uint32_t poly_low = kParams[IDX].poly_low;
    static constexpr int e_bytes = n / 8;
    static constexpr int groups = (n + 31) / 32;  // 32-position groups
};

}  // namespace cmd
