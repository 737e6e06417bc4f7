// Constants of the additive FFT used by the fused kernel's root search.
//
// The polynomial f (2^l coefficients) is evaluated at every element of GF(2^m) with the
// Gao-Mateer recursion over the basis (x^(m-1), ..., x, 1), so evaluation index c holds
// f(bitrev_m(c)).  Depth d works with basis B_d: B_0 = (x^(m-1), ..., 1); with lambda_d the
// last element of B_d and gamma_j = B_d[j] / lambda_d, B_{d+1}[j] = gamma_j^2 + gamma_j.
//  * twist at depth d: coefficient at node-local index i is multiplied by lambda_d^i;
//  * butterfly at depth d for the point index c' of the lower half:
//      G(c') = sum_j c'_j gamma_j,  w0 = u + G v,  w1 = w0 + v.
//    With c' = 32 i + b (word offset i in the half, bit b in the word), G splits into
//    low(b) (bits of b) and hi(i) (bits of i); lowpat stores low(b) bit-sliced over b, gpl the
//    sum of both as whole bit-sliced planes per word offset.
#include <cstring>

#include "host_gf.h"
#include "kernels.h"

namespace cmd {

void build_fft_tab(int m, int l, FftTab *out)
{
    std::memset(out, 0, sizeof *out);
    HostField F(m);
    int basis[16];
    for (int i = 0; i < m; i++) basis[i] = 1 << (m - 1 - i);
    int k = m;  // basis length at the current depth
    for (int d = 0; d < l; d++) {
        const uint16_t lam = (uint16_t)basis[k - 1];
        const uint16_t ilam = F.inv(lam);
        uint16_t gam[16];
        for (int j = 0; j < k - 1; j++) gam[j] = F.mul((uint16_t)basis[j], ilam);
        // twist constants for positions p < 2^l (node-local index p >> d)
        uint16_t pw = 1;
        int last_i = 0;
        for (int p = 0; p < (1 << l); p++) {
            const int i = p >> d;
            while (last_i < i) { pw = F.mul(pw, lam); last_i++; }
            if (i == 0) pw = 1;
            out->tw[d][p] = pw;
        }
        // butterfly multipliers
        for (int b = 0; b < 32; b++) {
            uint16_t lo = 0;
            for (int j = 0; j < 5 && j < k - 1; j++)
                if ((b >> j) & 1) lo ^= gam[j];
            for (int q = 0; q < m; q++) out->lowpat[d][q] |= (uint32_t)((lo >> q) & 1u) << b;
        }
        const int hwords = 1 << (k - 1 - 5);  // words in the lower half
        for (int i = 0; i < hwords && i < 128; i++) {
            uint16_t hv = 0;
            for (int j = 5; j < k - 1; j++)
                if ((i >> (j - 5)) & 1) hv ^= gam[j];
            out->hi[d][i] = hv;
        }
        for (int i = 0; i < 128; i++)
            for (int q = 0; q < m; q++)
                out->gpl[d][i][q] = out->lowpat[d][q] ^ (((out->hi[d][i] >> q) & 1u) ? ~0u : 0u);
        for (int j = 0; j < k - 1; j++) basis[j] = F.mul(gam[j], gam[j]) ^ gam[j];
        k--;
    }
}

}  // namespace cmd
