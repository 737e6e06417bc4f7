// Secret-key parsing and expansion (host).
#include "keys.h"

#include <cstring>

#include "../../include/cm_decrypt.h"
#include "host_gf.h"

namespace cmd {

void benes_apply(int m, uint16_t *v, const uint8_t *cond)
{
    const int count = 1 << m;
    const int layer_bytes = count / 16;  // count/2 switches
    int layer = 0;
    auto run_layer = [&](int lg) {
        const uint8_t *bits = cond + (size_t)layer * layer_bytes;
        const int s = 1 << lg;
        int sw = 0;
        for (int base = 0; base < count; base += 2 * s)
            for (int j = base; j < base + s; j++, sw++) {
                uint16_t take = (uint16_t)(0u - ((bits[sw >> 3] >> (sw & 7)) & 1u));
                uint16_t d = (uint16_t)((v[j] ^ v[j + s]) & take);
                v[j] ^= d;
                v[j + s] ^= d;
            }
        layer++;
    };
    for (int lg = 0; lg < m; lg++) run_layer(lg);
    for (int lg = m - 2; lg >= 0; lg--) run_layer(lg);
}

int support_gen(int pidx, const uint8_t *cond, uint16_t *L)
{
    if (pidx < 0 || pidx >= kNumParams) return CMD_ERR_PARAM;
    const ParamDesc &p = kParams[pidx];
    HostField F(p.m);
    std::vector<uint16_t> v((size_t)1 << p.m);
    for (size_t i = 0; i < v.size(); i++) v[i] = F.bitrev((uint16_t)i);
    benes_apply(p.m, v.data(), cond);
    std::memcpy(L, v.data(), sizeof(uint16_t) * p.n);
    return CMD_OK;
}

int expand_key(int pidx, const uint8_t *sk, size_t sk_len, ExpandedKey &out)
{
    if (pidx < 0 || pidx >= kNumParams) return CMD_ERR_PARAM;
    const ParamDesc &p = kParams[pidx];
    if (!sk || sk_len < (size_t)p.sk_bytes) return CMD_ERR_CRYPTO;
    HostField F(p.m);
    out.pidx = pidx;
    out.g.assign(p.t + 1, 0);
    out.L.assign(p.n, 0);
    const uint8_t *ptr = sk + 40;  // 32 B seed + 8 B pivot word (util.c:13, :63)
    for (int i = 0; i < p.t; i++, ptr += 2) out.g[i] = (uint16_t)((ptr[0] | (ptr[1] << 8)) & F.mask);
    out.g[p.t] = 1;
    support_gen(pidx, ptr, out.L.data());

    const int w = 2 * p.t;
    out.table.assign((size_t)p.sb * w, 0);
    for (int i = 0; i < p.sb; i++) {
        uint16_t v = F.eval(out.g.data(), p.t, out.L[i]);
        uint16_t *row = &out.table[(size_t)i * w];
        row[0] = F.inv(F.mul(v, v));
        for (int j = 1; j < w; j++) row[j] = F.mul(row[j - 1], out.L[i]);
    }

    // Benes swap masks per word: layer k (stride 2^s), switch j acts on (c, c + 2^s)
    const int words = 1 << (p.m - 5);
    const uint8_t *cond = sk + 40 + 2 * p.t;
    out.benes.assign((size_t)(2 * p.m - 1) * words, 0);
    for (int k = 0; k < 2 * p.m - 1; k++) {
        const int sft = k < p.m ? k : 2 * p.m - 2 - k;
        const uint8_t *bits = cond + (size_t)k * ((size_t)1 << (p.m - 4));
        for (int c = 0; c < (1 << p.m); c++) {
            if ((c >> sft) & 1) continue;
            const int j = ((c >> (sft + 1)) << sft) | (c & ((1 << sft) - 1));
            if (!((bits[j >> 3] >> (j & 7)) & 1)) continue;
            out.benes[(size_t)k * words + (c >> 5)] |= 1u << (c & 31);
            if (sft >= 5) out.benes[(size_t)k * words + ((c + (1 << sft)) >> 5)] |= 1u << (c & 31);
        }
    }

    const int stride = lbits_stride(pidx);
    out.lbits.assign((size_t)p.m * stride, 0);
    for (int i = 0; i < p.n; i++)
        for (int k = 0; k < p.m; k++)
            out.lbits[(size_t)k * stride + (i >> 5)] |= (uint32_t)((out.L[i] >> k) & 1u) << (i & 31);
    return CMD_OK;
}

}  // namespace cmd
