// Error positions from decoded error vectors: the numbers of the errorstream
// (write_error_positions, CPU/param/param6960119/src/decrypt.c:55-66) without the n/8-byte
// bitmap.  A decoded vector has at most t set bits (the reversed locator has degree t), so a
// row of t uint16 holds them in ascending order, padded with 0xFFFF; it is 2t bytes against
// n/8 for the bitmap (256 against 1024 at 8192128), which is what crosses PCIe.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "params.h"

namespace cmd {

namespace {

struct SetSizes {
    int n, t;
};
__device__ __forceinline__ SetSizes set_sizes(int pidx)
{
    switch (pidx) {
    case 0: return {P<0>::n, P<0>::t};
    case 1: return {P<1>::n, P<1>::t};
    case 2: return {P<2>::n, P<2>::t};
    case 3: return {P<3>::n, P<3>::t};
    default: return {P<4>::n, P<4>::t};
    }
}

// One wavefront per slot; lane l scans words 4l .. 4l+3 of the bitmap, an exclusive scan of
// the lanes' bit counts gives each lane its first output index.
__global__ __launch_bounds__(256) void k_positions(PositionsArgs a)
{
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= a.count) return;
    const int lane = threadIdx.x & 63;
    const int item = a.items[q];
    const SetSizes z = set_sizes(a.keys[a.item_key[q]].pidx);
    const uint32_t *ew = reinterpret_cast<const uint32_t *>(a.e + (size_t)item * a.e_stride);
    uint16_t *out = reinterpret_cast<uint16_t *>(a.out + (size_t)item * a.out_stride);
    const int nwords = (z.n + 31) / 32;
    static_assert((kMaxN + 31) / 32 <= 4 * 64, "four words per lane");
    uint32_t w[4];
    int cnt = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int wi = 4 * lane + r;
        uint32_t v = 0;
        if (wi < nwords) {
            v = ew[wi];
            const int valid = z.n - 32 * wi;  // bits of this word inside the vector
            if (valid < 32) v &= (1u << valid) - 1u;
        }
        w[r] = v;
        cnt += __builtin_popcount(v);
    }
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int up = __shfl_up(incl, off);
        if (lane >= off) incl += up;
    }
    const int total = __shfl(incl, 63);
    int k = incl - cnt;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        for (uint32_t v = w[r]; v; v &= v - 1, k++)
            if (k < z.t) out[k] = (uint16_t)(32 * (4 * lane + r) + __builtin_ctz(v));
    }
    for (int j = total + lane; j < z.t; j += 64) out[j] = 0xFFFFu;
}

}  // namespace

hipError_t launch_positions(const PositionsArgs &a, hipStream_t st)
{
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_positions, dim3((a.count + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace cmd
