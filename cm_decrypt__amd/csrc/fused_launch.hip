// Fused-kernel dispatch and the small helper kernel the launch sequence uses.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cmd {

hipError_t launch_fused_fft_occ1(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st);
hipError_t launch_fused_fft_occ2(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st);
hipError_t launch_fused_horner_occ1(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st);
hipError_t launch_fused_horner_occ2(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st);
hipError_t launch_fused_fft_pair(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st);

// Zero fill used instead of hipMemsetAsync so the launch sequence can be graph-captured.
__global__ void k_zero_u32(uint32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

bool fused_supported(const FusedSchedule &s)
{
    // the kernel is bit-sliced 32 wide; narrower tiles leave bit positions idle but spread a
    // small batch over more wavefronts
    const bool tile_ok = s.tile_cts >= 1 && s.tile_cts <= 32 && (s.tile_cts & (s.tile_cts - 1)) == 0;
    const bool wpg_ok = s.waves_per_group == 1 || s.waves_per_group == 2 || s.waves_per_group == 4;
    const bool alg_ok = s.root_alg == 0 || s.root_alg == 1;
    if (s.waves_per_ct == 2)  // one pair per workgroup, additive-FFT root search
        return tile_ok && s.waves_per_group == 2 && s.root_alg == 1;
    return tile_ok && wpg_ok && s.waves_per_ct == 1 && alg_ok;
}

hipError_t launch_fused(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st)
{
    if (!fused_supported(s)) return hipErrorInvalidValue;
    if (a.n_tiles <= 0) return hipSuccess;
    if (s.waves_per_ct == 2) return launch_fused_fft_pair(a, s, n_cu, st);
    const bool small = (long)a.n_tiles < (long)n_cu * s.waves_per_group * s.groups_per_cu;
    // Small launches of tiles of at most two ciphertexts (serving at low load) decode each
    // ciphertext with the per-ciphertext recurrence, a chain of dependent steps that one
    // wavefront issues slowly: the paired kernel splits it, and the syndrome, over the two
    // wavefronts of a workgroup.
    if (small && s.tile_cts <= 2 && s.root_alg == 1) {
        FusedSchedule p = s;
        p.waves_per_ct = 2;
        p.waves_per_group = 2;
        p.groups_per_cu = s.waves_per_group * s.groups_per_cu / 2 > 0 ? s.waves_per_group * s.groups_per_cu / 2 : 1;
        return launch_fused_fft_pair(a, p, n_cu, st);
    }
    // more than one wavefront per SIMD needs the 256-register build
    const bool occ2 = s.waves_per_group * s.groups_per_cu > 4;
    // Fewer tiles than resident wavefronts (small and serving launches): one wavefront per
    // workgroup, so the tiles land on as many CUs as there are tiles instead of sharing a few
    // CUs' scalar units, LDS and caches; the build and the wavefronts per CU stay the same.
    FusedSchedule t = s;
    if (small) {
        t.groups_per_cu = s.waves_per_group * s.groups_per_cu;
        t.waves_per_group = 1;
    }
    if (t.root_alg == 1) return occ2 ? launch_fused_fft_occ2(a, t, n_cu, st) : launch_fused_fft_occ1(a, t, n_cu, st);
    return occ2 ? launch_fused_horner_occ2(a, t, n_cu, st) : launch_fused_horner_occ1(a, t, n_cu, st);
}

}  // namespace cmd
