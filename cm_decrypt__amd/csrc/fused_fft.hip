// Fused kernel, additive-FFT root search (root_alg 1), built for one wavefront per SIMD.
#include "fused_impl.h"

namespace cmd {

hipError_t launch_fused_fft_occ1(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st)
{
    return launch_fused_variant<1, 1>(a, s, n_cu, st);
}

}  // namespace cmd
