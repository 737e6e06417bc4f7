// Fused kernel, additive-FFT root search, two wavefronts per tile, built for four
// wavefronts per SIMD (128 registers).
#include "fused_impl.h"

namespace cmd {

hipError_t launch_fused_fft_pair(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st)
{
    return launch_fused_pair_variant<1, 4>(a, s, n_cu, st);
}

}  // namespace cmd
