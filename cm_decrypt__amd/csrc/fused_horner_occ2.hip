// Fused kernel, Horner root search at each support element (root_alg 0), built for two wavefronts per SIMD (256 registers).
#include "fused_impl.h"

namespace cmd {

hipError_t launch_fused_horner_occ2(const FusedArgs &a, const FusedSchedule &s, int n_cu, hipStream_t st)
{
    return launch_fused_variant<0, 2>(a, s, n_cu, st);
}

}  // namespace cmd
