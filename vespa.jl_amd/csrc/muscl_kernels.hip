// Both arithmetic profiles of the MUSCL stage kernel (this file is compiled twice,
// VH_FAST_PROFILE = 0 and 1, into the exact and fast namespaces).
#include "../../include/vespahip.h"
#include "muscl.h"

namespace vh {
namespace VH_PROFILE_NS {

void muscl_launch_stage(int dtype, int variant, int nf, bool stage2, bool rec, const MusclDev& G, hipStream_t s,
                        const void* Us, const void* U0, void* Uout, void* frec) {
  if (dtype == VH_F64) {
    if (variant == 1) muscl_launch_t<double, 32, 8>(nf, stage2, rec, G, s, Us, U0, Uout, frec);
    else muscl_launch_t<double, 16, 16>(nf, stage2, rec, G, s, Us, U0, Uout, frec);
  } else {
    if (variant == 1) muscl_launch_t<float, 32, 8>(nf, stage2, rec, G, s, Us, U0, Uout, frec);
    else muscl_launch_t<float, 16, 16>(nf, stage2, rec, G, s, Us, U0, Uout, frec);
  }
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
