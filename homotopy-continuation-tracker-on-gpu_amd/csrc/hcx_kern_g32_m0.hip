// Generic tracker kernel instances: NMAX = 32 (any n <= 32, any G >= n), FP64 mode.
// One translation unit per term degree (hcx_kern_g32_m0_d*.hip) so that
// the instances compile in parallel; this unit only dispatches.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g32_m0(int deg) {
  switch (deg) {
    case 2: return kernel_g32_m0_d2();
    case 3: return kernel_g32_m0_d3();
    case 6: return kernel_g32_m0_d6();
    default: return nullptr;
  }
}

}  // namespace hcx
