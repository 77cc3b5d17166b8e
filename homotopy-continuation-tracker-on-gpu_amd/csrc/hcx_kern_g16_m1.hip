// Generic tracker kernel instances: NMAX = 16 (any n <= 16, any G >= n), mixed FP32/FP64 mode.
// One translation unit per term degree (hcx_kern_g16_m1_d*.hip) so that
// the instances compile in parallel; this unit only dispatches.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g16_m1(int deg) {
  switch (deg) {
    case 2: return kernel_g16_m1_d2();
    case 3: return kernel_g16_m1_d3();
    case 6: return kernel_g16_m1_d6();
    default: return nullptr;
  }
}

}  // namespace hcx
