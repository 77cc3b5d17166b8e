// Generic tracker kernel instances: NMAX = 16 (any n <= 16, any G >= n), mixed FP32/FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g16_m1(int deg) {
  switch (deg) {
    case 2: return track_kernel<16, 2, 1, false>;
    case 3: return track_kernel<16, 3, 1, false>;
    case 6: return track_kernel<16, 6, 1, false>;
    default: return nullptr;
  }
}

}  // namespace hcx
