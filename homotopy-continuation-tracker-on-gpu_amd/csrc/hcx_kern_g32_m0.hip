// Generic tracker kernel instances: NMAX = 32 (any n <= 32, any G >= n), FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g32_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<32, 2, 0, false>;
    case 3: return track_kernel<32, 3, 0, false>;
    case 6: return track_kernel<32, 6, 0, false>;
    default: return nullptr;
  }
}

}  // namespace hcx
