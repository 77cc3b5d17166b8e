// Generic tracker kernel instances: NMAX = 8 (any n <= 8, any G >= n), FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g8_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<8, 2, 0, false>;
    case 3: return track_kernel<8, 3, 0, false>;
    case 6: return track_kernel<8, 6, 0, false>;
    default: return nullptr;
  }
}

}  // namespace hcx
