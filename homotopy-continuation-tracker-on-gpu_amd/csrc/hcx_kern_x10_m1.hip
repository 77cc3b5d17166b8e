// Size-specialised tracker kernel instances: n = G = 10, mixed FP32/FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_x10_m1(int deg) {
  switch (deg) {
    case 2: return track_kernel<10, 2, 1, true>;
    case 3: return track_kernel<10, 3, 1, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
