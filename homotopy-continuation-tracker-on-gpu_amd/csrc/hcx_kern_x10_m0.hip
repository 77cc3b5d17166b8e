// Size-specialised tracker kernel instances: n = G = 10, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_x10_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<10, 2, 0, true>;
    case 3: return track_kernel<10, 3, 0, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
