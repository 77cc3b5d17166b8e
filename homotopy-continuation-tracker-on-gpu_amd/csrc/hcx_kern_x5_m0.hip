// Size-specialised tracker kernel instances: n = G = 5, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_x5_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<5, 2, 0, true>;
    case 3: return track_kernel<5, 3, 0, true>;
    case 4: return track_kernel<5, 4, 0, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
