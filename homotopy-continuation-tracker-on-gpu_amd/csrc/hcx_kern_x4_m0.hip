// Size-specialised tracker kernel instances: n = G = 4, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_x4_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<4, 2, 0, true>;
    case 3: return track_kernel<4, 3, 0, true>;
    case 4: return track_kernel<4, 4, 0, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
