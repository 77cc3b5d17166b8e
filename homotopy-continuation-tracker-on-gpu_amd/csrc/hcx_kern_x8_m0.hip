// Size-specialised tracker kernel instances: n = G = 8, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_x8_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<8, 2, 0, true>;
    case 3: return track_kernel<8, 3, 0, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
