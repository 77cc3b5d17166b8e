// Size-specialised tracker kernel instances: n = G = 12, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_x12_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<12, 2, 0, true>;
    case 3: return track_kernel<12, 3, 0, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
