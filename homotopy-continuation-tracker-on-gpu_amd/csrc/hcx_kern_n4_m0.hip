// Tracker kernel instances for NMAX = 4, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_n4_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<4, 2, 0>;
    case 3: return track_kernel<4, 3, 0>;
    case 4: return track_kernel<4, 4, 0>;
    case 6: return track_kernel<4, 6, 0>;
    default: return nullptr;
  }
}

}  // namespace hcx
