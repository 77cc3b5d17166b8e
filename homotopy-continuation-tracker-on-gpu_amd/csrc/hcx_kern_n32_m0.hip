// Tracker kernel instances for NMAX = 32, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_n32_m0(int deg) {
  switch (deg) {
    case 2: return track_kernel<32, 2, 0>;
    case 3: return track_kernel<32, 3, 0>;
    case 4: return track_kernel<32, 4, 0>;
    case 6: return track_kernel<32, 6, 0>;
    default: return nullptr;
  }
}

}  // namespace hcx
