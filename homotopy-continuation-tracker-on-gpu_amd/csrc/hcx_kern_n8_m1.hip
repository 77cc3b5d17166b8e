// Tracker kernel instances for NMAX = 8, mixed FP32/FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_n8_m1(int deg) {
  switch (deg) {
    case 2: return track_kernel<8, 2, 1>;
    case 3: return track_kernel<8, 3, 1>;
    case 4: return track_kernel<8, 4, 1>;
    case 6: return track_kernel<8, 6, 1>;
    default: return nullptr;
  }
}

}  // namespace hcx
