// Replica tracker kernel instances (FP64, R > 1 lane groups per path):
// generic, NMAX = 16.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_rep_g16(int deg) {
  switch (deg) {
    case 2: return track_kernel_rep<16, 2, false>;
    case 3: return track_kernel_rep<16, 3, false>;
    case 6: return track_kernel_rep<16, 6, false>;
    default: return nullptr;
  }
}

}  // namespace hcx
