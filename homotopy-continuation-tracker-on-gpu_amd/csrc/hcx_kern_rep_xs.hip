// Replica tracker kernel instances (FP64, R > 1 lane groups per path):
// size-specialised, n = G in {3, 5, 6}.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_rep_xs(int n, int deg) {
  const bool d2 = deg <= 2;
  switch (n) {
    case 3: return d2 ? track_kernel_rep<3, 2, true> : track_kernel_rep<3, 3, true>;
    case 5: return d2 ? track_kernel_rep<5, 2, true> : track_kernel_rep<5, 3, true>;
    case 6: return d2 ? track_kernel_rep<6, 2, true> : track_kernel_rep<6, 3, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
