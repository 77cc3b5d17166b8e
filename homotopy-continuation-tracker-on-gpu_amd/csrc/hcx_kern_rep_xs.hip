// Replica tracker kernel instances (FP64, R > 1 lane groups per path):
// size-specialised, n = G in {3, 4, 5, 6} (degree 4: n = 4, 5).
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_rep_xs(int n, int deg) {
  const bool d2 = deg <= 2;
  if (deg == 4 && n != 4 && n != 5) return nullptr;
  switch (n) {
    case 3: return d2 ? track_kernel_rep<3, 2, true> : track_kernel_rep<3, 3, true>;
    case 4: return d2 ? track_kernel_rep<4, 2, true> : deg == 3 ? track_kernel_rep<4, 3, true> : track_kernel_rep<4, 4, true>;
    case 5: return d2 ? track_kernel_rep<5, 2, true> : deg == 3 ? track_kernel_rep<5, 3, true> : track_kernel_rep<5, 4, true>;
    case 6: return d2 ? track_kernel_rep<6, 2, true> : track_kernel_rep<6, 3, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
