// Replica tracker kernel instances (FP64, R > 1 lane groups per path):
// size-specialised, n = G in {8, 9, 10, 12, 16}.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_rep_xl(int n, int deg) {
  const bool d2 = deg <= 2;
  if (deg > 3) return nullptr;
  switch (n) {
    case 8: return d2 ? track_kernel_rep<8, 2, true> : track_kernel_rep<8, 3, true>;
    case 9: return d2 ? track_kernel_rep<9, 2, true> : track_kernel_rep<9, 3, true>;
    case 10: return d2 ? track_kernel_rep<10, 2, true> : track_kernel_rep<10, 3, true>;
    case 12: return d2 ? track_kernel_rep<12, 2, true> : track_kernel_rep<12, 3, true>;
    case 16: return d2 ? track_kernel_rep<16, 2, true> : track_kernel_rep<16, 3, true>;
    default: return nullptr;
  }
}

}  // namespace hcx
