// Multi-size tracker kernel instance (n in {3, 5, 6, 9, 12}, quadratic terms), FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_set_m0() { return track_kernel_set<2, 0>; }

}  // namespace hcx
