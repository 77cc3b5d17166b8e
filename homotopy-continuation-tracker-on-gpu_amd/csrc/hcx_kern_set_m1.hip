// Multi-size tracker kernel instance (n in {3, 5, 6, 9, 12}, quadratic terms), mixed mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_set_m1() { return track_kernel_set<2, 1>; }

}  // namespace hcx
