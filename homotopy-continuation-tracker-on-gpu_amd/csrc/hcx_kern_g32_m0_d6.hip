// Generic tracker kernel instance: NMAX = 32, terms of up to 6 variable factors, FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g32_m0_d6() { return track_kernel<32, 6, 0, false>; }

}  // namespace hcx
