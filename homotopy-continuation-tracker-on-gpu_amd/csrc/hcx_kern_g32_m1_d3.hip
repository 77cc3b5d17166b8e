// Generic tracker kernel instance: NMAX = 32, terms of up to 3 variable factors, mixed FP32/FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g32_m1_d3() { return track_kernel<32, 3, 1, false>; }

}  // namespace hcx
