// Generic tracker kernel instance: NMAX = 16, terms of up to 2 variable factors, mixed FP32/FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g16_m1_d2() { return track_kernel<16, 2, 1, false>; }

}  // namespace hcx
