// Generic tracker kernel instance: NMAX = 8, terms of up to 2 variable factors, mixed FP32/FP64 mode.
#include "hcx_kernel.hpp"

namespace hcx {

KernelFn kernel_g8_m1_d2() { return track_kernel<8, 2, 1, false>; }

}  // namespace hcx
