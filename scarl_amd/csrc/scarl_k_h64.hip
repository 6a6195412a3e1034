// scarl_k_h64.hip -- FP64 Hermitian generator: the x-pass kernels without blend (compile-time planned
// and generic with 16 elements per thread), the kx = 0 plane and the launch dispatchers.
#include "scarl_modeh.h"
SCARL_MODEH_XPASS(double, f64, 8, 16, 16)
SCARL_MODEH_XPASS_PART(double, f64, 8, 16, 16, false, plain)
