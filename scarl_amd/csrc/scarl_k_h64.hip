// scarl_k_h64.hip -- FP64 instantiation of the Hermitian generator passes (16 elements per thread)
// plus the compile-time planned FP64 kernels.
#include "scarl_modeh.h"
SCARL_MODEH_WRAPPERS(double, f64, 8, 16, 16)
