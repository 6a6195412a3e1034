// scarl_k_h32.hip -- FP32 instantiation of the Hermitian generator passes: generic kernels with
// register budgets of 16 and 32 elements per thread, plus the compile-time planned FP32 kernels.
#include "scarl_modeh.h"
SCARL_MODEH_WRAPPERS(float, f32, 4, 16, 32)
