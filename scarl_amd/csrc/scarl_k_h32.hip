// scarl_k_h32.hip -- FP32 Hermitian generator: the x-pass kernels without blend (compile-time planned
// and generic with 16 or 32 elements per thread), the kx = 0 plane and the launch dispatchers.
#include "scarl_modeh.h"
SCARL_MODEH_XPASS(float, f32, 4, 16, 32)
SCARL_MODEH_XPASS_PART(float, f32, 4, 16, 32, false, plain)
