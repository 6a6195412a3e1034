// scarl_k_h64b.hip -- FP64 Hermitian generator: the x-pass kernels of a blend.
#include "scarl_modeh.h"
SCARL_MODEH_XPASS_PART(double, f64, 8, 16, 16, true, blend)
