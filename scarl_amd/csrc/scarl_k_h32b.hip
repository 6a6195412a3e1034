// scarl_k_h32b.hip -- FP32 Hermitian generator: the x-pass kernels of a blend.
#include "scarl_modeh.h"
SCARL_MODEH_XPASS_PART(float, f32, 4, 16, 32, true, blend)
