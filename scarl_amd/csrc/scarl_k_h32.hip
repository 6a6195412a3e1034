// scarl_k_h32.hip -- FP32 instantiation of the Hermitian generator passes.
#include "scarl_modeh.h"
SCARL_MODEH_WRAPPERS(float, f32)
