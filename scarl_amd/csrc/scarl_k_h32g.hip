// scarl_k_h32g.hip -- FP32 Hermitian generator: the strided passes that generate their input.
#include "scarl_modeh.h"
SCARL_MODEH_STRIDED(float, f32, 4, 16, 32, true, gen)
