// scarl_k_h32p.hip -- FP32 Hermitian generator: the strided passes over the workspace.
#include "scarl_modeh.h"
SCARL_MODEH_STRIDED(float, f32, 4, 16, 32, false, plain)
