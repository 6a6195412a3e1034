// scarl_k_h64p.hip -- FP64 Hermitian generator: the strided passes over the workspace.
#include "scarl_modeh.h"
SCARL_MODEH_STRIDED(double, f64, 8, 16, 16, false, plain)
