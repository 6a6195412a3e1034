// scarl_k_h64g.hip -- FP64 Hermitian generator: the strided passes that generate their input.
#include "scarl_modeh.h"
SCARL_MODEH_STRIDED(double, f64, 8, 16, 16, true, gen)
