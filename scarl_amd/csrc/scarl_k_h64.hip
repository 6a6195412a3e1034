// scarl_k_h64.hip -- FP64 instantiation of the Hermitian generator passes.
#include "scarl_modeh.h"
SCARL_MODEH_WRAPPERS(double, f64)
