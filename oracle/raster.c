/* Oracle: the differentiable compositor in FP64 (filled in with the adjoint). */
#include "oracle.h"
