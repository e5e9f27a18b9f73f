// mixed-precision kernels (split-BF16 MFMA); filled in below
#include <hip/hip_runtime.h>
#include "ctem_internal.h"
