// k_mixed.hip -- instantiates the mixed-precision PCG: FP32 inner iterations, FP64 refreshes, FP32 set-up
// for the formats listed in fvc_instances.h.
#define FVC_INSTANTIATE
#include "fvc_instances.h"

FVCI_FORMATS_MIXED(FVCI_MIXED, template)
FVCI_COLOFF(template, 3)
FVCI_COLOFF(template, 4)
