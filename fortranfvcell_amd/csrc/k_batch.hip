// k_batch.hip -- instantiates the batched Jacobi-PCG (FP64 and mixed) in one persistent launch
// for the formats listed in fvc_instances.h.
#define FVC_INSTANTIATE
#include "fvc_instances.h"

FVCI_FORMATS_BATCH(FVCI_BATCH, template)
