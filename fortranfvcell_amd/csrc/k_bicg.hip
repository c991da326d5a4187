// k_bicg.hip -- instantiates the Jacobi-BiCGSTAB: two-phase (bicg2) and reference-order (bicg) kernels
// for the formats listed in fvc_instances.h.
#define FVC_INSTANTIATE
#include "fvc_instances.h"

FVCI_FORMATS_ALL(FVCI_BICG, template)
