// k_gmres.hip -- instantiates the GMRES(k): modified Gram-Schmidt and CGS2 kernels
// for the formats listed in fvc_instances.h.
#define FVC_INSTANTIATE
#include "fvc_instances.h"

FVCI_FORMATS_ALL(FVCI_GMRES, template)
