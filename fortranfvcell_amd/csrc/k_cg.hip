// k_cg.hip -- instantiates the Jacobi-PCG: one-reduction (cg1) and reference-order (pcg) phases
// for the formats listed in fvc_instances.h.
#define FVC_INSTANTIATE
#include "fvc_instances.h"

FVCI_FORMATS_ALL(FVCI_CG, template)
