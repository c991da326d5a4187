// fvc_instances.h -- the solver kernel templates and the formats each is
// built for.  Every family is explicitly instantiated in its own translation
// unit (k_*.hip), so the library's device code compiles in parallel; the
// host code (fvcell.hip) sees explicit instantiation declarations and
// launches them through their kernel handles.
#pragma once
#include "fvc_krylov.h"
#include "fvc_mixed.h"
#include "fvc_mixed_krylov.h"
#include "fvc_batch.h"
#include "fvc_gmres_lag.h"

// Device formats (Fmt): every format for the FP64 solvers, ELLPACK and CSRD
// for mixed precision and for batches.
#define FVCI_FORMATS_ALL(X, P) X(P, 0) X(P, 1) X(P, 2) X(P, 3) X(P, 4) X(P, 5) X(P, 6)
#define FVCI_FORMATS_MIXED(X, P) X(P, 0) X(P, 1) X(P, 2) X(P, 3)
#define FVCI_FORMATS_BATCH(X, P) X(P, 0) X(P, 1) X(P, 2) X(P, 3)

#define FVCI_CG(P, F)                                      \
  P __global__ void fvc::cg1_persistent<F>(fvc::Cg1Args); \
  P __global__ void fvc::cg1_kernel<F>(fvc::Cg1Args);     \
  P __global__ void fvc::pcg_persistent<F>(fvc::PcgArgs); \
  P __global__ void fvc::pcg_kernel<F>(fvc::PcgArgs);
#define FVCI_BICG(P, F)                                        \
  P __global__ void fvc::bicg2_persistent<F>(fvc::Bicg2Args); \
  P __global__ void fvc::bicg2_kernel<F>(fvc::Bicg2Args);     \
  P __global__ void fvc::bicg_persistent<F>(fvc::BicgArgs);   \
  P __global__ void fvc::bicg_kernel<F>(fvc::BicgArgs);
#define FVCI_GMRES(P, F)                                            \
  P __global__ void fvc::gmres_persistent<F, 0>(fvc::GmresArgs); \
  P __global__ void fvc::gmres_kernel<F, 0>(fvc::GmresArgs);     \
  P __global__ void fvc::gmres_persistent<F, 1>(fvc::GmresArgs); \
  P __global__ void fvc::gmres_kernel<F, 1>(fvc::GmresArgs);     \
  P __global__ void fvc::gmres_lag_d_kernel<F>(fvc::GmresArgs); \
  P __global__ void fvc::gmres_lag_r_kernel<F>(fvc::GmresArgs); \
  P __global__ void fvc::gmres_lag_b_kernel<F>(fvc::GmresArgs);
#define FVCI_MIXED(P, F)                                           \
  P __global__ void fvc::mgmres_kernel<F>(fvc::MGmresArgs);      \
  P __global__ void fvc::mgmres_lag_d_kernel<F>(fvc::MGmresArgs); \
  P __global__ void fvc::mgmres_lag_b_kernel<F>(fvc::MGmresArgs); \
  P __global__ void fvc::mixed_persistent<F>(fvc::MixedArgs);   \
  P __global__ void fvc::mixed_kernel<F>(fvc::MixedArgs);       \
  P __global__ void fvc::mbicg_persistent<F>(fvc::MBicgArgs);   \
  P __global__ void fvc::mbicg_kernel<F>(fvc::MBicgArgs);       \
  P __global__ void fvc::mgmres_persistent<F>(fvc::MGmresArgs); \
  P __global__ void fvc::scale_to_f32_kernel<F>(fvc::Mat, const double *, float *, double *, float *);
#define FVCI_BATCH(P, F)                                               \
  P __global__ void fvc::pcg_batch_persistent<F>(fvc::BatchArgs);   \
  P __global__ void fvc::mixed_batch_persistent<F>(fvc::BatchArgs); \
  P __global__ void fvc::batch_scale_kernel<F>(fvc::BatchArgs, int, const int *, const int *);
#define FVCI_COLOFF(P, W) P __global__ void fvc::col_offsets_kernel<W>(int, const int *, const float *, short *, int *);

// In the host translation unit: declarations only.
#ifndef FVC_INSTANTIATE
FVCI_FORMATS_ALL(FVCI_CG, extern template)
FVCI_FORMATS_ALL(FVCI_BICG, extern template)
FVCI_FORMATS_ALL(FVCI_GMRES, extern template)
FVCI_FORMATS_MIXED(FVCI_MIXED, extern template)
FVCI_FORMATS_BATCH(FVCI_BATCH, extern template)
FVCI_COLOFF(extern template, 3)
FVCI_COLOFF(extern template, 4)
#endif
