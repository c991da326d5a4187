// fvc_plain.h -- the one-pass kernels of the C ABI: SpMV, dot, solver-state
// reset, the full-CSR Jacobi diagonal and getRes.  Included by the host
// translation unit only (they are not templates of a solver family).
#pragma once
#include "fvc_krylov.h"

namespace fvc {

// ---------------------------------------------------------- plain kernels --

template <int F>
__global__ __launch_bounds__(256) void spmv_kernel(Mat A, const double *__restrict__ x,
                                                   double *__restrict__ y) {
  const int i = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (i < A.neq) y[i] = row_apply<F>(A, i, XVec{x});
}

// a.b over G chunks, reduced by the last block into *out (device).
__global__ __launch_bounds__(1024) void dot_kernel(Grid g, const double *a, const double *b,
                                                   double *out) {
  __shared__ double lds[kMaxK * kMaxWaves];
  const int i0 = g.first(), i1 = g.row_end(), is = g.stride();
  double v[1] = {0.0};
  for (int i = i0; i < i1; i += is) v[0] += a[i] * b[i];
  grid_reduce<1>(g, v, lds, [&](const double *t) { *out = t[0]; });
}

// Resets the barrier counters and the solver state (a kernel rather than a
// memset, so that the sequence can be captured in a graph).
__global__ void reset_kernel(unsigned *ctr, int nctr, KState *st, int nst = 1) {
  for (int i = threadIdx.x; i < nctr; i += blockDim.x) ctr[i] = 0u;
  for (int i = threadIdx.x; i < nst; i += blockDim.x) {
    KState z{};
    z.status = S_MAXIT;
    st[i] = z;
  }
}

// The Jacobi diagonal of a full CSR matrix (diagonal stored in the row).
__global__ void csr_diag_kernel(int neq, const int *ia, const int *ja, const double *a,
                                double *m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= neq) return;
  double d = 0.0;
  for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k)
    if (ja[k] - 1 == i) d = a[k];
  m[i] = d;
}

// getRes (src/Solver.f:368-391)
__global__ void getres_kernel(double *u, double *x, const int *num, int numel, int flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numel) return;
  if (flag) x[num[i] - 1] = u[i];
  else u[i] = x[num[i] - 1];
}

}  // namespace fvc
