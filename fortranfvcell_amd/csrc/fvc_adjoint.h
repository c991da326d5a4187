// fvc_adjoint.h -- kernels of the differentiable solve.
//
// The adjoint of A T = b is A^T lam = dL/dT.  For an ELLPACK matrix whose
// pattern is symmetric (finite-volume adjacency), A^T has the same ja; only
// the values move: A^T(i, s) = A(j, t) with j = ja(s,i) and t the slot of row
// j that points back at i.  The adjoint solve is then the ordinary
// BiCGSTAB on (ad, a^T, ja).
//
// The gradient of L with respect to the per-cell conductivity k follows from
// dL/dk = -lam^T (dA/dk T - db/dk) and the reference's coefficient forms:
//   interior face (c, N): D_f = k_f * geo, 1/k_f = a/k_c + (1-a)/k_N
//     (src/CellibTrans.f:510-511,527); D_f enters row c as +D_f T_c - D_f T_N
//     and row N symmetrically, so its term is -(lam_c - lam_N)(T_c - T_N)
//     * geo * dk_f/dk_c with dk_f/dk_c = k_f^2 a / k_c^2;
//   Dirichlet face: a_P += k_c |S|/d and b += k_c |S|/d T_wall
//     (src/CellibTrans.f:488-491): term -lam_c (T_c - T_wall) |S|/d.
// One lane per cell sums its own faces; no atomics.
#pragma once
#include <hip/hip_runtime.h>

namespace fvc {

__global__ void ell_transpose_kernel(int neq, int w, const int *__restrict__ ja,
                                     const int *__restrict__ back, const double *__restrict__ a,
                                     double *__restrict__ at) {
  const int i = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (i >= neq) return;
  for (int s = 0; s < w; ++s) {
    const size_t e = (size_t)w * i + s;
    const int t = back[e];
    at[e] = (t >= 0) ? a[(size_t)w * (ja[e] - 1) + t] : 0.0;
  }
}

__global__ void grad_k_kernel(int neq, int w, const int *__restrict__ ja, const double *__restrict__ k,
                              const double *__restrict__ T, const double *__restrict__ lam,
                              const double *__restrict__ geo, const double *__restrict__ alpha,
                              const double *__restrict__ wall_g, const double *__restrict__ wall_T,
                              double *__restrict__ grad) {
  const int c = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (c >= neq) return;
  const double kc = k[c], Tc = T[c], lc = lam[c];
  double acc = 0.0;
  for (int s = 0; s < w; ++s) {
    const size_t e = (size_t)w * c + s;
    const double gf = geo[e];
    if (gf != 0.0) {
      const int j = ja[e] - 1;
      const double a = alpha[e];
      const double kf = 1.0 / (a / kc + (1.0 - a) / k[j]);
      const double dkf = kf * kf * a / (kc * kc);
      acc += -(lc - lam[j]) * (Tc - T[j]) * gf * dkf;
    }
    const double wg = wall_g[e];
    if (wg != 0.0) acc += -lc * (Tc - wall_T[e]) * wg;
  }
  grad[c] = acc;
}

// Face-based assembly of the heat / transport operator A(k) and b(k) in the
// reference's ELLPACK layout, from per-cell conductivities and fixed
// geometry and flux data (the coefficient forms of src/CellibTrans.f):
//   interior face s of cell c, neighbour N:
//     1/k_f = a/k_c + (1-a)/k_N                  (src/CellibTrans.f:510-511)
//     D_f = k_f * geo,  a_nb = D_f + conv        (:527, src/CellibSimple.f:789-816)
//     al(s, c) = -a_nb                           (src/EllPack.f:85)
//   Dirichlet face: ap_w = k_c * wall_g, added to the diagonal and, times
//     T_wall, to b                               (src/CellibTrans.f:488-491)
//   ad = (sum a_nb + convF) + sum ap_w           (src/CellibSimple.f:856-859)
// One lane per cell, slots summed left to right: the same operations in the
// same order as the host assembly, so the result is bit-identical to it.
__global__ void assemble_heat_kernel(int neq, int w, const int *__restrict__ ja, const double *__restrict__ k,
                                     const double *__restrict__ q, const double *__restrict__ geo,
                                     const double *__restrict__ alpha, const double *__restrict__ conv,
                                     const double *__restrict__ conv_f, const double *__restrict__ wall_g,
                                     const double *__restrict__ wall_T, double *__restrict__ ad,
                                     double *__restrict__ al, double *__restrict__ b) {
  const int c = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (c >= neq) return;
  const double kc = k[c];
  double sa = 0.0, sw = 0.0, sb = 0.0;
  for (int s = 0; s < w; ++s) {
    const size_t e = (size_t)w * c + s;
    const double gf = geo[e];
    double anb = 0.0;
    if (gf != 0.0) {
      const double a = alpha[e];
      const double kf = 1.0 / (a / kc + (1.0 - a) / k[ja[e] - 1]);
      anb = kf * gf + conv[e];
      al[e] = -anb;
    } else {
      al[e] = 0.0;
    }
    sa = sa + anb;
    const double apw = kc * wall_g[e];
    sw = sw + apw;
    sb = sb + apw * wall_T[e];
  }
  ad[c] = (sa + conv_f[c]) + sw;
  b[c] = q ? sb + q[c] : sb;
}

// Plain device copy, four 16-byte loads in flight per lane (bandwidth reference).
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace fvc
