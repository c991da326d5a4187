// fvc_post.h -- the post-solve step of a SIMPLE iteration, on the device:
// the pressure-correction gradient of gform (src/Pform.f:466-676) for the
// cells that take cell_si_p_i_ggl (types 63/65/73/75, src/Celllib.f:127-157)
// with iws = 2 (src/CellibPressure.f:390-463,611-650), and simpleUpdate
// (src/Simple.f:871-898), so that p', its gradient and the corrected
// velocities stay in HBM after the solve.
//
// A cell's new gradient depends only on its own value and old gradient, its
// neighbours' values and the vertices of it and its neighbours, so one
// thread computes one cell, in place.  Every operation is the reference's,
// in its order (the library builds with -ffp-contract=off), so results are
// bit-identical to it.  Arrays are the reference's Fortran arrays: x(2,
// nnode), ix(nen+1, numel) (1-based nodes, then the material), nelcon(nen,
// numel) (1-based neighbours, negative across a wall), pedge(nen+1, numel),
// u(numel), grad(2, numel).
#pragma once
#include "fvc_device.h"

namespace fvc {

// A 1-based index into [1, n], as a 0-based one; out of range: 0, and ok
// cleared (the cell then gets NaN instead of a read out of bounds).
__device__ __forceinline__ int checked(int one_based, int n, bool &ok) {
  const int i = one_based - 1;
  const bool in = i >= 0 && i < n;
  ok = ok && in;
  return in ? i : 0;
}

// cellGeom2D's centre: the vertex sum, times 1/nshared (src/Celllib.f:725-733)
template <int W>
__device__ __forceinline__ void cell_centre(const double *xn, int nnode, const int *ixc, double &cx,
                                            double &cy, bool &ok) {
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int no = checked(ixc[l], nnode, ok);
    sx = sx + xn[2 * (size_t)no];
    sy = sy + xn[2 * (size_t)no + 1];
  }
  cx = (1.0 / W) * sx;
  cy = (1.0 / W) * sy;
}

// |a x b| of in-plane vectors: vet (src/Matvec.f:304-316) with zero third
// components, then sqrt(dot_local(c, c, 3)) (src/Matvec.f:404-421)
__device__ __forceinline__ double cross_norm(double ax, double ay, double bx, double by) {
  const double c1 = ay * 0.0 - 0.0 * by;
  const double c2 = 0.0 * bx - ax * 0.0;
  const double c3 = ax * by - ay * bx;
  return sqrt(((0.0 + c1 * c1) + c2 * c2) + c3 * c3);
}

template <int W>
__global__ __launch_bounds__(256) void gform_p_kernel(int numel, int nnode, const double *xn,
                                                      const int *ix, const int *nelcon,
                                                      const int *pedge, const double *u,
                                                      double *grad, double *fluxl) {
  const int c = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (c >= numel) return;
  bool ok = true;
  const int *ixc = ix + (size_t)(W + 1) * c;
  double px[W], py[W];
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int no = checked(ixc[l], nnode, ok);
    px[l] = xn[2 * (size_t)no];
    py[l] = xn[2 * (size_t)no + 1];
  }
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int l = 0; l < W; ++l) {
    sx = sx + px[l];
    sy = sy + py[l];
  }
  const double xcx = (1.0 / W) * sx, xcy = (1.0 / W) * sy;
  const double uc = u[c];
  const double g0 = grad[2 * (size_t)c], g1 = grad[2 * (size_t)c + 1];
  // neighbour centres, or ghost centres across walls, and their values
  double qx[W], qy[W], ug[W];
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int a = l, b = (l + 1) % W;
    // the edge's unit normal and the centre-to-edge distance
    // (src/Celllib.f:737-808)
    const double ex = px[b] - px[a], ey = py[b] - py[a];
    const double meta = sqrt((0.0 + ex * ex) + ey * ey);
    const double nx = ey / meta, ny = -(ex / meta);
    const double xmx = 0.5 * (px[b] + px[a]), xmy = 0.5 * (py[b] + py[a]);
    const double ca = (xmx - xcx) * nx + (xmy - xcy) * ny;
    const int nb = nelcon[(size_t)W * c + l];
    if (nb > 0) {
      const int k = checked(nb, numel, ok);
      cell_centre<W>(xn, nnode, ix + (size_t)(W + 1) * k, qx[l], qy[l], ok);
      ug[l] = u[k];
    } else {
      // src/CellibPressure.f:438-458 with iws = 2
      const double gama = 2.0 * ca;
      qx[l] = xcx + gama * nx;
      qy[l] = xcy + gama * ny;
      const int pe = pedge[(size_t)(W + 1) * c + l];
      ug[l] = pe == 3 ? 0.0 : pe == 6 ? uc : uc + gama * (g0 * nx + g1 * ny);
    }
  }
  // Green-Gauss over the polygon of those centres (src/CellibPressure.f:611-650)
  double Ex[W], Ey[W], um[W];
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int a = l, b = (l + 1) % W;
    Ex[l] = qx[b] - qx[a];
    Ey[l] = qy[b] - qy[a];
    um[l] = 0.5 * (ug[b] + ug[a]);
  }
  double area;
  if constexpr (W == 3) {
    area = 0.5 * cross_norm(Ex[0], Ey[0], Ex[1], Ey[1]);
  } else {
    double s = 0.5 * cross_norm(Ex[0], Ey[0], Ex[1], Ey[1]);
    s = s + 0.5 * cross_norm(Ex[2], Ey[2], Ex[3], Ey[3]);
    s = s + 0.5 * cross_norm(Ex[0], Ey[0], Ex[3], Ey[3]);
    s = s + 0.5 * cross_norm(Ex[1], Ey[1], Ex[2], Ey[2]);
    area = 0.5 * s;
  }
  double gx = 0.0, gy = 0.0;
#pragma unroll
  for (int l = 0; l < W; ++l) {
    gx = gx + um[l] * Ey[l];
    gy = gy + um[l] * (-Ex[l]);
  }
  const double bad = __builtin_nan("");
  grad[2 * (size_t)c] = ok ? gx / area : bad;
  grad[2 * (size_t)c + 1] = ok ? gy / area : bad;
  fluxl[c] = ok ? 1.0 : bad;
}

// simpleUpdate (src/Simple.f:885-896): w = u* - d grad p'; p += underPc p'
// unless SIMPLEC
__global__ __launch_bounds__(256) void simple_update_kernel(int numel, double *w, double *pre,
                                                            const double *u1, const double *u2,
                                                            const double *prec, const double *gradpc,
                                                            const double *d, double under_pc,
                                                            int simple_r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numel) return;
  const size_t k = 2 * (size_t)i;
  w[k] = u1[i] - d[k] * gradpc[k];
  w[k + 1] = u2[i] - d[k + 1] * gradpc[k + 1];
  if (!simple_r) pre[i] = pre[i] + under_pc * prec[i];
}

}  // namespace fvc
