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

// The pressure-correction system as pform assembles it for SIMPLE
// (src/Pform.f:65-436, called at src/Simple.f:583-588 with iws = 1):
// cell_si_p_i_ggl (src/CellibPressure.f:390-571) per cell, into ELLPACK as
// ellPack does (src/EllPack.f:68-93).  One thread per cell: the cell's
// geometry (cellGeom2D, src/Celllib.f:695-808), where each centre line
// crosses its edge (vectorKm2d, src/Celllib.f:596-656), the face
// coefficients and the face velocity of cellVel (src/CellibSimple.f:
// 1947-2036; int_vel 1 linear, 2 Rhie-Chow, 5 Ferziger-Peric), in the
// reference's order.  Writes row num(c): ad, al(W, neq) (-a_nb, 0 where the
// neighbour has no equation), b, and rescell(c) = the cell's mass residual.
// Arrays beyond gform's: num(numel) (1-based equations), e(10, nmat)
// (density in row 2), sedge(3, W+1, numel), w, d, gradp (2, numel), p.
template <int W>
__global__ __launch_bounds__(256) void assemble_pc_kernel(
    int numel, int nnode, int neq, int nmat, int int_vel, const double *xn, const int *ix, const int *nelcon,
    const int *pedge, const int *num, const double *e, const double *sedge, const double *w, const double *d,
    const double *gradp, const double *p, double *ad, double *al, double *b, double *rescell) {
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
  const double rho = e[10 * (size_t)checked(ixc[W], nmat, ok) + 1];
  const double wc0 = w[2 * (size_t)c], wc1 = w[2 * (size_t)c + 1];
  const double dc0 = d[2 * (size_t)c], dc1 = d[2 * (size_t)c + 1];
  const double gc0 = gradp[2 * (size_t)c], gc1 = gradp[2 * (size_t)c + 1];
  const double pc = p[c];
  double pz = 0.0, sp = 0.0;
  double a[W];
  int ld[W];
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int ea = l, eb = (l + 1) % W;
    const double ex = px[eb] - px[ea], ey = py[eb] - py[ea];
    const double meta = sqrt((0.0 + ex * ex) + ey * ey);
    const double nx = ey / meta, ny = -(ex / meta);
    const double xmx = 0.5 * (px[eb] + px[ea]), xmy = 0.5 * (py[eb] + py[ea]);
    const double ca = (xmx - xcx) * nx + (xmy - xcy) * ny;
    const int nbr = nelcon[(size_t)W * c + l];
    a[l] = 0.0;
    ld[l] = 0;
    if (nbr > 0) {
      const int k = checked(nbr, numel, ok);
      double qx, qy;
      cell_centre<W>(xn, nnode, ix + (size_t)(W + 1) * k, qx, qy, ok);
      const double k0 = qx - xcx, k1 = qy - xcy;
      const double mksi = sqrt((0.0 + k0 * k0) + k1 * k1);
      const double ksi0 = k0 / mksi, ksi1 = k1 / mksi;
      // vectorKm2d: the centre line against the edge line
      const double a1 = qy - xcy, b1 = xcx - qx, c1 = qx * xcy - xcx * qy;
      const double a2 = py[ea] - py[eb], b2 = px[eb] - px[ea], c2 = px[ea] * py[eb] - px[eb] * py[ea];
      const double det = a1 * b2 - a2 * b1;
      ok = ok && det != 0.0;
      const double r = 1.0 / det;
      const double f1 = -c1, f2 = -c2;
      const double xi0 = (r * b2) * f1 + (-(r * b1)) * f2;
      const double xi1 = (-(r * a2)) * f1 + (r * a1) * f2;
      const double s0 = xi0 - xcx, s1 = xi1 - xcy;
      const double mksif = sqrt(s0 * s0 + s1 * s1);
      // src/CellibPressure.f:511-553
      const double alpha = mksif / mksi, om = 1.0 - alpha;
      const double dd0 = om * dc0 + alpha * d[2 * (size_t)k];
      const double dd1 = om * dc1 + alpha * d[2 * (size_t)k + 1];
      const double nk = nx * ksi0 + ny * ksi1;
      double kf, ndd;
      if (int_vel == 5) {
        kf = om * rho + alpha * e[10 * (size_t)checked(ix[(size_t)(W + 1) * k + W], nmat, ok) + 1];
        ndd = 0.5 * (dd0 + dd1);
      } else {
        kf = rho;
        ndd = (dd0 * nx) * nx + (dd1 * ny) * ny;
      }
      a[l] = ((kf * meta) * ndd) / (nk * mksi);
      // cellVel (src/CellibSimple.f:1956-2036)
      const double v0 = om * wc0 + alpha * w[2 * (size_t)k];
      const double v1 = om * wc1 + alpha * w[2 * (size_t)k + 1];
      double wf = v0 * nx + v1 * ny;
      if (int_vel != 1) {
        const double gf0 = om * gc0 + alpha * gradp[2 * (size_t)k];
        const double gf1 = om * gc1 + alpha * gradp[2 * (size_t)k + 1];
        if (int_vel == 2) {
          const double gfp1 = (p[k] - pc) / mksi;
          const double gfp2 = gf0 * ksi0 + gf1 * ksi1;
          wf = wf + (0.5 * (dd0 + dd1)) * (gfp2 - gfp1);
        } else {
          const double gfp1 = p[k] - pc;
          const double ks0 = mksi * ksi0, ks1 = mksi * ksi1;
          const double nks = ks0 * nx + ks1 * ny;
          double gfp2 = gf0 * ks0 + gf1 * ks1;
          gfp2 = (gfp2 - gfp1) / nks;
          wf = wf + (0.5 * (dd0 + dd1)) * gfp2;
        }
      }
      wf = wf * meta;
      pz = pz - kf * wf;
      ld[l] = num[k];
    } else {
      // walls (src/CellibPressure.f:476-507)
      const int pe = pedge[(size_t)(W + 1) * c + l];
      if (pe == 2 || pe == 4) {
        const double *se = sedge + 3 * ((size_t)(W + 1) * c + l);
        const double wfn = se[0] * nx + se[1] * ny;
        pz = pz - (rho * wfn) * meta;
      } else if (pe == 3) {
        const double ndd = (dc0 * nx) * nx + (dc1 * ny) * ny;
        sp = sp + ((rho * ndd) * meta) / ca;
        const double wfn = wc0 * nx + wc1 * ny;
        pz = pz - (rho * wfn) * meta;
      }
    }
  }
  // the pin of the pressure level (src/CellibPressure.f:557)
  if (pedge[(size_t)(W + 1) * c + W] == -1) sp = sp + 1.0e60;
  double apc = sp;
#pragma unroll
  for (int l = 0; l < W; ++l) apc = apc + a[l];
  const int eq = num[c] - 1;
  const double bad = __builtin_nan("");
  rescell[c] = ok ? pz : bad;
  if (eq < 0 || eq >= neq) return;
  ad[eq] = ok ? apc : bad;
  b[eq] = ok ? pz : bad;
#pragma unroll
  for (int l = 0; l < W; ++l) al[(size_t)W * eq + l] = !ok ? bad : ld[l] > 0 ? -a[l] : 0.0;
}

// The TVD limiters of the momentum convection, limit (src/Celllib.f:
// 370-431) by code 1..7; ok cleared for any other code (code 8 reads
// parameters the momentum cell does not pass).
__device__ __forceinline__ double tvd_limit(double r, int icod, bool &ok) {
  switch (icod) {
    case 1: return 0.0;
    case 2: return 1.0;
    case 3: return (r + fabs(r)) / (1.0 + fabs(r));
    case 4: return r > 0.0 ? (r + r * r) / (1.0 + r * r) : 0.0;
    case 5: return fmax(0.0, fmin(r, 1.0));
    case 6: return fmax(0.0, fmin(2.0, r));
    case 7: return fmax(0.0, fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0)));
    default: ok = false; return 0.0;
  }
}

// Arguments of assemble_mom_kernel: the mesh as for gform; num(numel);
// e(10, nmat) (viscosity, density, ..., row 6 the transient coefficient,
// row 7 the limiter code); sedge(3, W+1, numel); u, pres (numel): the
// velocity component comp and the pressure; w, grad1 (of u), grad2 (of the
// other component), gradp (2, numel); d (2, numel): row comp written.
struct MomArgs {
  int numel, nnode, neq, nmat, comp;   // comp: 0 (u1) or 1 (u2)
  int s_pressure, simple_c, bs;
  double under_u, dt;
  const double *x;
  const int *ix, *nelcon, *pedge, *num;
  const double *e, *sedge, *u, *pres, *w, *grad1, *grad2, *gradp;
  double *d, *ad, *al, *b, *rescell, *im;
};

// A momentum system as pform assembles it for SIMPLE (src/Pform.f:65-436,
// called at src/Simple.f:425-431,456-462 with iws = 1, lib = 1):
// cell_si_eb_i_ggl (src/CellibSimple.f:591-901) per cell, into ELLPACK as
// ellPack does (src/EllPack.f:68-93).  Per face: the harmonic-mean
// viscosity, the diffusion coefficient with Darwish's gradient correction,
// first-order upwind convection with the TVD deferred correction, the
// pressure on the face (second-order reconstruction with sPressure), and
// the non-linear viscous terms; walls by their condition (0 wall, 1 moving
// wall, 2 prescribed velocity, 3 prescribed pressure, 4 locally parabolic);
// then the pin (pedge -2), the transient diagonal and the under-relaxation.
// Writes row num(c) of ad, al, b; rescell(c); im(comp, c), im(comp+2, c);
// d(comp, c).  The reference's order throughout, so bit-identical.
template <int W>
__global__ __launch_bounds__(256) void assemble_mom_kernel(MomArgs A) {
  const int c = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (c >= A.numel) return;
  bool ok = true;
  const int cp = A.comp;
  const int *ixc = A.ix + (size_t)(W + 1) * c;
  double px[W], py[W];
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int no = checked(ixc[l], A.nnode, ok);
    px[l] = A.x[2 * (size_t)no];
    py[l] = A.x[2 * (size_t)no + 1];
  }
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int l = 0; l < W; ++l) {
    sx = sx + px[l];
    sy = sy + py[l];
  }
  const double xcx = (1.0 / W) * sx, xcy = (1.0 / W) * sy;
  // the cell's area from its raw edge vectors (areacell, src/Celllib.f:509-522)
  double area;
  {
    double Ex[W], Ey[W];
#pragma unroll
    for (int l = 0; l < W; ++l) {
      Ex[l] = px[(l + 1) % W] - px[l];
      Ey[l] = py[(l + 1) % W] - py[l];
    }
    if constexpr (W == 3) {
      area = 0.5 * cross_norm(Ex[0], Ey[0], Ex[1], Ey[1]);
    } else {
      double s = 0.5 * cross_norm(Ex[0], Ey[0], Ex[1], Ey[1]);
      s = s + 0.5 * cross_norm(Ex[2], Ey[2], Ex[3], Ey[3]);
      s = s + 0.5 * cross_norm(Ex[0], Ey[0], Ex[3], Ey[3]);
      s = s + 0.5 * cross_norm(Ex[1], Ey[1], Ex[2], Ey[2]);
      area = 0.5 * s;
    }
  }
  const double *ec = A.e + 10 * (size_t)checked(ixc[W], A.nmat, ok);
  const double visc = ec[0], rho = ec[1], cc = ec[5];
  const int icod = (int)ec[6];
  const double uc = A.u[c], prc = A.pres[c];
  const double w0 = A.w[2 * (size_t)c], w1 = A.w[2 * (size_t)c + 1];
  const double g1c0 = A.grad1[2 * (size_t)c], g1c1 = A.grad1[2 * (size_t)c + 1];
  const double g2c0 = A.grad2[2 * (size_t)c], g2c1 = A.grad2[2 * (size_t)c + 1];
  const double gpc0 = A.gradp[2 * (size_t)c], gpc1 = A.gradp[2 * (size_t)c + 1];
  const double eps = 1.0e-14;
  double p = 0.0, pf = 0.0, sp = 0.0;
  double a[W], un[W];
  int ld[W];
#pragma unroll
  for (int l = 0; l < W; ++l) {
    const int ea = l, eb = (l + 1) % W;
    const double ex = px[eb] - px[ea], ey = py[eb] - py[ea];
    const double meta = sqrt((0.0 + ex * ex) + ey * ey);
    const double nx = ey / meta, ny = -(ex / meta);
    const double ncp = cp == 0 ? nx : ny;
    const double xmx = 0.5 * (px[eb] + px[ea]), xmy = 0.5 * (py[eb] + py[ea]);
    const double df0 = xmx - xcx, df1 = xmy - xcy;
    const double ca = df0 * nx + df1 * ny;
    const int nbr = A.nelcon[(size_t)W * c + l];
    a[l] = 0.0;
    un[l] = 0.0;
    ld[l] = 0;
    if (nbr > 0) {
      const int k = checked(nbr, A.numel, ok);
      double qx, qy;
      cell_centre<W>(A.x, A.nnode, A.ix + (size_t)(W + 1) * k, qx, qy, ok);
      const double k0 = qx - xcx, k1 = qy - xcy;
      const double mksi = sqrt((0.0 + k0 * k0) + k1 * k1);
      const double ksi0 = k0 / mksi, ksi1 = k1 / mksi;
      // vectorKm2d (src/Celllib.f:596-656)
      const double a1 = qy - xcy, b1 = xcx - qx, c1 = qx * xcy - xcx * qy;
      const double a2 = py[ea] - py[eb], b2 = px[eb] - px[ea], c2 = px[ea] * py[eb] - px[eb] * py[ea];
      const double det = a1 * b2 - a2 * b1;
      ok = ok && det != 0.0;
      const double r_ = 1.0 / det;
      const double f1 = -c1, f2 = -c2;
      const double xi0 = (r_ * b2) * f1 + (-(r_ * b1)) * f2;
      const double xi1 = (-(r_ * a2)) * f1 + (r_ * a1) * f2;
      const double s0 = xi0 - xcx, s1 = xi1 - xcy;
      const double mksif = sqrt(s0 * s0 + s1 * s1);
      // src/CellibSimple.f:772-857
      const double alpha = mksif / mksi, om = 1.0 - alpha;
      const double *ek = A.e + 10 * (size_t)checked(A.ix[(size_t)(W + 1) * k + W], A.nmat, ok);
      double kf = alpha / visc + om / ek[0];
      kf = 1.0 / kf;
      const double uk = A.u[k];
      const double wf0 = om * w0 + alpha * A.w[2 * (size_t)k];
      const double wf1 = om * w1 + alpha * A.w[2 * (size_t)k + 1];
      const double g1k0 = A.grad1[2 * (size_t)k], g1k1 = A.grad1[2 * (size_t)k + 1];
      double gf0 = om * g1c0 + alpha * g1k0;
      double gf1 = om * g1c1 + alpha * g1k1;
      const double nk = nx * ksi0 + ny * ksi1;
      const double wfn = wf0 * nx + wf1 * ny;
      double gfk = gf0 * ksi0 + gf1 * ksi1;
      const double corr = (uk - uc) / mksi - gfk;
      gf0 = gf0 + corr * ksi0;
      gf1 = gf1 + corr * ksi1;
      const double gfn = gf0 * nx + gf1 * ny;
      gfk = gf0 * ksi0 + gf1 * ksi1;
      double ai = (kf * meta) / (nk * mksi);
      p = p + (kf * meta) * (gfn - gfk / nk);
      const double cv = (rho * wfn) * meta;
      double cvc;
      if (cv < 0.0) {
        const double du = (uc - uk) + eps;
        const double gpk = g1k0 * (xcx - qx) + g1k1 * (xcy - qy);
        const double r = (2.0 * gpk) / du - 1.0;
        cvc = (om * tvd_limit(r, icod, ok)) * du;
      } else {
        const double du = (uk - uc) + eps;
        const double gpk = (g1c0 * ksi0 + g1c1 * ksi1) * mksi;
        const double r = (2.0 * gpk) / du - 1.0;
        cvc = (alpha * tvd_limit(r, icod, ok)) * du;
      }
      cvc = cvc + gf0 * (xmx - xi0) + gf1 * (xmy - xi1);
      ai = ai + fmax(-cv, 0.0);
      sp = sp + cv;
      p = p - cv * cvc;
      const double prk = A.pres[k];
      double pface;
      if (A.s_pressure) {
        const double p1 = (prc + gpc0 * df0) + gpc1 * df1;
        const double p2 = (prk + A.gradp[2 * (size_t)k] * (xmx - qx)) + A.gradp[2 * (size_t)k + 1] * (xmy - qy);
        pface = 0.5 * (p1 + p2);
      } else {
        pface = om * prc + alpha * prk;
      }
      pf = pf + (pface * meta) * ncp;
      // the non-linear viscous terms
      const double g0 = om * (cp == 0 ? g1c0 : g1c1) + alpha * A.grad1[2 * (size_t)k + cp];
      const double g1 = om * (cp == 0 ? g2c0 : g2c1) + alpha * A.grad2[2 * (size_t)k + cp];
      p = p + (kf * meta) * (cp == 0 ? g0 * nx + g1 * ny : g0 * ny + g1 * nx);
      a[l] = ai;
      un[l] = uk;
      ld[l] = A.num[k];
    } else {
      // walls (src/CellibSimple.f:693-768)
      const double ap = visc * meta;
      p = p + ap * (cp == 0 ? g1c0 * nx + g2c0 * ny : g1c1 * ny + g2c1 * nx);
      const int pe = A.pedge[(size_t)(W + 1) * c + l];
      const double *se = A.sedge + 3 * ((size_t)(W + 1) * c + l);
      const double pface = pe == 3 ? se[2] : A.s_pressure ? prc + (gpc0 * df0 + gpc1 * df1) : prc;
      pf = pf + (pface * meta) * ncp;
      if (pe == 0 || pe == 1) {
        const double apw = (visc * meta) / ca;
        sp = sp + apw;
        if (pe == 1) p = p + apw * se[cp];
      } else if (pe == 2) {
        const double wfn = se[0] * nx + se[1] * ny;
        p = p - ((rho * wfn) * meta) * se[cp];
      } else if (pe == 3 || pe == 4) {
        const double wfn = w0 * nx + w1 * ny;
        const double cv = (rho * wfn) * meta;
        if (pe == 4) {
          sp = sp + cv;
        } else if (fabs(wfn) > 1.0e-7) {
          if (wfn > 0.0) sp = sp + cv;
          else if (wfn < 0.0) p = p + cv * wfn;
        }
      }
    }
  }
  // the pin, a_P, the transient and the under-relaxation (:861-883)
  if (A.pedge[(size_t)(W + 1) * c + W] == -2) sp = sp + 1.0e60;
  double apc = sp;
#pragma unroll
  for (int l = 0; l < W; ++l) apc = apc + a[l];
  double m = ((rho * cc) * area) / A.dt;
  if (A.bs) m = 1.5 * m;
  apc = apc + m;
  apc = apc / A.under_u;
  p = p + ((1.0 - A.under_u) * apc) * uc;
  double rm = 0.0, anb = 0.0;
#pragma unroll
  for (int l = 0; l < W; ++l) {
    rm = rm + a[l] * un[l];
    anb = anb + a[l];
  }
  const double bad = __builtin_nan("");
  A.im[4 * (size_t)c + cp] = ok ? rm + p : bad;
  A.im[4 * (size_t)c + cp + 2] = ok ? apc : bad;
  A.d[2 * (size_t)c + cp] = !ok ? bad : A.simple_c ? area / (apc - anb) : area / apc;
  p = p - pf;
  rm = (p + rm) - apc * uc;
  A.rescell[c] = ok ? rm : bad;
  const int eq = A.num[c] - 1;
  if (eq < 0 || eq >= A.neq) return;
  A.ad[eq] = ok ? apc : bad;
  A.b[eq] = ok ? p : bad;
#pragma unroll
  for (int l = 0; l < W; ++l) A.al[(size_t)W * eq + l] = !ok ? bad : ld[l] > 0 ? -a[l] : 0.0;
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
