// fvc_mixed.h -- mixed-precision Jacobi-PCG: FP32 matrix values and FP32
// inner iterations, FP64 solution and FP64 residual recomputation.
//
// The inner iteration is CG on the Jacobi-scaled system
//     (S A S) y = S r,   S = diag(1/sqrt(a_ii)),   x += S y,
// which is Jacobi-PCG on A in exact arithmetic (src/Iterativos.f:29-106 with
// m = ad).  Scaling first makes the unit-diagonal matrix safe in FP32: the
// reference's 1e60 pinning penalty (src/CellibPressure.f:414,557) overflows
// FP32, the scaled row does not.  The FP32 recursion drifts from the true
// residual, so whenever its unscaled estimate ||S^-1 r~|| falls by a factor
// `delta`, reaches the tolerance, or `max_inner` FP32 iterations have passed,
// the solution increment is folded into the FP64 x and the residual is
// recomputed in FP64 from the FP64 matrix ("reliable updates", van der Vorst
// & Ye 2000; Clark et al. 2010); the search direction is kept.  The solve
// stops when that FP64 residual satisfies ||b - A x|| <= tol ||b||.
#pragma once
#include "fvc_krylov.h"

namespace fvc {

struct Mat32 {
  int neq, w;
  const float *av;   // scaled off-diagonal values, same layout as the FP64 matrix
  const int *ia;     // shared with the FP64 matrix
  const int *ja;
};

// (S A S x)_i with unit diagonal; slots left to right, diagonal last.
template <int F, class X>
__device__ __forceinline__ float row_apply32(const Mat32 &A, int i, const X &xf) {
  if constexpr (F == F_ELL4) {
    const float *a = A.av + 4 * (size_t)i;
    const int *c = A.ja + 4 * (size_t)i;
    const float t1 = a[0] * xf(c[0] - 1), t2 = a[1] * xf(c[1] - 1);
    const float t3 = a[2] * xf(c[2] - 1), t4 = a[3] * xf(c[3] - 1);
    return (((t1 + t2) + t3) + t4) + xf(i);
  } else if constexpr (F == F_ELL3) {
    const float *a = A.av + 3 * (size_t)i;
    const int *c = A.ja + 3 * (size_t)i;
    const float t1 = a[0] * xf(c[0] - 1), t2 = a[1] * xf(c[1] - 1), t3 = a[2] * xf(c[2] - 1);
    return ((t1 + t2) + t3) + xf(i);
  } else if constexpr (F == F_ELLW) {
    const float *a = A.av + (size_t)A.w * (size_t)i;
    const int *c = A.ja + (size_t)A.w * (size_t)i;
    float t = a[0] * xf(c[0] - 1);
    for (int s = 1; s < A.w; ++s) t = t + a[s] * xf(c[s] - 1);
    return t + xf(i);
  } else {   // F_CSRD
    float t = xf(i);
    const int k1 = A.ia[i] - 1, k2 = A.ia[i + 1] - 1;
    for (int k = k1; k < k2; ++k) t = t + A.av[k] * xf(A.ja[k] - 1);
    return t;
  }
}

struct XVec32 {
  const float *x;
  __device__ __forceinline__ float operator()(int j) const { return x[j]; }
};
struct XDir32 {
  const float *r, *p;
  float beta;
  __device__ __forceinline__ float operator()(int j) const { return r[j] + beta * p[j]; }
};

// Builds S A S in FP32: av32 = a_ij * s_i * s_j, s_i = 1/sqrt(a_ii); also
// s (FP64) and dsq = sqrt(a_ii) (FP32, to unscale residual estimates).
template <int F>
__global__ void scale_to_f32_kernel(Mat A, const double *m, float *av32, double *s, float *dsq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.neq) return;
  const double si = 1.0 / sqrt(m[i]);
  s[i] = si;
  dsq[i] = (float)sqrt(m[i]);
  if constexpr (F == F_ELL4 || F == F_ELL3 || F == F_ELLW) {
    const int w = (F == F_ELL4) ? 4 : (F == F_ELL3) ? 3 : A.w;
    for (int k = 0; k < w; ++k) {
      const size_t e = (size_t)w * i + k;
      const int j = A.ja[e] - 1;
      av32[e] = (float)(A.av[e] * si * (1.0 / sqrt(m[j])));
    }
  } else {
    for (int k = A.ia[i] - 1; k < A.ia[i + 1] - 1; ++k) {
      const int j = A.ja[k] - 1;
      av32[k] = (float)(A.av[k] * si * (1.0 / sqrt(m[j])));
    }
  }
}

enum { MX_SETUP = 0, MX_A = 1, MX_B = 2, MX_X = 3, MX_R = 4, MX_FINAL = 5 };

__host__ __device__ constexpr int mixed_nred(int ph) {
  return ph == MX_A ? 1 : ph == MX_X ? 0 : (ph == MX_B || ph == MX_FINAL) ? 2 : 3;
}

struct MixedArgs {
  Mat A;         // FP64 matrix (residual recomputation, final energy)
  Mat32 A32;     // FP32 scaled matrix (inner iterations)
  Grid g;
  KState *st;
  const double *m;
  const double *s;    // 1/sqrt(a_ii)
  const float *dsq;   // sqrt(a_ii)
  double *x;
  const double *b;
  float *y, *r, *p0, *p1, *q;
  double tol, delta;
  int maxit, x0_zero, max_inner, phase;
  Defer d;
};

// KState use: d = r~.r~ (FP64 accumulation), d_prev = its value before the
// current iteration, conv = tol*||b||, last = the latest residual estimate,
// upd_est = the estimate at the last reliable update, cycle = number of
// reliable updates, inner = FP32 iterations since it, step_i = update due,
// next = the phase that follows (the order depends on the iterates: after
// B either the next iteration's A or an update's X and R).
// Row and scalar parts are split as for the one-reduction PCG, so that the
// persistent kernel, the batched kernel and the kernel-per-phase engine with
// deferred reductions all drive the same code.
template <int F>
__device__ void mixed_rows(const MixedArgs &a, int ph, const KState *S, double *v) {
  const int i0 = a.g.row_begin(), i1 = a.g.row_end();
  if (ph == MX_SETUP || ph == MX_R) {
    // r64 = b - A x (FP64); r~ = S r64 rounded to FP32; y = 0
    const bool zero = (ph == MX_SETUP) && a.x0_zero;
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
      double ax;
      if (zero) {
        a.x[i] = 0.0;
        ax = 0.0;
      } else {
        ax = row_apply<F>(a.A, i, XVec{a.x});
      }
      const double bi = a.b[i];
      const double ri = bi - ax;
      const float rt = (float)(ri * a.s[i]);
      a.r[i] = rt;
      a.y[i] = 0.0f;
      v[0] += (double)rt * (double)rt;
      v[1] += ri * ri;
      v[2] += bi * bi;
    }
  } else if (ph == MX_A) {
    // p = r~ + beta p (on the fly, ping-pong); q = (S A S) p; -> p.q
    const int j = ld_int(&S->iter) + 1;
    float *pn = (j & 1) ? a.p1 : a.p0;
    const float *po = (j & 1) ? a.p0 : a.p1;
    const bool restart = j == 1;
    const XDir32 xf{a.r, po, restart ? 0.0f : (float)ld_dbl(&S->beta)};
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
      float qi, pi;
      if (restart) {
        qi = row_apply32<F>(a.A32, i, XVec32{a.r});
        pi = a.r[i];
      } else {
        qi = row_apply32<F>(a.A32, i, xf);
        pi = xf(i);
      }
      pn[i] = pi;
      a.q[i] = qi;
      v[0] += (double)pi * (double)qi;
    }
  } else if (ph == MX_B) {
    // y += alpha p; r~ -= alpha q; -> r~.r~ and the unscaled estimate ||S^-1 r~||^2
    const int j = ld_int(&S->iter) + 1;
    const float *pn = (j & 1) ? a.p1 : a.p0;
    const float alpha = (float)ld_dbl(&S->alpha);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
      const float yi = a.y[i] + alpha * pn[i];
      const float ri = a.r[i] - alpha * a.q[i];
      a.y[i] = yi;
      a.r[i] = ri;
      v[0] += (double)ri * (double)ri;
      const double ru = (double)ri * (double)a.dsq[i];
      v[1] += ru * ru;
    }
  } else if (ph == MX_X) {
    // fold the FP32 increment into the FP64 solution: x += S y
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) a.x[i] = a.x[i] + a.s[i] * (double)a.y[i];
  } else {
    // MX_FINAL: energy x.Ax and ||b - Ax||^2 in FP64
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
      const double ax = row_apply<F>(a.A, i, XVec{a.x});
      const double ri = a.b[i] - ax;
      v[0] += a.x[i] * ax;
      v[1] += ri * ri;
    }
  }
}

__device__ inline void mixed_epi(const MixedArgs &a, int ph, KState *st, const double *t) {
  if (ph == MX_SETUP || ph == MX_R) {
    if (ph == MX_SETUP) {
      st->bnorm2 = t[2];
      st->conv = a.tol * sqrt(t[2]);
      st->iter = 0;
      st->cycle = 0;
      st->status = S_MAXIT;
      st->done = 0;
    } else {
      st->cycle = st->cycle + 1;
    }
    st->res2 = t[1];
    st->last = sqrt(t[1]);
    st->upd_est = st->last;
    // the recomputed residual replaces the recursive one; the direction
    // keeps its history: beta = r~new.r~new / r~old.r~old
    if (ph == MX_R) st->beta = t[0] / st->d_prev;
    st->d = t[0];
    st->inner = 0;
    st->step_i = 0;
    st->next = MX_A;
    if (st->last <= st->conv) { st->status = S_OK; st->done = 1; }
    else if (!finite(t[0]) || !finite(t[1])) { st->status = S_NONFINITE; st->done = 1; }
    else if (st->iter >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
  } else if (ph == MX_A) {
    st->alpha = st->d / t[0];
    st->next = MX_B;
    if (t[0] == 0.0) { st->status = S_BREAKDOWN; st->done = 1; }
    else if (!finite(st->alpha)) { st->status = S_NONFINITE; st->done = 1; }
  } else if (ph == MX_B) {
    const int j = st->iter + 1;
    st->iter = j;
    st->inner = st->inner + 1;
    const double est = sqrt(t[1]);
    st->last = est;
    st->d_prev = st->d;
    st->beta = t[0] / st->d;
    st->d = t[0];
    const bool upd = est <= st->conv || est < a.delta * st->upd_est || st->inner >= a.max_inner ||
                     j >= a.maxit || !finite(t[0]) || t[0] == 0.0;
    st->step_i = upd ? 1 : 0;
    st->next = upd ? MX_X : MX_A;
  } else if (ph == MX_X) {
    st->next = MX_R;
  } else {
    st->energy = t[0];
    st->res2 = t[1];
  }
}

template <int F>
__device__ bool mixed_phase(const MixedArgs &a, int ph, double *lds) {
  double v[kMaxK] = {0.0, 0.0, 0.0, 0.0};
  mixed_rows<F>(a, ph, a.st, v);
  auto epi = [&](const double *t) { mixed_epi(a, ph, a.st, t); };
  switch (mixed_nred(ph)) {
    case 0: return grid_sync<0>(a.g, v, lds, epi);
    case 1: return grid_sync<1>(a.g, v, lds, epi);
    case 2: return grid_sync<2>(a.g, v, lds, epi);
    default: return grid_sync<3>(a.g, v, lds, epi);
  }
}

// Persistent driver (one system, or one system's block group of a batch).
template <int F>
__device__ void mixed_run(const MixedArgs &a, double *lds) {
  if (!mixed_phase<F>(a, MX_SETUP, lds)) return;
  while (!ld_int(&a.st->done))
    if (!mixed_phase<F>(a, ld_int(&a.st->next), lds)) return;
  mixed_phase<F>(a, MX_FINAL, lds);
}

template <int F>
__global__ __launch_bounds__(1024) void mixed_persistent(MixedArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  mixed_run<F>(a, lds);
}

// Kernel-per-phase engine: a.phase is MX_SETUP, kPhaseNext, MX_FINAL or a
// flush; reductions deferred to the next launch (fvc_krylov.h, Defer).
template <int F>
__global__ __launch_bounds__(1024) void mixed_kernel(MixedArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ KState S;
  struct M {
    const MixedArgs &a;
    __device__ int nred(int ph) const { return mixed_nred(ph); }
    __device__ bool skip(int ph, const KState &s) const { return ph != MX_SETUP && ph != MX_FINAL && s.done; }
    __device__ void rows(int ph, const KState *s, double *v) const { mixed_rows<F>(a, ph, s, v); }
    __device__ void epi(int ph, KState *s, const double *t) const { mixed_epi(a, ph, s, t); }
  };
  deferred_phase(a.g, a.d, a.phase, M{a}, &S, lds);
}

}  // namespace fvc
