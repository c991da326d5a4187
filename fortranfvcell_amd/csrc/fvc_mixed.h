// fvc_mixed.h -- mixed-precision Jacobi-PCG: FP32 matrix values and FP32
// inner iterations, FP64 solution and FP64 residual recomputation.
//
// The inner iteration is CG on the Jacobi-scaled system
//     (S A S) y = S r,   S = diag(1/sqrt(a_ii)),   x += S y,
// which is Jacobi-PCG on A in exact arithmetic (src/Iterativos.f:29-106 with
// m = ad).  Scaling first makes the unit-diagonal matrix safe in FP32: the
// reference's 1e60 pinning penalty (src/CellibPressure.f:414,557) overflows
// FP32, the scaled row does not.  The FP32 recursion drifts from the true
// residual, so from time to time (see MX_A below) the solution increment is
// folded into the FP64 x and the residual is recomputed in FP64 from the FP64
// matrix ("reliable updates", van der Vorst & Ye 2000; Clark et al. 2010);
// the search direction is kept.  The solve stops when that FP64 residual
// satisfies ||b - A x|| <= tol ||b||.
#pragma once
#include "fvc_krylov.h"

namespace fvc {

struct Mat32 {
  int neq, w;
  const float *av;   // scaled off-diagonal values, same layout as the FP64 matrix
  const int *ia;     // shared with the FP64 matrix
  const int *ja;
  const short *jo;   // ELLPACK widths 3 and 4: column offsets j - i (0-based), or
                     // null (then ja); col_offsets_kernel
};

template <int F, class X>
__device__ __forceinline__ float row_apply32(const Mat32 &A, int i, const X &xf);

// One FP32 matrix row in registers, loaded ahead of its use for ELLPACK
// widths 3 and 4 (as RowRegs for FP64); apply() is the only definition of
// the row's arithmetic.
template <int F>
struct RowRegs32 {
  __device__ __forceinline__ void load(const Mat32 &, int) {}
  template <class X>
  __device__ __forceinline__ float apply(const Mat32 &A, int i, const X &xf) const {
    return row_apply32<F>(A, i, xf);
  }
};
// c0.. hold 0-based columns.
template <>
struct RowRegs32<F_ELL4> {
  float a0, a1, a2, a3;
  int c0, c1, c2, c3;
  __device__ __forceinline__ void load(const Mat32 &A, int i) {
    const float *a = A.av + 4 * (size_t)i;
    a0 = a[0]; a1 = a[1]; a2 = a[2]; a3 = a[3];
    if (A.jo) {
      const short4 o = *reinterpret_cast<const short4 *>(A.jo + 4 * (size_t)i);
      c0 = i + o.x; c1 = i + o.y; c2 = i + o.z; c3 = i + o.w;
    } else {
      const int *c = A.ja + 4 * (size_t)i;
      c0 = c[0] - 1; c1 = c[1] - 1; c2 = c[2] - 1; c3 = c[3] - 1;
    }
  }
  template <class X>
  __device__ __forceinline__ float apply(const Mat32 &, int i, const X &xf) const {
    const float t1 = a0 * xf(c0), t2 = a1 * xf(c1);
    const float t3 = a2 * xf(c2), t4 = a3 * xf(c3);
    return (((t1 + t2) + t3) + t4) + xf(i);
  }
};
template <>
struct RowRegs32<F_ELL3> {
  float a0, a1, a2;
  int c0, c1, c2;
  __device__ __forceinline__ void load(const Mat32 &A, int i) {
    const float *a = A.av + 3 * (size_t)i;
    a0 = a[0]; a1 = a[1]; a2 = a[2];
    if (A.jo) {
      const short *o = A.jo + 3 * (size_t)i;
      c0 = i + o[0]; c1 = i + o[1]; c2 = i + o[2];
    } else {
      const int *c = A.ja + 3 * (size_t)i;
      c0 = c[0] - 1; c1 = c[1] - 1; c2 = c[2] - 1;
    }
  }
  template <class X>
  __device__ __forceinline__ float apply(const Mat32 &, int i, const X &xf) const {
    const float t1 = a0 * xf(c0), t2 = a1 * xf(c1), t3 = a2 * xf(c2);
    return ((t1 + t2) + t3) + xf(i);
  }
};

// (S A S x)_i with unit diagonal; slots left to right, diagonal last.
template <int F, class X>
__device__ __forceinline__ float row_apply32(const Mat32 &A, int i, const X &xf) {
  if constexpr (F == F_ELL4 || F == F_ELL3) {
    RowRegs32<F> r;
    r.load(A, i);
    return r.apply(A, i, xf);
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

// Column offsets j - i of a width-W ELLPACK matrix as int16, read by the
// FP32 inner iterations instead of the int32 columns (8 instead of 16 bytes
// per row at width 4).  A slot whose scaled FP32 value is zero (ELLPACK
// padding: column 1, value 0) gets offset 0, the row itself, and still adds
// a zero.  A nonzero slot whose offset does not fit sets *wide, and the
// caller then keeps the int32 columns.
template <int W>
__global__ void col_offsets_kernel(int neq, const int *ja, const float *av32, short *jo, int *wide) {
  const int i = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (i >= neq) return;
  bool fits = true;
  for (int k = 0; k < W; ++k) {
    const size_t e = (size_t)W * i + k;
    const int d = (ja[e] - 1) - i;
    short o = 0;
    if (av32[e] != 0.0f) {
      if (d >= -32768 && d <= 32767) o = (short)d;
      else fits = false;
    }
    jo[e] = o;
  }
  if (!fits) *wide = 1;
}

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

enum { MX_SETUP = 0, MX_A = 1, MX_X = 3, MX_R = 4, MX_FINAL = 5 };

__host__ __device__ constexpr int mixed_nred(int ph) {
  return ph == MX_A ? 5 : ph == MX_X ? 0 : ph == MX_FINAL ? 2 : 3;
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
  // Records {r~, q, p, y} per row, two buffers alternating by iteration
  // parity, the second at +ps floats.  MX_A reads the current record (a
  // neighbour's r~, q, p are one 16-byte load) and writes the whole next one.
  float4 *rec;
  size_t ps;
  double tol, delta;
  int maxit, x0_zero, max_inner, phase;
  Defer d;
};

// The inner iteration is one phase, MX_A, with one reduction: CG in the form
// of D'Azevedo, Eijkhout & Romine ("Conjugate gradient algorithms with
// reduced synchronization overhead", 1993).  Iteration j, for its own rows:
//   completes iteration j-1:  y += alpha p,  r~ -= alpha q   (unless the
//                             residual is fresh from a refresh);
//   p = r~ + beta p,  q = (S A S) p  (p of the neighbours formed on the fly
//                             from r~, q and p of iteration j-1);
//   reduces p.q, r~.q, q.q, r~.r~ and ||S^-1 r~||^2.
// The recurrence then takes alpha = rho / p.q with rho = r~.r~ exact, and
// rho' = rho - 2 alpha r~.q + alpha^2 q.q for the next beta: a one-step
// recurrence anchored every iteration on the exact rho of this pass, so it
// does not drift (q is always a true product, never a recurrence).
// r~, p and q alternate between two buffers because the next iteration's
// gathers read the old values of other blocks' rows.
//
// A reliable update (MX_X: x += S (y + alpha p); MX_R: r~ = S (b - A x) in
// FP64) is due when
//   * the scaled residual ||r~'|| (from rho') has fallen by `delta` since the
//     last update,
//   * the unscaled estimate ||S^-1 r~|| of this pass reaches tol ||b|| times
//     `tscale` (not in the first pass after an update, whose residual was
//     just measured in FP64); tscale starts at 1 and, when an update so
//     triggered finds the true residual still above the tolerance, shrinks
//     by the ratio estimate / true residual it observed,
//   * `max_inner` FP32 iterations have passed since the last update.
// After an update the direction keeps its history: beta = r~new.r~new / rho.
//
// KState use: d = rho of the last pass, rr0 = r~.r~ at the last update,
// alpha, beta = the pending step and the next direction's coefficient,
// conv = tol*||b||, last = the latest residual estimate, tscale, cycle =
// number of updates, inner = FP32 iterations since the last one, step_i = 1
// while r~ is fresh from an update (nothing pending), ni = 1 to restart the
// direction (p = r~), inner_done = 1 when the tolerance test triggered the
// update, next = the phase that follows.
// Row and scalar parts are split as for the one-reduction PCG, so that the
// persistent kernel, the batched kernel and the kernel-per-phase engine with
// deferred reductions all drive the same code.

// p of iteration j at row k: (r~ - alpha q) + beta p, branch-free so that a
// row's gathers issue together.  alpha = 0 while r~ is fresh from an update
// (nothing pending) and beta = 0 when the direction restarts; the setup
// zero-fills the first record's q and p, so 0 * p and 0 * q are exact zeros.
struct XMixDir {
  const float4 *rc;
  float alpha, beta;
  __device__ __forceinline__ static float rnew(const float4 &c, float alpha) { return c.x - alpha * c.y; }
  __device__ __forceinline__ float operator()(int k) const {
    const float4 c = rc[k];
    return rnew(c, alpha) + beta * c.z;
  }
};

// The operands of an MX_A row that do not alternate between buffers: the
// FP32 matrix row and sqrt(a_ii).
template <int F>
struct MixRow {
  RowRegs32<F> mr;
  float dsq;
  bool ok;
};
template <int F>
__device__ __forceinline__ MixRow<F> mixed_load(const MixedArgs &a, int i) {
  MixRow<F> q;
  q.mr.load(a.A32, i);
  q.dsq = a.dsq[i];
  q.ok = true;
  return q;
}

template <int F>
__device__ __forceinline__ void mixed_rows(const MixedArgs &a, int ph, const KState *S, double *v,
                                           const MixRow<F> *pre = nullptr) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  if (ph == MX_SETUP || ph == MX_R) {
    // r64 = b - A x (FP64); r~ = S r64 rounded to FP32 into the buffer the
    // next MX_A reads; y = 0
    const bool setup = ph == MX_SETUP, zero = setup && a.x0_zero;
    const int j = setup ? 1 : ld_int(&S->iter) + 1;
    float4 *rc = a.rec + ((j & 1) ? 0 : a.ps);
    for (int i = i0; i < i1; i += is) {
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
      if (setup) {
        rc[i] = make_float4(rt, 0.0f, 0.0f, 0.0f);
      } else {
        // the refreshed residual; q and p stay (the direction keeps its history)
        float *f = reinterpret_cast<float *>(rc + i);
        f[0] = rt;
        f[3] = 0.0f;
      }
      v[0] += (double)rt * (double)rt;
      v[1] += ri * ri;
      v[2] += bi * bi;
    }
  } else if (ph == MX_A) {
    const int j = ld_int(&S->iter) + 1;
    const size_t oc = (j & 1) ? 0 : a.ps, on = a.ps - oc;
    const float4 *rc = a.rec + oc;
    float4 *rn = a.rec + on;
    // a fresh residual has nothing pending (alpha = 0), a restart no history
    // (beta = 0)
    const float alpha = ld_int(&S->step_i) ? 0.0f : (float)ld_dbl(&S->alpha);
    const float beta = ld_int(&S->ni) ? 0.0f : (float)ld_dbl(&S->beta);
    const XMixDir xf{rc, alpha, beta};
    auto row = [&](int i, const MixRow<F> &w) {
      const float4 c = rc[i];
      const float yi = c.w + xf.alpha * c.z;
      const float ri = XMixDir::rnew(c, xf.alpha);
      const float pi = ri + xf.beta * c.z;
      const float qi = w.mr.apply(a.A32, i, xf);
      rn[i] = make_float4(ri, qi, pi, yi);
      const double rd = ri, qd = qi;
      v[0] += (double)pi * qd;
      v[1] += rd * qd;
      v[2] += qd * qd;
      v[3] += rd * rd;
      const double ru = rd * (double)w.dsq;
      v[4] += ru * ru;
    };
    int i = i0;
    if (pre && pre->ok && i < i1) {
      row(i, *pre);
      i += is;
    }
    for (; i < i1; i += is) row(i, mixed_load<F>(a, i));
  } else if (ph == MX_X) {
    // complete the last step and fold the FP32 increment into the FP64
    // solution: x += S (y + alpha p)
    const int j = ld_int(&S->iter);
    const float4 *rj = a.rec + ((j & 1) ? a.ps : 0);
    const float alpha = (float)ld_dbl(&S->alpha);
    for (int i = i0; i < i1; i += is) {
      const float4 c = rj[i];
      const float yi = c.w + alpha * c.z;
      a.x[i] = a.x[i] + a.s[i] * (double)yi;
    }
  } else {
    // MX_FINAL: energy x.Ax and ||b - Ax||^2 in FP64
    for (int i = i0; i < i1; i += is) {
      const double ax = row_apply<F>(a.A, i, XVec{a.x});
      const double ri = a.b[i] - ax;
      v[0] += a.x[i] * ax;
      v[1] += ri * ri;
    }
  }
}

__device__ __forceinline__ void mixed_epi(const MixedArgs &a, int ph, KState *st, const double *t) {
  if (ph == MX_SETUP || ph == MX_R) {
    const double tr = sqrt(t[1]);
    if (ph == MX_SETUP) {
      st->bnorm2 = t[2];
      st->conv = a.tol * sqrt(t[2]);
      st->iter = 0;
      st->cycle = 0;
      st->status = S_MAXIT;
      st->done = 0;
      st->tscale = 1.0;
      st->ni = 1;
    } else {
      st->cycle = st->cycle + 1;
      if (st->inner_done && tr > st->conv && st->last < tr) st->tscale = st->tscale * (st->last / tr);
      // the recomputed residual replaces the recursive one; the direction
      // keeps its history
      const double beta = t[0] / st->d;
      st->ni = (st->d > 0.0 && finite(beta)) ? 0 : 1;
      st->beta = st->ni ? 0.0 : beta;
    }
    st->res2 = t[1];
    st->last = tr;
    st->rr0 = t[0];
    st->inner = 0;
    st->inner_done = 0;
    st->step_i = 1;
    st->next = MX_A;
    if (tr <= st->conv) { st->status = S_OK; st->done = 1; }
    else if (!finite(t[0]) || !finite(t[1])) { st->status = S_NONFINITE; st->done = 1; }
    else if (st->iter >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
  } else if (ph == MX_A) {
    const double sig = t[0], tau = t[1], th = t[2], rho = t[3];
    const int j = st->iter + 1;
    st->iter = j;
    st->inner = st->inner + 1;
    const double alpha = rho / sig;
    const double rho1 = rho - 2.0 * alpha * tau + alpha * alpha * th;
    const double beta = rho1 / rho;
    const double est = sqrt(t[4]);
    const bool ok = sig != 0.0 && rho != 0.0 && finite(alpha) && finite(beta);
    st->last = est;
    st->d = rho;
    st->alpha = ok ? alpha : 0.0;
    st->beta = ok ? beta : 0.0;
    st->step_i = 0;
    st->ni = 0;
    const bool tolhit = st->inner > 1 && est <= st->conv * st->tscale;
    const bool drop = sqrt(fmax(rho1, 0.0)) < a.delta * sqrt(st->rr0);
    st->inner_done = tolhit ? 1 : 0;
    const bool upd = !ok || tolhit || drop || st->inner >= a.max_inner || j >= a.maxit;
    st->next = upd ? MX_X : MX_A;
  } else if (ph == MX_X) {
    st->next = MX_R;
  } else {
    st->energy = t[0];
    st->res2 = t[1];
  }
}

template <int F>
__device__ __forceinline__ bool mixed_phase(const MixedArgs &a, int ph, double *lds) {
  double v[kMaxK] = {};
  mixed_rows<F>(a, ph, a.st, v);
  auto epi = [&](const double *t) { mixed_epi(a, ph, a.st, t); };
  switch (mixed_nred(ph)) {
    case 0: return grid_sync<0>(a.g, v, lds, epi);
    case 2: return grid_sync<2>(a.g, v, lds, epi);
    case 3: return grid_sync<3>(a.g, v, lds, epi);
    default: return grid_sync<5>(a.g, v, lds, epi);
  }
}

// Persistent driver (one system, or one system's block group of a batch).
template <int F>
__device__ bool mixed_run(const MixedArgs &a, double *lds) {
  if (!mixed_phase<F>(a, MX_SETUP, lds)) return false;
  while (!ld_int(&a.st->done))
    if (!mixed_phase<F>(a, ld_int(&a.st->next), lds)) return false;
  return mixed_phase<F>(a, MX_FINAL, lds);
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
    enum { kRed = 5 };
    using Pre = MixRow<F>;
    __device__ int nred(int ph) const { return mixed_nred(ph); }
    __device__ bool skip(int ph, const KState &s) const { return ph != MX_SETUP && ph != MX_FINAL && s.done; }
    // the phase a kPhaseNext launch runs is known only after the pending
    // recurrence: after X it is R, otherwise nearly always A (X follows an
    // A only at a reliable update), so A's operands are loaded ahead
    __device__ Pre prefetch(int ph, const KState &stored) const {
      Pre q{};
      const int i = a.g.first();
      const bool a_next = ph == MX_A || (ph == kPhaseNext && stored.pend - 1 != MX_X);
      if (a_next && i < a.g.row_end()) q = mixed_load<F>(a, i);
      return q;
    }
    __device__ void rows(int ph, const KState *s, double *v, const Pre &q) const { mixed_rows<F>(a, ph, s, v, &q); }
    __device__ void epi(int ph, KState *s, const double *t) const { mixed_epi(a, ph, s, t); }
  };
  deferred_phase(a.g, a.d, a.phase, M{a}, &S, lds);
}

}  // namespace fvc
