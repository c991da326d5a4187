// fvc_krylov.h -- the Krylov solvers of src/Iterativos.f / src/Iterativos_omp.f
// as sequences of row-parallel phases separated by grid-wide reductions.
//
// Each method is one phase function, pcg_phase / bicg_phase / gmres_phase,
// that runs one phase over this block's rows (Grid::first) and ends with a
// grid_reduce whose epilogue (run once, by the last block) advances the
// scalar recurrences held in KState.  Two engines drive the same functions:
//   persistent : one cooperative launch loops over the phases with grid
//                barriers (two barriers per PCG iteration);
//   kernels    : one launch per phase; the next launch sees the epilogue.
// Both engines use the same row order and reduction trees, so they produce
// bit-identical results.
#pragma once
#include "fvc_device.h"

namespace fvc {

// Status codes (mirrors include/fvcell.h).
constexpr int S_OK = 0, S_MAXIT = 1, S_BREAKDOWN = 2, S_NONFINITE = 3;

struct alignas(128) KState {
  double d, conv, alpha, beta, rr0, omega, energy, last;
  double bnorm2, res2, econv, rnorm2, upd_est, d_prev, tscale, pad3;   // tscale: mixed refreshes
  int iter, status, done, inner;      // inner: GMRES Arnoldi steps (nit)
  int step_i, mgs_j, inner_done, ni;  // GMRES position in the cycle
  int cycle, pend, next, pad6;         // pend: deferred reduction pending (Defer);
                                       // next: the phase a state-driven method runs next
};

__device__ __forceinline__ int ld_int(const int *p) {
  return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_dbl(const double *p) {
  unsigned long long u = __hip_atomic_load(reinterpret_cast<unsigned long long *>(const_cast<double *>(p)),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_bit_cast(double, u);
}
__device__ __forceinline__ bool finite(double v) { return __builtin_isfinite(v); }

// --------------------------------------------------- deferred reductions --
//
// The kernel-per-phase engine of the fused methods.  A launch never waits on
// another block: each block writes the block sums of its phase's values; the
// next launch's blocks each sum all G partials (in the same order as the
// grid barrier's epilogue, so both engines give bit-identical results) and
// apply the phase's scalar recurrence to their own copy of the state, in LDS.
// Kernel boundaries order the memory, so no fence and no atomic is needed.
// State and partials are double-buffered: a launch reads set `par` and
// writes set 1 - par; block 0 writes the updated state.  KState::pend is
// 1 + the phase whose partials wait for their recurrence (0: none).
struct Defer {
  KState *st;     // two states
  double *part;   // two arrays of kMaxK * G partials
  int par;
};

// M: nred(ph) values reduced by phase ph; skip(ph, state) whether the phase's
// rows are skipped (solve finished); rows(ph, state, v, pre); epi(ph, state,
// t); prefetch(ph, stored state) -> M::Pre, the operands of the thread's
// first row that do not depend on the state (an empty struct for methods
// that load nothing ahead).  ph = kPhaseNext: the phase the state names (KState::next, set by
// the recurrences of methods whose phase order depends on the iterates);
// other ph < 0: no rows, only the pending recurrence (a flush).
//
// The first row's operands are issued before the pending recurrence, so that
// HBM streams while the block sums the partials; the block's barriers order
// LDS only (no thread reads global data another thread wrote in the launch),
// so those loads and the rows' stores stay in flight across them.  A state
// that already says the solve is finished loads nothing ahead.
constexpr int kPhaseNext = -2;

struct NoPre {};

template <class M>
__device__ void deferred_phase(const Grid &g, const Defer &d, int ph, const M &m, KState *S,
                               double *lds) {
  constexpr int nw = (int)(sizeof(KState) / sizeof(unsigned));
  const KState *stored = d.st + d.par;
  const unsigned *in = reinterpret_cast<const unsigned *>(stored);
  const double *pin = d.part + (size_t)d.par * kMaxK * g.G;
  for (int k = threadIdx.x; k < nw; k += blockDim.x) reinterpret_cast<unsigned *>(S)[k] = in[k];
  // the partials of the pending phase, loaded before the state says which
  // phase that is (the unused values are never read)
  PartialRegs<M::kRed> pr;
  pr.load(pin, g.G);
  typename M::Pre pre{};
  if (ph != -1 && !stored->done) pre = m.prefetch(ph, *stored);
  lds_barrier();
  if (S->pend) {
    const int pp = S->pend - 1;
    double t[M::kRed] = {};
    switch (m.nred(pp)) {
      case 0: break;
      case 1: pr.template sum<1, true>(pin, g.G, t, lds); break;
      case 2: pr.template sum<2, true>(pin, g.G, t, lds); break;
      case 3: if constexpr (M::kRed >= 3) pr.template sum<3, true>(pin, g.G, t, lds); break;
      case 4: if constexpr (M::kRed >= 4) pr.template sum<4, true>(pin, g.G, t, lds); break;
      default: if constexpr (M::kRed >= 5) pr.template sum<5, true>(pin, g.G, t, lds); break;
    }
    if (threadIdx.x == 0) {
      m.epi(pp, S, t);
      S->pend = 0;
    }
    lds_barrier();
  }
  if (ph == kPhaseNext) ph = S->done ? -1 : S->next;
  if (ph >= 0 && !m.skip(ph, *S)) {
    double v[M::kRed] = {};
    m.rows(ph, S, v, pre);
    const int K = m.nred(ph);
    switch (K) {
      case 0: break;
      case 1: block_sum<1, true>(v, lds); break;
      case 2: block_sum<2, true>(v, lds); break;
      case 3: if constexpr (M::kRed >= 3) block_sum<3, true>(v, lds); break;
      case 4: if constexpr (M::kRed >= 4) block_sum<4, true>(v, lds); break;
      default: if constexpr (M::kRed >= 5) block_sum<5, true>(v, lds); break;
    }
    if (threadIdx.x == 0) {
      double *out = d.part + (size_t)(1 - d.par) * kMaxK * g.G;
      const int c = g.chunk();
      for (int k = 0; k < K; ++k) out[(size_t)k * g.G + c] = v[k];
      S->pend = ph + 1;
    }
  }
  if (g.bid() == 0) {
    lds_barrier();
    unsigned *out = reinterpret_cast<unsigned *>(d.st + 1 - d.par);
    for (int k = threadIdx.x; k < nw; k += blockDim.x) out[k] = reinterpret_cast<const unsigned *>(S)[k];
  }
}

// Gather functors: the vector an SpMV multiplies, element j (0-based).
struct XVec {
  const double *x;
  __device__ __forceinline__ double operator()(int j) const { return x[j]; }
};
// PCG direction formed on the fly: p_j = z_j + beta * p_old_j
// (src/Iterativos.f:69-71 evaluated where the SpMV needs it).
struct XPcgDir {
  const double *z, *p;
  double beta;
  __device__ __forceinline__ double operator()(int j) const { return z[j] + beta * p[j]; }
};
// GMRES basis vector normalised on the fly: g_j / nrm (src/Iterativos.f:359-361).
struct XScaled {
  const double *g;
  double nrm;
  __device__ __forceinline__ double operator()(int j) const { return g[j] / nrm; }
};

// ------------------------------------------------------------------ PCG --

enum { PCG_SETUP = 0, PCG_A = 1, PCG_B = 2, PCG_FINAL = 3 };

struct PcgArgs {
  Mat A;
  Grid g;
  KState *st;
  const double *m;   // Jacobi diagonal (m = ad, src/Solver.f:84)
  double *x;
  const double *b;
  double *r, *z, *p0, *p1, *q;
  double tol;
  int maxit, x0_zero, stop_rule, phase;
};

// src/Iterativos.f:29-106 / src/Iterativos_omp.f:30-127.
template <int F>
__device__ bool pcg_phase(const PcgArgs &a, int ph, double *lds) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  KState *st = a.st;
  if (ph == PCG_SETUP) {
    // z = A x; r = b - z; z = r/m; p = z; d = r.z; conv = tol*sqrt|d|
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0; i < i1; i += is) {
      double ax;
      if (a.x0_zero) {
        a.x[i] = 0.0;
        ax = 0.0;
      } else {
        ax = row_apply<F>(a.A, i, XVec{a.x});
      }
      const double bi = a.b[i];
      const double ri = bi - ax;
      const double zi = ri / a.m[i];
      a.r[i] = ri;
      a.z[i] = zi;
      v[0] += ri * zi;
      v[1] += bi * bi;
      v[2] += ri * ri;
    }
    return grid_sync<3>(a.g, v, lds, [&](const double *t) {
      st->d = t[0];
      st->bnorm2 = t[1];
      st->rnorm2 = t[2];
      st->iter = 0;
      st->inner = 0;
      st->status = S_MAXIT;
      st->done = 0;
      st->beta = 0.0;
      if (a.stop_rule == 1) {
        st->conv = a.tol * sqrt(t[1]);
        st->last = sqrt(t[2]);
        if (st->last <= st->conv) { st->status = S_OK; st->done = 1; }
      } else {
        st->conv = a.tol * sqrt(fabs(t[0]));
        st->last = sqrt(fabs(t[0]));
      }
      if (!finite(t[0])) { st->status = S_NONFINITE; st->done = 1; }
      if (a.maxit <= 0 && !st->done) st->done = 1;
    });
  }
  if (ph == PCG_A) {
    // q = A p with p = z + beta*p formed on the fly; alpha = d / p.q
    const int j = ld_int(&st->iter) + 1;
    double *pn = (j & 1) ? a.p1 : a.p0;
    const double *po = (j & 1) ? a.p0 : a.p1;
    double v[1] = {0.0};
    if (j == 1) {
      for (int i = i0; i < i1; i += is) {
        const double qi = row_apply<F>(a.A, i, XVec{a.z});
        const double pi = a.z[i];
        pn[i] = pi;
        a.q[i] = qi;
        v[0] += pi * qi;
      }
    } else {
      const XPcgDir xf{a.z, po, ld_dbl(&st->beta)};
      for (int i = i0; i < i1; i += is) {
        const double qi = row_apply<F>(a.A, i, xf);
        const double pi = xf(i);
        pn[i] = pi;
        a.q[i] = qi;
        v[0] += pi * qi;
      }
    }
    return grid_sync<1>(a.g, v, lds, [&](const double *t) {
      const double alpha = st->d / t[0];
      st->alpha = alpha;
      if (t[0] == 0.0) { st->status = S_BREAKDOWN; st->done = 1; st->iter = j; }
      else if (!finite(alpha)) { st->status = S_NONFINITE; st->done = 1; st->iter = j; }
    });
  }
  if (ph == PCG_B) {
    // x += alpha p; r -= alpha q; z = r/m; beta = r.z/d; d = beta*d
    const int j = ld_int(&st->iter) + 1;
    const double *pn = (j & 1) ? a.p1 : a.p0;
    const double alpha = ld_dbl(&st->alpha);
    double v[2] = {0.0, 0.0};
    for (int i = i0; i < i1; i += is) {
      const double xi = a.x[i] + alpha * pn[i];
      const double ri = a.r[i] - alpha * a.q[i];
      const double zi = ri / a.m[i];
      a.x[i] = xi;
      a.r[i] = ri;
      a.z[i] = zi;
      v[0] += ri * zi;
      v[1] += ri * ri;
    }
    return grid_sync<2>(a.g, v, lds, [&](const double *t) {
      const double beta = t[0] / st->d;
      const double d = beta * st->d;
      st->beta = beta;
      st->d = d;
      st->rnorm2 = t[1];
      st->iter = j;
      bool conv;
      if (a.stop_rule == 1) {
        st->last = sqrt(t[1]);
        conv = st->last <= st->conv;
      } else {
        st->last = sqrt(fabs(d));
        conv = st->last < st->conv;   // src/Iterativos.f:73
      }
      if (conv) { st->status = S_OK; st->done = 1; }
      else if (!finite(d) || !finite(beta)) { st->status = S_NONFINITE; st->done = 1; }
      else if (j >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
    });
  }
  // PCG_FINAL: energy norm x.Ax (src/Iterativos.f:82-83) and ||b - Ax||^2
  double v[2] = {0.0, 0.0};
  for (int i = i0; i < i1; i += is) {
    const double ax = row_apply<F>(a.A, i, XVec{a.x});
    const double xi = a.x[i];
    const double ri = a.b[i] - ax;
    v[0] += xi * ax;
    v[1] += ri * ri;
  }
  return grid_sync<2>(a.g, v, lds, [&](const double *t) {
    st->energy = t[0];
    st->res2 = t[1];
  });
}

template <int F>
__global__ __launch_bounds__(1024) void pcg_persistent(PcgArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  if (!pcg_phase<F>(a, PCG_SETUP, lds)) return;
  while (!ld_int(&a.st->done)) {
    if (!pcg_phase<F>(a, PCG_A, lds)) return;
    if (ld_int(&a.st->done)) break;
    if (!pcg_phase<F>(a, PCG_B, lds)) return;
  }
  pcg_phase<F>(a, PCG_FINAL, lds);
}

template <int F>
__global__ __launch_bounds__(1024) void pcg_kernel(PcgArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  if ((a.phase == PCG_A || a.phase == PCG_B) && ld_int(&a.st->done)) return;
  pcg_phase<F>(a, a.phase, lds);
}

// ------------------------------------------------- PCG, one reduction --
//
// The same Jacobi-PCG iterates with the single-reduction recurrence of
// Chronopoulos & Gear (1989).  With u = M^-1 r, w = A u and s = A p:
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,
//   u' = M^-1 r,  w' = A u',  gamma = r.u',  delta = w'.u',
//   beta' = gamma / gamma_old,  alpha' = gamma / (delta - beta' gamma / alpha).
// gamma is the reference's d = r.z (src/Iterativos.f:71-74), so the stop test
// and iteration count are the reference's; p.Ap is never formed, so one grid
// reduction per iteration suffices.  The SpMV operand u' at the neighbours is
// formed on the fly from r, w and the previous s there; those three vectors
// alternate between two buffers (odd iterations read set 0 and write set 1).
//
// Each phase is split into its row part (cg1_rows: this block's rows, the
// values to reduce) and its scalar part (cg1_epi: the recurrence on the
// reduced values), so that both engines can drive it: the persistent kernel
// reduces at a grid barrier, the kernel-per-phase engine defers the scalar
// part to the next launch (fvc_device.h, Defer).
enum { CG1_SETUP = 0, CG1_W0 = 1, CG1_ITER = 2, CG1_FINAL = 3 };

__host__ __device__ constexpr int cg1_nred(int ph) { return ph == CG1_W0 ? 1 : ph == CG1_FINAL ? 2 : 3; }

struct Cg1Args {
  Mat A;
  Grid g;
  KState *st;
  const double *m;
  double *x;
  const double *b;
  double *r0, *r1, *w0, *w1, *s0, *s1, *p;
  double tol;
  int maxit, x0_zero, stop_rule, phase;
  int lanex;   // west/east neighbour operands from the adjacent lanes (lane_exchange)
  Defer d;
};

// u' at row j: (r_j - alpha (w_j + beta s_j)) / m_j  (s = w in iteration 1)
struct XCgNext {
  const double *r, *w, *s, *m;
  double alpha, beta;
  int first;
  __device__ __forceinline__ double operator()(int j) const {
    const double sj = first ? w[j] : w[j] + beta * s[j];
    return (r[j] - alpha * sj) / m[j];
  }
};
struct XDiv {
  const double *r, *m;
  __device__ __forceinline__ double operator()(int j) const { return r[j] / m[j]; }
};

// A CG1_ITER row's matrix operands, loadable ahead of the row.
template <int F>
struct Cg1Row {
  RowRegs<F> mr;
  bool ok;
};
template <int F>
__device__ __forceinline__ Cg1Row<F> cg1_load(const Cg1Args &a, int i) {
  Cg1Row<F> q;
  q.mr.load(a.A, i);
  q.ok = true;
  return q;
}

template <int F>
__device__ void cg1_rows(const Cg1Args &a, int ph, const KState *S, double *v,
                         const Cg1Row<F> *pre = nullptr) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  if (ph == CG1_SETUP) {
    // r = b - A x (or r = b for x0 = 0); d = r.M^-1 r; ||b||^2, ||r||^2
    for (int i = i0; i < i1; i += is) {
      double ax;
      if (a.x0_zero) {
        a.x[i] = 0.0;
        ax = 0.0;
      } else {
        ax = row_apply<F>(a.A, i, XVec{a.x});
      }
      const double bi = a.b[i];
      const double ri = bi - ax;
      a.r0[i] = ri;
      v[0] += ri * (ri / a.m[i]);
      v[1] += bi * bi;
      v[2] += ri * ri;
    }
  } else if (ph == CG1_W0) {
    // w = A u, u = r/m on the fly; -> w.u
    const XDiv xf{a.r0, a.m};
    for (int i = i0; i < i1; i += is) {
      const double wi = row_apply<F>(a.A, i, xf);
      a.w0[i] = wi;
      v[0] += wi * xf(i);
    }
  } else if (ph == CG1_ITER) {
    const int j = ld_int(&S->iter) + 1;
    const bool odd = (j & 1) != 0;
    const double *rc = odd ? a.r0 : a.r1, *wc = odd ? a.w0 : a.w1, *sc = odd ? a.s0 : a.s1;
    double *rn = odd ? a.r1 : a.r0, *wn = odd ? a.w1 : a.w0, *sn = odd ? a.s1 : a.s0;
    const XCgNext xf{rc, wc, sc, a.m, ld_dbl(&S->alpha), ld_dbl(&S->beta), j == 1};
    auto row = [&](int i, const Cg1Row<F> &q, auto lx) {
      const double mi = a.m[i];
      const double ui = rc[i] / mi;
      const double pi = xf.first ? ui : ui + xf.beta * a.p[i];
      const double si = xf.first ? wc[i] : wc[i] + xf.beta * sc[i];
      a.x[i] = a.x[i] + xf.alpha * pi;
      const double ri = rc[i] - xf.alpha * si;
      const double un = ri / mi;   // = xf(i)
      double wi;
      if constexpr (decltype(lx)::value) wi = q.mr.apply(a.A, i, lane_exchange(xf, un, i, i1));
      else wi = q.mr.apply(a.A, i, xf);
      a.p[i] = pi;
      sn[i] = si;
      rn[i] = ri;
      wn[i] = wi;
      v[0] += ri * un;
      v[1] += wi * un;
      v[2] += ri * ri;
    };
    auto sweep = [&](auto lx) {
      int i = i0;
      if (pre && pre->ok && i < i1) {
        row(i, *pre, lx);
        i += is;
      }
      for (; i < i1; i += is) row(i, cg1_load<F>(a, i), lx);
    };
    if (a.lanex) sweep(LaneX<true>{});
    else sweep(LaneX<false>{});
  } else {
    // CG1_FINAL: energy norm x.Ax (src/Iterativos.f:82-83) and ||b - Ax||^2
    for (int i = i0; i < i1; i += is) {
      const double ax = row_apply<F>(a.A, i, XVec{a.x});
      const double ri = a.b[i] - ax;
      v[0] += a.x[i] * ax;
      v[1] += ri * ri;
    }
  }
}

__device__ inline void cg1_epi(const Cg1Args &a, int ph, KState *st, const double *t) {
  if (ph == CG1_SETUP) {
    st->d = t[0];
    st->bnorm2 = t[1];
    st->rnorm2 = t[2];
    st->iter = 0;
    st->inner = 0;
    st->status = S_MAXIT;
    st->done = 0;
    st->beta = 0.0;
    if (a.stop_rule == 1) {
      st->conv = a.tol * sqrt(t[1]);
      st->last = sqrt(t[2]);
      if (st->last <= st->conv) { st->status = S_OK; st->done = 1; }
    } else {
      st->conv = a.tol * sqrt(fabs(t[0]));
      st->last = sqrt(fabs(t[0]));
    }
    if (!finite(t[0])) { st->status = S_NONFINITE; st->done = 1; }
    if (a.maxit <= 0 && !st->done) st->done = 1;
  } else if (ph == CG1_W0) {
    st->alpha = st->d / t[0];
    if (t[0] == 0.0) { st->status = S_BREAKDOWN; st->done = 1; }
    else if (!finite(st->alpha)) { st->status = S_NONFINITE; st->done = 1; }
  } else if (ph == CG1_ITER) {
    const int j = st->iter + 1;
    const double gamma = t[0];
    const double beta = gamma / st->d;
    const double den = t[1] - beta * gamma / st->alpha;
    const double alpha = gamma / den;
    st->beta = beta;
    st->alpha = alpha;
    st->d = gamma;
    st->rnorm2 = t[2];
    st->iter = j;
    bool conv;
    if (a.stop_rule == 1) {
      st->last = sqrt(t[2]);
      conv = st->last <= st->conv;
    } else {
      st->last = sqrt(fabs(gamma));
      conv = st->last < st->conv;   // src/Iterativos.f:73
    }
    if (conv) { st->status = S_OK; st->done = 1; }
    else if (!finite(gamma) || !finite(beta)) { st->status = S_NONFINITE; st->done = 1; }
    else if (den == 0.0) { st->status = S_BREAKDOWN; st->done = 1; }
    else if (!finite(alpha)) { st->status = S_NONFINITE; st->done = 1; }
    else if (j >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
  } else {
    st->energy = t[0];
    st->res2 = t[1];
  }
}

template <int F>
__device__ bool cg1_phase(const Cg1Args &a, int ph, double *lds) {
  double v[kMaxK] = {0.0, 0.0, 0.0, 0.0};
  cg1_rows<F>(a, ph, a.st, v);
  auto epi = [&](const double *t) { cg1_epi(a, ph, a.st, t); };
  switch (cg1_nred(ph)) {
    case 1: return grid_sync<1>(a.g, v, lds, epi);
    case 2: return grid_sync<2>(a.g, v, lds, epi);
    default: return grid_sync<3>(a.g, v, lds, epi);
  }
}

template <int F>
__global__ __launch_bounds__(1024) void cg1_persistent(Cg1Args a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  if (!cg1_phase<F>(a, CG1_SETUP, lds)) return;
  if (!ld_int(&a.st->done) && !cg1_phase<F>(a, CG1_W0, lds)) return;
  while (!ld_int(&a.st->done))
    if (!cg1_phase<F>(a, CG1_ITER, lds)) return;
  cg1_phase<F>(a, CG1_FINAL, lds);
}

// Kernel-per-phase engine: one launch per phase, the reduction deferred.
template <int F>
__global__ __launch_bounds__(1024) void cg1_kernel(Cg1Args a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ KState S;
  struct M {
    const Cg1Args &a;
    enum { kRed = 3 };
    using Pre = Cg1Row<F>;
    __device__ int nred(int ph) const { return cg1_nred(ph); }
    __device__ bool skip(int ph, const KState &s) const { return (ph == CG1_W0 || ph == CG1_ITER) && s.done; }
    __device__ Pre prefetch(int ph, const KState &) const {
      Pre q{};
      const int i = a.g.first();
      if (ph == CG1_ITER && i < a.g.row_end()) q = cg1_load<F>(a, i);
      return q;
    }
    __device__ void rows(int ph, const KState *s, double *v, const Pre &q) const { cg1_rows<F>(a, ph, s, v, &q); }
    __device__ void epi(int ph, KState *s, const double *t) const { cg1_epi(a, ph, s, t); }
  };
  deferred_phase(a.g, a.d, a.phase, M{a}, &S, lds);
}

// ------------------------------------------------------------- BiCGSTAB --
//
// src/Iterativos.f:140-234 / src/Iterativos_omp.f:161-274 in three phases per
// iteration, one per reduction the recurrence needs (alpha, omega, beta):
//   F1  p = b + beta (p - w v), z = p / m    (the previous iteration's tail)
//       v = A z, with z of the neighbours formed on the fly;   -> v.r
//   F2  b' = b - alpha v, z2 = b' / m, t = A z2 (z2 on the fly); -> t.b', t.t
//   F3  x = (x + alpha z) + w z2, b' = b' - w t;   -> b'.z2, r.b', b'.b'
// The vectors the on-the-fly forms read (b, p, v) alternate between two
// buffers so that no block overwrites a value another block still gathers.
// Every value is the reference's expression in the reference's order, so the
// iterates are those of the five-loop form.

enum { BI_SETUP = 0, BI_F1 = 1, BI_F2 = 2, BI_F3 = 3, BI_FINAL = 4 };

struct BicgArgs {
  Mat A;
  Grid g;
  KState *st;
  const double *m;
  double *x;
  const double *b;
  double *bb0, *bb1;   // running residual (the reference's overwritten b), two buffers
  double *p0, *p1, *v0, *v1;
  double *r, *z, *z2, *t;
  double tol;
  int maxit, x0_zero, stop_rule, phase;
};

// z of the F1 gather: the new direction over m, or p/m in the first iteration.
struct XBiDir {
  const double *bb, *p, *v, *m;
  double beta, w;
  int first;
  __device__ __forceinline__ double operator()(int j) const {
    if (first) return p[j] / m[j];
    return (bb[j] + beta * (p[j] - w * v[j])) / m[j];
  }
};
// z2 of the F2 gather: (b - alpha v) / m.
struct XBiRes {
  const double *bb, *v, *m;
  double alpha;
  __device__ __forceinline__ double operator()(int j) const { return (bb[j] - alpha * v[j]) / m[j]; }
};

template <int F>
__device__ bool bicg_phase(const BicgArgs &a, int ph, double *lds) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  KState *st = a.st;
  if (ph == BI_SETUP) {
    // z = Ax; r = b - z; p = r; b = p; z = p/m; d = r.z
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0; i < i1; i += is) {
      double ax;
      if (a.x0_zero) {
        a.x[i] = 0.0;
        ax = 0.0;
      } else {
        ax = row_apply<F>(a.A, i, XVec{a.x});
      }
      const double bi = a.b[i];
      const double ri = bi - ax;
      const double zi = ri / a.m[i];
      a.r[i] = ri;
      a.p0[i] = ri;
      a.bb0[i] = ri;
      v[0] += ri * zi;
      v[1] += ri * ri;   // = b.r of the first iteration (b = p = r)
      v[2] += bi * bi;
    }
    return grid_sync<3>(a.g, v, lds, [&](const double *t) {
      st->d = t[0];
      st->rr0 = t[1];
      st->rnorm2 = t[1];
      st->bnorm2 = t[2];
      st->iter = 0;
      st->inner = 0;
      st->status = S_MAXIT;
      st->done = 0;
      if (a.stop_rule == 1) {
        st->conv = a.tol * sqrt(t[2]);
        st->last = sqrt(t[1]);
        if (st->last <= st->conv) { st->status = S_OK; st->done = 1; }
      } else {
        st->conv = a.tol * sqrt(fabs(t[0]));
        st->last = sqrt(fabs(t[0]));
      }
      if (!finite(t[0])) { st->status = S_NONFINITE; st->done = 1; }
      if (a.maxit <= 0 && !st->done) st->done = 1;
    });
  }
  const int j = ld_int(&st->iter) + 1;   // iteration in progress
  const bool odd = (j & 1) != 0;         // buffers: odd j reads set 0, writes set 1
  const double *bbc = odd ? a.bb0 : a.bb1;
  double *bbn = odd ? a.bb1 : a.bb0;
  const double *pc = odd ? a.p0 : a.p1;
  double *pn = odd ? a.p1 : a.p0;
  const double *vc = odd ? a.v0 : a.v1;
  double *vn = odd ? a.v1 : a.v0;
  if (ph == BI_F1) {
    const XBiDir zf{bbc, pc, vc, a.m, ld_dbl(&st->beta), ld_dbl(&st->omega), j == 1};
    double v[1] = {0.0};
    for (int i = i0; i < i1; i += is) {
      const double pi = (j == 1) ? pc[i] : bbc[i] + zf.beta * (pc[i] - zf.w * vc[i]);
      const double zi = pi / a.m[i];
      const double vi = row_apply<F>(a.A, i, zf);
      pn[i] = pi;
      a.z[i] = zi;
      vn[i] = vi;
      v[0] += vi * a.r[i];
    }
    return grid_sync<1>(a.g, v, lds, [&](const double *t) {
      const double alpha = st->rr0 / t[0];
      st->alpha = alpha;
      if (t[0] == 0.0) { st->status = S_BREAKDOWN; st->done = 1; st->iter = j; }
      else if (!finite(alpha)) { st->status = S_NONFINITE; st->done = 1; st->iter = j; }
    });
  }
  if (ph == BI_F2) {
    const XBiRes zf{bbc, vn, a.m, ld_dbl(&st->alpha)};
    double v[2] = {0.0, 0.0};
    for (int i = i0; i < i1; i += is) {
      const double bi = bbc[i] - zf.alpha * vn[i];
      const double ti = row_apply<F>(a.A, i, zf);
      bbn[i] = bi;
      a.z2[i] = bi / a.m[i];
      a.t[i] = ti;
      v[0] += ti * bi;
      v[1] += ti * ti;
    }
    return grid_sync<2>(a.g, v, lds, [&](const double *t) {
      const double w = t[0] / t[1];
      st->omega = w;
      if (t[1] == 0.0 || w == 0.0) { st->status = S_BREAKDOWN; st->done = 1; st->iter = j; }
      else if (!finite(w)) { st->status = S_NONFINITE; st->done = 1; st->iter = j; }
    });
  }
  if (ph == BI_F3) {
    const double alpha = ld_dbl(&st->alpha), w = ld_dbl(&st->omega);
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0; i < i1; i += is) {
      const double z2i = a.z2[i];
      a.x[i] = (a.x[i] + alpha * a.z[i]) + w * z2i;
      const double bi = bbn[i] - w * a.t[i];
      bbn[i] = bi;
      v[0] += bi * z2i;
      v[1] += a.r[i] * bi;
      v[2] += bi * bi;
    }
    return grid_sync<3>(a.g, v, lds, [&](const double *t) {
      st->iter = j;
      st->rnorm2 = t[2];
      const double beta1 = t[0] / st->d;
      bool conv;
      if (a.stop_rule == 1) {
        st->last = sqrt(t[2]);
        conv = st->last <= st->conv;
      } else {
        st->last = sqrt(fabs(beta1));
        conv = st->last < st->conv;   // src/Iterativos.f:189-190
      }
      const double beta = (t[1] / st->rr0) * (st->alpha / st->omega);
      st->beta = beta;
      st->rr0 = t[1];
      if (conv) { st->status = S_OK; st->done = 1; }
      else if (!finite(beta1) || !finite(beta)) { st->status = S_NONFINITE; st->done = 1; }
      else if (j >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
    });
  }
  // BI_FINAL: energy x.Ax (src/Iterativos.f:204-205) and ||b - Ax||^2
  double v[2] = {0.0, 0.0};
  for (int i = i0; i < i1; i += is) {
    const double ax = row_apply<F>(a.A, i, XVec{a.x});
    const double ri = a.b[i] - ax;
    v[0] += a.x[i] * ax;
    v[1] += ri * ri;
  }
  return grid_sync<2>(a.g, v, lds, [&](const double *t) {
    st->energy = t[0];
    st->res2 = t[1];
  });
}

template <int F>
__global__ __launch_bounds__(1024) void bicg_persistent(BicgArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  if (!bicg_phase<F>(a, BI_SETUP, lds)) return;
  while (!ld_int(&a.st->done)) {
    if (!bicg_phase<F>(a, BI_F1, lds)) return;
    if (ld_int(&a.st->done)) break;
    if (!bicg_phase<F>(a, BI_F2, lds)) return;
    if (ld_int(&a.st->done)) break;
    if (!bicg_phase<F>(a, BI_F3, lds)) return;
  }
  bicg_phase<F>(a, BI_FINAL, lds);
}

template <int F>
__global__ __launch_bounds__(1024) void bicg_kernel(BicgArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  if (a.phase != BI_SETUP && a.phase != BI_FINAL && ld_int(&a.st->done)) return;
  bicg_phase<F>(a, a.phase, lds);
}

// ------------------------------------------------ BiCGSTAB, two phases --
//
// The reference's BiCGSTAB (src/Iterativos.f:140-234) with two grid
// reductions per iteration instead of three.  Iteration j:
//   P(j)  completes iteration j-1 for the own rows -- b = s - w t and
//         x = (x + alpha p/m) + w s/m -- then forms p = b + beta (p - w v)
//         and v = A (p/m), gathering p/m of the neighbours from s, t, p, v;
//         reduces rhat.v (alpha of j) and, exactly, the quantities of
//         iteration j-1's stop test: b.(s/m) (src/Iterativos.f:188) and b.b.
//   Q(j)  s = b - alpha v, t = A (s/m) (s/m of the neighbours on the fly);
//         reduces t.s, t.t (omega, src/Iterativos.f:180-181) and rhat.s,
//         rhat.t, which give rhat.b_new = rhat.s - omega rhat.t: the
//         reference's dot(r, b) (src/Iterativos.f:192) without a third pass.
// The stop test of iteration j-1 is decided in P(j)'s recurrence, one phase
// later than the reference, with the same exact dot products; x is then
// x_(j-1) and the extra direction P(j) built is dropped.  p and v alternate
// between two buffers (odd iterations read set 0 and write set 1).  Row and
// scalar parts are split as for the one-reduction PCG.
enum { B2_SETUP = 0, B2_P = 1, B2_Q = 2, B2_FINAL = 3 };

__host__ __device__ constexpr int bicg2_nred(int ph) { return ph == B2_Q ? 4 : ph == B2_FINAL ? 2 : 3; }

struct Bicg2Args {
  Mat A;
  Grid g;
  KState *st;
  const double *m;
  double *x;
  const double *b;
  double *bb;               // the reference's running residual b (own rows)
  double *p0, *p1, *v0, *v1;
  double *s, *t, *r;        // half-step residual, A M^-1 s, shadow residual
  double tol;
  int maxit, x0_zero, stop_rule, phase;
  int lanex;   // west/east neighbour operands from the adjacent lanes (lane_exchange)
  Defer d;
};

// p/m at row j for P(j), j > 1: ((s - w t) + beta (p - w v)) / m
struct XBi2Dir {
  const double *s, *t, *p, *v, *m;
  double beta, w;
  __device__ __forceinline__ double operator()(int j) const {
    return ((s[j] - w * t[j]) + beta * (p[j] - w * v[j])) / m[j];
  }
};

// A P or Q row's matrix operands, loadable ahead of the row.
template <int F>
struct Bi2Row {
  RowRegs<F> mr;
  bool ok;
};
template <int F>
__device__ __forceinline__ Bi2Row<F> bicg2_load(const Bicg2Args &a, int i) {
  Bi2Row<F> q;
  q.mr.load(a.A, i);
  q.ok = true;
  return q;
}

template <int F>
__device__ void bicg2_rows(const Bicg2Args &a, int ph, const KState *S, double *v,
                           const Bi2Row<F> *pre = nullptr) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  if (ph == B2_SETUP) {
    // r = b - Ax; rhat = r; p = r; b = r; d = r.(r/m); rho = b.rhat = r.r
    for (int i = i0; i < i1; i += is) {
      double ax;
      if (a.x0_zero) {
        a.x[i] = 0.0;
        ax = 0.0;
      } else {
        ax = row_apply<F>(a.A, i, XVec{a.x});
      }
      const double bi = a.b[i];
      const double ri = bi - ax;
      a.r[i] = ri;
      a.p0[i] = ri;
      a.bb[i] = ri;
      v[0] += ri * (ri / a.m[i]);
      v[1] += ri * ri;
      v[2] += bi * bi;
    }
    return;
  }
  if (ph == B2_FINAL) {
    // energy x.Ax (src/Iterativos.f:204-205) and ||b - Ax||^2
    for (int i = i0; i < i1; i += is) {
      const double ax = row_apply<F>(a.A, i, XVec{a.x});
      const double ri = a.b[i] - ax;
      v[0] += a.x[i] * ax;
      v[1] += ri * ri;
    }
    return;
  }
  const int j = ld_int(&S->iter) + 1;
  const bool odd = (j & 1) != 0;
  const double *pc = odd ? a.p0 : a.p1, *vc = odd ? a.v0 : a.v1;
  double *pn = odd ? a.p1 : a.p0, *vn = odd ? a.v1 : a.v0;
  if (ph == B2_P) {
    if (j == 1) {
      const XDiv xf{pc, a.m};
      for (int i = i0; i < i1; i += is) {
        const double vi = row_apply<F>(a.A, i, xf);
        pn[i] = pc[i];
        vn[i] = vi;
        v[0] += vi * a.r[i];
      }
    } else {
      const double alpha = ld_dbl(&S->alpha);
      const XBi2Dir xf{a.s, a.t, pc, vc, a.m, ld_dbl(&S->beta), ld_dbl(&S->omega)};
      auto row = [&](int i, const Bi2Row<F> &q, auto lx) {
        const double mi = a.m[i];
        const double si = a.s[i];
        const double z2i = si / mi;
        const double bi = si - xf.w * a.t[i];
        a.x[i] = (a.x[i] + alpha * (pc[i] / mi)) + xf.w * z2i;
        a.bb[i] = bi;
        const double pi = bi + xf.beta * (pc[i] - xf.w * vc[i]);
        double vi;
        if constexpr (decltype(lx)::value) vi = q.mr.apply(a.A, i, lane_exchange(xf, pi / mi, i, i1));
        else vi = q.mr.apply(a.A, i, xf);
        pn[i] = pi;
        vn[i] = vi;
        v[0] += vi * a.r[i];
        v[1] += bi * z2i;
        v[2] += bi * bi;
      };
      auto sweep = [&](auto lx) {
        int i = i0;
        if (pre && pre->ok && i < i1) {
          row(i, *pre, lx);
          i += is;
        }
        for (; i < i1; i += is) row(i, bicg2_load<F>(a, i), lx);
      };
      if (a.lanex) sweep(LaneX<true>{});
      else sweep(LaneX<false>{});
    }
    return;
  }
  // B2_Q
  const XBiRes xf{a.bb, vn, a.m, ld_dbl(&S->alpha)};
  auto row = [&](int i, const Bi2Row<F> &q, auto lx) {
    const double si = a.bb[i] - xf.alpha * vn[i];
    double ti;
    if constexpr (decltype(lx)::value) ti = q.mr.apply(a.A, i, lane_exchange(xf, si / a.m[i], i, i1));
    else ti = q.mr.apply(a.A, i, xf);
    const double ri = a.r[i];
    a.s[i] = si;
    a.t[i] = ti;
    v[0] += ti * si;
    v[1] += ti * ti;
    v[2] += ri * si;
    v[3] += ri * ti;
  };
  auto sweep = [&](auto lx) {
    int i = i0;
    if (pre && pre->ok && i < i1) {
      row(i, *pre, lx);
      i += is;
    }
    for (; i < i1; i += is) row(i, bicg2_load<F>(a, i), lx);
  };
  if (a.lanex) sweep(LaneX<true>{});
  else sweep(LaneX<false>{});
}

__device__ inline void bicg2_epi(const Bicg2Args &a, int ph, KState *st, const double *t) {
  if (ph == B2_SETUP) {
    st->d = t[0];
    st->rr0 = t[1];
    st->rnorm2 = t[1];
    st->bnorm2 = t[2];
    st->iter = 0;
    st->inner = 0;
    st->status = S_MAXIT;
    st->done = 0;
    if (a.stop_rule == 1) {
      st->conv = a.tol * sqrt(t[2]);
      st->last = sqrt(t[1]);
      if (st->last <= st->conv) { st->status = S_OK; st->done = 1; }
    } else {
      st->conv = a.tol * sqrt(fabs(t[0]));
      st->last = sqrt(fabs(t[0]));
    }
    if (!finite(t[0])) { st->status = S_NONFINITE; st->done = 1; }
    if (a.maxit <= 0 && !st->done) st->done = 1;
  } else if (ph == B2_P) {
    const int j = st->iter + 1;
    if (j > 1) {
      // the stop test of iteration j-1 (src/Iterativos.f:188-190)
      const double beta1 = t[1] / st->d;
      st->rnorm2 = t[2];
      bool conv;
      if (a.stop_rule == 1) {
        st->last = sqrt(t[2]);
        conv = st->last <= st->conv;
      } else {
        st->last = sqrt(fabs(beta1));
        conv = st->last < st->conv;
      }
      if (conv) { st->status = S_OK; st->done = 1; return; }
      if (!finite(beta1)) { st->status = S_NONFINITE; st->done = 1; return; }
      if (j - 1 >= a.maxit) { st->status = S_MAXIT; st->done = 1; return; }
    }
    const double alpha = st->rr0 / t[0];
    st->alpha = alpha;
    if (t[0] == 0.0) { st->status = S_BREAKDOWN; st->done = 1; }
    else if (!finite(alpha)) { st->status = S_NONFINITE; st->done = 1; }
  } else if (ph == B2_Q) {
    const double w = t[0] / t[1];
    const double rho = t[2] - w * t[3];
    const double beta = (rho / st->rr0) * (st->alpha / w);
    st->omega = w;
    st->beta = beta;
    st->rr0 = rho;
    st->iter = st->iter + 1;
    if (t[1] == 0.0 || w == 0.0) { st->status = S_BREAKDOWN; st->done = 1; }
    else if (!finite(w) || !finite(beta)) { st->status = S_NONFINITE; st->done = 1; }
  } else {
    st->energy = t[0];
    st->res2 = t[1];
  }
}

template <int F>
__device__ bool bicg2_phase(const Bicg2Args &a, int ph, double *lds) {
  double v[kMaxK] = {0.0, 0.0, 0.0, 0.0};
  bicg2_rows<F>(a, ph, a.st, v);
  auto epi = [&](const double *t) { bicg2_epi(a, ph, a.st, t); };
  switch (bicg2_nred(ph)) {
    case 2: return grid_sync<2>(a.g, v, lds, epi);
    case 3: return grid_sync<3>(a.g, v, lds, epi);
    default: return grid_sync<4>(a.g, v, lds, epi);
  }
}

template <int F>
__global__ __launch_bounds__(1024) void bicg2_persistent(Bicg2Args a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  if (!bicg2_phase<F>(a, B2_SETUP, lds)) return;
  while (!ld_int(&a.st->done)) {
    if (!bicg2_phase<F>(a, B2_P, lds)) return;
    if (ld_int(&a.st->done)) break;
    if (!bicg2_phase<F>(a, B2_Q, lds)) return;
  }
  bicg2_phase<F>(a, B2_FINAL, lds);
}

// Kernel-per-phase engine: one launch per phase, the reduction deferred.
template <int F>
__global__ __launch_bounds__(1024) void bicg2_kernel(Bicg2Args a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ KState S;
  struct M {
    const Bicg2Args &a;
    enum { kRed = 4 };
    using Pre = Bi2Row<F>;
    __device__ int nred(int ph) const { return bicg2_nred(ph); }
    __device__ bool skip(int ph, const KState &s) const { return (ph == B2_P || ph == B2_Q) && s.done; }
    __device__ Pre prefetch(int ph, const KState &) const {
      Pre q{};
      const int i = a.g.first();
      if ((ph == B2_P || ph == B2_Q) && i < a.g.row_end()) q = bicg2_load<F>(a, i);
      return q;
    }
    __device__ void rows(int ph, const KState *s, double *v, const Pre &q) const { bicg2_rows<F>(a, ph, s, v, &q); }
    __device__ void epi(int ph, KState *s, const double *t) const { bicg2_epi(a, ph, s, t); }
  };
  deferred_phase(a.g, a.d, a.phase, M{a}, &S, lds);
}

// ---------------------------------------------------------------- GMRES --

constexpr int kMaxGmresK = 256;
constexpr int kCgsWide = 32;   // CGS2: largest restart length
constexpr int kMaxGmresLag = 64;  // lagged CGS2 (fvc_gmres_lag.h): largest restart length
constexpr int kCgsChunk = 16;  // CGS2: dot products reduced together

enum { GM_START = 0, GM_MV = 1, GM_MGS = 2, GM_XUPD = 3, GM_FINAL = 4, GM_C1 = 5, GM_C2 = 6, GM_C3 = 7 };

struct GmresArgs {
  Mat A;
  Grid g;
  KState *st;
  const double *m;
  double *x;
  const double *b;
  double *gv;      // basis g(neq, k+1), column-major
  double *h;       // h(k+1, k)
  double *cs;      // c(k)  (1-based, k+1 entries)
  double *sn;      // s(k)
  double *e;       // e(k+1) (k+2 entries)
  double *y;       // y(k)
  double *nrm;     // normaliser of each basis column (k+2 entries)
  double tol;
  int k, maxit, phase;
  int stop_rule;   // 0 the reference's test on the preconditioned residual, 1 the true residual
  int orth;        // 0 modified Gram-Schmidt (the reference's order), 1 CGS2,
                   // 2 CGS2 lagged by one step (fvc_gmres_lag.h)
  int chunk;       // CGS2, one kernel per phase: the chunk of dot products
  size_t ld;       // leading dimension of gv (= neq)
  // lagged CGS2 only: the candidate u and w = M^-1 A u (neq each), the
  // Hessenberg before the Givens rotations h(k+1, k), the coefficients of the
  // update step (nu, a, c), the per-block partials (lag_kp values per block)
  // and their totals
  double *gu, *gw, *hraw, *lag, *lag_part, *lag_tot;
  int lag_kp;
};

// The end of Arnoldi step i once h(1..i+1, i) is known: the previous Givens
// rotations applied to column i, the new one, e, the convergence test and,
// when the cycle ends, the back substitution (src/Iterativos.f:363-389).
// Runs on one thread.
__device__ __forceinline__ void gmres_close_step(const GmresArgs &a, KState *st, int i) {
  const int k = a.k;
  auto H = [&](int r, int c) -> double & { return a.h[(size_t)(c - 1) * (k + 1) + (r - 1)]; };
  for (int jj = 1; jj <= i - 1; ++jj) {
    const double aux1 = a.cs[jj] * H(jj, i) + a.sn[jj] * H(jj + 1, i);
    const double aux2 = -a.sn[jj] * H(jj, i) + a.cs[jj] * H(jj + 1, i);
    H(jj, i) = aux1;
    H(jj + 1, i) = aux2;
  }
  const double rr = sqrt(H(i, i) * H(i, i) + H(i + 1, i) * H(i + 1, i));
  a.cs[i] = H(i, i) / rr;
  a.sn[i] = H(i + 1, i) / rr;
  H(i, i) = rr;
  H(i + 1, i) = 0.0;
  a.e[i + 1] = -a.sn[i] * a.e[i];
  a.e[i] = a.cs[i] * a.e[i];
  st->inner = st->inner + 1;
  const bool conv = fabs(a.e[i + 1]) <= st->econv;   // src/Iterativos.f:376
  if (conv || i == k || !finite(a.e[i + 1])) {
    const int nI = i;
    st->ni = nI;
    st->inner_done = 1;
    // back substitution h y = e (src/Iterativos.f:382-389)
    a.y[nI] = a.e[nI] / H(nI, nI);
    for (int ii = nI - 1; ii >= 1; --ii) {
      double yi = 0.0;
      for (int jj = ii + 1; jj <= nI; ++jj) yi = yi - H(ii, jj) * a.y[jj];
      a.y[ii] = (yi + a.e[ii]) / H(ii, ii);
    }
  } else {
    st->step_i = i + 1;
    st->mgs_j = 1;
  }
}

// Classical Gram-Schmidt with one re-orthogonalisation ("twice is enough",
// Giraud, Langou, Rozloznik 2005): per Arnoldi step i two passes of dot
// products against columns 1..i and one norm, instead of the i+1 sequential
// reductions of modified Gram-Schmidt:
//   C1  w = A (g_i / nrm_i) / m into column i+1;  h1_j = w . g_j, j <= i
//   C2  column i stored normalised; w -= sum_j h1_j g_j;  h2_j = w . g_j
//   C3  w -= sum_j h2_j g_j;  ||w||;  h(j,i) = h1_j + h2_j; Givens as MGS
// The dot products of a pass are reduced kCgsChunk at a time (chunk c covers
// columns c*kCgsChunk+1 ..), so a step takes 2 ceil(i/kCgsChunk) + 1 grid
// reductions and no thread holds more than kCgsChunk partial sums.  Chunk 0
// of C1 also does the SpMV and of C2 the first projection; later chunks read
// w back (own rows).  h2 waits in y, which is free until the cycle's back
// substitution.
template <int F>
__device__ __forceinline__ bool gmres_cgs_phase(const GmresArgs &a, int ph, int c, double *lds) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  KState *st = a.st;
  const int k = a.k;
  const int i = ld_int(&st->step_i);
  const double ni = ld_dbl(&a.nrm[i]);
  double *gi = a.gv + (size_t)(i - 1) * a.ld;
  double *gn = a.gv + (size_t)i * a.ld;
  auto H = [&](int r, int col) -> double & { return a.h[(size_t)(col - 1) * (k + 1) + (r - 1)]; };
  // Row r of columns p0+1 .. p0+kCgsLoads, loaded together: a column past i
  // reads column i again (an L1 hit), so no load waits behind a branch.
  constexpr int kCgsLoads = 8;
  auto load_group = [&](double *gq, int p0, int r) {
#pragma unroll
    for (int q = 0; q < kCgsLoads; ++q) {
      const int col = min(p0 + q + 1, i);
      gq[q] = a.gv[(size_t)(col - 1) * a.ld + r];
    }
  };
  const int q0 = c * kCgsChunk;
  // w . g_col for the chunk's columns; column i is stored normalised from C2 on
  auto dots = [&](double *v, double w, int r) {
#pragma unroll
    for (int u = 0; u < kCgsChunk; u += kCgsLoads) {
      if (q0 + u >= i) break;
      double gq[kCgsLoads];
      load_group(gq, q0 + u, r);
#pragma unroll
      for (int q = 0; q < kCgsLoads; ++q) {
        const int col = q0 + u + q + 1;
        const double g = (ph == GM_C1 && col == i) ? gq[q] / ni : gq[q];
        if (col <= i) v[u + q] += w * g;
      }
    }
  };
  // w - h_1 g_1 - h_2 g_2 - ... - h_i g_i, in that order; gir = g_i (normalised)
  auto project = [&](double w, const double *h, double gir, int r) {
    for (int p0 = 0; p0 < i; p0 += kCgsLoads) {
      double gq[kCgsLoads];
      load_group(gq, p0, r);
#pragma unroll
      for (int q = 0; q < kCgsLoads; ++q) {
        const int col = p0 + q + 1;
        if (col <= i) w = w - h[col - 1] * (col == i ? gir : gq[q]);
      }
    }
    return w;
  };
  double v[kCgsChunk];
#pragma unroll
  for (int q = 0; q < kCgsChunk; ++q) v[q] = 0.0;
  if (ph == GM_C1 || (ph == GM_C2 && c > 0)) {
    if (ph == GM_C1 && c == 0) {
      const XScaled xf{gi, ni};
      for (int r = i0; r < i1; r += is) {
        const double w = row_apply<F>(a.A, r, xf) / a.m[r];
        gn[r] = w;
        dots(v, w, r);
      }
    } else {
      for (int r = i0; r < i1; r += is) dots(v, gn[r], r);
    }
    return grid_sync<kCgsChunk>(a.g, v, lds, [&](const double *t) {
#pragma unroll
      for (int q = 0; q < kCgsChunk; ++q) {
        if (q0 + q < i) {
          if (ph == GM_C1) H(q0 + q + 1, i) = t[q];
          else a.y[q0 + q + 1] = t[q];
        }
      }
    });
  }
  // the projection coefficients of C2 (h1) and C3 (h2), shared by the block
  __shared__ double h[kCgsWide];
  {
    const double *src = ph == GM_C2 ? a.h + (size_t)(i - 1) * (k + 1) : a.y + 1;
    for (int q = (int)threadIdx.x; q < i; q += (int)blockDim.x) h[q] = ld_dbl(src + q);
    __syncthreads();
  }
  if (ph == GM_C2) {
    // chunk 0: normalise column i, project, first dot products
    for (int r = i0; r < i1; r += is) {
      const double gir = gi[r] / ni;
      gi[r] = gir;
      const double w = project(gn[r], h, gir, r);
      gn[r] = w;
      dots(v, w, r);
    }
    return grid_sync<kCgsChunk>(a.g, v, lds, [&](const double *t) {
#pragma unroll
      for (int q = 0; q < kCgsChunk; ++q)
        if (q < i) a.y[q + 1] = t[q];
    });
  }
  // GM_C3
  double nv[1] = {0.0};
  for (int r = i0; r < i1; r += is) {
    const double w = project(gn[r], h, gi[r], r);
    gn[r] = w;
    nv[0] += w * w;
  }
  return grid_sync<1>(a.g, nv, lds, [&](const double *t) {
    for (int q = 1; q <= i; ++q) H(q, i) = H(q, i) + a.y[q];
    const double norm = sqrt(t[0]);
    H(i + 1, i) = norm;
    a.nrm[i + 1] = norm;
    gmres_close_step(a, st, i);
  });
}

template <int F>
__device__ __forceinline__ bool gmres_phase(const GmresArgs &a, int ph, double *lds) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  KState *st = a.st;
  const int k = a.k;
  auto col = [&](int c) { return a.gv + (size_t)(c - 1) * a.ld; };
  auto H = [&](int i, int j) -> double & { return a.h[(size_t)(j - 1) * (k + 1) + (i - 1)]; };
  if (ph == GM_START) {
    // g1 = (b - A x)/m; e(1) = ||g1||; cycle 1 starts from x = 0 and sets econv
    const int first = ld_int(&st->cycle) == 0;
    double *g1 = a.orth == 2 ? a.gu : col(1);   // lagged CGS2: the first candidate
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0; i < i1; i += is) {
      double ax;
      const double bi = a.b[i];
      if (first) {
        a.x[i] = 0.0;
        ax = 0.0;
        v[1] += bi * bi;
      } else {
        ax = row_apply<F>(a.A, i, XVec{a.x});
      }
      const double ri = bi - ax;
      const double gi = ri / a.m[i];
      g1[i] = gi;
      v[0] += gi * gi;
      v[2] += ri * ri;
    }
    return grid_sync<3>(a.g, v, lds, [&](const double *t) {
      if (first) {
        st->bnorm2 = t[1];
        st->econv = a.tol * sqrt(t[1]);
        st->conv = st->econv;
        st->inner = 0;
        st->status = S_MAXIT;
        st->done = 0;
      }
      const double e1 = sqrt(t[0]);
      if (a.stop_rule == 1) {
        const double rt = sqrt(t[2]);
        st->rnorm2 = t[2];
        st->last = rt;
        st->iter = st->cycle;
        const bool stop = rt <= st->conv || !finite(rt) || !finite(e1) || st->cycle >= a.maxit;
        if (rt <= st->conv) st->status = S_OK;
        else if (!finite(rt) || !finite(e1)) st->status = S_NONFINITE;
        if (stop) {
          st->done = 1;
          st->inner_done = 1;
          return;
        }
        // this cycle's target for the preconditioned residual
        st->econv = e1 * (st->conv / rt);
      }
      st->cycle = st->cycle + 1;
      a.e[1] = e1;
      a.nrm[1] = e1;
      st->step_i = 1;
      st->mgs_j = 1;
      st->inner_done = 0;
      st->ni = 0;
      if (e1 == 0.0) {          // exact solution already: nothing to orthogonalise
        st->status = S_OK;
        st->done = 1;
        st->inner_done = 1;
        st->last = 0.0;
        st->iter = st->cycle;
      } else if (!finite(e1)) {
        st->status = S_NONFINITE;
        st->done = 1;
        st->inner_done = 1;
      }
    });
  }
  if (ph == GM_MV) {
    // g(i+1) = A g(i) / m; h(1,i) = g(i+1).g(1)
    const int i = ld_int(&st->step_i);
    const double ni = ld_dbl(&a.nrm[i]);
    const double n1 = ld_dbl(&a.nrm[1]);
    const double *gi = col(i);
    const double *g1 = col(1);
    double *gn = col(i + 1);
    const XScaled xf{gi, ni};
    double v[1] = {0.0};
    for (int r = i0; r < i1; r += is) {
      const double w = row_apply<F>(a.A, r, xf) / a.m[r];
      gn[r] = w;
      const double g1r = (i == 1) ? g1[r] / n1 : g1[r];
      v[0] += w * g1r;
    }
    return grid_sync<1>(a.g, v, lds, [&](const double *t) {
      H(1, i) = t[0];
      st->mgs_j = 1;
    });
  }
  if (ph == GM_MGS) {
    // g(i+1) -= h(j,i) g(j); then h(j+1,i) = g(i+1).g(j+1), or the norm at j = i
    const int i = ld_int(&st->step_i);
    const int j = ld_int(&st->mgs_j);
    const double hj = ld_dbl(&a.h[(size_t)(i - 1) * (k + 1) + (j - 1)]);
    double *gi = col(i);
    double *gn = col(i + 1);
    const double *gj = col(j);
    const double *gj1 = col(j + 1);
    const double ni = ld_dbl(&a.nrm[i]);
    double v[1] = {0.0};
    for (int r = i0; r < i1; r += is) {
      if (j == 1) gi[r] = gi[r] / ni;      // store column i normalised
      const double w = gn[r] - hj * gj[r];
      gn[r] = w;
      v[0] += (j < i) ? w * gj1[r] : w * w;
    }
    return grid_sync<1>(a.g, v, lds, [&](const double *t) {
      if (j < i) {
        H(j + 1, i) = t[0];
        st->mgs_j = j + 1;
        return;
      }
      const double norm = sqrt(t[0]);
      H(i + 1, i) = norm;
      a.nrm[i + 1] = norm;
      gmres_close_step(a, st, i);
    });
  }
  if (ph == GM_XUPD) {
    // x += sum_j y(j) g(j), j in order (src/Iterativos.f:393-397); convergence check
    const int nI = ld_int(&st->ni);
    __shared__ double ysh[kMaxGmresK + 2];
    for (int j = (int)threadIdx.x + 1; j <= nI; j += blockDim.x) ysh[j] = ld_dbl(&a.y[j]);
    __syncthreads();
    double v[1] = {0.0};
    for (int r = i0; r < i1; r += is) {
      double xr = a.x[r];
      for (int j = 1; j <= nI; ++j) xr = xr + ysh[j] * a.gv[(size_t)(j - 1) * a.ld + r];
      a.x[r] = xr;
    }
    return grid_sync<0>(a.g, v, lds, [&](const double *) {
      const double en = a.e[nI + 1];
      st->last = fabs(en);
      st->iter = st->cycle;
      if (!finite(en)) { st->status = S_NONFINITE; st->done = 1; }
      else if (a.stop_rule == 1) {}   // the next cycle's start tests the true residual
      else if (fabs(en) <= st->econv) { st->status = S_OK; st->done = 1; }   // :402
      else if (st->cycle >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
      st->inner_done = 0;
    });
  }
  // GM_FINAL: energy x.b (src/Iterativos.f:412) and ||b - Ax||^2
  double v[2] = {0.0, 0.0};
  for (int i = i0; i < i1; i += is) {
    const double ax = row_apply<F>(a.A, i, XVec{a.x});
    const double bi = a.b[i];
    const double ri = bi - ax;
    v[0] += a.x[i] * bi;
    v[1] += ri * ri;
  }
  return grid_sync<2>(a.g, v, lds, [&](const double *t) {
    st->energy = t[0];
    st->res2 = t[1];
  });
}

// One persistent launch per solve.  Orth 0: modified Gram-Schmidt; 1: CGS2,
// whose 16 partial sums per pass would not fit the registers of 1024 threads
// (it spilled there), so it runs 512 per block, at most 128 registers so
// that two blocks share a CU.
template <int F, int Orth>
__global__ __launch_bounds__(Orth == 1 ? 512 : 1024) __attribute__((amdgpu_waves_per_eu(Orth == 1 ? 4 : 1)))
void gmres_persistent(GmresArgs a) {
  __shared__ double lds[(Orth == 1 ? kCgsChunk : kMaxK) * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  while (!ld_int(&a.st->done)) {
    if (!gmres_phase<F>(a, GM_START, lds)) return;
    if (ld_int(&a.st->done)) break;
    for (int i = 1; i <= a.k; ++i) {
      if constexpr (Orth == 1) {
        const int nch = (i + kCgsChunk - 1) / kCgsChunk;
        for (int c = 0; c < nch; ++c)
          if (!gmres_cgs_phase<F>(a, GM_C1, c, lds)) return;
        for (int c = 0; c < nch; ++c)
          if (!gmres_cgs_phase<F>(a, GM_C2, c, lds)) return;
        if (!gmres_cgs_phase<F>(a, GM_C3, 0, lds)) return;
      } else {
        if (!gmres_phase<F>(a, GM_MV, lds)) return;
        for (int j = 1; j <= i; ++j)
          if (!gmres_phase<F>(a, GM_MGS, lds)) return;
      }
      if (ld_int(&a.st->inner_done)) break;
    }
    if (!gmres_phase<F>(a, GM_XUPD, lds)) return;
  }
  gmres_phase<F>(a, GM_FINAL, lds);
}

template <int F, int Orth>
__global__ __launch_bounds__(1024) void gmres_kernel(GmresArgs a) {
  __shared__ double lds[(Orth == 1 ? kCgsChunk : kMaxK) * kMaxWaves];
  const int ph = a.phase;
  if (ph != GM_FINAL && ld_int(&a.st->done)) return;
  if ((ph == GM_MV || ph == GM_MGS || ph >= GM_C1) && ld_int(&a.st->inner_done)) return;
  if constexpr (Orth == 1) {
    if (ph >= GM_C1) {
      gmres_cgs_phase<F>(a, ph, a.chunk, lds);
      return;
    }
  }
  gmres_phase<F>(a, ph, lds);
}

}  // namespace fvc
