// fvc_mixed_krylov.h -- mixed-precision BiCGSTAB and GMRES(k): FP32 matrix
// values and FP32 inner iterations on the Jacobi-scaled system
//     (S A S) y = S r,   S = diag(1/sqrt(a_ii)),   x += S y,
// (the matrix of the mixed PCG, fvc_mixed.h, with its unit diagonal), FP64
// solution, FP64 residual recomputation, FP64 reductions, and for GMRES the
// Hessenberg matrix, Givens rotations and back substitution in FP64 on the
// device.  Both stop only on the FP64 true residual ||b - A x|| <= tol ||b||.
// In exact arithmetic the inner iterations are those of Jacobi-preconditioned
// BiCGSTAB / GMRES on A (src/Iterativos.f:140-234, 271-446) with the Jacobi
// preconditioner split symmetrically.
#pragma once
#include "fvc_mixed.h"

namespace fvc {

// ------------------------------------------------------ mixed BiCGSTAB --
//
// The two-phase iteration of bicg2 (fvc_krylov.h) in FP32 on B = S A S:
//   P(j)  completes iteration j-1 (r = s - w t, y += alpha p + w s), forms
//         p = r + beta (p - w v) and v = B p (p of the neighbours formed on
//         the fly from s, t, p, v); reduces rhat.v (alpha), ||S^-1 r||^2 (the
//         true residual of x + S y, estimated) and ||r||^2;
//   Q(j)  s = r - alpha v, t = B s (s of the neighbours on the fly);
//         reduces t.s, t.t, rhat.s, rhat.t: omega, and rho = rhat.r_new =
//         rhat.s - omega rhat.t without a third pass.
// Reliable updates (van der Vorst & Ye 2000): P's recurrence sends the solve
// to X (x += S y in FP64) and R (r = S (b - A x) from the FP64 matrix, y = 0)
// when the estimate reaches tol ||b|| times tscale (as the mixed PCG), when
// ||r|| has fallen by kMbDrop since the last update, after max_inner
// iterations, or on a breakdown.  The shadow residual and the direction are
// kept: P(j) runs again with the recomputed r and beta = (rhat.r / rho_j-1)
// (alpha / omega); its first run wrote only the other buffers of p and v.  A
// breakdown restarts the shadow residual and the direction instead (rhat = p
// = r).  KState: rr0 = rho_j, d = rho_j-1, alpha, omega, beta, conv =
// tol ||b||, last = the latest estimate, upd_est = ||r|| at the last update,
// tscale, cycle = updates, inner = FP32 iterations since the last update,
// step_i = 1 while r is fresh from an update (P has nothing to complete), ni =
// 1 to restart the direction (p = r), pad6 = 1 to restart the shadow
// residual in R, inner_done = 1 when the tolerance test sent the solve to R,
// next = the phase that follows.
constexpr double kMbDrop = 1e-4;   // FP32 residual drop between reliable updates

enum { MB_SETUP = 0, MB_P = 1, MB_Q = 2, MB_X = 3, MB_R = 4, MB_FINAL = 5 };

__host__ __device__ constexpr int mbicg_nred(int ph) {
  return ph == MB_P ? 3 : ph == MB_X ? 0 : ph == MB_FINAL ? 2 : 4;
}

struct MBicgArgs {
  Mat A;          // FP64 matrix (residual recomputation, final energy)
  Mat32 A32;      // FP32 scaled matrix
  Grid g;
  KState *st;
  const double *s;    // 1/sqrt(a_ii)
  const float *dsq;   // sqrt(a_ii)
  double *x;
  const double *b;
  float *r, *rh, *p0, *p1, *v0, *v1, *sv, *t, *y;
  double tol;
  int maxit, x0_zero, max_inner, phase;
  Defer d;
};

// p at row k for P: the residual r (fresh) or s - w t (completing), plus
// beta (p - w v) unless the direction restarts
struct XMbDir {
  const float *r, *sv, *t, *p, *v;
  float w, beta;
  int fresh, restart;
  __device__ __forceinline__ float operator()(int k) const {
    const float rk = fresh ? r[k] : sv[k] - w * t[k];
    return restart ? rk : rk + beta * (p[k] - w * v[k]);
  }
};
// s at row k for Q: r - alpha v
struct XMbRes {
  const float *r, *v;
  float alpha;
  __device__ __forceinline__ float operator()(int k) const { return r[k] - alpha * v[k]; }
};

template <int F>
__device__ __forceinline__ void mbicg_rows(const MBicgArgs &a, int ph, const KState *S, double *v) {
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  if (ph == MB_SETUP || ph == MB_R) {
    // r64 = b - A x; r = S r64 in FP32; y = 0; a restart (or the setup) sets
    // rhat = r
    const bool setup = ph == MB_SETUP, zero = setup && a.x0_zero;
    const bool newrh = setup || ld_int(&S->pad6);
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
      a.r[i] = rt;
      a.y[i] = 0.0f;
      float rh;
      if (newrh) {
        a.rh[i] = rt;
        rh = rt;
      } else {
        rh = a.rh[i];
      }
      v[0] += (double)rt * (double)rt;
      v[1] += ri * ri;
      v[2] += bi * bi;
      v[3] += (double)rh * (double)rt;
    }
    return;
  }
  if (ph == MB_X) {
    for (int i = i0; i < i1; i += is) a.x[i] = a.x[i] + a.s[i] * (double)a.y[i];
    return;
  }
  if (ph == MB_FINAL) {
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
  const float *pc = odd ? a.p0 : a.p1, *vc = odd ? a.v0 : a.v1;
  float *pn = odd ? a.p1 : a.p0, *vn = odd ? a.v1 : a.v0;
  if (ph == MB_P) {
    const XMbDir xf{a.r, a.sv, a.t, pc, vc, (float)ld_dbl(&S->omega), (float)ld_dbl(&S->beta),
                    ld_int(&S->step_i), ld_int(&S->ni)};
    const float alpha = (float)ld_dbl(&S->alpha);
    for (int i = i0; i < i1; i += is) {
      float ri;
      if (xf.fresh) {
        ri = a.r[i];
      } else {
        const float si = a.sv[i];
        ri = si - xf.w * a.t[i];
        a.r[i] = ri;
        a.y[i] = (a.y[i] + alpha * pc[i]) + xf.w * si;
      }
      const float pi = xf.restart ? ri : ri + xf.beta * (pc[i] - xf.w * vc[i]);
      const float vi = row_apply32<F>(a.A32, i, xf);
      pn[i] = pi;
      vn[i] = vi;
      const double rd = ri, ru = rd * (double)a.dsq[i];
      v[0] += (double)a.rh[i] * (double)vi;
      v[1] += ru * ru;
      v[2] += rd * rd;
    }
    return;
  }
  // MB_Q
  const XMbRes xf{a.r, vn, (float)ld_dbl(&S->alpha)};
  for (int i = i0; i < i1; i += is) {
    const float si = xf(i);
    const float ti = row_apply32<F>(a.A32, i, xf);
    a.sv[i] = si;
    a.t[i] = ti;
    const double sd = si, td = ti, rh = a.rh[i];
    v[0] += td * sd;
    v[1] += td * td;
    v[2] += rh * sd;
    v[3] += rh * td;
  }
}

__device__ __forceinline__ void mbicg_epi(const MBicgArgs &a, int ph, KState *st, const double *t) {
  if (ph == MB_SETUP || ph == MB_R) {
    const double tr = sqrt(t[1]);
    if (ph == MB_SETUP) {
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
      // the recomputed residual continues the recurrence, unless it restarts
      const double beta = (t[3] / st->d) * (st->alpha / st->omega);
      st->ni = (st->pad6 || !finite(beta)) ? 1 : 0;
      st->beta = st->ni ? 0.0 : beta;
    }
    st->rr0 = t[3];
    st->res2 = t[1];
    st->last = tr;
    st->upd_est = sqrt(t[0]);
    st->inner = 0;
    st->inner_done = 0;
    st->step_i = 1;
    st->pad6 = 0;
    st->next = MB_P;
    if (tr <= st->conv) { st->status = S_OK; st->done = 1; }
    else if (!finite(t[0]) || !finite(t[1]) || !finite(t[3])) { st->status = S_NONFINITE; st->done = 1; }
    else if (st->iter >= a.maxit) { st->status = S_MAXIT; st->done = 1; }
    else if (t[3] == 0.0) { st->status = S_BREAKDOWN; st->done = 1; }
    return;
  }
  if (ph == MB_P) {
    const bool fresh = st->step_i != 0;
    const double est = sqrt(t[1]);
    st->last = est;
    const double alpha = st->rr0 / t[0];
    const bool broke = t[0] == 0.0 || !finite(alpha);
    if (!fresh) {
      const bool tolhit = est <= st->conv * st->tscale;
      const bool drop = sqrt(t[2]) < kMbDrop * st->upd_est;
      if (tolhit || drop || broke || st->inner >= a.max_inner || st->iter >= a.maxit) {
        st->inner_done = tolhit ? 1 : 0;
        st->pad6 = broke ? 1 : 0;
        st->next = MB_X;
        return;
      }
    } else if (broke) {
      // right after an update: restart the shadow residual, or give up if it
      // was just restarted
      if (st->ni) { st->status = S_BREAKDOWN; st->done = 1; return; }
      st->pad6 = 1;
      st->next = MB_X;
      return;
    }
    st->alpha = alpha;
    st->next = MB_Q;
    return;
  }
  if (ph == MB_Q) {
    const double w = t[0] / t[1];
    const double rho = t[2] - w * t[3];
    const double beta = (rho / st->rr0) * (st->alpha / w);
    st->iter = st->iter + 1;
    st->inner = st->inner + 1;
    st->step_i = 0;
    st->ni = 0;
    if (t[1] == 0.0 || w == 0.0 || !finite(w) || !finite(beta)) {
      // the step of this iteration is dropped; restart from x + S y
      st->pad6 = 1;
      st->inner_done = 0;
      st->next = MB_X;
      return;
    }
    st->omega = w;
    st->d = st->rr0;
    st->rr0 = rho;
    st->beta = beta;
    st->next = MB_P;
    return;
  }
  if (ph == MB_X) {
    st->next = MB_R;
    return;
  }
  st->energy = t[0];
  st->res2 = t[1];
}

template <int F>
__device__ __forceinline__ bool mbicg_phase(const MBicgArgs &a, int ph, double *lds) {
  double v[kMaxK] = {};
  mbicg_rows<F>(a, ph, a.st, v);
  auto epi = [&](const double *t) { mbicg_epi(a, ph, a.st, t); };
  switch (mbicg_nred(ph)) {
    case 0: return grid_sync<0>(a.g, v, lds, epi);
    case 2: return grid_sync<2>(a.g, v, lds, epi);
    case 3: return grid_sync<3>(a.g, v, lds, epi);
    default: return grid_sync<4>(a.g, v, lds, epi);
  }
}

template <int F>
__global__ __launch_bounds__(1024) void mbicg_persistent(MBicgArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  a.g.ep = &s_ep;
  if (!mbicg_phase<F>(a, MB_SETUP, lds)) return;
  while (!ld_int(&a.st->done))
    if (!mbicg_phase<F>(a, ld_int(&a.st->next), lds)) return;
  mbicg_phase<F>(a, MB_FINAL, lds);
}

// Kernel-per-phase engine: a.phase is MB_SETUP, kPhaseNext, MB_FINAL or a
// flush; reductions deferred to the next launch (fvc_krylov.h, Defer).
template <int F>
__global__ __launch_bounds__(1024) void mbicg_kernel(MBicgArgs a) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ KState S;
  struct M {
    const MBicgArgs &a;
    enum { kRed = 4 };
    using Pre = NoPre;
    __device__ int nred(int ph) const { return mbicg_nred(ph); }
    __device__ bool skip(int ph, const KState &s) const { return ph != MB_SETUP && ph != MB_FINAL && s.done; }
    __device__ Pre prefetch(int, const KState &) const { return Pre{}; }
    __device__ void rows(int ph, const KState *s, double *v, const Pre &) const { mbicg_rows<F>(a, ph, s, v); }
    __device__ void epi(int ph, KState *s, const double *t) const { mbicg_epi(a, ph, s, t); }
  };
  deferred_phase(a.g, a.d, a.phase, M{a}, &S, lds);
}

// -------------------------------------------------------- mixed GMRES --
//
// GMRES(k) whose restart cycles are the refinement steps: a cycle starts from
// the FP64 residual r = b - A x, runs the Arnoldi process on B = S A S with
// the basis in FP32 (CGS2 orthogonalisation as gmres_cgs_phase, dot products
// accumulated in FP64), keeps the Hessenberg matrix, the Givens rotations
// and the back substitution in FP64 (gmres_close_step), and adds S sum_j y_j
// g_j to x in FP64.  Stopping is the true-residual rule of the FP64 GMRES:
// the cycle start ends the solve when ||r|| <= tol ||b||, and otherwise sets
// the cycle's target |e(i+1)| <= e(1) tol ||b|| / ||r||.
enum { MG_START = 0, MG_C1 = 1, MG_C2 = 2, MG_C3 = 3, MG_XUPD = 4, MG_FINAL = 5 };

struct MGmresArgs {
  GmresArgs q;        // h, cs, sn, e, y, nrm, k, maxit, tol, st, g (gv unused)
  Mat32 A32;
  const double *s;    // 1/sqrt(a_ii)
  float *gv;          // basis g(neq, k+1) in FP32, column-major
  float *gu, *gw;     // lagged CGS2 (q.orth 2): the candidate u and w = B u in FP32
};

template <int F>
__device__ __forceinline__ bool mgmres_phase(const MGmresArgs &m, int ph, int c, double *lds) {
  const GmresArgs &a = m.q;
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  KState *st = a.st;
  const int k = a.k;
  auto col = [&](int cc) { return m.gv + (size_t)(cc - 1) * a.ld; };
  auto H = [&](int r, int cc) -> double & { return a.h[(size_t)(cc - 1) * (k + 1) + (r - 1)]; };
  if (ph == MG_START) {
    const int first = ld_int(&st->cycle) == 0;
    float *g1 = a.orth == 2 ? m.gu : col(1);   // lagged CGS2: the first candidate
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
      const float gi = (float)(ri * m.s[i]);
      g1[i] = gi;
      v[0] += (double)gi * (double)gi;
      v[2] += ri * ri;
    }
    return grid_sync<3>(a.g, v, lds, [&](const double *t) {
      if (first) {
        st->bnorm2 = t[1];
        st->conv = a.tol * sqrt(t[1]);
        st->inner = 0;
        st->status = S_MAXIT;
        st->done = 0;
      }
      const double e1 = sqrt(t[0]), rt = sqrt(t[2]);
      st->rnorm2 = t[2];
      st->last = rt;
      st->iter = st->cycle;
      const bool stop = rt <= st->conv || !finite(rt) || !finite(e1) || e1 == 0.0 || st->cycle >= a.maxit;
      if (rt <= st->conv) st->status = S_OK;
      else if (!finite(rt) || !finite(e1)) st->status = S_NONFINITE;
      else if (e1 == 0.0) st->status = S_BREAKDOWN;
      if (stop) {
        st->done = 1;
        st->inner_done = 1;
        return;
      }
      st->econv = e1 * (st->conv / rt);
      st->cycle = st->cycle + 1;
      a.e[1] = e1;
      a.nrm[1] = e1;
      st->step_i = 1;
      st->mgs_j = 1;
      st->inner_done = 0;
      st->ni = 0;
    });
  }
  if (ph == MG_XUPD) {
    // x += S sum_j y(j) g(j), j in order, in FP64
    const int nI = ld_int(&st->ni);
    __shared__ double ysh[kMaxGmresLag + 2];   // k <= kCgsWide (CGS2) or kMaxGmresLag (lagged)
    for (int j = (int)threadIdx.x + 1; j <= nI; j += blockDim.x) ysh[j] = ld_dbl(&a.y[j]);
    __syncthreads();
    double v[1] = {0.0};
    for (int r = i0; r < i1; r += is) {
      double d = 0.0;
      for (int j = 1; j <= nI; ++j) d = d + ysh[j] * (double)m.gv[(size_t)(j - 1) * a.ld + r];
      a.x[r] = a.x[r] + m.s[r] * d;
    }
    return grid_sync<0>(a.g, v, lds, [&](const double *) {
      const double en = a.e[nI + 1];
      if (!finite(en)) { st->status = S_NONFINITE; st->done = 1; }
      st->inner_done = 0;
    });
  }
  if (ph == MG_FINAL) {
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
  // Arnoldi step i, CGS2 as gmres_cgs_phase: C1 (SpMV + dots), C2 (normalise
  // column i, project, dots), C3 (project, norm, Givens); dots kCgsChunk at a time
  const int i = ld_int(&st->step_i);
  const double ni = ld_dbl(&a.nrm[i]);
  float *gi = col(i);
  float *gn = col(i + 1);
  constexpr int kLoads = 8;
  auto load_group = [&](double *gq, int p0, int r) {
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      const int cc = min(p0 + q + 1, i);
      gq[q] = (double)m.gv[(size_t)(cc - 1) * a.ld + r];
    }
  };
  const int q0 = c * kCgsChunk;
  auto dots = [&](double *v, double w, int r) {
#pragma unroll
    for (int u = 0; u < kCgsChunk; u += kLoads) {
      if (q0 + u >= i) break;
      double gq[kLoads];
      load_group(gq, q0 + u, r);
#pragma unroll
      for (int q = 0; q < kLoads; ++q) {
        const int cc = q0 + u + q + 1;
        const double g = (ph == MG_C1 && cc == i) ? gq[q] / ni : gq[q];
        if (cc <= i) v[u + q] += w * g;
      }
    }
  };
  auto project = [&](double w, const double *h, double gir, int r) {
    for (int p0 = 0; p0 < i; p0 += kLoads) {
      double gq[kLoads];
      load_group(gq, p0, r);
#pragma unroll
      for (int q = 0; q < kLoads; ++q) {
        const int cc = p0 + q + 1;
        if (cc <= i) w = w - h[cc - 1] * (cc == i ? gir : gq[q]);
      }
    }
    return w;
  };
  double v[kCgsChunk];
#pragma unroll
  for (int q = 0; q < kCgsChunk; ++q) v[q] = 0.0;
  if (ph == MG_C1 || (ph == MG_C2 && c > 0)) {
    if (ph == MG_C1 && c == 0) {
      // w = B (g_i / nrm_i), the division in FP32 as the basis is stored
      const float nf = (float)ni;
      struct XScaled32 {
        const float *g;
        float n;
        __device__ __forceinline__ float operator()(int j) const { return g[j] / n; }
      } xf{gi, nf};
      for (int r = i0; r < i1; r += is) {
        const float w = row_apply32<F>(m.A32, r, xf);
        gn[r] = w;
        dots(v, (double)w, r);
      }
    } else {
      for (int r = i0; r < i1; r += is) dots(v, (double)gn[r], r);
    }
    return grid_sync<kCgsChunk>(a.g, v, lds, [&](const double *t) {
#pragma unroll
      for (int q = 0; q < kCgsChunk; ++q) {
        if (q0 + q < i) {
          if (ph == MG_C1) H(q0 + q + 1, i) = t[q];
          else a.y[q0 + q + 1] = t[q];
        }
      }
    });
  }
  __shared__ double h[kCgsWide];
  {
    const double *src = ph == MG_C2 ? a.h + (size_t)(i - 1) * (k + 1) : a.y + 1;
    for (int q = (int)threadIdx.x; q < i; q += (int)blockDim.x) h[q] = ld_dbl(src + q);
    __syncthreads();
  }
  if (ph == MG_C2) {
    for (int r = i0; r < i1; r += is) {
      const float gir = (float)((double)gi[r] / ni);
      gi[r] = gir;
      const float w = (float)project((double)gn[r], h, (double)gir, r);
      gn[r] = w;
      dots(v, (double)w, r);
    }
    return grid_sync<kCgsChunk>(a.g, v, lds, [&](const double *t) {
#pragma unroll
      for (int q = 0; q < kCgsChunk; ++q)
        if (q < i) a.y[q + 1] = t[q];
    });
  }
  // MG_C3
  double nv[1] = {0.0};
  for (int r = i0; r < i1; r += is) {
    const float w = (float)project((double)gn[r], h, (double)gi[r], r);
    gn[r] = w;
    nv[0] += (double)w * (double)w;
  }
  return grid_sync<1>(a.g, nv, lds, [&](const double *t) {
    for (int q = 1; q <= i; ++q) H(q, i) = H(q, i) + a.y[q];
    const double norm = sqrt(t[0]);
    H(i + 1, i) = norm;
    a.nrm[i + 1] = norm;
    gmres_close_step(a, st, i);
  });
}

// The cycle start, x update and final phases of the mixed GMRES with lagged
// CGS2 (fvc_gmres_lag.h), one launch each (their reductions run in the last
// block to arrive, grid_reduce).
template <int F>
__global__ __launch_bounds__(1024) void mgmres_kernel(MGmresArgs m) {
  __shared__ double lds[kCgsChunk * kMaxWaves];
  const int ph = m.q.phase;
  if (ph != MG_FINAL && ld_int(&m.q.st->done)) return;
  mgmres_phase<F>(m, ph, 0, lds);
}

// One persistent launch per solve, shaped as the FP64 CGS2 GMRES.
template <int F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void mgmres_persistent(MGmresArgs m) {
  __shared__ double lds[kCgsChunk * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  m.q.g.ep = &s_ep;
  const GmresArgs &a = m.q;
  while (!ld_int(&a.st->done)) {
    if (!mgmres_phase<F>(m, MG_START, 0, lds)) return;
    if (ld_int(&a.st->done)) break;
    for (int i = 1; i <= a.k; ++i) {
      const int nch = (i + kCgsChunk - 1) / kCgsChunk;
      for (int c = 0; c < nch; ++c)
        if (!mgmres_phase<F>(m, MG_C1, c, lds)) return;
      for (int c = 0; c < nch; ++c)
        if (!mgmres_phase<F>(m, MG_C2, c, lds)) return;
      if (!mgmres_phase<F>(m, MG_C3, 0, lds)) return;
      if (ld_int(&a.st->inner_done)) break;
    }
    if (!mgmres_phase<F>(m, MG_XUPD, 0, lds)) return;
  }
  mgmres_phase<F>(m, MG_FINAL, 0, lds);
}

}  // namespace fvc
