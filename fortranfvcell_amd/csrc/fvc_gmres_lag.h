// fvc_gmres_lag.h -- GMRES(k) with one grid reduction per Arnoldi step:
// classical Gram-Schmidt with one re-orthogonalisation whose second pass and
// normalisation are lagged by one step (the "DCGS2" Arnoldi of Swirydowicz,
// Langou, Ananthan, Yang & Thomas, Numer. Linear Algebra Appl. 2021, and
// Bielich et al., Parallel Computing 2022).  The restart loop, the Givens
// rotations, the convergence test and the back substitution are those of the
// reference's gmres (src/Iterativos.f:271-446, Givens :363-375, back
// substitution :382-389), shared with the MGS and CGS2 variants
// (gmres_close_step); only the orthogonalisation differs.
//
// Step j holds V_{j-1} = [v_1 .. v_{j-1}] (orthonormal, final) and the
// candidate u_j, already projected once against V_{j-1}.  Three launches:
//   D_j  w = M^-1 A u_j and, in the same sweep, a = V_{j-1}' u_j, beta = u_j'u_j,
//        b = V_{j-1}' w, gamma = u_j'w: 2j dot products, written per block;
//   R_j  one block sums the partials in a fixed order and runs the scalar
//        part: nu = sqrt(beta - a'a) (u_j - V a is orthogonal to V), the
//        finished column j-1 of H (first pass + a, and nu below it), its
//        Givens rotation and the convergence test; then, from the Arnoldi
//        relation M^-1 A V_{j-1} = V_j H(1:j, 1:j-1), the first-pass
//        coefficients of column j, h = (V_j'w - H a) / nu with
//        V_j'w = [b; (gamma - a'b) / nu], and c = H a / nu + h;
//   B_j  v_j = (u_j - V_{j-1} a) / nu into column j, and the next candidate
//        u_{j+1} = w / nu - V_j c, in place.
// Per step the basis is read twice (D and B) instead of 2j+ times (MGS), and
// the only grid-wide dependences are the kernel boundaries.  Step k+1 only
// re-orthogonalises u_{k+1} (column k), without its SpMV.
//
// Two storages share these kernels: FP64 (GmresArgs: left-Jacobi operator
// M^-1 A, basis in FP64) and mixed precision (MGmresArgs: the Jacobi-scaled
// operator S A S in FP32, basis, u and w in FP32, every dot product and the
// scalar part in FP64), each an access policy below.
#pragma once
#include "fvc_mixed_krylov.h"

namespace fvc {

constexpr int kLagChunk = 8;      // basis columns loaded together by B
constexpr int kLagBlock = 256;    // threads per block of D and B
constexpr int kLagTile = 1024;    // rows of u and w a D block stages in LDS at a time
constexpr int kLagRows = 8;       // rows per lane of a basis column D loads together
constexpr int kLagPerWave = (kMaxGmresLag + kLagBlock / kWave - 1) / (kLagBlock / kWave);
// arrival counter of the R blocks: a word of the grid counters that the
// kernel-per-phase reductions leave unused (zeroed with them per solve)
constexpr int kCtrLag = kCtrErr + 16;

// Value q of step j's reduction: a_1..a_{j-1} at 0..j-2, beta at j-1,
// b_1..b_{j-1} at j..2j-2, gamma at 2j-1.  Partials value-major:
// part[q * G + block]; totals tot[q].
__host__ __device__ inline int lag_kp(int k) { return 2 * (k + 1); }

// FP64 storage: w = M^-1 A u (left Jacobi, src/Iterativos.f:333-336).
template <int F>
struct LagF64 {
  const GmresArgs &a;
  __device__ const GmresArgs &q() const { return a; }
  __device__ double u(int r) const { return a.gu[r]; }
  __device__ void set_u(int r, double v) const { a.gu[r] = v; }
  __device__ double w(int r) const { return a.gw[r]; }
  __device__ double set_w(int r, double v) const { a.gw[r] = v; return v; }   // returns the stored value
  __device__ double op(int r) const { return row_apply<F>(a.A, r, XVec{a.gu}) / a.m[r]; }
  __device__ double col(int c, int r) const { return a.gv[(size_t)c * a.ld + r]; }
  __device__ void set_col(int c, int r, double v) const { a.gv[(size_t)c * a.ld + r] = v; }
};
// Mixed storage: w = S A S u in FP32 (the scaled FP32 matrix), FP32 vectors.
template <int F>
struct LagF32 {
  const MGmresArgs &m;
  struct XU {
    const float *g;
    __device__ __forceinline__ float operator()(int j) const { return g[j]; }
  };
  __device__ const GmresArgs &q() const { return m.q; }
  __device__ double u(int r) const { return (double)m.gu[r]; }
  __device__ void set_u(int r, double v) const { m.gu[r] = (float)v; }
  __device__ double w(int r) const { return (double)m.gw[r]; }
  __device__ double set_w(int r, double v) const {   // returns the stored value
    const float f = (float)v;
    m.gw[r] = f;
    return (double)f;
  }
  __device__ double op(int r) const { return (double)row_apply32<F>(m.A32, r, XU{m.gu}); }
  __device__ double col(int c, int r) const { return (double)m.gv[(size_t)c * m.q.ld + r]; }
  __device__ void set_col(int c, int r, double v) const { m.gv[(size_t)c * m.q.ld + r] = (float)v; }
};

// D_j.  A block owns a contiguous range of rows (Grid chunks) and takes it
// kLagTile rows at a time: its threads form w = A u / m for the tile (u of
// the neighbours gathered), store w, and stage u and w in LDS; then wave v
// streams basis columns v, v + 4, ... over the tile, multiplying with u and
// w from LDS, so that each basis element is read once and u, w once.  Each
// column's sum over a tile is reduced over the wave and added to the wave's
// running total in LDS, in tile order: a fixed order.
template <class P>
__device__ __forceinline__ void gmres_lag_d(const P &pol) {
  constexpr int nwv = kLagBlock / kWave;
  __shared__ double us[kLagTile], ws[kLagTile];
  __shared__ double lds[2 * kMaxWaves];
  __shared__ double cw[nwv][2 * kLagPerWave];   // each wave's running column totals
  const GmresArgs &a = pol.q();
  const KState *st = a.st;
  if (ld_int(&st->done) || ld_int(&st->inner_done)) return;
  const int j = ld_int(&st->step_i);
  const int nv = j - 1, G = a.g.G;
  const bool last = j == a.k + 1;   // column k's re-orthogonalisation only
  const int blk = a.g.chunk();
  const int r0 = blk * a.g.R;
  const int r1 = min(r0 + a.g.R, a.g.neq);
  const int wid = (int)threadIdx.x / kWave, lane = (int)threadIdx.x % kWave;
  if (lane == 0)
    for (int p = 0; p < 2 * kLagPerWave; ++p) cw[wid][p] = 0.0;
  double uu = 0.0, uw = 0.0;
  for (int t0 = r0; t0 < r1; t0 += kLagTile) {
    const int tn = min(kLagTile, r1 - t0);
    for (int q = threadIdx.x; q < tn; q += kLagBlock) {
      const int r = t0 + q;
      const double ur = pol.u(r);
      double wr = 0.0;
      if (!last) wr = pol.set_w(r, pol.op(r));
      us[q] = ur;
      ws[q] = wr;
      uu += ur * ur;
      uw += ur * wr;
    }
    __syncthreads();
    // wave wid: columns wid, wid + nwv, ... over the tile (lane-strided
    // rows), reduced over the wave and added to the wave's running totals
    for (int c = wid; c < nv; c += nwv) {
      double sa = 0.0, sb = 0.0;
      int q = lane;
      // kLagRows rows per lane loaded together, then the rest one at a time
      for (; q + (kLagRows - 1) * kWave < tn; q += kLagRows * kWave) {
        double g[kLagRows];
#pragma unroll
        for (int u = 0; u < kLagRows; ++u) g[u] = pol.col(c, t0 + q + u * kWave);
#pragma unroll
        for (int u = 0; u < kLagRows; ++u) {
          sa += g[u] * us[q + u * kWave];
          sb += g[u] * ws[q + u * kWave];
        }
      }
      for (; q < tn; q += kWave) {
        const double g = pol.col(c, t0 + q);
        sa += g * us[q];
        sb += g * ws[q];
      }
      sa = wave_sum(sa);
      sb = wave_sum(sb);
      if (lane == 0) {
        const int p = c / nwv;
        cw[wid][2 * p] += sa;
        cw[wid][2 * p + 1] += sb;
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    for (int c = wid; c < nv; c += nwv) {
      const int p = c / nwv;
      a.lag_part[(size_t)c * G + blk] = cw[wid][2 * p];
      a.lag_part[(size_t)(j + c) * G + blk] = cw[wid][2 * p + 1];
    }
  }
  double v[2] = {uu, uw};
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    a.lag_part[(size_t)nv * G + blk] = v[0];
    a.lag_part[(size_t)(2 * j - 1) * G + blk] = v[1];
  }
}

template <int F>
__global__ __launch_bounds__(kLagBlock) void gmres_lag_d_kernel(GmresArgs a) {
  gmres_lag_d(LagF64<F>{a});
}
template <int F>
__global__ __launch_bounds__(kLagBlock) void mgmres_lag_d_kernel(MGmresArgs m) {
  gmres_lag_d(LagF32<F>{m});
}

// R_j, launched with 2(k+1) blocks: block q < 2j sums value q over the G
// blocks of D (thread t takes blocks t, t + 256, ... in order, then the
// block's fixed tree) into tot[q]; the last block to arrive then runs the
// scalar part of step j.  Everything the scalar part reads is first loaded
// into LDS by all the block's threads together, so the one thread that runs
// the sequential part (the Givens rotations, the back substitution) waits on
// no global memory.  (Templated on the format like D and B only so that
// every family of solver kernels is instantiated in one translation unit.)
template <int F>
__global__ __launch_bounds__(kLagBlock) void gmres_lag_r_kernel(GmresArgs a) {
  constexpr int L = kMaxGmresLag + 1;
  __shared__ double lds[2 * kMaxWaves];
  __shared__ double tot[2 * L];
  __shared__ double hr[L * L];          // H before the rotations, rows and columns 1..j
  __shared__ double hc[L + 1];          // H(1:j, 1:j-1) a
  __shared__ double cs[L + 1], sn[L + 1], ee[L + 2];
  __shared__ double col[L + 1];         // column j-1 under rotation; y at the end
  __shared__ int s_last, s_ni, s_ended;
  __shared__ double s_aa, s_ab;
  KState *st = a.st;
  if (ld_int(&st->done) || ld_int(&st->inner_done)) return;
  const int j = ld_int(&st->step_i);
  const int nv = j - 1, K = 2 * j, k = a.k;
  const int G = a.g.G;
  const int qv = (int)blockIdx.x;
  if (qv >= K) return;
  {
    const double *p = a.lag_part + (size_t)qv * G;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < G; b += blockDim.x) v[0] += p[b];
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) {
      a.lag_tot[qv] = v[0];
      fence_release();
      const unsigned t = add_relaxed(a.g.ctr + kCtrLag, 1u);
      s_last = (int)t + 1 == K;
      if (s_last) {
        st_relaxed(a.g.ctr + kCtrLag, 0u);
        fence_acquire();
      }
    }
    __syncthreads();
    if (!s_last) return;
  }
  auto HR = [&](int r, int c) -> double & { return hr[(c - 1) * L + (r - 1)]; };
  auto Hg = [&](double *h, int r, int c) -> double & { return h[(size_t)(c - 1) * (k + 1) + (r - 1)]; };
  // the small state, loaded by all threads at once
  for (int t = threadIdx.x; t < K; t += blockDim.x) tot[t] = a.lag_tot[t];
  for (int t = threadIdx.x; t < j * nv; t += blockDim.x) {
    const int r = t % j + 1, c = t / j + 1;
    HR(r, c) = r <= c + 1 ? Hg(a.hraw, r, c) : 0.0;
  }
  for (int t = threadIdx.x; t <= j; t += blockDim.x) {
    cs[t] = a.cs[t];
    sn[t] = a.sn[t];
    ee[t] = a.e[t];
  }
  __syncthreads();
  // H(1:j, 1:j-1) a with column j-1 as it stands (its correction follows),
  // one row per thread; a'a and a'b on the first wave
  for (int r = threadIdx.x + 1; r <= j; r += blockDim.x) {
    double sh = 0.0;
    for (int c = r > 1 ? r - 1 : 1; c <= nv; ++c) sh += HR(r, c) * tot[c - 1];   // Hessenberg: c >= r - 1
    hc[r - 1] = sh;
  }
  if (threadIdx.x < kWave) {
    double pa = 0.0, pb = 0.0;
    for (int q = (int)threadIdx.x; q < nv; q += kWave) {
      pa += tot[q] * tot[q];
      pb += tot[q] * tot[j + q];
    }
    pa = wave_sum(pa);
    pb = wave_sum(pb);
    if (threadIdx.x == 0) {
      s_aa = pa;
      s_ab = pb;
    }
  }
  __syncthreads();
  const double beta = tot[nv], gamma = tot[2 * j - 1];
  const double nu2 = beta - s_aa;
  const double nu = nu2 > 0.0 ? sqrt(nu2) : 0.0;
  if (threadIdx.x == 0) s_ni = 0;
  if (j == 1) {
    if (threadIdx.x == 0) {
      a.e[1] = nu;   // ||r0 / m||: the least-squares right-hand side
      a.nrm[1] = nu;
      s_ended = 0;
    }
  } else {
    // column i = j-1 is finished: first pass + re-orthogonalisation, nu below;
    // (H a)_r gains a_{j-1} a_r from the correction
    const int i = nv;
    for (int r = threadIdx.x + 1; r <= i; r += blockDim.x) {
      HR(r, i) = HR(r, i) + tot[r - 1];
      hc[r - 1] += tot[r - 1] * tot[i - 1];
    }
    if (threadIdx.x == 0) {
      HR(j, i) = nu;
      hc[i] = nu * tot[i - 1];
    }
    __syncthreads();
    for (int r = threadIdx.x + 1; r <= j; r += blockDim.x) {
      Hg(a.hraw, r, i) = HR(r, i);
      col[r] = HR(r, i);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      // the Givens rotations of src/Iterativos.f:363-375, as gmres_close_step
      for (int jj = 1; jj <= i - 1; ++jj) {
        const double aux1 = cs[jj] * col[jj] + sn[jj] * col[jj + 1];
        const double aux2 = -sn[jj] * col[jj] + cs[jj] * col[jj + 1];
        col[jj] = aux1;
        col[jj + 1] = aux2;
      }
      const double rr = sqrt(col[i] * col[i] + col[i + 1] * col[i + 1]);
      cs[i] = col[i] / rr;
      sn[i] = col[i + 1] / rr;
      col[i] = rr;
      col[i + 1] = 0.0;
      ee[i + 1] = -sn[i] * ee[i];
      ee[i] = cs[i] * ee[i];
      a.cs[i] = cs[i];
      a.sn[i] = sn[i];
      a.e[i] = ee[i];
      a.e[i + 1] = ee[i + 1];
      st->inner = st->inner + 1;
      const bool conv = fabs(ee[i + 1]) <= st->econv;   // src/Iterativos.f:376
      s_ended = conv || i == k || !finite(ee[i + 1]);
      if (s_ended) {
        st->ni = i;
        st->inner_done = 1;
        s_ni = i;
      }
    }
    __syncthreads();
    for (int r = threadIdx.x + 1; r <= j; r += blockDim.x) Hg(a.h, r, i) = col[r];
  }
  __syncthreads();
  if (!s_ended) {
    if (!(nu > 0.0) || !finite(nu)) {
      // u_j lies in span(V_{j-1}) (or is not finite): no column j
      if (threadIdx.x == 0) {
        st->ni = nv;
        st->inner_done = 1;
        s_ni = nv;
        if (!finite(nu)) {
          st->status = S_NONFINITE;
          st->done = 1;
          s_ni = 0;
        }
      }
    } else {
      // column j, first pass: h = (V_j' w - H a) / nu; next candidate's c
      double *la = a.lag;   // la[0] = nu, la[1..j-1] = a, la[k+1..k+j] = c
      for (int r = threadIdx.x + 1; r <= j; r += blockDim.x) {
        const double vtw = r < j ? tot[j + r - 1] : (gamma - s_ab) / nu;
        const double h1 = (vtw - hc[r - 1]) / nu;
        Hg(a.hraw, r, j) = h1;
        la[k + r] = hc[r - 1] / nu + h1;
        if (r <= nv) la[r] = tot[r - 1];
      }
      if (threadIdx.x == 0) {
        la[0] = nu;
        a.nrm[j] = nu;
        st->step_i = j + 1;
      }
    }
  }
  __syncthreads();
  const int ni = s_ni;
  if (ni == 0) return;
  // the cycle ended at ni steps: back substitution h y = e
  // (src/Iterativos.f:382-389) on the rotated H, loaded into LDS first
  for (int t = threadIdx.x; t < ni * ni; t += blockDim.x) {
    const int r = t % ni + 1, c = t / ni + 1;
    HR(r, c) = r <= c ? Hg(a.h, r, c) : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *y = col;
    y[ni] = ee[ni] / HR(ni, ni);
    for (int ii = ni - 1; ii >= 1; --ii) {
      double yi = 0.0;
      for (int jj = ii + 1; jj <= ni; ++jj) yi = yi - HR(ii, jj) * y[jj];
      y[ii] = (yi + ee[ii]) / HR(ii, ii);
    }
    for (int ii = 1; ii <= ni; ++ii) a.y[ii] = y[ii];
  }
}

// v_j = (u_j - V_{j-1} a) / nu into column j; u <- w / nu - V_j c.
template <class P>
__device__ __forceinline__ void gmres_lag_b(const P &pol) {
  __shared__ double ca[kMaxGmresLag + 1], cc[kMaxGmresLag + 2];
  const GmresArgs &a = pol.q();
  const KState *st = a.st;
  if (ld_int(&st->done) || ld_int(&st->inner_done)) return;
  const int j = ld_int(&st->step_i) - 1;
  const int nv = j - 1, k = a.k;
  for (int q = threadIdx.x; q <= k + 1; q += blockDim.x) {
    if (q <= nv) ca[q] = ld_dbl(a.lag + q);        // ca[0] = nu
    if (q >= 1 && q <= j) cc[q] = ld_dbl(a.lag + k + q);
  }
  __syncthreads();
  const double nu = ca[0];
  const int i0 = a.g.first(), i1 = a.g.row_end(), is = a.g.stride();
  for (int r = i0; r < i1; r += is) {
    double s1 = pol.u(r), s2 = 0.0;
    for (int c0 = 0; c0 < nv; c0 += kLagChunk) {
      double g[kLagChunk];
#pragma unroll
      for (int q = 0; q < kLagChunk; ++q) g[q] = pol.col(min(c0 + q, nv - 1), r);
#pragma unroll
      for (int q = 0; q < kLagChunk; ++q) {
        if (c0 + q < nv) {
          s1 = s1 - ca[c0 + q + 1] * g[q];
          s2 = s2 + cc[c0 + q + 1] * g[q];
        }
      }
    }
    const double v = s1 / nu;
    pol.set_col(nv, r, v);
    pol.set_u(r, pol.w(r) / nu - (s2 + cc[j] * v));
  }
}

template <int F>
__global__ __launch_bounds__(kLagBlock) void gmres_lag_b_kernel(GmresArgs a) {
  gmres_lag_b(LagF64<F>{a});
}
template <int F>
__global__ __launch_bounds__(kLagBlock) void mgmres_lag_b_kernel(MGmresArgs m) {
  gmres_lag_b(LagF32<F>{m});
}

}  // namespace fvc
