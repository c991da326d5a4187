// fvc_batch.h -- many independent systems in one persistent launch.
//
// The systems are concatenated: system s owns rows [row_off[s], row_off[s+1])
// of the concatenated ELLPACK arrays, with its own 1-based local column
// indices (each system in the reference's layout, back to back).  Every
// system gets its own group of consecutive blocks (sized by the host from a
// work estimate), its own barrier counters, reduction partials and solver
// state, and runs the unchanged single-system phase functions on pointers
// offset to its rows.  Groups never wait on each other.  A system too small
// to keep a block busy for the batch's duration shares its block with other
// small systems: the block solves them one after another, each as a group of
// one block.  All blocks are co-resident (the grid fits the device at once).
#pragma once
#include "fvc_mixed.h"

namespace fvc {

struct SysDesc {
  int row_off;   // first row of the system in the concatenated arrays
  int neq;
  int g_first;   // first block of its group
  int g_count;   // blocks in its group
  int val_off;   // CSRD: first stored value of the system (ELLPACK: w row_off)
  int ia_off;    // CSRD: first entry of its row pointer (row_off + system index)
};

struct BatchArgs {
  int w;                    // ELLPACK width (CSRD: 0)
  const double *ad;         // concatenated ad (Jacobi diagonal as well)
  const double *av;         // concatenated a(w, .), or the CSRD values
  const int *ia;            // CSRD: concatenated row pointers, each local and 1-based
  const int *ja;            // concatenated ja, local 1-based indices
  const SysDesc *desc;
  const int *block_ptr;     // block b runs systems block_list[block_ptr[b] ..
  const int *block_list;    //   block_ptr[b + 1]) in turn
  KState *st;               // per system
  unsigned *ctr;            // kCtrWords per system
  unsigned *flags;          // one arrival flag per block
  int flagbar;
  double *partials;         // kMaxK per block
  double *x;
  const double *b;
  double *v[5];                     // FP64 work vectors (concatenated)
  float4 *rec;                      // mixed: {r~, q, p, y} records, two buffers,
  size_t fs;                        // the second at +fs
  float *av32;                      // FP32 scaled values (mixed)
  short *jo;                        // mixed, ELLPACK widths 3 and 4: local column
                                    // offsets j - i, or null (then ja)
  int *wide;                        // set when an offset does not fit int16
  double *s;                        // 1/sqrt(a_ii) (mixed)
  float *dsq;                       // sqrt(a_ii) (mixed)
  double tol;
  double delta;                     // mixed: residual drop that triggers an FP64 update
  int maxit, x0_zero, stop_rule, mixed;
  int max_inner;                    // mixed: FP32 iterations between updates, at most
};

__device__ __forceinline__ void batch_grid(const BatchArgs &B, int sys, const SysDesc &D, Grid &g) {
  g.G = D.g_count;
  g.tiles = 0;
  g.R = grid_rows(D.neq, D.g_count, 0);
  g.neq = D.neq;
  g.ctr = B.ctr + (size_t)sys * kCtrWords;
  g.partials = B.partials + (size_t)kMaxK * D.g_first;
  g.persistent = true;
  g.base = D.g_first;
  g.flagbar = B.flagbar;
  g.flags = B.flags + D.g_first;
}

template <int F>
__device__ __forceinline__ void batch_mat(const BatchArgs &B, const SysDesc &D, Mat &A) {
  const size_t off = (size_t)D.row_off;
  A.neq = D.neq;
  A.w = B.w;
  A.ad = B.ad + off;
  A.av = B.av + D.val_off;
  A.au = nullptr;
  A.ia = F == F_CSRD ? B.ia + D.ia_off : nullptr;
  A.ja = B.ja + D.val_off;
  A.tptr = A.tsrc = A.tpos = nullptr;
}

// FP64: Jacobi-PCG in the reference's two-reduction form (pcg_phase).  With
// a few blocks per system a group barrier is cheap, and the one-reduction
// form's extra neighbour gathers cost more than its saved barrier.
template <int F>
__global__ __launch_bounds__(1024) void pcg_batch_persistent(BatchArgs B) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  // s_ep keeps counting across the systems of a block: its arrival flag only
  // grows, while each system's counters start from zero
  for (int k = B.block_ptr[blockIdx.x]; k < B.block_ptr[blockIdx.x + 1]; ++k) {
    const int sys = B.block_list[k];
    const SysDesc D = B.desc[sys];
    const size_t off = (size_t)D.row_off;
    PcgArgs a{};
    batch_mat<F>(B, D, a.A);
    batch_grid(B, sys, D, a.g);
    a.g.ep = &s_ep;
    a.st = B.st + sys;
    a.m = a.A.ad;
    a.x = B.x + off;
    a.b = B.b + off;
    a.r = B.v[0] + off;
    a.z = B.v[1] + off;
    a.p0 = B.v[2] + off;
    a.p1 = B.v[3] + off;
    a.q = B.v[4] + off;
    a.tol = B.tol;
    a.maxit = B.maxit;
    a.x0_zero = B.x0_zero;
    a.stop_rule = B.stop_rule;
    if (!pcg_phase<F>(a, PCG_SETUP, lds)) return;
    bool ok = true;
    while (ok && !ld_int(&a.st->done)) {
      if (!(ok = pcg_phase<F>(a, PCG_A, lds))) break;
      if (ld_int(&a.st->done)) break;
      ok = pcg_phase<F>(a, PCG_B, lds);
    }
    if (!ok || !pcg_phase<F>(a, PCG_FINAL, lds)) return;
    __syncthreads();   // every thread is past this system before LDS is reused
  }
}

// Mixed precision (fvc_mixed.h) per system.
template <int F>
__global__ __launch_bounds__(1024) void mixed_batch_persistent(BatchArgs B) {
  __shared__ double lds[kMaxK * kMaxWaves];
  __shared__ unsigned s_ep;
  if (threadIdx.x == 0) s_ep = 0u;
  __syncthreads();
  for (int k = B.block_ptr[blockIdx.x]; k < B.block_ptr[blockIdx.x + 1]; ++k) {
    const int sys = B.block_list[k];
    const SysDesc D = B.desc[sys];
    const size_t off = (size_t)D.row_off;
    MixedArgs a{};
    batch_mat<F>(B, D, a.A);
    a.A32.neq = D.neq;
    a.A32.w = B.w;
    a.A32.av = B.av32 + D.val_off;
    a.A32.ia = a.A.ia;
    a.A32.ja = a.A.ja;
    a.A32.jo = B.jo ? B.jo + off * B.w : nullptr;
    batch_grid(B, sys, D, a.g);
    a.g.ep = &s_ep;
    a.st = B.st + sys;
    a.m = a.A.ad;
    a.s = B.s + off;
    a.dsq = B.dsq + off;
    a.x = B.x + off;
    a.b = B.b + off;
    a.rec = B.rec + off;
    a.ps = B.fs;
    a.tol = B.tol;
    a.delta = B.delta;
    a.maxit = B.maxit;
    a.x0_zero = B.x0_zero;
    a.max_inner = B.max_inner;
    if (!mixed_run<F>(a, lds)) return;
    __syncthreads();
  }
}

// S A S in FP32 for every system of the batch (rows are independent; the
// column indices are local to each system), and with B.jo the int16 column
// offsets of col_offsets_kernel (local row and column, the same rules).
// val_off: CSRD, the first value of each system (nsys entries).
template <int F>
__global__ void batch_scale_kernel(BatchArgs B, int nsys, const int *row_off, const int *val_off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  // system of row i: binary search in row_off[0..nsys]
  int lo = 0, hi = nsys;
  const int total = row_off[nsys];
  if (i >= total) return;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (row_off[mid] <= i) lo = mid; else hi = mid;
  }
  const size_t off = (size_t)row_off[lo];
  const double si = 1.0 / sqrt(B.ad[i]);
  B.s[i] = si;
  B.dsq[i] = (float)sqrt(B.ad[i]);
  if constexpr (F == F_CSRD) {
    // the system's row pointer starts at row_off + lo; positions are local
    const int *ia = B.ia + off + lo;
    const int il = i - (int)off;
    const size_t v0 = (size_t)val_off[lo];
    for (int k = ia[il] - 1; k < ia[il + 1] - 1; ++k) {
      const size_t e = v0 + k;
      const int j = B.ja[e] - 1;
      B.av32[e] = (float)(B.av[e] * si * (1.0 / sqrt(B.ad[off + j])));
    }
    return;
  }
  const int w = (F == F_ELL4) ? 4 : (F == F_ELL3) ? 3 : B.w;
  bool fits = true;
  for (int k = 0; k < w; ++k) {
    const size_t e = (size_t)w * i + k;
    const int j = B.ja[e] - 1;
    const float v = (float)(B.av[e] * si * (1.0 / sqrt(B.ad[off + j])));
    B.av32[e] = v;
    if constexpr (F == F_ELL4 || F == F_ELL3) {
      if (B.jo) {
        const int d = j - (i - (int)off);
        short o = 0;
        if (v != 0.0f) {
          if (d >= -32768 && d <= 32767) o = (short)d;
          else fits = false;
        }
        B.jo[e] = o;
      }
    }
  }
  if (!fits) *B.wide = 1;
}

}  // namespace fvc
