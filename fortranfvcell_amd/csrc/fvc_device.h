// fvc_device.h -- device-side building blocks of the gfx950 solver library:
//   * the matrix view and per-row operators for every reference storage
//     format, each summing in the reference's sequential order;
//   * deterministic block and grid reductions (fixed trees, no FP atomics);
//   * a two-level grid barrier fused with the reduction epilogue, used by
//     the persistent (one launch per solve) engine, whose waits are bounded.
//
// The library is compiled with -ffp-contract=off: every product below is
// rounded before it is added, exactly like the reference's x86-64 code, so a
// row result is bit-identical to the reference's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvc {

constexpr int kWave = 64;
constexpr int kMaxWaves = 16;   // up to 1024 threads per block
constexpr int kMaxK = 5;        // partial sums per reduction

// Row-operator kinds (device formats).
enum Fmt : int {
  F_ELL4 = 0,   // ELLPACK width 4, reference layout a(4,neq) slot-fastest
  F_ELL3 = 1,   // ELLPACK width 3
  F_ELLW = 2,   // ELLPACK any other width (the reference skips these)
  F_CSRD = 3,   // CSR without diagonal + ad
  F_CSR = 4,    // full CSR, diagonal inside the values, t starts at 0
  F_LSYM = 5,   // lower triangle + transposed gather (csrd_sym / csrc_sym)
  F_CSRC = 6,   // CSRC unsymmetric: al row values, au transposed upper
};

struct Mat {
  int neq, w;
  const double *ad;    // diagonal (unused by F_CSR)
  const double *av;    // off-diagonal values: ELL a(w,neq); CSR-like values
  const double *au;    // F_CSRC transposed-upper values
  const int *ia;       // row pointer, 1-based (CSR-like formats)
  const int *ja;       // column indices, 1-based
  const int *tptr;     // F_LSYM/F_CSRC: 0-based offsets per row into tsrc/tpos
  const int *tsrc;     // source row of each transposed entry (0-based, ascending)
  const int *tpos;     // position (0-based) of that entry in al/au
};

template <int F, class X>
__device__ __forceinline__ double row_apply(const Mat &A, int i, const X &xf);

// One matrix row held in registers.  ELLPACK rows of width 3 and 4 are loaded
// by load() ahead of their use (a phase issues them before its deferred
// scalar recurrence, so HBM streams while the block reduces); other formats
// load in apply().  apply() is the only definition of the row's arithmetic.
template <int F>
struct RowRegs {
  __device__ __forceinline__ void load(const Mat &, int) {}
  template <class X>
  __device__ __forceinline__ double apply(const Mat &A, int i, const X &xf) const {
    return row_apply<F>(A, i, xf);
  }
};
template <>
struct RowRegs<F_ELL4> {
  double a0, a1, a2, a3, d;
  int c0, c1, c2, c3;
  __device__ __forceinline__ void load(const Mat &A, int i) {
    const double *a = A.av + 4 * (size_t)i;
    const int *c = A.ja + 4 * (size_t)i;
    a0 = a[0]; a1 = a[1]; a2 = a[2]; a3 = a[3];
    c0 = c[0]; c1 = c[1]; c2 = c[2]; c3 = c[3];
    d = A.ad[i];
  }
  // src/Matvec.f:289-296: y = t1+t2+t3+t4+t5, t5 = ad*x(i)
  template <class X>
  __device__ __forceinline__ double apply(const Mat &, int i, const X &xf) const {
    const double t5 = d * xf(i);
    const double t1 = a0 * xf(c0 - 1);
    const double t2 = a1 * xf(c1 - 1);
    const double t3 = a2 * xf(c2 - 1);
    const double t4 = a3 * xf(c3 - 1);
    return (((t1 + t2) + t3) + t4) + t5;
  }
};
template <>
struct RowRegs<F_ELL3> {
  double a0, a1, a2, d;
  int c0, c1, c2;
  __device__ __forceinline__ void load(const Mat &A, int i) {
    const double *a = A.av + 3 * (size_t)i;
    const int *c = A.ja + 3 * (size_t)i;
    a0 = a[0]; a1 = a[1]; a2 = a[2];
    c0 = c[0]; c1 = c[1]; c2 = c[2];
    d = A.ad[i];
  }
  // src/Matvec.f:281-287: y = t1+t2+t3+t4, t4 = ad*x(i)
  template <class X>
  __device__ __forceinline__ double apply(const Mat &, int i, const X &xf) const {
    const double t4 = d * xf(i);
    const double t1 = a0 * xf(c0 - 1);
    const double t2 = a1 * xf(c1 - 1);
    const double t3 = a2 * xf(c2 - 1);
    return ((t1 + t2) + t3) + t4;
  }
};

// (A x)_i with x supplied by a gather functor xf(j), j 0-based.
template <int F, class X>
__device__ __forceinline__ double row_apply(const Mat &A, int i, const X &xf) {
  if constexpr (F == F_ELL4 || F == F_ELL3) {
    RowRegs<F> r;
    r.load(A, i);
    return r.apply(A, i, xf);
  } else if constexpr (F == F_ELLW) {
    // Same pattern for any width: slots left to right, diagonal last.
    const double *a = A.av + (size_t)A.w * (size_t)i;
    const int *c = A.ja + (size_t)A.w * (size_t)i;
    double t = a[0] * xf(c[0] - 1);
    for (int s = 1; s < A.w; ++s) t = t + a[s] * xf(c[s] - 1);
    return t + A.ad[i] * xf(i);
  } else if constexpr (F == F_CSRD) {
    // src/Matvec.f:28-36
    double t = A.ad[i] * xf(i);
    const int k1 = A.ia[i] - 1, k2 = A.ia[i + 1] - 1;
    for (int k = k1; k < k2; ++k) t = t + A.av[k] * xf(A.ja[k] - 1);
    return t;
  } else if constexpr (F == F_CSR) {
    // src/Matvec.f:68-75
    double t = 0.0;
    const int k1 = A.ia[i] - 1, k2 = A.ia[i + 1] - 1;
    for (int k = k1; k < k2; ++k) t = t + A.av[k] * xf(A.ja[k] - 1);
    return t;
  } else {
    // F_LSYM (src/Matvec.f:157-192,216-250) and F_CSRC (src/Matvec.f:102-135):
    // the row's own lower products, then the contributions the reference
    // scatters into y(i) from later rows, in increasing row order.
    double t = A.ad[i] * xf(i);
    const int k1 = A.ia[i] - 1, k2 = A.ia[i + 1] - 1;
    for (int k = k1; k < k2; ++k) t = t + A.av[k] * xf(A.ja[k] - 1);
    const double *up = (F == F_CSRC) ? A.au : A.av;
    const int e1 = A.tptr[i], e2 = A.tptr[i + 1];
    for (int e = e1; e < e2; ++e) t = t + up[A.tpos[e]] * xf(A.tsrc[e]);
    return t;
  }
}

// ------------------------------------------------- lane neighbour exchange --
//
// A wavefront's lanes hold consecutive rows (Grid::first), so on a mesh's
// natural numbering a row's west and east neighbours i-1 and i+1 are the
// rows of lanes t-1 and t+1.  A fused phase forms its SpMV operand on the
// fly from several vectors (p/m from s, t, p, v and m in BiCGSTAB), and it
// forms that operand for its own row anyway; the DPP wave shifts hand it to
// the lanes on either side, so those two neighbours cost two register moves
// instead of a gather of every vector.  The operand is the same expression on
// the same values, so results are bit-identical.  The rows of one pass that
// are below the block's row end are a prefix of the wavefront's lanes: lane
// t-1 of an active lane is active; lane t+1 is when row i+1 is below the end.
// Lane 0's west and lane 63's east neighbours, and columns that are not i+-1,
// are gathered as before.
constexpr int kDppWaveShl1 = 0x130;   // lane t reads lane t+1
constexpr int kDppWaveShr1 = 0x138;   // lane t reads lane t-1

template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, Ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), Ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <class X>
struct XLane {
  const X &xf;
  double own, west, east;
  int i;
  bool wok, eok;
  __device__ __forceinline__ double operator()(int j) const {
    if (j == i) return own;
    if (wok && j == i - 1) return west;
    if (eok && j == i + 1) return east;
    return xf(j);
  }
};

// Compile-time switch of a row sweep: with or without the exchange.
template <bool B>
struct LaneX {
  static constexpr bool value = B;
};

// xf(own row i) = own; i1 = the row end of this pass.
template <class X>
__device__ __forceinline__ XLane<X> lane_exchange(const X &xf, double own, int i, int i1) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  return XLane<X>{xf, own, dpp_f64<kDppWaveShr1>(own), dpp_f64<kDppWaveShl1>(own), i, lane > 0,
                  lane < kWave - 1 && i + 1 < i1};
}

// ------------------------------------------------------------- reductions --

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;   // valid in lane 0
}

// A block barrier that orders LDS only: the block's outstanding global loads
// and stores stay in flight across it.  For code in which no thread reads
// global data another thread of the block wrote in the same launch.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool LdsOnly>
__device__ __forceinline__ void block_barrier() {
  if constexpr (LdsOnly) lds_barrier();
  else __syncthreads();
}

// Sums K values over the block; the totals land in thread 0's v[].
// LdsOnly: the barriers order LDS only (see lds_barrier).
// Each value is summed over the waves in wave order by its own thread (k),
// so the dependent LDS reads form K chains of nw in parallel instead of one
// chain of K nw on thread 0; thread 0 then reads the K totals at once.
template <int K, bool LdsOnly = false>
__device__ __forceinline__ void block_sum(double *v, double *lds) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * kMaxWaves + wid] = v[k];
  }
  block_barrier<LdsOnly>();
  double s = 0.0;
  const int q = (int)threadIdx.x;
  if (q < K) {
    s = lds[q * kMaxWaves];
    for (int w = 1; w < nw; ++w) s += lds[q * kMaxWaves + w];
  }
  block_barrier<LdsOnly>();
  if (q < K) lds[q * kMaxWaves] = s;
  block_barrier<LdsOnly>();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = lds[k * kMaxWaves];
  }
  block_barrier<LdsOnly>();
}

// Deterministic sum of G partials per value (partials[k*G + c]): thread t
// adds c = t, t + nt, t + 2 nt, ... in that order, then the block sums the
// threads.  The loads of U consecutive terms are issued together, so a
// thread waits for one round trip per U terms, not one per term (U = 1 for
// kernels whose registers are already scarce).
template <int K, bool LdsOnly = false, int U = 4>
__device__ __forceinline__ void sum_partials(const double *partials, int G, double *v,
                                             double *lds) {
  const int nt = blockDim.x;
  double s[K];
#pragma unroll
  for (int k = 0; k < K; ++k) s[k] = 0.0;
  if constexpr (U == 1) {
    for (int c = threadIdx.x; c < G; c += nt) {
#pragma unroll
      for (int k = 0; k < K; ++k) s[k] += partials[(size_t)k * G + c];
    }
  } else {
  for (int c = threadIdx.x; c < G; c += U * nt) {
    double q[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int cu = c + u * nt;
#pragma unroll
      for (int k = 0; k < K; ++k) q[u][k] = cu < G ? partials[(size_t)k * G + cu] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c + u * nt < G) {
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] += q[u][k];
      }
    }
  }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = s[k];
  block_sum<K, LdsOnly>(v, lds);
}

// The same sum with the loads issued ahead: load() takes the thread's terms
// of the first KM values (at most four terms each, G <= 4 nt) into registers
// before the caller knows which K <= KM it needs; sum<K>() then adds the
// first K values in sum_partials' order.  With G > 4 nt load() holds nothing
// and sum<K>() falls back to sum_partials.
template <int KM>
struct PartialRegs {
  double q[4][KM];
  bool held;
  __device__ __forceinline__ void load(const double *partials, int G) {
    const int nt = blockDim.x, c = threadIdx.x;
    held = G <= 4 * nt;
    if (!held) return;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int cu = c + u * nt;
#pragma unroll
      for (int k = 0; k < KM; ++k) q[u][k] = cu < G ? partials[(size_t)k * G + cu] : 0.0;
    }
  }
  template <int K, bool LdsOnly>
  __device__ __forceinline__ void sum(const double *partials, int G, double *v, double *lds) const {
    static_assert(K <= KM, "more values than loaded");
    if (!held) {
      sum_partials<K, LdsOnly>(partials, G, v, lds);
      return;
    }
    const int nt = blockDim.x, c = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (c + u * nt < G) s += q[u][k];
      v[k] = s;
    }
    block_sum<K, LdsOnly>(v, lds);
  }
};

// -------------------------------------------------- grid barrier + reduce --

// Terms per load batch in the grid barrier's partial sums: one, the
// persistent kernels (the batched ones above all) have no registers to spare.
constexpr int kGridSumU = 1;
//
// ctr layout (unsigned, 128-byte strides): [g*32] arrival counter of group g
// (g = block % 8, matching the round-robin placement of blocks on the eight
// XCDs), [8*32] top-level counter, [9*32] generation, [10*32] error flag.
constexpr int kCtrStride = 32;
constexpr int kCtrTop = 8 * kCtrStride;
constexpr int kCtrGen = 9 * kCtrStride;
constexpr int kCtrErr = 10 * kCtrStride;
constexpr int kCtrWords = 11 * kCtrStride;
// Bounded wait: ~2^23 polls with s_sleep in between (about a second);
// a barrier that never completes sets the error flag and every block leaves.
constexpr unsigned kSpinLimit = 1u << 23;

__device__ __forceinline__ unsigned ld_relaxed(unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned add_relaxed(unsigned *p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acq_rel() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent"); }

struct Grid {
  int G;                 // blocks = row chunks
  int R;                 // rows per chunk (tiles: rows per group of blocks)
  int neq;
  int tiles;             // row order: 0 contiguous chunk per block, 1 XCD tiles
  unsigned *ctr;
  double *partials;      // kMaxK * G
  bool persistent;
  int base;              // first block of this grid in the launch (batches)
  int flagbar;           // persistent barrier: 1 = per-block flags, 0 = counters
  unsigned *flags;       // G arrival flags (flag barrier)
  unsigned *ep;          // this block's barrier count (LDS), persistent kernels

  __device__ __forceinline__ int bid() const { return (int)blockIdx.x - base; }
  // XCD-aware chunk: consecutive block ids go to different XCDs, so give each
  // XCD a contiguous 1/8 of the rows (neighbour gathers stay in its L2).
  __device__ __forceinline__ int chunk() const {
    const int b = bid();
    return ((G & 7) == 0) ? (b & 7) * (G >> 3) + (b >> 3) : b;
  }
  // This thread's rows are first(), first() + stride(), ... below row_end().
  // Chunks: block b's thread t takes chunk() R + t, + bs, ...  A block then
  // meets a row's neighbours one grid row (nx rows) later in its own sweep,
  // which stays in L2 only while nx is a few block sizes.
  // Tiles: the 8 groups of blocks (one per XCD) own [x R, (x+1) R) each,
  // and a group's G/8 blocks sweep it together, bs rows per block: block j
  // of group x takes x R + j bs + t, + (G/8) bs, ...  The group's rows in
  // flight are then (G/8) bs consecutive ones, so the neighbours nx rows
  // away are read by the XCD at about the same time at any nx below that.
  __device__ __forceinline__ int ngroups() const { return ((G & 7) == 0) ? 8 : 1; }
  __device__ __forceinline__ int first() const {
    const int t = (int)threadIdx.x;
    if (!tiles) return chunk() * R + t;
    const int b = bid(), ng = ngroups();
    return (b % ng) * R + (b / ng) * (int)blockDim.x + t;
  }
  __device__ __forceinline__ int stride() const {
    return tiles ? (G / ngroups()) * (int)blockDim.x : (int)blockDim.x;
  }
  __device__ __forceinline__ int row_end() const {
    const long e = (long)((tiles ? bid() % ngroups() : chunk()) + 1) * R;
    return e < neq ? (int)e : neq;
  }
};

// Rows per chunk (per group of blocks with tiles) for a grid of G blocks
__host__ __device__ inline int grid_rows(int neq, int G, int tiles) {
  const int parts = tiles ? (((G & 7) == 0) ? 8 : 1) : G;
  return (int)(((long)neq + parts - 1) / parts);
}

__device__ __forceinline__ int group_size(int G, int g) { return (G - g + 7) / 8; }

// One-item-per-thread launches over cells or rows: this block's place when
// each XCD (the blocks b = x mod 8 under round-robin placement) takes a
// contiguous eighth of the blocks, so the neighbours a grid row away are
// gathered through the same L2.  A block count that is not a multiple of 8
// keeps the identity; launches round up with xcd_blocks.
__device__ __forceinline__ int xcd_block() {
  const int b = (int)blockIdx.x, nb = (int)gridDim.x;
  return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}
inline unsigned xcd_blocks(long items, int bs) {
  const long nb = (items + bs - 1) / bs;
  return (unsigned)((nb + 7) & ~7L);
}

// Reduce K per-thread values over the whole grid and run `epi(tot)` once, on
// thread 0 of the last block to arrive.  In persistent mode every block then
// waits for the epilogue to be published; in kernel-per-phase mode the other
// blocks return at once and the kernel boundary publishes it.
// Returns false if a bounded wait expired (the caller must leave).
//
// Memory ordering: each block releases once (its vector writes and its
// partial) before a relaxed arrival; group leaders and the last block
// acquire once; waiters poll with relaxed loads and acquire once when the
// generation moves.  On gfx950 an agent-scope acquire invalidates the L2 of
// the XCD, so it happens once per block per barrier, by thread 0 only; the
// block's other waves order behind it with the workgroup barrier.
template <int K, class Epi>
__device__ bool grid_reduce(const Grid &g, double *v, double *lds, Epi &&epi) {
  __shared__ int s_last;
  __shared__ unsigned s_gen;
  __shared__ int s_ok;
  // every wave's vector writes precede thread 0's release (block_sum's
  // workgroup barrier does it when there is something to sum)
  if constexpr (K > 0) block_sum<K>(v, lds);
  else __syncthreads();
  if (threadIdx.x == 0) {
    const int c = g.chunk();
#pragma unroll
    for (int k = 0; k < K; ++k) g.partials[(size_t)k * g.G + c] = v[k];
    const unsigned gen = g.persistent ? ld_relaxed(g.ctr + kCtrGen) : 0u;
    fence_release();
    const int grp = g.bid() & 7;
    const int ngroups = g.G < 8 ? g.G : 8;
    int last = 0;
    const unsigned t = add_relaxed(g.ctr + grp * kCtrStride, 1u);
    if ((int)t + 1 == group_size(g.G, grp)) {
      st_relaxed(g.ctr + grp * kCtrStride, 0u);
      fence_acq_rel();
      const unsigned t2 = add_relaxed(g.ctr + kCtrTop, 1u);
      if ((int)t2 + 1 == ngroups) {
        st_relaxed(g.ctr + kCtrTop, 0u);
        fence_acquire();
        last = 1;
      }
    }
    s_last = last;
    s_gen = gen;
    s_ok = 1;
  }
  __syncthreads();
  if (s_last) {
    double tot[K > 0 ? K : 1];
    if constexpr (K > 0) {
      double w[K];
      sum_partials<K, false, kGridSumU>(g.partials, g.G, w, lds);
#pragma unroll
      for (int k = 0; k < K; ++k) tot[k] = w[k];
    }
    if (threadIdx.x == 0) epi(tot);
    __syncthreads();
    if (threadIdx.x == 0 && g.persistent) {
      fence_release();
      st_relaxed(g.ctr + kCtrGen, s_gen + 1u);
    }
  } else if (g.persistent) {
    if (threadIdx.x == 0) {
      unsigned spins = 0;
      while (ld_relaxed(g.ctr + kCtrGen) == s_gen) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit || ld_relaxed(g.ctr + kCtrErr) != 0u) {
          st_relaxed(g.ctr + kCtrErr, 1u);
          s_ok = 0;
          break;
        }
      }
      fence_acquire();
    }
  }
  if (g.persistent) {
    __syncthreads();
    return s_ok != 0;
  }
  return true;
}

// Persistent grid reduction with per-block arrival flags: every block stores
// its partial, releases, and writes its own flag (no atomic read-modify-write
// to serialise on); block 0 polls all G flags in parallel, acquires, sums the
// partials in fixed order, runs the epilogue, releases and publishes the
// generation the others poll.  Same partials, same summation order as
// grid_reduce, so the results are identical.
template <int K, class Epi>
__device__ bool grid_reduce_flags(const Grid &g, double *v, double *lds, Epi &&epi) {
  __shared__ int s_ok;
  if constexpr (K > 0) block_sum<K>(v, lds);
  else __syncthreads();   // as in grid_reduce: all waves' writes before the release
  const unsigned e = *g.ep + 1u;          // barriers passed so far by every block, plus one
  const int b = g.bid();
  if (threadIdx.x == 0) {
    const int c = g.chunk();
#pragma unroll
    for (int k = 0; k < K; ++k) g.partials[(size_t)k * g.G + c] = v[k];
    fence_release();
    st_relaxed(g.flags + b, e);
    s_ok = 1;
  }
  if (b == 0) {
    __syncthreads();
    int ok = 1;
    for (int q = threadIdx.x; q < g.G; q += blockDim.x) {
      unsigned spins = 0;
      while (ld_relaxed(g.flags + q) < e) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit || ld_relaxed(g.ctr + kCtrErr) != 0u) {
          st_relaxed(g.ctr + kCtrErr, 1u);
          ok = 0;
          break;
        }
      }
      if (!ok) break;
    }
    if (!ok) s_ok = 0;
    __syncthreads();
    if (threadIdx.x == 0) fence_acquire();
    __syncthreads();
    double tot[K > 0 ? K : 1];
    if constexpr (K > 0) {
      double w[K];
      sum_partials<K, false, kGridSumU>(g.partials, g.G, w, lds);
#pragma unroll
      for (int k = 0; k < K; ++k) tot[k] = w[k];
    }
    if (threadIdx.x == 0) {
      if (s_ok) epi(tot);
      fence_release();
      st_relaxed(g.ctr + kCtrGen, e);
    }
  } else if (threadIdx.x == 0) {
    unsigned spins = 0;
    while (ld_relaxed(g.ctr + kCtrGen) < e) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kSpinLimit || ld_relaxed(g.ctr + kCtrErr) != 0u) {
        st_relaxed(g.ctr + kCtrErr, 1u);
        s_ok = 0;
        break;
      }
    }
    fence_acquire();
  }
  __syncthreads();
  if (threadIdx.x == 0) *g.ep = e;
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}

// The persistent engines' barrier: flags or counters as configured.
template <int K, class Epi>
__device__ __forceinline__ bool grid_sync(const Grid &g, double *v, double *lds, Epi &&epi) {
  if (g.persistent && g.flagbar) return grid_reduce_flags<K>(g, v, lds, epi);
  return grid_reduce<K>(g, v, lds, epi);
}

}  // namespace fvc
