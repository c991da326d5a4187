// fvcell.hip -- host side of the gfx950 solver library: the C ABI declared
// in include/fvcell.h, device workspaces, launch configuration and the two
// engines (persistent cooperative kernel, kernel per phase).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fvcell.h"
#include "fvc_instances.h"   // the solver kernels, built in k_*.hip
#include "fvc_plain.h"
#include "fvc_host.h"   // host-only algorithms (also built with sanitizers on the CPU)
#include "fvc_adjoint.h"
#include "fvc_post.h"

using namespace fvc;

// The workspace of fvc_solve_dev_noalloc: one fixed array per device, which
// that device holds until the process ends.  Host code never passes this
// name to a kernel or a copy (in host code it is only a stand-in for the
// symbol); it uses the device address that hipGetSymbolAddress returned at
// set-up (DevState::scratch_dev).
constexpr size_t kScratchBytes = (size_t)FVC_NOALLOC_WORKSPACE_BYTES;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// A workspace whose pieces outgrow the size computed for it is a defect of
// this library, reported before any kernel would write past its end.
int carve_overflow(const Carve &cv) {
  return fail(FVC_DEVICE_ERROR, "internal error: workspace layout needs %zu bytes of %zu", cv.off, cv.cap);
}

#define HIPCHK(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(FVC_DEVICE_ERROR, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                  __FILE__, __LINE__);                                                      \
  } while (0)

constexpr int kDefaultBlock = 512;
// Mixed precision: an FP64 residual recomputation at least every 200 FP32
// iterations, or when the residual estimate fell by 0.3 -- frequent
// refreshes keep the FP32 recursion close to FP64 CG (measured on MI355X:
// 8.4M-cell lognormal triangles, 41.5k FP32 iterations at 200 against 49k at
// 1000 and 194k at 5000; FP64 needs 23.6k)
constexpr int kRefreshEvery = 200;
constexpr double kRefreshDrop = 0.3;

// ------------------------------------------------------- device workspace --

struct MatDesc {
  int F = -1;
  Mat m{};
  int nval = 0;          // entries of av (and au)
  bool needs_trans = false;
  bool csr_diag = false; // Jacobi diagonal must be extracted (full CSR)
};

// Device buffers the host-array path keeps between calls (grow-only): its
// uploads reuse them instead of allocating and freeing per call.
struct DevBufs {
  static constexpr int kSlots = 12;
  void *p[kSlots] = {};
  size_t cap[kSlots] = {};
  int next = 0;
};

// What identifies a host matrix for fvc_keep_matrix: its format, shape and
// the addresses of its arrays (the caller vouches for the contents).
struct HostMatKey {
  int matriz = -1, unsym = 0, neq = 0, w = 0;
  size_t nval = 0;
  const void *ad = nullptr, *au = nullptr, *al = nullptr, *iax = nullptr, *ja = nullptr;
  bool operator==(const HostMatKey &o) const {
    return matriz == o.matriz && unsym == o.unsym && neq == o.neq && w == o.w && nval == o.nval &&
           ad == o.ad && au == o.au && al == o.al && iax == o.iax && ja == o.ja;
  }
};

struct DevState {
  bool ready = false;
  int device = -1;
  int cus = 0;
  char name[256] = {0};
  std::mutex mu;
  // one grow-only device arena for the solver's scratch vectors and state
  char *arena = nullptr;
  size_t arena_bytes = 0;
  KState *host_st = nullptr;   // pinned
  KState *host_ring = nullptr; // pinned, two slots (overlapped done-checks)
  double *host_scalar = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evq[2] = {nullptr, nullptr};
  // host-array solves: the uploaded matrix (kept for fvc_keep_matrix), x, b
  DevBufs mat_bufs, vec_bufs;
  HostMatKey mat_key;
  MatDesc mat_md;
  const double *mat_m = nullptr;
  bool mat_valid = false;
  bool keep_next = false;
  // fvc_solve_dev_noalloc: the device address of `scratch` on this device, and
  // whether the call in progress must take its workspace from there
  char *scratch_dev = nullptr;
  bool noalloc = false;
};

DevState g_dev[64];
std::mutex g_init_mu;

int current_device(DevState **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FVC_DEVICE_ERROR, "device index %d out of range", dev);
  DevState &d = g_dev[dev];
  if (!d.ready) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (!d.ready) {
      hipDeviceProp_t prop;
      HIPCHK(hipGetDeviceProperties(&prop, dev));
      d.device = dev;
      d.cus = prop.multiProcessorCount;
      std::snprintf(d.name, sizeof d.name, "%s (%s)", prop.name, prop.gcnArchName);
      HIPCHK(hipHostMalloc((void **)&d.host_st, sizeof(KState), hipHostMallocDefault));
      HIPCHK(hipHostMalloc((void **)&d.host_ring, 2 * sizeof(KState), hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&d.evq[0], hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&d.evq[1], hipEventDisableTiming));
      HIPCHK(hipHostMalloc((void **)&d.host_scalar, 64, hipHostMallocDefault));
      HIPCHK(hipEventCreate(&d.ev0));
      HIPCHK(hipEventCreate(&d.ev1));
      // this device's copy of the fixed workspace (its own address per device)
      HIPCHK(hipGetSymbolAddress((void **)&d.scratch_dev, HIP_SYMBOL(scratch)));
      d.ready = true;
    }
  }
  *out = &d;
  return FVC_OK;
}

int ensure_arena(DevState &d, size_t bytes) {
  if (d.arena_bytes >= bytes) return FVC_OK;
  if (d.arena) {
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(d.arena));
    d.arena = nullptr;
    d.arena_bytes = 0;
  }
  size_t want = std::max(bytes, d.arena_bytes * 5 / 4);
  HIPCHK(hipMalloc((void **)&d.arena, want));
  d.arena_bytes = want;
  return FVC_OK;
}

// The workspace of a solve: the grow-only arena, or for fvc_solve_dev_noalloc
// the fixed per-device array, which must hold it.
int workspace(DevState &d, size_t bytes, char **base) {
  if (d.noalloc) {
    if (bytes > kScratchBytes)
      return fail(FVC_UNSUPPORTED,
                  "fvc_solve_dev_noalloc: this solve needs %zu bytes of workspace, the fixed workspace "
                  "holds %zu (FVC_NOALLOC_WORKSPACE_BYTES); use fvc_solve_dev", bytes, kScratchBytes);
    *base = d.scratch_dev;
    return FVC_OK;
  }
  const int rc = ensure_arena(d, bytes);
  *base = d.arena;
  return rc;
}

// ------------------------------------------------------------ matrix view --

int fmt_for(int matriz, int unsym, int w, int *F) {
  switch (matriz) {
    case FVC_ELLPACK: *F = (w == 4) ? F_ELL4 : (w == 3) ? F_ELL3 : F_ELLW; return FVC_OK;
    case FVC_CSRD: *F = unsym ? F_CSRD : F_LSYM; return FVC_OK;
    case FVC_CSR: *F = F_CSR; return FVC_OK;
    case FVC_CSRC: *F = unsym ? F_CSRC : F_LSYM; return FVC_OK;
  }
  return fail(FVC_UNSUPPORTED, "matriz %d is not a storage format (1 CSRD, 2 CSR, 3 CSRC, 4 ELLPACK)", matriz);
}

// --------------------------------------------------------- launch config --

template <class K>
int occupancy(K kernel, int bs, int *nb) {
  *nb = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, kernel, bs, 0));
  return FVC_OK;
}

// ctr must hold kCtrWords counters followed by G arrival flags.
Grid make_grid(int neq, int G, unsigned *ctr, double *partials, bool persistent, int barrier = 0,
               int tiles = 0) {
  Grid g;
  g.G = G;
  g.neq = neq;
  g.tiles = tiles;
  g.R = grid_rows(neq, G, tiles);
  g.ctr = ctr;
  g.partials = partials;
  g.persistent = persistent;
  g.base = 0;
  g.flagbar = barrier == 0 ? 1 : 0;
  g.flags = ctr + kCtrWords;
  g.ep = nullptr;
  return g;
}

// A thread's row index runs up to the last row plus one stride (at most
// G bs), which must stay an int.
int check_rows(int neq, int G, int bs) {
  if ((long)neq + (long)G * bs > (long)INT_MAX)
    return fail(FVC_UNSUPPORTED, "%d rows with %d blocks of %d threads overflow 32-bit row indices", neq, G, bs);
  return FVC_OK;
}

// Row order of a solve (Grid::first).  Default: tiles once every thread of
// the grid has a row; below that, chunks give every block some rows.  On
// MI355X tiles took BiCGSTAB at 1M cells from 2.80 to 2.71 ms, PCG from 32.0
// to 29.6 us per iteration, and at 33.5M cells PCG from 1353 to 1017 us per
// iteration (FP32 inner iterations: 649 to 458).  GMRES, whose Arnoldi phases
// are mostly dot products over whole vectors, keeps chunks (GMRES(30) at 1M
// cells: 17.1 ms against 17.5 with tiles).
int use_tiles(int row_order, int code, int neq, int G, int bs) {
  if (row_order == 1) return 0;
  if (row_order == 2) return 1;
  if (code == FVC_GMRES) return 0;
  return (long)neq >= (long)G * bs ? 1 : 0;
}

// The kernel-per-phase engine is not bound to co-resident blocks.  Its
// deferred reductions have every block sum all G partials, a cost that grows
// as G^2; on MI355X at 1M cells four blocks per CU were best (BiCGSTAB 3.13 ms
// at 512 threads against 3.24 ms at two blocks per CU; PCG 34.9 us per
// iteration at 256 threads against 38.1).
int kernels_blocks(int neq, int cus, int requested, int per_cu = 4) {
  int G = requested > 0 ? requested : per_cu * cus;
  G = std::min(G, std::max(1, (neq + 63) / 64));
  if (G >= 64) G -= G % 8;
  return std::max(G, 1);
}
int kernels_block(int code, bool fused) { return (fused && code != FVC_GMRES) ? 256 : 512; }

int choose_blocks(int neq, int cus, int occ_blocks, int requested) {
  int G = requested > 0 ? requested : cus * std::max(1, std::min(occ_blocks, 2));
  G = std::min(G, std::max(1, (neq + 63) / 64));
  if (G >= 64) G -= G % 8;
  return std::max(G, 1);
}

// ------------------------------------------------------------ the solves --

struct SolveIn {
  int code, F;
  Mat A;
  const double *m;
  double *x;
  const double *b;
  double tol;
  int maxit, x0_zero, stop_rule, gmres_k, engine, bs, G_req, precision, barrier, recurrence;
  int row_order;          // FVC options' row_order: 0 rule below, 1 chunks, 2 XCD tiles
  int lanex;              // fused PCG / BiCGSTAB: neighbour operands from adjacent lanes
  int orth;               // GMRES: FVC_ORTH_*
  int refresh_every;      // mixed: FP32 iterations between FP64 residual recomputations, at most
  double refresh_drop;    // mixed: residual drop that triggers one
};

// Persistent launches are ordinary launches whose grid fits the device at
// once (G <= co-resident blocks, checked with the occupancy API before every
// persistent launch); the runtime's cooperative-launch queue is not used (it
// is not shareable between processes).  Should blocks nevertheless not be
// co-resident, the kernels' bounded waits expire, every block leaves, and the
// solve reports it (the AUTO engine then reruns it one kernel per phase).
template <class KArgs>
int launch(const void *fn, bool persistent, int G, int bs, KArgs &args, hipStream_t s) {
  (void)persistent;
  void *params[] = {&args};
  HIPCHK(hipLaunchKernel(fn, dim3(G), dim3(bs), params, 0, s));
  return FVC_OK;
}

bool barrier_expired(unsigned *ctr) {
  unsigned err = 0;
  if (hipMemcpy(&err, ctr + kCtrErr, sizeof err, hipMemcpyDeviceToHost) != hipSuccess) return true;
  return err != 0;
}

int read_state(DevState &d, KState *st, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(d.host_st, st, sizeof(KState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FVC_OK;
}

void fill_stats(fvc_stats *stats, const KState &hs, int code, bool persistent, float ms, int G, int bs,
                int flags);
int finish(const KState &hs, unsigned *ctr);

// The size of the next batch: doubling up to kMaxBatch, capped near the end
// of the solve by a convergence-rate model.  From the first observed state
// (it0, l0) and the latest one, the residual measure's geometric rate per
// iteration predicts the iterations still needed; the batch covers them, less
// those already enqueued, plus one (at least one, so that the GPU always has
// a batch while the host waits).  Without a usable rate it just doubles; a
// wrong prediction costs host round trips, never results.
int next_batch(const KState &h, int launched, int batch, int &it0, double &l0) {
  constexpr int kMaxBatch = 8;
  const int nb = std::min(batch * 2, kMaxBatch);
  const double l = h.last, c = h.conv;
  if (!(l > 0.0) || !(c > 0.0) || !std::isfinite(l) || !std::isfinite(c)) return nb;
  if (it0 < 0) {
    it0 = h.iter;
    l0 = l;
    return nb;
  }
  if (h.iter <= it0 || !(l < l0)) return nb;
  const double rate = std::pow(l / l0, 1.0 / (h.iter - it0));
  const double remaining = std::log(c / l) / std::log(rate);
  if (!std::isfinite(remaining)) return nb;
  const double need = remaining - (launched - h.iter);
  const int want = need <= 0.0 ? 1 : (int)std::ceil(need) + 1;
  return std::max(1, std::min(nb, want));
}

// Kernel-per-phase driver: enqueues iterations in batches and checks for
// convergence one batch behind, so the GPU never idles on the host.  After
// each batch the latest state goes to pinned host memory behind an event;
// the host then waits only for the previous batch's event, while the GPU
// runs the current one.  Launches after convergence find `done` set and
// return at once (at most one batch of them, kept short by next_batch).
// `iteration()` enqueues one iteration; `state()` is the device state the
// last launch wrote.
template <class Iter, class State>
int drive_batches(DevState &d, hipStream_t s, int maxit, Iter &&iteration, State &&state) {
  int launched = 0, batch = 2, k = 0, it0 = -1;
  double l0 = 0.0;
  bool waiting[2] = {false, false};
  while (launched < maxit) {
    const int nb = std::min(batch, maxit - launched);
    for (int it = 0; it < nb; ++it) {
      const int rc = iteration();
      if (rc) return rc;
    }
    launched += nb;
    const int slot = k & 1, prev = slot ^ 1;
    HIPCHK(hipMemcpyAsync(d.host_ring + slot, state(), sizeof(KState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(d.evq[slot], s));
    waiting[slot] = true;
    ++k;
    if (waiting[prev]) {
      HIPCHK(hipEventSynchronize(d.evq[prev]));
      waiting[prev] = false;
      if (d.host_ring[prev].done) break;
      batch = next_batch(d.host_ring[prev], launched, batch, it0, l0);
    } else {
      batch = std::min(batch * 2, 8);
    }
  }
  return FVC_OK;
}

template <int F>
int run_solve(DevState &d, const SolveIn &in, fvc_stats *stats, hipStream_t s) {
  const int neq = in.A.neq;
  const bool fused = in.recurrence == FVC_REC_FUSED;
  // AUTO: PCG and BiCGSTAB run one kernel per phase (with kernel boundaries
  // for ordering, a phase needs no fence and no atomic; measured faster on
  // MI355X than the grid barrier of one persistent launch); GMRES, whose
  // Arnoldi step is a chain of short phases, runs persistent.
  // GMRES with lagged CGS2 runs one kernel per phase only (fvc_gmres_lag.h)
  const bool lag = in.code == FVC_GMRES && in.orth == FVC_ORTH_CGS2_LAG;
  if (lag && in.engine == FVC_ENGINE_PERSISTENT)
    return fail(FVC_UNSUPPORTED, "GMRES with lagged CGS2 runs one kernel per phase (engine %d)", in.engine);
  bool persistent = in.engine == FVC_ENGINE_PERSISTENT ||
                    (in.engine == FVC_ENGINE_AUTO && in.code == FVC_GMRES && !lag);
  // Persistent GMRES: its Arnoldi step is j+1 short phases, each ended by a
  // grid barrier whose cost grows with the blocks taking part, so fewer and
  // larger blocks win: 1024 threads on three quarters of the CUs (1M cells,
  // GMRES(30): 17.0 ms against 22.7 at 2 x 512 per CU, 17.9 at 128 blocks,
  // 19.1 at 256).
  const bool gmres_p = persistent && in.code == FVC_GMRES;
  const bool cgs = in.code == FVC_GMRES && in.orth == FVC_ORTH_CGS2;
  // CGS2 holds 16 partial sums per thread: 512 threads per block (see gmres_persistent)
  const int bs = in.bs > 0 ? in.bs : gmres_p ? (cgs ? 512 : 1024) : persistent ? kDefaultBlock
                                                                   : lag ? 256 : kernels_block(in.code, fused);
  if (bs % 64 != 0 || bs > 1024) return fail(FVC_BAD_ARGUMENT, "block size %d must be a multiple of 64 <= 1024", bs);
  if (lag && bs != kLagBlock)
    return fail(FVC_BAD_ARGUMENT, "GMRES with lagged CGS2 runs blocks of %d threads (asked %d)", kLagBlock, bs);
  const void *pfn = nullptr, *kfn = nullptr;
  if (in.code == FVC_PCG && fused) {
    pfn = (const void *)&cg1_persistent<F>;
    kfn = (const void *)&cg1_kernel<F>;
  } else if (in.code == FVC_PCG) {
    pfn = (const void *)&pcg_persistent<F>;
    kfn = (const void *)&pcg_kernel<F>;
  } else if (in.code == FVC_BICGSTAB && fused) {
    pfn = (const void *)&bicg2_persistent<F>;
    kfn = (const void *)&bicg2_kernel<F>;
  } else if (in.code == FVC_BICGSTAB) {
    pfn = (const void *)&bicg_persistent<F>;
    kfn = (const void *)&bicg_kernel<F>;
  } else if (cgs) {
    pfn = (const void *)&gmres_persistent<F, 1>;
    kfn = (const void *)&gmres_kernel<F, 1>;
  } else {
    pfn = (const void *)&gmres_persistent<F, 0>;
    kfn = (const void *)&gmres_kernel<F, 0>;
  }
  int occ = 0;
  int rc = occupancy(pfn, bs, &occ);
  if (rc) return rc;
  if (occ < 1) return fail(FVC_DEVICE_ERROR, "persistent kernel does not fit a CU at block size %d", bs);
  // GMRES, persistent: MGS on three quarters of the CUs (above); CGS2, with
  // three to five reductions per Arnoldi step instead of j + 1, three blocks
  // of 512 per two CUs, and no more blocks than rows / 512 (1M cells,
  // GMRES(30) to true 1e-10: 19.8 ms against 20.0 at one block per CU, 21.8
  // at two, 23.7 for MGS)
  const int G_cgs = std::max(8, std::min(3 * d.cus / 2, (neq / 512) & ~7));
  const int G_req = (gmres_p && in.G_req == 0) ? (cgs ? G_cgs : std::max(8, (3 * d.cus / 4) & ~7)) : in.G_req;
  // Fused BiCGSTAB from 24 rows per thread on: two blocks per CU (MI355X,
  // us per iteration at 1024 -> 512 blocks: 8.4M cells 491 -> 464, 16.7M
  // 1040 -> 945; equal at 4M).  PCG stays at four (16.7M: 553 against 722).
  const int per_cu = (in.code == FVC_BICGSTAB && fused && (long)neq >= 24L * 4 * d.cus * bs) ? 2 : 4;
  const int G = persistent ? choose_blocks(neq, d.cus, occ, G_req) : kernels_blocks(neq, d.cus, G_req, per_cu);
  if (persistent && G > occ * d.cus) {
    if (in.engine == FVC_ENGINE_PERSISTENT)
      return fail(FVC_BAD_ARGUMENT, "%d blocks exceed the %d co-resident blocks of a persistent launch", G, occ * d.cus);
    persistent = false;
  }
  if ((rc = check_rows(neq, G, bs))) return rc;
  const int k = in.gmres_k;
  if (in.code == FVC_GMRES && (k < 1 || k > kMaxGmresK))
    return fail(FVC_BAD_ARGUMENT, "GMRES restart k = %d outside [1, %d]", k, kMaxGmresK);
  if (cgs && k > kCgsWide)
    return fail(FVC_UNSUPPORTED, "GMRES with CGS2 orthogonalisation takes k <= %d (asked %d)", kCgsWide, k);
  if (lag && k > kMaxGmresLag)
    return fail(FVC_UNSUPPORTED, "GMRES with lagged CGS2 takes k <= %d (asked %d)", kMaxGmresLag, k);
  // hraw, (nu, a, c), totals
  const size_t lag_small = lag ? (size_t)(k + 1) * k + 2 * (size_t)k + 3 + (size_t)lag_kp(k) : 0;
  const size_t lag_parts = lag ? (size_t)lag_kp(k) * G : 0;
  // partial sums per block: CGS2 reduces kCgsChunk dot products at once
  const size_t npart = (size_t)(cgs ? kCgsChunk : kMaxK) * G;

  // workspace
  const size_t n = (size_t)neq;
  // GMRES: basis columns 1..k+1; lagged CGS2: columns 1..k, the candidate u and w
  int nvec = in.code == FVC_PCG ? (fused ? 7 : 5) : in.code == FVC_BICGSTAB ? (fused ? 8 : 10) : (k + 1 + (lag ? 1 : 0));
  size_t bytes = align_up((kCtrWords + (size_t)G) * sizeof(unsigned)) + align_up(sizeof(KState)) +
                 align_up(npart * sizeof(double)) + (size_t)nvec * align_up(n * sizeof(double)) +
                 align_up((size_t)(k + 2) * (k + 2) * sizeof(double)) * 2 + align_up(n * sizeof(double)) +
                 align_up(2 * sizeof(KState)) + align_up(2 * kMaxK * (size_t)G * sizeof(double)) +
                 align_up(lag_small * sizeof(double)) + align_up(lag_parts * sizeof(double)) + 4096;
  char *ws = nullptr;
  if ((rc = workspace(d, bytes, &ws))) return rc;
  Carve cv{ws, bytes};
  unsigned *ctr = cv.take<unsigned>(kCtrWords + (size_t)G);
  KState *st = cv.take<KState>(1);
  double *partials = cv.take<double>(npart);
  // deferred reductions (kernel-per-phase engine of the fused methods)
  Defer dfr{cv.take<KState>(2), cv.take<double>(2 * kMaxK * (size_t)G), 0};
  double *vec[kMaxGmresK + 1];
  for (int i = 0; i < nvec; ++i) vec[i] = cv.take<double>(n);
  double *gsmall = cv.take<double>((size_t)(k + 2) * (k + 2) * 2);
  double *xbackup = cv.take<double>(n);
  double *lagsmall = cv.take<double>(lag_small);
  double *lagpart = cv.take<double>(lag_parts);
  if (!cv.fits()) return carve_overflow(cv);
  const bool may_fall_back = persistent && in.engine == FVC_ENGINE_AUTO;
  if (may_fall_back && !in.x0_zero)
    HIPCHK(hipMemcpyAsync(xbackup, in.x, n * sizeof(double), hipMemcpyDeviceToDevice, s));

  // lagged CGS2's D kernel takes contiguous row chunks
  Grid g = make_grid(neq, G, ctr, partials, persistent, in.barrier,
                     lag ? 0 : use_tiles(in.row_order, in.code, neq, G, bs));
  HIPCHK(hipEventRecord(d.ev0, s));
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, (int)kCtrWords + G, st, 1);
  HIPCHK(hipGetLastError());

  if (in.code == FVC_PCG && fused) {
    Cg1Args a{};
    a.A = in.A; a.g = g; a.st = st; a.m = in.m; a.x = in.x; a.b = in.b;
    a.r0 = vec[0]; a.r1 = vec[1]; a.w0 = vec[2]; a.w1 = vec[3]; a.s0 = vec[4]; a.s1 = vec[5]; a.p = vec[6];
    a.tol = in.tol; a.maxit = in.maxit; a.x0_zero = in.x0_zero; a.stop_rule = in.stop_rule;
    a.lanex = in.lanex;
    if (persistent) {
      a.phase = -1;
      if ((rc = launch(pfn, true, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    } else {
      // each launch reads state/partials set par and writes set 1 - par
      auto phase = [&](int ph) {
        a.phase = ph;
        a.d = dfr;
        dfr.par ^= 1;
        return launch(kfn, false, G, bs, a, s);
      };
      hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, 0, dfr.st, 2);
      if ((rc = phase(CG1_SETUP))) return rc;
      if ((rc = phase(CG1_W0))) return rc;
      if ((rc = drive_batches(d, s, in.maxit, [&] { return phase(CG1_ITER); },
                              [&] { return dfr.st + dfr.par; })))
        return rc;
      if ((rc = phase(CG1_ITER))) return rc;   // applies the last iteration's recurrence
      if ((rc = phase(CG1_FINAL))) return rc;
      if ((rc = phase(-1))) return rc;          // applies FINAL's
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, dfr.st + dfr.par, s))) return rc;
    }
  } else if (in.code == FVC_BICGSTAB && fused) {
    Bicg2Args a{};
    a.A = in.A; a.g = g; a.st = st; a.m = in.m; a.x = in.x; a.b = in.b;
    a.bb = vec[0]; a.p0 = vec[1]; a.p1 = vec[2]; a.v0 = vec[3]; a.v1 = vec[4];
    a.s = vec[5]; a.t = vec[6]; a.r = vec[7];
    a.tol = in.tol; a.maxit = in.maxit; a.x0_zero = in.x0_zero; a.stop_rule = in.stop_rule;
    a.lanex = in.lanex;
    if (persistent) {
      a.phase = -1;
      if ((rc = launch(pfn, true, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    } else {
      auto phase = [&](int ph) {
        a.phase = ph;
        a.d = dfr;
        dfr.par ^= 1;
        return launch(kfn, false, G, bs, a, s);
      };
      hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, 0, dfr.st, 2);
      if ((rc = phase(B2_SETUP))) return rc;
      auto iteration = [&] {
        const int r = phase(B2_P);
        return r ? r : phase(B2_Q);
      };
      if ((rc = drive_batches(d, s, in.maxit, iteration, [&] { return dfr.st + dfr.par; }))) return rc;
      // the P phase that completes the last iteration and decides its stop
      // test, then one that applies that decision
      if ((rc = phase(B2_P))) return rc;
      if ((rc = phase(B2_Q))) return rc;
      if ((rc = phase(B2_FINAL))) return rc;
      if ((rc = phase(-1))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, dfr.st + dfr.par, s))) return rc;
    }
  } else if (in.code == FVC_PCG) {
    PcgArgs a{};
    a.A = in.A; a.g = g; a.st = st; a.m = in.m; a.x = in.x; a.b = in.b;
    a.r = vec[0]; a.z = vec[1]; a.p0 = vec[2]; a.p1 = vec[3]; a.q = vec[4];
    a.tol = in.tol; a.maxit = in.maxit; a.x0_zero = in.x0_zero; a.stop_rule = in.stop_rule;
    if (persistent) {
      a.phase = -1;
      rc = launch(pfn, true, G, bs, a, s);
      if (rc) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      rc = read_state(d, st, s);
      if (rc) return rc;
    } else {
      a.phase = PCG_SETUP;
      if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
      auto iteration = [&] {
        a.phase = PCG_A;
        const int r = launch(kfn, false, G, bs, a, s);
        if (r) return r;
        a.phase = PCG_B;
        return launch(kfn, false, G, bs, a, s);
      };
      if ((rc = drive_batches(d, s, in.maxit, iteration, [&] { return st; }))) return rc;
      a.phase = PCG_FINAL;
      if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    }
  } else if (in.code == FVC_BICGSTAB) {
    BicgArgs a{};
    a.A = in.A; a.g = g; a.st = st; a.m = in.m; a.x = in.x; a.b = in.b;
    a.bb0 = vec[0]; a.bb1 = vec[1]; a.p0 = vec[2]; a.p1 = vec[3]; a.v0 = vec[4]; a.v1 = vec[5];
    a.r = vec[6]; a.z = vec[7]; a.z2 = vec[8]; a.t = vec[9];
    a.tol = in.tol; a.maxit = in.maxit; a.x0_zero = in.x0_zero; a.stop_rule = in.stop_rule;
    if (persistent) {
      a.phase = -1;
      if ((rc = launch(pfn, true, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    } else {
      a.phase = BI_SETUP;
      if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
      auto iteration = [&] {
        for (int ph = BI_F1; ph <= BI_F3; ++ph) {
          a.phase = ph;
          const int r = launch(kfn, false, G, bs, a, s);
          if (r) return r;
        }
        return 0;
      };
      if ((rc = drive_batches(d, s, in.maxit, iteration, [&] { return st; }))) return rc;
      a.phase = BI_FINAL;
      if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    }
  } else {
    GmresArgs a{};
    a.A = in.A; a.g = g; a.st = st; a.m = in.m; a.x = in.x; a.b = in.b;
    a.gv = vec[0];
    a.ld = n;
    // vec[] are carved consecutively with 256-byte alignment: use one stride
    if (nvec > 1) a.ld = (size_t)(vec[1] - vec[0]);
    double *sm = gsmall;
    a.h = sm; sm += (size_t)(k + 1) * k;
    a.cs = sm; sm += k + 2;
    a.sn = sm; sm += k + 2;
    a.e = sm; sm += k + 2;
    a.y = sm; sm += k + 2;
    a.nrm = sm; sm += k + 2;
    a.tol = in.tol; a.k = k; a.maxit = in.maxit; a.stop_rule = in.stop_rule; a.orth = lag ? 2 : cgs ? 1 : 0;
    if (lag) {
      a.gu = vec[k];
      a.gw = vec[k + 1];
      a.hraw = lagsmall;
      a.lag = lagsmall + (size_t)(k + 1) * k;
      a.lag_tot = a.lag + 2 * (size_t)k + 3;
      a.lag_part = lagpart;
      a.lag_kp = lag_kp(k);
    }
    if (lag) {
      // lagged CGS2: per Arnoldi step D (SpMV and dot products), R (one
      // block: the reduction and the scalar part) and B (basis column and
      // next candidate).  A cycle's steps are enqueued in batches; the host
      // stops enqueueing one batch after the device reports the cycle done.
      for (int cyc = 0; cyc <= in.maxit; ++cyc) {
        a.phase = GM_START;
        if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
        int j = 1, batch = 4, q = 0;
        bool waiting[2] = {false, false};
        while (j <= k + 1) {
          const int nb = std::min(batch, k + 2 - j);
          for (int t = 0; t < nb; ++t, ++j) {
            hipLaunchKernelGGL(gmres_lag_d_kernel<F>, dim3(G), dim3(bs), 0, s, a);
            hipLaunchKernelGGL(gmres_lag_r_kernel<F>, dim3(2 * j), dim3(kLagBlock), 0, s, a);
            if (j <= k) hipLaunchKernelGGL(gmres_lag_b_kernel<F>, dim3(G), dim3(bs), 0, s, a);
          }
          HIPCHK(hipGetLastError());
          const int slot = q & 1, prev = slot ^ 1;
          HIPCHK(hipMemcpyAsync(d.host_ring + slot, st, sizeof(KState), hipMemcpyDeviceToHost, s));
          HIPCHK(hipEventRecord(d.evq[slot], s));
          waiting[slot] = true;
          ++q;
          if (waiting[prev]) {
            HIPCHK(hipEventSynchronize(d.evq[prev]));
            waiting[prev] = false;
            if (d.host_ring[prev].inner_done || d.host_ring[prev].done) break;
          }
          batch = std::min(batch * 2, 8);
        }
        a.phase = GM_XUPD;
        if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
        if ((rc = read_state(d, st, s))) return rc;
        if (d.host_st->done) break;
      }
      a.phase = GM_FINAL;
      if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    } else if (persistent) {
      a.phase = -1;
      if ((rc = launch(pfn, true, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    } else {
      // under the true-residual rule the start of cycle maxit + 1 decides
      for (int cyc = 0; cyc <= in.maxit; ++cyc) {
        a.phase = GM_START;
        if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
        for (int i = 1; i <= k; ++i) {
          if (cgs) {
            const int nch = (i + kCgsChunk - 1) / kCgsChunk;
            for (int ph = GM_C1; ph <= GM_C3; ++ph)
              for (int c = 0; c < (ph == GM_C3 ? 1 : nch); ++c) {
                a.phase = ph;
                a.chunk = c;
                if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
              }
            continue;
          }
          a.phase = GM_MV;
          if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
          a.phase = GM_MGS;
          for (int j = 1; j <= i; ++j)
            if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
        }
        a.phase = GM_XUPD;
        if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
        if ((rc = read_state(d, st, s))) return rc;
        if (d.host_st->done) break;
      }
      a.phase = GM_FINAL;
      if ((rc = launch(kfn, false, G, bs, a, s))) return rc;
      HIPCHK(hipEventRecord(d.ev1, s));
      if ((rc = read_state(d, st, s))) return rc;
    }
  }
  HIPCHK(hipGetLastError());
  if (may_fall_back && barrier_expired(ctr)) {
    // blocks were not co-resident: rerun one kernel per phase from the same x0
    if (!in.x0_zero)
      HIPCHK(hipMemcpyAsync(in.x, xbackup, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SolveIn again = in;
    again.engine = FVC_ENGINE_KERNELS;
    return run_solve<F>(d, again, stats, s);
  }
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
  fill_stats(stats, *d.host_st, in.code, persistent, ms, G, bs, g.tiles ? FVC_FLAG_TILES : 0);
  // A bounded grid wait that expired is a device error, never a result.
  return finish(*d.host_st, ctr);
}

void fill_stats(fvc_stats *stats, const KState &hs, int code, bool persistent, float ms, int G, int bs,
                int flags) {
  if (!stats) return;
  std::memset(stats, 0, sizeof *stats);
  stats->status = hs.status;
  stats->iterations = hs.iter;
  stats->inner_iterations = code == FVC_GMRES ? hs.inner : hs.iter;
  stats->engine = persistent ? FVC_ENGINE_PERSISTENT : FVC_ENGINE_KERNELS;
  stats->flags = flags;
  stats->energy = hs.energy;
  stats->conv = hs.conv;
  stats->last = hs.last;
  stats->true_relres = hs.bnorm2 > 0 ? std::sqrt(hs.res2 / hs.bnorm2) : std::sqrt(hs.res2);
  stats->seconds = ms * 1e-3;
  stats->grid_blocks = G;
  stats->block_size = bs;
}

int finish(const KState &hs, unsigned *ctr) {
  unsigned err = 0;
  HIPCHK(hipMemcpy(&err, ctr + kCtrErr, sizeof err, hipMemcpyDeviceToHost));
  if (err) return fail(FVC_DEVICE_ERROR, "persistent solve: a grid barrier wait expired");
  if (hs.status != S_OK) {
    static const char *names[] = {"converged", "maximum iterations reached", "breakdown", "non-finite value"};
    return fail(hs.status, "%s after %d iterations", names[std::min(std::max(hs.status, 0), 3)], hs.iter);
  }
  return FVC_OK;
}

// Mixed-precision PCG (fvc_mixed.h): FP32 scaled matrix built per solve.
template <int F>
int run_mixed(DevState &d, const SolveIn &in, fvc_stats *stats, hipStream_t s) {
  if constexpr (F != F_ELL4 && F != F_ELL3 && F != F_ELLW && F != F_CSRD) {
    return fail(FVC_UNSUPPORTED, "mixed precision supports ELLPACK and unsymmetric CSRD storage");
  } else {
    const int neq = in.A.neq;
    // AUTO runs one kernel per phase, as for FP64 PCG
    const bool persistent = in.engine == FVC_ENGINE_PERSISTENT;
    // one-launch inner iterations: blocks of 256, four per CU as for FP64
    // PCG, measured best on MI355X at 1M cells (19.3 us per inner iteration
    // against 21.8 at two blocks of 512 and 29.0 at four of 512).  From 16
    // rows per thread on, three per CU: 60.3 against 62.6 us at 4M cells,
    // 237 against 253 at 16.7M, 418 against 454 at 33.5M (1.5 % slower at 1M)
    const int bs = in.bs > 0 ? in.bs : persistent ? kDefaultBlock : kernels_block(FVC_PCG, true);
    const int per_cu = (long)neq >= 16L * 4 * d.cus * bs ? 3 : 4;
    if (bs % 64 != 0 || bs > 1024) return fail(FVC_BAD_ARGUMENT, "block size %d must be a multiple of 64 <= 1024", bs);
    const void *pfn = (const void *)&mixed_persistent<F>;
    const void *kfn = (const void *)&mixed_kernel<F>;
    int occ = 0;
    int rc = occupancy(pfn, bs, &occ);
    if (rc) return rc;
    const int G = persistent ? choose_blocks(neq, d.cus, occ, in.G_req) : kernels_blocks(neq, d.cus, in.G_req, per_cu);
    if (persistent && G > occ * d.cus)
      return fail(FVC_BAD_ARGUMENT, "%d blocks exceed the %d co-resident blocks of a persistent launch", G, occ * d.cus);
    if ((rc = check_rows(neq, G, bs))) return rc;
    const size_t n = (size_t)neq;
    const size_t nval = (F == F_CSRD) ? (size_t)0 : (size_t)in.A.w * n;
    size_t nv = nval;
    if (F == F_CSRD) {
      int last = 0;
      HIPCHK(hipMemcpy(&last, in.A.ia + neq, sizeof(int), hipMemcpyDeviceToHost));
      nv = (size_t)std::max(last - 1, 0);
    }
    // int16 column offsets for ELLPACK widths 3 and 4 (col_offsets_kernel)
    constexpr int kW16 = (F == F_ELL4) ? 4 : (F == F_ELL3) ? 3 : 0;
    const size_t n16 = kW16 ? (size_t)kW16 * n : 0;
    size_t bytes = align_up((kCtrWords + (size_t)G + 1) * sizeof(unsigned)) + align_up(sizeof(KState)) +
                   align_up(kMaxK * (size_t)G * sizeof(double)) + align_up(nv * sizeof(float)) +
                   align_up(n * sizeof(double)) + align_up(n * sizeof(float)) + 2 * align_up(n * sizeof(float4)) +
                   align_up(2 * sizeof(KState)) + align_up(2 * kMaxK * (size_t)G * sizeof(double)) +
                   align_up(std::max<size_t>(n16, 1) * sizeof(short)) + 4096;
    char *ws = nullptr;
    if ((rc = workspace(d, bytes, &ws))) return rc;
    Carve cv{ws, bytes};
    // barrier counters, G arrival flags, then the offsets' "does not fit" flag
    unsigned *ctr = cv.take<unsigned>(kCtrWords + (size_t)G + 1);
    int *wide = reinterpret_cast<int *>(ctr + kCtrWords + G);
    KState *st = cv.take<KState>(1);
    double *partials = cv.take<double>(kMaxK * (size_t)G);
    Defer dfr{cv.take<KState>(2), cv.take<double>(2 * kMaxK * (size_t)G), 0};
    float *av32 = cv.take<float>(std::max<size_t>(nv, 1));
    short *jo = cv.take<short>(std::max<size_t>(n16, 1));
    double *sc = cv.take<double>(n);
    float *dsq = cv.take<float>(n);
    MixedArgs a{};
    a.ps = align_up(n * sizeof(float4)) / sizeof(float4);
    a.rec = cv.take<float4>(2 * a.ps);
    if (!cv.fits()) return carve_overflow(cv);
    a.A = in.A;
    a.A32.neq = neq;
    a.A32.w = in.A.w;
    a.A32.av = av32;
    a.A32.ia = in.A.ia;
    a.A32.ja = in.A.ja;
    a.g = make_grid(neq, G, ctr, partials, persistent, in.barrier, use_tiles(in.row_order, FVC_PCG, neq, G, bs));
    a.st = st;
    a.m = in.m;
    a.s = sc;
    a.dsq = dsq;
    a.x = in.x;
    a.b = in.b;
    a.tol = in.tol;
    a.delta = in.refresh_drop;
    a.maxit = in.maxit;
    a.x0_zero = in.x0_zero;
    a.max_inner = in.refresh_every;
    HIPCHK(hipEventRecord(d.ev0, s));
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, (int)kCtrWords + G + 1, st, 1);
    hipLaunchKernelGGL(scale_to_f32_kernel<F>, dim3((neq + 255) / 256), dim3(256), 0, s, in.A, in.m,
                       av32, sc, dsq);
    HIPCHK(hipGetLastError());
    if constexpr (kW16 > 0) {
      hipLaunchKernelGGL(col_offsets_kernel<kW16>, dim3(xcd_blocks(neq, 256)), dim3(256), 0, s, neq, in.A.ja,
                         av32, jo, wide);
      HIPCHK(hipGetLastError());
      int h_wide = 1;
      HIPCHK(hipMemcpyAsync(&h_wide, wide, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (!h_wide) a.A32.jo = jo;
    }
    const KState *final_state = st;
    if (persistent) {
      a.phase = -1;
      if ((rc = launch(pfn, true, G, bs, a, s))) return rc;
    } else {
      auto phase = [&](int ph) {
        a.phase = ph;
        a.d = dfr;
        dfr.par ^= 1;
        return launch(kfn, false, G, bs, a, s);
      };
      hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, 0, dfr.st, 2);
      if ((rc = phase(MX_SETUP))) return rc;
      // an inner iteration is two launches (A, B), a reliable update two more
      // (X, R), and the device stops at maxit: this many launches reach `done`
      const int cap = 4 * std::max(in.maxit, 0) + 16;
      if ((rc = drive_batches(d, s, cap, [&] { return phase(kPhaseNext); }, [&] { return dfr.st + dfr.par; })))
        return rc;
      if ((rc = phase(MX_FINAL))) return rc;
      if ((rc = phase(-1))) return rc;
      final_state = dfr.st + dfr.par;
    }
    HIPCHK(hipEventRecord(d.ev1, s));
    if ((rc = read_state(d, const_cast<KState *>(final_state), s))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
    fill_stats(stats, *d.host_st, FVC_PCG, persistent, ms, G, bs,
               (a.g.tiles ? FVC_FLAG_TILES : 0) | (a.A32.jo ? FVC_FLAG_COL16 : 0));
    if (stats) stats->inner_iterations = d.host_st->cycle;   // FP64 residual recomputations
    return finish(*d.host_st, ctr);
  }
}

// Mixed-precision BiCGSTAB and GMRES(k) (fvc_mixed_krylov.h): the FP32
// scaled matrix and int16 column offsets as for the mixed PCG, FP32 vectors,
// FP64 solution and residuals.
template <int F>
int run_mixed_krylov(DevState &d, const SolveIn &in, fvc_stats *stats, hipStream_t s) {
  if constexpr (F != F_ELL4 && F != F_ELL3 && F != F_ELLW && F != F_CSRD) {
    return fail(FVC_UNSUPPORTED, "mixed precision supports ELLPACK and unsymmetric CSRD storage");
  } else {
    const int neq = in.A.neq;
    const bool gm = in.code == FVC_GMRES;
    // GMRES with lagged CGS2: one kernel per phase (fvc_gmres_lag.h, FP32 storage)
    const bool lag = gm && in.orth == FVC_ORTH_CGS2_LAG;
    const int k = in.gmres_k;
    const int kmax = lag ? kMaxGmresLag : kCgsWide;
    if (gm && (k < 1 || k > kmax))
      return fail(FVC_UNSUPPORTED, "mixed-precision GMRES takes k <= %d (asked %d)", kmax, k);
    if (gm && !lag && in.engine == FVC_ENGINE_KERNELS)
      return fail(FVC_UNSUPPORTED, "mixed-precision GMRES with CGS2 runs as one persistent launch");
    if (lag && in.engine == FVC_ENGINE_PERSISTENT)
      return fail(FVC_UNSUPPORTED, "GMRES with lagged CGS2 runs one kernel per phase (engine %d)", in.engine);
    // BiCGSTAB: one kernel per phase by default, shaped as the FP64 two-phase
    // BiCGSTAB; GMRES: persistent, shaped as the FP64 CGS2 GMRES, or with
    // lagged CGS2 one kernel per phase as the FP64 one
    const bool persistent = (gm && !lag) || in.engine == FVC_ENGINE_PERSISTENT;
    const int bs = in.bs > 0 ? in.bs : lag ? kLagBlock : gm ? 512 : persistent ? kDefaultBlock
                                                                   : kernels_block(FVC_BICGSTAB, true);
    if (bs % 64 != 0 || bs > (gm ? 512 : 1024))
      return fail(FVC_BAD_ARGUMENT, "block size %d must be a multiple of 64 <= %d", bs, gm ? 512 : 1024);
    if (lag && bs != kLagBlock)
      return fail(FVC_BAD_ARGUMENT, "GMRES with lagged CGS2 runs blocks of %d threads (asked %d)", kLagBlock, bs);
    const void *pfn = gm ? (const void *)&mgmres_persistent<F> : (const void *)&mbicg_persistent<F>;
    const void *kfn = (const void *)&mbicg_kernel<F>;
    int occ = 0;
    int rc = occupancy(pfn, bs, &occ);
    if (rc) return rc;
    if (occ < 1) return fail(FVC_DEVICE_ERROR, "persistent kernel does not fit a CU at block size %d", bs);
    const int G_req = (gm && !lag && in.G_req == 0) ? std::max(8, std::min(3 * d.cus / 2, (neq / 512) & ~7)) : in.G_req;
    const int per_cu = (long)neq >= 24L * 4 * d.cus * bs ? 2 : 4;
    const int G = persistent ? choose_blocks(neq, d.cus, occ, G_req) : kernels_blocks(neq, d.cus, G_req, per_cu);
    if (persistent && G > occ * d.cus)
      return fail(FVC_BAD_ARGUMENT, "%d blocks exceed the %d co-resident blocks of a persistent launch", G, occ * d.cus);
    if ((rc = check_rows(neq, G, bs))) return rc;
    const size_t n = (size_t)neq;
    size_t nv = (F == F_CSRD) ? 0 : (size_t)in.A.w * n;
    if (F == F_CSRD) {
      int last = 0;
      HIPCHK(hipMemcpy(&last, in.A.ia + neq, sizeof(int), hipMemcpyDeviceToHost));
      nv = (size_t)std::max(last - 1, 0);
    }
    constexpr int kW16 = (F == F_ELL4) ? 4 : (F == F_ELL3) ? 3 : 0;
    const size_t n16 = kW16 ? (size_t)kW16 * n : 0;
    // GMRES: basis columns 1..k+1; lagged CGS2: columns 1..k, the candidate u and w
    const int nvec = gm ? k + 1 + (lag ? 1 : 0) : 9;
    const size_t npart = (size_t)(gm ? kCgsChunk : kMaxK) * G;
    // lagged CGS2: hraw, (nu, a, c), totals; per-block partials
    const size_t lag_small = lag ? (size_t)(k + 1) * k + 2 * (size_t)k + 3 + (size_t)lag_kp(k) : 0;
    const size_t lag_parts = lag ? (size_t)lag_kp(k) * G : 0;
    size_t bytes = align_up((kCtrWords + (size_t)G + 1) * sizeof(unsigned)) + align_up(sizeof(KState)) +
                   align_up(npart * sizeof(double)) + align_up(std::max<size_t>(nv, 1) * sizeof(float)) +
                   align_up(n * sizeof(double)) + align_up(n * sizeof(float)) +
                   (size_t)nvec * align_up(n * sizeof(float)) + align_up(2 * sizeof(KState)) +
                   align_up(2 * kMaxK * (size_t)G * sizeof(double)) +
                   align_up(std::max<size_t>(n16, 1) * sizeof(short)) +
                   align_up((size_t)(k + 2) * (k + 2) * 2 * sizeof(double)) +
                   align_up(lag_small * sizeof(double)) + align_up(lag_parts * sizeof(double)) + 4096;
    char *ws = nullptr;
    if ((rc = workspace(d, bytes, &ws))) return rc;
    Carve cv{ws, bytes};
    unsigned *ctr = cv.take<unsigned>(kCtrWords + (size_t)G + 1);
    int *wide = reinterpret_cast<int *>(ctr + kCtrWords + G);
    KState *st = cv.take<KState>(1);
    double *partials = cv.take<double>(npart);
    Defer dfr{cv.take<KState>(2), cv.take<double>(2 * kMaxK * (size_t)G), 0};
    float *av32 = cv.take<float>(std::max<size_t>(nv, 1));
    short *jo = cv.take<short>(std::max<size_t>(n16, 1));
    double *sc = cv.take<double>(n);
    float *dsq = cv.take<float>(n);
    const size_t fstride = align_up(n * sizeof(float)) / sizeof(float);
    float *fv = cv.take<float>((size_t)nvec * fstride);
    double *gsmall = cv.take<double>((size_t)(k + 2) * (k + 2) * 2);
    double *lagsmall = cv.take<double>(lag_small);
    double *lagpart = cv.take<double>(lag_parts);
    if (!cv.fits()) return carve_overflow(cv);
    Mat32 A32{neq, in.A.w, av32, in.A.ia, in.A.ja, nullptr};
    Grid g = make_grid(neq, G, ctr, partials, persistent, in.barrier,
                       gm ? 0 : use_tiles(in.row_order, FVC_BICGSTAB, neq, G, bs));
    HIPCHK(hipEventRecord(d.ev0, s));
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, (int)kCtrWords + G + 1, st, 1);
    hipLaunchKernelGGL(scale_to_f32_kernel<F>, dim3((neq + 255) / 256), dim3(256), 0, s, in.A, in.m, av32, sc, dsq);
    HIPCHK(hipGetLastError());
    if constexpr (kW16 > 0) {
      hipLaunchKernelGGL(col_offsets_kernel<kW16>, dim3(xcd_blocks(neq, 256)), dim3(256), 0, s, neq, in.A.ja,
                         av32, jo, wide);
      HIPCHK(hipGetLastError());
      int h_wide = 1;
      HIPCHK(hipMemcpyAsync(&h_wide, wide, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (!h_wide) A32.jo = jo;
    }
    const KState *final_state = st;
    if (gm) {
      MGmresArgs m{};
      GmresArgs &a = m.q;
      a.A = in.A; a.g = g; a.st = st; a.m = in.m; a.x = in.x; a.b = in.b;
      a.ld = fstride;
      double *sm = gsmall;
      a.h = sm; sm += (size_t)(k + 1) * k;
      a.cs = sm; sm += k + 2;
      a.sn = sm; sm += k + 2;
      a.e = sm; sm += k + 2;
      a.y = sm; sm += k + 2;
      a.nrm = sm; sm += k + 2;
      a.tol = in.tol; a.k = k; a.maxit = in.maxit; a.stop_rule = 1; a.orth = lag ? 2 : 1;
      m.A32 = A32;
      m.s = sc;
      m.gv = fv;
      if (lag) {
        m.gu = fv + (size_t)k * fstride;
        m.gw = fv + (size_t)(k + 1) * fstride;
        a.hraw = lagsmall;
        a.lag = lagsmall + (size_t)(k + 1) * k;
        a.lag_tot = a.lag + 2 * (size_t)k + 3;
        a.lag_part = lagpart;
        a.lag_kp = lag_kp(k);
        // as the FP64 lagged GMRES: per Arnoldi step D, R, B; a cycle's steps
        // enqueued in batches, stopped one batch after the device reports
        // the cycle done
        const void *mfn = (const void *)&mgmres_kernel<F>;
        for (int cyc = 0; cyc <= in.maxit; ++cyc) {
          a.phase = MG_START;
          if ((rc = launch(mfn, false, G, bs, m, s))) return rc;
          int j = 1, batch = 4, q = 0;
          bool waiting[2] = {false, false};
          while (j <= k + 1) {
            const int nb = std::min(batch, k + 2 - j);
            for (int t = 0; t < nb; ++t, ++j) {
              hipLaunchKernelGGL(mgmres_lag_d_kernel<F>, dim3(G), dim3(bs), 0, s, m);
              hipLaunchKernelGGL(gmres_lag_r_kernel<F>, dim3(2 * j), dim3(kLagBlock), 0, s, a);
              if (j <= k) hipLaunchKernelGGL(mgmres_lag_b_kernel<F>, dim3(G), dim3(bs), 0, s, m);
            }
            HIPCHK(hipGetLastError());
            const int slot = q & 1, prev = slot ^ 1;
            HIPCHK(hipMemcpyAsync(d.host_ring + slot, st, sizeof(KState), hipMemcpyDeviceToHost, s));
            HIPCHK(hipEventRecord(d.evq[slot], s));
            waiting[slot] = true;
            ++q;
            if (waiting[prev]) {
              HIPCHK(hipEventSynchronize(d.evq[prev]));
              waiting[prev] = false;
              if (d.host_ring[prev].inner_done || d.host_ring[prev].done) break;
            }
            batch = std::min(batch * 2, 8);
          }
          a.phase = MG_XUPD;
          if ((rc = launch(mfn, false, G, bs, m, s))) return rc;
          if ((rc = read_state(d, st, s))) return rc;
          if (d.host_st->done) break;
        }
        a.phase = MG_FINAL;
        if ((rc = launch(mfn, false, G, bs, m, s))) return rc;
      } else if ((rc = launch(pfn, true, G, bs, m, s))) {
        return rc;
      }
    } else {
      MBicgArgs a{};
      a.A = in.A; a.A32 = A32; a.g = g; a.st = st; a.s = sc; a.dsq = dsq; a.x = in.x; a.b = in.b;
      float **vecs[9] = {&a.r, &a.rh, &a.p0, &a.p1, &a.v0, &a.v1, &a.sv, &a.t, &a.y};
      for (int i = 0; i < 9; ++i) *vecs[i] = fv + (size_t)i * fstride;
      a.tol = in.tol; a.maxit = in.maxit; a.x0_zero = in.x0_zero; a.max_inner = in.refresh_every;
      if (persistent) {
        a.phase = -1;
        if ((rc = launch(pfn, true, G, bs, a, s))) return rc;
      } else {
        auto phase = [&](int ph) {
          a.phase = ph;
          a.d = dfr;
          dfr.par ^= 1;
          return launch(kfn, false, G, bs, a, s);
        };
        hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, 0, dfr.st, 2);
        if ((rc = phase(MB_SETUP))) return rc;
        // an iteration is two launches (P, Q), a reliable update two more
        const int cap = 4 * std::max(in.maxit, 0) + 16;
        if ((rc = drive_batches(d, s, cap, [&] { return phase(kPhaseNext); }, [&] { return dfr.st + dfr.par; })))
          return rc;
        if ((rc = phase(MB_FINAL))) return rc;
        if ((rc = phase(-1))) return rc;
        final_state = dfr.st + dfr.par;
      }
    }
    HIPCHK(hipEventRecord(d.ev1, s));
    if ((rc = read_state(d, const_cast<KState *>(final_state), s))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
    fill_stats(stats, *d.host_st, in.code, persistent, ms, G, bs,
               (g.tiles ? FVC_FLAG_TILES : 0) | (A32.jo ? FVC_FLAG_COL16 : 0));
    if (stats && !gm) stats->inner_iterations = d.host_st->cycle;   // FP64 residual recomputations
    return finish(*d.host_st, ctr);
  }
}

int dispatch_fp64(DevState &d, const SolveIn &in, fvc_stats *stats, hipStream_t s);

int dispatch_solve(DevState &d, const SolveIn &in, fvc_stats *stats, hipStream_t s) {
  if (in.precision == FVC_PREC_MIXED) {
    if (in.code != FVC_PCG) {
      switch (in.F) {
        case F_ELL4: return run_mixed_krylov<F_ELL4>(d, in, stats, s);
        case F_ELL3: return run_mixed_krylov<F_ELL3>(d, in, stats, s);
        case F_ELLW: return run_mixed_krylov<F_ELLW>(d, in, stats, s);
        case F_CSRD: return run_mixed_krylov<F_CSRD>(d, in, stats, s);
      }
      return fail(FVC_UNSUPPORTED, "mixed precision supports ELLPACK and unsymmetric CSRD storage");
    }
    switch (in.F) {
      case F_ELL4: return run_mixed<F_ELL4>(d, in, stats, s);
      case F_ELL3: return run_mixed<F_ELL3>(d, in, stats, s);
      case F_ELLW: return run_mixed<F_ELLW>(d, in, stats, s);
      case F_CSRD: return run_mixed<F_CSRD>(d, in, stats, s);
    }
    return fail(FVC_UNSUPPORTED, "mixed precision supports ELLPACK and unsymmetric CSRD storage");
  }
  if (in.precision != FVC_PREC_FP64) return fail(FVC_BAD_ARGUMENT, "precision %d", in.precision);
  if (in.stop_rule == FVC_STOP_RELRES && in.code != FVC_GMRES) {
    // The rule promises ||b - A x|| <= tol ||b|| for the true residual; over
    // many iterations the recursive residual drifts below the true one, so
    // the solve restarts from its x (a fresh r = b - A x) while the true
    // residual, computed in FP64 at the end of every solve, misses it.
    fvc_stats local, part;
    fvc_stats *st = stats ? stats : &local;
    int rc = dispatch_fp64(d, in, st, s);
    SolveIn again = in;
    again.x0_zero = 0;
    // A restart that does not halve the true residual has reached the FP64
    // floor of this system (a tolerance near machine precision): stop there;
    // stats.true_relres reports what was reached.
    double prev = 2.0 * st->true_relres;
    for (int restart = 0; restart < 8 && rc == FVC_OK && st->true_relres > in.tol &&
                          st->true_relres <= 0.5 * prev && st->iterations < in.maxit; ++restart) {
      prev = st->true_relres;
      again.maxit = in.maxit - st->iterations;
      rc = dispatch_fp64(d, again, &part, s);
      const int its = st->iterations + part.iterations;
      const double secs = st->seconds + part.seconds;
      *st = part;
      st->iterations = its;
      st->inner_iterations = its;
      st->seconds = secs;
    }
    return rc;
  }
  return dispatch_fp64(d, in, stats, s);
}

int dispatch_fp64(DevState &d, const SolveIn &in, fvc_stats *stats, hipStream_t s) {
  switch (in.F) {
    case F_ELL4: return run_solve<F_ELL4>(d, in, stats, s);
    case F_ELL3: return run_solve<F_ELL3>(d, in, stats, s);
    case F_ELLW: return run_solve<F_ELLW>(d, in, stats, s);
    case F_CSRD: return run_solve<F_CSRD>(d, in, stats, s);
    case F_CSR: return run_solve<F_CSR>(d, in, stats, s);
    case F_LSYM: return run_solve<F_LSYM>(d, in, stats, s);
    case F_CSRC: return run_solve<F_CSRC>(d, in, stats, s);
  }
  return fail(FVC_UNSUPPORTED, "no kernel for format %d", in.F);
}

// The reference's x0 choice per path (src/Solver.f:89-263, src/Iterativos.f:288).
bool x0_is_zero(int code, int matriz, int unsym, int rule) {
  if (code == FVC_GMRES) return true;
  if (rule == FVC_X0_ZERO) return true;
  if (rule == FVC_X0_GIVEN) return false;
  const bool omp = rule == FVC_X0_AS_OPENMP;
  if (matriz == FVC_ELLPACK) return false;
  if (matriz == FVC_CSR || matriz == FVC_CSRC) return true;
  // CSRD
  if (code == FVC_PCG) return !omp;   // pcg_omp init=.false. / pcg init=.true.
  return true;                        // pbicgstab[_omp] init=.true.
}

fvc_options defaults() {
  fvc_options o;
  std::memset(&o, 0, sizeof o);
  o.x0_rule = FVC_X0_AS_OPENMP;
  o.gmres_k = 15;
  o.stop_rule = FVC_STOP_REFERENCE;
  o.true_residual = 1;
  o.engine = FVC_ENGINE_AUTO;
  return o;
}

int check_common(int code, int matriz, int neq, const void *ad, const void *x, const void *b,
                 const void *iax, const void *ja) {
  if (code < FVC_PCG || code > FVC_GMRES) return fail(FVC_UNSUPPORTED, "solver code %d (1 PCG, 2 BiCGSTAB, 3 GMRES)", code);
  if (matriz < FVC_CSRD || matriz > FVC_ELLPACK) return fail(FVC_UNSUPPORTED, "matriz %d", matriz);
  if (neq <= 0) return fail(FVC_BAD_ARGUMENT, "neq = %d", neq);
  if (!ad || !x || !b || !iax || !ja) return fail(FVC_BAD_ARGUMENT, "null array argument");
  return FVC_OK;
}

// Builds the device matrix from host arrays (uploading them) or device
// arrays (used in place).  Owned uploads land in `own`.
// With `persist` set, uploads take the next slot of those device buffers
// (grown when too small) and outlive this object; otherwise they are freed
// with it.
struct Owned {
  std::vector<void *> ptrs;
  DevBufs *persist = nullptr;
  ~Owned() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int up(T **dst, const T *src, size_t n) {
    if (n == 0) { *dst = nullptr; return FVC_OK; }
    const size_t bytes = n * sizeof(T);
    void *p = nullptr;
    if (persist) {
      if (persist->next >= DevBufs::kSlots) return fail(FVC_DEVICE_ERROR, "out of upload slots");
      const int k = persist->next++;
      if (persist->cap[k] < bytes) {
        if (persist->p[k]) HIPCHK(hipFree(persist->p[k]));
        persist->p[k] = nullptr;
        persist->cap[k] = 0;
        HIPCHK(hipMalloc(&persist->p[k], bytes));
        persist->cap[k] = bytes;
      }
      p = persist->p[k];
    } else {
      HIPCHK(hipMalloc(&p, bytes));
      ptrs.push_back(p);
    }
    if (src) HIPCHK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = (T *)p;
    return FVC_OK;
  }
};

// Device (or managed) memory?  Plain host memory reports an error or an
// unregistered type; either way it is read directly.
bool is_device_pointer(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int build_device_matrix(int matriz, int unsym, int neq, int nad, const double *ad,
                        const double *au, const double *al, const int32_t *iax,
                        const int32_t *ja, bool on_device, Owned &own, MatDesc &md,
                        const double **m_out, hipStream_t s) {
  int w = 0;
  std::vector<int> h_ia;
  const int *ia_host = nullptr;
  if (matriz == FVC_ELLPACK) {
    int nl[2] = {0, 0};
    if (on_device && is_device_pointer(iax)) HIPCHK(hipMemcpy(nl, iax, sizeof nl, hipMemcpyDeviceToHost));
    else std::memcpy(nl, iax, sizeof nl);
    w = nl[0];
    if (w < 1 || w > 64) return fail(FVC_BAD_ARGUMENT, "ELLPACK width iax(1) = %d", w);
  } else {
    if (on_device) {
      h_ia.resize(neq + 1);
      HIPCHK(hipMemcpy(h_ia.data(), iax, (neq + 1) * sizeof(int), hipMemcpyDeviceToHost));
      ia_host = h_ia.data();
    } else {
      ia_host = iax;
    }
    if (ia_host[0] != 1) return fail(FVC_BAD_ARGUMENT, "ia(1) = %d, expected 1", ia_host[0]);
  }
  int F = 0;
  int rc = fmt_for(matriz, unsym, w, &F);
  if (rc) return rc;
  md.F = F;
  md.m.neq = neq;
  md.m.w = w;
  const size_t nval = matriz == FVC_ELLPACK ? (size_t)w * neq : (size_t)(ia_host[neq] - 1);
  md.nval = (int)nval;
  (void)nad;
  const double *vals = (matriz == FVC_CSR) ? ad : (F == F_CSRC ? al : (unsym ? (au ? au : al) : al));
  if (matriz == FVC_ELLPACK) vals = au ? au : al;
  if (!vals && nval > 0) return fail(FVC_BAD_ARGUMENT, "null off-diagonal array");
  if (F == F_CSRC && !au) return fail(FVC_BAD_ARGUMENT, "CSRC unsymmetric needs au");
  if (on_device) {
    md.m.ad = ad;
    md.m.av = (matriz == FVC_CSR) ? ad : vals;
    md.m.au = au;
    md.m.ia = matriz == FVC_ELLPACK ? nullptr : iax;
    md.m.ja = ja;
  } else {
    double *d_ad = nullptr, *d_av = nullptr, *d_au = nullptr;
    int *d_ia = nullptr, *d_ja = nullptr;
    if (matriz == FVC_CSR) {
      if ((rc = own.up(&d_av, ad, nval))) return rc;
      d_ad = d_av;
    } else {
      if ((rc = own.up(&d_ad, ad, (size_t)neq))) return rc;
      if ((rc = own.up(&d_av, vals, nval))) return rc;
    }
    if (F == F_CSRC && (rc = own.up(&d_au, au, nval))) return rc;
    if (matriz != FVC_ELLPACK && (rc = own.up(&d_ia, iax, (size_t)neq + 1))) return rc;
    if ((rc = own.up(&d_ja, ja, nval))) return rc;
    md.m.ad = d_ad;
    md.m.av = d_av;
    md.m.au = d_au;
    md.m.ia = d_ia;
    md.m.ja = d_ja;
  }
  if (F == F_LSYM || F == F_CSRC) {
    std::vector<int> hja;
    const int *ja_host = ja;
    if (on_device) {
      hja.resize(std::max<size_t>(nval, 1));
      if (nval) HIPCHK(hipMemcpy(hja.data(), ja, nval * sizeof(int), hipMemcpyDeviceToHost));
      ja_host = hja.data();
    }
    std::vector<int> tptr, tsrc, tpos;
    build_transpose(neq, ia_host, ja_host, tptr, tsrc, tpos);
    int *p1, *p2, *p3;
    if ((rc = own.up(&p1, tptr.data(), tptr.size()))) return rc;
    if ((rc = own.up(&p2, tsrc.data(), tsrc.size()))) return rc;
    if ((rc = own.up(&p3, tpos.data(), tpos.size()))) return rc;
    md.m.tptr = p1;
    md.m.tsrc = p2;
    md.m.tpos = p3;
  }
  // Jacobi preconditioner m = diag(A)
  if (matriz == FVC_CSR) {
    double *dm = nullptr;
    if ((rc = own.up(&dm, (const double *)nullptr, (size_t)neq))) return rc;
    // on the solve's stream, so the diagonal is written before the solve reads it
    hipLaunchKernelGGL(csr_diag_kernel, dim3((neq + 255) / 256), dim3(256), 0, s, neq,
                       md.m.ia, md.m.ja, md.m.av, dm);
    HIPCHK(hipGetLastError());
    *m_out = dm;
  } else {
    *m_out = md.m.ad;
  }
  return FVC_OK;
}

// Marks a device's solve as taking its workspace from the fixed array for the
// duration of one call (under the device's mutex).
struct NoallocScope {
  DevState *d;
  NoallocScope(DevState *dd, bool on) : d(dd) { d->noalloc = on; }
  ~NoallocScope() { d->noalloc = false; }
};

int solve_impl(int code, int matriz, int unsym, int neq, int nad, const double *ad,
               const double *au, const double *al, double *x, const double *b,
               const int32_t *iax, const int32_t *ja, double tol, int maxit,
               const fvc_options *opt, fvc_stats *stats, bool on_device, hipStream_t s,
               bool noalloc = false) {
  int rc = check_common(code, matriz, neq, ad, x, b, iax, ja);
  if (stats) { std::memset(stats, 0, sizeof *stats); stats->status = rc; stats->true_relres = -1; }
  if (rc) return rc;
  fvc_options o = opt ? *opt : defaults();
  if (o.gmres_k <= 0) o.gmres_k = 15;
  DevState *d = nullptr;
  if (noalloc) {
    // no set-up here: it allocates (pinned host words, events)
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64 || !g_dev[dev].ready)
      return fail(FVC_BAD_ARGUMENT, "fvc_solve_dev_noalloc: call fvc_init(%d) first on this device", dev);
    if (matriz != FVC_ELLPACK)
      return fail(FVC_UNSUPPORTED, "fvc_solve_dev_noalloc takes ELLPACK storage (matriz 4): the other "
                                   "formats build index arrays per call");
    d = &g_dev[dev];
  } else if ((rc = current_device(&d))) {
    return rc;
  }
  std::lock_guard<std::mutex> lk(d->mu);
  NoallocScope scope(d, noalloc);
  Owned own;
  MatDesc md;
  const double *m = nullptr;
  double *dx = x;
  const double *db = b;
  if (on_device) {
    if ((rc = build_device_matrix(matriz, unsym, neq, nad, ad, au, al, iax, ja, true, own, md, &m, s))) {
      if (stats) stats->status = rc;
      return rc;
    }
  } else {
    // host arrays: uploaded into the device's kept buffers, or not uploaded
    // at all when the caller kept the matrix of the previous host solve
    const bool keep = d->keep_next;
    d->keep_next = false;
    HostMatKey key;
    key.matriz = matriz;
    key.unsym = unsym ? 1 : 0;
    key.neq = neq;
    key.w = matriz == FVC_ELLPACK ? iax[0] : 0;
    key.nval = matriz == FVC_ELLPACK ? (size_t)key.w * neq : (size_t)std::max(iax[neq] - 1, 0);
    key.ad = ad; key.au = au; key.al = al; key.iax = iax; key.ja = ja;
    if (keep && d->mat_valid && key == d->mat_key) {
      md = d->mat_md;
      m = d->mat_m;
    } else {
      d->mat_valid = false;
      d->mat_bufs.next = 0;
      Owned kept;
      kept.persist = &d->mat_bufs;
      if ((rc = build_device_matrix(matriz, unsym, neq, nad, ad, au, al, iax, ja, false, kept, md, &m, s))) {
        if (stats) stats->status = rc;
        return rc;
      }
      d->mat_key = key;
      d->mat_md = md;
      d->mat_m = m;
      d->mat_valid = true;
    }
    d->vec_bufs.next = 0;
    Owned vecs;
    vecs.persist = &d->vec_bufs;
    double *tx = nullptr, *tb = nullptr;
    if ((rc = vecs.up(&tx, x, (size_t)neq))) return rc;
    if ((rc = vecs.up(&tb, b, (size_t)neq))) return rc;
    dx = tx;
    db = tb;
  }
  SolveIn in{};
  in.code = code;
  in.F = md.F;
  in.A = md.m;
  in.m = m;
  in.x = dx;
  in.b = db;
  in.tol = tol;
  in.maxit = maxit;
  in.x0_zero = x0_is_zero(code, matriz, unsym, o.x0_rule) ? 1 : 0;
  in.stop_rule = o.stop_rule;
  in.gmres_k = o.gmres_k;
  in.engine = o.engine;
  in.bs = o.block_size;
  in.G_req = o.grid_blocks;
  in.precision = o.precision;
  in.barrier = o.barrier;
  in.recurrence = o.recurrence;
  in.row_order = o.row_order;
  if (o.lane_exchange < 0 || o.lane_exchange > 2)
    return fail(FVC_BAD_ARGUMENT, "lane_exchange %d (0 default, 1 on, 2 off)", o.lane_exchange);
  // Off by default: at 1M cells the exchange made BiCGSTAB 2.86 against 2.72 ms
  // per solve and PCG 31.6 against 29.7 us per iteration on MI355X.  The
  // west/east gathers it saves hit the L1 cache (adjacent rows), and the
  // per-column selection splits a row's gathers into branches that no longer
  // issue together.
  in.lanex = o.lane_exchange == 1;
  if (o.gmres_orth != FVC_ORTH_MGS && o.gmres_orth != FVC_ORTH_CGS2 && o.gmres_orth != FVC_ORTH_CGS2_LAG)
    return fail(FVC_BAD_ARGUMENT, "gmres_orth %d (0 MGS, 1 CGS2, 2 lagged CGS2)", o.gmres_orth);
  in.orth = o.gmres_orth;
  if (o.row_order < 0 || o.row_order > 2)
    return fail(FVC_BAD_ARGUMENT, "row_order %d (0 default, 1 chunks, 2 XCD tiles)", o.row_order);
  in.refresh_every = o.refresh_every > 0 ? o.refresh_every : kRefreshEvery;
  in.refresh_drop = o.refresh_drop_e3 > 0 ? o.refresh_drop_e3 * 1e-3 : kRefreshDrop;
  if (in.refresh_drop >= 1.0) return fail(FVC_BAD_ARGUMENT, "refresh_drop_e3 %d must be below 1000", o.refresh_drop_e3);
  if (o.recurrence != FVC_REC_FUSED && o.recurrence != FVC_REC_REFERENCE)
    return fail(FVC_BAD_ARGUMENT, "recurrence %d (0 fused, 1 reference order)", o.recurrence);
  if (o.stop_rule != FVC_STOP_REFERENCE && o.stop_rule != FVC_STOP_RELRES)
    return fail(FVC_BAD_ARGUMENT, "stop_rule %d (0 reference, 1 true relative residual)", o.stop_rule);
  rc = dispatch_solve(*d, in, stats, s);
  if (!on_device && (rc == FVC_OK || rc == FVC_MAXIT || rc == FVC_BREAKDOWN || rc == FVC_NONFINITE))
    HIPCHK(hipMemcpy(x, dx, (size_t)neq * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

template <int F>
int spmv_launch(const Mat &A, const double *x, double *y, hipStream_t s) {
  if (A.neq > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d rows overflow 32-bit row indices", A.neq);
  hipLaunchKernelGGL(spmv_kernel<F>, dim3(xcd_blocks(A.neq, 256)), dim3(256), 0, s, A, x, y);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int matvec_impl(int matriz, int unsym, int neq, int nad, const double *ad, const double *au,
                const double *al, const int32_t *iax, const int32_t *ja, const double *x,
                double *y, bool on_device, hipStream_t s) {
  if (neq <= 0 || !ad || !iax || !ja || !x || !y) return fail(FVC_BAD_ARGUMENT, "bad matvec arguments");
  DevState *d = nullptr;
  int rc = current_device(&d);
  if (rc) return rc;
  Owned own;
  MatDesc md;
  const double *m = nullptr;
  if ((rc = build_device_matrix(matriz, unsym, neq, nad, ad, au, al, iax, ja, on_device, own, md, &m, s))) return rc;
  const double *dx = x;
  double *dy = y;
  if (!on_device) {
    double *tx = nullptr, *ty = nullptr;
    if ((rc = own.up(&tx, x, (size_t)neq))) return rc;
    if ((rc = own.up(&ty, (const double *)nullptr, (size_t)neq))) return rc;
    dx = tx;
    dy = ty;
  }
  switch (md.F) {
    case F_ELL4: rc = spmv_launch<F_ELL4>(md.m, dx, dy, s); break;
    case F_ELL3: rc = spmv_launch<F_ELL3>(md.m, dx, dy, s); break;
    case F_ELLW: rc = spmv_launch<F_ELLW>(md.m, dx, dy, s); break;
    case F_CSRD: rc = spmv_launch<F_CSRD>(md.m, dx, dy, s); break;
    case F_CSR: rc = spmv_launch<F_CSR>(md.m, dx, dy, s); break;
    case F_LSYM: rc = spmv_launch<F_LSYM>(md.m, dx, dy, s); break;
    case F_CSRC: rc = spmv_launch<F_CSRC>(md.m, dx, dy, s); break;
  }
  if (rc) return rc;
  if (!on_device) HIPCHK(hipMemcpy(y, dy, (size_t)neq * sizeof(double), hipMemcpyDeviceToHost));
  else if (!own.ptrs.empty()) HIPCHK(hipStreamSynchronize(s));   // temporaries die with `own`
  return FVC_OK;
}

int dot_impl(const double *a, const double *b, int n, double *result, bool on_device, hipStream_t s) {
  if (n < 0 || !result || (n > 0 && (!a || !b))) return fail(FVC_BAD_ARGUMENT, "bad dot arguments");
  if (n == 0) { *result = 0.0; return FVC_OK; }
  DevState *d = nullptr;
  int rc = current_device(&d);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(d->mu);
  Owned own;
  const double *da = a, *db = b;
  if (!on_device) {
    double *ta, *tb;
    if ((rc = own.up(&ta, a, (size_t)n))) return rc;
    if ((rc = own.up(&tb, b, (size_t)n))) return rc;
    da = ta;
    db = tb;
  }
  const int bs = kDefaultBlock;
  const int G = choose_blocks(n, d->cus, 2, 0);
  size_t bytes = align_up((kCtrWords + (size_t)G) * sizeof(unsigned)) + align_up(sizeof(KState)) +
                 align_up(kMaxK * (size_t)G * sizeof(double)) + 4096;
  if ((rc = ensure_arena(*d, bytes))) return rc;
  Carve cv{d->arena, bytes};
  unsigned *ctr = cv.take<unsigned>(kCtrWords + (size_t)G);
  KState *st = cv.take<KState>(1);
  double *partials = cv.take<double>(kMaxK * (size_t)G);
  double *out = cv.take<double>(1);
  if (!cv.fits()) return carve_overflow(cv);
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, (int)kCtrWords, st);
  Grid g = make_grid(n, G, ctr, partials, false);
  hipLaunchKernelGGL(dot_kernel, dim3(G), dim3(bs), 0, s, g, da, db, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(d->host_scalar, out, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *result = d->host_scalar[0];
  return FVC_OK;
}

// ------------------------------------------------------------- batches --

template <int F>
int run_batch(DevState &d, int nsys, const std::vector<int> &row_off, const std::vector<int> &val_off, int w,
              const double *ad, const double *av, const int *ia, const int *ja, double *x, const double *b,
              const int *d_row_off, const int *d_val_off, double tol, int maxit, const fvc_options &o,
              fvc_stats *stats, hipStream_t s) {
  const bool mixed = o.precision == FVC_PREC_MIXED;
  const int bs = o.block_size > 0 ? o.block_size : kDefaultBlock;
  if (bs % 64 != 0 || bs > 1024) return fail(FVC_BAD_ARGUMENT, "block size %d", bs);
  const void *fn = mixed ? (const void *)&mixed_batch_persistent<F> : (const void *)&pcg_batch_persistent<F>;
  int occ = 0;
  int rc = occupancy(fn, bs, &occ);
  if (rc) return rc;
  // every block of a batch must be co-resident: up to four per CU, as the
  // occupancy API allows
  const int cap = std::max(1, o.grid_blocks > 0 ? std::min(o.grid_blocks, occ * d.cus)
                                               : d.cus * std::max(1, std::min(occ, 4)));
  const int total = row_off[nsys];
  // systems go in launches ("waves") of at most cap/2 systems, or of
  // batch_launch_systems: each launch runs on all blocks, so fewer systems
  // per launch give each more blocks (a shorter chain of iterations for the
  // largest), at the cost of a launch tail each (the plan packs small
  // systems onto shared blocks)
  if (o.batch_launch_systems < 0)
    return fail(FVC_BAD_ARGUMENT, "batch_launch_systems %d", o.batch_launch_systems);
  const int wave = std::max(1, std::min(nsys, o.batch_launch_systems > 0 ? std::min(o.batch_launch_systems, 4096)
                                                                          : (cap / 2 > 0 ? cap / 2 : 1)));
  const double expo = o.batch_work_exp10 > 0 ? o.batch_work_exp10 / 10.0 : 1.5;
  if (o.batch_rows_per_wave < 0) return fail(FVC_BAD_ARGUMENT, "batch_rows_per_wave %d", o.batch_rows_per_wave);
  const int min_rows = o.batch_rows_per_wave > 0 ? o.batch_rows_per_wave * (bs / 64) : 64;
  const size_t n = (size_t)total;
  const size_t nv = (size_t)val_off[nsys];   // stored off-diagonal values
  size_t bytes = align_up(((size_t)wave * kCtrWords + cap) * sizeof(unsigned)) + align_up((size_t)wave * sizeof(KState)) +
                 align_up((size_t)kMaxK * cap * sizeof(double)) + align_up((size_t)wave * sizeof(SysDesc)) +
                 align_up((size_t)(2 * cap + 1 + wave) * sizeof(int)) + 4096;
  constexpr bool kOff16 = (F == F_ELL4 || F == F_ELL3);
  if (mixed) bytes += 2 * align_up(n * sizeof(float4)) + align_up(std::max<size_t>(nv, 1) * sizeof(float)) +
                      align_up(n * sizeof(double)) + align_up(n * sizeof(float)) +
                      (kOff16 ? align_up(n * w * sizeof(short)) + align_up(sizeof(int)) : 0);
  else bytes += 5 * align_up(n * sizeof(double));
  if ((rc = ensure_arena(d, bytes))) return rc;
  Carve cv{d.arena, bytes};
  unsigned *ctr = cv.take<unsigned>((size_t)wave * kCtrWords + (size_t)cap);
  KState *st = cv.take<KState>(wave);
  double *partials = cv.take<double>((size_t)kMaxK * cap);
  SysDesc *ddesc = cv.take<SysDesc>(wave);
  int *dplan = cv.take<int>((size_t)2 * cap + 1 + wave);   // block_ptr, then block_list
  BatchArgs B{};
  B.w = w;
  B.ad = ad;
  B.av = av;
  B.ia = ia;
  B.ja = ja;
  B.x = x;
  B.b = b;
  B.tol = tol;
  B.maxit = maxit;
  B.x0_zero = o.x0_rule == FVC_X0_GIVEN ? 0 : 1;
  B.stop_rule = o.stop_rule;
  B.mixed = mixed;
  B.max_inner = o.refresh_every > 0 ? o.refresh_every : kRefreshEvery;
  B.delta = o.refresh_drop_e3 > 0 ? o.refresh_drop_e3 * 1e-3 : kRefreshDrop;
  if (B.delta >= 1.0) return fail(FVC_BAD_ARGUMENT, "refresh_drop_e3 %d must be below 1000", o.refresh_drop_e3);
  if (mixed) {
    B.fs = align_up(n * sizeof(float4)) / sizeof(float4);
    B.rec = cv.take<float4>(2 * B.fs);
    B.av32 = cv.take<float>(std::max<size_t>(nv, 1));
    B.s = cv.take<double>(n);
    B.dsq = cv.take<float>(n);
    if (kOff16) {
      B.jo = cv.take<short>(n * w);
      B.wide = cv.take<int>(1);
    }
  } else {
    for (double *&v : B.v) v = cv.take<double>(n);
  }
  if (!cv.fits()) return carve_overflow(cv);
  B.ctr = ctr;
  B.flags = ctr + (size_t)wave * kCtrWords;
  B.flagbar = o.barrier == 0 ? 1 : 0;
  B.st = st;
  B.partials = partials;
  B.desc = ddesc;
  B.block_ptr = dplan;
  HIPCHK(hipEventRecord(d.ev0, s));
  if (mixed) {
    if (B.jo) hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned *>(B.wide), 1, st, 0);
    hipLaunchKernelGGL(batch_scale_kernel<F>, dim3((total + 255) / 256), dim3(256), 0, s, B, nsys, d_row_off,
                       d_val_off);
    HIPCHK(hipGetLastError());
    if (B.jo) {
      int h_wide = 1;
      HIPCHK(hipMemcpyAsync(&h_wide, B.wide, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (h_wide) B.jo = nullptr;   // keep the int32 columns
    }
  }
  std::vector<KState> host_st(nsys);
  for (int s0 = 0; s0 < nsys; s0 += wave) {
    const int m = std::min(wave, nsys - s0);
    std::vector<int> sizes(m);
    for (int i = 0; i < m; ++i) sizes[i] = row_off[s0 + i + 1] - row_off[s0 + i];
    const BlockPlan plan = plan_blocks(sizes, cap, expo, min_rows);
    std::vector<SysDesc> desc(m);
    for (int i = 0; i < m; ++i)
      desc[i] = SysDesc{row_off[s0 + i], sizes[i], plan.first[i], plan.count[i], val_off[s0 + i],
                        row_off[s0 + i] + s0 + i};
    const int G = plan.blocks;
    // block_ptr (G + 1 entries) then block_list (one per block of a group,
    // one per packed system: at most cap + m)
    std::vector<int> hplan(plan.ptr);
    hplan.insert(hplan.end(), plan.list.begin(), plan.list.end());
    if (G > cap || hplan.size() > (size_t)2 * cap + 1 + wave)
      return fail(FVC_DEVICE_ERROR, "batch plan of %d blocks for capacity %d", G, cap);
    B.block_list = dplan + G + 1;
    HIPCHK(hipMemcpyAsync(ddesc, desc.data(), m * sizeof(SysDesc), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dplan, hplan.data(), hplan.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(256), 0, s, ctr, wave * kCtrWords + cap, st, m);
    HIPCHK(hipGetLastError());
    if ((rc = launch(fn, true, G, bs, B, s))) return rc;
    HIPCHK(hipMemcpyAsync(host_st.data() + s0, st, m * sizeof(KState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));   // desc/hplan host buffers are reused next wave
    // every system's barrier words in one copy, then its error word
    std::vector<unsigned> hctr((size_t)m * kCtrWords);
    HIPCHK(hipMemcpy(hctr.data(), ctr, hctr.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
    for (int i = 0; i < m; ++i)
      if (hctr[(size_t)i * kCtrWords + kCtrErr]) return fail(FVC_DEVICE_ERROR, "batched solve: a grid barrier wait expired");
  }
  HIPCHK(hipEventRecord(d.ev1, s));
  HIPCHK(hipEventSynchronize(d.ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
  int worst = S_OK;
  for (int i = 0; i < nsys; ++i) {
    if (stats) {
      fill_stats(&stats[i], host_st[i], FVC_PCG, true, ms, 0, bs, B.jo ? FVC_FLAG_COL16 : 0);
      if (mixed) stats[i].inner_iterations = host_st[i].cycle;
    }
    if (host_st[i].status != S_OK && worst == S_OK) worst = host_st[i].status;
  }
  if (worst != S_OK) return fail(worst, "a system of the batch did not converge");
  return FVC_OK;
}

// Batched Jacobi-PCG over ELLPACK (matriz 4, width w) or CSRD (matriz 1,
// row pointers ia and value offsets val_offsets) systems.
int batch_impl(int code, int matriz, int nsys, const int32_t *row_offsets, const int32_t *val_offsets, int w,
               const double *ad, const double *al, double *x, const double *b, const int32_t *ia,
               const int32_t *ja, double tol, int maxit, const fvc_options *opt, fvc_stats *stats,
               bool on_device, hipStream_t s) {
  if (code != FVC_PCG) return fail(FVC_UNSUPPORTED, "batched solves run Jacobi-PCG (code 1)");
  const bool csrd = matriz == FVC_CSRD;
  if (nsys <= 0 || !row_offsets || !ad || !al || !x || !b || !ja || (csrd && (!ia || !val_offsets)))
    return fail(FVC_BAD_ARGUMENT, "bad batch arguments");
  if (!csrd && (w < 1 || w > 64)) return fail(FVC_BAD_ARGUMENT, "ELLPACK width %d", w);
  fvc_options o = opt ? *opt : defaults();
  DevState *d = nullptr;
  int rc = current_device(&d);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(d->mu);
  std::vector<int> row_off(nsys + 1), val_off(nsys + 1);
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
  HIPCHK(hipMemcpy(row_off.data(), row_offsets, (nsys + 1) * sizeof(int), kind));
  if (row_off[0] != 0) return fail(FVC_BAD_ARGUMENT, "row_offsets[0] must be 0");
  for (int i = 0; i < nsys; ++i)
    if (row_off[i + 1] <= row_off[i]) return fail(FVC_BAD_ARGUMENT, "system %d is empty", i);
  if (row_off[nsys] > INT_MAX - 1024 - nsys)
    return fail(FVC_UNSUPPORTED, "%d rows overflow 32-bit row indices", row_off[nsys]);
  if (csrd) {
    HIPCHK(hipMemcpy(val_off.data(), val_offsets, (nsys + 1) * sizeof(int), kind));
    if (val_off[0] != 0) return fail(FVC_BAD_ARGUMENT, "val_offsets[0] must be 0");
    for (int i = 0; i < nsys; ++i)
      if (val_off[i + 1] < val_off[i]) return fail(FVC_BAD_ARGUMENT, "val_offsets decrease at system %d", i);
  } else {
    if ((long long)row_off[nsys] * w > INT_MAX - 1024)
      return fail(FVC_UNSUPPORTED, "%d rows of width %d overflow 32-bit value indices", row_off[nsys], w);
    for (int i = 0; i <= nsys; ++i) val_off[i] = row_off[i] * w;
  }
  const size_t n = (size_t)row_off[nsys], nv = (size_t)val_off[nsys], nia = n + nsys;
  if (csrd && !on_device) {
    // each system's row pointer: 1 at its first row, nondecreasing, ending at
    // its value count + 1 (so no row reads outside its own values)
    for (int sys = 0; sys < nsys; ++sys) {
      const int32_t *p = ia + row_off[sys] + sys;
      const int ns = row_off[sys + 1] - row_off[sys];
      if (p[0] != 1 || p[ns] != val_off[sys + 1] - val_off[sys] + 1)
        return fail(FVC_BAD_ARGUMENT, "system %d: row pointer does not span its %d values", sys,
                    val_off[sys + 1] - val_off[sys]);
      for (int r = 0; r < ns; ++r)
        if (p[r + 1] < p[r]) return fail(FVC_BAD_ARGUMENT, "system %d: row pointer decreases at row %d", sys, r);
    }
  }
  Owned own;
  const double *dad = ad, *dav = al, *db = b;
  const int *dja = ja, *dia = ia, *droff = row_offsets, *dvoff = val_offsets;
  double *dx = x;
  if (!on_device) {
    double *t1, *t2, *t3, *t4;
    int *i1, *i2;
    if ((rc = own.up(&t1, ad, n)) || (rc = own.up(&t2, al, std::max<size_t>(nv, 1))) || (rc = own.up(&t3, b, n)) ||
        (rc = own.up(&t4, x, n)) || (rc = own.up(&i1, ja, std::max<size_t>(nv, 1))) ||
        (rc = own.up(&i2, (const int *)row_off.data(), (size_t)nsys + 1)))
      return rc;
    dad = t1; dav = t2; db = t3; dx = t4; dja = i1; droff = i2;
    if (csrd) {
      int *i3;
      if ((rc = own.up(&i3, ia, nia))) return rc;
      dia = i3;
    }
  } else if (o.precision == FVC_PREC_MIXED) {
    int *i2;
    if ((rc = own.up(&i2, (const int *)row_off.data(), (size_t)nsys + 1))) return rc;
    droff = i2;
  }
  if (csrd && o.precision == FVC_PREC_MIXED) {
    int *i4;
    if ((rc = own.up(&i4, (const int *)val_off.data(), (size_t)nsys + 1))) return rc;
    dvoff = i4;
  }
  if (csrd) {
    rc = run_batch<F_CSRD>(*d, nsys, row_off, val_off, 0, dad, dav, dia, dja, dx, db, droff, dvoff, tol, maxit, o,
                           stats, s);
  } else {
    switch (w) {
      case 4: rc = run_batch<F_ELL4>(*d, nsys, row_off, val_off, w, dad, dav, nullptr, dja, dx, db, droff, nullptr,
                                     tol, maxit, o, stats, s); break;
      case 3: rc = run_batch<F_ELL3>(*d, nsys, row_off, val_off, w, dad, dav, nullptr, dja, dx, db, droff, nullptr,
                                     tol, maxit, o, stats, s); break;
      default: rc = run_batch<F_ELLW>(*d, nsys, row_off, val_off, w, dad, dav, nullptr, dja, dx, db, droff,
                                      nullptr, tol, maxit, o, stats, s); break;
    }
  }
  if (!on_device && (rc == FVC_OK || rc == FVC_MAXIT || rc == FVC_BREAKDOWN || rc == FVC_NONFINITE))
    HIPCHK(hipMemcpy(x, dx, n * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

}  // namespace

// =================================================================== C ABI ==

extern "C" {

void fvc_default_options(fvc_options *opt) {
  if (opt) *opt = defaults();
}

int fvc_init(int32_t device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (n <= 0) return fail(FVC_DEVICE_ERROR, "no HIP device");
  if (device < 0 || device >= n) return fail(FVC_BAD_ARGUMENT, "device %d of %d", device, n);
  HIPCHK(hipSetDevice(device));
  DevState *d = nullptr;
  return current_device(&d);
}

const char *fvc_last_error(void) { return g_err.c_str(); }

int fvc_solver(const double *ad, const double *au, const double *al, double *x, double *b,
               const int32_t *iax, const int32_t *ja, int32_t neq, int32_t nad, int32_t unsym,
               double tol, int32_t maxit, int32_t code, int32_t matriz, int32_t nlit,
               int32_t omp_solver, fvc_stats *stats) {
  (void)nlit;
  fvc_options o = defaults();
  o.x0_rule = omp_solver ? FVC_X0_AS_OPENMP : FVC_X0_AS_SEQUENTIAL;
  return solve_impl(code, matriz, unsym, neq, nad, ad, au, al, x, b, iax, ja, tol, maxit, &o,
                    stats, false, nullptr);
}

int fvc_keep_matrix(int32_t keep) {
  DevState *d = nullptr;
  const int rc = current_device(&d);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(d->mu);
  d->keep_next = keep != 0;
  return FVC_OK;
}

int fvc_solve(int32_t code, int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
              const double *ad, const double *au, const double *al, double *x, const double *b,
              const int32_t *iax, const int32_t *ja, double tol, int32_t maxit,
              const fvc_options *opt, fvc_stats *stats) {
  return solve_impl(code, matriz, unsym, neq, nad, ad, au, al, x, b, iax, ja, tol, maxit, opt,
                    stats, false, nullptr);
}

int fvc_solve_dev(int32_t code, int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
                  const double *ad, const double *au, const double *al, double *x,
                  const double *b, const int32_t *iax, const int32_t *ja, double tol,
                  int32_t maxit, const fvc_options *opt, fvc_stats *stats, void *stream) {
  return solve_impl(code, matriz, unsym, neq, nad, ad, au, al, x, b, iax, ja, tol, maxit, opt,
                    stats, true, (hipStream_t)stream);
}

int fvc_solve_dev_noalloc(int32_t code, int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
                          const double *ad, const double *au, const double *al, double *x,
                          const double *b, const int32_t *iax, const int32_t *ja, double tol,
                          int32_t maxit, const fvc_options *opt, fvc_stats *stats, void *stream) {
  return solve_impl(code, matriz, unsym, neq, nad, ad, au, al, x, b, iax, ja, tol, maxit, opt,
                    stats, true, (hipStream_t)stream, true);
}

int fvc_solve_batch(int32_t code, int32_t nsys, const int32_t *row_offsets, int32_t w,
                    const double *ad, const double *al, double *x, const double *b,
                    const int32_t *ja, double tol, int32_t maxit, const fvc_options *opt,
                    fvc_stats *stats) {
  return batch_impl(code, FVC_ELLPACK, nsys, row_offsets, nullptr, w, ad, al, x, b, nullptr, ja, tol, maxit, opt,
                    stats, false, nullptr);
}

int fvc_solve_batch_dev(int32_t code, int32_t nsys, const int32_t *row_offsets, int32_t w,
                        const double *ad, const double *al, double *x, const double *b,
                        const int32_t *ja, double tol, int32_t maxit, const fvc_options *opt,
                        fvc_stats *stats, void *stream) {
  return batch_impl(code, FVC_ELLPACK, nsys, row_offsets, nullptr, w, ad, al, x, b, nullptr, ja, tol, maxit, opt,
                    stats, true, (hipStream_t)stream);
}

int fvc_solve_batch_csrd(int32_t code, int32_t nsys, const int32_t *row_offsets,
                         const int32_t *val_offsets, const double *ad, const double *a, double *x,
                         const double *b, const int32_t *ia, const int32_t *ja, double tol,
                         int32_t maxit, const fvc_options *opt, fvc_stats *stats) {
  return batch_impl(code, FVC_CSRD, nsys, row_offsets, val_offsets, 0, ad, a, x, b, ia, ja, tol, maxit, opt, stats,
                    false, nullptr);
}

int fvc_solve_batch_csrd_dev(int32_t code, int32_t nsys, const int32_t *row_offsets,
                             const int32_t *val_offsets, const double *ad, const double *a, double *x,
                             const double *b, const int32_t *ia, const int32_t *ja, double tol,
                             int32_t maxit, const fvc_options *opt, fvc_stats *stats, void *stream) {
  return batch_impl(code, FVC_CSRD, nsys, row_offsets, val_offsets, 0, ad, a, x, b, ia, ja, tol, maxit, opt, stats,
                    true, (hipStream_t)stream);
}

int fvc_ell_transpose_dev(int32_t neq, int32_t w, const int32_t *ja, const int32_t *back,
                          const double *al, double *alt, void *stream) {
  if (neq <= 0 || w < 1 || !ja || !back || !al || !alt) return fail(FVC_BAD_ARGUMENT, "bad transpose arguments");
  if (neq > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d rows overflow 32-bit row indices", neq);
  hipLaunchKernelGGL(ell_transpose_kernel, dim3(xcd_blocks(neq, 256)), dim3(256), 0, (hipStream_t)stream,
                     neq, w, ja, back, al, alt);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_grad_k_dev(int32_t neq, int32_t w, const int32_t *ja, const double *k, const double *t,
                   const double *lam, const double *geo, const double *alpha, const double *wall_g,
                   const double *wall_t, double *grad_k, void *stream) {
  if (neq <= 0 || w < 1 || !ja || !k || !t || !lam || !geo || !alpha || !wall_g || !wall_t || !grad_k)
    return fail(FVC_BAD_ARGUMENT, "bad gradient arguments");
  if (neq > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d cells overflow 32-bit indices", neq);
  hipLaunchKernelGGL(grad_k_kernel, dim3(xcd_blocks(neq, 256)), dim3(256), 0, (hipStream_t)stream, neq, w,
                     ja, k, t, lam, geo, alpha, wall_g, wall_t, grad_k);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_assemble_heat_dev(int32_t neq, int32_t w, const int32_t *ja, const double *k, const double *q,
                          const double *geo, const double *alpha, const double *conv, const double *conv_f,
                          const double *wall_g, const double *wall_t, double *ad, double *al, double *b,
                          void *stream) {
  if (neq <= 0 || w < 1 || !ja || !k || !geo || !alpha || !conv || !conv_f || !wall_g || !wall_t || !ad ||
      !al || !b)
    return fail(FVC_BAD_ARGUMENT, "bad assembly arguments");
  if (neq > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d cells overflow 32-bit indices", neq);
  hipLaunchKernelGGL(assemble_heat_kernel, dim3(xcd_blocks(neq, 256)), dim3(256), 0, (hipStream_t)stream, neq, w,
                     ja, k, q, geo, alpha, conv, conv_f, wall_g, wall_t, ad, al, b);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_gform_dev(int32_t numel, int32_t nnode, int32_t nen, const double *x, const int32_t *ix,
                  const int32_t *nelcon, const int32_t *pedge, const double *u, double *grad,
                  double *fluxl, void *stream) {
  if (numel <= 0 || nnode <= 0 || !x || !ix || !nelcon || !pedge || !u || !grad || !fluxl)
    return fail(FVC_BAD_ARGUMENT, "bad gform arguments");
  if (numel > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d cells overflow 32-bit indices", numel);
  const dim3 grid(xcd_blocks(numel, 256)), block(256);
  if (nen == 4)
    hipLaunchKernelGGL(gform_p_kernel<4>, grid, block, 0, (hipStream_t)stream, numel, nnode, x, ix, nelcon,
                       pedge, u, grad, fluxl);
  else if (nen == 3)
    hipLaunchKernelGGL(gform_p_kernel<3>, grid, block, 0, (hipStream_t)stream, numel, nnode, x, ix, nelcon,
                       pedge, u, grad, fluxl);
  else
    return fail(FVC_UNSUPPORTED, "gform: cells of %d edges (triangles and quads only)", nen);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_assemble_pc_dev(int32_t numel, int32_t nnode, int32_t nen, int32_t neq, int32_t nmat,
                        int32_t int_vel, const double *x, const int32_t *ix, const int32_t *nelcon,
                        const int32_t *pedge, const int32_t *num, const double *e, const double *sedge,
                        const double *w, const double *d, const double *gradp, const double *p,
                        double *ad, double *al, double *b, double *rescell, void *stream) {
  if (numel <= 0 || nnode <= 0 || neq <= 0 || nmat <= 0 || !x || !ix || !nelcon || !pedge || !num || !e ||
      !sedge || !w || !d || !gradp || !p || !ad || !al || !b || !rescell)
    return fail(FVC_BAD_ARGUMENT, "bad pressure-correction assembly arguments");
  if (int_vel != 1 && int_vel != 2 && int_vel != 5)
    return fail(FVC_UNSUPPORTED, "intVel %d (1 linear, 2 Rhie-Chow, 5 Ferziger-Peric)", int_vel);
  if (numel > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d cells overflow 32-bit indices", numel);
  const dim3 grid(xcd_blocks(numel, 256)), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (nen == 4)
    hipLaunchKernelGGL(assemble_pc_kernel<4>, grid, block, 0, s, numel, nnode, neq, nmat, int_vel, x, ix, nelcon,
                       pedge, num, e, sedge, w, d, gradp, p, ad, al, b, rescell);
  else if (nen == 3)
    hipLaunchKernelGGL(assemble_pc_kernel<3>, grid, block, 0, s, numel, nnode, neq, nmat, int_vel, x, ix, nelcon,
                       pedge, num, e, sedge, w, d, gradp, p, ad, al, b, rescell);
  else
    return fail(FVC_UNSUPPORTED, "pressure-correction assembly: cells of %d edges (triangles and quads only)", nen);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_assemble_mom_dev(int32_t numel, int32_t nnode, int32_t nen, int32_t neq, int32_t nmat, int32_t comp,
                         double dt, double under_u, int32_t s_pressure, int32_t simple_c, int32_t bs,
                         const double *x, const int32_t *ix, const int32_t *nelcon, const int32_t *pedge,
                         const int32_t *num, const double *e, const double *sedge, const double *u,
                         const double *pres, const double *w, const double *grad1, const double *grad2,
                         const double *gradp, double *d, double *ad, double *al, double *b, double *rescell,
                         double *im, void *stream) {
  if (numel <= 0 || nnode <= 0 || neq <= 0 || nmat <= 0 || !x || !ix || !nelcon || !pedge || !num || !e ||
      !sedge || !u || !pres || !w || !grad1 || !grad2 || !gradp || !d || !ad || !al || !b || !rescell || !im)
    return fail(FVC_BAD_ARGUMENT, "bad momentum assembly arguments");
  if (comp != 1 && comp != 2) return fail(FVC_BAD_ARGUMENT, "momentum component %d (1 or 2)", comp);
  if (!(dt != 0.0) || !(under_u > 0.0))
    return fail(FVC_BAD_ARGUMENT, "momentum assembly: dt %g, underU %g", dt, under_u);
  if (numel > INT_MAX - 2048) return fail(FVC_UNSUPPORTED, "%d cells overflow 32-bit indices", numel);
  MomArgs A{numel, nnode, neq, nmat, comp - 1, s_pressure ? 1 : 0, simple_c ? 1 : 0, bs ? 1 : 0, under_u, dt,
            x, ix, nelcon, pedge, num, e, sedge, u, pres, w, grad1, grad2, gradp, d, ad, al, b, rescell, im};
  const dim3 grid(xcd_blocks(numel, 256)), block(256);
  hipStream_t s = (hipStream_t)stream;
  if (nen == 4)
    hipLaunchKernelGGL(assemble_mom_kernel<4>, grid, block, 0, s, A);
  else if (nen == 3)
    hipLaunchKernelGGL(assemble_mom_kernel<3>, grid, block, 0, s, A);
  else
    return fail(FVC_UNSUPPORTED, "momentum assembly: cells of %d edges (triangles and quads only)", nen);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_simple_update_dev(int32_t numel, int32_t ndm, double *w, double *pre, const double *u1,
                          const double *u2, const double *prec, const double *gradpc, const double *d,
                          double under_u, double under_pc, int32_t simple_r, void *stream) {
  (void)under_u;
  if (numel <= 0 || !w || !pre || !u1 || !u2 || !prec || !gradpc || !d)
    return fail(FVC_BAD_ARGUMENT, "bad simpleUpdate arguments");
  if (ndm != 2) return fail(FVC_UNSUPPORTED, "simpleUpdate: ndm = %d (2-D only)", ndm);
  hipLaunchKernelGGL(simple_update_kernel, dim3((numel + 255) / 256), dim3(256), 0, (hipStream_t)stream, numel,
                     w, pre, u1, u2, prec, gradpc, d, under_pc, simple_r ? 1 : 0);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_matvec(int32_t matriz, int32_t unsym, int32_t neq, int32_t nad, const double *ad,
               const double *au, const double *al, const int32_t *iax, const int32_t *ja,
               const double *x, double *y) {
  return matvec_impl(matriz, unsym, neq, nad, ad, au, al, iax, ja, x, y, false, nullptr);
}

int fvc_matvec_dev(int32_t matriz, int32_t unsym, int32_t neq, int32_t nad, const double *ad,
                   const double *au, const double *al, const int32_t *iax, const int32_t *ja,
                   const double *x, double *y, void *stream) {
  return matvec_impl(matriz, unsym, neq, nad, ad, au, al, iax, ja, x, y, true, (hipStream_t)stream);
}

int fvc_dot(const double *a, const double *b, int32_t n, double *result) {
  return dot_impl(a, b, n, result, false, nullptr);
}

int fvc_dot_dev(const double *a, const double *b, int32_t n, double *result, void *stream) {
  return dot_impl(a, b, n, result, true, (hipStream_t)stream);
}

int fvc_getres(double *u, double *x, const int32_t *num, int32_t neq, int32_t numel, int32_t flag) {
  if (!u || !x || !num || numel < 0) return fail(FVC_BAD_ARGUMENT, "bad getres arguments");
  for (int32_t i = 0; i < numel; ++i) {
    const int32_t e = num[i];
    if (e < 1 || e > neq) return fail(FVC_BAD_ARGUMENT, "num(%d) = %d outside 1..%d", i + 1, e, neq);
  }
  if (flag) for (int32_t i = 0; i < numel; ++i) x[num[i] - 1] = u[i];
  else for (int32_t i = 0; i < numel; ++i) u[i] = x[num[i] - 1];
  return FVC_OK;
}

int fvc_getres_dev(double *u, double *x, const int32_t *num, int32_t neq, int32_t numel,
                   int32_t flag, void *stream) {
  (void)neq;
  if (!u || !x || !num || numel < 0) return fail(FVC_BAD_ARGUMENT, "bad getres arguments");
  if (numel == 0) return FVC_OK;
  hipLaunchKernelGGL(getres_kernel, dim3((numel + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     u, x, num, numel, flag);
  HIPCHK(hipGetLastError());
  return FVC_OK;
}

int fvc_malloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(FVC_BAD_ARGUMENT, "null pointer");
  HIPCHK(hipMalloc(ptr, bytes));
  return FVC_OK;
}
int fvc_free(void *ptr) {
  HIPCHK(hipFree(ptr));
  return FVC_OK;
}
int fvc_memcpy_htod(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FVC_OK;
}
int fvc_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FVC_OK;
}
int fvc_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
  if (!dst || !src) return fail(FVC_BAD_ARGUMENT, "null pointer");
  if (bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0)
    return fail(FVC_BAD_ARGUMENT, "fvc_memcpy_dtod needs 16-byte aligned pointers and sizes");
  DevState *d = nullptr;
  int rc = current_device(&d);
  if (rc) return rc;
  const size_t n16 = bytes / 16;
  const int blocks = (int)std::min<size_t>((n16 + 255) / 256, (size_t)d->cus * 16);
  if (n16) {
    hipLaunchKernelGGL(copy_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, (hipStream_t)stream,
                       (const uint4 *)src, (uint4 *)dst, n16);
    HIPCHK(hipGetLastError());
  }
  return FVC_OK;
}

int fvc_memset_dev(void *dst, int value, size_t bytes) {
  HIPCHK(hipMemset(dst, value, bytes));
  return FVC_OK;
}
int fvc_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return FVC_OK;
}
int fvc_device_info(int32_t *compute_units, char *name, int32_t name_len) {
  DevState *d = nullptr;
  int rc = current_device(&d);
  if (rc) return rc;
  if (compute_units) *compute_units = d->cus;
  if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", d->name);
  return FVC_OK;
}
int fvc_event_create(void **ev) {
  HIPCHK(hipEventCreate((hipEvent_t *)ev));
  return FVC_OK;
}
int fvc_event_destroy(void *ev) {
  HIPCHK(hipEventDestroy((hipEvent_t)ev));
  return FVC_OK;
}
int fvc_event_record(void *ev, void *stream) {
  HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return FVC_OK;
}
int fvc_event_elapsed_ms(void *start, void *stop, float *ms) {
  HIPCHK(hipEventSynchronize((hipEvent_t)stop));
  HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return FVC_OK;
}

}  // extern "C"
