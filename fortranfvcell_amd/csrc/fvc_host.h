// fvc_host.h -- the host-only algorithms of the library, free of HIP so that
// they also build with the CPU sanitizers (tests/sanitize): the carving of a
// workspace into aligned pieces, the transposed-gather index of a
// lower-triangle pattern, and the block allocation of a batch.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <utility>
#include <vector>

namespace fvc {

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Carves a workspace of cap bytes into 256-byte aligned pieces.  take()
// only computes addresses (the workspace is device memory); the caller
// checks fits() once every piece is taken, before anything is launched.
struct Carve {
  char *base;
  size_t cap;
  size_t off = 0;
  template <class T>
  T *take(size_t n) {
    T *p = reinterpret_cast<T *>(base + off);
    off = align_up(off + n * sizeof(T));
    return p;
  }
  bool fits() const { return off <= cap; }
};

// Transposed-gather structure of a lower-triangle pattern: for each row c,
// the entries (row i > c, position k) with ja(k) = c, in increasing i.
inline void build_transpose(int neq, const int *ia, const int *ja, std::vector<int> &tptr,
                     std::vector<int> &tsrc, std::vector<int> &tpos) {
  const int nnz = ia[neq] - ia[0];
  tptr.assign(neq + 1, 0);
  for (int k = 0; k < nnz; ++k) {
    const int c = ja[k] - 1;
    if (c >= 0 && c < neq) tptr[c + 1]++;
  }
  for (int i = 0; i < neq; ++i) tptr[i + 1] += tptr[i];
  tsrc.assign(std::max(nnz, 1), 0);
  tpos.assign(std::max(nnz, 1), 0);
  std::vector<int> fill(tptr.begin(), tptr.end() - 1);
  for (int i = 0; i < neq; ++i)
    for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
      const int c = ja[k] - 1;
      if (c < 0 || c >= neq) continue;
      const int e = fill[c]++;
      tsrc[e] = i;
      tpos[e] = k;
    }
}

// Block groups per system: g_s proportional to neq_s^exp (PCG work ~ rows x
// iterations, iterations ~ sqrt(rows) on these meshes), so that the systems
// finish together; at least one block each, at most one per
// min_rows_per_block rows, the total within the co-resident capacity.  Shares are rounded by largest
// remainder, so the large systems that set the batch's time are not starved
// by the one-block floor of the small ones.
inline std::vector<int> allocate_blocks(const std::vector<int> &neq, int cap, double expo,
                                        int min_rows_per_block = 64) {
  const int n = (int)neq.size();
  const long long mr = std::max(min_rows_per_block, 1);
  std::vector<double> w(n);
  double tot = 0;
  for (int i = 0; i < n; ++i) tot += (w[i] = std::pow((double)std::max(neq[i], 1), expo));
  std::vector<int> g(n, 1), lim(n);
  for (int i = 0; i < n; ++i) lim[i] = (int)std::max(1LL, ((long long)neq[i] + mr - 1) / mr);
  int used = n;
  if (cap > n) {
    // systems whose share is below one block keep one; the rest share what is left
    double rest = tot;
    int avail = cap;
    std::vector<char> floor1(n, 0);
    for (bool changed = true; changed;) {
      changed = false;
      for (int i = 0; i < n; ++i)
        if (!floor1[i] && (w[i] / rest) * avail < 1.0) {
          floor1[i] = 1;
          rest -= w[i];
          --avail;
          changed = true;
        }
      if (rest <= 0.0) break;
    }
    std::vector<std::pair<double, int>> frac;
    used = 0;
    for (int i = 0; i < n; ++i) {
      if (floor1[i]) {
        g[i] = 1;
      } else {
        const double t = (w[i] / rest) * avail;
        g[i] = std::min(lim[i], std::max(1, (int)std::floor(t)));
        if (g[i] < lim[i]) frac.push_back({t - std::floor(t), i});
      }
      used += g[i];
    }
    std::sort(frac.begin(), frac.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    for (const auto &f : frac) {
      if (used >= cap) break;
      ++g[f.second];
      ++used;
    }
  }
  return g;
}

}  // namespace fvc
