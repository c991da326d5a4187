// fvc_host.h -- the host-only algorithms of the library, free of HIP so that
// they also build with the CPU sanitizers (tests/sanitize): the carving of a
// workspace into aligned pieces, the transposed-gather index of a
// lower-triangle pattern, and the block plan of a batch.
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

// The block plan of a batch.  A system's work is taken as neq^expo (PCG
// work ~ rows x iterations, iterations ~ sqrt(rows) on these meshes: expo
// 1.5), and every block should carry the same work, so that the systems
// finish together.  A system worth at least one block's share gets a group
// of its own, sized to minimise the largest work per block (blocks go one at
// a time to the group with the most work per block, at most one per
// min_rows_per_block rows).  The systems worth less than a block share
// blocks: each such block runs its systems one after another, and they are
// packed largest first onto the least loaded of nb blocks, with nb chosen so
// that the most loaded block of the whole plan carries as little as it can.
// Without the packing every small system would hold a block for the whole
// batch while doing a fraction of a block's work.
struct BlockPlan {
  std::vector<int> first, count;   // per system: its group's first block and size
  std::vector<int> ptr, list;      // block b runs systems list[ptr[b] .. ptr[b+1]) in turn
  int blocks = 0;
  double max_load = 0.0;           // the largest work per block, in neq^expo units
};

namespace detail {

// Largest work per block when the systems idx (work w) get groups within
// `blocks` blocks: one block each, then one more at a time to the largest
// w / g below its limit.  g must hold idx.size() entries.
inline double minimax_groups(const std::vector<double> &w, const std::vector<int> &lim, const std::vector<int> &idx,
                             int blocks, std::vector<int> &g) {
  const int m = (int)idx.size();
  std::vector<std::pair<double, int>> heap;
  heap.reserve(m);
  for (int k = 0; k < m; ++k) {
    g[k] = 1;
    heap.push_back({w[idx[k]], k});
  }
  std::make_heap(heap.begin(), heap.end());
  for (int left = blocks - m; left > 0 && !heap.empty();) {
    std::pop_heap(heap.begin(), heap.end());
    const int k = heap.back().second;
    heap.pop_back();
    if (g[k] >= lim[idx[k]]) continue;   // at its limit: leave it out of the heap
    ++g[k];
    --left;
    heap.push_back({w[idx[k]] / g[k], k});
    std::push_heap(heap.begin(), heap.end());
  }
  double mx = 0.0;
  for (int k = 0; k < m; ++k) mx = std::max(mx, w[idx[k]] / g[k]);
  return mx;
}

// Largest load when the systems idx (sorted by work, largest first) go one
// at a time onto the least loaded of nb blocks; bin[k] receives the block of
// idx[k].
inline double pack_lpt(const std::vector<double> &w, const std::vector<int> &idx, int nb, std::vector<int> &bin) {
  std::vector<std::pair<double, int>> heap;   // (-load, block): a min-heap on load
  heap.reserve(nb);
  for (int b = 0; b < nb; ++b) heap.push_back({-0.0, -b});
  std::make_heap(heap.begin(), heap.end());
  double mx = 0.0;
  for (size_t k = 0; k < idx.size(); ++k) {
    std::pop_heap(heap.begin(), heap.end());
    auto &h = heap.back();
    bin[k] = -h.second;
    h.first -= w[idx[k]];
    mx = std::max(mx, -h.first);
    std::push_heap(heap.begin(), heap.end());
  }
  return mx;
}

}  // namespace detail

inline BlockPlan plan_blocks(const std::vector<int> &neq, int cap, double expo, int min_rows_per_block = 64) {
  const int n = (int)neq.size();
  cap = std::max(cap, 1);
  const long long mr = std::max(min_rows_per_block, 1);
  std::vector<double> w(n);
  std::vector<int> lim(n);
  double tot = 0.0;
  for (int i = 0; i < n; ++i) {
    tot += (w[i] = std::pow((double)std::max(neq[i], 1), expo));
    lim[i] = (int)std::max(1LL, ((long long)std::max(neq[i], 1) + mr - 1) / mr);
  }
  const double share = tot / cap;
  std::vector<int> big, small;
  for (int i = 0; i < n; ++i) (w[i] >= share ? big : small).push_back(i);
  // at most cap groups (w >= tot / cap holds for at most cap systems)
  std::stable_sort(small.begin(), small.end(), [&](int a, int b) { return w[a] > w[b]; });
  const int nbig = (int)big.size(), nsmall = (int)small.size();
  std::vector<int> g(nbig), bin(nsmall), best_g(nbig), best_bin(nsmall);
  int best_nb = 0;
  double best = -1.0;
  if (nsmall == 0) {
    best = detail::minimax_groups(w, lim, big, cap, best_g);
  } else {
    const int hi = std::min(nsmall, cap - nbig);   // >= 1: nbig < cap when a system is small
    // the packed blocks' load falls as nb grows and the groups' load rises:
    // bisect for the first nb where the packed load is no longer the larger,
    // then take the best nb around it
    auto cost = [&](int nb, double &lb) {
      lb = detail::pack_lpt(w, small, nb, bin);
      const double lg = nbig ? detail::minimax_groups(w, lim, big, cap - nb, g) : 0.0;
      return std::max(lb, lg);
    };
    int lo = 1, up = hi;
    while (lo < up) {
      const int mid = (lo + up) / 2;
      double lb;
      const double c = cost(mid, lb);
      if (lb < c || (lb == c && nbig == 0)) up = mid;   // the groups dominate (or nothing else)
      else lo = mid + 1;
    }
    // ... and hi, one block per small system, so that sharing never loses
    // to not sharing
    std::vector<int> cand;
    for (int nb = std::max(1, lo - 3); nb <= std::min(hi, lo + 3); ++nb) cand.push_back(nb);
    if (cand.back() != hi) cand.push_back(hi);
    for (int nb : cand) {
      double lb;
      const double c = cost(nb, lb);
      if (best < 0.0 || c < best) {
        best = c;
        best_nb = nb;
        best_g = g;
        best_bin = bin;
      }
    }
  }
  BlockPlan p;
  p.first.assign(n, 0);
  p.count.assign(n, 1);
  p.max_load = best;
  p.ptr.push_back(0);
  int b = 0;
  for (int k = 0; k < nbig; ++k) {
    const int i = big[k];
    p.first[i] = b;
    p.count[i] = best_g[k];
    for (int q = 0; q < best_g[k]; ++q, ++b) {
      p.list.push_back(i);
      p.ptr.push_back((int)p.list.size());
    }
  }
  std::vector<std::vector<int>> bins(best_nb);
  for (int k = 0; k < nsmall; ++k) bins[best_bin[k]].push_back(small[k]);
  for (const auto &q : bins) {
    for (int i : q) {
      p.first[i] = b;
      p.list.push_back(i);
    }
    p.ptr.push_back((int)p.list.size());
    ++b;
  }
  p.blocks = b;
  return p;
}

}  // namespace fvc
