// host_checks.cpp -- the library's host-only algorithms (fortranfvcell_amd/
// csrc/fvc_host.h) under AddressSanitizer and UndefinedBehaviorSanitizer:
// built and run by tests/test_sanitize.py with g++ on the CPU.  Each check
// compares against a brute-force statement of what the algorithm promises.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../fortranfvcell_amd/csrc/fvc_host.h"

using namespace fvc;

static int failures = 0;
#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);          \
      std::fprintf(stderr, "\n");                 \
      ++failures;                                 \
    }                                             \
  } while (0)

// A random strictly-lower pattern, 1-based as the solver's CSRC input: row i
// holds up to maxw distinct columns below i, in increasing order.
static void lower_pattern(std::mt19937 &rng, int neq, int maxw, std::vector<int> &ia, std::vector<int> &ja) {
  ia.assign(neq + 1, 1);
  ja.clear();
  for (int i = 0; i < neq; ++i) {
    std::vector<char> used(i, 0);
    const int w = i == 0 ? 0 : (int)(rng() % (unsigned)(std::min(maxw, i) + 1));
    for (int t = 0; t < w; ++t) used[rng() % (unsigned)i] = 1;
    for (int c = 0; c < i; ++c)
      if (used[c]) ja.push_back(c + 1);
    ia[i + 1] = (int)ja.size() + 1;
  }
}

static void check_transpose(std::mt19937 &rng, int neq, int maxw) {
  std::vector<int> ia, ja, tptr, tsrc, tpos;
  lower_pattern(rng, neq, maxw, ia, ja);
  build_transpose(neq, ia.data(), ja.empty() ? nullptr : ja.data(), tptr, tsrc, tpos);
  const int nnz = (int)ja.size();
  CHECK((int)tptr.size() == neq + 1 && tptr[0] == 0 && tptr[neq] == nnz, "tptr bounds neq=%d", neq);
  // brute force: column c's entries are the (i, k) with ja(k) = c, i ascending
  for (int c = 0; c < neq; ++c) {
    std::vector<std::pair<int, int>> want;
    for (int i = 0; i < neq; ++i)
      for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k)
        if (ja[k] - 1 == c) want.push_back({i, k});
    CHECK(tptr[c + 1] - tptr[c] == (int)want.size(), "column %d count", c);
    if (tptr[c + 1] - tptr[c] != (int)want.size()) continue;
    for (size_t e = 0; e < want.size(); ++e) {
      CHECK(tsrc[tptr[c] + e] == want[e].first && tpos[tptr[c] + e] == want[e].second,
            "column %d entry %zu", c, e);
    }
  }
}

static void check_blocks(std::mt19937 &rng, int nsys, int cap) {
  std::vector<int> neq(nsys);
  for (int &v : neq) v = 1 + (int)(rng() % 200000u) * ((rng() & 3) ? 1 : 0) + (int)(rng() % 64u);
  for (double expo : {0.5, 1.0, 1.5, 2.0}) {
    const BlockPlan p = plan_blocks(neq, cap, expo);
    CHECK((int)p.first.size() == nsys && (int)p.count.size() == nsys, "size");
    CHECK(p.blocks >= 1 && p.blocks <= std::max(cap, 1), "%d blocks for capacity %d", p.blocks, cap);
    CHECK((int)p.ptr.size() == p.blocks + 1 && p.ptr[0] == 0 && p.ptr.back() == (int)p.list.size(), "ptr");
    if ((int)p.ptr.size() != p.blocks + 1) continue;
    // every system runs exactly once per block of its group: a group of g
    // blocks is g blocks running that system alone; a packed system is one
    // entry in its shared block's list
    std::vector<int> seen(nsys, 0);
    std::vector<double> load(p.blocks, 0.0);
    for (int b = 0; b < p.blocks; ++b) {
      CHECK(p.ptr[b] < p.ptr[b + 1], "block %d runs nothing", b);
      for (int k = p.ptr[b]; k < p.ptr[b + 1]; ++k) {
        const int i = p.list[k];
        CHECK(i >= 0 && i < nsys, "system %d", i);
        if (i < 0 || i >= nsys) continue;
        ++seen[i];
        CHECK(b >= p.first[i] && b < p.first[i] + p.count[i], "system %d on block %d outside its group", i, b);
        if (p.count[i] > 1) CHECK(p.ptr[b + 1] - p.ptr[b] == 1, "group block %d shared", b);
        load[b] += std::pow((double)neq[i], expo) / p.count[i];
      }
    }
    double mx = 0.0, tot = 0.0;
    for (int i = 0; i < nsys; ++i) {
      CHECK(seen[i] == p.count[i], "system %d runs on %d blocks, group of %d", i, seen[i], p.count[i]);
      const int lim = std::max(1, (neq[i] + 63) / 64);
      CHECK(p.count[i] >= 1 && p.count[i] <= lim, "system %d: %d blocks for %d rows", i, p.count[i], neq[i]);
      tot += std::pow((double)neq[i], expo);
    }
    for (double l : load) mx = std::max(mx, l);
    CHECK(std::fabs(mx - p.max_load) <= 1e-9 * mx, "reported load %g, actual %g", p.max_load, mx);
    // no plan beats the average load, nor a system's work over its limit
    CHECK(mx >= tot / std::max(cap, 1) * (1 - 1e-12), "load %g below the average %g", mx, tot / cap);
    // against the plan without sharing (each system its own group, blocks
    // one at a time to the largest work per block): never worse when every
    // system can have a block
    if (cap >= nsys) {
      std::vector<double> w(nsys);
      std::vector<int> lim(nsys), idx(nsys), g(nsys);
      for (int i = 0; i < nsys; ++i) {
        w[i] = std::pow((double)neq[i], expo);
        lim[i] = std::max(1, (neq[i] + 63) / 64);
        idx[i] = i;
      }
      const double solo = detail::minimax_groups(w, lim, idx, cap, g);
      CHECK(mx <= solo * (1 + 1e-12), "plan load %g above the unshared plan's %g", mx, solo);
    }
  }
  // brute force on tiny cases: the groups of the unshared plan are optimal
  for (int t = 0; t < 50; ++t) {
    const int n = 1 + (int)(rng() % 3u), c = n + (int)(rng() % 5u);
    std::vector<double> w(n);
    std::vector<int> lim(n), idx(n), g(n);
    for (int i = 0; i < n; ++i) {
      w[i] = 1.0 + (double)(rng() % 1000u);
      lim[i] = 1 + (int)(rng() % 4u);
      idx[i] = i;
    }
    const double got = detail::minimax_groups(w, lim, idx, c, g);
    double best = 1e300;
    std::vector<int> q(n, 1);
    for (;;) {
      int used = 0;
      double m = 0.0;
      for (int i = 0; i < n; ++i) {
        used += q[i];
        m = std::max(m, w[i] / q[i]);
      }
      if (used <= c) best = std::min(best, m);
      int i = 0;
      while (i < n && ++q[i] > lim[i]) q[i++] = 1;
      if (i == n) break;
    }
    CHECK(std::fabs(got - best) <= 1e-12 * best, "minimax %g, brute force %g", got, best);
  }
}

static void check_carve(std::mt19937 &rng) {
  for (int t = 0; t < 200; ++t) {
    std::vector<size_t> sizes(1 + rng() % 12);
    size_t want = 0;
    for (size_t &s : sizes) {
      s = rng() % 5000;
      want += align_up(s * sizeof(double));
    }
    std::vector<char> buf(want + 256);
    char *base = buf.data() + (256 - ((uintptr_t)buf.data() % 256)) % 256;
    Carve cv{base, want};
    std::vector<double *> p;
    for (size_t s : sizes) p.push_back(cv.take<double>(s));
    CHECK(cv.fits() && cv.off == want, "carve %zu of %zu", cv.off, want);
    for (size_t i = 0; i < sizes.size(); ++i) {
      CHECK(((uintptr_t)p[i] - (uintptr_t)base) % 256 == 0, "alignment of piece %zu", i);
      // every piece can be written in full without touching the next
      for (size_t k = 0; k < sizes[i]; ++k) p[i][k] = (double)i;
    }
    for (size_t i = 0; i < sizes.size(); ++i)
      for (size_t k = 0; k < sizes[i]; ++k) CHECK(p[i][k] == (double)i, "piece %zu overwritten", i);
    Carve small{base, want - 1};
    for (size_t s : sizes) small.take<double>(s);
    CHECK(want == 0 || !small.fits(), "overflow not reported");
  }
}

int main() {
  std::mt19937 rng(20261203u);
  for (int neq : {1, 2, 3, 7, 64, 300}) check_transpose(rng, neq, 6);
  check_transpose(rng, 2000, 3);
  for (int nsys : {1, 2, 5, 17, 100})
    for (int cap : {1, 4, 64, 256, 1024}) check_blocks(rng, nsys, cap);
  check_carve(rng);
  std::printf("host checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
