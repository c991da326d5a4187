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
    const std::vector<int> g = allocate_blocks(neq, cap, expo);
    CHECK((int)g.size() == nsys, "size");
    long sum = 0;
    for (int i = 0; i < nsys; ++i) {
      const int lim = std::max(1, (neq[i] + 63) / 64);
      CHECK(g[i] >= 1, "system %d gets %d blocks", i, g[i]);
      if (cap > nsys) CHECK(g[i] <= lim, "system %d: %d blocks for %d rows", i, g[i], neq[i]);
      sum += g[i];
    }
    CHECK(sum <= std::max(cap, nsys), "%ld blocks for capacity %d and %d systems", sum, cap, nsys);
    // shares grow with size: a larger system gets at most one block fewer
    // than a smaller one (the one a largest-remainder rounding can move)
    for (int i = 0; i < nsys; ++i)
      for (int j = 0; j < nsys; ++j)
        if (cap > nsys && neq[i] > neq[j]) CHECK(g[i] + 1 >= g[j], "monotone %d vs %d", i, j);
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
