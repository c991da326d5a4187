/* oracle_checks.c -- the C oracle (oracle/fvc_oracle.c) under
 * AddressSanitizer and UndefinedBehaviorSanitizer: every SpMV kind, dot,
 * getres and every branch of the solver dispatch (method x storage x OpenMP
 * flag) on small grid systems whose arrays are allocated at their exact
 * sizes, so that a read or write one past an array's end is reported.
 * Built and run by tests/test_sanitize.py with gcc on the CPU. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t iterations, inner, status, pad;
  double energy, conv, last;
} orc_stats;

void orc_matvec(int32_t kind, int32_t neq, const double *ad, const double *au, const double *al,
                const int32_t *ia, const int32_t *ja, const double *x, double *y);
double orc_dot(const double *a, const double *b, int32_t n);
double orc_dot_L4(const double *a, const double *b, int32_t n);
void orc_solver(int32_t code, int32_t matriz, int32_t unsym, int32_t omp, int32_t neq, const double *ad,
                const double *au, const double *al, double *x, double *b, const int32_t *iax,
                const int32_t *ja, double tol, int32_t maxit, orc_stats *st);
void orc_gmres(int32_t kind, int32_t neq, const double *ad, const double *au, const double *al,
               const int32_t *ia, const int32_t *ja, const double *m, const double *b, double *x,
               int32_t k, double tol, int32_t maxit, orc_stats *st);
void orc_getres(double *u, double *x, const int32_t *num, int32_t neq, int32_t numel, int32_t flag);

static int failures = 0;
#define CHECK(cond, ...)                                         \
  do {                                                           \
    if (!(cond)) {                                               \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);       \
      fprintf(stderr, __VA_ARGS__);                              \
      fprintf(stderr, "\n");                                     \
      ++failures;                                                \
    }                                                            \
  } while (0)

static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) abort();
  return p;
}

/* A 5-point finite-volume operator on nx x ny cells: the off-diagonal
 * coefficient to neighbour (di, dj) is -(1 + c*sign) (c = 0: symmetric),
 * the diagonal 4.5 + |c| so that the matrix is diagonally dominant.  The
 * dense matrix is kept for the checks. */
typedef struct {
  int nx, ny, neq;
  double *dense;                    /* neq x neq, row-major */
  /* CSRD: ad, off-diagonals row by row (au), iad/jad; lower-only (al_l) */
  double *ad, *au, *csr;            /* csr: full rows incl. the diagonal */
  int32_t *iad, *jad, *icsr, *jcsr;
  /* CSRC: lower pattern il/jl, al lower values, aut transposed upper */
  double *al, *aut;
  int32_t *il, *jl;
  /* ELLPACK w = 4 */
  double *ae;
  int32_t nl[2], *je;
} sys_t;

static double coef(double c, int di, int dj) {
  return -(1.0 + c * (di + 0.5 * dj));
}

static void build(sys_t *S, int nx, int ny, double c) {
  const int n = nx * ny;
  S->nx = nx; S->ny = ny; S->neq = n;
  S->dense = xmalloc(sizeof(double) * (size_t)n * n);
  memset(S->dense, 0, sizeof(double) * (size_t)n * n);
  const int di[4] = {0, -1, 1, 0}, dj[4] = {-1, 0, 0, 1};   /* south, west, east, north: column order */
  int nnz = 0, nlow = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int r = j * nx + i;
      S->dense[(size_t)r * n + r] = 4.5 + fabs(c);
      for (int q = 0; q < 4; ++q) {
        const int ii = i + di[q], jj = j + dj[q];
        if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
        S->dense[(size_t)r * n + jj * nx + ii] = coef(c, di[q], dj[q]);
        ++nnz;
        if (jj * nx + ii < r) ++nlow;
      }
    }
  S->ad = xmalloc(sizeof(double) * n);
  S->au = xmalloc(sizeof(double) * nnz);
  S->iad = xmalloc(sizeof(int32_t) * (n + 1));
  S->jad = xmalloc(sizeof(int32_t) * nnz);
  S->csr = xmalloc(sizeof(double) * (nnz + n));
  S->icsr = xmalloc(sizeof(int32_t) * (n + 1));
  S->jcsr = xmalloc(sizeof(int32_t) * (nnz + n));
  S->al = xmalloc(sizeof(double) * nlow);
  S->aut = xmalloc(sizeof(double) * nlow);
  S->il = xmalloc(sizeof(int32_t) * (n + 1));
  S->jl = xmalloc(sizeof(int32_t) * nlow);
  S->ae = xmalloc(sizeof(double) * 4 * n);
  S->je = xmalloc(sizeof(int32_t) * 4 * n);
  S->nl[0] = 4; S->nl[1] = n + 1;
  int k = 0, kc = 0, kl = 0;
  S->iad[0] = S->icsr[0] = S->il[0] = 1;
  for (int r = 0; r < n; ++r) {
    S->ad[r] = S->dense[(size_t)r * n + r];
    int ke = 0;
    for (int col = 0; col < n; ++col) {
      const double v = S->dense[(size_t)r * n + col];
      if (v == 0.0) continue;
      S->csr[kc] = v; S->jcsr[kc++] = col + 1;
      if (col == r) continue;
      S->au[k] = v; S->jad[k++] = col + 1;
      S->ae[4 * r + ke] = v; S->je[4 * r + ke++] = col + 1;
      if (col < r) {
        S->al[kl] = v; S->aut[kl] = S->dense[(size_t)col * n + r]; S->jl[kl++] = col + 1;
      }
    }
    for (; ke < 4; ++ke) { S->ae[4 * r + ke] = 0.0; S->je[4 * r + ke] = r + 1; }
    S->iad[r + 1] = k + 1; S->icsr[r + 1] = kc + 1; S->il[r + 1] = kl + 1;
  }
}

static void release(sys_t *S) {
  free(S->dense); free(S->ad); free(S->au); free(S->iad); free(S->jad); free(S->csr);
  free(S->icsr); free(S->jcsr); free(S->al); free(S->aut); free(S->il); free(S->jl);
  free(S->ae); free(S->je);
}

static double relres(const sys_t *S, const double *x, const double *b) {
  double num = 0.0, den = 0.0;
  for (int r = 0; r < S->neq; ++r) {
    double t = 0.0;
    for (int c = 0; c < S->neq; ++c) t += S->dense[(size_t)r * S->neq + c] * x[c];
    num += (b[r] - t) * (b[r] - t);
    den += b[r] * b[r];
  }
  return sqrt(num / den);
}

static void check_matvec(const sys_t *S, int sym) {
  const int n = S->neq;
  double *x = xmalloc(sizeof(double) * n), *y = xmalloc(sizeof(double) * n), *want = xmalloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) x[i] = sin(0.37 * i) + 0.1 * i;
  for (int r = 0; r < n; ++r) {
    want[r] = 0.0;
    for (int c = 0; c < n; ++c) want[r] += S->dense[(size_t)r * n + c] * x[c];
  }
  for (int kind = 1; kind <= 7; ++kind) {
    if ((kind == 4 || kind == 5) && !sym) continue;   /* symmetric-storage kinds */
    memset(y, 0, sizeof(double) * n);
    switch (kind) {
      case 1: case 7: orc_matvec(kind, n, S->ad, S->au, NULL, S->iad, S->jad, x, y); break;
      case 2: orc_matvec(kind, n, S->csr, NULL, NULL, S->icsr, S->jcsr, x, y); break;
      case 3: orc_matvec(kind, n, S->ad, S->aut, S->al, S->il, S->jl, x, y); break;
      case 4: case 5: orc_matvec(kind, n, S->ad, NULL, S->al, S->il, S->jl, x, y); break;
      case 6: orc_matvec(kind, n, S->ad, S->ae, NULL, S->nl, S->je, x, y); break;
    }
    for (int r = 0; r < n; ++r)
      CHECK(fabs(y[r] - want[r]) <= 1e-12 * (1.0 + fabs(want[r])), "matvec kind %d row %d: %g vs %g", kind, r,
            y[r], want[r]);
  }
  free(x); free(y); free(want);
}

/* Every (code, matriz, omp) branch of the dispatch; the supported ones must
 * converge on these diagonally dominant systems, the rest report status 4. */
static void check_solver(const sys_t *S, int unsym) {
  const int n = S->neq;
  double *b0 = xmalloc(sizeof(double) * n), *b = xmalloc(sizeof(double) * n), *x = xmalloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) b0[i] = 1.0 + cos(0.11 * i);
  for (int code = 1; code <= 3; ++code)
    for (int matriz = 1; matriz <= 4; ++matriz)
      for (int omp = 0; omp <= 1; ++omp) {
        if (code == 1 && unsym) continue;                /* PCG takes symmetric systems */
        memcpy(b, b0, sizeof(double) * n);
        memset(x, 0, sizeof(double) * n);
        orc_stats st;
        const double *ad = S->ad, *au = S->au, *al = S->al;
        const int32_t *ia = S->iad, *ja = S->jad;
        if (matriz == 1 && !unsym) { au = NULL; ia = S->il; ja = S->jl; }
        if (matriz == 2) { ad = S->csr; au = NULL; al = NULL; ia = S->icsr; ja = S->jcsr; }
        if (matriz == 3) { au = S->aut; ia = S->il; ja = S->jl; }
        if (matriz == 4) { au = S->ae; al = NULL; ia = S->nl; ja = S->je; }
        orc_solver(code, matriz, unsym, omp, n, ad, au, al, x, b, ia, ja, 1e-10, 2000, &st);
        const int supported = !(code >= 2 && matriz == 2) && !(code == 3 && matriz == 4);
        if (!supported) {
          CHECK(st.status == 4, "code %d matriz %d: status %d, expected unsupported", code, matriz, st.status);
          continue;
        }
        if (matriz == 2) {   /* PCG on CSR: the preconditioner is the first neq CSR values */
          CHECK(st.status == 0 || st.status == 1, "code %d matriz 2 status %d", code, st.status);
          continue;
        }
        if (code == 2 && n <= 2 && st.status == 1) {
          /* the half step can solve these exactly: s = 0, omega = 0/0
           * (src/Iterativos.f:184), and the reference's BiCGSTAB, which
           * tests only after the full step (:190), runs on with NaN */
          CHECK(isnan(st.last), "n %d BiCGSTAB: maxit without the half-step breakdown", n);
          continue;
        }
        CHECK(st.status == 0, "n %d code %d matriz %d omp %d unsym %d: status %d", n, code, matriz, omp, unsym, st.status);
        const double r = relres(S, x, b0);
        CHECK(r < 1e-8, "code %d matriz %d omp %d unsym %d: relres %g", code, matriz, omp, unsym, r);
      }
  /* GMRES restart lengths 1 .. beyond n */
  const int ks[] = {1, 2, 5, 15, 30, 200};
  for (size_t t = 0; t < sizeof ks / sizeof ks[0]; ++t) {
    memset(x, 0, sizeof(double) * n);
    orc_stats st;
    orc_gmres(1, n, S->ad, S->au, NULL, S->iad, S->jad, S->ad, b0, x, ks[t], 1e-10, 5000, &st);
    CHECK(st.status == 0, "gmres k=%d status %d", ks[t], st.status);
    CHECK(relres(S, x, b0) < 1e-8, "gmres k=%d relres %g", ks[t], relres(S, x, b0));
  }
  free(b0); free(b); free(x);
}

static void check_dots(void) {
  for (int n = 0; n <= 41; ++n) {
    double *a = xmalloc(sizeof(double) * n), *b = xmalloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) { a[i] = i + 1; b[i] = 2.0 * (i + 1); }
    const double want = (double)n * (n + 1) * (2 * n + 1) / 3.0;
    CHECK(orc_dot(a, b, n) == want, "dot n=%d", n);
    CHECK(orc_dot_L4(a, b, n) == want, "dot_L4 n=%d", n);
    free(a); free(b);
  }
}

static void check_getres(void) {
  const int neq = 50, numel = 37;
  double *x = xmalloc(sizeof(double) * neq), *u = xmalloc(sizeof(double) * numel),
         *v = xmalloc(sizeof(double) * numel);
  int32_t *num = xmalloc(sizeof(int32_t) * numel);
  for (int i = 0; i < numel; ++i) num[i] = 1 + (i * 7) % neq;   /* distinct: 7 is prime to 50 */
  for (int i = 0; i < neq; ++i) x[i] = -1.0;
  for (int i = 0; i < numel; ++i) u[i] = 0.5 * i;
  orc_getres(u, x, num, neq, numel, 1);
  orc_getres(v, x, num, neq, numel, 0);
  for (int i = 0; i < numel; ++i) CHECK(v[i] == u[i], "getres round trip %d", i);
  free(x); free(u); free(v); free(num);
}

int main(void) {
  const int shapes[][2] = {{1, 1}, {2, 1}, {1, 5}, {3, 3}, {8, 5}, {17, 12}};
  for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; ++s) {
    sys_t S;
    build(&S, shapes[s][0], shapes[s][1], 0.0);
    check_matvec(&S, 1);
    check_solver(&S, 0);
    release(&S);
    build(&S, shapes[s][0], shapes[s][1], 0.4);
    check_matvec(&S, 0);
    check_solver(&S, 1);
    release(&S);
  }
  check_dots();
  check_getres();
  printf("oracle checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
