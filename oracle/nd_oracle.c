/*
 * C restatement of socu's nested-dissection block-tridiagonal Cholesky
 * (socu/block_tridiag_solver.py), float64, for the CPU baseline timing and
 * as a second, independently coded oracle.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ (checked against oracle/nd.py
 * and the dense factor) and by bench.py's CPU-baseline leg.  The product
 * library never links or calls it.
 *
 * Same elimination order, same E layout, same per-node operations as
 * oracle/nd.py (which cites the reference lines):
 *   level l, stride s = 2^l, nodes i = s-1+2st, t < (N+s)/(2s)      (:174, :798)
 *   L_i <- chol(L_i)                                                  (potrf.py:97-119)
 *   R  = E[off+i/s]      <- R L_i^{-T};  L_{i+s} -= R R^T              (trsm.py:94-120, syrk.py:106-158)
 *   Lc = E[off+(i-s)/s]  <- L_i^{-1} Lc; L_{i-s} -= Lc^T Lc            (trsm.py:202-228, syrk.py:264-315)
 *   F  = E[next+(i-s)/(2s)] -= R Lc                                    (gemm.py:151-201)
 * The Schur updates of a level are applied per receiving block (owner
 * computes), so the OpenMP loops have no write conflicts.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int levels_of(int N) {
  int L = 0;
  while (N > 0) {
    ++L;
    N >>= 1;
  }
  return L;
}

long long ndo_off_diag_storage_len(int N) {
  long long t = 0;
  for (int l = 0; l < levels_of(N); ++l) t += (N >> l) - 1;
  return t;
}

static int level_off(int N, int l) {
  int off = 0;
  for (int q = 0; q < l; ++q) off += (N >> q) - 1;
  return off;
}

/* In-place lower Cholesky of an n x n row-major block (lower triangle read). */
static void potrf(double* A, int n) {
  for (int k = 0; k < n; ++k) {
    double d = A[(size_t)k * n + k];
    for (int j = 0; j < k; ++j) d -= A[(size_t)k * n + j] * A[(size_t)k * n + j];
    d = sqrt(d);
    A[(size_t)k * n + k] = d;
    const double inv = 1.0 / d;
    for (int r = k + 1; r < n; ++r) {
      double v = A[(size_t)r * n + k];
      const double* ar = A + (size_t)r * n;
      const double* ak = A + (size_t)k * n;
      for (int j = 0; j < k; ++j) v -= ar[j] * ak[j];
      A[(size_t)r * n + k] = v * inv;
    }
  }
}

/* X <- X L^{-T} (X n x n row-major): each row x solves L z = x. */
static void trsm_rltn(const double* Lb, double* X, int n) {
  for (int r = 0; r < n; ++r) {
    double* x = X + (size_t)r * n;
    for (int k = 0; k < n; ++k) {
      const double* lk = Lb + (size_t)k * n;
      double v = x[k];
      for (int j = 0; j < k; ++j) v -= lk[j] * x[j];
      x[k] = v / lk[k];
    }
  }
}

/* X <- L^{-1} X (X n x w row-major, row pitch ld). */
static void trsm_llnn(const double* Lb, double* X, int n, int w, int ld) {
  for (int k = 0; k < n; ++k) {
    const double inv = 1.0 / Lb[(size_t)k * n + k];
    double* xk = X + (size_t)k * ld;
    for (int j = 0; j < w; ++j) xk[j] *= inv;
    for (int r = k + 1; r < n; ++r) {
      const double l = Lb[(size_t)r * n + k];
      double* xr = X + (size_t)r * ld;
      for (int j = 0; j < w; ++j) xr[j] -= l * xk[j];
    }
  }
}

/* X <- L^{-T} X (X n x w row-major). */
static void trsm_lltn(const double* Lb, double* X, int n, int w, int ld) {
  for (int k = n - 1; k >= 0; --k) {
    const double inv = 1.0 / Lb[(size_t)k * n + k];
    double* xk = X + (size_t)k * ld;
    for (int j = 0; j < w; ++j) xk[j] *= inv;
    for (int r = 0; r < k; ++r) {
      const double l = Lb[(size_t)k * n + r];
      double* xr = X + (size_t)r * ld;
      for (int j = 0; j < w; ++j) xr[j] -= l * xk[j];
    }
  }
}

/* C -= A A^T (lower triangle incl. diagonal). */
static void syrk_ln(double* C, const double* A, int n) {
  for (int r = 0; r < n; ++r) {
    const double* ar = A + (size_t)r * n;
    for (int c = 0; c <= r; ++c) {
      const double* ac = A + (size_t)c * n;
      double v = 0.0;
      for (int k = 0; k < n; ++k) v += ar[k] * ac[k];
      C[(size_t)r * n + c] -= v;
    }
  }
}

/* C -= A^T A (lower triangle incl. diagonal). */
static void syrk_lt(double* C, const double* A, int n) {
  for (int k = 0; k < n; ++k) {
    const double* ak = A + (size_t)k * n;
    for (int r = 0; r < n; ++r) {
      const double a = ak[r];
      double* cr = C + (size_t)r * n;
      for (int c = 0; c <= r; ++c) cr[c] -= a * ak[c];
    }
  }
}

/* C -= A B (all n x n). */
static void gemm_nn(double* C, const double* A, const double* Bm, int n, int w, int ldb) {
  for (int r = 0; r < n; ++r) {
    double* cr = C + (size_t)r * ldb;
    const double* ar = A + (size_t)r * n;
    for (int k = 0; k < n; ++k) {
      const double a = ar[k];
      const double* bk = Bm + (size_t)k * ldb;
      for (int c = 0; c < w; ++c) cr[c] -= a * bk[c];
    }
  }
}

/* C -= A^T B (A n x n, B n x w). */
static void gemm_tn(double* C, const double* A, const double* Bm, int n, int w, int ldb) {
  for (int k = 0; k < n; ++k) {
    const double* ak = A + (size_t)k * n;
    const double* bk = Bm + (size_t)k * ldb;
    for (int r = 0; r < n; ++r) {
      const double a = ak[r];
      double* cr = C + (size_t)r * ldb;
      for (int c = 0; c < w; ++c) cr[c] -= a * bk[c];
    }
  }
}

static int active(const unsigned char* mask, int b) { return mask == NULL || mask[b] != 0; }

void ndo_factor(double* L, double* E, const unsigned char* mask, int B, int N, int n) {
  const size_t nn = (size_t)n * n;
  const long long nE = ndo_off_diag_storage_len(N);
  const int Lv = levels_of(N);
  for (int l = 0; l < Lv; ++l) {
    const int s = 1 << l;
    const int dim = (N + s) / (2 * s);
    const int off = level_off(N, l);
    const int next = off + (N >> l) - 1;
    const int nT = N / (2 * s);
    /* potrf + both trsm per node */
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int b = 0; b < B; ++b)
      for (int t = 0; t < dim; ++t) {
        if (!active(mask, b)) continue;
        const int i = s - 1 + 2 * s * t;
        double* Li = L + ((size_t)b * N + i) * nn;
        double* Eb = E + (size_t)b * nE * nn;
        potrf(Li, n);
        if (i + s < N) trsm_rltn(Li, Eb + (size_t)(off + i / s) * nn, n);
        if (i >= s) trsm_llnn(Li, Eb + (size_t)(off + (i - s) / s) * nn, n, n, n);
      }
    if (l == Lv - 1) break;
    /* receiving blocks L_j, j = 2s(u+1)-1, and fill-ins */
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int b = 0; b < B; ++b)
      for (int u = 0; u < nT + dim; ++u) {
        if (!active(mask, b)) continue;
        double* Eb = E + (size_t)b * nE * nn;
        if (u < nT) {
          const int j = 2 * s * (u + 1) - 1;
          double* Lj = L + ((size_t)b * N + j) * nn;
          syrk_ln(Lj, Eb + (size_t)(off + (j - s) / s) * nn, n);
          if (j + s < N) syrk_lt(Lj, Eb + (size_t)(off + j / s) * nn, n);
        } else {
          const int t = u - nT;
          const int i = s - 1 + 2 * s * t;
          if (t >= 1 && i + s < N)
            gemm_nn(Eb + (size_t)(next + (i - s) / (2 * s)) * nn, Eb + (size_t)(off + i / s) * nn,
                    Eb + (size_t)(off + (i - s) / s) * nn, n, n, n);
        }
      }
  }
}

void ndo_forward(const double* L, const double* E, double* x, const unsigned char* mask, int B, int N, int n, int m) {
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const long long nE = ndo_off_diag_storage_len(N);
  const int Lv = levels_of(N);
  for (int l = 0; l < Lv; ++l) {
    const int s = 1 << l;
    const int dim = (N + s) / (2 * s);
    const int off = level_off(N, l);
    const int nT = N / (2 * s);
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int t = 0; t < dim; ++t) {
        if (!active(mask, b)) continue;
        const int i = s - 1 + 2 * s * t;
        trsm_llnn(L + ((size_t)b * N + i) * nn, x + ((size_t)b * N + i) * nm, n, m, m);
      }
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int u = 0; u < nT; ++u) {
        if (!active(mask, b)) continue;
        const int j = 2 * s * (u + 1) - 1;
        const double* Eb = E + (size_t)b * nE * nn;
        double* xb = x + (size_t)b * N * nm;
        gemm_nn(xb + (size_t)j * nm, Eb + (size_t)(off + (j - s) / s) * nn, xb + (size_t)(j - s) * nm, n, m, m);
        if (j + s < N) gemm_tn(xb + (size_t)j * nm, Eb + (size_t)(off + j / s) * nn, xb + (size_t)(j + s) * nm, n, m, m);
      }
  }
}

void ndo_backward(const double* L, const double* E, double* x, const unsigned char* mask, int B, int N, int n, int m) {
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const long long nE = ndo_off_diag_storage_len(N);
  const int Lv = levels_of(N);
  for (int l = Lv - 1; l >= 0; --l) {
    const int s = 1 << l;
    const int dim = (N + s) / (2 * s);
    const int off = level_off(N, l);
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int t = 0; t < dim; ++t) {
        if (!active(mask, b)) continue;
        const int i = s - 1 + 2 * s * t;
        const double* Eb = E + (size_t)b * nE * nn;
        double* xb = x + (size_t)b * N * nm;
        if (i + s < N) gemm_tn(xb + (size_t)i * nm, Eb + (size_t)(off + i / s) * nn, xb + (size_t)(i + s) * nm, n, m, m);
        if (i >= s) gemm_nn(xb + (size_t)i * nm, Eb + (size_t)(off + (i - s) / s) * nn, xb + (size_t)(i - s) * nm, n, m, m);
        trsm_lltn(L + ((size_t)b * N + i) * nn, xb + (size_t)i * nm, n, m, m);
      }
  }
}

void ndo_factor_and_solve(double* L, double* E, double* x, const unsigned char* mask, int B, int N, int n, int m) {
  ndo_factor(L, E, mask, B, N, n);
  ndo_forward(L, E, x, mask, B, N, n, m);
  ndo_backward(L, E, x, mask, B, N, n, m);
}

int ndo_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

/*
 * The same computations behind the argument lists of libsocu_hip.so's direct
 * entry points (include/socu_hip.h: socu_cholesky_factor, _solve,
 * _forward_substitution, _backward_substitution, _factor_and_solve), on host
 * arrays, so that one harness can drive either library with the same call:
 * dtype must be 1 (float64; 0 = float32 is not restated and returns -2),
 * work and stream are accepted and ignored (the arrays are host memory; the
 * call returns when done), validation returns -1 like the product's
 * SOCU_EINVAL.  Test infrastructure and CPU baseline only.
 */
static int capi_check(const void* L, const void* E, const void* x, int need_x, int B, int N, int n, int m, int dtype) {
  if (dtype != 1) return dtype == 0 ? -2 : -1;
  if (B < 0 || N < 1 || n < 1 || (need_x && m < 1)) return -1;
  if (B > 0 && (L == NULL || (E == NULL && N > 1) || (need_x && x == NULL))) return -1;
  return 0;
}

size_t ndo_cholesky_solve_workspace_size(int B, int N, int n, int dtype) {
  (void)B, (void)N, (void)n, (void)dtype;
  return 0;
}

int ndo_cholesky_factor(void* L, void* E, const unsigned char* mask, int B, int N, int n, int dtype, void* stream) {
  (void)stream;
  const int rc = capi_check(L, E, NULL, 0, B, N, n, 1, dtype);
  if (rc == 0 && B > 0) ndo_factor((double*)L, (double*)E, mask, B, N, n);
  return rc;
}

int ndo_cholesky_forward_substitution(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                                      int n, int m, int dtype, void* work, void* stream) {
  (void)work, (void)stream;
  const int rc = capi_check(L, E, x, 1, B, N, n, m, dtype);
  if (rc == 0 && B > 0) ndo_forward((const double*)L, (const double*)E, (double*)x, mask, B, N, n, m);
  return rc;
}

int ndo_cholesky_backward_substitution(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N,
                                       int n, int m, int dtype, void* work, void* stream) {
  (void)work, (void)stream;
  const int rc = capi_check(L, E, x, 1, B, N, n, m, dtype);
  if (rc == 0 && B > 0) ndo_backward((const double*)L, (const double*)E, (double*)x, mask, B, N, n, m);
  return rc;
}

int ndo_cholesky_solve(const void* L, const void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                       int dtype, void* work, void* stream) {
  (void)work, (void)stream;
  const int rc = capi_check(L, E, x, 1, B, N, n, m, dtype);
  if (rc == 0 && B > 0) {
    ndo_forward((const double*)L, (const double*)E, (double*)x, mask, B, N, n, m);
    ndo_backward((const double*)L, (const double*)E, (double*)x, mask, B, N, n, m);
  }
  return rc;
}

int ndo_cholesky_factor_and_solve(void* L, void* E, void* x, const unsigned char* mask, int B, int N, int n, int m,
                                  int dtype, void* stream) {
  (void)stream;
  const int rc = capi_check(L, E, x, 1, B, N, n, m, dtype);
  if (rc == 0 && B > 0) ndo_factor_and_solve((double*)L, (double*)E, (double*)x, mask, B, N, n, m);
  return rc;
}
