/* sanitize_main.c -- TEST INFRASTRUCTURE: runs every solver path of the CPU
 * oracle (oracle.c) on small cases so that a build under
 * -fsanitize=address,undefined (the `sanitize` target of the Makefile beside
 * this file) checks it for out-of-bounds access, use after free, leaks and
 * undefined behaviour. Exit status 0 when every solve returns 0 and ends with
 * finite histories; the sanitizers abort on their own findings.
 *
 * Paths covered: Rodas5P dense in FP64, FP32 and FP32 with refinement
 * (Robertson, the BBN network), Rodas5P with the sparse LU (a tridiagonal
 * linear-quadratic system, whose CSR needs no fill and is written here), Tsit5
 * (Lorenz), the right-hand side and Jacobian entry points, and the OpenMP
 * ensemble loop with two threads. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

int oracle_rodas5p(const int* sys_i, const double* sys_d, const int* off_rc,
                   const double* off_val, const double* inv_eps, int N, const double* y0,
                   const double* times, int n_save, const double* params,
                   const double* weights, double dt0, double rtol, double atol, int max_steps,
                   int lu_fp32, double pcoeff, double icoeff, double dcoeff, int tf_index,
                   int nnz, const int* row_ptr, const int* col_ind, const int* diag_ptr,
                   const int* order, const int* slot_of, const int* pattern, double* hist,
                   int* accepted, int* rejected, int* loop, int n_threads);
int oracle_tsit5(const int* sys_i, const double* sys_d, const int* off_rc, const double* off_val,
                 const double* inv_eps, int N, const double* y0, const double* times, int n_save,
                 const double* params, const double* weights, double dt0, double rtol,
                 double atol, int max_steps, double pcoeff, double icoeff, double dcoeff,
                 double* hist, int* accepted, int* rejected, int* loop, int n_threads);
int oracle_rhs(const int* sys_i, const double* sys_d, const int* off_rc, const double* off_val,
               const double* inv_eps, const double* y, double t, const double* p, double* f);
int oracle_jacobian(const int* sys_i, const double* sys_d, const int* off_rc,
                    const double* off_val, const double* inv_eps, const double* y, double t,
                    const double* p, double* J, double* dT);

enum { LORENZ = 0, ROBERTSON = 1, LINQUAD = 7, BBN = 11 };
#define NT 4

static int failures = 0;

static void check(const char* what, int rc, const double* hist, int count) {
  int finite = 1;
  for (int i = 0; i < count; ++i) finite = finite && isfinite(hist[i]);
  if (rc != 0 || !finite) {
    fprintf(stderr, "sanitize: %s failed (rc %d, finite %d)\n", what, rc, finite);
    ++failures;
  } else {
    printf("sanitize: %s ok\n", what);
  }
}

static void dense(const char* name, int kind, int n, int n_p, const double* y0row,
                  const double* prow, const double* times, int n_save, double dt0, double rtol,
                  double atol) {
  int si[7] = {kind, n, n_p, 0, 0, 1, 0};
  double sd[4] = {0.0, 0.0, 0.0, 0.0};
  int off_rc[1] = {0};
  double off_val[1] = {0.0}, inv_eps[1] = {1.0};
  double* y0 = malloc(sizeof(double) * NT * n);
  double* p = malloc(sizeof(double) * NT * n_p);
  double* w = malloc(sizeof(double) * NT * n);
  double* hist = malloc(sizeof(double) * NT * n_save * n);
  int acc[NT], rej[NT], loop[NT];
  for (int i = 0; i < NT; ++i) {
    for (int j = 0; j < n; ++j) y0[i * n + j] = y0row[j], w[i * n + j] = 1.0;
    for (int j = 0; j < n_p; ++j) p[i * n_p + j] = prow[j] * (1.0 + 0.05 * i);
  }
  const char* modes[3] = {"fp64", "fp32", "fp32 refined"};
  for (int lu = 0; lu < 3; ++lu) {
    int rc = oracle_rodas5p(si, sd, off_rc, off_val, inv_eps, NT, y0, times, n_save, p, w, dt0,
                            rtol, atol, 100000, lu, 0.0, 1.0, 0.0, -1, 0, NULL, NULL, NULL,
                            NULL, NULL, NULL, hist, acc, rej, loop, 2);
    char what[96];
    snprintf(what, sizeof what, "rodas5p %s dense %s", name, modes[lu]);
    check(what, rc, hist, NT * n_save * n);
  }
  double f[16], J[256], dT[16];
  oracle_rhs(si, sd, off_rc, off_val, inv_eps, y0, times[0] + 0.5 * (times[1] - times[0]), p, f);
  oracle_jacobian(si, sd, off_rc, off_val, inv_eps, y0, times[0] + 0.1, p, J, dT);
  free(y0), free(p), free(w), free(hist);
}

/* dy/dt = diag(p) y + 0.5 (y_{j-1} + y_{j+1}) + 0.01 y*y on a tridiagonal
 * pattern of n variables: no fill in the natural order, so L+U is the
 * pattern, written here in CSR with the caller slot table */
static void sparse_tridiagonal(void) {
  enum { n = 8 };
  int n_off = 2 * (n - 1);
  int off_rc[3 * 2 * (n - 1)];
  double off_val[2 * (n - 1)];
  int k = 0;
  for (int j = 0; j < n - 1; ++j) {
    off_rc[k] = j, off_rc[n_off + k] = j + 1, off_rc[2 * n_off + k] = -1, off_val[k] = 0.5, ++k;
    off_rc[k] = j + 1, off_rc[n_off + k] = j, off_rc[2 * n_off + k] = -1, off_val[k] = 0.5, ++k;
  }
  int si[7] = {LINQUAD, n, n, 0, 0, 1, n_off};
  double sd[4] = {0.0, 0.0, 0.0, 0.01};
  double inv_eps[1] = {1.0};
  int row_ptr[n + 1], col_ind[3 * n], diag_ptr[n], order[n], slot_of[n * n], pattern[n * n];
  int nnz = 0;
  for (int r = 0; r < n * n; ++r) slot_of[r] = -1, pattern[r] = 0;
  for (int r = 0; r < n; ++r) {
    order[r] = r;
    row_ptr[r] = nnz;
    for (int c = r - 1; c <= r + 1; ++c) {
      if (c < 0 || c >= n) continue;
      if (c == r) diag_ptr[r] = nnz;
      col_ind[nnz] = c;
      slot_of[r * n + c] = nnz;
      pattern[r * n + c] = 1;
      ++nnz;
    }
  }
  row_ptr[n] = nnz;
  double y0[NT * n], p[NT * n], w[NT * n], times[3] = {0.0, 0.5, 2.0};
  double hist[NT * 3 * n];
  int acc[NT], rej[NT], loop[NT];
  for (int i = 0; i < NT; ++i)
    for (int j = 0; j < n; ++j) {
      y0[i * n + j] = 0.1 * (j + 1) - 0.3 * i;
      p[i * n + j] = -pow(10.0, (double)(j % 4) + 0.2 * i);
      w[i * n + j] = 1.0;
    }
  const char* modes[3] = {"fp64", "fp32", "fp32 refined"};
  for (int lu = 0; lu < 3; ++lu) {
    int rc = oracle_rodas5p(si, sd, off_rc, off_val, inv_eps, NT, y0, times, 3, p, w, 0.0,
                            1e-7, 1e-9, 100000, lu, 0.0, 1.0, 0.0, -1, nnz, row_ptr, col_ind,
                            diag_ptr, order, slot_of, pattern, hist, acc, rej, loop, 2);
    char what[96];
    snprintf(what, sizeof what, "rodas5p tridiagonal sparse %s", modes[lu]);
    check(what, rc, hist, NT * 3 * n);
  }
}

static void tsit5_lorenz(void) {
  int si[7] = {LORENZ, 3, 1, 0, 0, 1, 0};
  double sd[4] = {0}, off_val[1] = {0}, inv_eps[1] = {1.0};
  int off_rc[1] = {0};
  double y0[NT * 3], p[NT], w[NT * 3], times[5] = {0.0, 0.25, 0.5, 0.75, 1.0};
  double hist[NT * 5 * 3];
  int acc[NT], rej[NT], loop[NT];
  for (int i = 0; i < NT; ++i) {
    y0[3 * i] = 1.0, y0[3 * i + 1] = 0.0, y0[3 * i + 2] = 0.0;
    p[i] = 5.0 + 6.0 * i;
    for (int j = 0; j < 3; ++j) w[3 * i + j] = 1.0;
  }
  int rc = oracle_tsit5(si, sd, off_rc, off_val, inv_eps, NT, y0, times, 5, p, w, 0.0, 1e-8,
                        1e-10, 100000, 0.0, 1.0, 0.0, hist, acc, rej, loop, 2);
  check("tsit5 lorenz", rc, hist, NT * 5 * 3);
}

int main(void) {
  const double rob_y0[3] = {1.0, 0.0, 0.0}, rob_p[3] = {0.04, 1e4, 3e7};
  const double rob_t[4] = {0.0, 1e-3, 1.0, 1e3};
  dense("robertson", ROBERTSON, 3, 3, rob_y0, rob_p, rob_t, 4, 0.0, 1e-6, 1e-8);
  const double bbn_y0[4] = {0.4677, 0.5323, 1e-20, 0.0}, bbn_p[2] = {0.785, 3.044};
  const double bbn_x[2] = {0.1293, 129.3};
  dense("bbn", BBN, 4, 2, bbn_y0, bbn_p, bbn_x, 2, 0.1, 1e-4, 1e-9);
  sparse_tridiagonal();
  tsit5_lorenz();
  if (failures) fprintf(stderr, "sanitize: %d failures\n", failures);
  return failures ? 1 : 0;
}
