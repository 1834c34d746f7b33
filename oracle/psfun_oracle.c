/*
 * psfun_oracle.c -- CPU restatement of PSFUN's f(A)x path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library (psfun_amd/libpsfun_hip.so) never calls it.
 *
 * Every routine follows a reference file/line range (paths relative to the
 * PSFUN tree) and runs unfused, one pass per PSBLAS-like primitive, exactly in
 * the reference's operation order.  PSBLAS itself is not part of the reference
 * tree; its primitives are restated from their textbook definitions (CSR row
 * sums in storage order, BLAS-1 with left-to-right accumulation).
 *
 * Pinning (what checks each routine; see tests/):
 *   or_tridiag_eig / or_exp_tridiag   -- scipy.linalg.expm fixtures, closed-form 2x2
 *   or_cram_apply                      -- the approximation's known uniform error on [0, inf)
 *                                         (max |exp(-x) - r(x)| about 9.4e-15, equioscillating),
 *                                         scipy expm on negative definite matrices
 *   or_dense_expm                      -- scipy.linalg.expm fixtures, diagonal/nilpotent cases
 *   or_pade_expm                       -- scalar closed form r_p(x/2^ns)^(2^ns) from the factorial
 *                                         Pade coefficients, diagonal matrices, scipy expm at p = 6..13
 *   or_sym_eig / or_userf_sym_apply    -- numpy.linalg.eigh (LAPACK) eigenvalues and reconstruction,
 *                                         scipy expm / sqrtm of the upper-triangle matrix
 *   or_lanczos_f / or_arnoldi_f /
 *   or_sai_arnoldi_f                   -- f = exp equals the EXP restatements; f = sqrt against
 *                                         scipy sqrtm on small SPD matrices
 *   or_lanczos                         -- DST closed form exp(-tA) for Dirichlet Laplacians,
 *                                         scipy dense expm on small matrices, breakdown cases
 *   or_lanczos_batch                   -- bitwise equal to or_lanczos run instance by instance
 *   or_arnoldi                         -- same oracles as or_lanczos
 *   or_sai_arnoldi / or_dense_inverse  -- DST closed form, dense expm, A*inverse = I
 *   or_cg_shifted                      -- dense solve, residual check, PSBLAS stop rules
 *   or_ilu0 / or_ilu0_apply            -- L U equals A on A's pattern; exact LU (M^{-1} A = I) for
 *                                         tridiagonal and dense patterns; the breakdown on a zero pivot
 *   or_mscg                            -- dense complex solve per shift
 *   or_pcocg / or_mscg_prec            -- dense complex solve per shift for each preconditioner; with
 *                                         M = I one COCG run equals or_mscg's iterate for one shift
 *                                         (the shared Krylov space); warm start from a solution
 *   or_ellipkkp / or_ellipj / horner  -- the control values of examples/utils/utiltest.f90,
 *                                         mpmath elliptic functions
 *   or_hhtmethod1 / or_quad_apply      -- scipy sqrtm / expm on small SPD matrices
 *   or_talbot                          -- scalar exp(-x) on a grid (max error at N=16)
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zdouble;

static int g_threads = 1;
void or_set_threads(int t) {
  g_threads = t < 1 ? 1 : t;
#ifdef _OPENMP
  omp_set_num_threads(g_threads);
#endif
}
int or_get_threads(void) { return g_threads; }
int or_get_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* PSBLAS-like primitives (0-based CSR)                                */
/* ------------------------------------------------------------------ */
/* y = alpha*A*x + beta*y  (psb_spmm) */
void or_spmv(int n, const int *rp, const int *ci, const double *va, double alpha, const double *x,
             double beta, double *y) {
#pragma omp parallel for schedule(static) if (g_threads > 1)
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) s += va[j] * x[ci[j]];
    y[i] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[i];
  }
}

double or_dot(int n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (g_threads > 1)
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

double or_nrm2(int n, const double *x) { return sqrt(or_dot(n, x, x)); }

double or_amax(int n, const double *x) {
  double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static) if (g_threads > 1)
  for (int i = 0; i < n; ++i) m = fmax(m, fabs(x[i]));
  return m;
}

/* y = alpha*x + beta*y  (psb_geaxpby) */
void or_axpby(int n, double alpha, const double *x, double beta, double *y) {
#pragma omp parallel for schedule(static) if (g_threads > 1)
  for (int i = 0; i < n; ++i) y[i] = (beta == 0.0) ? alpha * x[i] : alpha * x[i] + beta * y[i];
}

/* max_i sum_j |a_ij|  (psb_spnrmi) */
double or_spnrmi(int n, const int *rp, const int *ci, const double *va) {
  (void)ci;
  double m = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) s += fabs(va[j]);
    m = fmax(m, s);
  }
  return m;
}

/* ------------------------------------------------------------------ */
/* Small dense f(H) step                                               */
/* ------------------------------------------------------------------ */
/* Symmetric tridiagonal eigen-decomposition by implicit QL with Wilkinson-type
 * shifts (Golub & Van Loan, Matrix Computations, sec. 8.3; the classical tql2
 * scheme).  d[k] diagonal (overwritten by eigenvalues), e[k-1] off-diagonal,
 * Z[k*k] column-major eigenvectors.  Returns 0 on success. */
int or_tridiag_eig(int k, double *d, const double *e_in, double *Z) {
  double *e = (double *)calloc((size_t)k + 1, sizeof(double));
  for (int i = 0; i + 1 < k; ++i) e[i] = e_in[i];
  for (int i = 0; i < k * k; ++i) Z[i] = 0.0;
  for (int i = 0; i < k; ++i) Z[i + (size_t)i * k] = 1.0;
  for (int l = 0; l < k; ++l) {
    int iter = 0, m;
    for (;;) {
      for (m = l; m < k - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++iter > 100) {
        free(e);
        return 1;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      int early = 0;
      for (i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          early = 1;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int q = 0; q < k; ++q) {
          double *zi = Z + (size_t)i * k, *zi1 = Z + (size_t)(i + 1) * k;
          f = zi1[q];
          zi1[q] = s * zi[q] + c * f;
          zi[q] = c * zi[q] - s * f;
        }
      }
      if (early) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  free(e);
  return 0;
}

typedef double (*or_f_t)(double);

static int tridiag_fapply(int k, const double *d, const double *e, double s, or_f_t f, const double *x, double *y);

/* y = exp(s*T) x for symmetric tridiagonal T (d: k, e: k-1), via the
 * eigen-decomposition T = Z diag(lambda) Z^T. */
int or_exp_tridiag(int k, const double *d, const double *e, double s, const double *x, double *y) {
  return tridiag_fapply(k, d, e, s, NULL, x, y);
}

/* EXPOKIT's DSCHBV (serial/exponential/dschbv.f:1-95, called for CHBSYM at
 * serial/psfun_d_serial_mod.F90:357-375): y = exp(t*H) y for a symmetric H of
 * order k by the partial fraction form of the uniform rational Chebyshev
 * approximation of type (14,14) to exp(-x) on [0, inf) (Gallopoulos and
 * Saad, SIAM J. Sci. Stat. Comput. 13 (1992), the seven poles theta_p of
 * positive imaginary part and their residues alpha_p, conjugate pairs
 * folded):
 *   y <- alpha0 y + Re sum_p alpha_p (-t H - theta_p I)^{-1} y.
 * The reference solves each complex symmetric system with LAPACK's ZSYSV
 * (Bunch-Kaufman); here Gaussian elimination with partial pivoting on the
 * same matrix (the same solution; the rounding of the solve differs, far
 * below the approximation's own error of about 1e-14).  H is dense,
 * column-major with leading dimension ldh.  Returns 0, or 1 on a zero pivot
 * (the reference stops there). */
static const double cram_alpha0 = 0.183216998528140087e-11;
static const double cram_alpha[7][2] = {
    {0.557503973136501826e+02, -0.204295038779771857e+03}, {-0.938666838877006739e+02, 0.912874896775456363e+02},
    {0.469965415550370835e+02, -0.116167609985818103e+02}, {-0.961424200626061065e+01, -0.264195613880262669e+01},
    {0.752722063978321642e+00, 0.670367365566377770e+00},  {-0.188781253158648576e-01, -0.343696176445802414e-01},
    {0.143086431411801849e-03, 0.287221133228814096e-03}};
static const double cram_theta[7][2] = {
    {-0.562314417475317895e+01, 0.119406921611247440e+01}, {-0.508934679728216110e+01, 0.358882439228376881e+01},
    {-0.399337136365302569e+01, 0.600483209099604664e+01}, {-0.226978543095856366e+01, 0.846173881758693369e+01},
    {0.208756929753827868e+00, 0.109912615662209418e+02},  {0.370327340957595652e+01, 0.136563731924991884e+02},
    {0.889777151877331107e+01, 0.166309842834712071e+02}};

int or_cram_apply(int k, const double *H, int ldh, double t, double *y) {
  zdouble *A = (zdouble *)malloc(sizeof(zdouble) * (size_t)k * k);
  zdouble *b = (zdouble *)malloc(sizeof(zdouble) * (size_t)k);
  double *z = (double *)malloc(sizeof(double) * (size_t)k);
  int rc = 0;
  for (int j = 0; j < k; ++j) {
    z[j] = y[j];
    y[j] = y[j] * cram_alpha0;
  }
  for (int p = 0; p < 7 && rc == 0; ++p) {
    const zdouble th = cram_theta[p][0] + I * cram_theta[p][1];
    const zdouble al = cram_alpha[p][0] + I * cram_alpha[p][1];
    for (int j = 0; j < k; ++j) {
      for (int i = 0; i < k; ++i) A[i + (size_t)j * k] = -t * H[i + (size_t)j * ldh];
      A[j + (size_t)j * k] -= th;
      b[j] = z[j];
    }
    for (int c = 0; c < k && rc == 0; ++c) { /* elimination with partial pivoting */
      int piv = c;
      for (int i = c + 1; i < k; ++i)
        if (cabs(A[i + (size_t)c * k]) > cabs(A[piv + (size_t)c * k])) piv = i;
      if (cabs(A[piv + (size_t)c * k]) == 0.0) {
        rc = 1;
        break;
      }
      if (piv != c) {
        for (int j = c; j < k; ++j) {
          const zdouble tmp = A[c + (size_t)j * k];
          A[c + (size_t)j * k] = A[piv + (size_t)j * k];
          A[piv + (size_t)j * k] = tmp;
        }
        const zdouble tb = b[c];
        b[c] = b[piv];
        b[piv] = tb;
      }
      for (int i = c + 1; i < k; ++i) {
        const zdouble m = A[i + (size_t)c * k] / A[c + (size_t)c * k];
        if (m == 0.0) continue;
        for (int j = c + 1; j < k; ++j) A[i + (size_t)j * k] -= m * A[c + (size_t)j * k];
        b[i] -= m * b[c];
      }
    }
    if (rc) break;
    for (int i = k - 1; i >= 0; --i) {
      zdouble a = b[i];
      for (int j = i + 1; j < k; ++j) a -= A[i + (size_t)j * k] * b[j];
      b[i] = a / A[i + (size_t)i * k];
    }
    for (int i = 0; i < k; ++i) y[i] += creal(al * b[i]);
  }
  free(A);
  free(b);
  free(z);
  return rc;
}

/* Which exp the EXP variants of the Krylov restatements use for f(T):
 * 0 (default) the eigen-decomposition below, exact to rounding; 1 DSCHBV's
 * CRAM(14,14), as the reference's CHBSYM computes it, error floor included. */
static int g_exp_cram = 0;
void or_set_exp_cram(int on) { g_exp_cram = on ? 1 : 0; }

/* y = Z diag(f(s*lambda)) Z^T x for symmetric tridiagonal T = Z diag(lambda) Z^T;
 * f = NULL is exp (or CRAM(14,14) after or_set_exp_cram(1)). */
static int tridiag_fapply(int k, const double *d, const double *e, double s, or_f_t f, const double *x, double *y) {
  if (!f && g_exp_cram) {
    double *H = (double *)calloc((size_t)k * k, sizeof(double));
    for (int i = 0; i < k; ++i) {
      H[i + (size_t)i * k] = d[i];
      if (i + 1 < k) H[i + (size_t)(i + 1) * k] = H[(i + 1) + (size_t)i * k] = e[i];
    }
    memcpy(y, x, sizeof(double) * (size_t)k);
    const int rc = or_cram_apply(k, H, k, s, y);
    free(H);
    return rc;
  }
  double *lam = (double *)malloc(sizeof(double) * (size_t)k);
  double *Z = (double *)malloc(sizeof(double) * (size_t)k * k);
  double *t = (double *)malloc(sizeof(double) * (size_t)k);
  memcpy(lam, d, sizeof(double) * (size_t)k);
  int rc = or_tridiag_eig(k, lam, e, Z);
  if (rc == 0) {
    for (int j = 0; j < k; ++j) {
      double a = 0.0;
      for (int i = 0; i < k; ++i) a += Z[i + (size_t)j * k] * x[i];
      t[j] = a * (f ? f(s * lam[j]) : exp(s * lam[j]));
    }
    for (int i = 0; i < k; ++i) {
      double a = 0.0;
      for (int j = 0; j < k; ++j) a += Z[i + (size_t)j * k] * t[j];
      y[i] = a;
    }
  }
  free(lam);
  free(Z);
  free(t);
  return rc;
}

/* USERF/SYM f(H) step (serial/psfun_d_serial_mod.F90:420-485, DSYEV('V','U')
 * then DGEMV / f / DGEMV).  The symmetric matrix S of H's upper triangle is
 * reduced to tridiagonal form by Householder reflections (Golub & Van Loan,
 * Matrix Computations, Alg. 8.3.1) accumulated into Q, T's eigenvectors come
 * from or_tridiag_eig, and Z = Q Z_T.  w receives the eigenvalues, Z (k x k,
 * column-major) the eigenvectors.  Returns 0 on success. */
int or_sym_eig(int k, const double *H, int ldh, double *w, double *Z) {
  double *A = (double *)malloc(sizeof(double) * (size_t)k * k);
  double *Q = (double *)calloc((size_t)k * k, sizeof(double));
  double *Zt = (double *)malloc(sizeof(double) * (size_t)k * k);
  double *v = (double *)malloc(sizeof(double) * (size_t)k), *pv = (double *)malloc(sizeof(double) * (size_t)k);
  double *e = (double *)calloc((size_t)k + 1, sizeof(double));
#define AA(i, j) A[(i) + (size_t)(j)*k]
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) AA(i, j) = (i <= j) ? H[i + (size_t)j * ldh] : H[j + (size_t)i * ldh];
  for (int i = 0; i < k; ++i) Q[i + (size_t)i * k] = 1.0;
  for (int j = 0; j + 2 < k; ++j) {
    /* reflector P = I - beta v v^T with P x = alpha e1, x = A(j+1:k, j) */
    const int m = k - j - 1;
    double sig = 0.0;
    for (int i = 1; i < m; ++i) sig += AA(j + 1 + i, j) * AA(j + 1 + i, j);
    const double x0 = AA(j + 1, j);
    if (sig == 0.0) continue; /* already tridiagonal in this column */
    const double mu = sqrt(x0 * x0 + sig);
    const double alpha = (x0 <= 0.0) ? mu : -mu;
    v[0] = x0 - alpha;
    for (int i = 1; i < m; ++i) v[i] = AA(j + 1 + i, j);
    double vv = 0.0;
    for (int i = 0; i < m; ++i) vv += v[i] * v[i];
    const double beta = 2.0 / vv;
    /* A22 <- P A22 P: p = beta A22 v, q = p - (beta p.v / 2) v, A22 -= v q^T + q v^T */
    double pvv = 0.0;
    for (int r = 0; r < m; ++r) {
      double acc = 0.0;
      for (int c = 0; c < m; ++c) acc += AA(j + 1 + r, j + 1 + c) * v[c];
      pv[r] = beta * acc;
      pvv += pv[r] * v[r];
    }
    for (int r = 0; r < m; ++r) pv[r] -= 0.5 * beta * pvv * v[r];
    for (int c = 0; c < m; ++c)
      for (int r = 0; r < m; ++r) AA(j + 1 + r, j + 1 + c) -= v[r] * pv[c] + pv[r] * v[c];
    AA(j + 1, j) = alpha;
    AA(j, j + 1) = alpha;
    for (int i = 1; i < m; ++i) AA(j + 1 + i, j) = AA(j, j + 1 + i) = 0.0;
    /* Q <- Q P */
    for (int r = 0; r < k; ++r) {
      double acc = 0.0;
      for (int c = 0; c < m; ++c) acc += Q[r + (size_t)(j + 1 + c) * k] * v[c];
      acc *= beta;
      for (int c = 0; c < m; ++c) Q[r + (size_t)(j + 1 + c) * k] -= acc * v[c];
    }
  }
  for (int i = 0; i < k; ++i) w[i] = AA(i, i);
  for (int i = 0; i + 1 < k; ++i) e[i] = AA(i + 1, i);
#undef AA
  int rc = or_tridiag_eig(k, w, e, Zt);
  if (rc == 0)
    for (int c = 0; c < k; ++c)
      for (int r = 0; r < k; ++r) {
        double acc = 0.0;
        for (int l = 0; l < k; ++l) acc += Q[r + (size_t)l * k] * Zt[l + (size_t)c * k];
        Z[r + (size_t)c * k] = acc;
      }
  free(A); free(Q); free(Zt); free(v); free(pv); free(e);
  return rc;
}

/* y = Z diag(f(s*lambda)) Z^T x for the symmetric matrix of H's upper triangle. */
int or_userf_sym_apply(int k, const double *H, int ldh, double s, or_f_t f, const double *x, double *y) {
  double *w = (double *)malloc(sizeof(double) * (size_t)k), *Z = (double *)malloc(sizeof(double) * (size_t)k * k);
  double *t = (double *)malloc(sizeof(double) * (size_t)k);
  int rc = or_sym_eig(k, H, ldh, w, Z);
  if (rc == 0) {
    for (int j = 0; j < k; ++j) { /* t = Z^T x, scaled by f(s lambda_j) */
      double a = 0.0;
      for (int i = 0; i < k; ++i) a += Z[i + (size_t)j * k] * x[i];
      t[j] = f(s * w[j]) * a;
    }
    for (int i = 0; i < k; ++i) {
      double a = 0.0;
      for (int j = 0; j < k; ++j) a += Z[i + (size_t)j * k] * t[j];
      y[i] = a;
    }
  }
  free(w); free(Z); free(t);
  return rc;
}

/* ---- dense helpers (column-major) ---- */
static void mm(int k, const double *A, const double *B, double *C) {
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) {
      double s = 0.0;
      for (int l = 0; l < k; ++l) s += A[i + (size_t)l * k] * B[l + (size_t)j * k];
      C[i + (size_t)j * k] = s;
    }
}

/* Solve A X = B (k x k, nrhs columns) by Gaussian elimination with partial
 * pivoting; A and B are overwritten. */
static int gesv(int k, int nrhs, double *A, double *B) {
  for (int c = 0; c < k; ++c) {
    int p = c;
    for (int i = c + 1; i < k; ++i)
      if (fabs(A[i + (size_t)c * k]) > fabs(A[p + (size_t)c * k])) p = i;
    if (A[p + (size_t)c * k] == 0.0) return 1;
    if (p != c) {
      for (int j = 0; j < k; ++j) {
        double t = A[c + (size_t)j * k];
        A[c + (size_t)j * k] = A[p + (size_t)j * k];
        A[p + (size_t)j * k] = t;
      }
      for (int j = 0; j < nrhs; ++j) {
        double t = B[c + (size_t)j * k];
        B[c + (size_t)j * k] = B[p + (size_t)j * k];
        B[p + (size_t)j * k] = t;
      }
    }
    const double piv = A[c + (size_t)c * k];
    for (int i = c + 1; i < k; ++i) {
      const double f = A[i + (size_t)c * k] / piv;
      if (f == 0.0) continue;
      for (int j = c; j < k; ++j) A[i + (size_t)j * k] -= f * A[c + (size_t)j * k];
      for (int j = 0; j < nrhs; ++j) B[i + (size_t)j * k] -= f * B[c + (size_t)j * k];
    }
  }
  for (int j = 0; j < nrhs; ++j)
    for (int i = k - 1; i >= 0; --i) {
      double s = B[i + (size_t)j * k];
      for (int l = i + 1; l < k; ++l) s -= A[i + (size_t)l * k] * B[l + (size_t)j * k];
      B[i + (size_t)j * k] = s / A[i + (size_t)i * k];
    }
  return 0;
}

/* E = exp(s*H), H k x k column-major: [13/13] Pade approximant with scaling
 * and squaring (Higham, SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193). */
int or_dense_expm(int k, const double *H, double s, double *E) {
  static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                               1187353796428800.0,  129060195264000.0,   10559470521600.0,
                               670442572800.0,      33522128640.0,       1323241920.0,
                               40840800.0,          960960.0,            16380.0,
                               182.0,               1.0};
  const size_t kk = (size_t)k * k;
  double *A = (double *)malloc(sizeof(double) * kk), *A2 = (double *)malloc(sizeof(double) * kk),
         *A4 = (double *)malloc(sizeof(double) * kk), *A6 = (double *)malloc(sizeof(double) * kk),
         *U = (double *)malloc(sizeof(double) * kk), *V = (double *)malloc(sizeof(double) * kk),
         *T = (double *)malloc(sizeof(double) * kk), *P = (double *)malloc(sizeof(double) * kk);
  double nrm1 = 0.0;
  for (int j = 0; j < k; ++j) {
    double cs = 0.0;
    for (int i = 0; i < k; ++i) {
      A[i + (size_t)j * k] = s * H[i + (size_t)j * k];
      cs += fabs(A[i + (size_t)j * k]);
    }
    nrm1 = fmax(nrm1, cs);
  }
  int sq = 0;
  const double theta13 = 5.371920351148152;
  if (nrm1 > theta13) sq = (int)ceil(log2(nrm1 / theta13));
  if (sq < 0) sq = 0;
  const double sc = ldexp(1.0, -sq);
  for (size_t i = 0; i < kk; ++i) A[i] *= sc;
  mm(k, A, A, A2);
  mm(k, A2, A2, A4);
  mm(k, A4, A2, A6);
  /* U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I] */
  for (size_t i = 0; i < kk; ++i) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
  mm(k, A6, T, P);
  for (size_t i = 0; i < kk; ++i) P[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
  for (int i = 0; i < k; ++i) P[i + (size_t)i * k] += b[1];
  mm(k, A, P, U);
  /* V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I */
  for (size_t i = 0; i < kk; ++i) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
  mm(k, A6, T, V);
  for (size_t i = 0; i < kk; ++i) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
  for (int i = 0; i < k; ++i) V[i + (size_t)i * k] += b[0];
  /* (V - U) E = (V + U) */
  for (size_t i = 0; i < kk; ++i) {
    T[i] = V[i] - U[i];
    E[i] = V[i] + U[i];
  }
  int rc = gesv(k, k, T, E);
  for (int q = 0; q < sq && rc == 0; ++q) {
    mm(k, E, E, T);
    memcpy(E, T, sizeof(double) * kk);
  }
  free(A); free(A2); free(A4); free(A6); free(U); free(V); free(T); free(P);
  return rc;
}

/* E = exp(t*H), H k x k column-major, by the diagonal Pade approximant
 * r_p = N/D of degree p = ideg >= 2 with scaling and squaring, the method of
 * EXPOKIT's DGPADM (Sidje, ACM TOMS 24 (1998) 130-156; the reference calls it
 * with fun%padedegree, serial/psfun_d_serial_mod.F90:271):
 *   ns = max(0, int(log2(|t| ||H||_inf)) + 2), X = (t / 2^ns) H,
 *   c_0 = 1, c_j = c_{j-1} (p+1-j) / (j (2p+1-j)),  N(X) = sum_j c_j X^j,
 *   D(X) = N(-X);  E = (D^{-1} N)^(2^ns).
 * Evaluated here as N = Ev + X Od, D = Ev - X Od with Ev, Od polynomials in
 * X^2 (Horner), one solve D E = N, then ns squarings.  Returns 1 on a
 * singular D, 2 for ideg < 2; H = 0 gives I (where DGPADM stops). */
int or_pade_expm(int ideg, int k, const double *H, double t, double *E) {
  if (ideg < 2) return 2;
  const size_t kk = (size_t)k * k;
  double hn = 0.0;
  for (int i = 0; i < k; ++i) {
    double rs = 0.0;
    for (int j = 0; j < k; ++j) rs += fabs(H[i + (size_t)j * k]);
    hn = fmax(hn, rs);
  }
  hn = fabs(t * hn);
  if (hn == 0.0) {
    for (size_t i = 0; i < kk; ++i) E[i] = (i % (size_t)(k + 1) == 0) ? 1.0 : 0.0;
    return 0;
  }
  int ns = (int)(log(hn) / log(2.0)) + 2;
  if (ns < 0) ns = 0;
  double *c = (double *)malloc(sizeof(double) * (ideg + 1));
  c[0] = 1.0;
  for (int j = 1; j <= ideg; ++j) c[j] = c[j - 1] * (double)(ideg + 1 - j) / (double)(j * (2 * ideg + 1 - j));
  double *X = (double *)malloc(sizeof(double) * kk), *X2 = (double *)malloc(sizeof(double) * kk),
         *Ev = (double *)calloc(kk, sizeof(double)), *Od = (double *)calloc(kk, sizeof(double)),
         *T = (double *)malloc(sizeof(double) * kk), *D = (double *)malloc(sizeof(double) * kk);
  const double sc = ldexp(t, -ns);
  for (size_t i = 0; i < kk; ++i) X[i] = sc * H[i];
  mm(k, X, X, X2);
  /* Ev = sum_{j even} c_j X^j and Od = sum_{j odd} c_j X^(j-1), Horner in X^2 */
  const int je = (ideg % 2 == 0) ? ideg : ideg - 1, jo = (ideg % 2 == 1) ? ideg : ideg - 1;
  for (int i = 0; i < k; ++i) {
    Ev[i + (size_t)i * k] = c[je];
    Od[i + (size_t)i * k] = c[jo];
  }
  for (int j = je - 2; j >= 0; j -= 2) {
    mm(k, Ev, X2, T);
    for (int i = 0; i < k; ++i) T[i + (size_t)i * k] += c[j];
    memcpy(Ev, T, sizeof(double) * kk);
  }
  for (int j = jo - 2; j >= 1; j -= 2) {
    mm(k, Od, X2, T);
    for (int i = 0; i < k; ++i) T[i + (size_t)i * k] += c[j];
    memcpy(Od, T, sizeof(double) * kk);
  }
  mm(k, X, Od, T); /* the odd part X Od */
  for (size_t i = 0; i < kk; ++i) {
    D[i] = Ev[i] - T[i];
    E[i] = Ev[i] + T[i];
  }
  int rc = gesv(k, k, D, E) ? 1 : 0;
  for (int q = 0; q < ns && rc == 0; ++q) {
    mm(k, E, E, T);
    memcpy(E, T, sizeof(double) * kk);
  }
  free(c); free(X); free(X2); free(Ev); free(Od); free(T); free(D);
  return rc;
}

/* The dense exp step of the restatements below follows the GENPADE degree
 * when one other than 6 is set (or_set_pade_degree), else or_dense_expm. */
static int g_pade = 0;
void or_set_pade_degree(int p) { g_pade = p; }

/* y = exp(s*H) x for a dense k x k column-major H. */
int or_dense_expm_apply(int k, const double *H, double s, const double *x, double *y) {
  double *E = (double *)malloc(sizeof(double) * (size_t)k * k);
  int rc = (g_pade >= 2 && g_pade != 6) ? or_pade_expm(g_pade, k, H, s, E) : or_dense_expm(k, H, s, E);
  if (rc == 0)
    for (int i = 0; i < k; ++i) {
      double a = 0.0;
      for (int j = 0; j < k; ++j) a += E[i + (size_t)j * k] * x[j];
      y[i] = a;
    }
  free(E);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Lanczos f(A)x  (krylov/polynomial/psfun_d_lanczos.f90:97-386)       */
/* ------------------------------------------------------------------ */
/* f = exp(fscale * .), Krylov operator opscale*A (the reference passes 1).
 * res must hold itmax+1 doubles.  Returns 0 or an error code:
 * 2 = invalid istop (lines 108-113), 1 = other invalid argument. */
/* f = NULL: EXP (CHBSYM); otherwise the USERF/SYM step with that f. */
int or_lanczos_f(int n, const int *rp, const int *ci, const double *va, double fscale, double opscale,
                 const double *x, double *y, double eps, int itmax, int istop, int *iter, double *err,
                 double *res, int *nres, or_f_t f) {
  if (istop < 1 || istop > 2) return 2;
  if (itmax < 2) return 1;
  const int nl = itmax;
  double *V = (double *)malloc(sizeof(double) * (size_t)n * (nl + 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *T = (double *)calloc((size_t)(nl + 1) * (nl + 1), sizeof(double)); /* T(i,j) -> T[(i-1)+(j-1)*(nl+1)] */
  double *rs = (double *)calloc((size_t)nl + 1, sizeof(double));
  double *yk = NULL;
  int nyk = 0;
#define TT(i, j) T[((i)-1) + (size_t)((j)-1) * (nl + 1)]
#define VV(i) (V + (size_t)((i)-1) * n)
  double errnum = 0.0, errden = 1.0;
  /* v(1) = x; rs(1) = ||v(1)||; v(1) *= 1/rs(1)   (183-196) */
  or_axpby(n, 1.0, x, 0.0, VV(1));
  rs[0] = or_nrm2(n, VV(1));
  or_axpby(n, 1.0 / rs[0], VV(1), 0.0, VV(1));
  /* w = A v(1); t(1,1) = w.v(1); w -= t(1,1) v(1)   (198-217) */
  or_spmv(n, rp, ci, va, opscale, VV(1), 0.0, w);
  TT(1, 1) = or_dot(n, w, VV(1));
  or_axpby(n, -TT(1, 1), VV(1), 1.0, w);
  int itx = 0;
  for (int i = 2; i <= nl; ++i) { /* 219-326 */
    itx++;
    TT(i - 1, i) = or_nrm2(n, w);
    TT(i, i - 1) = TT(i - 1, i);
    if (fabs(TT(i - 1, i)) < 1e-16) break;
    or_axpby(n, 1.0 / TT(i - 1, i), w, 0.0, VV(i));
    or_spmv(n, rp, ci, va, opscale, VV(i), 0.0, w);
    TT(i, i) = or_dot(n, w, VV(i));
    or_axpby(n, -TT(i, i), VV(i), 1.0, w);
    or_axpby(n, -TT(i - 1, i), VV(i - 1), 1.0, w);
    if (istop == 1) {
      const int k = i - 1;
      double *d = (double *)malloc(sizeof(double) * k), *e = (double *)malloc(sizeof(double) * (k + 1));
      double *e1 = (double *)calloc((size_t)k, sizeof(double));
      free(yk);
      yk = (double *)malloc(sizeof(double) * k);
      nyk = k;
      for (int q = 1; q <= k; ++q) d[q - 1] = TT(q, q);
      for (int q = 1; q < k; ++q) e[q - 1] = TT(q, q + 1);
      e1[0] = 1.0;
      tridiag_fapply(k, d, e, fscale, f, e1, yk); /* fun%apply(t(1:i-1,1:i-1),yk,e)  (292) */
      free(d); free(e); free(e1);
      errnum = rs[0] * TT(i, i - 1) * fabs(yk[i - 2]);
      double s2 = 0.0;
      for (int q = 0; q < k; ++q) s2 += yk[q] * yk[q];
      errden = sqrt(s2);
      rs[i - 1] = errnum / errden;
      if (rs[i - 1] < eps) break;
      free(yk);
      yk = NULL;
      nyk = 0;
    }
  }
  if (!yk) { /* 330-362: yk = f(T(1:itx-1,1:itx-1)) e1; itx-1 == 0 is undefined in
                the reference and is taken as the 1x1 block here */
    int k = itx - 1;
    if (k < 1) k = 1;
    double *d = (double *)malloc(sizeof(double) * k), *e = (double *)malloc(sizeof(double) * (k + 1));
    double *e1 = (double *)calloc((size_t)k, sizeof(double));
    yk = (double *)malloc(sizeof(double) * k);
    nyk = k;
    for (int q = 1; q <= k; ++q) d[q - 1] = TT(q, q);
    for (int q = 1; q < k; ++q) e[q - 1] = TT(q, q + 1);
    e1[0] = 1.0;
    tridiag_fapply(k, d, e, fscale, f, e1, yk);
    free(d); free(e); free(e1);
  }
  /* log_end: err = errnum/errden, iter = itx   (365-371) */
  const double derr = (errden == 0.0) ? errnum : errnum / errden;
  if (iter) *iter = itx;
  if (err) *err = derr;
  if (res) {
    for (int q = 0; q < itx; ++q) res[q] = rs[q];
    res[itx] = derr;
  }
  if (nres) *nres = itx + 1;
  /* y = rs(1) (yk(1) v(1) + sum_{i=2}^{itx-1} yk(i) v(i))   (373-377) */
  or_axpby(n, yk[0], VV(1), 0.0, y);
  for (int q = 2; q <= itx - 1 && q <= nyk; ++q) or_axpby(n, yk[q - 1], VV(q), 1.0, y);
  or_axpby(n, rs[0], y, 0.0, y);
#undef TT
#undef VV
  free(V); free(w); free(T); free(rs); free(yk);
  return 0;
}

int or_lanczos(int n, const int *rp, const int *ci, const double *va, double fscale, double opscale,
               const double *x, double *y, double eps, int itmax, int istop, int *iter, double *err,
               double *res, int *nres) {
  return or_lanczos_f(n, rp, ci, va, fscale, opscale, x, y, eps, itmax, istop, iter, err, res, nres, NULL);
}

/* Many independent or_lanczos runs: graph g owns rows off[g]..off[g+1]-1 of
 * one block-diagonal 0-based CSR (columns global, inside the graph's rows);
 * X and Y are row-major nrows x K, instance g*K+k takes column k of graph g.
 * Each instance is exactly or_lanczos on the graph's own CSR (local column
 * numbers); the loop over instances is what runs in parallel, every instance
 * itself single-threaded.  iters/errs (ngraph*K) may be NULL.  Returns the
 * first nonzero status, else 0. */
int or_lanczos_batch(int ngraph, const int64_t *off, const int *rp, const int *ci, const double *va, int K,
                     double fscale, double opscale, const double *X, double *Y, double eps, int itmax, int istop,
                     int *iters, double *errs) {
  int status = 0;
  const int saved = g_threads;
  g_threads = 1; /* the primitives run serially inside each instance */
#pragma omp parallel for schedule(dynamic, 1) reduction(max : status)
  for (int64_t inst = 0; inst < (int64_t)ngraph * K; ++inst) {
    const int g = (int)(inst / K), k = (int)(inst % K);
    const int64_t r0 = off[g];
    const int n = (int)(off[g + 1] - r0);
    const int base = rp[r0], nz = rp[r0 + n] - base;
    int *lrp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *lci = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    double *x = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *res = (double *)malloc(sizeof(double) * (size_t)(itmax + 2));
    for (int i = 0; i <= n; ++i) lrp[i] = rp[r0 + i] - base;
    for (int j = 0; j < nz; ++j) lci[j] = ci[base + j] - (int)r0;
    for (int i = 0; i < n; ++i) x[i] = X[(r0 + i) * K + k];
    int it = 0, nres = 0;
    double er = 0.0;
    const int rc = or_lanczos(n, lrp, lci, va + base, fscale, opscale, x, y, eps, itmax, istop, &it, &er, res, &nres);
    for (int i = 0; i < n; ++i) Y[(r0 + i) * K + k] = y[i];
    if (iters) iters[inst] = it;
    if (errs) errs[inst] = er;
    if (rc > status) status = rc;
    free(lrp); free(lci); free(x); free(y); free(res);
  }
  g_threads = saved;
  return status;
}

/* ------------------------------------------------------------------ */
/* Arnoldi f(A)x  (krylov/polynomial/psfun_d_arnoldi.f90:172-329)      */
/* ------------------------------------------------------------------ */
/* The Krylov operator is scaling*A (line 202) and f(H) uses scaling 1
 * (line 174); fscale multiplies H inside f for callers that want it. */
/* f = NULL: EXP of the full H; otherwise USERF/SYM (upper triangle of H). */
int or_arnoldi_f(int n, const int *rp, const int *ci, const double *va, double fscale, double scaling,
                 const double *x, double *y, double eps, int itmax, int istop, int *iter, double *err,
                 double *res, int *nres, or_f_t f) {
  if (istop < 1 || istop > 2) return 2;
  if (itmax < 2) return 1;
  const int nl = itmax;
  double *V = (double *)malloc(sizeof(double) * (size_t)n * (nl + 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *Hm = (double *)calloc((size_t)(nl + 1) * (nl + 1), sizeof(double));
  double *rs = (double *)calloc((size_t)nl + 1, sizeof(double));
  double *yk = NULL;
  int nyk = 0;
#define HH(i, j) Hm[((i)-1) + (size_t)((j)-1) * (nl + 1)]
#define VV(i) (V + (size_t)((i)-1) * n)
  double errnum = 0.0, errden = 1.0;
  or_axpby(n, 1.0, x, 0.0, VV(1));
  rs[0] = or_nrm2(n, VV(1));
  or_axpby(n, 1.0 / rs[0], VV(1), 0.0, VV(1));
  int itx = 0;
  for (int i = 2; i <= nl; ++i) {
    itx++;
    or_spmv(n, rp, ci, va, scaling, VV(i - 1), 0.0, w);
    for (int k = 1; k <= i - 1; ++k) {
      HH(k, i - 1) = or_dot(n, VV(k), w);
      or_axpby(n, -HH(k, i - 1), VV(k), 1.0, w);
    }
    HH(i, i - 1) = or_nrm2(n, w);
    or_axpby(n, 1.0 / HH(i, i - 1), w, 0.0, VV(i));
    if (istop == 1) {
      const int k = i - 1;
      double *Hk = (double *)malloc(sizeof(double) * k * k), *e1 = (double *)calloc((size_t)k, sizeof(double));
      for (int c = 1; c <= k; ++c)
        for (int r = 1; r <= k; ++r) Hk[(r - 1) + (size_t)(c - 1) * k] = HH(r, c);
      e1[0] = 1.0;
      free(yk);
      yk = (double *)malloc(sizeof(double) * k);
      nyk = k;
      if (f) or_userf_sym_apply(k, Hk, k, fscale, f, e1, yk);
      else or_dense_expm_apply(k, Hk, fscale, e1, yk);
      free(Hk); free(e1);
      errnum = rs[0] * HH(i, i - 1) * fabs(yk[i - 2]);
      double s2 = 0.0;
      for (int q = 0; q < k; ++q) s2 += yk[q] * yk[q];
      errden = sqrt(s2);
      rs[i - 1] = errnum / errden;
      if (rs[i - 1] < eps) break;
      free(yk);
      yk = NULL;
      nyk = 0;
    }
  }
  if (!yk) {
    int k = itx - 1;
    if (k < 1) k = 1;
    double *Hk = (double *)malloc(sizeof(double) * k * k), *e1 = (double *)calloc((size_t)k, sizeof(double));
    for (int c = 1; c <= k; ++c)
      for (int r = 1; r <= k; ++r) Hk[(r - 1) + (size_t)(c - 1) * k] = HH(r, c);
    e1[0] = 1.0;
    yk = (double *)malloc(sizeof(double) * k);
    nyk = k;
    if (f) or_userf_sym_apply(k, Hk, k, fscale, f, e1, yk);
      else or_dense_expm_apply(k, Hk, fscale, e1, yk);
    free(Hk); free(e1);
  }
  const double derr = (errden == 0.0) ? errnum : errnum / errden;
  if (iter) *iter = itx;
  if (err) *err = derr;
  if (res) {
    for (int q = 0; q < itx; ++q) res[q] = rs[q];
    res[itx] = derr;
  }
  if (nres) *nres = itx + 1;
  or_axpby(n, yk[0], VV(1), 0.0, y);
  for (int q = 2; q <= itx - 1 && q <= nyk; ++q) or_axpby(n, yk[q - 1], VV(q), 1.0, y);
  or_axpby(n, rs[0], y, 0.0, y);
#undef HH
#undef VV
  free(V); free(w); free(Hm); free(rs); free(yk);
  return 0;
}

int or_arnoldi(int n, const int *rp, const int *ci, const double *va, double fscale, double scaling,
               const double *x, double *y, double eps, int itmax, int istop, int *iter, double *err,
               double *res, int *nres) {
  return or_arnoldi_f(n, rp, ci, va, fscale, scaling, x, y, eps, itmax, istop, iter, err, res, nres, NULL);
}

/* ------------------------------------------------------------------ */
/* Shift-and-invert Arnoldi (krylov/rational/psfun_d_sai.f90:183-390)  */
/* ------------------------------------------------------------------ */
/* Inverse of a k x k column-major matrix by Gauss-Jordan elimination with
 * partial pivoting (replaces the reference's dgetrf/dgetri, sai:269-289).
 * Returns 0 on success, 1 on a zero pivot. */
int or_dense_inverse(int k, const double *H, double *Hinv) {
  double *W = (double *)malloc(sizeof(double) * (size_t)k * k);
  memcpy(W, H, sizeof(double) * (size_t)k * k);
  for (int i = 0; i < k * k; ++i) Hinv[i] = 0.0;
  for (int i = 0; i < k; ++i) Hinv[i + (size_t)i * k] = 1.0;
#define WW(i, j) W[(i) + (size_t)(j) * k]
#define II(i, j) Hinv[(i) + (size_t)(j) * k]
  for (int c = 0; c < k; ++c) {
    int p = c;
    for (int r = c + 1; r < k; ++r)
      if (fabs(WW(r, c)) > fabs(WW(p, c))) p = r;
    if (WW(p, c) == 0.0) {
      free(W);
      return 1;
    }
    if (p != c)
      for (int j = 0; j < k; ++j) {
        double t = WW(c, j); WW(c, j) = WW(p, j); WW(p, j) = t;
        t = II(c, j); II(c, j) = II(p, j); II(p, j) = t;
      }
    const double inv = 1.0 / WW(c, c);
    for (int j = 0; j < k; ++j) {
      WW(c, j) *= inv;
      II(c, j) *= inv;
    }
    for (int r = 0; r < k; ++r) {
      if (r == c) continue;
      const double f = WW(r, c);
      if (f == 0.0) continue;
      for (int j = 0; j < k; ++j) {
        WW(r, j) -= f * WW(c, j);
        II(r, j) -= f * II(c, j);
      }
    }
  }
#undef WW
#undef II
  free(W);
  return 0;
}

/* y = f(H_k^{-1} - tau I) e1 (the SAI projection of A). */
static int sai_fproj(int k, const double *Hm, int ldh, double tau, double fscale, or_f_t f, double *yk) {
  double *Hk = (double *)malloc(sizeof(double) * k * k), *Hi = (double *)malloc(sizeof(double) * k * k);
  double *e1 = (double *)calloc((size_t)k, sizeof(double));
  for (int c = 0; c < k; ++c)
    for (int r = 0; r < k; ++r) Hk[r + (size_t)c * k] = Hm[r + (size_t)c * ldh];
  int rc = or_dense_inverse(k, Hk, Hi);
  for (int i = 0; i < k; ++i) Hi[i + (size_t)i * k] -= tau;
  e1[0] = 1.0;
  if (!rc) rc = f ? or_userf_sym_apply(k, Hi, k, fscale, f, e1, yk) : or_dense_expm_apply(k, Hi, fscale, e1, yk);
  free(Hk); free(Hi); free(e1);
  return rc;
}

int or_cg_shifted(int n, const int *rp, const int *ci, const double *va, int prec, const double *b, double eta,
                  double zeta, double *x, double eps, int itmax, int istop, int *iter, double *err, double *cond);

/* Krylov space of (A + tau I)^{-1}, each inner solve by PCG (prec 0/1) with
 * tolerance eps, inner itmax/istop, zero initial guess; MGS as Arnoldi; the
 * estimate and the fallback both from f(H_k^{-1} - tau I) e1 with the shift
 * on every diagonal entry; assembly as Arnoldi. */
int or_sai_arnoldi_f(int n, const int *rp, const int *ci, const double *va, double fscale, double tau, int prec,
                     int in_itmax, int in_istop, const double *x, double *y, double eps, int itmax, int istop,
                     int *iter, double *err, double *res, int *nres, or_f_t f) {
  if (istop < 1 || istop > 2) return 2;
  if (itmax < 2) return 1;
  const int nl = itmax;
  double *V = (double *)malloc(sizeof(double) * (size_t)n * (nl + 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *Hm = (double *)calloc((size_t)(nl + 1) * (nl + 1), sizeof(double));
  double *rs = (double *)calloc((size_t)nl + 1, sizeof(double));
  double *yk = NULL;
  int nyk = 0;
#define HH(i, j) Hm[((i)-1) + (size_t)((j)-1) * (nl + 1)]
#define VV(i) (V + (size_t)((i)-1) * n)
  double errnum = 0.0, errden = 1.0;
  or_axpby(n, 1.0, x, 0.0, VV(1));
  rs[0] = or_nrm2(n, VV(1));
  or_axpby(n, 1.0 / rs[0], VV(1), 0.0, VV(1));
  int itx = 0;
  for (int i = 2; i <= nl; ++i) {
    itx++;
    int it_in = 0;
    double err_in = 0.0;
    memset(w, 0, sizeof(double) * (size_t)n);
    or_cg_shifted(n, rp, ci, va, prec, VV(i - 1), 1.0, tau, w, eps, in_itmax, in_istop, &it_in, &err_in, NULL);
    for (int k = 1; k <= i - 1; ++k) {
      HH(k, i - 1) = or_dot(n, VV(k), w);
      or_axpby(n, -HH(k, i - 1), VV(k), 1.0, w);
    }
    HH(i, i - 1) = or_nrm2(n, w);
    or_axpby(n, 1.0 / HH(i, i - 1), w, 0.0, VV(i));
    if (istop == 1) {
      const int k = i - 1;
      free(yk);
      yk = (double *)malloc(sizeof(double) * k);
      nyk = k;
      sai_fproj(k, Hm, nl + 1, tau, fscale, f, yk);
      errnum = rs[0] * HH(i, i - 1) * fabs(yk[k - 1]);
      double s2 = 0.0;
      for (int q = 0; q < k; ++q) s2 += yk[q] * yk[q];
      errden = sqrt(s2);
      rs[i - 1] = errnum / errden;
      if (rs[i - 1] < eps) break;
      free(yk);
      yk = NULL;
      nyk = 0;
    }
  }
  if (!yk) {
    int k = itx - 1;
    if (k < 1) k = 1;
    yk = (double *)malloc(sizeof(double) * k);
    nyk = k;
    sai_fproj(k, Hm, nl + 1, tau, fscale, f, yk);
  }
  const double derr = (errden == 0.0) ? errnum : errnum / errden;
  if (iter) *iter = itx;
  if (err) *err = derr;
  if (res) {
    for (int q = 0; q < itx; ++q) res[q] = rs[q];
    res[itx] = derr;
  }
  if (nres) *nres = itx + 1;
  or_axpby(n, yk[0], VV(1), 0.0, y);
  for (int q = 2; q <= itx - 1 && q <= nyk; ++q) or_axpby(n, yk[q - 1], VV(q), 1.0, y);
  or_axpby(n, rs[0], y, 0.0, y);
#undef HH
#undef VV
  free(V); free(w); free(Hm); free(rs); free(yk);
  return 0;
}

int or_sai_arnoldi(int n, const int *rp, const int *ci, const double *va, double fscale, double tau, int prec,
                   int in_itmax, int in_istop, const double *x, double *y, double eps, int itmax, int istop,
                   int *iter, double *err, double *res, int *nres) {
  return or_sai_arnoldi_f(n, rp, ci, va, fscale, tau, prec, in_itmax, in_istop, x, y, eps, itmax, istop, iter, err,
                          res, nres, NULL);
}

/* ------------------------------------------------------------------ */
/* Shifted PCG (krylov/shifted/psfun_dcg.F90:249-357)                  */
/* ------------------------------------------------------------------ */
/* ILU(0) on A's own sparsity pattern: the BJAC preconditioner PSBLAS builds
 * for prec%init(ctxt, 'BJAC', info) with one process (one diagonal block, the
 * whole matrix; default block solver ILU(0)).  PSBLAS is not part of the
 * reference tree: this follows the textbook IKJ algorithm (Saad, Iterative
 * Methods for Sparse Linear Systems, 2nd ed., Alg. 10.4) on rows sorted by
 * column with repeated entries summed in input order, with explicit fma.
 * f->lu holds the strict lower part of L (unit diagonal) and U in one CSR;
 * f->dp the position of each diagonal.  Returns 0, or 10 on a zero or
 * missing pivot. */
typedef struct {
  int n;
  int *rp, *ci, *dp;
  double *lu;
} or_ilu_t;

static void ilu0_free(or_ilu_t *f) {
  free(f->rp); free(f->ci); free(f->dp); free(f->lu);
}

static int ilu0_build(int n, const int *rp, const int *ci, const double *va, or_ilu_t *f) {
  f->n = n;
  f->rp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  f->ci = (int *)malloc(sizeof(int) * (size_t)(rp[n] > 0 ? rp[n] : 1));
  f->lu = (double *)malloc(sizeof(double) * (size_t)(rp[n] > 0 ? rp[n] : 1));
  f->dp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *where = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) where[i] = -1;
  int nz = 0;
  f->rp[0] = 0;
  for (int i = 0; i < n; ++i) { /* sorted row, repeats summed in input order (stable insertion) */
    const int start = nz;
    for (int q = rp[i]; q < rp[i + 1]; ++q) {
      int pos = start;
      while (pos < nz && f->ci[pos] <= ci[q]) ++pos;
      if (pos > start && f->ci[pos - 1] == ci[q]) {
        f->lu[pos - 1] += va[q];
        continue;
      }
      for (int t = nz; t > pos; --t) {
        f->ci[t] = f->ci[t - 1];
        f->lu[t] = f->lu[t - 1];
      }
      f->ci[pos] = ci[q];
      f->lu[pos] = va[q];
      ++nz;
    }
    f->rp[i + 1] = nz;
  }
  int rc = 0;
  for (int i = 0; i < n && !rc; ++i) {
    f->dp[i] = -1;
    for (int q = f->rp[i]; q < f->rp[i + 1]; ++q) where[f->ci[q]] = q;
    for (int q = f->rp[i]; q < f->rp[i + 1]; ++q) {
      const int k = f->ci[q];
      if (k >= i) break;
      const int dk = f->dp[k];
      if (dk < 0 || f->lu[dk] == 0.0) {
        rc = 10;
        break;
      }
      const double lik = f->lu[q] / f->lu[dk];
      f->lu[q] = lik;
      for (int t = dk + 1; t < f->rp[k + 1]; ++t) {
        const int wj = where[f->ci[t]];
        if (wj >= 0) f->lu[wj] = fma(-lik, f->lu[t], f->lu[wj]);
      }
    }
    for (int q = f->rp[i]; q < f->rp[i + 1]; ++q) {
      if (f->ci[q] == i) f->dp[i] = q;
      where[f->ci[q]] = -1;
    }
    if (!rc && (f->dp[i] < 0 || f->lu[f->dp[i]] == 0.0)) rc = 10;
  }
  free(where);
  return rc;
}

/* z = U^{-1} L^{-1} r: forward then backward substitution, terms in column order */
static void ilu0_solve(const or_ilu_t *f, const double *r, double *z) {
  const int n = f->n;
  for (int i = 0; i < n; ++i) {
    double t = r[i];
    for (int q = f->rp[i]; q < f->dp[i]; ++q) t = fma(-f->lu[q], z[f->ci[q]], t);
    z[i] = t;
  }
  for (int i = n - 1; i >= 0; --i) {
    double t = z[i];
    for (int q = f->dp[i] + 1; q < f->rp[i + 1]; ++q) t = fma(-f->lu[q], z[f->ci[q]], t);
    z[i] = t / f->lu[f->dp[i]];
  }
}

/* The factors as a sorted CSR (out_rp n+1, out_ci/out_lu up to nnz); returns
 * the status of ilu0_build and the stored count in *nnz_out. */
int or_ilu0(int n, const int *rp, const int *ci, const double *va, int *out_rp, int *out_ci, double *out_lu,
            int *nnz_out) {
  or_ilu_t f;
  const int rc = ilu0_build(n, rp, ci, va, &f);
  for (int i = 0; i <= n; ++i) out_rp[i] = f.rp[i];
  for (int q = 0; q < f.rp[n]; ++q) {
    out_ci[q] = f.ci[q];
    out_lu[q] = f.lu[q];
  }
  *nnz_out = f.rp[n];
  ilu0_free(&f);
  return rc;
}

/* z = M^{-1} r with M the ILU(0) of A */
int or_ilu0_apply(int n, const int *rp, const int *ci, const double *va, const double *r, double *z) {
  or_ilu_t f;
  const int rc = ilu0_build(n, rp, ci, va, &f);
  if (!rc) ilu0_solve(&f, r, z);
  ilu0_free(&f);
  return rc;
}

/* (eta*A + zeta*I) x = b, x holds the initial guess.  prec: 0 = none,
 * 1 = diagonal of the unshifted A (the DIAG preconditioner built on A),
 * 2 = ILU(0) of the unshifted A (BJAC).
 * Stopping follows PSBLAS psb_init_conv/psb_check_conv/psb_end_conv:
 * istop 1: ||r||_inf <= eps (||A||_inf ||x||_inf + ||b||_inf);
 * istop 2: ||r||_2 <= eps ||b||_2; and it >= itmax stops.
 * cond (if non-null) is the ratio of extreme eigenvalues of the CG
 * tridiagonal (dstebz in the reference, lines 314-354). */
int or_cg_shifted(int n, const int *rp, const int *ci, const double *va, int prec, const double *b, double eta,
                  double zeta, double *x, double eps, int itmax, int istop, int *iter, double *err, double *cond) {
  if (istop < 1 || istop > 2) return 2;
  double *p = (double *)calloc((size_t)n, sizeof(double)), *q = (double *)calloc((size_t)n, sizeof(double)),
         *r = (double *)calloc((size_t)n, sizeof(double)), *z = (double *)calloc((size_t)n, sizeof(double)),
         *dinv = NULL;
  or_ilu_t ilu;
  int have_ilu = 0;
  if (prec == 2) {
    if (ilu0_build(n, rp, ci, va, &ilu)) {
      ilu0_free(&ilu);
      free(p); free(q); free(r); free(z);
      return 10;
    }
    have_ilu = 1;
  }
  if (prec == 1) {
    dinv = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
      double dv = 0.0;
      for (int j = rp[i]; j < rp[i + 1]; ++j)
        if (ci[j] == i) dv += va[j];
      dinv[i] = (dv != 0.0) ? 1.0 / dv : 1.0;
    }
  }
  double *td = cond ? (double *)calloc((size_t)itmax + 2, sizeof(double)) : NULL;
  double *tu = cond ? (double *)calloc((size_t)itmax + 2, sizeof(double)) : NULL;
  int istebz = 0;
  int itx = 0;
  double alpha = 0.0, beta = 0.0, rho = 0.0, rho_old, sigma, alpha_old;
  double ani = 0.0, bni = 0.0, bn2 = 0.0, errnum = 0.0, errden = 1.0;
  int conv = 0;
  for (;;) { /* restart */
    if (itx >= itmax) break;
    int it = 0;
    or_axpby(n, 1.0, b, 0.0, r);
    or_spmv(n, rp, ci, va, -eta, x, 1.0, r);
    or_axpby(n, -zeta, x, 1.0, r);
    rho = 0.0;
    /* psb_init_conv (uses ||A|| of the unshifted A, line 272) */
    if (istop == 1) {
      ani = or_spnrmi(n, rp, ci, va);
      bni = or_amax(n, b);
    } else {
      bn2 = or_nrm2(n, b);
    }
    for (;;) { /* iteration */
      it++;
      itx++;
      if (prec == 1)
        for (int i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
      else if (prec == 2)
        ilu0_solve(&ilu, r, z);
      else
        or_axpby(n, 1.0, r, 0.0, z);
      rho_old = rho;
      rho = or_dot(n, r, z);
      if (it == 1) {
        or_axpby(n, 1.0, z, 0.0, p);
      } else {
        if (rho_old == 0.0) break;
        beta = rho / rho_old;
        or_axpby(n, 1.0, z, beta, p);
      }
      or_spmv(n, rp, ci, va, eta, p, 0.0, q);
      or_axpby(n, zeta, p, 1.0, q);
      sigma = or_dot(n, p, q);
      if (sigma == 0.0) break;
      alpha_old = alpha;
      alpha = rho / sigma;
      if (cond) {
        istebz++;
        if (istebz == 1)
          td[0] = 1.0 / alpha;
        else {
          td[istebz - 1] = 1.0 / alpha + beta / alpha_old;
          tu[istebz - 2] = sqrt(beta) / alpha_old;
        }
      }
      or_axpby(n, alpha, p, 1.0, x);
      or_axpby(n, -alpha, q, 1.0, r);
      /* psb_check_conv */
      if (istop == 1) {
        errnum = or_amax(n, r);
        errden = ani * or_amax(n, x) + bni;
      } else {
        errnum = or_nrm2(n, r);
        errden = bn2;
      }
      const int ok = (errden == 0.0) ? (errnum <= eps) : (errnum <= eps * errden);
      if (ok || itx >= itmax) {
        conv = 1;
        break;
      }
    }
    if (conv) break;
  }
  if (cond) {
    *cond = 0.0;
    if (istebz > 0) {
      double *Z = (double *)malloc(sizeof(double) * (size_t)istebz * istebz);
      if (or_tridiag_eig(istebz, td, tu, Z) == 0) {
        double mn = td[0], mx = td[0];
        for (int i = 1; i < istebz; ++i) {
          mn = fmin(mn, td[i]);
          mx = fmax(mx, td[i]);
        }
        *cond = mx / mn;
      }
      free(Z);
    }
    free(td);
    free(tu);
  }
  /* psb_end_conv */
  if (err) *err = (errden == 0.0) ? errnum : errnum / errden;
  if (iter) *iter = itx;
  free(p); free(q); free(r); free(z); free(dinv);
  if (have_ilu) ilu0_free(&ilu);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Multi-shift CG for (A + sigma_j I) x_j = b, complex sigma_j         */
/* (no reference counterpart: psfun_dcg.F90 is a single real shift).   */
/* ------------------------------------------------------------------ */
/* Seed: real CG on the SPD matrix A with x0 = 0.  Shifted iterates follow the
 * collinear-residual recurrences r_k^s = zeta_k^s r_k (Jegerlehner 1996;
 * Frommer 2003):
 *   zeta_{k+1} = zeta_k zeta_{k-1} a_{k-1} /
 *                (a_k b_{k-1} (zeta_{k-1} - zeta_k) + zeta_{k-1} a_{k-1} (1 + s a_k))
 *   a_k^s = a_k zeta_{k+1}/zeta_k,  b_k^s = b_k (zeta_{k+1}/zeta_k)^2
 *   x^s += a_k^s p^s,  p^s = zeta_{k+1} r_{k+1} + b_k^s p^s.
 * A shift stops updating once |zeta_k^s| ||r_k|| <= eps ||b||.  X is
 * nshift x n complex (interleaved re/im), row j = shift j.  iters[j] is the
 * iteration at which shift j converged (or itmax). */
int or_mscg(int n, const int *rp, const int *ci, const double *va, int nshift, const double *shifts,
            const double *b, double *X, double eps, int itmax, int *iters, double *resid) {
  zdouble *sg = (zdouble *)malloc(sizeof(zdouble) * nshift);
  for (int s = 0; s < nshift; ++s) sg[s] = shifts[2 * s] + I * shifts[2 * s + 1];
  zdouble *Xs = (zdouble *)X;
  zdouble *Ps = (zdouble *)malloc(sizeof(zdouble) * (size_t)nshift * n);
  double *r = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n),
         *q = (double *)malloc(sizeof(double) * n);
  zdouble *zc = (zdouble *)malloc(sizeof(zdouble) * nshift), *zp = (zdouble *)malloc(sizeof(zdouble) * nshift);
  int *act = (int *)malloc(sizeof(int) * nshift);
  const double bn = or_nrm2(n, b);
  for (int s = 0; s < nshift; ++s) {
    zc[s] = 1.0;
    zp[s] = 1.0;
    act[s] = 1;
    iters[s] = 0;
    if (resid) resid[s] = (bn == 0.0) ? 0.0 : 1.0;   /* |zeta_0| ||r_0|| / ||b|| */
    for (int i = 0; i < n; ++i) {
      Xs[(size_t)s * n + i] = 0.0;
      Ps[(size_t)s * n + i] = b[i];
    }
  }
  memcpy(r, b, sizeof(double) * n);
  memcpy(p, b, sizeof(double) * n);
  double rr = or_dot(n, r, r);
  double a_prev = 1.0, b_prev = 0.0;
  int nact = nshift;
  if (bn == 0.0) nact = 0;
  for (int k = 0; k < itmax && nact > 0; ++k) {
    or_spmv(n, rp, ci, va, 1.0, p, 0.0, q);
    const double pq = or_dot(n, p, q);
    const double a = rr / pq;
    for (int i = 0; i < n; ++i) r[i] -= a * q[i];
    const double rr_new = or_dot(n, r, r);
    const double bk = rr_new / rr;
    const double rnorm = sqrt(rr_new);
    for (int s = 0; s < nshift; ++s) {
      if (!act[s]) continue;
      const zdouble zn = zc[s] * zp[s] * a_prev /
                         (a * b_prev * (zp[s] - zc[s]) + zp[s] * a_prev * (1.0 + sg[s] * a));
      const zdouble as = a * zn / zc[s];
      const zdouble bs = bk * (zn / zc[s]) * (zn / zc[s]);
      zdouble *xs = Xs + (size_t)s * n, *ps = Ps + (size_t)s * n;
      for (int i = 0; i < n; ++i) {
        xs[i] += as * ps[i];
        ps[i] = zn * r[i] + bs * ps[i];
      }
      zp[s] = zc[s];
      zc[s] = zn;
      iters[s] = k + 1;
      if (resid) resid[s] = cabs(zn) * rnorm / bn;
      if (cabs(zn) * rnorm <= eps * bn) {
        act[s] = 0;
        nact--;
      }
    }
    for (int i = 0; i < n; ++i) p[i] = r[i] + bk * p[i];
    a_prev = a;
    b_prev = bk;
    rr = rr_new;
    if (rr == 0.0) break;
  }
  free(sg); free(Ps); free(r); free(p); free(q); free(zc); free(zp); free(act);
  return 0;
}

/* Preconditioned COCG (van der Vorst and Melissen 1990: CG with the
 * bilinear form x^T y in place of x^H y, valid for complex symmetric
 * systems) for (A + sigma I) x = b, b real, x complex: the initial guess on
 * entry, the solution on exit.  M is complex symmetric:
 *   prec 0: M = I
 *   prec 1: M = diag(A) + sigma I, the Jacobi matrix of the shifted system
 *           (an entry equal to 0 is replaced by 1)
 *   prec 2: M = L U, the ILU(0) of the real A (L D L^T for a symmetric A),
 *           applied to the real and the imaginary part.
 *   r = b - (A + sigma I) x;  ||r|| <= eps ||b|| already: no iteration
 *   z = M^{-1} r; p = z; rho = r^T z
 *   repeat: q = (A + sigma I) p; mu = p^T q; alpha = rho / mu
 *           x += alpha p; r -= alpha q; it++
 *           stop when ||r||_2 <= eps ||b||_2 or it == itmax
 *           z = M^{-1} r; rho' = r^T z; p = z + (rho'/rho) p; rho = rho'
 * mu == 0 or rho == 0 ends the run (a breakdown; the iterate is kept).
 * b = 0 gives x = 0.  iter = iterations done, resid = ||r||_2 / ||b||_2 of
 * the last residual.  Returns 10 on a zero ILU(0) pivot. */
static int pcocg_run(int n, const int *rp, const int *ci, const double *va, int prec, const or_ilu_t *ilu,
                     zdouble sigma, const double *b, zdouble *x, double eps, int itmax, int *iter, double *resid) {
  const double bn = or_nrm2(n, b);
  *iter = 0;
  if (bn == 0.0) {
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    *resid = 0.0;
    return 0;
  }
  zdouble *r = (zdouble *)malloc(sizeof(zdouble) * n), *z = (zdouble *)malloc(sizeof(zdouble) * n),
          *p = (zdouble *)malloc(sizeof(zdouble) * n), *q = (zdouble *)malloc(sizeof(zdouble) * n),
          *dm = NULL;
  double *tr = (double *)malloc(sizeof(double) * n), *ti = (double *)malloc(sizeof(double) * n),
         *ur = (double *)malloc(sizeof(double) * n), *ui = (double *)malloc(sizeof(double) * n);
  if (prec == 1) {
    dm = (zdouble *)malloc(sizeof(zdouble) * n);
    for (int i = 0; i < n; ++i) {
      double dv = 0.0;
      for (int j = rp[i]; j < rp[i + 1]; ++j)
        if (ci[j] == i) dv += va[j];
      const zdouble m = dv + sigma;
      dm[i] = (m != 0.0) ? m : 1.0;
    }
  }
#define PC_AX(src, dst)                                                             \
  for (int i = 0; i < n; ++i) {                                                     \
    zdouble t = 0.0;                                                                \
    for (int j = rp[i]; j < rp[i + 1]; ++j) t += va[j] * src[ci[j]];                \
    dst[i] = t + sigma * src[i];                                                    \
  }
#define PC_PREC()                                                                   \
  if (prec == 1) {                                                                  \
    for (int i = 0; i < n; ++i) z[i] = r[i] / dm[i];                                \
  } else if (prec == 2) {                                                           \
    for (int i = 0; i < n; ++i) { tr[i] = creal(r[i]); ti[i] = cimag(r[i]); }       \
    ilu0_solve(ilu, tr, ur);                                                        \
    ilu0_solve(ilu, ti, ui);                                                        \
    for (int i = 0; i < n; ++i) z[i] = ur[i] + I * ui[i];                           \
  } else {                                                                          \
    for (int i = 0; i < n; ++i) z[i] = r[i];                                        \
  }
  PC_AX(x, q);
  double rn2 = 0.0;
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    rn2 += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
  }
  *resid = sqrt(rn2) / bn;
  if (sqrt(rn2) > eps * bn && itmax > 0) {
    PC_PREC();
    zdouble rho = 0.0;
    for (int i = 0; i < n; ++i) {
      rho += r[i] * z[i];
      p[i] = z[i];
    }
    int it = 0;
    while (rho != 0.0) {
      PC_AX(p, q);
      zdouble mu = 0.0;
      for (int i = 0; i < n; ++i) mu += p[i] * q[i];
      if (mu == 0.0) break;
      const zdouble alpha = rho / mu;
      rn2 = 0.0;
      for (int i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
        rn2 += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
      }
      it++;
      *resid = sqrt(rn2) / bn;
      if (sqrt(rn2) <= eps * bn || it >= itmax) break;
      PC_PREC();
      zdouble rho_new = 0.0;
      for (int i = 0; i < n; ++i) rho_new += r[i] * z[i];
      const zdouble beta = rho_new / rho;
      for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      rho = rho_new;
    }
    *iter = it;
  }
#undef PC_AX
#undef PC_PREC
  free(r); free(z); free(p); free(q); free(dm); free(tr); free(ti); free(ur); free(ui);
  return 0;
}

int or_pcocg(int n, const int *rp, const int *ci, const double *va, int prec, double sre, double sim,
             const double *b, double *x, double eps, int itmax, int *iter, double *resid) {
  or_ilu_t ilu;
  int have = 0;
  if (prec == 2) {
    if (ilu0_build(n, rp, ci, va, &ilu)) {
      ilu0_free(&ilu);
      return 10;
    }
    have = 1;
  }
  const int rc = pcocg_run(n, rp, ci, va, prec, have ? &ilu : NULL, sre + I * sim, b, (zdouble *)x, eps, itmax,
                           iter, resid);
  if (have) ilu0_free(&ilu);
  return rc;
}

/* or_mscg, then each shift it left above eps (resid > eps) continues from
 * its multi-shift iterate by preconditioned COCG (prec as in or_pcocg) for
 * at most itmax more iterations: iters = multi-shift + COCG iterations,
 * resid = COCG's final relative residual for those shifts. */
int or_mscg_prec(int n, const int *rp, const int *ci, const double *va, int prec, int nshift, const double *shifts,
                 const double *b, double *X, double eps, int itmax, int *iters, double *resid) {
  or_ilu_t ilu;
  int have = 0;
  if (prec == 2) {
    if (ilu0_build(n, rp, ci, va, &ilu)) {
      ilu0_free(&ilu);
      return 10;
    }
    have = 1;
  }
  or_mscg(n, rp, ci, va, nshift, shifts, b, X, eps, itmax, iters, resid);
  for (int s = 0; s < nshift; ++s) {
    if (!(resid[s] > eps)) continue;
    int it = 0;
    double rs = 0.0;
    pcocg_run(n, rp, ci, va, prec, have ? &ilu : NULL, shifts[2 * s] + I * shifts[2 * s + 1], b,
              (zdouble *)X + (size_t)s * n, eps, itmax, &it, &rs);
    iters[s] += it;
    resid[s] = rs;
  }
  if (have) ilu0_free(&ilu);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Special functions for contour rules                                 */
/* ------------------------------------------------------------------ */
/* Complete elliptic K(M), K(1-M) at M = exp(-2 pi L) by the AGM
 * (utils/psfun_d_utils_mod.f90:37-99, including the L > 10 asymptotic). */
void or_ellipkkp(double L, double *K) {
  const double pi = 4.0 * atan(1.0);
  if (L > 10.0) {
    K[0] = pi / 2.0;
    K[1] = pi * L + log(4.0);
    return;
  }
  const double M = exp(-2.0 * pi * L);
  for (int pass = 0; pass < 2; ++pass) {
    double a0 = 1.0, b0 = pass == 0 ? sqrt(1.0 - M) : sqrt(M), a1 = a0, mm = 1.0;
    int i1 = 0;
    while (mm > DBL_EPSILON) {
      a1 = (a0 + b0) / 2.0;
      const double b1 = sqrt(a0 * b0);
      const double c1 = (a0 - b0) / 2.0;
      i1++;
      mm = ldexp(1.0, i1) * c1 * c1;
      a0 = a1;
      b0 = b1;
    }
    K[pass] = pi / (2.0 * a1);
  }
}

/* Horner evaluation sum_i c_i x^(i-1)  (utils/psfun_d_utils_mod.f90:119-138) */
double or_horner(int nc, const double *c, double x) {
  double r = 0.0;
  for (int i = nc - 1; i >= 0; --i) r = r * x + c[i];
  return r;
}

/* Jacobi sn, cn, dn at complex u, parameter m = exp(-2 pi L) (flag = 1) or
 * m = L directly (flag = 0), by descending Landen transformations
 * (utils/psfun_z_utils_mod.f90:39-107).  As in the reference, the argument
 * reflected for Im u > K'/2 is not used afterwards (lines 76-77, 88). */
static void z_ellipj(zdouble u, double L, zdouble *sn, zdouble *cn, zdouble *dn, int flag) {
  const double pi = 4.0 * atan(1.0);
  static const double coeffs[7] = {0.0, 1.0, 2.0, 5.0, 14.0, 42.0, 132.0};
  double m, K[2];
  int high;
  zdouble ucp = u;
  if (flag) {
    or_ellipkkp(L, K);
    high = cimag(ucp) > K[1] / 2.0;
    if (high) ucp = I * K[1] - ucp;
    m = exp(-2.0 * pi * L);
  } else {
    high = 0;
    m = L;
  }
  (void)ucp;
  if (m < 4.0 * DBL_EPSILON) {
    const zdouble s = csin(u), c = ccos(u);
    *sn = s + m / 4.0 * (s * c - u) * c;
    *cn = c + m / 4.0 * (-s * c + u) * s;
    *dn = 1.0 + m / 4.0 * (c * c - s * s - 1.0);
  } else {
    double kappa;
    if (m > 1e-3)
      kappa = (1.0 - sqrt(1.0 - m)) / (1.0 + sqrt(1.0 - m));
    else
      kappa = or_horner(7, coeffs, m / 4.0);
    const double mu = kappa * kappa;
    const zdouble v = u / (1.0 + kappa);
    zdouble sn1, cn1, dn1;
    z_ellipj(v, mu, &sn1, &cn1, &dn1, 0);
    const zdouble den = 1.0 + kappa * sn1 * sn1;
    *sn = (1.0 + kappa) * sn1 / den;
    *cn = cn1 * dn1 / den;
    *dn = (1.0 - kappa * sn1 * sn1) / den;
  }
  if (high) {
    const zdouble snh = *sn, cnh = *cn, dnh = *dn;
    *sn = -1.0 / (sqrt(m) * snh);
    *cn = I * dnh / (sqrt(m) * snh);
    *dn = I * cnh / snh;
  }
}

void or_ellipj(double ure, double uim, double L, double *out /* sn, cn, dn as re,im pairs */) {
  zdouble sn, cn, dn;
  z_ellipj(ure + I * uim, L, &sn, &cn, &dn, 1);
  out[0] = creal(sn); out[1] = cimag(sn);
  out[2] = creal(cn); out[3] = cimag(cn);
  out[4] = creal(dn); out[5] = cimag(dn);
}

typedef void (*zfun_t)(double re, double im, double *out_re, double *out_im);

/* HHT 2008 Method 1 (quadratures/psfun_z_quadrules_mod.f90:63-83).
 * xi, c: N complex (interleaved). */
int or_hhtmethod1(zfun_t zfun, int N, double r1, double r2, double *xi, double *c, double *eta, double *sign) {
  const double pi = 4.0 * atan(1.0);
  const double mineig = fmin(r1, r2), maxeig = fmax(r1, r2);
  const double cnd = (sqrt(maxeig / mineig) - 1.0) / (sqrt(maxeig / mineig) + 1.0);
  const double L = -log(cnd) / pi;
  double K[2];
  or_ellipkkp(L, K);
  for (int j = 1; j <= N; ++j) {
    const zdouble t = 0.5 * I * K[1] - K[0] + (2.0 * (j - 1) + 1.0) * K[0] / N;
    zdouble u, cn, dn;
    z_ellipj(t, L, &u, &cn, &dn, 1);
    const zdouble x = sqrt(mineig * maxeig) * ((1.0 / cnd + u) / (1.0 / cnd - u));
    double fr, fi;
    zfun(creal(x), cimag(x), &fr, &fi);
    const zdouble cc = ((fr + I * fi) / x) * (cn * dn) / ((1.0 / cnd - u) * (1.0 / cnd - u));
    xi[2 * (j - 1)] = creal(x);
    xi[2 * (j - 1) + 1] = cimag(x);
    c[2 * (j - 1)] = creal(cc);
    c[2 * (j - 1) + 1] = cimag(cc);
  }
  *eta = -(4.0 * K[0] * sqrt(mineig * maxeig)) / (cnd * pi * N);
  *sign = -1.0;
  return 0;
}

/* Modified Talbot contour for exp(-t x), x >= 0 (Trefethen, Weideman &
 * Schmelzer, BIT 46 (2006) 653-670):  z(th) = N (0.5017 th cot(0.6407 th)
 * - 0.6122 + 0.2645 i th), midpoints th_j = -pi + (j - 1/2) 2 pi / N,
 * exp(-t x) ~ Re sum_j w_j (z_j/t + x)^{-1}, w_j = e^{z_j} z'(th_j)/(i N t).
 * Returned as shifts sigma_j = z_j/t (solve (A + sigma_j I) v = x) and
 * weights w_j. */
int or_talbot(int N, double t, double *sigma, double *w) {
  const double pi = 4.0 * atan(1.0);
  for (int j = 1; j <= N; ++j) {
    const double th = -pi + (j - 0.5) * 2.0 * pi / N;
    const double a = 0.6407 * th;
    const double cot = cos(a) / sin(a);
    const zdouble z = N * (0.5017 * th * cot - 0.6122 + 0.2645 * I * th);
    const double sn = sin(a);
    const zdouble dz = N * (0.5017 * cot - 0.5017 * 0.6407 * th / (sn * sn) + 0.2645 * I);
    const zdouble ww = cexp(z) * dz / (I * N * t);
    sigma[2 * (j - 1)] = creal(z) / t;
    sigma[2 * (j - 1) + 1] = cimag(z) / t;
    w[2 * (j - 1)] = creal(ww);
    w[2 * (j - 1) + 1] = cimag(ww);
  }
  return 0;
}

/* Quadrature apply.  v_j = (A + sigma_j I)^{-1} x by or_mscg (prec < 0) or
 * by or_mscg_prec (prec 0, 1, 2: the preconditioned per-shift fallback), then
 *   form 0 (resolvent sum): y = gamma * Re( sum_j w_j v_j )
 *   form 1 (HHT method 1):  y = gamma * A * Im( sum_j w_j v_j ). */
int or_quad_apply_prec(int n, const int *rp, const int *ci, const double *va, int prec, int nshift,
                       const double *sigma, const double *w, double gamma, int form, const double *x, double *y,
                       double eps, int itmax, int *iters, double *resid) {
  double *X = (double *)malloc(sizeof(double) * 2 * (size_t)nshift * n);
  double *rs = resid ? resid : (double *)malloc(sizeof(double) * (nshift > 0 ? nshift : 1));
  int rc = 0;
  if (prec < 0)
    or_mscg(n, rp, ci, va, nshift, sigma, x, X, eps, itmax, iters, rs);
  else
    rc = or_mscg_prec(n, rp, ci, va, prec, nshift, sigma, x, X, eps, itmax, iters, rs);
  if (!resid) free(rs);
  if (rc) {
    free(X);
    return rc;
  }
  double *acc = (double *)calloc((size_t)n, sizeof(double));
  for (int s = 0; s < nshift; ++s) {
    const zdouble ws = w[2 * s] + I * w[2 * s + 1];
    const zdouble *xs = (const zdouble *)X + (size_t)s * n;
    for (int i = 0; i < n; ++i) {
      const zdouble t = ws * xs[i];
      acc[i] += (form == 0) ? creal(t) : cimag(t);
    }
  }
  if (form == 0) {
    for (int i = 0; i < n; ++i) y[i] = gamma * acc[i];
  } else {
    or_spmv(n, rp, ci, va, gamma, acc, 0.0, y);
  }
  free(X);
  free(acc);
  return 0;
}

int or_quad_apply(int n, const int *rp, const int *ci, const double *va, int nshift, const double *sigma,
                  const double *w, double gamma, int form, const double *x, double *y, double eps, int itmax,
                  int *iters) {
  return or_quad_apply_prec(n, rp, ci, va, -1, nshift, sigma, w, gamma, form, x, y, eps, itmax, iters, NULL);
}
