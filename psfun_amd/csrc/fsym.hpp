// Symmetric eigen-decomposition S = Q diag(w) Q^T of a small dense k x k
// matrix by ONE block, for the USERF/SYM f(H) step
// (serial/psfun_d_serial_mod.F90:420-485: DSYEV('V','U') on H, then
// y = Q diag(f(scaling*w)) Q^T x with a user function f).
//
// Method: cyclic Jacobi with the round-robin parallel ordering (Golub & Van
// Loan, Matrix Computations, 4th ed., sec. 8.5: the symmetric 2x2 Schur
// rotation and the parallel ordering).  A sweep has m-1 rounds, m = k rounded
// up to even; in a round the m/2 disjoint index pairs are rotated together,
// S <- J^T S J and Q <- Q J, the columns first, then the rows, with the 2x2
// pivot blocks set to their exact rotated values.  A pair is skipped when its
// off-diagonal entry is below 2^-53 sqrt(|s_pp s_qq|) (or subnormal); sweeps
// repeat until one rotates nothing (at most PSFUN_SYM_SWEEPS).  Jacobi's
// backward error is a few ulps of ||S||, as for the QR-based LAPACK driver.
//
// S and Q are k x k column-major scratch in global memory (they stay in the
// CU's caches at the sizes the Krylov loops meet); the pair table lives in LDS.
#pragma once
#include "common.hpp"
#include "fdense.hpp"

#define PSFUN_SYM_SWEEPS 60

struct SymLds {
  int p[PSFUN_DENSE_KMAX / 2];
  int q[PSFUN_DENSE_KMAX / 2];
  double c[PSFUN_DENSE_KMAX / 2];
  double s[PSFUN_DENSE_KMAX / 2];
  double dp[PSFUN_DENSE_KMAX / 2];
  double dq[PSFUN_DENSE_KMAX / 2];
  double z[PSFUN_DENSE_KMAX];
  double red[PSFUN_BS / 64 + 1];
  int any;
};

// Input forms of the symmetric matrix.
enum { SYM_FROM_TRIDIAG = 0, SYM_FROM_UPPER = 1 };

// S = the symmetric matrix given by (a, b): SYM_FROM_TRIDIAG: diagonal a[0..k-1],
// off-diagonal b[0..k-2]; SYM_FROM_UPPER: S(i,j) = a[min(i,j) + max(i,j)*ld].
__device__ inline void block_sym_load(int k, int form, const double *a, const double *b, int ld, double *S) {
  for (int e = threadIdx.x; e < k * k; e += PSFUN_BS) {
    const int r = e % k, cc = e / k;
    double v;
    if (form == SYM_FROM_TRIDIAG) {
      v = (r == cc) ? a[r] : (r == cc + 1) ? b[cc] : (cc == r + 1) ? b[r] : 0.0;
    } else {
      const int i = r < cc ? r : cc, j = r < cc ? cc : r;
      v = a[(int64_t)j * ld + i];
    }
    S[e] = v;
  }
  __syncthreads();
}

// Jacobi on S (overwritten: diagonal on exit); Q receives the eigenvectors as
// columns and w the eigenvalues (w[i] pairs with column i).  Returns the
// number of sweeps run.
__device__ inline int block_sym_eig(int k, double *S, double *Q, double *w, SymLds *L) {
  const int tid = threadIdx.x;
  for (int e = tid; e < k * k; e += PSFUN_BS) Q[e] = (e % k == e / k) ? 1.0 : 0.0;
  __syncthreads();
  const int m = k + (k & 1);
  const int np = m / 2;
  int sweep = 0;
  for (; sweep < PSFUN_SYM_SWEEPS; ++sweep) {
    if (tid == 0) L->any = 0;
    __syncthreads();
    for (int rd = 0; rd < m - 1; ++rd) {
      // pair i holds positions (i, m-1-i) of the round-robin table, where
      // position 0 keeps index 0 and position t >= 1 holds 1 + (t-1+rd) mod (m-1)
      for (int i = tid; i < np; i += PSFUN_BS) {
        const int a = (i == 0) ? 0 : 1 + (i - 1 + rd) % (m - 1);
        const int b = 1 + (m - 2 - i + rd) % (m - 1);
        const int p = a < b ? a : b, q = a < b ? b : a;
        int pp = -1;
        if (q < k) {
          const double app = S[(int64_t)p * k + p], aqq = S[(int64_t)q * k + q];
          const double apq = 0.5 * (S[(int64_t)q * k + p] + S[(int64_t)p * k + q]);
          const double aa = fabs(apq);
          if (aa >= DBL_MIN && aa > 0x1p-53 * sqrt(fabs(app)) * sqrt(fabs(aqq))) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (fabs(theta) > 1e150) ? 0.5 / theta
                                                   : copysign(1.0, theta) / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
            const double c = 1.0 / sqrt(fma(t, t, 1.0));
            pp = p;
            L->c[i] = c;
            L->s[i] = t * c;
            L->dp[i] = fma(-t, apq, app);
            L->dq[i] = fma(t, apq, aqq);
            L->any = 1;
          }
        }
        L->p[i] = pp;
        L->q[i] = q;
      }
      __syncthreads();
      // columns: S <- S J, Q <- Q J
      for (int e = tid; e < np * k; e += PSFUN_BS) {
        const int i = e / k, r = e % k;
        const int p = L->p[i];
        if (p < 0) continue;
        const int q = L->q[i];
        const double c = L->c[i], s = L->s[i];
        double *sp = S + (int64_t)p * k, *sq = S + (int64_t)q * k;
        const double x = sp[r], y = sq[r];
        sp[r] = c * x - s * y;
        sq[r] = fma(s, x, c * y);
        double *qp = Q + (int64_t)p * k, *qq = Q + (int64_t)q * k;
        const double u = qp[r], v = qq[r];
        qp[r] = c * u - s * v;
        qq[r] = fma(s, u, c * v);
      }
      __syncthreads();
      // rows: S <- J^T S, with the pivot block set exactly
      for (int e = tid; e < np * k; e += PSFUN_BS) {
        const int i = e / k, r = e % k;
        const int p = L->p[i];
        if (p < 0) continue;
        const int q = L->q[i];
        double *cr = S + (int64_t)r * k;
        if (r == p) {
          cr[p] = L->dp[i];
          cr[q] = 0.0;
        } else if (r == q) {
          cr[p] = 0.0;
          cr[q] = L->dq[i];
        } else {
          const double c = L->c[i], s = L->s[i];
          const double x = cr[p], y = cr[q];
          cr[p] = c * x - s * y;
          cr[q] = fma(s, x, c * y);
        }
      }
      __syncthreads();
    }
    const int any = L->any;
    __syncthreads();
    if (!any) break;
  }
  for (int i = tid; i < k; i += PSFUN_BS) w[i] = S[(int64_t)i * k + i];
  __syncthreads();
  return sweep;
}

// y = Q diag(fv) Q^T x (x == nullptr: x = e1), all threads; returns ||y||_2
// (every thread).  y may be global memory; z in LDS.
__device__ inline double block_sym_apply(int k, const double *Q, const double *fv, const double *x, double *y,
                                         SymLds *L) {
  const int tid = threadIdx.x;
  for (int i = tid; i < k; i += PSFUN_BS) {
    const double *qi = Q + (int64_t)i * k;
    double z;
    if (x) {
      z = 0.0;
      for (int r = 0; r < k; ++r) z = fma(qi[r], x[r], z);
    } else {
      z = qi[0];
    }
    L->z[i] = z * fv[i];
  }
  __syncthreads();
  double ss = 0.0;
  for (int r = tid; r < k; r += PSFUN_BS) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc = fma(Q[(int64_t)i * k + r], L->z[i], acc);
    y[r] = acc;
    ss = fma(acc, acc, ss);
  }
  return sqrt(block_sum<PSFUN_BS>(ss, L->red));
}

// Device side of one USERF step, shared by the serial, Lanczos and Arnoldi
// paths: fills S and takes its eigen-decomposition (k_sym_eig), then, after
// the host has evaluated fv = f(fscale*w), applies Q diag(fv) Q^T.
struct SymScratch {
  double *S, *Q, *w, *fv;
};
