// The small f(T) step on the device: y = exp(s*T) v for a symmetric
// tridiagonal T of order k <= 64*R, evaluated by ONE wavefront.
//
// Method: Chebyshev expansion of exp on the Gershgorin interval [lo, hi] of
// s*T (Clenshaw-Curtis/Chebyshev series with modified-Bessel coefficients,
// exp(c + r x) = e^{c+r} [b_0 + 2 sum_m b_m T_m(x)], b_m = e^{-r} I_m(r);
// the b_m come from Miller's backward recurrence normalised by
// e^{-r}(I_0 + 2 sum I_m) = 1).  The three-term recurrence t_{m+1} = 2 S t_m -
// t_{m-1}, S = (sT - cI)/r, keeps |t_m| <= 1, and because T is tridiagonal the
// j-th component of T_m(S) e1 is exactly zero for m < j-1.  The trailing
// components that drive the Lanczos stopping estimate are therefore obtained
// with an error proportional to their own size instead of to ||y||.
//
// Lane l owns rows l*R .. l*R+R-1; neighbour rows are exchanged with one
// shuffle up and one shuffle down per term.  The coefficient table lives in
// LDS (`coef`, capacity `cap` doubles) and is written by lane 0 only; the
// wave-scope fence orders it for the other lanes of the same wavefront.
#pragma once
#include "common.hpp"

#define PSFUN_CHEB_CAP 2048     // coefficient table capacity (doubles)
#define PSFUN_CHEB_RMAX 600.0   // largest half-width per sub-step

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bessel table b_m = e^{-r} I_m(r), m = 0..M, with the factor 2 folded into
// m >= 1.  Returns M (number of terms - 1).  Lane 0 only.
__device__ inline int cheb_exp_coefficients(double r, double *coef, int cap) {
  int ms = (int)ceil(r + 12.0 * sqrt(r) + 40.0);
  if (ms > cap - 2) ms = cap - 2;
  const double twoinv = 2.0 / r;
  double ip1 = 0.0, i0 = 1e-280;
  coef[ms] = i0;
  for (int m = ms; m >= 1; --m) {
    const double im1 = fma((double)m * twoinv, i0, ip1);
    coef[m - 1] = im1;
    ip1 = i0;
    i0 = im1;
    if (im1 > 1e250) {
      for (int q = m - 1; q <= ms; ++q) coef[q] *= 1e-250;
      ip1 *= 1e-250;
      i0 *= 1e-250;
    }
  }
  double s = 0.0;
  for (int m = ms; m >= 1; --m) s += coef[m];
  s = coef[0] + 2.0 * s;
  const double inv = 1.0 / s;
  const double b0 = coef[0] * inv;
  coef[0] = b0;
  int M = 0;
  for (int m = 1; m <= ms; ++m) {
    const double b = coef[m] * inv;
    coef[m] = 2.0 * b;
    if (b > 1e-26 * b0) M = m;
  }
  return M;
}

// y <- exp(s*T) y in place (y distributed as described above).  d, e point to
// the diagonal (k) and off-diagonal (k-1) of T in global or LDS memory.
template <int R>
__device__ void wave_exp_tridiag(int k, const double *d, const double *e, double s, double (&y)[R],
                                 double *coef) {
  const int lane = threadIdx.x & 63;
  double a[R], lo_c[R], hi_c[R];
  double glo = 1e308, ghi = -1e308;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane * R + q;
    double di = 0.0, el = 0.0, eh = 0.0;
    if (i < k) {
      di = s * d[i];
      if (i > 0) el = s * e[i - 1];
      if (i < k - 1) eh = s * e[i];
      glo = fmin(glo, di - fabs(el) - fabs(eh));
      ghi = fmax(ghi, di + fabs(el) + fabs(eh));
    }
    a[q] = di;
    lo_c[q] = el;
    hi_c[q] = eh;
  }
  glo = wave_min(glo);
  ghi = wave_max(ghi);
  const double width = ghi - glo;
  if (!(width > 1e-300)) {  // s*T is a multiple of the identity
    const double f = exp(ghi);
#pragma unroll
    for (int q = 0; q < R; ++q) y[q] *= f;
    return;
  }
  int nsub = (int)ceil(0.5 * width / PSFUN_CHEB_RMAX);
  if (nsub < 1) nsub = 1;
  const double ssub = 1.0 / nsub;
  const double c = 0.5 * (glo + ghi) * ssub, r = 0.5 * width * ssub, hs = ghi * ssub;
  __shared__ int M_sh;
  if (lane == 0) M_sh = cheb_exp_coefficients(r, coef, PSFUN_CHEB_CAP);
  wave_lds_sync();
  const int M = M_sh;
  const double rinv = 1.0 / r;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    a[q] = (a[q] * ssub - c) * rinv;
    lo_c[q] *= ssub * rinv;
    hi_c[q] *= ssub * rinv;
  }
  const double escale = exp(hs);
  for (int sub = 0; sub < nsub; ++sub) {
    double t0[R], t1[R], acc[R];
    const double b0 = coef[0];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      t0[q] = y[q];
      acc[q] = b0 * y[q];
    }
    // t1 = S t0
    {
      const double up = __shfl_up(t0[R - 1], 1, 64);
      const double dn = __shfl_down(t0[0], 1, 64);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const double tm = (q == 0) ? up : t0[q - 1];
        const double tp = (q == R - 1) ? dn : t0[q + 1];
        t1[q] = fma(a[q], t0[q], fma(lo_c[q], tm, hi_c[q] * tp));
      }
    }
    if (M >= 1) {
      const double b1 = coef[1];
#pragma unroll
      for (int q = 0; q < R; ++q) acc[q] = fma(b1, t1[q], acc[q]);
    }
    for (int m = 2; m <= M; ++m) {
      const double bm = coef[m];
      const double up = __shfl_up(t1[R - 1], 1, 64);
      const double dn = __shfl_down(t1[0], 1, 64);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const double tm = (q == 0) ? up : t1[q - 1];
        const double tp = (q == R - 1) ? dn : t1[q + 1];
        const double st = fma(a[q], t1[q], fma(lo_c[q], tm, hi_c[q] * tp));
        const double t2 = fma(2.0, st, -t0[q]);
        t0[q] = t1[q];
        t1[q] = t2;
        acc[q] = fma(bm, t2, acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) y[q] = acc[q] * escale;
  }
}

// exp(s*T) e1 for k <= 64*R; returns ||y||_2 and y[k-1] in every lane and, if
// yout != nullptr, stores y[0..k-1].
template <int R>
__device__ void wave_exp_tridiag_e1(int k, const double *d, const double *e, double s, double *coef,
                                    double *yout, double &nrm, double &last) {
  const int lane = threadIdx.x & 63;
  double y[R];
#pragma unroll
  for (int q = 0; q < R; ++q) y[q] = (lane == 0 && q == 0) ? 1.0 : 0.0;
  wave_exp_tridiag<R>(k, d, e, s, y, coef);
  double ss = 0.0, lv = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane * R + q;
    if (i < k) {
      ss = fma(y[q], y[q], ss);
      if (i == k - 1) lv = y[q];
      if (yout) yout[i] = y[q];
    }
  }
  nrm = sqrt(wave_sum(ss));
  last = wave_sum(lv);
}

// Dispatch on k (k <= 512).
__device__ inline void wave_exp_tridiag_e1_any(int k, const double *d, const double *e, double s, double *coef,
                                               double *yout, double &nrm, double &last) {
  if (k <= 64) wave_exp_tridiag_e1<1>(k, d, e, s, coef, yout, nrm, last);
  else if (k <= 128) wave_exp_tridiag_e1<2>(k, d, e, s, coef, yout, nrm, last);
  else if (k <= 256) wave_exp_tridiag_e1<4>(k, d, e, s, coef, yout, nrm, last);
  else wave_exp_tridiag_e1<8>(k, d, e, s, coef, yout, nrm, last);
}
