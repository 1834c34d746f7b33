// The small f(T) step on the device: y = exp(s*T) v for a symmetric
// tridiagonal T of order k <= PSFUN_FT_KMAX, evaluated by ONE wavefront.
//
// Method: Chebyshev expansion of exp on an interval [lo, hi] holding the
// spectrum of s*T, exp(c + r x) = e^{c+r} [b_0 + 2 sum_m b_m T_m(x)] with
// b_m = e^{-r} I_m(r) (modified Bessel functions).  The b_m come from Miller's
// backward recurrence I_{m-1} = (2m/r) I_m + I_{m+1}, normalised by the
// identity e^{-r}(I_0 + 2 sum_m I_m) = 1.  The three-term recurrence
// t_{m+1} = 2 S t_m - t_{m-1}, S = (sT - cI)/r, keeps |t_m| <= 1, and because
// T is tridiagonal the j-th component of T_m(S) e1 is exactly zero for
// m < j-1, so the trailing components that drive the Lanczos stopping
// estimate come out with an error proportional to their own size.  Widths
// above 2*PSFUN_CHEB_RMAX are split into equal sub-steps
// exp(sT) = (exp(sT/p))^p.
//
// lo is the Gershgorin lower bound; hi is the largest eigenvalue of s*T to
// within PSFUN_CHEB_HI_TOL (multisection on Sturm counts), not the Gershgorin
// upper bound.  The bracketed sum is e^{-hi} exp(sT) e1, so an upper end d
// above the largest eigenvalue makes it e^{-d} times smaller than its terms
// and costs e^{d} in relative rounding error: the Gershgorin bound of a
// Lanczos T with large off-diagonals can exceed the spectrum by tens (30 on a
// 1024^2 anisotropic diffusion operator scaled by t = 0.1, an error of 1e-3).
//
// All vectors live in LDS (one wavefront, rows i = lane, lane+64, ...), so the
// routine costs few registers and can sit in a control block of a streaming
// kernel without lowering that kernel's occupancy.  The wave-scope fences
// order LDS traffic between the lanes of the wavefront.
#pragma once
#include "common.hpp"

#define PSFUN_CHEB_CAP 704      // coefficient table capacity (doubles; r + 12 sqrt(r) + 42 at r = RMAX)
#define PSFUN_CHEB_RMAX 400.0   // largest half-width per sub-step
#define PSFUN_FT_KMAX 512       // largest order of T
#define PSFUN_CHEB_HI_TOL 0.25  // bracket width of the largest eigenvalue (e^0.25 of roundoff at most)

// The coefficient table alone: what the register path (k <= 64) needs, so a
// kernel that only takes that path holds 5.6 KB of LDS instead of 18 KB.
struct ChebCoefLds {
  double coef[PSFUN_CHEB_CAP];
  int M;
};

struct ChebLds {
  double coef[PSFUN_CHEB_CAP];
  double y[PSFUN_FT_KMAX];
  double t0[PSFUN_FT_KMAX];
  double t1[PSFUN_FT_KMAX];
  int M;
};


// Bessel table b_m = e^{-r} I_m(r), m = 0..M, with the factor 2 folded into
// m >= 1.  Returns M (number of terms - 1).  Lane 0 only.
__device__ inline int cheb_exp_coefficients(double r, double *coef, int cap) {
  int ms = (int)ceil(r + 12.0 * sqrt(r) + 40.0);
  if (ms > cap - 2) ms = cap - 2;
  const double twoinv = 2.0 / r;
  double ip1 = 0.0, i0 = 1e-280;
  coef[ms] = i0;
#pragma unroll 1
  for (int m = ms; m >= 1; --m) {
    const double im1 = fma((double)m * twoinv, i0, ip1);
    coef[m - 1] = im1;
    ip1 = i0;
    i0 = im1;
    if (im1 > 1e250) {
#pragma unroll 1
      for (int q = m - 1; q <= ms; ++q) coef[q] *= 1e-250;
      ip1 *= 1e-250;
      i0 *= 1e-250;
    }
  }
  double s = 0.0;
#pragma unroll 1
  for (int m = ms; m >= 1; --m) s += coef[m];
  s = coef[0] + 2.0 * s;
  const double inv = 1.0 / s;
  const double b0 = coef[0] * inv;
  coef[0] = b0;
  int M = 0;
#pragma unroll 1
  for (int m = 1; m <= ms; ++m) {
    const double b = coef[m] * inv;
    coef[m] = 2.0 * b;
    if (b > 1e-26 * b0) M = m;
  }
  return M;
}

// Number of eigenvalues of a symmetric tridiagonal below x (Sturm count of
// the LDL^T pivots of the matrix minus x I, a zero pivot replaced by -1e-300
// as in LAPACK's dstebz).  diag(i) and off2(i) return the i-th diagonal entry
// and the square of the off-diagonal entry (i-1, i) for i >= 1.
template <typename Diag, typename Off2>
__device__ inline int tridiag_count_below(int k, Diag diag, Off2 off2, double x) {
  int cnt = 0;
  double q = 1.0;
#pragma unroll 1
  for (int i = 0; i < k; ++i) {
    double v = diag(i) - x;
    if (i > 0) v -= off2(i) / q;
    if (fabs(v) < 1e-300) v = -1e-300;
    q = v;
    cnt += (q < 0.0);
  }
  return cnt;
}

// The largest eigenvalue from above: the bracket [glo, ghi] (Gershgorin) is
// cut at 64 interior points per round, one Sturm count per lane, and narrowed
// to the cell where the count reaches k, until it is no wider than
// PSFUN_CHEB_HI_TOL (at most 6 rounds).  Every lane of the wavefront computes
// the same points and returns the same upper end, which is >= the largest
// eigenvalue up to the count's rounding.
template <typename Diag, typename Off2>
__device__ __attribute__((noinline)) double wave_tridiag_upper(int k, Diag diag, Off2 off2, double glo, double ghi) {
  const int lane = threadIdx.x & 63;
  double lo = glo, hi = ghi;
#pragma unroll 1
  for (int round = 0; round < 6 && hi - lo > PSFUN_CHEB_HI_TOL; ++round) {
    const double step = (hi - lo) / 65.0;
    const double x = fma((double)(lane + 1), step, lo);
    const unsigned long long all = __ballot(tridiag_count_below(k, diag, off2, x) == k);
    if (all == 0ull) {
      lo = fma(64.0, step, lo);
    } else {
      const int L = __ffsll((long long)all) - 1;
      const double nlo = (L > 0) ? fma((double)L, step, lo) : lo;
      hi = fma((double)(L + 1), step, lo);
      lo = nlo;
    }
  }
  return hi;
}

// y <- exp(s*T) y, y = W->y[0..k-1].  d (k) and e (k-1) are the diagonal and
// off-diagonal of T in global memory.  Called by all 64 lanes of one wave.
__device__ void wave_exp_tridiag(int k, const double *d, const double *e, double s, ChebLds *W) {
  const int lane = threadIdx.x & 63;
  double glo = 1e308, ghi = -1e308;
  for (int i = lane; i < k; i += 64) {
    const double di = s * d[i];
    const double el = (i > 0) ? fabs(s * e[i - 1]) : 0.0;
    const double eh = (i < k - 1) ? fabs(s * e[i]) : 0.0;
    glo = fmin(glo, di - el - eh);
    ghi = fmax(ghi, di + el + eh);
  }
  glo = wave_min(glo);
  ghi = wave_max(ghi);
  if (!(ghi - glo > 1e-300)) {  // s*T is a multiple of the identity
    const double f = exp(ghi);
    for (int i = lane; i < k; i += 64) W->y[i] *= f;
    wave_lds_sync();
    return;
  }
  {  // s*d and (s*e)^2 staged in t0/t1 (free until the recurrence) for the Sturm counts
    double *sdv = W->t0, *se2 = W->t1;
    for (int i = lane; i < k; i += 64) {
      sdv[i] = s * d[i];
      const double ei = (i > 0) ? s * e[i - 1] : 0.0;
      se2[i] = ei * ei;
    }
    wave_lds_sync();
    ghi = wave_tridiag_upper(
        k, [&](int i) { return sdv[i]; }, [&](int i) { return se2[i]; }, glo, ghi);
    wave_lds_sync();
  }
  const double width = ghi - glo;
  int nsub = (int)ceil(0.5 * width / PSFUN_CHEB_RMAX);
  if (nsub < 1) nsub = 1;
  const double ssub = 1.0 / nsub;
  const double c = 0.5 * (glo + ghi) * ssub, r = 0.5 * width * ssub, hs = ghi * ssub;
  if (lane == 0) W->M = cheb_exp_coefficients(r, W->coef, PSFUN_CHEB_CAP);
  wave_lds_sync();
  const int M = W->M;
  const double rinv = 1.0 / r;
  const double sd = s * ssub * rinv, cd = c * rinv;  // S_ii = sd*d_i - cd, S_i,i+1 = sd*e_i
  const double escale = exp(hs);
  for (int sub = 0; sub < nsub; ++sub) {
    double *t0 = W->t0, *t1 = W->t1;
    const double b0 = W->coef[0];
    for (int i = lane; i < k; i += 64) {
      const double v = W->y[i];
      t0[i] = v;
      W->y[i] = b0 * v;
    }
    wave_lds_sync();
    for (int i = lane; i < k; i += 64) {  // t1 = S t0
      double v = (sd * d[i] - cd) * t0[i];
      if (i > 0) v = fma(sd * e[i - 1], t0[i - 1], v);
      if (i < k - 1) v = fma(sd * e[i], t0[i + 1], v);
      t1[i] = v;
      if (M >= 1) W->y[i] = fma(W->coef[1], v, W->y[i]);
    }
    wave_lds_sync();
    for (int m = 2; m <= M; ++m) {
      const double bm = W->coef[m];
      for (int i = lane; i < k; i += 64) {  // t0 <- 2 S t1 - t0 (row i only touches t0[i])
        double v = (sd * d[i] - cd) * t1[i];
        if (i > 0) v = fma(sd * e[i - 1], t1[i - 1], v);
        if (i < k - 1) v = fma(sd * e[i], t1[i + 1], v);
        const double t2 = fma(2.0, v, -t0[i]);
        t0[i] = t2;
        W->y[i] = fma(bm, t2, W->y[i]);
      }
      wave_lds_sync();
      double *tmp = t0;
      t0 = t1;
      t1 = tmp;
    }
    for (int i = lane; i < k; i += 64) W->y[i] *= escale;
    wave_lds_sync();
  }
}

// exp(s*T) e1 for k <= 64: row i = lane lives in registers (its entries of
// the scaled S and of the Chebyshev vectors); the neighbours t[i-1], t[i+1]
// come from DPP wavefront shifts, so the recurrence needs no LDS traffic or
// fences (a wavefront that runs beside LDS-heavy kernels on its CU is not
// held up by their LDS queue).  Same formulas, in the same order, as wave_exp_tridiag.  Returns
// y[lane] (0 for lane >= k).
template <typename Lds>
__device__ inline double wave_exp_tridiag_e1_reg(int k, const double *d, const double *e, double s, Lds *W) {
  const int lane = threadIdx.x & 63;
  const bool in = lane < k;
  const double dl = in ? d[lane] : 0.0;
  const double elo = (in && lane > 0) ? e[lane - 1] : 0.0;
  const double ehi = (lane < k - 1) ? e[lane] : 0.0;
  const double di = s * dl;
  const double al = fabs(s * elo), ah = fabs(s * ehi);
  const double glo = wave_min(in ? di - al - ah : 1e308);
  double ghi = wave_max(in ? di + al + ah : -1e308);
  double y = (lane == 0) ? 1.0 : 0.0;
  if (!(ghi - glo > 1e-300)) return y * exp(ghi);  // s*T is a multiple of the identity
  // the Sturm counts read row i's entries from lane i (no memory traffic)
  ghi = wave_tridiag_upper(
      k, [&](int i) { return wave_read_f64(di, i); },
      [&](int i) {
        const double ei = wave_read_f64(al, i);
        return ei * ei;
      },
      glo, ghi);
  const double width = ghi - glo;
  int nsub = (int)ceil(0.5 * width / PSFUN_CHEB_RMAX);
  if (nsub < 1) nsub = 1;
  const double ssub = 1.0 / nsub;
  const double c = 0.5 * (glo + ghi) * ssub, r = 0.5 * width * ssub, hs = ghi * ssub;
  if (lane == 0) W->M = cheb_exp_coefficients(r, W->coef, PSFUN_CHEB_CAP);
  wave_lds_sync();
  const int M = W->M;
  const double rinv = 1.0 / r;
  const double sd = s * ssub * rinv, cd = c * rinv;
  const double sdi = in ? sd * dl - cd : 0.0;   // S_ii
  const double slo = in ? sd * elo : 0.0;       // S_i,i-1
  const double shi = in ? sd * ehi : 0.0;       // S_i,i+1
  const double escale = exp(hs);
  const double b0 = W->coef[0], b1 = (M >= 1) ? W->coef[1] : 0.0;
  for (int sub = 0; sub < nsub; ++sub) {
    double t0 = y;
    y = b0 * t0;
    double t1 = sdi * t0;
    t1 = fma(slo, wave_from_prev(t0), t1);
    t1 = fma(shi, wave_from_next(t0), t1);
    if (M >= 1) y = fma(b1, t1, y);
    // the coefficients 64 at a time, one per lane; term m reads its own by
    // readlane, so the recurrence waits on no LDS load
    for (int mb = 2; mb <= M; mb += 64) {
      const double cb = (mb + lane <= M) ? W->coef[mb + lane] : 0.0;
      const int me = min(M, mb + 63);
#pragma unroll 2
      for (int m = mb; m <= me; ++m) {
        const double bm = wave_read_f64(cb, m - mb);
        double v = sdi * t1;
        v = fma(slo, wave_from_prev(t1), v);
        v = fma(shi, wave_from_next(t1), v);
        const double t2 = fma(2.0, v, -t0);
        y = fma(bm, t2, y);
        t0 = t1;
        t1 = t2;
      }
    }
    y *= escale;
  }
  return in ? y : 0.0;
}

// exp(s*T) e1 for 64 < k <= 128: lane l holds rows 2l and 2l+1 in registers.
// Row 2l's upper neighbour and row 2l+1's lower one are in the same lane; the
// other two come from one wavefront shift each, as many as the one-row
// routine needs.  Each row is updated by the same formulas, in the same
// order, as in wave_exp_tridiag.  Returns the two components in ya (row 2l)
// and yb (row 2l+1), 0 past k.
template <typename Lds>
__device__ inline void wave_exp_tridiag_e1_reg2(int k, const double *d, const double *e, double s, Lds *W,
                                                double &ya, double &yb) {
  const int lane = threadIdx.x & 63;
  const int ia = 2 * lane, ib = ia + 1;
  const bool ina = ia < k, inb = ib < k;
  const double da = ina ? d[ia] : 0.0, db = inb ? d[ib] : 0.0;
  const double ea_lo = (ina && ia > 0) ? e[ia - 1] : 0.0;  // T(ia, ia-1)
  const double eab = (ib < k) ? e[ia] : 0.0;               // T(ia, ib) = T(ib, ia)
  const double eb_hi = (ib < k - 1) ? e[ib] : 0.0;         // T(ib, ib+1)
  const double dia = s * da, dib = s * db;
  const double ala = fabs(s * ea_lo), aab = fabs(s * eab), ahb = fabs(s * eb_hi);
  const double glo = wave_min(fmin(ina ? dia - ala - aab : 1e308, inb ? dib - aab - ahb : 1e308));
  double ghi = wave_max(fmax(ina ? dia + ala + aab : -1e308, inb ? dib + aab + ahb : -1e308));
  ya = (lane == 0) ? 1.0 : 0.0;
  yb = 0.0;
  if (!(ghi - glo > 1e-300)) {  // s*T is a multiple of the identity
    ya *= exp(ghi);
    return;
  }
  // the Sturm counts read row i's entries from lane i / 2 (no memory traffic)
  ghi = wave_tridiag_upper(
      k, [&](int i) { return (i & 1) ? wave_read_f64(dib, i >> 1) : wave_read_f64(dia, i >> 1); },
      [&](int i) {
        const double ei = (i & 1) ? wave_read_f64(aab, i >> 1) : wave_read_f64(ala, i >> 1);
        return ei * ei;
      },
      glo, ghi);
  const double width = ghi - glo;
  int nsub = (int)ceil(0.5 * width / PSFUN_CHEB_RMAX);
  if (nsub < 1) nsub = 1;
  const double ssub = 1.0 / nsub;
  const double c = 0.5 * (glo + ghi) * ssub, r = 0.5 * width * ssub, hs = ghi * ssub;
  if (lane == 0) W->M = cheb_exp_coefficients(r, W->coef, PSFUN_CHEB_CAP);
  wave_lds_sync();
  const int M = W->M;
  const double rinv = 1.0 / r;
  const double sd = s * ssub * rinv, cd = c * rinv;
  const double sda = ina ? sd * da - cd : 0.0, sdb = inb ? sd * db - cd : 0.0;  // S_ii
  const double sla = ina ? sd * ea_lo : 0.0;                                  // S(ia, ia-1)
  const double sha = ina ? sd * eab : 0.0, slb = inb ? sd * eab : 0.0;        // S(ia, ib), S(ib, ia)
  const double shb = inb ? sd * eb_hi : 0.0;                                  // S(ib, ib+1)
  const double escale = exp(hs);
  const double b0 = W->coef[0], b1 = (M >= 1) ? W->coef[1] : 0.0;
  for (int sub = 0; sub < nsub; ++sub) {
    double t0a = ya, t0b = yb;
    ya = b0 * t0a;
    yb = b0 * t0b;
    double t1a = sda * t0a, t1b = sdb * t0b;
    t1a = fma(sla, wave_from_prev(t0b), t1a);
    t1a = fma(sha, t0b, t1a);
    t1b = fma(slb, t0a, t1b);
    t1b = fma(shb, wave_from_next(t0a), t1b);
    if (M >= 1) {
      ya = fma(b1, t1a, ya);
      yb = fma(b1, t1b, yb);
    }
    for (int mb = 2; mb <= M; mb += 64) {  // coefficients as in the one-row routine
      const double cb = (mb + lane <= M) ? W->coef[mb + lane] : 0.0;
      const int me = min(M, mb + 63);
#pragma unroll 2
      for (int m = mb; m <= me; ++m) {
        const double bm = wave_read_f64(cb, m - mb);
        double va = sda * t1a, vb = sdb * t1b;
        va = fma(sla, wave_from_prev(t1b), va);
        va = fma(sha, t1b, va);
        vb = fma(slb, t1a, vb);
        vb = fma(shb, wave_from_next(t1a), vb);
        const double t2a = fma(2.0, va, -t0a), t2b = fma(2.0, vb, -t0b);
        ya = fma(bm, t2a, ya);
        yb = fma(bm, t2b, yb);
        t0a = t1a;
        t0b = t1b;
        t1a = t2a;
        t1b = t2b;
      }
    }
    ya *= escale;
    yb *= escale;
  }
  if (!ina) ya = 0.0;
  if (!inb) yb = 0.0;
}

// exp(s*T) e1 for 64 < k <= 128 through the two-row register routine; returns
// ||y||_2 and y[k-1] in every lane and, if yout is not null, stores y[0..k-1].
template <typename Lds>
__device__ inline void wave_exp_tridiag_e1_mid(int k, const double *d, const double *e, double s, Lds *W,
                                               double *yout, double &nrm, double &last) {
  const int lane = threadIdx.x & 63;
  double ya, yb;
  wave_exp_tridiag_e1_reg2(k, d, e, s, W, ya, yb);
  if (yout) {
    if (2 * lane < k) yout[2 * lane] = ya;
    if (2 * lane + 1 < k) yout[2 * lane + 1] = yb;
  }
  nrm = sqrt(wave_sum(fma(ya, ya, yb * yb)));
  const double la = __shfl(ya, (k - 1) >> 1, 64), lb = __shfl(yb, (k - 1) >> 1, 64);
  last = ((k - 1) & 1) ? lb : la;
}

// exp(s*T) e1 for k <= 64 through the register routine; returns ||y||_2 and
// y[k-1] in every lane and, if yout is not null, stores y[0..k-1].
template <typename Lds>
__device__ inline void wave_exp_tridiag_e1_small(int k, const double *d, const double *e, double s, Lds *W,
                                                 double *yout, double &nrm, double &last) {
  const int lane = threadIdx.x & 63;
  const double y = wave_exp_tridiag_e1_reg(k, d, e, s, W);
  if (yout && lane < k) yout[lane] = y;
  nrm = sqrt(wave_sum(y * y));
  last = __shfl(y, k - 1, 64);
}

// Any k <= PSFUN_FT_KMAX through the LDS routine.
__device__ void wave_exp_tridiag_e1_lds(int k, const double *d, const double *e, double s, ChebLds *W,
                                        double *yout, double &nrm, double &last) {
  const int lane = threadIdx.x & 63;
  for (int i = lane; i < k; i += 64) W->y[i] = (i == 0) ? 1.0 : 0.0;
  wave_lds_sync();
  wave_exp_tridiag(k, d, e, s, W);
  double ss = 0.0;
  for (int i = lane; i < k; i += 64) {
    const double v = W->y[i];
    ss = fma(v, v, ss);
    if (yout) yout[i] = v;
  }
  nrm = sqrt(wave_sum(ss));
  last = W->y[k - 1];
}

// exp(s*T) e1 for any k <= PSFUN_FT_KMAX (registers up to 128, LDS beyond).
__device__ inline void wave_exp_tridiag_e1(int k, const double *d, const double *e, double s, ChebLds *W,
                                           double *yout, double &nrm, double &last) {
  if (k <= 64)
    wave_exp_tridiag_e1_small(k, d, e, s, W, yout, nrm, last);
  else if (k <= 128)
    wave_exp_tridiag_e1_mid(k, d, e, s, W, yout, nrm, last);
  else
    wave_exp_tridiag_e1_lds(k, d, e, s, W, yout, nrm, last);
}

