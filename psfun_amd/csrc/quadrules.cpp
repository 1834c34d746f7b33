// Host-side scalar special functions and quadrature rules (pole/weight
// generation).  These run once per rule on the host, as in the reference; no
// device is needed.
//
//   psfun_horner     -- utils/psfun_d_utils_mod.f90:119-138
//   psfun_ellipkkp   -- utils/psfun_d_utils_mod.f90:37-99 (AGM, L > 10 asymptotic)
//   psfun_ellipj     -- utils/psfun_z_utils_mod.f90:39-107 (descending Landen)
//   psfun_hhtmethod1 -- quadratures/psfun_z_quadrules_mod.f90:40-85
//   psfun_talbot     -- modified Talbot contour for exp(-t z) (Trefethen,
//                       Weideman & Schmelzer, BIT 46 (2006) 653-670)
#include <cfloat>
#include <cmath>
#include <complex>

#include "../../include/psfun_hip.h"

using zd = std::complex<double>;
static const double kPi = 3.14159265358979323846264338327950288;

extern "C" double psfun_horner(int nc, const double *c, double x) {
  double r = 0.0;
  for (int i = nc - 1; i >= 0; --i) r = r * x + c[i];
  return r;
}

// Arithmetic-geometric mean of (1, b0) run until 2^i c_i^2 <= machine epsilon.
static double agm_quarter_period(double b0) {
  double a0 = 1.0, a1 = 1.0, mm = 1.0;
  int i = 0;
  while (mm > DBL_EPSILON) {
    a1 = 0.5 * (a0 + b0);
    const double b1 = std::sqrt(a0 * b0);
    const double c1 = 0.5 * (a0 - b0);
    ++i;
    mm = std::ldexp(c1 * c1, i);
    a0 = a1;
    b0 = b1;
  }
  return kPi / (2.0 * a1);
}

extern "C" int psfun_ellipkkp(double L, double *K) {
  if (!K || !(L > 0.0)) return PSFUN_ERR_INVALID_ARG;
  if (L > 10.0) {  // M = exp(-2 pi L) ~ 0
    K[0] = kPi / 2.0;
    K[1] = kPi * L + std::log(4.0);
    return PSFUN_SUCCESS;
  }
  const double M = std::exp(-2.0 * kPi * L);
  K[0] = agm_quarter_period(std::sqrt(1.0 - M));
  K[1] = agm_quarter_period(std::sqrt(M));
  return PSFUN_SUCCESS;
}

// flag: L is the nome exponent (m = exp(-2 pi L)); otherwise L is m itself.
static void zellipj(zd u, double L, bool flag, zd &sn, zd &cn, zd &dn) {
  static const double kappa_series[7] = {0.0, 1.0, 2.0, 5.0, 14.0, 42.0, 132.0};
  double m;
  bool high = false;
  if (flag) {
    double K[2];
    psfun_ellipkkp(L, K);
    high = u.imag() > 0.5 * K[1];
    // The reflected argument i K' - u is computed by the reference but the
    // Landen step below keeps using u (psfun_z_utils_mod.f90:64-67, 76-88).
    m = std::exp(-2.0 * kPi * L);
  } else {
    m = L;
  }
  if (m < 4.0 * DBL_EPSILON) {
    const zd s = std::sin(u), c = std::cos(u);
    sn = s + m / 4.0 * (s * c - u) * c;
    cn = c + m / 4.0 * (-s * c + u) * s;
    dn = 1.0 + m / 4.0 * (c * c - s * s - 1.0);
  } else {
    const double kappa = (m > 1e-3) ? (1.0 - std::sqrt(1.0 - m)) / (1.0 + std::sqrt(1.0 - m))
                                    : psfun_horner(7, kappa_series, m / 4.0);
    zd s1, c1, d1;
    zellipj(u / (1.0 + kappa), kappa * kappa, false, s1, c1, d1);
    const zd den = 1.0 + kappa * s1 * s1;
    sn = (1.0 + kappa) * s1 / den;
    cn = c1 * d1 / den;
    dn = (1.0 - kappa * s1 * s1) / den;
  }
  if (high) {
    const zd sh = sn, ch = cn, dh = dn;
    sn = -1.0 / (std::sqrt(m) * sh);
    cn = zd(0.0, 1.0) * dh / (std::sqrt(m) * sh);
    dn = zd(0.0, 1.0) * ch / sh;
  }
}

extern "C" int psfun_ellipj(double ure, double uim, double L, double *out) {
  if (!out || !(L > 0.0)) return PSFUN_ERR_INVALID_ARG;
  zd sn, cn, dn;
  zellipj(zd(ure, uim), L, true, sn, cn, dn);
  out[0] = sn.real();
  out[1] = sn.imag();
  out[2] = cn.real();
  out[3] = cn.imag();
  out[4] = dn.real();
  out[5] = dn.imag();
  return PSFUN_SUCCESS;
}

extern "C" int psfun_hhtmethod1(psfun_zfun_t zfun, void *user, int N, const double *rparams, double *xi,
                                double *c, double *eta, double *sign) {
  if (!zfun || N < 1 || !rparams || !xi || !c || !eta || !sign) return PSFUN_ERR_INVALID_ARG;
  const double mineig = std::fmin(rparams[0], rparams[1]), maxeig = std::fmax(rparams[0], rparams[1]);
  if (!(mineig > 0.0) || !(maxeig > mineig)) return PSFUN_ERR_INVALID_ARG;
  const double q = std::sqrt(maxeig / mineig);
  const double k = (q - 1.0) / (q + 1.0);
  const double L = -std::log(k) / kPi;
  double K[2];
  psfun_ellipkkp(L, K);
  const double smM = std::sqrt(mineig * maxeig);
  for (int j = 1; j <= N; ++j) {
    const zd t = zd(0.0, 0.5 * K[1]) - K[0] + (2.0 * (j - 1) + 1.0) * K[0] / N;
    zd u, cn, dn;
    zellipj(t, L, true, u, cn, dn);
    const zd x = smM * ((1.0 / k + u) / (1.0 / k - u));
    double z[2] = {x.real(), x.imag()}, fz[2] = {0.0, 0.0};
    zfun(z, fz, user);
    const zd cc = (zd(fz[0], fz[1]) / x) * (cn * dn) / ((1.0 / k - u) * (1.0 / k - u));
    xi[2 * (j - 1)] = x.real();
    xi[2 * (j - 1) + 1] = x.imag();
    c[2 * (j - 1)] = cc.real();
    c[2 * (j - 1) + 1] = cc.imag();
  }
  *eta = -(4.0 * K[0] * smM) / (k * kPi * N);
  *sign = -1.0;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_talbot(int N, double t, double *xi, double *c, double *eta, double *sign) {
  if (N < 1 || !(t > 0.0) || !xi || !c || !eta || !sign) return PSFUN_ERR_INVALID_ARG;
  for (int j = 1; j <= N; ++j) {
    const double th = -kPi + (j - 0.5) * 2.0 * kPi / N;
    const double a = 0.6407 * th;
    const double sn = std::sin(a), cot = std::cos(a) / sn;
    const zd z = double(N) * zd(0.5017 * th * cot - 0.6122, 0.2645 * th);
    const zd dz = double(N) * zd(0.5017 * cot - 0.5017 * 0.6407 * th / (sn * sn), 0.2645);
    const zd w = std::exp(z) * dz / (zd(0.0, 1.0) * double(N) * t);
    xi[2 * (j - 1)] = z.real() / t;
    xi[2 * (j - 1) + 1] = z.imag() / t;
    c[2 * (j - 1)] = w.real();
    c[2 * (j - 1) + 1] = w.imag();
  }
  *eta = 1.0;
  *sign = 1.0;
  return PSFUN_SUCCESS;
}
