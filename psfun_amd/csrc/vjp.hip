// One-call vector-Jacobian product of Y = exp(-tA) X for C and Fortran
// callers (no reference counterpart; the reference is not differentiable).
// It runs the same device sequence as the torch.autograd backward of the
// Python layer with the forward basis recomputed:
//   dL/dX    = exp(-tA) G              batched Lanczos on G, keeping its basis W
//   U_s      = (xi_s I + A)^{-1} G      Galerkin resolvents from W
//   V_s      = (xi_s I + A)^{-1} X      forward Lanczos on X again, resolvents from V
//   dL/da_q  = -Re sum_s w_s sum_col U_s[i,col] V_s[k,col]   (SDDMM over A's pattern)
//   dL/dt    = -<G, A Y>
// with (xi_s, w_s) the modified Talbot rule for exp(-t z) whose conjugate
// pairs are folded into one pole of twice the weight (Re assembly).
// A must be symmetric; the batch holds that one matrix in its own numbering.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "common.hpp"

namespace {

// Device buffer released on every return path.
struct DevBuf {
  double *p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  int alloc(size_t count) {
    if (hipMalloc(&p, sizeof(double) * std::max<size_t>(count, 1)) != hipSuccess) {
      p = nullptr;
      return PSFUN_ERR_ALLOC;
    }
    return PSFUN_SUCCESS;
  }
};

// Talbot poles and weights with each conjugate pair (s, w), (conj s, conj w)
// merged into (s, 2w): the first pole of a pair is kept, its partner is the
// unused pole nearest to (conj s, conj w) within a relative 1e-12.
int folded_talbot(int npoles, double t, std::vector<double> &shifts, std::vector<double> &weights) {
  using zd = std::complex<double>;
  std::vector<double> xi(2 * npoles), cw(2 * npoles);
  double eta = 0.0, sign = 0.0;
  PSFUN_TRY(psfun_talbot(npoles, t, xi.data(), cw.data(), &eta, &sign));
  std::vector<zd> s(npoles), w(npoles);
  for (int j = 0; j < npoles; ++j) {
    s[j] = sign * zd(xi[2 * j], xi[2 * j + 1]);
    w[j] = zd(cw[2 * j], cw[2 * j + 1]);
  }
  const double tol = 1e-13;
  std::vector<char> used(npoles, 0);
  shifts.clear();
  weights.clear();
  for (int i = 0; i < npoles; ++i) {
    if (used[i]) continue;
    used[i] = 1;
    const double scale = std::max(std::abs(s[i]), 1e-300);
    double mult = 1.0;
    if (std::fabs(s[i].imag()) > tol * scale) {
      int best = -1;
      double dbest = INFINITY;
      for (int j = 0; j < npoles; ++j) {
        if (used[j]) continue;
        const double d = std::abs(s[j] - std::conj(s[i])) +
                         std::abs(w[j] - std::conj(w[i])) / std::max(std::abs(w[i]), 1e-300) * scale;
        if (d < dbest) {
          dbest = d;
          best = j;
        }
      }
      if (best >= 0 && dbest <= tol * scale * 10) {
        used[best] = 1;
        mult = 2.0;
      }
    }
    shifts.push_back(s[i].real());
    shifts.push_back(s[i].imag());
    weights.push_back(mult * w[i].real());
    weights.push_back(mult * w[i].imag());
  }
  return PSFUN_SUCCESS;
}

// Galerkin resolvents of every folded pole from a fresh Lanczos run on the
// n x K block R (the operator -t*A, so the stored tridiagonal maps back to A
// with tscale = -1/t).  Y receives exp(-tA) R.
int resolvents_of(psfun_ctx *c, psfun_batch_t B, int64_t n, int K, double t, const double *R, double *Y, double eps,
                  int itmax, int S, const double *shifts, double *out) {
  const int ld = itmax + 2;
  DevBuf V, T;
  PSFUN_TRY(V.alloc((size_t)(itmax - 1) * n * K));
  PSFUN_TRY(T.alloc((size_t)K * (2 * ld + 1) + (K + 1) / 2));  // T_alpha, T_beta, rs1, nbasis
  double *Ta = T.p, *Tb = Ta + (size_t)K * ld, *rs1 = Tb + (size_t)K * ld;
  int *nb = reinterpret_cast<int *>(rs1 + K);
  PSFUN_TRY(psfun_batch_lanczos_basis(c, B, K, 1.0, -t, R, Y, eps, itmax, 1, V.p, Ta, Tb, nb, rs1));
  std::vector<int> hnb(K);
  PSFUN_HIP_TRY(hipMemcpyAsync(hnb.data(), nb, sizeof(int) * K, hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  const int mmax = std::max(1, *std::max_element(hnb.begin(), hnb.end()));
  PSFUN_TRY(psfun_krylov_resolvents(c, n, K, mmax, V.p, Ta, Tb, ld, -1.0 / t, nb, rs1, S, shifts, out));
  // the basis is freed on return; the stream must be done reading it
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSFUN_SUCCESS;
}

}  // namespace

extern "C" int psfun_expv_vjp(psfun_ctx_t c, psfun_batch_t B, int K, double t, const double *X, const double *Y,
                              const double *G, double eps, int itmax, int npoles, double *gX, double *gt,
                              double *gvals) {
  if (!c || !B || !G || K < 1 || K > 64 || itmax < 2 || !(eps >= 0.0)) return PSFUN_ERR_INVALID_ARG;
  if ((gt && !Y) || (gvals && !X)) return PSFUN_ERR_INVALID_ARG;
  if (gvals && (!(t > 0.0) || npoles < 1 || npoles > 64)) return PSFUN_ERR_INVALID_ARG;
  int ngraph = 0;
  int64_t n = 0, nnz = 0;
  PSFUN_TRY(psfun_batch_info(B, &ngraph, &n, &nnz, nullptr, nullptr));
  if (ngraph != 1) return PSFUN_ERR_INVALID_ARG;
  const int32_t *rowptr = nullptr, *colind = nullptr;
  if (gvals) PSFUN_TRY(psfun_batch_csr(B, &rowptr, &colind, nullptr));  // refuses a reordered batch
  if (n == 0) {
    if (gt) *gt = 0.0;
    return PSFUN_SUCCESS;
  }
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  const size_t nk = (size_t)n * K;

  if (gt) {  // dL/dt = -<G, A Y>
    DevBuf AY;
    PSFUN_TRY(AY.alloc(nk));
    PSFUN_TRY(psfun_batch_spmm(c, B, K, 1.0, Y, 0.0, AY.p));
    double d = 0.0;
    PSFUN_TRY(psfun_dot(c, (int64_t)nk, G, AY.p, &d));
    *gt = -d;
  }
  if (!gvals) {
    if (gX) PSFUN_TRY(psfun_batch_lanczos(c, B, K, 1.0, -t, G, gX, eps, itmax, 1, nullptr, nullptr));
    return PSFUN_SUCCESS;
  }

  std::vector<double> shifts, weights;
  PSFUN_TRY(folded_talbot(npoles, t, shifts, weights));
  const int S = (int)shifts.size() / 2;
  DevBuf U, Vr, scratch;
  PSFUN_TRY(U.alloc(2 * (size_t)S * nk));
  PSFUN_TRY(scratch.alloc(nk));
  // adjoint run: dL/dX (into gX, or a scratch block) and the resolvents U of G
  PSFUN_TRY(resolvents_of(c, B, n, K, t, G, gX ? gX : scratch.p, eps, itmax, S, shifts.data(), U.p));
  PSFUN_TRY(Vr.alloc(2 * (size_t)S * nk));
  // forward run again (its output exp(-tA)X goes to scratch): resolvents of X
  PSFUN_TRY(resolvents_of(c, B, n, K, t, X, scratch.p, eps, itmax, S, shifts.data(), Vr.p));
  if (nnz > 0) PSFUN_HIP_TRY(hipMemsetAsync(gvals, 0, sizeof(double) * nnz, c->stream));
  PSFUN_TRY(psfun_sddmm_cols(c, (int)n, rowptr, colind, K, S, weights.data(), U.p, Vr.p, -1.0, gvals));
  // U and Vr are freed on return
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSFUN_SUCCESS;
}
