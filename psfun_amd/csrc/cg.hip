// Shifted preconditioned CG, (eta*A + zeta*I) x = b, with the semantics of
// psfun_dcg_vect (krylov/shifted/psfun_dcg.F90:249-357) and PSBLAS's stopping
// rules (istop 1: ||r||_inf <= eps (||A||_inf ||x||_inf + ||b||_inf) with the
// norm of the UNshifted A, line 272; istop 2: ||r||_2 <= eps ||b||_2; the
// iteration also ends when itx >= itmax).  Preconditioner: none, or the
// diagonal of the unshifted A (the DIAG preconditioner built on A).
//
// Device schedule per iteration `it` (global count itx):
//   k_cg_a(it): every block folds the rho partials of the previous k_cg_b
//               (of k_cg_resid, which runs on the SpMV grid, when it == 1)
//               (rho_it = r.z), beta = rho_it / rho_{it-1}; p_new = z + beta
//               p_old for its rows (z = M^{-1} r and p_old gathered, p_new
//               formed on the fly at every column), q = eta*A p_new +
//               zeta*p_new, partials of sigma = p.q.  A control block runs the reference's checks
//               for the previous iteration: psb_check_conv (from the
//               ||r||/||x|| partials) and the rho_old == 0 breakdown.
//   k_cg_b(it): sigma from partials; sigma == 0 is a breakdown; alpha =
//               rho/sigma; x += alpha p; r -= alpha q; z = M^{-1} r for DIAG
//               (the preconditioner object writes z for BJAC); partials of
//               rho_next = r.z and of the convergence norms.
// A breakdown ends the device loop; the host then restarts the way the
// reference does (r = b - (eta A + zeta I) x, it = 0) while itx < itmax.
#include "common.hpp"
#include "kernels.hpp"

enum { CG_RUN = 0, CG_CONV = 1, CG_BRK = 2, CG_MAXIT = 3 };

struct CgCtl {
  int done;
  int itx;
  int it;
  int pad;
  double errnum, errden;
  double ani, bni, bn2;   // ||A||_inf, ||b||_inf, ||b||_2
  double rho[2];          // rho of the last two iterations (ring by parity)
  double alpha_hist[PSFUN_MAX_GRID];  // CG alphas/betas for the condition estimate
  double beta_hist[PSFUN_MAX_GRID];
};

#define CG_HIST_MAX PSFUN_MAX_GRID
#define CG_UB 4  // elements per thread in flight in k_cg_b

__device__ __forceinline__ double prec_apply(const double *__restrict__ dinv, int64_t i, double r) {
  return dinv ? r * dinv[i] : r;
}

template <int L, bool SELL>
__global__ __launch_bounds__(PSFUN_BS) void k_cg_a(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                   const double *__restrict__ val, double eta, double zeta,
                                                   const double *__restrict__ r, const double *__restrict__ dinv,
                                                   const double *__restrict__ p_old, double *__restrict__ p_new,
                                                   double *__restrict__ q, int it, int G,
                                                   const double *__restrict__ p_rho, int g_rho,
                                                   const double *__restrict__ p_conv, int g_in,
                                                   double *__restrict__ p_sig, CgCtl *ctl, int istop, double eps,
                                                   int itmax, int *hf, const double *__restrict__ zv) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  // block-uniform: the control block of this same grid may set the flag
  if (__syncthreads_or(ctl->done)) return;
  const double rho = block_sum_partials<PSFUN_BS>(p_rho, g_rho, lds);
  const double rho_old = (it > 1) ? ctl->rho[(it - 1) & 1] : 0.0;
  if (blockIdx.x == G) {  // control block
    double cnum = 0.0, cden = 0.0;
    if (it > 1) {  // psb_check_conv for the previous iteration
      if (istop == 1) {
        double rmax = 0.0, xmax = 0.0;
        for (int b = threadIdx.x; b < g_in; b += PSFUN_BS) {
          rmax = fmax(rmax, p_conv[b]);
          xmax = fmax(xmax, p_conv[g_in + b]);
        }
        rmax = block_max<PSFUN_BS>(rmax, lds);
        xmax = block_max<PSFUN_BS>(xmax, lds);
        cnum = rmax;
        cden = ctl->ani * xmax + ctl->bni;
      } else {
        cnum = sqrt(block_sum_partials<PSFUN_BS>(p_conv, g_in, lds));
        cden = ctl->bn2;
      }
    }
    if (threadIdx.x == 0) {
      ctl->rho[it & 1] = rho;
      ctl->it = it;
      if (it > 1) {
        ctl->errnum = cnum;
        ctl->errden = cden;
        const int ok = (cden == 0.0) ? (cnum <= eps) : (cnum <= eps * cden);
        if (ok) ctl->done = CG_CONV;
        else if (ctl->itx >= itmax) ctl->done = CG_MAXIT;
        else if (rho_old == 0.0) {  // psfun_dcg.F90:290-295
          ctl->done = CG_BRK;
          ctl->itx += 1;  // the reference counts the iteration that breaks down (itx = itx + 1 at 281)
        }
      }
      if (!ctl->done) ctl->itx += 1;
      host_flag_store(hf, it);
      if (ctl->done) host_flag_store(hf + 1, ctl->done);
    }
    return;
  }
  if (it > 1 && rho_old == 0.0) return;
  const double beta = (it > 1) ? rho / rho_old : 0.0;
  // direction p = M^{-1} r + beta p_old, formed on the fly at every gathered
  // column; z = M^{-1} r is read from zv where it is stored (DIAG, BJAC)
  auto zk = [&](int c) -> double { return zv ? zv[c] : prec_apply(dinv, c, r[c]); };
  auto pk = [&](int c) -> double { return (it > 1) ? fma(beta, p_old[c], zk(c)) : zk(c); };
  constexpr int RPB = SELL ? PSFUN_BS : PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range_g(nchunks, G);
  const int sub = SELL ? threadIdx.x : threadIdx.x / L, lane = SELL ? 0 : threadIdx.x % L;
  double sig = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    const double sum = row_dot<L, SELL>(row, n, rowptr, col, val, lane, pk);
    if (row < n && lane == 0) {
      const double pr = pk(row);
      const double qr = fma(zeta, pr, eta * sum);
      p_new[row] = pr;
      q[row] = qr;
      sig = fma(pr, qr, sig);
    }
  }
  sig = block_sum<PSFUN_BS>(sig, lds);
  if (threadIdx.x == 0) p_sig[blockIdx.x] = sig;
}

__global__ __launch_bounds__(PSFUN_BS) void k_cg_b(int n, double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   const double *__restrict__ dinv, int it,
                                                   const double *__restrict__ p_sig, int g_sig,
                                                   double *__restrict__ p_rho, double *__restrict__ p_conv,
                                                   CgCtl *ctl, int istop, int *hf, double *__restrict__ zout) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  const double sigma = block_sum_partials<PSFUN_BS>(p_sig, g_sig, lds);
  const double rho = ctl->rho[it & 1];
  if (sigma == 0.0) {  // psfun_dcg.F90:305-310
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->done = CG_BRK;
      host_flag_store(hf + 1, CG_BRK);  // the later k_cg_a launches return before their control block
    }
    return;
  }
  const double alpha = rho / sigma;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int h = ctl->itx - 1;
    if (h >= 0 && h < CG_HIST_MAX) {
      ctl->alpha_hist[h] = alpha;
      ctl->beta_hist[h] = (it > 1) ? rho / ctl->rho[(it - 1) & 1] : 0.0;
    }
  }
  double arho = 0.0, a1 = 0.0, a2 = 0.0;
  // CG_UB elements per thread at a time, their loads in flight together;
  // each thread still visits its elements in the same order, so the
  // partials are unchanged
  const int64_t gs = (int64_t)gridDim.x * PSFUN_BS;
  int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x;
  for (; i + (CG_UB - 1) * gs < n; i += CG_UB * gs) {
    double xv[CG_UB], pv[CG_UB], rv[CG_UB], qv[CG_UB], dv[CG_UB];
#pragma unroll
    for (int u = 0; u < CG_UB; ++u) {
      const int64_t k = i + u * gs;
      xv[u] = x[k];
      pv[u] = p[k];
      rv[u] = r[k];
      qv[u] = q[k];
      dv[u] = dinv ? dinv[k] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < CG_UB; ++u) {
      const int64_t k = i + u * gs;
      const double xi = fma(alpha, pv[u], xv[u]);
      const double ri = fma(-alpha, qv[u], rv[u]);
      x[k] = xi;
      r[k] = ri;
      const double zi = dinv ? ri * dv[u] : ri;
      if (zout) zout[k] = zi;
      arho = fma(ri, zi, arho);
      if (istop == 1) {
        a1 = fmax(a1, fabs(ri));
        a2 = fmax(a2, fabs(xi));
      } else {
        a1 = fma(ri, ri, a1);
      }
    }
  }
  for (; i < n; i += gs) {
    const double xi = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    x[i] = xi;
    r[i] = ri;
    const double zi = prec_apply(dinv, i, ri);
    if (zout) zout[i] = zi;
    arho = fma(ri, zi, arho);
    if (istop == 1) {
      a1 = fmax(a1, fabs(ri));
      a2 = fmax(a2, fabs(xi));
    } else {
      a1 = fma(ri, ri, a1);
    }
  }
  arho = block_sum<PSFUN_BS>(arho, lds);
  if (istop == 1) {
    a1 = block_max<PSFUN_BS>(a1, lds);
    a2 = block_max<PSFUN_BS>(a2, lds);
  } else {
    a1 = block_sum<PSFUN_BS>(a1, lds);
  }
  if (threadIdx.x == 0) {
    p_rho[blockIdx.x] = arho;
    p_conv[blockIdx.x] = a1;
    if (istop == 1) p_conv[gridDim.x + blockIdx.x] = a2;
  }
}

// Partials of rho = r.z (z = M^{-1} r formed by a preconditioner object).
__global__ __launch_bounds__(PSFUN_BS) void k_cg_rz(int n, const double *__restrict__ r, const double *__restrict__ z,
                                                    double *__restrict__ p_rho, const CgCtl *ctl) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS)
    acc = fma(r[i], z[i], acc);
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_rho[blockIdx.x] = acc;
}

// r = b - eta*A x - zeta*x; partials of rho = r.M^{-1}r (restart entry, lines 259-262).
template <int L, bool SELL>
__global__ __launch_bounds__(PSFUN_BS) void k_cg_resid(int n, const int *__restrict__ rowptr,
                                                       const int *__restrict__ col, const double *__restrict__ val,
                                                       double eta, double zeta, const double *__restrict__ b,
                                                       const double *__restrict__ x, double *__restrict__ r,
                                                       const double *__restrict__ dinv, double *__restrict__ p_rho,
                                                       double *__restrict__ zout) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  constexpr int RPB = SELL ? PSFUN_BS : PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range(nchunks);
  const int sub = SELL ? threadIdx.x : threadIdx.x / L, lane = SELL ? 0 : threadIdx.x % L;
  double acc = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    const double sum = row_dot<L, SELL>(row, n, rowptr, col, val, lane, [&](int c) { return x[c]; });
    if (row < n && lane == 0) {
      const double rr = fma(-zeta, x[row], fma(-eta, sum, b[row]));
      r[row] = rr;
      const double zr = prec_apply(dinv, row, rr);
      if (zout) zout[row] = zr;
      acc = fma(rr, zr, acc);
    }
  }
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_rho[blockIdx.x] = acc;
}

// Row statistics of A: dinv = 1/diag(A) (1 where the diagonal is 0), and the
// per-block maxima of the absolute row sums (||A||_inf partials); in the same
// pass the per-block partials of ||b||_inf (p_b[block]) and ||b||_2^2
// (p_b[gridDim + block]).
__global__ __launch_bounds__(PSFUN_BS) void k_cg_rowstats(int n, const int *__restrict__ rowptr,
                                                          const int *__restrict__ col, const double *__restrict__ val,
                                                          const double *__restrict__ b, double *__restrict__ dinv,
                                                          double *__restrict__ p_ani, double *__restrict__ p_b) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double mx = 0.0, bmx = 0.0, bss = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    double s = 0.0, dg = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      s += fabs(val[k]);
      if (col[k] == i) dg += val[k];
    }
    mx = fmax(mx, s);
    if (dinv) dinv[i] = (dg != 0.0) ? 1.0 / dg : 1.0;
    const double bi = b[i];
    bmx = fmax(bmx, fabs(bi));
    bss = fma(bi, bi, bss);
  }
  mx = block_max<PSFUN_BS>(mx, lds);
  bmx = block_max<PSFUN_BS>(bmx, lds);
  bss = block_sum<PSFUN_BS>(bss, lds);
  if (threadIdx.x == 0) {
    p_ani[blockIdx.x] = mx;
    p_b[blockIdx.x] = bmx;
    p_b[gridDim.x + blockIdx.x] = bss;
  }
}

// psb_init_conv: ||A||_inf, ||b||_inf, ||b||_2 from the k_cg_rowstats
// partials (g of each; one block, a fixed fold order).
__global__ __launch_bounds__(PSFUN_BS) void k_cg_initconv(const double *__restrict__ p_ani,
                                                          const double *__restrict__ p_b, int g, CgCtl *ctl,
                                                          int reset) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double mx = 0.0, s2 = 0.0, a = 0.0;
  for (int i = threadIdx.x; i < g; i += PSFUN_BS) {
    mx = fmax(mx, p_b[i]);
    s2 += p_b[g + i];
    a = fmax(a, p_ani[i]);
  }
  mx = block_max<PSFUN_BS>(mx, lds);
  s2 = block_sum<PSFUN_BS>(s2, lds);
  a = block_max<PSFUN_BS>(a, lds);
  if (threadIdx.x == 0) {
    ctl->bni = mx;
    ctl->bn2 = sqrt(s2);
    ctl->ani = a;
    if (reset) {
      ctl->itx = 0;
      ctl->errnum = 0.0;
      ctl->errden = 1.0;
    }
    ctl->done = CG_RUN;
  }
}

// Final convergence check after the last iteration the host launched.
__global__ __launch_bounds__(PSFUN_BS) void k_cg_lastcheck(const double *__restrict__ p_conv, int g_in, CgCtl *ctl,
                                                           int istop, double eps, int itmax) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  double cnum, cden;
  if (istop == 1) {
    double rmax = 0.0, xmax = 0.0;
    for (int b = threadIdx.x; b < g_in; b += PSFUN_BS) {
      rmax = fmax(rmax, p_conv[b]);
      xmax = fmax(xmax, p_conv[g_in + b]);
    }
    cnum = block_max<PSFUN_BS>(rmax, lds);
    cden = ctl->ani * block_max<PSFUN_BS>(xmax, lds) + ctl->bni;
  } else {
    cnum = sqrt(block_sum_partials<PSFUN_BS>(p_conv, g_in, lds));
    cden = ctl->bn2;
  }
  if (threadIdx.x == 0) {
    ctl->errnum = cnum;
    ctl->errden = cden;
    const int ok = (cden == 0.0) ? (cnum <= eps) : (cnum <= eps * cden);
    ctl->done = ok ? CG_CONV : CG_MAXIT;
  }
}

// Extreme eigenvalues of the symmetric tridiagonal (d, e) by bisection on
// Sturm counts (host); used for the CG condition estimate.
static double sturm_count_lt(const std::vector<double> &d, const std::vector<double> &e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); ++i) {
    const double e2 = (i > 0) ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - ((i > 0) ? e2 / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

static double tridiag_extreme(const std::vector<double> &d, const std::vector<double> &e, bool largest) {
  double lo = 1e308, hi = -1e308;
  for (size_t i = 0; i < d.size(); ++i) {
    const double rad = ((i > 0) ? fabs(e[i - 1]) : 0.0) + ((i + 1 < d.size()) ? fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - rad);
    hi = std::max(hi, d[i] + rad);
  }
  const int target = largest ? (int)d.size() - 1 : 0;  // eigenvalue index (0-based, ascending)
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(fabs(lo), fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_count_lt(d, e, mid) > target) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// The solver: prec NONE or DIAG (formed in the kernels), or the
// preconditioner object P (BJAC: z = M^{-1} r by its triangular solves after
// every residual update, then the rho partials r.z).
static int cg_run(psfun_ctx *c, psfun_csr *A, int prec, const psfun_prec *P, const double *b, double eta,
                  double zeta, double *x, double eps, int itmax, int istop, int *iter, double *err, double *cond) {
  const int n = A->n;
  if (cond) *cond = 0.0;
  if (n == 0) {
    if (iter) *iter = 0;
    if (err) *err = 0.0;
    return PSFUN_SUCCESS;
  }
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  const bool zmode = prec == PSFUN_PREC_BJAC;
  const size_t ctl_bytes = align_up(sizeof(CgCtl), 256);
  const size_t vec = align_up((size_t)n, 64);
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, ctl_bytes + sizeof(double) * vec * 6, &ws));
  CgCtl *ctl = (CgCtl *)ws;
  double *r = (double *)((char *)ws + ctl_bytes);
  double *pA = r + vec, *pB = pA + vec, *q = pB + vec, *dinv = q + vec, *zv = dinv + vec;
  if (prec != PSFUN_PREC_DIAG) dinv = nullptr;
  // z = M^{-1} r is stored for DIAG (by the residual updates) and for BJAC
  // (by the preconditioner object), so the SpMV step gathers z and p_old only
  double *zdiag = dinv ? zv : nullptr;
  if (!zmode && !zdiag) zv = nullptr;

  double *Psig = c->d_partials, *Prho = c->d_partials + PSFUN_MAX_GRID,
         *Pconv = c->d_partials + 2 * PSFUN_MAX_GRID, *Pani = c->d_partials + 4 * PSFUN_MAX_GRID,
         *Pb = c->d_partials + 7 * PSFUN_MAX_GRID;  // two lists of Gu (no BLAS-1 call runs inside)
  const int Gs = psfun_grid(c, psfun_spmv_chunks(A), psfun_row_per_thread(A) ? 8 : 4);
  // the update grid: CG_UB elements per thread where n allows (k_cg_b)
  const int Gu = std::min(psfun_grid(c, (n + CG_UB * PSFUN_BS - 1) / (CG_UB * PSFUN_BS), 4), PSFUN_MAX_GRID / 2);

  hipLaunchKernelGGL(k_cg_rowstats, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, A->rowptr, A->colind, A->val, b,
                     dinv, Pani, Pb);
  PSFUN_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_cg_initconv, dim3(1), dim3(PSFUN_BS), 0, c->stream, Pani, Pb, Gu, ctl, 1);
  PSFUN_HIP_TRY(hipGetLastError());

  CgCtl *hctl = (CgCtl *)((char *)c->h_pinned + 256);
  int itx_host = 0;
  int done = CG_RUN;
  for (int restart = 0;; ++restart) {
    if (itx_host >= itmax) break;  // psfun_dcg.F90:256
    // r = b - (eta A + zeta I) x, rho partials
    PSFUN_DISPATCH_STORAGE(A, hipLaunchKernelGGL((k_cg_resid<LL, SS>), dim3(Gs), dim3(PSFUN_BS), 0, c->stream, n,
                                                 MRP, MCOL, MVAL, eta, zeta, b, x, r, dinv, Prho, zdiag));
    PSFUN_HIP_TRY(hipGetLastError());
    if (restart > 0) {
      hipLaunchKernelGGL(k_cg_initconv, dim3(1), dim3(PSFUN_BS), 0, c->stream, Pani, Pb, Gu, ctl, 0);
      PSFUN_HIP_TRY(hipGetLastError());
    }
    if (zmode) {  // z = M^{-1} r and rho = r.z partials over Gu blocks
      PSFUN_TRY(psfun_prec_launch(c, P, r, zv));
      hipLaunchKernelGGL(k_cg_rz, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, r, zv, Prho, ctl);
      PSFUN_HIP_TRY(hipGetLastError());
    }
    c->h_flags[0] = 0;
    c->h_flags[1] = 0;
    int it;
    double *pold = pA, *pnew = pB;
    for (it = 1; it <= itmax - itx_host + 1; ++it) {
      // algorithmic bytes: val+colind, rowptr, z (and p_old) read, p_new and q written
      PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_CG, 12.0 * A->nnz + 4.0 * (n + 1) + (it > 1 ? 32.0 : 24.0) * n));
      PSFUN_DISPATCH_STORAGE(A, hipLaunchKernelGGL((k_cg_a<LL, SS>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n,
                                                   MRP, MCOL, MVAL, eta, zeta, r, zv ? nullptr : dinv, pold, pnew, q,
                                                   it, Gs, Prho,
                                                   (it == 1 && !zmode) ? Gs : Gu, Pconv, Gu, Psig, ctl, istop, eps,
                                                   itmax, c->d_flags, zv));
      PSFUN_HIP_TRY(hipGetLastError());
      PSFUN_TRY(psfun_prof_end(c));
      hipLaunchKernelGGL(k_cg_b, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, x, r, pnew, q, dinv, it, Psig, Gs, Prho,
                         Pconv, ctl, istop, c->d_flags, zdiag);
      PSFUN_HIP_TRY(hipGetLastError());
      if (zmode) {  // the next iteration's z = M^{-1} r and rho = r.z
        PSFUN_TRY(psfun_prec_launch(c, P, r, zv));
        hipLaunchKernelGGL(k_cg_rz, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, r, zv, Prho, ctl);
        PSFUN_HIP_TRY(hipGetLastError());
      }
      std::swap(pold, pnew);
      // pacing on the host flags: about two iterations queued ahead of the
      // one whose control block has reported
      if (it >= 3) {
        PSFUN_TRY(psfun_host_wait(c, it - 2));
        if (c->h_flags[1] != 0) break;
      }
    }
    hipLaunchKernelGGL(k_cg_lastcheck, dim3(1), dim3(PSFUN_BS), 0, c->stream, Pconv, Gu, ctl, istop, eps, itmax);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_HIP_TRY(hipMemcpyAsync(hctl, ctl, sizeof(CgCtl), hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    itx_host = hctl->itx;
    done = hctl->done;
    if (done != CG_BRK) break;
  }
  if (iter) *iter = itx_host;
  if (err) *err = (hctl->errden == 0.0) ? hctl->errnum : hctl->errnum / hctl->errden;
  if (cond) {
    // Lanczos tridiagonal of CG (psfun_dcg.F90:314-322): td(1) = 1/a1,
    // td(i) = 1/a_i + b_i/a_{i-1}, tu(i-1) = sqrt(b_i)/a_{i-1}.
    const int m = std::min(itx_host, (int)CG_HIST_MAX);
    if (m > 0) {
      std::vector<double> td(m), tu(m > 1 ? m - 1 : 0);
      for (int i = 0; i < m; ++i) {
        const double a = hctl->alpha_hist[i];
        td[i] = (i == 0) ? 1.0 / a : 1.0 / a + hctl->beta_hist[i] / hctl->alpha_hist[i - 1];
        if (i > 0) tu[i - 1] = sqrt(hctl->beta_hist[i]) / hctl->alpha_hist[i - 1];
      }
      *cond = tridiag_extreme(td, tu, true) / tridiag_extreme(td, tu, false);
    }
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_cg_shifted(psfun_ctx_t c, psfun_csr_t A, int prec, const double *b, double eta, double zeta,
                                double *x, double eps, int itmax, int istop, int *iter, double *err, double *cond) {
  if (!c || !A) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!b || !x) return PSFUN_ERR_INVALID_VECT_STATE;
  if (prec != PSFUN_PREC_NONE && prec != PSFUN_PREC_DIAG && prec != PSFUN_PREC_BJAC) return PSFUN_ERR_INVALID_ARG;
  if (itmax < 0) return PSFUN_ERR_INVALID_ARG;
  if (prec != PSFUN_PREC_BJAC || A->n == 0)
    return cg_run(c, A, prec, nullptr, b, eta, zeta, x, eps, itmax, istop, iter, err, cond);
  // BJAC without a prebuilt object: build it for this call
  psfun_prec_t P = nullptr;
  PSFUN_TRY(psfun_prec_create(c, A, PSFUN_PREC_BJAC, &P));
  const int st = cg_run(c, A, prec, P, b, eta, zeta, x, eps, itmax, istop, iter, err, cond);
  psfun_prec_destroy(P);
  return st;
}

extern "C" int psfun_cg_shifted_prec(psfun_ctx_t c, psfun_csr_t A, psfun_prec_t P, const double *b, double eta,
                                     double zeta, double *x, double eps, int itmax, int istop, int *iter, double *err,
                                     double *cond) {
  if (!c || !A || !P) return PSFUN_ERR_INVALID_ARG;
  if (psfun_prec_size(P) != A->n) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!b || !x) return PSFUN_ERR_INVALID_VECT_STATE;
  if (itmax < 0) return PSFUN_ERR_INVALID_ARG;
  return cg_run(c, A, psfun_prec_kind(P), P, b, eta, zeta, x, eps, itmax, istop, iter, err, cond);
}
