// Multi-shift CG for (A + sigma_s I) x_s = b, s = 1..S, complex shifts,
// real SPD A, all shifts in one Krylov space and one launch sequence, and the
// quadrature apply built on it.  No reference counterpart: the reference's
// krylov/shifted/psfun_dcg.F90 solves one real shift, and the quadrature
// module (quadratures/psfun_z_quadrature_mod.F90) only stores poles/weights.
//
// Seed: real CG on A with x0 = 0 (r_0 = p_0 = b).  Shifted iterates follow
// the collinear-residual recurrences r_k^s = zeta_k^s r_k (Jegerlehner 1996;
// Frommer 2003):
//   zeta_{k+1} = zeta_k zeta_{k-1} a_{k-1} /
//                (a_k b_{k-1} (zeta_{k-1} - zeta_k) + zeta_{k-1} a_{k-1} (1 + s a_k))
//   alpha_k^s = a_k zeta_{k+1}/zeta_k,  beta_k^s = b_k (zeta_{k+1}/zeta_k)^2
//   x^s += alpha_k^s p^s,  p^s = zeta_{k+1} r_{k+1} + beta_k^s p^s
// with a_{-1} = 1, b_{-1} = 0, zeta_{-1} = zeta_0 = 1.  Shift s stops when
// |zeta_{k+1}^s| ||r_{k+1}|| <= eps ||b||; the loop ends when every shift has
// stopped or after itmax seed iterations.
//
// Per seed iteration k (all scalars are folded from per-block partials by
// every consumer block in the same order, so they are bit-identical
// everywhere; per-shift state is double-buffered by the parity of k):
//   k_ms_spmv(k):  b_{k-1} = rr_k/rr_{k-1}; p_k = r_k + b_{k-1} p_{k-1} (own rows
//                  written, neighbours formed on the fly); q = A p_k; p.q partials.
//   k_ms_step(k):  a_k = rr_k / (p.q); r_{k+1} = r_k - a_k q and rr_{k+1} partials;
//                  the stop test of iteration k-1 for every shift (it needs
//                  ||r_k||); for the shifts still running p^s_k = zeta_k r_k +
//                  beta^s_{k-1} p^s_{k-1}, zeta_{k+1}, alpha^s_k and x^s +=
//                  alpha^s_k p^s_k.  The control block commits the new shift
//                  state and the stop flag.
//   k_ms_finish:   after a loop that reached itmax, the last test's bookkeeping.
// Complex vectors are stored shift-major, n (re, im) pairs per shift.
#include "common.hpp"
#include "kernels.hpp"

#define MS_MAX 64

enum { MS_RUN = 0, MS_DONE = 1 };

struct MsState {
  double zc[MS_MAX][2];  // zeta_k
  double zp[MS_MAX][2];  // zeta_{k-1}
  int active[MS_MAX];
};

struct MsCtl {
  int done;
  int k;
  int nshift;
  int pad;
  double bnorm;
  double rr[2];   // rr_k at [k & 1]
  double a[2];    // a_k at [k & 1]
  double b[2];    // b_k at [k & 1]
  double sig[MS_MAX][2];
  MsState st[2];  // state at the start of iteration k: st[k & 1]
  int iters[MS_MAX];
  double resid[MS_MAX];
};

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

__global__ __launch_bounds__(PSFUN_BS) void k_ms_init(int64_t n, const double *__restrict__ b,
                                                      double *__restrict__ r, double *__restrict__ p_rr) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    const double v = b[i];
    r[i] = v;
    acc = fma(v, v, acc);
  }
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_rr[blockIdx.x] = acc;
}

__global__ void k_ms_ctl_init(MsCtl *ctl, int nshift, const double *__restrict__ shifts) {
  const int s = threadIdx.x;
  if (s == 0) {
    ctl->done = MS_RUN;
    ctl->k = 0;
    ctl->nshift = nshift;
    ctl->a[1] = 1.0;  // a_{-1}
    ctl->b[1] = 0.0;  // b_{-1}
  }
  if (s < MS_MAX) {
    const bool on = s < nshift;
    ctl->st[0].zc[s][0] = 1.0;
    ctl->st[0].zc[s][1] = 0.0;
    ctl->st[0].zp[s][0] = 1.0;
    ctl->st[0].zp[s][1] = 0.0;
    ctl->st[0].active[s] = on ? 1 : 0;
    ctl->iters[s] = 0;
    ctl->resid[s] = on ? 1.0 : 0.0;
    ctl->sig[s][0] = on ? shifts[2 * s] : 0.0;
    ctl->sig[s][1] = on ? shifts[2 * s + 1] : 0.0;
  }
}

template <int L, bool SELL>
__global__ __launch_bounds__(PSFUN_BS) void k_ms_spmv(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                      const double *__restrict__ val, const double *__restrict__ r,
                                                      const double *__restrict__ p_old, double *__restrict__ p_new,
                                                      double *__restrict__ q, int k, int G,
                                                      const double *__restrict__ p_rr, int g_rr,
                                                      double *__restrict__ p_pq, MsCtl *ctl) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  const double rr = block_sum_partials<PSFUN_BS>(p_rr, g_rr, lds);
  if (blockIdx.x == G) {
    if (threadIdx.x == 0) {
      ctl->rr[k & 1] = rr;
      ctl->k = k;
      if (k == 0) ctl->bnorm = sqrt(rr);
    }
    return;
  }
  const double bk = (k > 0) ? rr / ctl->rr[(k - 1) & 1] : 0.0;
  // seed direction p = r + bk p_old, formed on the fly at every gathered column
  auto pk = [&](int c) -> double { return (k > 0) ? fma(bk, p_old[c], r[c]) : r[c]; };
  constexpr int RPB = SELL ? PSFUN_BS : PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range_g(nchunks, G);
  const int sub = SELL ? threadIdx.x : threadIdx.x / L, lane = SELL ? 0 : threadIdx.x % L;
  double pq = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    const double sum = row_dot<L, SELL>(row, n, rowptr, col, val, lane, pk);
    if (row < n && lane == 0) {
      const double pr = pk(row);
      p_new[row] = pr;
      q[row] = sum;
      pq = fma(pr, sum, pq);
    }
  }
  pq = block_sum<PSFUN_BS>(pq, lds);
  if (threadIdx.x == 0) p_pq[blockIdx.x] = pq;
}

// One seed iteration after its SpMV, in one pass over the rows:
//   a_k = rr_k / (p_k.q_k);  r_{k+1} = r_k - a_k q_k and the rr_{k+1} partials;
//   for every shift still running (the test of iteration k-1 below):
//     p^s_k = zeta_k r_k + beta^s_{k-1} p^s_{k-1}   (p^s_0 = b, never stored)
//     zeta_{k+1}, alpha^s_k = a_k zeta_{k+1}/zeta_k;  x^s += alpha^s_k p^s_k
// The test of iteration k-1 (|zeta_k| ||r_k|| <= eps ||b||, or the iteration
// limit) needs ||r_k||, which the launch before produced, so each shift's
// direction update comes one launch after the classical place (at the step
// that uses it); every value is the same formula of the same inputs, so the
// iterates do not change.  A shift that stops contributes its last x^s
// update at the step it stops after, as its x^s would.
// ACC: the x^s are not kept; acc += sum_s Re(wq_s alpha^s p^s) (form 0) or
// Im(...) (form 1), the quadrature's combination accumulated one iteration
// at a time (acc is not read at k = 0).
template <bool ACC>
__global__ __launch_bounds__(PSFUN_BS) void k_ms_step(int n, double *__restrict__ r, const double *__restrict__ q,
                                                      const double *__restrict__ b, double2 *__restrict__ X,
                                                      int64_t ldx, double *__restrict__ acc,
                                                      const double *__restrict__ wq, int form,
                                                      double2 *__restrict__ P, int64_t ldp, int k, int G,
                                                      const double *__restrict__ p_pq, int g_pq,
                                                      double *__restrict__ p_rr, MsCtl *ctl, double eps, int itmax,
                                                      int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  __shared__ double s_alpha[MS_MAX][2], s_beta[MS_MAX][2], s_zc[MS_MAX][2], s_w[MS_MAX][2];
  __shared__ int s_x[MS_MAX];
  const int t = threadIdx.x;
  if (ACC && t < MS_MAX) {
    s_w[t][0] = t < ctl->nshift ? wq[2 * t] : 0.0;
    s_w[t][1] = t < ctl->nshift ? wq[2 * t + 1] : 0.0;
  }
  // block-uniform: the control block of this same grid may set the flag
  if (__syncthreads_or(ctl->done)) return;
  const double pq = block_sum_partials<PSFUN_BS>(p_pq, g_pq, lds);
  const int nshift = ctl->nshift;
  const int par = k & 1;
  const MsState *st = &ctl->st[par];  // zc = zeta_k, zp = zeta_{k-1}; active: updated at k-1 (running at k = 0)
  const double rrk = ctl->rr[par];
  const double ak = rrk / pq;
  const double akm1 = ctl->a[par ^ 1];                        // a_{k-1} (a_{-1} = 1)
  const double bkm1 = (k > 0) ? rrk / ctl->rr[par ^ 1] : 0.0;  // b_{k-1} (b_{-1} = 0)
  const double bnorm = ctl->bnorm;
  const double rnorm = sqrt(rrk);
  if (t < MS_MAX) {
    int xk = 0;
    cplx zn = {0.0, 0.0}, al = {0.0, 0.0}, be = {0.0, 0.0};
    const cplx zc = {st->zc[t][0], st->zc[t][1]}, zp = {st->zp[t][0], st->zp[t][1]};
    bool tested = false;
    double res = 0.0;
    if (t < nshift && st->active[t]) {
      if (!(bnorm > 0.0)) {
        xk = (k == 0);  // b = 0: x^s = 0 (alpha^s = 0 on the zero start), no iteration
      } else {
        bool run = true;
        if (k > 0) {  // the test of iteration k-1
          res = hypot(zc.re, zc.im) * rnorm;
          tested = true;
          run = !(res <= eps * bnorm || k >= itmax);
          if (run) {  // beta^s_{k-1} = b_{k-1} (zeta_k / zeta_{k-1})^2
            const cplx ratio = cdiv(zc, zp);
            const cplx r2 = cmul(ratio, ratio);
            be = {bkm1 * r2.re, bkm1 * r2.im};
          }
        }
        if (run) {
          // zeta_{k+1}: denominator a_k b_{k-1} (zp - zc) + zp a_{k-1} (1 + s a_k)
          const cplx sg = {ctl->sig[t][0], ctl->sig[t][1]};
          const cplx d1 = {ak * bkm1 * (zp.re - zc.re), ak * bkm1 * (zp.im - zc.im)};
          const cplx one_sa = {1.0 + sg.re * ak, sg.im * ak};
          const cplx d2 = cmul(zp, one_sa);
          const cplx den = {d1.re + akm1 * d2.re, d1.im + akm1 * d2.im};
          const cplx num0 = cmul(zc, zp);
          const cplx num = {num0.re * akm1, num0.im * akm1};
          zn = cdiv(num, den);
          const cplx ratio = cdiv(zn, zc);
          al = {ak * ratio.re, ak * ratio.im};
          xk = 1;
        }
      }
    }
    s_alpha[t][0] = al.re;
    s_alpha[t][1] = al.im;
    s_beta[t][0] = be.re;
    s_beta[t][1] = be.im;
    s_zc[t][0] = zc.re;
    s_zc[t][1] = zc.im;
    s_x[t] = xk;
    if (blockIdx.x == G) {  // control block: commit the state for iteration k+1
      MsState *nx = &ctl->st[par ^ 1];
      if (tested) {  // the bookkeeping of iteration k-1
        ctl->iters[t] = k;
        ctl->resid[t] = res / bnorm;
      } else if (xk && !(bnorm > 0.0)) {
        ctl->iters[t] = 0;
        ctl->resid[t] = 0.0;
      }
      if (xk && bnorm > 0.0) {
        nx->zp[t][0] = zc.re;
        nx->zp[t][1] = zc.im;
        nx->zc[t][0] = zn.re;
        nx->zc[t][1] = zn.im;
      } else {
        nx->zp[t][0] = zp.re;
        nx->zp[t][1] = zp.im;
        nx->zc[t][0] = zc.re;
        nx->zc[t][1] = zc.im;
      }
      nx->active[t] = (xk && bnorm > 0.0) ? 1 : 0;
    }
  }
  __syncthreads();
  int any = 0;
  for (int s = 0; s < nshift; ++s) any |= s_x[s];
  if (blockIdx.x == G) {
    if (t == 0) {
      ctl->a[par] = ak;
      host_flag_store(hf, k + 1);
      // nothing runs at k, or b = 0 (its only step is this one)
      if (!any || !(bnorm > 0.0)) {
        ctl->done = MS_DONE;
        host_flag_store(hf + 1, MS_DONE);
      }
    }
    return;
  }
  if (!any) return;
  double arr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + t; i < n; i += (int64_t)G * PSFUN_BS) {
    const double ri = r[i];
    const double rn = fma(-ak, q[i], ri);
    r[i] = rn;
    arr = fma(rn, rn, arr);
    const double bi = (k <= 1) ? b[i] : 0.0;
    double a = (ACC && k > 0) ? acc[i] : 0.0;
    for (int s = 0; s < nshift; ++s) {
      if (!s_x[s]) continue;
      double2 pv;
      if (k == 0) {
        pv = make_double2(bi, 0.0);
      } else {
        double2 *pp = P + (int64_t)s * ldp + i;
        const double2 po = (k == 1) ? make_double2(bi, 0.0) : *pp;
        const double br = s_beta[s][0], bim = s_beta[s][1];
        pv = make_double2(fma(s_zc[s][0], ri, br * po.x - bim * po.y), fma(s_zc[s][1], ri, br * po.y + bim * po.x));
        *pp = pv;
      }
      const double ar = s_alpha[s][0], ai = s_alpha[s][1];
      const double dxr = ar * pv.x - ai * pv.y, dxi = ar * pv.y + ai * pv.x;  // alpha^s p^s
      if (ACC) {
        a += (form == 0) ? (s_w[s][0] * dxr - s_w[s][1] * dxi) : (s_w[s][0] * dxi + s_w[s][1] * dxr);
      } else {
        double2 *xp = X + (int64_t)s * ldx + i;
        const double2 xv = (k == 0) ? make_double2(0.0, 0.0) : *xp;
        *xp = make_double2(xv.x + dxr, xv.y + dxi);
      }
    }
    if (ACC) acc[i] = a;
  }
  arr = block_sum<PSFUN_BS>(arr, lds);
  if (t == 0) p_rr[blockIdx.x] = arr;
}

// The bookkeeping of the last iteration after the loop ran to itmax: the
// shifts that were updated at k = itmax-1 get iters = itmax and their
// residual |zeta_itmax| ||r_itmax|| / ||b||.
__global__ __launch_bounds__(PSFUN_BS) void k_ms_finish(int k, const double *__restrict__ p_rr, int g_rr,
                                                        MsCtl *ctl) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  const double rr = block_sum_partials<PSFUN_BS>(p_rr, g_rr, lds);
  const int t = threadIdx.x;
  const MsState *st = &ctl->st[k & 1];
  if (t < ctl->nshift && st->active[t]) {
    ctl->iters[t] = k;
    ctl->resid[t] = hypot(st->zc[t][0], st->zc[t][1]) * sqrt(rr) / ctl->bnorm;
  }
}

struct MsWork {
  MsCtl *ctl;
  double *r, *pa, *pb, *q, *dshift, *dw, *tmp;
  double2 *P, *X;
  int64_t ld;
};

// One workspace holds the control block, the seed vectors, the shifted
// directions and (optionally) the shifted solutions.
static int ms_workspace(psfun_ctx *c, int n, int nshift, bool with_x, MsWork *w) {
  const size_t ctl_bytes = align_up(sizeof(MsCtl), 256);
  const int64_t ld = (int64_t)align_up((size_t)n, 64);
  const size_t nsh = (size_t)std::max(nshift, 1);
  const size_t bytes = ctl_bytes + sizeof(double) * (5 * (size_t)ld + 4 * MS_MAX) +
                       sizeof(double2) * (size_t)ld * nsh * (with_x ? 2 : 1);
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, bytes, &ws));
  w->ctl = (MsCtl *)ws;
  w->ld = ld;
  w->r = (double *)((char *)ws + ctl_bytes);
  w->pa = w->r + ld;
  w->pb = w->pa + ld;
  w->q = w->pb + ld;
  w->tmp = w->q + ld;
  w->dshift = w->tmp + ld;
  w->dw = w->dshift + 2 * MS_MAX;
  w->P = (double2 *)(w->dw + 2 * MS_MAX);
  w->X = with_x ? w->P + (size_t)ld * nsh : nullptr;
  return PSFUN_SUCCESS;
}

// With acc != nullptr the solutions are not stored: acc receives the
// combination sum_s Re (form 0) or Im (form 1) of wq_s x^s (wq on the device).
static int ms_run(psfun_ctx *c, const psfun_csr *A, int nshift, const double *shifts_host, const double *b,
                  double2 *X, int64_t ldx, const MsWork &w, double eps, int itmax, int *iters, double *resid,
                  double *acc = nullptr, const double *wq = nullptr, int form = 0) {
  const int n = A->n;
  MsCtl *ctl = w.ctl;
  double *r = w.r, *q = w.q;
  double2 *P = w.P;
  PSFUN_HIP_TRY(hipMemcpyAsync(w.dshift, shifts_host, sizeof(double) * 2 * nshift, hipMemcpyHostToDevice,
                               c->stream));
  double *Ppq = c->d_partials, *Prr = c->d_partials + PSFUN_MAX_GRID;
  const int Gs = psfun_grid(c, psfun_spmv_chunks(A), psfun_row_per_thread(A) ? 8 : 4);
  const int Gu = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 4);

  c->h_flags[0] = 0;
  c->h_flags[1] = 0;
  hipLaunchKernelGGL(k_ms_ctl_init, dim3(1), dim3(MS_MAX), 0, c->stream, ctl, nshift, (const double *)w.dshift);
  hipLaunchKernelGGL(k_ms_init, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, (int64_t)n, b, r, Prr);
  PSFUN_HIP_TRY(hipGetLastError());
  double *pold = w.pa, *pnew = w.pb;
  for (int k = 0; k < itmax; ++k) {
    PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_MS_SPMV, 12.0 * A->nnz + 4.0 * (n + 1) + (k ? 32.0 : 24.0) * n, k));
    PSFUN_DISPATCH_STORAGE(A, hipLaunchKernelGGL((k_ms_spmv<LL, SS>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n,
                                                 MRP, MCOL, MVAL, r, pold, pnew, q, k, Gs, Prr, Gu, Ppq, ctl));
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(psfun_prof_end(c));
    // algorithmic bytes of the step kernel: r and q read, r written (the seed), p^s read+written
    // per running shift, and x^s read+written per shift, or the one real accumulator read+written
    PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_MS_UPDATE,
                               acc ? 24.0 * n + 32.0 * n * nshift + 16.0 * n : 24.0 * n + 64.0 * n * nshift, k));
    if (acc)
      hipLaunchKernelGGL(k_ms_step<true>, dim3(Gu + 1), dim3(PSFUN_BS), 0, c->stream, n, r, q, b, X, ldx, acc, wq,
                         form, P, w.ld, k, Gu, Ppq, Gs, Prr, ctl, eps, itmax, c->d_flags);
    else
      hipLaunchKernelGGL(k_ms_step<false>, dim3(Gu + 1), dim3(PSFUN_BS), 0, c->stream, n, r, q, b, X, ldx, acc, wq,
                         form, P, w.ld, k, Gu, Ppq, Gs, Prr, ctl, eps, itmax, c->d_flags);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(psfun_prof_end(c));
    std::swap(pold, pnew);
    // pacing on the host flags: keep about two iterations queued ahead of
    // the one whose shift kernel has reported
    if (k >= 2 && k + 1 < itmax) {
      PSFUN_TRY(psfun_host_wait(c, k - 1));
      if (c->h_flags[1] != 0) break;
    }
  }
  // a loop that ran to itmax still owes its last iteration's bookkeeping
  hipLaunchKernelGGL(k_ms_finish, dim3(1), dim3(PSFUN_BS), 0, c->stream, itmax, Prr, Gu, ctl);
  PSFUN_HIP_TRY(hipGetLastError());
  MsCtl *h = (MsCtl *)((char *)c->h_pinned + 256);
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(MsCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  // iterations after the one that stopped the loop (h->k) did no work
  psfun_prof_discard_after(c, PSFUN_PROF_MS_SPMV, h->k);
  psfun_prof_discard_after(c, PSFUN_PROF_MS_UPDATE, h->k);
  for (int s = 0; s < nshift; ++s) {
    if (iters) iters[s] = h->iters[s];
    if (resid) resid[s] = h->resid[s];
  }
  return PSFUN_SUCCESS;
}

int psfun_mscg_run(psfun_ctx *c, const psfun_csr *A, int nshift, const double *shifts, const double *b, double *X,
                   int64_t ldx, double eps, int itmax, int *iters, double *resid) {
  MsWork w;
  PSFUN_TRY(ms_workspace(c, A->n, nshift, false, &w));
  return ms_run(c, A, nshift, shifts, b, (double2 *)X, ldx, w, eps, itmax, iters, resid);
}

extern "C" int psfun_mscg(psfun_ctx_t c, psfun_csr_t A, int nshift, const double *shifts, const double *b,
                          double *X, int64_t ldx, double eps, int itmax, int *iters, double *resid) {
  if (!c || !A || nshift < 0 || nshift > MS_MAX || (nshift && !shifts)) return PSFUN_ERR_INVALID_ARG;
  if (!b || !X) return PSFUN_ERR_INVALID_VECT_STATE;
  if (itmax < 0 || ldx < A->n) return PSFUN_ERR_INVALID_ARG;
  if (A->n == 0 || nshift == 0) return PSFUN_SUCCESS;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  if (itmax == 0) {
    PSFUN_HIP_TRY(hipMemsetAsync(X, 0, sizeof(double2) * (size_t)ldx * nshift, c->stream));
    return psfun_ctx_synchronize(c);
  }
  MsWork w;
  PSFUN_TRY(ms_workspace(c, A->n, nshift, false, &w));
  return ms_run(c, A, nshift, shifts, b, (double2 *)X, ldx, w, eps, itmax, iters, resid);
}

extern "C" int psfun_quad_apply(psfun_ctx_t c, psfun_csr_t A, int nshift, const double *shifts,
                                const double *weights, double gamma, int form, const double *x, double *y,
                                double eps, int itmax, int *iters, double *resid) {
  if (!c || !A || nshift < 0 || nshift > MS_MAX || (nshift && (!shifts || !weights))) return PSFUN_ERR_INVALID_ARG;
  if (form != PSFUN_QUAD_RE_SUM && form != PSFUN_QUAD_A_IM_SUM) return PSFUN_ERR_INVALID_ARG;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  if (itmax < 1) return PSFUN_ERR_INVALID_ARG;
  const int n = A->n;
  if (n == 0) return PSFUN_SUCCESS;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  if (nshift == 0) {
    PSFUN_HIP_TRY(hipMemsetAsync(y, 0, sizeof(double) * n, c->stream));
    return psfun_ctx_synchronize(c);
  }
  // the solutions are not stored: the shift kernel accumulates
  // gamma * (Re or Im) sum_s w_s x^s itself, into y (form 0) or into a work
  // vector that one SpMV then maps to y (form 1)
  MsWork w;
  PSFUN_TRY(ms_workspace(c, n, nshift, false, &w));
  double *wg = (double *)c->h_pinned;  // gamma * w_s, staged in pinned memory
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  for (int s = 0; s < 2 * nshift; ++s) wg[s] = gamma * weights[s];
  PSFUN_HIP_TRY(hipMemcpyAsync(w.dw, wg, sizeof(double) * 2 * nshift, hipMemcpyHostToDevice, c->stream));
  double *acc = form == PSFUN_QUAD_RE_SUM ? y : w.tmp;
  PSFUN_TRY(ms_run(c, A, nshift, shifts, x, nullptr, 0, w, eps, itmax, iters, resid, acc, w.dw,
                   form == PSFUN_QUAD_RE_SUM ? 0 : 1));
  if (form == PSFUN_QUAD_A_IM_SUM) PSFUN_TRY(launch_spmv(c, A, 1.0, w.tmp, 0.0, y));
  return psfun_ctx_synchronize(c);
}

// grad[q] += scale * sum_s Re(w_s U_s[i] V_s[k]) for every nonzero q = (i, k)
// of A: the sampled (SDDMM) form of the derivative of
// Re sum_s w_s g^T (A + sigma_s I)^{-1} x with respect to the entries of A.
__global__ __launch_bounds__(PSFUN_BS) void k_sddmm_shifted(int n, const int *__restrict__ rowptr,
                                                            const int *__restrict__ col, int nshift,
                                                            const double *__restrict__ w, const double2 *__restrict__ U,
                                                            int64_t ldu, const double2 *__restrict__ V, int64_t ldv,
                                                            double scale, double *__restrict__ grad) {
  __shared__ double sw[MS_MAX][2];
  if (threadIdx.x < nshift) {
    sw[threadIdx.x][0] = w[2 * threadIdx.x];
    sw[threadIdx.x][1] = w[2 * threadIdx.x + 1];
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    const int s0 = rowptr[i], s1 = rowptr[i + 1];
    for (int q = s0; q < s1; ++q) {
      const int k = col[q];
      double acc = 0.0;
      for (int s = 0; s < nshift; ++s) {
        const double2 a = U[(int64_t)s * ldu + i], b = V[(int64_t)s * ldv + k];
        // w * a * b, real part
        const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
        acc += sw[s][0] * pr - sw[s][1] * pi;
      }
      grad[q] = fma(scale, acc, grad[q]);
    }
  }
}

extern "C" int psfun_sddmm_shifted(psfun_ctx_t c, int n, const int32_t *rowptr, const int32_t *colind, int nshift,
                                   const double *weights, const double *U, int64_t ldu, const double *V, int64_t ldv,
                                   double scale, double *grad) {
  if (!c || n < 0 || nshift < 0 || nshift > MS_MAX || (nshift && !weights)) return PSFUN_ERR_INVALID_ARG;
  if (n == 0 || nshift == 0) return PSFUN_SUCCESS;
  if (!rowptr || !colind || !U || !V || !grad || ldu < n || ldv < n) return PSFUN_ERR_INVALID_ARG;
  // the weight table (2*nshift <= 128 doubles) goes to the context's small parameter area
  PSFUN_HIP_TRY(hipMemcpyAsync(c->d_scalar, weights, sizeof(double) * 2 * nshift, hipMemcpyHostToDevice, c->stream));
  const int G = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_SDDMM, 0.0));
  hipLaunchKernelGGL(k_sddmm_shifted, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, rowptr, colind, nshift,
                     (const double *)c->d_scalar, (const double2 *)U, ldu, (const double2 *)V, ldv, scale, grad);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));
  return psfun_ctx_synchronize(c);
}
