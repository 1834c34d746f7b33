// Device-resident Lanczos f(A)x with the reference's semantics
// (krylov/polynomial/psfun_d_lanczos.f90:183-377).
//
// Reference loop (i = 2..nl):  itx = i-1; beta = ||w|| -> T(i-1,i) = T(i,i-1);
// exit if |beta| < 1e-16; v_i = (1/beta) w; w = A v_i; T(i,i) = w.v_i;
// w -= T(i,i) v_i; w -= beta v_{i-1}; if istop == 1: yk = f(T(1:i-1,1:i-1)) e1,
// rs(i) = rs(1) beta |yk(i-1)| / ||yk||, exit if rs(i) < eps.
// Assembly: y = rs(1) (yk(1) v_1 + sum_{l=2}^{itx-1} yk(l) v_l).
//
// Device schedule per Lanczos index j (j = i-1):
//   k_lz_update(j):  w = u - alpha_j v_j - beta_{j-1} v_{j-1},  beta_j = ||w||,
//                    then (last block) itx = j, breakdown test, and the
//                    a-posteriori estimate rs(j+1) from T_j in one wavefront.
//   k_lz_spmv(j+1):  v_{j+1} = (1/beta_j) w (written once), u = opscale*A v_{j+1}
//                    (x gathered from w and scaled on the fly),
//                    alpha_{j+1} = u.v_{j+1}.
// The reference's SpMV at the exit index feeds nothing it returns, so it is
// skipped.  Every kernel returns at once when ctl->done is set; the host only
// polls that flag every few iterations, one chunk behind the GPU.
#include "common.hpp"
#include "kernels.hpp"
#include "fexp.hpp"
#include "lanczos.hpp"

template <int L>
__global__ __launch_bounds__(PSFUN_BS) void k_lz_spmv(int n, const int *__restrict__ rowptr,
                                                      const int *__restrict__ col,
                                                      const double *__restrict__ val, double opscale,
                                                      const double *__restrict__ src, const double *sigma,
                                                      double *__restrict__ v_out, double *__restrict__ u_out,
                                                      double *partials, unsigned *counter, double *alpha_out,
                                                      LzCtl *ctl) {
  if (ctl->done) return;
  const double sg = *sigma;
  if (sg == 0.0) {  // zero start vector: y = 0
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = LZ_DONE_ZERO;
    return;
  }
  const double inv = 1.0 / sg;
  constexpr int RPB = PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range(nchunks);
  const int sub = threadIdx.x / L, lane = threadIdx.x % L;
  double dot = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    double sum = 0.0;
    if (row < n) {
      const int s = rowptr[row], e = rowptr[row + 1];
      for (int j = s + lane; j < e; j += L) sum = fma(val[j], src[col[j]] * inv, sum);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
    if (row < n && lane == 0) {
      const double v = src[row] * inv;
      const double u = opscale * sum;
      v_out[row] = v;
      u_out[row] = u;
      dot = fma(u, v, dot);
    }
  }
  double mine[1] = {dot}, tot[1];
  if (grid_reduce_last<PSFUN_BS, 1>(mine, partials, counter, tot)) {
    if (threadIdx.x == 0) *alpha_out = tot[0];
  }
}

__global__ __launch_bounds__(PSFUN_BS) void k_lz_update(int n, double *__restrict__ u,
                                                        const double *__restrict__ vj,
                                                        const double *__restrict__ vjm1, int j, int nl,
                                                        int istop, double eps, double fscale,
                                                        double *partials, unsigned *counter, LzCtl *ctl) {
  if (ctl->done) return;
  const double al = ctl->alpha[j];
  const double be = (j > 1) ? ctl->beta[j - 1] : 0.0;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    double w = fma(-al, vj[i], u[i]);
    if (j > 1) w = fma(-be, vjm1[i], w);
    u[i] = w;
    acc = fma(w, w, acc);
  }
  double mine[1] = {acc}, tot[1];
  if (!grid_reduce_last<PSFUN_BS, 1>(mine, partials, counter, tot)) return;
  if (threadIdx.x >= 64) return;
  // ---- one wavefront: loop control of the reference at index i = j+1 ----
  __shared__ double coef[PSFUN_CHEB_CAP];
  const double beta = sqrt(tot[0]);
  const int lane = threadIdx.x;
  if (lane == 0) {
    ctl->beta[j] = beta;
    ctl->itx = j;
  }
  if (fabs(beta) < 1e-16) {  // psfun_d_lanczos.f90:225-230
    if (lane == 0) ctl->done = LZ_DONE_BREAKDOWN;
    return;
  }
  if (istop == 1) {
    // yk = f(T(1:j,1:j)) e1; rs(j+1) = rs1 * beta * |yk(j)| / ||yk||
    double nrm, last;
    wave_exp_tridiag_e1_any(j, ctl->alpha + 1, ctl->beta + 1, fscale, coef, ctl->yk, nrm, last);
    const double errnum = ctl->rs1 * beta * fabs(last);
    const double rsj = errnum / nrm;
    if (lane == 0) {
      ctl->errnum = errnum;
      ctl->errden = nrm;
      ctl->rs[j + 1] = rsj;
      if (rsj < eps) {
        ctl->done = LZ_DONE_CONVERGED;
        ctl->m = j;  // yk holds f(T_j) e1
      }
    }
    if (rsj < eps) return;
  }
  if (j >= nl - 1 && lane == 0) ctl->done = LZ_DONE_MAXIT;
}

// Coefficients for the assembly when the loop did not end by convergence:
// yk = f(T(1:itx-1,1:itx-1)) e1 (psfun_d_lanczos.f90:330-362).  The reference
// leaves itx-1 = 0 undefined; here the 1x1 block T(1,1) is used.
__global__ __launch_bounds__(64) void k_lz_final(LzCtl *ctl, double fscale) {
  __shared__ double coef[PSFUN_CHEB_CAP];
  const int done = ctl->done;
  if (done == LZ_DONE_CONVERGED || done == LZ_DONE_ZERO) return;
  int k = ctl->itx - 1;
  if (k < 1) k = 1;
  double nrm, last;
  wave_exp_tridiag_e1_any(k, ctl->alpha + 1, ctl->beta + 1, fscale, coef, ctl->yk, nrm, last);
  if (threadIdx.x == 0) ctl->m = k;
}

// y = rs1 * (yk(1) v_1 + sum_{l=2}^{m'} yk(l) v_l), m' = max(1, m-1) after
// convergence (the last coefficient is dropped, psfun_d_lanczos.f90:373-377),
// m' = m otherwise.
__global__ __launch_bounds__(PSFUN_BS) void k_lz_assemble(int n, const double *__restrict__ V, int64_t ldv,
                                                          const LzCtl *ctl, double *__restrict__ y) {
  const int done = ctl->done;
  int m = ctl->m;
  if (done == LZ_DONE_CONVERGED) m = (m - 1 < 1) ? 1 : m - 1;
  if (done == LZ_DONE_ZERO) m = 0;
  const double rs1 = ctl->rs1;
  __shared__ double yk[LZ_NL_MAX];
  for (int l = threadIdx.x; l < m; l += PSFUN_BS) yk[l] = ctl->yk[l];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    double acc = 0.0;
    if (m > 0) {
      acc = yk[0] * V[i];
      for (int l = 1; l < m; ++l) acc = fma(yk[l], V[(int64_t)l * ldv + i], acc);
    }
    y[i] = rs1 * acc;
  }
}

__global__ void k_lz_init(LzCtl *ctl) {
  if (threadIdx.x == 0) {
    ctl->done = 0;
    ctl->itx = 0;
    ctl->m = 0;
    ctl->errnum = 0.0;
    ctl->errden = 1.0;
    ctl->beta[0] = 0.0;
  }
  for (int i = threadIdx.x; i < LZ_NL_MAX + 2; i += blockDim.x) ctl->rs[i] = 0.0;
}

// ---------------------------------------------------------------------------
int lanczos_enqueue_step_spmv(psfun_ctx *c, const psfun_csr *A, double opscale, const double *src,
                              const double *sigma, double *v_out, double *u_out, double *alpha_out, LzCtl *ctl) {
  const int L = A->lanes;
  const int64_t chunks = ((int64_t)A->n + PSFUN_BS / L - 1) / (PSFUN_BS / L);
  const int G = psfun_grid(c, chunks);
  double *part = c->d_partials;
  unsigned *cnt = c->d_counters + CNT_LZ_SPMV;
  // algorithmic bytes: val+colind, rowptr, src rows, v and u written
  const double bytes = 12.0 * A->nnz + 4.0 * (A->n + 1) + 24.0 * A->n;
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_LZ_SPMV, bytes));
#define LZ_LAUNCH(LL)                                                                                  \
  hipLaunchKernelGGL((k_lz_spmv<LL>), dim3(G), dim3(PSFUN_BS), 0, c->stream, A->n, A->rowptr, A->colind, \
                     A->val, opscale, src, sigma, v_out, u_out, part, cnt, alpha_out, ctl)
  switch (L) {
    case 1: LZ_LAUNCH(1); break;
    case 2: LZ_LAUNCH(2); break;
    case 4: LZ_LAUNCH(4); break;
    case 8: LZ_LAUNCH(8); break;
    case 16: LZ_LAUNCH(16); break;
    case 32: LZ_LAUNCH(32); break;
    default: LZ_LAUNCH(64); break;
  }
#undef LZ_LAUNCH
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));
  return PSFUN_SUCCESS;
}

extern "C" int psfun_lanczos_apply(psfun_ctx_t c, psfun_csr_t A, int fname, double fscale, double opscale,
                                   const double *x, double *y, double eps, int itmax, int istop, int *iter,
                                   double *err, double *res, int *nres) {
  if (!c || !A) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  if (fname != PSFUN_FN_EXP) return PSFUN_ERR_UNSUPPORTED;
  if (itmax < 2 || itmax > LZ_NL_MAX) return PSFUN_ERR_INVALID_ARG;
  const int n = A->n;
  const int nl = itmax;
  if (n == 0) {
    if (iter) *iter = 0;
    if (err) *err = 0.0;
    if (nres) *nres = 0;
    return PSFUN_SUCCESS;
  }
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  const int64_t ldv = (int64_t)align_up((size_t)n, 64);
  const size_t ctl_bytes = align_up(sizeof(LzCtl), 256);
  const size_t vbytes = sizeof(double) * (size_t)ldv * (size_t)(nl + 2);  // v_1..v_{nl-1}, u0, u1
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, ctl_bytes + vbytes, &ws));
  LzCtl *ctl = (LzCtl *)ws;
  double *V = (double *)((char *)ws + ctl_bytes);
  double *U[2] = {V + (size_t)ldv * nl, V + (size_t)ldv * (nl + 1)};

  hipLaunchKernelGGL(k_lz_init, dim3(1), dim3(256), 0, c->stream, ctl);
  PSFUN_HIP_TRY(hipGetLastError());
  // rs(1) = ||x||; v_1 = (1/rs1) x; u = A v_1; T(1,1) = u.v_1
  PSFUN_TRY(launch_reduce(c, 1, n, x, nullptr, &ctl->rs1));
  PSFUN_TRY(lanczos_enqueue_step_spmv(c, A, opscale, x, &ctl->rs1, V, U[0], &ctl->alpha[1], ctl));

  const int G = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  volatile int *hflag = (volatile int *)c->h_pinned;
  const int poll = 4;
  int pending = -1;  // slot with an outstanding flag copy
  for (int j = 1; j <= nl - 1; ++j) {
    double *u = U[(j - 1) & 1];
    const double *vj = V + (size_t)ldv * (j - 1);
    const double *vjm1 = (j > 1) ? V + (size_t)ldv * (j - 2) : nullptr;
    // algorithmic bytes: u read + w written, v_j (and v_{j-1}) read
    PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_LZ_UPDATE, (j > 1 ? 32.0 : 24.0) * n));
    hipLaunchKernelGGL(k_lz_update, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, u, vj, vjm1, j, nl, istop, eps,
                       fscale, c->d_partials, c->d_counters + CNT_LZ_UPD, ctl);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(psfun_prof_end(c));
    if (j < nl - 1) {
      PSFUN_TRY(lanczos_enqueue_step_spmv(c, A, opscale, u, &ctl->beta[j], V + (size_t)ldv * j, U[j & 1],
                                          &ctl->alpha[j + 1], ctl));
    }
    if (j % poll == 0 && j < nl - 1) {
      const int slot = (j / poll) & 1;
      PSFUN_HIP_TRY(hipMemcpyAsync((void *)(hflag + slot), &ctl->done, sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
      PSFUN_HIP_TRY(hipEventRecord(c->ev[slot], c->stream));
      if (pending >= 0) {
        PSFUN_HIP_TRY(hipEventSynchronize(c->ev[pending]));
        if (hflag[pending] != 0) {
          pending = slot;
          break;
        }
      }
      pending = slot;
    }
  }
  hipLaunchKernelGGL(k_lz_final, dim3(1), dim3(64), 0, c->stream, ctl, fscale);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_LZ_ASSEMBLE, 0.0));
  hipLaunchKernelGGL(k_lz_assemble, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, V, ldv, ctl, y);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));

  // results back to the host
  LzCtl *h = (LzCtl *)((char *)c->h_pinned + 256);
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(LzCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  const int itx = h->itx;
  const double e = (h->errden == 0.0) ? h->errnum : h->errnum / h->errden;
  if (iter) *iter = itx;
  if (err) *err = e;
  if (res) {
    for (int i = 1; i <= itx; ++i) res[i - 1] = (i == 1) ? h->rs1 : h->rs[i];
    res[itx] = e;
  }
  if (nres) *nres = itx + 1;
  return PSFUN_SUCCESS;
}
