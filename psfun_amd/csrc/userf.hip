// USERF/SYM f(H) step (serial/psfun_d_serial_mod.F90:420-485): device
// eigen-decomposition, host user function, device combination.
#include "userf.hpp"

// pinned host staging of the eigenvalues and of f at them
#define USERF_PINNED_OFFSET (512 << 10)

__global__ __launch_bounds__(PSFUN_BS) void k_sym_eig(int k, int form, const double *a, const double *b, int ld,
                                                      double *S, double *Q, double *w) {
  __shared__ SymLds L;
  block_sym_load(k, form, a, b, ld, S);
  block_sym_eig(k, S, Q, w, &L);
}

__global__ __launch_bounds__(PSFUN_BS) void k_sym_apply(int k, const double *Q, const double *fv, const double *x,
                                                        double *y) {
  __shared__ SymLds L;
  block_sym_apply(k, Q, fv, x, y, &L);
}

// Same stop bookkeeping as the EXP control steps (psfun_d_lanczos.f90:276-326,
// psfun_d_arnoldi.f90:215-257).
__global__ __launch_bounds__(PSFUN_BS) void k_userf_est(LzCtl *ctl, int j, const double *sig, const double *Q,
                                                        const double *fv, int nl, double eps, int *hf) {
  __shared__ SymLds L;
  if (ctl->done) return;
  const int k = j - 1;
  const double nrm = block_sym_apply(k, Q, fv, nullptr, ctl->yk, &L);
  if (threadIdx.x != 0) return;
  const double errnum = ctl->rs1 * (*sig) * fabs(ctl->yk[k - 1]);
  const double rsj = errnum / nrm;
  ctl->errnum = errnum;
  ctl->errden = nrm;
  ctl->rs[j] = rsj;
  int st = 0;
  if (rsj < eps) {
    ctl->m = k;
    st = LZ_DONE_CONVERGED;
  } else if (j >= nl) {
    st = LZ_DONE_MAXIT;
  }
  if (st) {
    ctl->done = st;
    host_flag_store(hf + 1, st);
  }
}

__global__ __launch_bounds__(PSFUN_BS) void k_userf_final(LzCtl *ctl, int k, const double *Q, const double *fv) {
  __shared__ SymLds L;
  block_sym_apply(k, Q, fv, nullptr, ctl->yk, &L);
  if (threadIdx.x == 0) ctl->m = k;
}

size_t userf_scratch_bytes(int kmax) {
  const size_t kk = align_up(sizeof(double) * (size_t)kmax * kmax, 256);
  const size_t kv = align_up(sizeof(double) * (size_t)kmax, 256);
  return 2 * kk + 2 * kv;
}

UserfScratch userf_scratch_carve(void *base, int kmax) {
  const size_t kk = align_up(sizeof(double) * (size_t)kmax * kmax, 256);
  const size_t kv = align_up(sizeof(double) * (size_t)kmax, 256);
  char *p = (char *)base;
  UserfScratch sc;
  sc.S = (double *)p;
  sc.Q = (double *)(p + kk);
  sc.w = (double *)(p + 2 * kk);
  sc.fv = (double *)(p + 2 * kk + kv);
  return sc;
}

int userf_check(const psfun_ctx *c, int variant) {
  if (variant != PSFUN_VAR_SYM) return PSFUN_ERR_UNSUPPORTED;
  if (!c->userf) return PSFUN_ERR_INVALID_ARG;
  return PSFUN_SUCCESS;
}

int userf_eval(psfun_ctx *c, int k, int form, const double *a, const double *b, int ld, double fscale,
               const UserfScratch &sc, bool *stopped) {
  if (k < 1 || k > PSFUN_DENSE_KMAX) return PSFUN_ERR_INVALID_ARG;
  hipLaunchKernelGGL(k_sym_eig, dim3(1), dim3(PSFUN_BS), 0, c->stream, k, form, a, b, ld, sc.S, sc.Q, sc.w);
  PSFUN_HIP_TRY(hipGetLastError());
  double *hw = (double *)((char *)c->h_pinned + USERF_PINNED_OFFSET);
  double *hf = hw + PSFUN_DENSE_KMAX;
  PSFUN_HIP_TRY(hipMemcpyAsync(hw, sc.w, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  if (stopped) {
    *stopped = c->h_flags[1] != 0;
    if (*stopped) return PSFUN_SUCCESS;
  }
  for (int i = 0; i < k; ++i) hf[i] = c->userf(fscale * hw[i], c->userf_data);
  PSFUN_HIP_TRY(hipMemcpyAsync(sc.fv, hf, sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
  return PSFUN_SUCCESS;
}

int userf_launch_est(psfun_ctx *c, LzCtl *ctl, int j, const double *sig, const UserfScratch &sc, int nl,
                     double eps) {
  hipLaunchKernelGGL(k_userf_est, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, j, sig, sc.Q, sc.fv, nl, eps,
                     c->d_flags);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

int userf_launch_final(psfun_ctx *c, LzCtl *ctl, int k, const UserfScratch &sc) {
  hipLaunchKernelGGL(k_userf_final, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, k, sc.Q, sc.fv);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

int userf_launch_apply(psfun_ctx *c, int k, const UserfScratch &sc, const double *x, double *y) {
  hipLaunchKernelGGL(k_sym_apply, dim3(1), dim3(PSFUN_BS), 0, c->stream, k, sc.Q, sc.fv, x, y);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}
