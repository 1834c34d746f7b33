// The f(H) step as a stand-alone call (psfun_d_serial%apply on a dense array,
// serial/psfun_d_serial_mod.F90:172-523), evaluated on the device.
#include "common.hpp"
#include "fdense.hpp"
#include "fexp.hpp"
#include "kernels.hpp"
#include "userf.hpp"

// y = exp(s*T) x for symmetric tridiagonal T (d: k, e: k-1), one wavefront.
__global__ __launch_bounds__(64) void k_serial_tridiag(int k, const double *d, const double *e, double s,
                                                       const double *x, double *y) {
  __shared__ ChebLds W;
  for (int i = threadIdx.x; i < k; i += 64) W.y[i] = x[i];
  wave_lds_sync();
  wave_exp_tridiag(k, d, e, s, &W);
  for (int i = threadIdx.x; i < k; i += 64) y[i] = W.y[i];
}

// y = exp(s*H) x for a dense k x k column-major H read under `model`
// (full, Hessenberg part, or symmetric from the upper triangle), one block.
__global__ __launch_bounds__(PSFUN_BS) void k_serial_dense(int k, const double *H, int model, double s,
                                                          const double *x, double *y) {
  __shared__ DenseLds W;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W.y[i] = x[i];
  __syncthreads();
  block_expmv(k, H, k, model, s, &W);
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) y[i] = W.y[i];
}

// y = exp(s*H) x by the degree-`ideg` Pade approximant (GENPADE with a
// degree other than 6), one block; S: 5 k*k doubles; *bad = 1 on a singular
// Pade denominator.
__global__ __launch_bounds__(PSFUN_BS) void k_serial_pade(int k, const double *H, double s, int ideg, double *S,
                                                         const double *x, double *y, int *bad) {
  __shared__ DenseLds W;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W.y[i] = x[i];
  __syncthreads();
  const bool ok = block_pade_expmv(k, H, k, s, ideg, S, &W);
  if (threadIdx.x == 0) *bad = ok ? 0 : 1;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) y[i] = W.y[i];
}

// The same for k <= 64 by one wavefront (wave_expmv).
__global__ __launch_bounds__(64) void k_serial_dense_wave(int k, const double *H, int model, double s,
                                                         const double *x, double *y) {
  __shared__ WaveDenseLds L;
  const int lane = threadIdx.x;
  const double v = wave_expmv(k, H, k, model, s, lane < k ? x[lane] : 0.0, &L);
  if (lane < k) y[lane] = v;
}

int launch_serial_tridiag(psfun_ctx *c, int k, const double *d, const double *e, double s, const double *x,
                          double *y) {
  hipLaunchKernelGGL(k_serial_tridiag, dim3(1), dim3(64), 0, c->stream, k, d, e, s, x, y);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

extern "C" int psfun_serial_apply_host(psfun_ctx_t c, int fname, int variant, double scaling, int k,
                                       const double *H, const double *x, double *y) {
  if (!c || k < 0 || (k && (!H || !x || !y))) return PSFUN_ERR_INVALID_ARG;
  if (fname == PSFUN_FN_USERF) {
    PSFUN_TRY(userf_check(c, variant));
  } else {
    if (fname != PSFUN_FN_EXP) return PSFUN_ERR_UNSUPPORTED;
    if (variant < PSFUN_VAR_CHBSYM || variant > PSFUN_VAR_SASQ) return PSFUN_ERR_UNSUPPORTED;
  }
  if (k == 0) return PSFUN_SUCCESS;
  if (k > PSFUN_DENSE_KMAX) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  if (fname == PSFUN_FN_USERF) {
    // DSYEV('V','U') on H, y = Q diag(f(scaling*lambda)) Q^T x  (serial_mod:420-485)
    const size_t hk = align_up(sizeof(double) * (size_t)k * k, 256);
    const size_t kv = align_up(sizeof(double) * (size_t)k, 256);
    void *ws = nullptr;
    PSFUN_TRY(psfun_workspace2(c, hk + 2 * kv + userf_scratch_bytes(k), &ws));
    double *dH = (double *)ws;
    double *dx = (double *)((char *)ws + hk), *dy = (double *)((char *)ws + hk + kv);
    const UserfScratch sc = userf_scratch_carve((char *)ws + hk + 2 * kv, k);
    PSFUN_HIP_TRY(hipMemcpyAsync(dH, H, sizeof(double) * k * k, hipMemcpyHostToDevice, c->stream));
    PSFUN_HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
    PSFUN_TRY(userf_eval(c, k, SYM_FROM_UPPER, dH, nullptr, k, scaling, sc));
    PSFUN_TRY(userf_launch_apply(c, k, sc, dx, dy));
    PSFUN_HIP_TRY(hipMemcpyAsync(y, dy, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    return PSFUN_SUCCESS;
  }
  // CHBSYM reads H as symmetric from its upper triangle (DSCHBV factors with
  // ZSYSV 'U'); when that matrix is tridiagonal the Chebyshev LDS routine is
  // used, otherwise the dense block routine, as for every other variant.
  bool tridiag = false;
  if (variant == PSFUN_VAR_CHBSYM && k <= PSFUN_FT_KMAX) {
    tridiag = true;
    for (int j = 0; j < k && tridiag; ++j)
      for (int i = 0; i + 1 < j; ++i)
        if (H[(size_t)j * k + i] != 0.0) {
          tridiag = false;
          break;
        }
  }
  void *ws = nullptr;
  if (tridiag) {
    std::vector<double> buf(3 * (size_t)k, 0.0);
    for (int i = 0; i < k; ++i) buf[i] = H[(size_t)i * k + i];
    for (int i = 0; i + 1 < k; ++i) buf[k + i] = H[(size_t)(i + 1) * k + i];  // upper triangle
    for (int i = 0; i < k; ++i) buf[2 * k + i] = x[i];
    PSFUN_TRY(psfun_workspace(c, sizeof(double) * 4 * (size_t)k, &ws));
    double *dd = (double *)ws;
    PSFUN_HIP_TRY(hipMemcpyAsync(dd, buf.data(), sizeof(double) * 3 * k, hipMemcpyHostToDevice, c->stream));
    PSFUN_TRY(launch_serial_tridiag(c, k, dd, dd + k, scaling, dd + 2 * k, dd + 3 * k));
    PSFUN_HIP_TRY(hipMemcpyAsync(y, dd + 3 * k, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
  } else if (variant == PSFUN_VAR_GENPADE && c->pade_degree != 6) {
    const size_t hk = (size_t)k * k;
    PSFUN_TRY(psfun_workspace(c, sizeof(double) * (6 * hk + 2 * (size_t)k) + 256, &ws));
    double *dH = (double *)ws;
    double *dx = dH + hk, *dy = dx + k, *S = dy + k;
    int *dbad = (int *)(S + 5 * hk);
    PSFUN_HIP_TRY(hipMemcpyAsync(dH, H, sizeof(double) * hk, hipMemcpyHostToDevice, c->stream));
    PSFUN_HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_serial_pade, dim3(1), dim3(PSFUN_BS), 0, c->stream, k, (const double *)dH, scaling,
                       c->pade_degree, S, (const double *)dx, dy, dbad);
    PSFUN_HIP_TRY(hipGetLastError());
    int *hb = (int *)c->h_pinned;
    PSFUN_HIP_TRY(hipMemcpyAsync(y, dy, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipMemcpyAsync(hb, dbad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    return *hb ? PSFUN_ERR_BREAKDOWN : PSFUN_SUCCESS;
  } else {
    const int model = variant == PSFUN_VAR_CHBSYM   ? DENSE_SYM_UPPER
                      : variant == PSFUN_VAR_CHBHES ? DENSE_HESSENBERG
                                                    : DENSE_FULL;
    const size_t hk = (size_t)k * k;
    PSFUN_TRY(psfun_workspace(c, sizeof(double) * (hk + 2 * (size_t)k), &ws));
    double *dH = (double *)ws;
    double *dx = dH + hk, *dy = dx + k;
    PSFUN_HIP_TRY(hipMemcpyAsync(dH, H, sizeof(double) * hk, hipMemcpyHostToDevice, c->stream));
    PSFUN_HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
    if (k <= PSFUN_WDENSE_K)
      hipLaunchKernelGGL(k_serial_dense_wave, dim3(1), dim3(64), 0, c->stream, k, dH, model, scaling, dx, dy);
    else
      hipLaunchKernelGGL(k_serial_dense, dim3(1), dim3(PSFUN_BS), 0, c->stream, k, dH, model, scaling, dx, dy);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_HIP_TRY(hipMemcpyAsync(y, dy, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
  }
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSFUN_SUCCESS;
}
