// Device-resident Arnoldi f(A)x with the reference's semantics
// (krylov/polynomial/psfun_d_arnoldi.f90:183-329).
//
// Reference loop (i = 2..nl):  itx = i-1; w = scaling*A v_{i-1};
// for k = 1..i-1: h(k,i-1) = <v_k, w>; w -= h(k,i-1) v_k   (modified Gram-Schmidt);
// h(i,i-1) = ||w||; v_i = (1/h(i,i-1)) w; if istop == 1: yk = f(H(1:i-1,1:i-1)) e1,
// rs(i) = rs(1) h(i,i-1) |yk(i-1)| / ||yk||, exit if rs(i) < eps.  No breakdown
// test (an exact breakdown gives rs(i) = 0 and ends the loop by convergence).
// Assembly: y = rs(1) (yk(1) v_1 + sum_{l=2}^{itx-1} yk(l) v_l)  (lines 318-324).
//
// Device schedule per column j = 1, 2, ...:
//   k_ar_spmv(j):   all blocks fold the ||w||^2 partials into h(j,j-1) (||x||
//                   for j = 1), write v_j = (1/h) w once and u = scaling*A v_j
//                   (gathering w*(1/h) on the fly), and leave partials of
//                   <v_1, u>.  Block 0 is the control block: it runs the
//                   reference's control for loop index i = j (the a-posteriori
//                   estimate from f(H_{j-1}) e1, by the dense block exp of
//                   fdense.hpp) while the streaming blocks work.
//   k_ar_mgs(k), k = 1..j: fold <v_k, w> into h(k,j), w -= h(k,j) v_k in place,
//                   leave partials of <v_{k+1}, w> (k < j) or ||w||^2 (k = j).
// Scalars cross kernels through kernel boundaries only; the control block
// stores its loop index and stop state to host-visible flags on which the
// host paces its launches (as in the Lanczos path).
#include "common.hpp"
#include "kernels.hpp"
#include "fdense.hpp"
#include "lanczos.hpp"
#include "userf.hpp"

// Reference loop bookkeeping at index i = j >= 2, with h(j,j-1) = sigma:
// the subdiagonal entry, itx, the host progress flag and, when no estimate is
// taken (istop 2), the iteration limit.  The istop-1 estimate runs in
// k_ar_est right after the step, so the dense f(H) code stays out of the
// streaming kernel (and out of its register budget).
__device__ void ar_control(LzCtl *ctl, double *H, int ldh, int j, double sigma, int nl, int istop, int *hf) {
  if (threadIdx.x != 0) return;
  host_flag_store(hf, j);
  H[(int64_t)(j - 2) * ldh + (j - 1)] = sigma;  // h(j, j-1)
  ctl->itx = j - 1;
  if (istop != 1 && j >= nl) {
    ctl->done = LZ_DONE_MAXIT;
    host_flag_store(hf + 1, LZ_DONE_MAXIT);
  }
}

// The reference's estimate at index j (istop 1, psfun_d_arnoldi.f90:215-257):
// yk = f(H(1:j-1,1:j-1)) e1, rs(j) = rs1 h(j,j-1) |yk(j-1)| / ||yk||, then
// the convergence test and the iteration limit.  WAVE: one wavefront
// (order <= 64); otherwise one block.
template <bool WAVE>
__global__ __launch_bounds__(WAVE ? 64 : PSFUN_BS) void k_ar_est(LzCtl *ctl, const double *H, int ldh, int model,
                                                                 int j, int nl, double eps, double fscale, int *hf) {
  if (ctl->done) return;
  const int tid = threadIdx.x;
  const int k = j - 1;
  double nrm, last;
  if (WAVE) {
    __shared__ WaveDenseLds L;
    const double y = wave_expmv(k, H, ldh, model, fscale, tid == 0 ? 1.0 : 0.0, &L);
    if (tid < k) ctl->yk[tid] = y;
    nrm = sqrt(wave_sum(tid < k ? y * y : 0.0));
    last = __shfl(y, k - 1, 64);
  } else {
    __shared__ DenseLds W;
    for (int i = tid; i < k; i += PSFUN_BS) W.y[i] = (i == 0) ? 1.0 : 0.0;
    __syncthreads();
    block_expmv(k, H, ldh, model, fscale, &W);
    double ss = 0.0;
    for (int i = tid; i < k; i += PSFUN_BS) {
      const double v = W.y[i];
      ss = fma(v, v, ss);
      ctl->yk[i] = v;
    }
    nrm = sqrt(dense_block_sum(ss, &W));
    last = W.y[k - 1];
  }
  if (tid != 0) return;
  const double sigma = H[(int64_t)(j - 2) * ldh + (j - 1)];
  const double errnum = ctl->rs1 * sigma * fabs(last);
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

static int ar_launch_est(psfun_ctx *c, LzCtl *ctl, const double *H, int ldh, int model, int j, int nl, double eps,
                         double fscale) {
  if (j - 1 <= PSFUN_WDENSE_K)
    hipLaunchKernelGGL(k_ar_est<true>, dim3(1), dim3(64), 0, c->stream, ctl, H, ldh, model, j, nl, eps, fscale,
                       c->d_flags);
  else
    hipLaunchKernelGGL(k_ar_est<false>, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, H, ldh, model, j, nl, eps,
                       fscale, c->d_flags);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

template <int L, bool SELL>
__global__ __launch_bounds__(PSFUN_BS) void k_ar_spmv(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                                      const double *__restrict__ val, double scaling,
                                                      const double *__restrict__ src, int j, int G,
                                                      const double *__restrict__ p_in, int g_in,
                                                      double *__restrict__ V, int64_t ldv, double *__restrict__ u_out,
                                                      double *__restrict__ p_out, LzCtl *ctl, double *H, int ldh,
                                                      int model, int nl, int istop, double eps, double fscale,
                                                      int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  // block-uniform: the control block of this same grid may set the flag
  if (__syncthreads_or(ctl->done)) return;
  const double sigma = sqrt(block_sum_partials<PSFUN_BS>(p_in, g_in, lds));
  if (blockIdx.x == 0) {  // control block
    if (j == 1) {
      if (threadIdx.x == 0) {
        host_flag_store(hf, 1);
        ctl->rs1 = sigma;
        ctl->rs[1] = sigma;
        if (sigma == 0.0) {  // x = 0: y = 0 (the reference would divide by zero)
          ctl->done = LZ_DONE_ZERO;
          host_flag_store(hf + 1, LZ_DONE_ZERO);
        }
      }
      return;
    }
    ar_control(ctl, H, ldh, j, sigma, nl, istop, hf);
    return;
  }
  if (sigma == 0.0 && j == 1) return;
  const double inv = 1.0 / sigma;
  double *vj = V + (int64_t)(j - 1) * ldv;
  constexpr int RPB = SELL ? PSFUN_BS : PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range_c0(nchunks, G);
  const int sub = SELL ? threadIdx.x : threadIdx.x / L, lane = SELL ? 0 : threadIdx.x % L;
  double dot = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    double sum = 0.0;
    if (SELL) {
      if (row < n) sum = sell_row_dot(row, rp, col, val, src, inv);
    } else {
      if (row < n) {
        const int s = rp[row], e = rp[row + 1];
        for (int q = s + lane; q < e; q += L) sum = fma(val[q], src[col[q]] * inv, sum);
      }
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
    }
    if (row < n && lane == 0) {
      const double v = src[row] * inv;
      const double u = scaling * sum;
      vj[row] = v;
      u_out[row] = u;
      const double v1 = (j == 1) ? v : V[row];
      dot = fma(v1, u, dot);
    }
  }
  dot = block_sum<PSFUN_BS>(dot, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x - 1] = dot;
}

// One modified Gram-Schmidt step k of column j (1-based).  Each thread keeps
// AR_UB rows' (w, v_k, v_{k+1}) loads in flight and visits its rows in
// grid-stride order, so the partials do not depend on AR_UB.
#define AR_UB 4
__global__ __launch_bounds__(PSFUN_BS) void k_ar_mgs(int n, int k, int j, const double *__restrict__ V, int64_t ldv,
                                                     double *__restrict__ w, const double *__restrict__ p_in, int g_in,
                                                     double *__restrict__ p_out, LzCtl *ctl, double *H, int ldh) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  const double *vk = V + (int64_t)(k - 1) * ldv;
  const double *vn = V + (int64_t)k * ldv;  // v_{k+1} (read only when k < j)
  const bool last = (k == j);
  const int64_t stride = (int64_t)gridDim.x * PSFUN_BS;
  const int64_t i0 = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x;
  // the first rows' loads are issued before the fold of the partials waits
  double ww[AR_UB], aa[AR_UB], bb[AR_UB];
#pragma unroll
  for (int q = 0; q < AR_UB; ++q) {
    const int64_t i = i0 + q * stride;
    ww[q] = aa[q] = bb[q] = 0.0;
    if (i < n) {
      ww[q] = w[i];
      aa[q] = vk[i];
      if (!last) bb[q] = vn[i];
    }
  }
  const double h = block_sum_partials<PSFUN_BS>(p_in, g_in, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) H[(int64_t)(j - 1) * ldh + (k - 1)] = h;  // h(k, j)
  double acc = 0.0;
  for (int64_t base = i0; base < n; base += AR_UB * stride) {
    if (base != i0) {
#pragma unroll
      for (int q = 0; q < AR_UB; ++q) {
        const int64_t i = base + q * stride;
        if (i < n) {
          ww[q] = w[i];
          aa[q] = vk[i];
          if (!last) bb[q] = vn[i];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < AR_UB; ++q) {
      const int64_t i = base + q * stride;
      if (i < n) {
        const double x = fma(-h, aa[q], ww[q]);
        w[i] = x;
        acc = last ? fma(x, x, acc) : fma(bb[q], x, acc);
      }
    }
  }
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = acc;
}

// Control at the last loop index i = nl (no SpMV needed there).
__global__ __launch_bounds__(PSFUN_BS) void k_ar_ctrl(const double *__restrict__ p_in, int g_in, int j, LzCtl *ctl,
                                                      double *H, int ldh, int model, int nl, int istop, double eps,
                                                      double fscale, int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  const double sigma = sqrt(block_sum_partials<PSFUN_BS>(p_in, g_in, lds));
  ar_control(ctl, H, ldh, j, sigma, nl, istop, hf);
}

// Coefficients when the loop did not end by convergence:
// yk = f(H(1:itx-1,1:itx-1)) e1 (psfun_d_arnoldi.f90:259-296); the 1x1 block
// stands in for itx-1 = 0.
__global__ __launch_bounds__(PSFUN_BS) void k_ar_final(LzCtl *ctl, const double *H, int ldh, int model,
                                                       double fscale) {
  __shared__ DenseLds W;
  const int done = ctl->done;
  if (done == LZ_DONE_CONVERGED || done == LZ_DONE_ZERO) return;
  int k = ctl->itx - 1;
  if (k < 1) k = 1;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W.y[i] = (i == 0) ? 1.0 : 0.0;
  __syncthreads();
  block_expmv(k, H, ldh, model, fscale, &W);
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) ctl->yk[i] = W.y[i];
  if (threadIdx.x == 0) ctl->m = k;
}

__global__ __launch_bounds__(PSFUN_BS) void k_ar_zero(double *H, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < count; i += (int64_t)gridDim.x * PSFUN_BS)
    H[i] = 0.0;
}

// y = rs1 * (yk(1) v_1 + sum_{l=2}^{m'} yk(l) v_l); m' = max(1, m-1) after
// convergence (the last coefficient is dropped), m otherwise.
__global__ __launch_bounds__(PSFUN_BS) void k_ar_assemble(int n, const double *__restrict__ V, int64_t ldv,
                                                          const LzCtl *ctl, double *__restrict__ y) {
  const int done = ctl->done;
  int m = ctl->m;
  if (done == LZ_DONE_CONVERGED) m = (m - 1 < 1) ? 1 : m - 1;
  if (done == LZ_DONE_ZERO) m = 0;
  const double rs1 = ctl->rs1;
  __shared__ double yk[LZ_NL_MAX];
  for (int l = threadIdx.x; l < m; l += PSFUN_BS) yk[l] = ctl->yk[l];
  __syncthreads();
  basis_combine<false>(n, V, ldv, m, yk, nullptr, rs1, y);
}

__global__ void k_ar_init(LzCtl *ctl) {
  if (threadIdx.x == 0) {
    ctl->done = 0;
    ctl->itx = 0;
    ctl->m = 0;
    ctl->errnum = 0.0;
    ctl->errden = 1.0;
  }
  for (int i = threadIdx.x; i < LZ_NL_MAX + 2; i += blockDim.x) ctl->rs[i] = 0.0;
}

__global__ __launch_bounds__(PSFUN_BS) void k_ar_sumsq(int64_t n, const double *__restrict__ x,
                                                       double *__restrict__ p_out) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS)
    acc = fma(x[i], x[i], acc);
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = acc;
}

// ---------------------------------------------------------------------------
static int dense_model(int variant) {
  if (variant == PSFUN_VAR_CHBSYM) return DENSE_SYM_UPPER;
  if (variant == PSFUN_VAR_CHBHES) return DENSE_HESSENBERG;
  return DENSE_FULL;
}

extern "C" int psfun_arnoldi_apply(psfun_ctx_t c, psfun_csr_t A, int fname, int variant, double fscale,
                                   double scaling, const double *x, double *y, double eps, int itmax, int istop,
                                   int *iter, double *err, double *res, int *nres) {
  if (!c || !A) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  const bool userf = fname == PSFUN_FN_USERF;
  if (userf) {
    PSFUN_TRY(userf_check(c, variant));
  } else {
    if (fname != PSFUN_FN_EXP) return PSFUN_ERR_UNSUPPORTED;
    if (variant < PSFUN_VAR_CHBSYM || variant > PSFUN_VAR_SASQ) return PSFUN_ERR_UNSUPPORTED;
  }
  if (itmax < 2 || itmax > LZ_NL_MAX || itmax > PSFUN_DENSE_KMAX) return PSFUN_ERR_INVALID_ARG;
  const int n = A->n;
  const int nl = itmax;
  if (n == 0) {
    if (iter) *iter = 0;
    if (err) *err = 0.0;
    if (nres) *nres = 0;
    return PSFUN_SUCCESS;
  }
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  const int model = userf ? DENSE_SYM_UPPER : dense_model(variant);
  // USERF: the control blocks keep the bookkeeping; with istop = 1 the
  // estimate and the iteration limit run in a host-synchronous step after
  // each control index (the user function runs on the host).
  UserfScratch sc;
  if (userf) {
    void *ws2 = nullptr;
    PSFUN_TRY(psfun_workspace2(c, userf_scratch_bytes(nl), &ws2));
    sc = userf_scratch_carve(ws2, nl);
  }
  const bool fsync = userf && istop == 1;
  const int istop_k = userf ? 2 : istop;
  const int nl_k = fsync ? (1 << 30) : nl;
  const int ldh = nl + 1;
  const int64_t ldv = (int64_t)align_up((size_t)n, 64);
  const size_t ctl_bytes = align_up(sizeof(LzCtl), 256);
  const size_t h_bytes = align_up(sizeof(double) * (size_t)ldh * ldh, 256);
  const size_t vbytes = sizeof(double) * (size_t)ldv * (size_t)(nl + 2);  // v_1..v_{nl-1}, u0, u1
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, ctl_bytes + h_bytes + vbytes, &ws));
  LzCtl *ctl = (LzCtl *)ws;
  double *H = (double *)((char *)ws + ctl_bytes);
  double *V = (double *)((char *)ws + ctl_bytes + h_bytes);
  double *U[2] = {V + (size_t)ldv * nl, V + (size_t)ldv * (nl + 1)};
  double *P[2] = {c->d_partials, c->d_partials + PSFUN_MAX_GRID};

  const bool sell = A->format == PSFUN_FORMAT_SELL;
  const int L = sell ? 1 : A->lanes;
  const int64_t chunks = ((int64_t)n + PSFUN_BS / L - 1) / (PSFUN_BS / L);
  int Gs = psfun_grid(c, chunks, sell ? 8 : 4);
  if (Gs + 1 > 8 * c->num_cu) Gs = (8 * c->num_cu - 1) / 8 * 8;
  const int Gu = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 4);

  c->h_flags[0] = 0;
  c->h_flags[1] = 0;
  hipLaunchKernelGGL(k_ar_init, dim3(1), dim3(256), 0, c->stream, ctl);
  hipLaunchKernelGGL(k_ar_zero, dim3(psfun_grid(c, ((int64_t)ldh * ldh + PSFUN_BS - 1) / PSFUN_BS, 1)),
                     dim3(PSFUN_BS), 0, c->stream, H, (int64_t)ldh * ldh);
  hipLaunchKernelGGL(k_ar_sumsq, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, (int64_t)n, x, P[0]);
  PSFUN_HIP_TRY(hipGetLastError());
  int cur = 0, gcur = Gu;  // slot holding the newest partials, and their count
  const double spmv_bytes = 12.0 * A->nnz + 4.0 * (A->n + 1) + 24.0 * A->n;
  bool ran_all = true;
  for (int j = 1; j <= nl - 1; ++j) {
    const double *src = (j == 1) ? x : U[j & 1];
    double *u = U[(j + 1) & 1];
    PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_AR_SPMV, spmv_bytes + (j > 1 ? 8.0 * n : 0.0), j));
    if (sell) {
      hipLaunchKernelGGL((k_ar_spmv<1, true>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n, A->sell_ptr,
                         A->sell_col, A->sell_val, scaling, src, j, Gs, P[cur], gcur, V, ldv, u, P[1 - cur], ctl, H,
                         ldh, model, nl_k, istop_k, eps, fscale, c->d_flags);
    } else {
#define AR_LAUNCH(LL)                                                                                               \
  hipLaunchKernelGGL((k_ar_spmv<LL, false>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n, A->rowptr, A->colind, \
                     A->val, scaling, src, j, Gs, P[cur], gcur, V, ldv, u, P[1 - cur], ctl, H, ldh, model, nl_k,   \
                     istop_k, eps, fscale, c->d_flags)
      switch (L) {
        case 1: AR_LAUNCH(1); break;
        case 2: AR_LAUNCH(2); break;
        case 4: AR_LAUNCH(4); break;
        case 8: AR_LAUNCH(8); break;
        case 16: AR_LAUNCH(16); break;
        case 32: AR_LAUNCH(32); break;
        default: AR_LAUNCH(64); break;
      }
#undef AR_LAUNCH
    }
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(psfun_prof_end(c));
    cur = 1 - cur;
    gcur = Gs;
    if (istop_k == 1 && j >= 2) PSFUN_TRY(ar_launch_est(c, ctl, H, ldh, model, j, nl_k, eps, fscale));
    if (fsync && j >= 2) {
      // control index j: yk = f(H(1:j-1,1:j-1)) e1, rs(j) from h(j, j-1)
      bool stopped = false;
      PSFUN_TRY(userf_eval(c, j - 1, SYM_FROM_UPPER, H, nullptr, ldh, fscale, sc, &stopped));
      if (stopped) {  // stopped at index j-1 (seen after the wait)
        ran_all = false;
        break;
      }
      PSFUN_TRY(userf_launch_est(c, ctl, j, H + (size_t)(j - 2) * ldh + (j - 1), sc, nl, eps));
    }
    for (int k = 1; k <= j; ++k) {
      PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_AR_MGS, (k < j ? 32.0 : 24.0) * n, j));
      hipLaunchKernelGGL(k_ar_mgs, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, k, j, V, ldv, u, P[cur], gcur,
                         P[1 - cur], ctl, H, ldh);
      PSFUN_HIP_TRY(hipGetLastError());
      PSFUN_TRY(psfun_prof_end(c));
      cur = 1 - cur;
      gcur = Gu;
    }
    if (!fsync && j >= 2 && j < nl - 1) {
      PSFUN_TRY(psfun_host_wait(c, j - 1));
      if (c->h_flags[1] != 0) {
        ran_all = false;
        break;
      }
    }
  }
  if (ran_all) {
    hipLaunchKernelGGL(k_ar_ctrl, dim3(1), dim3(PSFUN_BS), 0, c->stream, P[cur], gcur, nl, ctl, H, ldh, model,
                       nl_k, istop_k, eps, fscale, c->d_flags);
    PSFUN_HIP_TRY(hipGetLastError());
    if (istop_k == 1) PSFUN_TRY(ar_launch_est(c, ctl, H, ldh, model, nl, nl_k, eps, fscale));
    if (fsync) {
      bool stopped = false;
      PSFUN_TRY(userf_eval(c, nl - 1, SYM_FROM_UPPER, H, nullptr, ldh, fscale, sc, &stopped));
      if (!stopped) PSFUN_TRY(userf_launch_est(c, ctl, nl, H + (size_t)(nl - 2) * ldh + (nl - 1), sc, nl, eps));
    }
  }
  if (userf) {
    // yk = f(H(1:itx-1,1:itx-1)) e1 unless the loop converged (or x = 0)
    LzCtl *hs = (LzCtl *)((char *)c->h_pinned + 256);
    PSFUN_HIP_TRY(hipMemcpyAsync(hs, ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    if (hs->done != LZ_DONE_CONVERGED && hs->done != LZ_DONE_ZERO) {
      const int k = hs->itx - 1 < 1 ? 1 : hs->itx - 1;
      PSFUN_TRY(userf_eval(c, k, SYM_FROM_UPPER, H, nullptr, ldh, fscale, sc));
      PSFUN_TRY(userf_launch_final(c, ctl, k, sc));
    }
  } else {
    hipLaunchKernelGGL(k_ar_final, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, H, ldh, model, fscale);
    PSFUN_HIP_TRY(hipGetLastError());
  }
  const int Ga = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  hipLaunchKernelGGL(k_ar_assemble, dim3(Ga), dim3(PSFUN_BS), 0, c->stream, n, V, ldv, ctl, y);
  PSFUN_HIP_TRY(hipGetLastError());

  LzCtl *h = (LzCtl *)((char *)c->h_pinned + 256);
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(LzCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  const int itx = h->itx;
  psfun_prof_discard_after(c, PSFUN_PROF_AR_SPMV, itx + 1);
  psfun_prof_discard_after(c, PSFUN_PROF_AR_MGS, itx);
  const double e = (h->errden == 0.0) ? h->errnum : h->errnum / h->errden;
  if (iter) *iter = itx;
  if (err) *err = e;
  if (res) {
    for (int i = 1; i <= itx; ++i) res[i - 1] = h->rs[i];
    res[itx] = e;
  }
  if (nres) *nres = itx + 1;
  return PSFUN_SUCCESS;
}

// ===========================================================================
// Shift-and-invert Arnoldi (krylov/rational/psfun_d_sai.f90:183-390): the
// Krylov space of (A + tau I)^{-1}, every inner solve by the device CG of
// psfun_cg_shifted; MGS as in Arnoldi; the estimate from
// f(H_k^{-1} - tau I) e1, H_k^{-1} by an in-block Gauss-Jordan inverse.
// ===========================================================================

// v_out = w / ||w|| from the partials of ||w||^2; with first != 0 this is
// v_1 = x / ||x|| and also records rs(1) (zero x: stop with y = 0).
__global__ __launch_bounds__(PSFUN_BS) void k_sai_normalize(int n, const double *__restrict__ w,
                                                            double *__restrict__ v_out,
                                                            const double *__restrict__ p_in, int g_in, LzCtl *ctl,
                                                            int first) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  const double sigma = sqrt(block_sum_partials<PSFUN_BS>(p_in, g_in, lds));
  if (first && blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->rs1 = sigma;
    ctl->rs[1] = sigma;
    if (sigma == 0.0) ctl->done = LZ_DONE_ZERO;
  }
  if (sigma == 0.0) return;
  const double inv = 1.0 / sigma;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS)
    v_out[i] = w[i] * inv;
}

__global__ __launch_bounds__(PSFUN_BS) void k_sai_dot(int64_t n, const double *__restrict__ a,
                                                      const double *__restrict__ b, double *__restrict__ p_out) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS)
    acc = fma(a[i], b[i], acc);
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = acc;
}

// yk = f(H_k^{-1} - tau I) e1 into W->y; false when H_k is singular.
__device__ bool sai_fproj(int k, const double *H, int ldh, int model, double *M, double *Hi, double tau,
                          double fscale, DenseLds *W) {
  if (!block_gj_inverse(k, H, ldh, M, Hi, W)) return false;
  for (int d = threadIdx.x; d < k; d += PSFUN_BS) Hi[d * k + d] -= tau;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W->y[i] = (i == 0) ? 1.0 : 0.0;
  __syncthreads();
  block_expmv(k, Hi, k, model, fscale, W);
  return true;
}

// Reference control at loop index i (h(i,i-1) = ||w|| from the partials).
__global__ __launch_bounds__(PSFUN_BS) void k_sai_ctrl(const double *__restrict__ p_in, int g_in, int i, LzCtl *ctl,
                                                       double *H, int ldh, int model, double *M, double *Hi,
                                                       double tau, int nl, int istop, double eps, double fscale) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  __shared__ DenseLds W;
  if (ctl->done) return;
  const int tid = threadIdx.x;
  const double sigma = sqrt(block_sum_partials<PSFUN_BS>(p_in, g_in, lds));
  if (tid == 0) {
    H[(int64_t)(i - 2) * ldh + (i - 1)] = sigma;  // h(i, i-1)
    ctl->itx = i - 1;
  }
  if (istop == 1) {
    const int k = i - 1;
    if (!sai_fproj(k, H, ldh, model, M, Hi, tau, fscale, &W)) {
      if (tid == 0) ctl->done = LZ_DONE_BREAKDOWN;
      return;
    }
    double ss = 0.0;
    for (int q = tid; q < k; q += PSFUN_BS) {
      const double v = W.y[q];
      ss = fma(v, v, ss);
      ctl->yk[q] = v;
    }
    const double nrm = sqrt(dense_block_sum(ss, &W));
    const double errnum = ctl->rs1 * sigma * fabs(W.y[k - 1]);
    const double rsi = errnum / nrm;
    if (tid == 0) {
      ctl->errnum = errnum;
      ctl->errden = nrm;
      ctl->rs[i] = rsi;
      if (rsi < eps) {
        ctl->m = k;
        ctl->done = LZ_DONE_CONVERGED;
      }
    }
    if (rsi < eps) return;
  }
  if (i >= nl && tid == 0) ctl->done = LZ_DONE_MAXIT;
}

// Coefficients when the loop did not end by convergence: the same projection
// f(H_{itx-1}^{-1} - tau I) e1 (the 1x1 block stands in for itx-1 = 0).
__global__ __launch_bounds__(PSFUN_BS) void k_sai_final(LzCtl *ctl, const double *H, int ldh, int model, double *M,
                                                        double *Hi, double tau, double fscale) {
  __shared__ DenseLds W;
  const int done = ctl->done;
  if (done == LZ_DONE_CONVERGED || done == LZ_DONE_ZERO || done == LZ_DONE_BREAKDOWN) return;
  int k = ctl->itx - 1;
  if (k < 1) k = 1;
  if (!sai_fproj(k, H, ldh, model, M, Hi, tau, fscale, &W)) {
    if (threadIdx.x == 0) ctl->done = LZ_DONE_BREAKDOWN;
    return;
  }
  for (int q = threadIdx.x; q < k; q += PSFUN_BS) ctl->yk[q] = W.y[q];
  if (threadIdx.x == 0) ctl->m = k;
}

// USERF projection matrix Hi = H_k^{-1} - tau I (k x k, leading dimension k)
// for the symmetric eigen-step; a singular H_k stops the loop with a
// breakdown.  Inside the loop (final == 0) a stopped loop skips it.
__global__ __launch_bounds__(PSFUN_BS) void k_sai_inv(LzCtl *ctl, int k, const double *H, int ldh, double *M,
                                                      double *Hi, double tau, int final) {
  __shared__ DenseLds W;
  if (!final && ctl->done) return;
  if (!block_gj_inverse(k, H, ldh, M, Hi, &W)) {
    if (threadIdx.x == 0) ctl->done = LZ_DONE_BREAKDOWN;
    return;
  }
  for (int d = threadIdx.x; d < k; d += PSFUN_BS) Hi[d * k + d] -= tau;
}

extern "C" int psfun_sai_arnoldi_apply(psfun_ctx_t c, psfun_csr_t A, int fname, int variant, double fscale,
                                       double tau, int prec, int inner_itmax, int inner_istop, const double *x,
                                       double *y, double eps, int itmax, int istop, int *iter, double *err,
                                       double *res, int *nres) {
  if (!c || !A) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  const bool userf = fname == PSFUN_FN_USERF;
  if (userf) {
    PSFUN_TRY(userf_check(c, variant));
  } else {
    if (fname != PSFUN_FN_EXP) return PSFUN_ERR_UNSUPPORTED;
    if (variant < PSFUN_VAR_CHBSYM || variant > PSFUN_VAR_SASQ) return PSFUN_ERR_UNSUPPORTED;
  }
  if (itmax < 2 || itmax > LZ_NL_MAX || itmax > PSFUN_DENSE_KMAX) return PSFUN_ERR_INVALID_ARG;
  if (prec != PSFUN_PREC_NONE && prec != PSFUN_PREC_DIAG) return PSFUN_ERR_INVALID_ARG;
  if (inner_istop < 1 || inner_istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  const int n = A->n;
  const int nl = itmax;
  if (n == 0) {
    if (iter) *iter = 0;
    if (err) *err = 0.0;
    if (nres) *nres = 0;
    return PSFUN_SUCCESS;
  }
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  const int model = userf ? DENSE_SYM_UPPER : dense_model(variant);
  const bool fsync = userf && istop == 1;
  const int istop_k = userf ? 2 : istop;
  const int nl_k = fsync ? (1 << 30) : nl;
  const int ldh = nl + 1;
  const int64_t ldv = (int64_t)align_up((size_t)n, 64);
  const size_t ctl_bytes = align_up(sizeof(LzCtl), 256);
  const size_t h_bytes = align_up(sizeof(double) * (size_t)ldh * ldh, 256);
  const size_t m_bytes = align_up(sizeof(double) * (size_t)nl * nl, 256);
  const size_t vbytes = align_up(sizeof(double) * (size_t)ldv * (size_t)(nl + 1), 256);  // v_1..v_nl, w
  const size_t u_bytes = userf ? userf_scratch_bytes(nl) : 0;
  // the basis lives in the second workspace: the inner CG takes the first
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace2(c, ctl_bytes + h_bytes + 2 * m_bytes + vbytes + u_bytes, &ws));
  UserfScratch sc;
  if (userf) sc = userf_scratch_carve((char *)ws + ctl_bytes + h_bytes + 2 * m_bytes + vbytes, nl);
  LzCtl *ctl = (LzCtl *)ws;
  double *H = (double *)((char *)ws + ctl_bytes);
  double *M = (double *)((char *)ws + ctl_bytes + h_bytes);
  double *Hi = (double *)((char *)ws + ctl_bytes + h_bytes + m_bytes);
  double *V = (double *)((char *)ws + ctl_bytes + h_bytes + 2 * m_bytes);
  double *w = V + (size_t)ldv * nl;
  double *P[2] = {c->d_partials + 5 * PSFUN_MAX_GRID, c->d_partials + 6 * PSFUN_MAX_GRID};
  const int Gu = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 4);
  LzCtl *h = (LzCtl *)((char *)c->h_pinned + 256);

  hipLaunchKernelGGL(k_ar_init, dim3(1), dim3(256), 0, c->stream, ctl);
  hipLaunchKernelGGL(k_ar_zero, dim3(psfun_grid(c, ((int64_t)ldh * ldh + PSFUN_BS - 1) / PSFUN_BS, 1)),
                     dim3(PSFUN_BS), 0, c->stream, H, (int64_t)ldh * ldh);
  hipLaunchKernelGGL(k_ar_sumsq, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, (int64_t)n, x, P[0]);
  hipLaunchKernelGGL(k_sai_normalize, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, x, V, P[0], Gu, ctl, 1);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  const int Gz = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 4);
  if (h->done == 0) {
    for (int i = 2; i <= nl; ++i) {
      const int j = i - 1;
      // w = (A + tau I)^{-1} v_j, zero initial guess (psfun_d_sai.f90:216-218)
      hipLaunchKernelGGL(k_ar_zero, dim3(Gz), dim3(PSFUN_BS), 0, c->stream, w, (int64_t)n);
      PSFUN_HIP_TRY(hipGetLastError());
      int it_in = 0;
      double err_in = 0.0;
      const int st = psfun_cg_shifted(c, A, prec, V + (size_t)ldv * (j - 1), 1.0, tau, w, eps, inner_itmax,
                                      inner_istop, &it_in, &err_in, nullptr);
      if (st != PSFUN_SUCCESS) return st;
      // modified Gram-Schmidt against v_1..v_j, then v_{j+1} = w / h(j+1, j)
      hipLaunchKernelGGL(k_sai_dot, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, (int64_t)n, V, w, P[0]);
      int cur = 0;
      for (int k = 1; k <= j; ++k) {
        hipLaunchKernelGGL(k_ar_mgs, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, k, j, V, ldv, w, P[cur], Gu,
                           P[1 - cur], ctl, H, ldh);
        cur = 1 - cur;
      }
      hipLaunchKernelGGL(k_sai_normalize, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, w, V + (size_t)ldv * j,
                         P[cur], Gu, ctl, 0);
      hipLaunchKernelGGL(k_sai_ctrl, dim3(1), dim3(PSFUN_BS), 0, c->stream, P[cur], Gu, i, ctl, H, ldh, model, M,
                         Hi, tau, nl_k, istop_k, eps, fscale);
      PSFUN_HIP_TRY(hipGetLastError());
      if (fsync) {
        // yk = f(H_{i-1}^{-1} - tau I) e1 with f on the host, rs(i) from h(i, i-1)
        hipLaunchKernelGGL(k_sai_inv, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, i - 1, H, ldh, M, Hi, tau, 0);
        PSFUN_HIP_TRY(hipGetLastError());
        PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
        if (h->done != 0) break;
        PSFUN_TRY(userf_eval(c, i - 1, SYM_FROM_UPPER, Hi, nullptr, i - 1, fscale, sc));
        PSFUN_TRY(userf_launch_est(c, ctl, i, H + (size_t)(i - 2) * ldh + (i - 1), sc, nl, eps));
      }
      PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
      if (h->done != 0) break;
    }
  }
  if (userf) {
    PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    if (h->done != LZ_DONE_CONVERGED && h->done != LZ_DONE_ZERO && h->done != LZ_DONE_BREAKDOWN) {
      const int k = h->itx - 1 < 1 ? 1 : h->itx - 1;
      hipLaunchKernelGGL(k_sai_inv, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, k, H, ldh, M, Hi, tau, 1);
      PSFUN_HIP_TRY(hipGetLastError());
      PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
      if (h->done != LZ_DONE_BREAKDOWN) {
        PSFUN_TRY(userf_eval(c, k, SYM_FROM_UPPER, Hi, nullptr, k, fscale, sc));
        PSFUN_TRY(userf_launch_final(c, ctl, k, sc));
      }
    }
  } else {
    hipLaunchKernelGGL(k_sai_final, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, H, ldh, model, M, Hi, tau, fscale);
    PSFUN_HIP_TRY(hipGetLastError());
  }
  const int Ga = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  hipLaunchKernelGGL(k_ar_assemble, dim3(Ga), dim3(PSFUN_BS), 0, c->stream, n, V, ldv, ctl, y);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(LzCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  if (h->done == LZ_DONE_BREAKDOWN) return PSFUN_ERR_BREAKDOWN;
  const int itx = h->itx;
  const double e = (h->errden == 0.0) ? h->errnum : h->errnum / h->errden;
  if (iter) *iter = itx;
  if (err) *err = e;
  if (res) {
    for (int q = 1; q <= itx; ++q) res[q - 1] = h->rs[q];
    res[itx] = e;
  }
  if (nres) *nres = itx + 1;
  return PSFUN_SUCCESS;
}
