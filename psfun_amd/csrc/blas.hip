// CSR SpMV/SpMM and BLAS-1 (the PSBLAS primitives psb_spmm, psb_gedot,
// psb_genrm2, psb_geamax, psb_geaxpby, v%scal) for gfx950, FP64.
//
// SpMV is vector-CSR: a group of L lanes (L a power of two, L <= 64) owns one
// row, reads that row's val/colind contiguously (so a wave reads 64/L rows'
// worth of consecutive nonzeros: coalesced), gathers x through L2/MALL and
// reduces the row with cross-lane shuffles.  Row chunks are split over the
// XCDs so that neighbouring rows, which gather neighbouring x, share an L2.
#include "common.hpp"
#include "kernels.hpp"

// ---------------------------------------------------------------------------
// SpMV
// ---------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(PSFUN_BS) void k_csr_spmv(int n, const int *__restrict__ rowptr,
                                                       const int *__restrict__ col,
                                                       const double *__restrict__ val, double alpha,
                                                       const double *__restrict__ x, double beta,
                                                       double *__restrict__ y) {
  constexpr int RPB = PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range(nchunks);
  const int sub = threadIdx.x / L, lane = threadIdx.x % L;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    double sum = 0.0;
    if (row < n) {
      const int s = rowptr[row], e = rowptr[row + 1];
      for (int j = s + lane; j < e; j += L) sum = fma(val[j], x[col[j]], sum);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
    if (row < n && lane == 0) {
      double r = alpha * sum;
      if (beta != 0.0) r = fma(beta, y[row], r);
      y[row] = r;
    }
  }
}

// Sliced-ELL SpMV: one thread per row (chunks of PSFUN_BS rows = 4 slices),
// each entry step reads one full cache line of colind and two of val per wave.
__global__ __launch_bounds__(PSFUN_BS) void k_sell_spmv(int n, const int *__restrict__ sp,
                                                        const int *__restrict__ col,
                                                        const double *__restrict__ val, double alpha,
                                                        const double *__restrict__ x, double beta,
                                                        double *__restrict__ y) {
  const int64_t nchunks = (n + PSFUN_BS - 1) / PSFUN_BS;
  const XcdRange xr = xcd_range(nchunks);
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * PSFUN_BS + threadIdx.x);
    if (row < n) {
      double r = alpha * sell_row_dot(row, sp, col, val, x, 1.0);
      if (beta != 0.0) r = fma(beta, y[row], r);
      y[row] = r;
    }
  }
}

// The same on the diagonal slices (chunks of PSFUN_BS rows keep every
// wavefront inside one 64-row slice).
__global__ __launch_bounds__(PSFUN_BS) void k_hdia_spmv(int n, const int *__restrict__ hp,
                                                        const int *__restrict__ off,
                                                        const double *__restrict__ hv, double alpha,
                                                        const double *__restrict__ x, double beta,
                                                        double *__restrict__ y) {
  const int64_t nchunks = (n + PSFUN_BS - 1) / PSFUN_BS;
  const XcdRange xr = xcd_range(nchunks);
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * PSFUN_BS + threadIdx.x);
    if (row < n) {
      double r = alpha * hdia_row_dot(row, n, hp, off, hv, x);
      if (beta != 0.0) r = fma(beta, y[row], r);
      y[row] = r;
    }
  }
}

// Column-major multi-RHS: Y(:,c) = alpha*A*X(:,c) + beta*Y(:,c), KB columns per pass.
template <int L, int KB>
__global__ __launch_bounds__(PSFUN_BS) void k_csr_spmm(int n, const int *__restrict__ rowptr,
                                                       const int *__restrict__ col,
                                                       const double *__restrict__ val, int k,
                                                       double alpha, const double *__restrict__ X,
                                                       int64_t ldx, double beta,
                                                       double *__restrict__ Y, int64_t ldy) {
  constexpr int RPB = PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range(nchunks);
  const int sub = threadIdx.x / L, lane = threadIdx.x % L;
  const int c0 = blockIdx.y * KB;
  const int kc = min(KB, k - c0);
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    double sum[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) sum[q] = 0.0;
    if (row < n) {
      const int s = rowptr[row], e = rowptr[row + 1];
      for (int j = s + lane; j < e; j += L) {
        const double a = val[j];
        const int64_t cj = col[j];
#pragma unroll
        for (int q = 0; q < KB; ++q)
          if (q < kc) sum[q] = fma(a, X[cj + (int64_t)(c0 + q) * ldx], sum[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < KB; ++q)
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1) sum[q] += __shfl_xor(sum[q], off, L);
    if (row < n && lane == 0) {
#pragma unroll
      for (int q = 0; q < KB; ++q)
        if (q < kc) {
          double *yp = Y + row + (int64_t)(c0 + q) * ldy;
          double r = alpha * sum[q];
          if (beta != 0.0) r = fma(beta, *yp, r);
          *yp = r;
        }
    }
  }
}

#define PSFUN_DISPATCH_LANES(L, ...)        \
  switch (L) {                              \
    case 1: { constexpr int LL = 1; __VA_ARGS__; } break;   \
    case 2: { constexpr int LL = 2; __VA_ARGS__; } break;   \
    case 4: { constexpr int LL = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int LL = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int LL = 16; __VA_ARGS__; } break; \
    case 32: { constexpr int LL = 32; __VA_ARGS__; } break; \
    default: { constexpr int LL = 64; __VA_ARGS__; } break; \
  }

int launch_spmv(psfun_ctx *c, const psfun_csr *A, double alpha, const double *x, double beta, double *y) {
  if (A->n == 0) return PSFUN_SUCCESS;
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_SPMV,
                             12.0 * A->nnz + 4.0 * (A->n + 1) + (beta != 0.0 ? 24.0 : 16.0) * A->n));
  if (psfun_reads_hdia(A)) {
    const int G = psfun_grid(c, ((int64_t)A->n + PSFUN_BS - 1) / PSFUN_BS);
    hipLaunchKernelGGL(k_hdia_spmv, dim3(G), dim3(PSFUN_BS), 0, c->stream, A->n, A->hdia_ptr, A->hdia_off,
                       A->hdia_val, alpha, x, beta, y);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(psfun_prof_end(c));
    return PSFUN_SUCCESS;
  }
  if (psfun_reads_sell(A)) {
    const int G = psfun_grid(c, ((int64_t)A->n + PSFUN_BS - 1) / PSFUN_BS);
    hipLaunchKernelGGL(k_sell_spmv, dim3(G), dim3(PSFUN_BS), 0, c->stream, A->n, A->sell_ptr, A->sell_col,
                       A->sell_val, alpha, x, beta, y);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(psfun_prof_end(c));
    return PSFUN_SUCCESS;
  }
  const int L = A->lanes;
  const int64_t chunks = ((int64_t)A->n + PSFUN_BS / L - 1) / (PSFUN_BS / L);
  const int G = psfun_grid(c, chunks);
  PSFUN_DISPATCH_LANES(L, hipLaunchKernelGGL((k_csr_spmv<LL>), dim3(G), dim3(PSFUN_BS), 0, c->stream, A->n,
                                             A->rowptr, A->colind, A->val, alpha, x, beta, y));
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));
  return PSFUN_SUCCESS;
}

extern "C" int psfun_spmv(psfun_ctx_t c, psfun_csr_t A, double alpha, const double *x, double beta,
                          double *y) {
  if (!c || !A || (A->n && (!x || !y))) return PSFUN_ERR_INVALID_ARG;
  return launch_spmv(c, A, alpha, x, beta, y);
}

extern "C" int psfun_spmm(psfun_ctx_t c, psfun_csr_t A, int k, double alpha, const double *X, int64_t ldx,
                          double beta, double *Y, int64_t ldy) {
  if (!c || !A || k < 0 || (A->n && k && (!X || !Y)) || ldx < A->n || ldy < A->n) return PSFUN_ERR_INVALID_ARG;
  if (A->n == 0 || k == 0) return PSFUN_SUCCESS;
  constexpr int KB = 8;
  const int L = A->lanes;
  const int64_t chunks = ((int64_t)A->n + PSFUN_BS / L - 1) / (PSFUN_BS / L);
  const int G = psfun_grid(c, chunks, 8);
  dim3 grid(G, (k + KB - 1) / KB);
  PSFUN_DISPATCH_LANES(L, hipLaunchKernelGGL((k_csr_spmm<LL, KB>), grid, dim3(PSFUN_BS), 0, c->stream, A->n,
                                             A->rowptr, A->colind, A->val, k, alpha, X, ldx, beta, Y, ldy));
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

// ---------------------------------------------------------------------------
// BLAS-1
// ---------------------------------------------------------------------------
// op: 0 dot(x,y), 1 sum of squares -> sqrt, 2 max |x|.  Two launches: per-block
// partials, then one block folds them in a fixed order (deterministic, and no
// grid-wide atomics or fences).
template <int OP>
__global__ __launch_bounds__(PSFUN_BS) void k_reduce_partials(int64_t n, const double *__restrict__ x,
                                                              const double *__restrict__ y,
                                                              double *__restrict__ partials) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    const double a = x[i];
    if (OP == 0) acc = fma(a, y[i], acc);
    else if (OP == 1) acc = fma(a, a, acc);
    else acc = fmax(acc, fabs(a));
  }
  acc = (OP == 2) ? block_max<PSFUN_BS>(acc, lds) : block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

template <int OP>
__global__ __launch_bounds__(PSFUN_BS) void k_reduce_final(const double *__restrict__ partials, int G,
                                                           double *__restrict__ out) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double acc = 0.0;
  for (int b = threadIdx.x; b < G; b += PSFUN_BS) acc = (OP == 2) ? fmax(acc, partials[b]) : acc + partials[b];
  acc = (OP == 2) ? block_max<PSFUN_BS>(acc, lds) : block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) *out = (OP == 1) ? sqrt(acc) : acc;
}

int launch_reduce(psfun_ctx *c, int op, int64_t n, const double *x, const double *y, double *d_out) {
  const int64_t chunks = (n + PSFUN_BS - 1) / PSFUN_BS;
  const int G = psfun_grid(c, chunks > 0 ? chunks : 1, 16);
  double *part = c->d_partials + 7 * PSFUN_MAX_GRID;
  // algorithmic bytes: x (and y for a dot) read once
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_REDUCE, (op == 0 ? 16.0 : 8.0) * n));
  if (op == 0) hipLaunchKernelGGL(k_reduce_partials<0>, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, x, y, part);
  else if (op == 1) hipLaunchKernelGGL(k_reduce_partials<1>, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, x, y, part);
  else hipLaunchKernelGGL(k_reduce_partials<2>, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, x, y, part);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));
  if (op == 0) hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(PSFUN_BS), 0, c->stream, (const double *)part, G, d_out);
  else if (op == 1) hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(PSFUN_BS), 0, c->stream, (const double *)part, G, d_out);
  else hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), dim3(PSFUN_BS), 0, c->stream, (const double *)part, G, d_out);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

static int reduce_to_host(psfun_ctx *c, int op, int64_t n, const double *x, const double *y, double *res) {
  if (!c || !res || n < 0 || (n && (!x || (op == 0 && !y)))) return PSFUN_ERR_INVALID_ARG;
  PSFUN_TRY(launch_reduce(c, op, n, x, y, c->d_scalar));
  double *h = (double *)c->h_pinned;
  PSFUN_HIP_TRY(hipMemcpyAsync(h, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  *res = h[0];
  return PSFUN_SUCCESS;
}

extern "C" int psfun_dot(psfun_ctx_t c, int64_t n, const double *x, const double *y, double *r) {
  return reduce_to_host(c, 0, n, x, y, r);
}
extern "C" int psfun_nrm2(psfun_ctx_t c, int64_t n, const double *x, double *r) {
  return reduce_to_host(c, 1, n, x, nullptr, r);
}
extern "C" int psfun_amax(psfun_ctx_t c, int64_t n, const double *x, double *r) {
  return reduce_to_host(c, 2, n, x, nullptr, r);
}

__global__ __launch_bounds__(PSFUN_BS) void k_axpby(int64_t n, double alpha, const double *__restrict__ x,
                                                    double beta, double *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS) {
    const double ax = alpha * x[i];
    y[i] = (beta == 0.0) ? ax : fma(beta, y[i], ax);
  }
}

__global__ __launch_bounds__(PSFUN_BS) void k_scal(int64_t n, double alpha, double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS)
    x[i] = alpha * x[i];
}

int launch_axpby(psfun_ctx *c, int64_t n, double alpha, const double *x, double beta, double *y) {
  if (n == 0) return PSFUN_SUCCESS;
  const int G = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 16);
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_AXPBY, (beta == 0.0 ? 16.0 : 24.0) * n));
  hipLaunchKernelGGL(k_axpby, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, alpha, x, beta, y);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));
  return PSFUN_SUCCESS;
}

extern "C" int psfun_axpby(psfun_ctx_t c, int64_t n, double alpha, const double *x, double beta, double *y) {
  if (!c || n < 0 || (n && (!x || !y))) return PSFUN_ERR_INVALID_ARG;
  return launch_axpby(c, n, alpha, x, beta, y);
}

extern "C" int psfun_scal(psfun_ctx_t c, int64_t n, double alpha, double *x) {
  if (!c || n < 0 || (n && !x)) return PSFUN_ERR_INVALID_ARG;
  if (n == 0) return PSFUN_SUCCESS;
  const int G = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  hipLaunchKernelGGL(k_scal, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, alpha, x);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}
