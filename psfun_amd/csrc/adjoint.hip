// Adjoint helpers for the differentiable exp(-tA)X (no reference
// counterpart).  The derivative of L = <G, exp(-tA) X> with respect to the
// stored nonzeros of a symmetric A follows from the resolvent form
// exp(-tA) ~ Re sum_s w_s (xi_s I + A)^{-1}:
//   dL/da_ik = -Re sum_s w_s sum_col u_s[i,col] v_s[k,col],
//   v_s = (xi_s I + A)^{-1} x_col,  u_s = (xi_s I + A)^{-1} g_col.
// Both resolvent families are taken from Krylov bases that the Lanczos runs
// already built (the forward one on X, the adjoint one on G): in the basis
// V_m of K_m(A, x) the Galerkin solution is v_s = ||x|| V_m (xi_s I + T_m)^{-1} e1,
// which is the multi-shift CG iterate of step m.  No further solves over n.
// The stored tridiagonal belongs to the operator the Lanczos run used
// (opscale*A); tscale = 1/opscale maps it back to A in the same basis.
//
//   k_fom_coef:          per (instance, shift) complex tridiagonal solve of
//                        order m (Thomas), scaled by ||x||.
//   k_krylov_resolvents: out_s[r, col] = sum_l coef[s][l][col] V_l[r, col]
//                        (one pass over the basis for all shifts).
//   k_sddmm_cols:        grad[q] += scale * sum_s Re(w_s sum_col U_s[i,col] V_s[k,col]).
#include "common.hpp"
#include "kernels.hpp"

#define ADJ_SMAX 32

// coef layout: [(s * mmax + l) * K + col] complex (re, im pairs).
__global__ void k_fom_coef(int K, int nshift, int mmax, const double *__restrict__ T_alpha,
                           const double *__restrict__ T_beta, int ldt, double tscale, const int *__restrict__ nbasis,
                           const double *__restrict__ rs1, const double *__restrict__ shifts, double *__restrict__ coef,
                           double *__restrict__ scratch) {
  const int col = blockIdx.x;  // instance = column (one matrix)
  const int s = threadIdx.x;
  if (s >= nshift) return;
  const int m = nbasis[col];
  const double *al = T_alpha + (int64_t)col * ldt + 1;  // alpha_1..alpha_m
  const double *be = T_beta + (int64_t)col * ldt + 1;   // beta_1..beta_{m-1}
  const double xr = shifts[2 * s], xi = shifts[2 * s + 1];
  // scratch: c'_l and rhs'_l, complex, per (col, s)
  double *cp = scratch + ((int64_t)col * nshift + s) * 4 * mmax;
  double *rp = cp + 2 * mmax;
  double pcr = 0.0, pci = 0.0, prr = 0.0, pri = 0.0;  // previous c', rhs'
  for (int l = 0; l < m; ++l) {
    // den = (xi + alpha_l) - beta_{l-1} c'_{l-1}
    double dr = fma(tscale, al[l], xr), di = xi;
    const double bm = (l > 0) ? tscale * be[l - 1] : 0.0;  // T(l-1, l) of A
    if (l > 0) {
      dr -= bm * pcr;
      di -= bm * pci;
    }
    const double dd = dr * dr + di * di;
    const double ir = dr / dd, ii = -di / dd;  // 1/den
    // c'_l = beta_l / den ; rhs'_l = (delta_l0 - beta_{l-1} rhs'_{l-1}) / den
    const double bl = (l < m - 1) ? tscale * be[l] : 0.0;
    const double ncr = bl * ir, nci = bl * ii;
    double nr = (l == 0) ? 1.0 : -bm * prr;
    double ni = (l == 0) ? 0.0 : -bm * pri;
    const double nrr = nr * ir - ni * ii, nri = nr * ii + ni * ir;
    cp[2 * l] = ncr;
    cp[2 * l + 1] = nci;
    rp[2 * l] = nrr;
    rp[2 * l + 1] = nri;
    pcr = ncr;
    pci = nci;
    prr = nrr;
    pri = nri;
  }
  const double sc = rs1[col];
  double yr = 0.0, yi = 0.0;
  for (int l = m - 1; l >= 0; --l) {
    // y_l = rhs'_l - c'_l y_{l+1}
    const double cr = cp[2 * l], ci = cp[2 * l + 1];
    const double nyr = rp[2 * l] - (cr * yr - ci * yi), nyi = rp[2 * l + 1] - (cr * yi + ci * yr);
    yr = nyr;
    yi = nyi;
    // the stored basis is unnormalised: v_{l+1} = slot l / beta_l (beta_0 = rs1)
    const double bl = (l == 0) ? sc : be[l - 1];
    const double ib = (bl != 0.0) ? 1.0 / bl : 0.0;
    double *o = coef + 2 * (((int64_t)s * mmax + l) * K + col);
    o[0] = sc * yr * ib;
    o[1] = sc * yi * ib;
  }
  for (int l = m; l < mmax; ++l) {
    double *o = coef + 2 * (((int64_t)s * mmax + l) * K + col);
    o[0] = 0.0;
    o[1] = 0.0;
  }
}

// out[(s*nrows + r)*K + col] (complex) = sum_{l < mmax} coef[s][l][col] * V[l][r][col]
template <int S>
__global__ __launch_bounds__(PSFUN_BS) void k_krylov_resolvents(int64_t nrows, int K, int mmax, int nshift,
                                                                const double *__restrict__ V,
                                                                const double *__restrict__ coef,
                                                                double2 *__restrict__ out) {
  const int64_t tot = nrows * K;
  const int64_t vs = tot;
  for (int64_t idx = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * PSFUN_BS) {
    const int col = (int)(idx % K);
    double ar[S], ai[S];
#pragma unroll
    for (int s = 0; s < S; ++s) ar[s] = ai[s] = 0.0;
    for (int l = 0; l < mmax; ++l) {
      const double v = V[(int64_t)l * vs + idx];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (s < nshift) {
          const double2 c = reinterpret_cast<const double2 *>(coef)[((int64_t)s * mmax + l) * K + col];
          ar[s] = fma(c.x, v, ar[s]);
          ai[s] = fma(c.y, v, ai[s]);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (s < nshift) out[(int64_t)s * tot + idx] = make_double2(ar[s], ai[s]);
  }
}

// Same sum with R elements per thread that share one column: for K | PSFUN_BS
// the elements e, e + PSFUN_BS, ... of a tile of R*PSFUN_BS have the same
// column, so each coefficient read serves R rows (one coefficient load per V
// value made the coefficient table the limit).
// Each element's sum runs over l in the same order as k_krylov_resolvents.
#define RS_LDS 2048  // double2 coefficients staged per block (32 KiB)
template <int S, int R>
__global__ __launch_bounds__(PSFUN_BS) void k_krylov_resolvents_rows(int64_t nrows, int K, int mmax, int nshift,
                                                                     const double *__restrict__ V,
                                                                     const double *__restrict__ coef,
                                                                     double2 *__restrict__ out) {
  const int64_t tot = nrows * K;
  const int64_t vs = tot;
  const double2 *cf = reinterpret_cast<const double2 *>(coef);
  // The coefficients of lt consecutive l (all shifts, all K columns) are
  // staged in LDS by the whole block, so the four waves share one load of
  // them instead of each lane fetching its column's through the L1 at every l
  // (those loads were four fifths of the kernel's L1 accesses).
  __shared__ double2 scoef[RS_LDS];
  const int lt = RS_LDS / (S * K);  // l per staged chunk (>= 2: S*K <= 16*64)
  const int col = (int)(threadIdx.x % K);
  for (int64_t t0 = (int64_t)blockIdx.x * (R * PSFUN_BS); t0 < tot; t0 += (int64_t)gridDim.x * (R * PSFUN_BS)) {
    const int64_t e0 = t0 + threadIdx.x;
    double ar[R][S], ai[R][S];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < S; ++q) ar[r][q] = ai[r][q] = 0.0;
    for (int l0 = 0; l0 < mmax; l0 += lt) {
      const int nl = min(lt, mmax - l0);
      __syncthreads();  // the previous chunk has been read
      // for each shift the chunk is nl*K consecutive coefficients in both
      // arrays: a straight copy (no index divisions, which cost about as many
      // VALU instructions as the sums they feed)
      for (int q = 0; q < nshift; ++q) {
        const double2 *src = cf + ((int64_t)q * mmax + l0) * K;
        double2 *dst = scoef + q * lt * K;
        for (int idx = threadIdx.x; idx < nl * K; idx += PSFUN_BS) dst[idx] = src[idx];
      }
      __syncthreads();
      for (int dl = 0; dl < nl; ++dl) {
        const int l = l0 + dl;
        double v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int64_t e = e0 + (int64_t)r * PSFUN_BS;
          v[r] = (e < tot) ? V[(int64_t)l * vs + e] : 0.0;
        }
        // all S shifts unguarded (the sums of shifts >= nshift read stale LDS
        // and are never stored), each coefficient read issued two shifts
        // ahead of its sums: a branch per shift, or one register slot for
        // the coefficient, waited for every read in turn
        double2 cq[S];
        const double2 *sc = scoef + dl * K + col;
        cq[0] = sc[0];
        cq[1] = sc[lt * K];
#pragma unroll
        for (int q = 0; q < S; ++q) {
          if (q + 2 < S) cq[q + 2] = sc[(q + 2) * lt * K];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            ar[r][q] = fma(cq[q].x, v[r], ar[r][q]);
            ai[r][q] = fma(cq[q].y, v[r], ai[r][q]);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t e = e0 + (int64_t)r * PSFUN_BS;
      if (e < tot) {
#pragma unroll
        for (int q = 0; q < S; ++q)
          if (q < nshift) out[(int64_t)q * tot + e] = make_double2(ar[r][q], ai[r][q]);
      }
    }
  }
}

#define SD_QMAX 8  // row length up to which the SDDMM keeps a row's sums in registers
template <int KP>
__global__ __launch_bounds__(PSFUN_BS) void k_sddmm_cols(int n, const int *__restrict__ rowptr,
                                                         const int *__restrict__ col, int K, int nshift,
                                                         const double *__restrict__ w, const double2 *__restrict__ U,
                                                         const double2 *__restrict__ V, double scale,
                                                         double *__restrict__ grad) {
  constexpr int GPB = PSFUN_BS / KP;  // nonzero groups per block pass
  __shared__ double sw[ADJ_SMAX][2];
  if (threadIdx.x < nshift) {
    sw[threadIdx.x][0] = w[2 * threadIdx.x];
    sw[threadIdx.x][1] = w[2 * threadIdx.x + 1];
  }
  __syncthreads();
  const int lane = threadIdx.x % KP, grp = threadIdx.x / KP;
  const int64_t tot = (int64_t)n * K;
  // rows are handled group-by-group; each KP-lane group owns one row at a time
  for (int64_t i = (int64_t)blockIdx.x * GPB + grp; i < n; i += (int64_t)gridDim.x * GPB) {
    const int q0 = rowptr[i], q1 = rowptr[i + 1];
    if (q1 - q0 <= SD_QMAX) {
      // rows of up to SD_QMAX entries: U_s[i,:] is loaded once per shift and
      // serves all of the row's entries; each entry's sum still runs over s
      // in order, as below
      int kq[SD_QMAX];
      double acc[SD_QMAX];
#pragma unroll
      for (int m = 0; m < SD_QMAX; ++m) {
        kq[m] = (q0 + m < q1) ? col[q0 + m] : 0;
        acc[m] = 0.0;
      }
      if (lane < K) {
        for (int s = 0; s < nshift; ++s) {
          const double2 a = U[(int64_t)s * tot + i * K + lane];
          const double wr = sw[s][0], wi = sw[s][1];
#pragma unroll
          for (int m = 0; m < SD_QMAX; ++m) {
            if (q0 + m < q1) {
              const double2 b = V[(int64_t)s * tot + (int64_t)kq[m] * K + lane];
              const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
              acc[m] += wr * pr - wi * pi;
            }
          }
        }
      }
#pragma unroll
      for (int m = 0; m < SD_QMAX; ++m) {
        double v = acc[m];
#pragma unroll
        for (int off = KP / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, KP);
        if (lane == 0 && q0 + m < q1) grad[q0 + m] = fma(scale, v, grad[q0 + m]);
      }
      continue;
    }
    for (int q = q0; q < q1; ++q) {
      const int k = col[q];
      double acc = 0.0;
      if (lane < K) {
        for (int s = 0; s < nshift; ++s) {
          const double2 a = U[(int64_t)s * tot + i * K + lane], b = V[(int64_t)s * tot + (int64_t)k * K + lane];
          const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
          acc += sw[s][0] * pr - sw[s][1] * pi;
        }
      }
#pragma unroll
      for (int off = KP / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, KP);
      if (lane == 0) grad[q] = fma(scale, acc, grad[q]);
    }
  }
}

extern "C" int psfun_krylov_resolvents(psfun_ctx_t c, int64_t nrows, int K, int mmax, const double *V,
                                       const double *T_alpha, const double *T_beta, int ldt, double tscale,
                                       const int *nbasis,
                                       const double *rs1, int nshift, const double *shifts, double *out) {
  if (!c || nrows < 0 || K < 1 || K > 64 || mmax < 0 || nshift < 0 || nshift > ADJ_SMAX || ldt < mmax + 1)
    return PSFUN_ERR_INVALID_ARG;
  if (nrows == 0 || nshift == 0 || mmax == 0) return PSFUN_SUCCESS;
  if (!V || !T_alpha || !T_beta || !nbasis || !rs1 || !shifts || !out) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  // workspace: shifts, coefficient table, Thomas scratch
  const size_t coef_n = 2 * (size_t)nshift * mmax * K;
  const size_t scr_n = 4 * (size_t)K * nshift * mmax;
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, sizeof(double) * (2 * ADJ_SMAX + coef_n + scr_n), &ws));
  double *dsh = (double *)ws, *coef = dsh + 2 * ADJ_SMAX, *scr = coef + coef_n;
  PSFUN_HIP_TRY(hipMemcpyAsync(dsh, shifts, sizeof(double) * 2 * nshift, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_fom_coef, dim3(K), dim3(ADJ_SMAX), 0, c->stream, K, nshift, mmax, T_alpha, T_beta, ldt, tscale,
                     nbasis,
                     rs1, (const double *)dsh, coef, scr);
  PSFUN_HIP_TRY(hipGetLastError());
  const int G = psfun_grid(c, (nrows * K + PSFUN_BS - 1) / PSFUN_BS, 8);
  if (PSFUN_BS % K == 0 && nshift <= 16) {
    // several rows per thread share each coefficient load
    if (nshift <= 8) {
      const int Gr = psfun_grid(c, (nrows * K + 4 * PSFUN_BS - 1) / (4 * PSFUN_BS), 8);
      hipLaunchKernelGGL((k_krylov_resolvents_rows<8, 4>), dim3(Gr), dim3(PSFUN_BS), 0, c->stream, nrows, K, mmax,
                         nshift, V, (const double *)coef, (double2 *)out);
    } else {
      const int Gr = psfun_grid(c, (nrows * K + 2 * PSFUN_BS - 1) / (2 * PSFUN_BS), 8);
      hipLaunchKernelGGL((k_krylov_resolvents_rows<16, 2>), dim3(Gr), dim3(PSFUN_BS), 0, c->stream, nrows, K, mmax,
                         nshift, V, (const double *)coef, (double2 *)out);
    }
  } else if (nshift <= 8)
    hipLaunchKernelGGL(k_krylov_resolvents<8>, dim3(G), dim3(PSFUN_BS), 0, c->stream, nrows, K, mmax, nshift, V,
                       (const double *)coef, (double2 *)out);
  else if (nshift <= 16)
    hipLaunchKernelGGL(k_krylov_resolvents<16>, dim3(G), dim3(PSFUN_BS), 0, c->stream, nrows, K, mmax, nshift, V,
                       (const double *)coef, (double2 *)out);
  else
    hipLaunchKernelGGL(k_krylov_resolvents<ADJ_SMAX>, dim3(G), dim3(PSFUN_BS), 0, c->stream, nrows, K, mmax, nshift,
                       V, (const double *)coef, (double2 *)out);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

extern "C" int psfun_sddmm_cols(psfun_ctx_t c, int n, const int32_t *rowptr, const int32_t *colind, int K, int nshift,
                                const double *weights, const double *U, const double *V, double scale, double *grad) {
  if (!c || n < 0 || K < 1 || K > 64 || nshift < 0 || nshift > ADJ_SMAX || (nshift && !weights))
    return PSFUN_ERR_INVALID_ARG;
  if (n == 0 || nshift == 0) return PSFUN_SUCCESS;
  if (!rowptr || !colind || !U || !V || !grad) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipMemcpyAsync(c->d_scalar, weights, sizeof(double) * 2 * nshift, hipMemcpyHostToDevice, c->stream));
  int KP = 1;
  while (KP < K) KP <<= 1;
  const int G = psfun_grid(c, ((int64_t)n * KP + PSFUN_BS - 1) / PSFUN_BS, 8);
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_SDDMM, 0.0));
#define SD_CASE(KPV)                                                                                           \
  case KPV:                                                                                                    \
    hipLaunchKernelGGL((k_sddmm_cols<KPV>), dim3(G), dim3(PSFUN_BS), 0, c->stream, n, rowptr, colind, K, nshift, \
                       (const double *)c->d_scalar, (const double2 *)U, (const double2 *)V, scale, grad);      \
    break;
  switch (KP) {
    SD_CASE(1) SD_CASE(2) SD_CASE(4) SD_CASE(8) SD_CASE(16) SD_CASE(32) SD_CASE(64)
    default: return PSFUN_ERR_INVALID_ARG;
  }
#undef SD_CASE
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(psfun_prof_end(c));
  return PSFUN_SUCCESS;
}
