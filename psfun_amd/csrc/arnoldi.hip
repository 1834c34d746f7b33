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
//                   (gathering w*(1/h) on the fly).  Block 0 is the control block: it runs the
//                   reference's control for loop index i = j (the a-posteriori
//                   estimate from f(H_{j-1}) e1, by the dense block exp of
//                   fdense.hpp) while the streaming blocks work.
//   k_ar_gram, k_ar_hsolve, k_ar_lsu: the column's modified Gram-Schmidt in
//                   its one-reduction form (below): h(1:j, j), w in place of
//                   u, partials of ||w||^2.
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
                                                                 int j, int nl, double eps, double fscale, int *hf,
                                                                 int pade, double *S) {
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
    block_expmv_var(k, H, ldh, model, fscale, pade, S, &W);
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

// pade > 0: the GENPADE step by the Pade approximant of that degree (always
// the block kernel), with its 5 nl^2 work doubles at S.
static int ar_launch_est(psfun_ctx *c, hipStream_t st, LzCtl *ctl, const double *H, int ldh, int model, int j, int nl,
                         double eps, double fscale, int pade, double *S) {
  if (j - 1 <= PSFUN_WDENSE_K && pade == 0)
    hipLaunchKernelGGL(k_ar_est<true>, dim3(1), dim3(64), 0, st, ctl, H, ldh, model, j, nl, eps, fscale, c->d_flags,
                       pade, S);
  else
    hipLaunchKernelGGL(k_ar_est<false>, dim3(1), dim3(PSFUN_BS), 0, st, ctl, H, ldh, model, j, nl, eps, fscale,
                       c->d_flags, pade, S);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

// The estimate of loop index j reads only H(1:j-1, 1:j-1) and h(j, j-1),
// which neither the orthogonalisation of column j nor the SpMV step of index
// j+1 writes, so it runs on the auxiliary stream beside both: forked after
// the SpMV step of index j, joined before the orthogonalisation of column j+1.
// A stop it finds leaves that one SpMV step's work unused, and the step's
// control block may have advanced itx past the stop index; the host then
// takes itx = m (the reference exits with itx = m on convergence).
static int ar_fork_est(psfun_ctx *c, LzCtl *ctl, const double *H, int ldh, int model, int j, int nl, double eps,
                       double fscale, int pade, double *S) {
  PSFUN_HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
  PSFUN_HIP_TRY(hipStreamWaitEvent(c->aux_stream, c->ev_fork, 0));
  PSFUN_TRY(ar_launch_est(c, c->aux_stream, ctl, H, ldh, model, j, nl, eps, fscale, pade, S));
  PSFUN_HIP_TRY(hipEventRecord(c->ev_join, c->aux_stream));
  return PSFUN_SUCCESS;
}
static int ar_join_est(psfun_ctx *c) {
  PSFUN_HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return PSFUN_SUCCESS;
}

template <int L, bool SELL>
__global__ __launch_bounds__(PSFUN_BS) void k_ar_spmv(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                                      const double *__restrict__ val, double scaling,
                                                      const double *__restrict__ src, int j, int G,
                                                      const double *__restrict__ p_in, int g_in,
                                                      double *__restrict__ V, int64_t ldv, double *__restrict__ u_out,
                                                      LzCtl *ctl, double *H, int ldh,
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
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    double sum = 0.0;
    if (SELL) {
      if (row < n)
        sum = (L == PSFUN_L_HDIA) ? hdia_row_dot<true>(row, n, rp, col, val, src, inv)
                                  : sell_row_dot(row, rp, col, val, src, inv);
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
    }
  }
}

// ---------------------------------------------------------------------------
// Orthogonalisation of column j against v_1..v_j.  The reference runs modified
// Gram-Schmidt (psfun_d_arnoldi.f90:209-213, psfun_d_sai.f90 likewise):
// for k = 1..j: h(k,j) = <v_k, w>; w -= h(k,j) v_k.  Substituting the earlier
// updates into each inner product gives
//   h(k,j) = <v_k, u> - sum_{i<k} <v_k, v_i> h(i,j)      (u = the SpMV output),
// i.e. h = (I + L)^{-1} V^T u with L the strictly lower part of the basis Gram
// matrix: the "inverse compact WY" form of MGS (Swirydowicz, Langou,
// Ananthan, Yang, Thomas, "Low synchronization Gram-Schmidt and generalized
// minimal residual algorithms", Numer. Linear Algebra Appl. 28 (2021) e2343),
// which keeps MGS's bound on the loss of orthogonality.  One reduction per
// column instead of j:
//   k_ar_gram:   partials of <v_l, u> (l = 1..j) and <v_l, v_j> (l < j);
//   k_ar_hsolve: one block folds them in a fixed order, keeps row j of the
//                Gram matrix and solves (I + L) h = V^T u by forward
//                substitution; h is column j of H;
//   k_ar_lsu:    w = u - h(1,j) v_1 - ... - h(j,j) v_j, row by row in that
//                order (the reference's sequence of updates), and the
//                partials of ||w||^2.
// Each basis vector is read twice per column instead of twice per MGS step.
// ---------------------------------------------------------------------------

// Partial sums of one gram launch over basis vectors l0 <= l < l1.  A block
// takes tiles of AR_GR*PSFUN_BS rows (u and v_j of the tile in registers),
// streams v_l of the tile for every l (the next v_l's loads issued before the
// current one is reduced), reduces both products over the wavefront and adds
// them to per-wave LDS sums; the block's sums go to the partials, list-major:
// list l-1 holds <v_l, u>, list j+l-1 holds <v_l, v_j>.  Fixed order
// throughout, so the partials do not depend on timing.
#define AR_GR 8
#define AR_LG 128
__global__ __launch_bounds__(PSFUN_BS) void k_ar_gram(int n, const double *__restrict__ V, int64_t ldv, int j,
                                                      int l0, int l1, const double *__restrict__ u,
                                                      double *__restrict__ P, int Gg, const LzCtl *ctl) {
  __shared__ double az[PSFUN_BS / 64][AR_LG], ag[PSFUN_BS / 64][AR_LG];
  if (ctl->done) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nv = l1 - l0;
  for (int i = lane; i < nv; i += 64) {
    az[wv][i] = 0.0;
    ag[wv][i] = 0.0;
  }
  wave_lds_sync();  // lane 0 of each wave owns its row of sums from here on
  const double *vj = V + (int64_t)(j - 1) * ldv;
  constexpr int64_t TILE = (int64_t)AR_GR * PSFUN_BS;
  const int64_t ntiles = ((int64_t)n + TILE - 1) / TILE;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r0 = t * TILE + threadIdx.x;
    double uu[AR_GR], jj[AR_GR], a[AR_GR];
#pragma unroll
    for (int q = 0; q < AR_GR; ++q) {
      const int64_t r = r0 + q * PSFUN_BS;
      uu[q] = r < n ? u[r] : 0.0;
      jj[q] = r < n ? vj[r] : 0.0;
      a[q] = r < n ? V[(int64_t)(l0 - 1) * ldv + r] : 0.0;
    }
    for (int l = l0; l < l1; ++l) {
      double b[AR_GR];
      const double *vn = V + (int64_t)l * ldv;  // v_{l+1}
#pragma unroll
      for (int q = 0; q < AR_GR; ++q) {
        const int64_t r = r0 + q * PSFUN_BS;
        b[q] = (l + 1 < l1 && r < n) ? vn[r] : 0.0;
      }
      double z = 0.0, g = 0.0;
#pragma unroll
      for (int q = 0; q < AR_GR; ++q) {
        z = fma(a[q], uu[q], z);
        g = fma(a[q], jj[q], g);
      }
      z = wave_sum_dpp(z);
      g = wave_sum_dpp(g);
      if (lane == 0) {
        az[wv][l - l0] += z;
        ag[wv][l - l0] += g;
      }
#pragma unroll
      for (int q = 0; q < AR_GR; ++q) a[q] = b[q];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nv; i += PSFUN_BS) {
    const int l = l0 + i;
    double z = 0.0, g = 0.0;
#pragma unroll
    for (int w = 0; w < PSFUN_BS / 64; ++w) {
      z += az[w][i];
      g += ag[w][i];
    }
    P[(int64_t)(l - 1) * Gg + blockIdx.x] = z;
    if (l < j) P[(int64_t)(j + l - 1) * Gg + blockIdx.x] = g;
  }
}

// Fold of the gram partials and the triangular solve for h(1:j, j), one block.
// Gm keeps the Gram matrix's strictly lower part, G(k,i) = <v_k, v_i> at
// Gm[(i-1)*ldg + (k-1)] (row k written at column k).  Forward substitution by
// columns in one wavefront: h_i = z_i - acc_i, then acc_k += G(k,i) h_i for
// k > i, the Gram columns staged through LDS AR_HS_STAGE doubles at a time.
#define AR_HS_BS 1024
#define AR_HS_STAGE 4096
#define AR_HS_QB 16  // lists whose loads are in flight together in the fold
__global__ __launch_bounds__(AR_HS_BS) void k_ar_hsolve(int j, const double *__restrict__ P, int Gg,
                                                        double *__restrict__ Gm, int ldg, const LzCtl *ctl, double *H,
                                                        int ldh) {
  __shared__ double z[LZ_NL_MAX], gj[LZ_NL_MAX], acc[LZ_NL_MAX], h[LZ_NL_MAX];
  __shared__ double gs[AR_HS_STAGE];
  if (ctl->done) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int NW = AR_HS_BS / 64;
  // Fold: thread t takes partial t (and t + AR_HS_BS, ...) of AR_HS_QB lists at
  // a time, all its loads issued before any sum (the partials come from
  // blocks on every XCD, so each round trip is long); each list is then
  // summed over the wavefront (DPP) and over the 16 wavefronts in order.
  __shared__ double wsum[NW][64];
  const int nlist = 2 * j - 1;
  for (int q0 = 0; q0 < nlist; q0 += 64) {
    const int nq = min(64, nlist - q0);
    for (int qb = 0; qb < nq; qb += AR_HS_QB) {
      double v[AR_HS_QB];
      if (Gg <= AR_HS_BS) {  // one partial per thread and list: the loads issue together
#pragma unroll
        for (int r = 0; r < AR_HS_QB; ++r)
          v[r] = (qb + r < nq && (int)threadIdx.x < Gg) ? P[(int64_t)(q0 + qb + r) * Gg + threadIdx.x] : 0.0;
      } else {
#pragma unroll
        for (int r = 0; r < AR_HS_QB; ++r) {
          v[r] = 0.0;
          if (qb + r < nq)
            for (int b = threadIdx.x; b < Gg; b += AR_HS_BS) v[r] += P[(int64_t)(q0 + qb + r) * Gg + b];
        }
      }
#pragma unroll
      for (int r = 0; r < AR_HS_QB; ++r) {
        const double t = wave_sum_dpp(v[r]);
        if (lane == 0 && qb + r < nq) wsum[wv][qb + r] = t;
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nq; q += AR_HS_BS) {
      double t = 0.0;
      for (int w = 0; w < NW; ++w) t += wsum[w][q];
      if (q0 + q < j)
        z[q0 + q] = t;
      else
        gj[q0 + q - j] = t;
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < j; k += AR_HS_BS) acc[k] = 0.0;
  __syncthreads();
  for (int q = threadIdx.x; q < j - 1; q += AR_HS_BS) Gm[(int64_t)q * ldg + (j - 1)] = gj[q];  // row j
  if (j <= 64) {
    // all columns staged at once (j*j <= AR_HS_STAGE); lane k of wavefront 0
    // keeps z_k and acc_k in registers and h_i comes from lane i, so a step is
    // one broadcast and one fma (the sums run in the same order as below)
    for (int e = threadIdx.x; e < j * j; e += AR_HS_BS) {
      const int i = e / j, k = e % j;
      double g = 0.0;
      if (k > i) g = (k == j - 1) ? gj[i] : Gm[(int64_t)i * ldg + k];
      gs[e] = g;
    }
    __syncthreads();
    if (wv == 0) {
      const bool own = lane < j;
      const double zk = own ? z[lane] : 0.0;
      double acck = 0.0, hk = 0.0;
      for (int i = 0; i < j; ++i) {
        const double g = (own && lane > i) ? gs[i * j + lane] : 0.0;
        const double hi = wave_read_f64(zk - acck, i);
        if (lane == i) hk = hi;
        acck = fma(g, hi, acck);
      }
      if (own) H[(int64_t)(j - 1) * ldh + lane] = hk;  // h(lane+1, j)
    }
    return;
  }
  const int cw = (j * j <= AR_HS_STAGE) ? j : AR_HS_STAGE / j;
  for (int i0 = 0; i0 < j; i0 += cw) {
    const int i1 = min(j, i0 + cw);
    for (int e = threadIdx.x; e < (i1 - i0) * j; e += AR_HS_BS) {
      const int i = i0 + e / j, k = e % j;
      double g = 0.0;
      if (k > i) g = (k == j - 1) ? gj[i] : Gm[(int64_t)i * ldg + k];
      gs[e] = g;
    }
    __syncthreads();
    if (wv == 0) {
      for (int i = i0; i < i1; ++i) {
        const double hi = z[i] - acc[i];
        if (lane == 0) h[i] = hi;
        for (int k = i + 1 + lane; k < j; k += 64) acc[k] = fma(gs[(i - i0) * j + k], hi, acc[k]);
        wave_lds_sync();
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < j; k += AR_HS_BS) H[(int64_t)(j - 1) * ldh + k] = h[k];  // h(k+1, j)
}

// w = u - sum_{l=1..j} h(l,j) v_l in place (row by row, l ascending) and the
// partials of ||w||^2.  Rows in pairs (16-byte loads; w and the slots are
// 16-byte aligned), the basis with nontemporal loads; the odd last row goes
// to thread 0 of the last block.
__global__ __launch_bounds__(PSFUN_BS) void k_ar_lsu(int n, int j, const double *__restrict__ V, int64_t ldv,
                                                     double *__restrict__ w, const double *__restrict__ H, int ldh,
                                                     double *__restrict__ p_out, const LzCtl *ctl) {
  __shared__ double h[LZ_NL_MAX];
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->done) return;
  for (int k = threadIdx.x; k < j; k += PSFUN_BS) h[k] = H[(int64_t)(j - 1) * ldh + k];
  __syncthreads();
  const int64_t n2 = n >> 1;
  psfun_d2 *w2 = (psfun_d2 *)w;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n2; i += (int64_t)gridDim.x * PSFUN_BS) {
    psfun_d2 x = w2[i];
#pragma unroll 8
    for (int l = 0; l < j; ++l) {
      const psfun_d2 v = __builtin_nontemporal_load((const psfun_d2 *)(V + (int64_t)l * ldv) + i);
      x.x = fma(-h[l], v.x, x.x);
      x.y = fma(-h[l], v.y, x.y);
    }
    w2[i] = x;
    acc = fma(x.x, x.x, acc);
    acc = fma(x.y, x.y, acc);
  }
  if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    const int64_t i = n - 1;
    double x = w[i];
    for (int l = 0; l < j; ++l) x = fma(-h[l], V[(int64_t)l * ldv + i], x);
    w[i] = x;
    acc = fma(x, x, acc);
  }
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = acc;
}

// Host side: orthogonalise the device vector w (u on entry) against basis
// slots 1..j; leaves ||w||^2 partials (gl of them) in p_w.
struct ArOrtho {
  double *Pg;   // gram partials, 2*nl lists of Gg
  double *Gm;   // Gram matrix, nl x nl
  int Gg, Gl, ldg;
};
static size_t ar_ortho_bytes(const psfun_ctx *c, int n, int nl) {
  const int64_t tiles = ((int64_t)n + AR_GR * PSFUN_BS - 1) / (AR_GR * PSFUN_BS);
  const int Gg = psfun_grid(c, tiles, 4);
  return align_up(sizeof(double) * 2 * (size_t)nl * Gg, 256) + align_up(sizeof(double) * (size_t)nl * nl, 256);
}
static ArOrtho ar_ortho_carve(const psfun_ctx *c, void *base, int n, int nl) {
  ArOrtho o;
  const int64_t tiles = ((int64_t)n + AR_GR * PSFUN_BS - 1) / (AR_GR * PSFUN_BS);
  o.Gg = psfun_grid(c, tiles, 4);
  o.Gl = psfun_grid(c, ((int64_t)(n >> 1) + PSFUN_BS - 1) / PSFUN_BS, 4);
  o.ldg = nl;
  o.Pg = (double *)base;
  o.Gm = (double *)((char *)base + align_up(sizeof(double) * 2 * (size_t)nl * o.Gg, 256));
  return o;
}
static int ar_orthogonalize(psfun_ctx *c, const ArOrtho &o, int n, int j, const double *V, int64_t ldv, double *w,
                            double *p_w, LzCtl *ctl, double *H, int ldh) {
  for (int l0 = 1; l0 <= j; l0 += AR_LG) {
    const int l1 = min(j + 1, l0 + AR_LG);
    hipLaunchKernelGGL(k_ar_gram, dim3(o.Gg), dim3(PSFUN_BS), 0, c->stream, n, V, ldv, j, l0, l1, w, o.Pg, o.Gg,
                       ctl);
  }
  hipLaunchKernelGGL(k_ar_hsolve, dim3(1), dim3(AR_HS_BS), 0, c->stream, j, o.Pg, o.Gg, o.Gm, o.ldg, ctl, H, ldh);
  hipLaunchKernelGGL(k_ar_lsu, dim3(o.Gl), dim3(PSFUN_BS), 0, c->stream, n, j, V, ldv, w, H, ldh, p_w, ctl);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
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
                                                       double fscale, int pade, double *S) {
  __shared__ DenseLds W;
  const int done = ctl->done;
  if (done == LZ_DONE_CONVERGED || done == LZ_DONE_ZERO) return;
  int k = ctl->itx - 1;
  if (k < 1) k = 1;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W.y[i] = (i == 0) ? 1.0 : 0.0;
  __syncthreads();
  block_expmv_var(k, H, ldh, model, fscale, pade, S, &W);
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
  const size_t vbytes = align_up(sizeof(double) * (size_t)ldv * (size_t)(nl + 2), 256);  // v_1..v_{nl-1}, u0, u1
  const int pade = (!userf && variant == PSFUN_VAR_GENPADE && c->pade_degree != 6) ? c->pade_degree : 0;
  const size_t pade_bytes = pade ? align_up(sizeof(double) * 5 * (size_t)nl * nl, 256) : 0;
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, ctl_bytes + h_bytes + vbytes + ar_ortho_bytes(c, n, nl) + pade_bytes, &ws));
  double *S = pade ? (double *)((char *)ws + ctl_bytes + h_bytes + vbytes + ar_ortho_bytes(c, n, nl)) : nullptr;
  LzCtl *ctl = (LzCtl *)ws;
  double *H = (double *)((char *)ws + ctl_bytes);
  double *V = (double *)((char *)ws + ctl_bytes + h_bytes);
  double *U[2] = {V + (size_t)ldv * nl, V + (size_t)ldv * (nl + 1)};
  double *P[2] = {c->d_partials, c->d_partials + PSFUN_MAX_GRID};
  const ArOrtho orth = ar_ortho_carve(c, (char *)ws + ctl_bytes + h_bytes + vbytes, n, nl);

  const bool sell = psfun_reads_sell(A);
  const int L = psfun_row_per_thread(A) ? 1 : A->lanes;
  const int64_t chunks = ((int64_t)n + PSFUN_BS / L - 1) / (PSFUN_BS / L);
  // diagonal slices: 6 streaming blocks per CU measured about 2% faster per
  // f(A)x than 8 at 128^3
  int Gs = psfun_grid(c, chunks, psfun_reads_hdia(A) ? 6 : (psfun_row_per_thread(A) ? 8 : 4));
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
  bool est_pending = false;
  for (int j = 1; j <= nl - 1; ++j) {
    const double *src = (j == 1) ? x : U[j & 1];
    double *u = U[(j + 1) & 1];
    PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_AR_SPMV, spmv_bytes, j));
    if (psfun_reads_hdia(A)) {
      hipLaunchKernelGGL((k_ar_spmv<PSFUN_L_HDIA, true>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n,
                         A->hdia_ptr, A->hdia_off, A->hdia_val, scaling, src, j, Gs, P[cur], gcur, V, ldv, u, ctl, H,
                         ldh, model, nl_k, istop_k, eps, fscale, c->d_flags);
    } else if (sell) {
      hipLaunchKernelGGL((k_ar_spmv<1, true>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n, A->sell_ptr,
                         A->sell_col, A->sell_val, scaling, src, j, Gs, P[cur], gcur, V, ldv, u, ctl, H, ldh,
                         model, nl_k, istop_k, eps, fscale, c->d_flags);
    } else {
#define AR_LAUNCH(LL)                                                                                               \
  hipLaunchKernelGGL((k_ar_spmv<LL, false>), dim3(Gs + 1), dim3(PSFUN_BS), 0, c->stream, n, A->rowptr, A->colind, \
                     A->val, scaling, src, j, Gs, P[cur], gcur, V, ldv, u, ctl, H, ldh, model, nl_k, istop_k, eps, \
                     fscale, c->d_flags)
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
    if (est_pending) {  // the estimate of index j-1 ends before column j's work
      PSFUN_TRY(ar_join_est(c));
      est_pending = false;
    }
    if (istop_k == 1 && j >= 2) {
      PSFUN_TRY(ar_fork_est(c, ctl, H, ldh, model, j, nl_k, eps, fscale, pade, S));
      est_pending = true;
    }
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
    // orthogonalise u against v_1..v_j: gram pass (u and j slots read),
    // update pass (j slots and u read, w written)
    PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_AR_MGS, (2.0 * j + 3.0) * 8.0 * n, j));
    PSFUN_TRY(ar_orthogonalize(c, orth, n, j, V, ldv, u, P[1 - cur], ctl, H, ldh));
    PSFUN_TRY(psfun_prof_end(c));
    cur = 1 - cur;
    gcur = orth.Gl;
    if (!fsync && j >= 2 && j < nl - 1) {
      PSFUN_TRY(psfun_host_wait(c, j - 1));
      if (c->h_flags[1] != 0) {
        ran_all = false;
        break;
      }
    }
  }
  if (est_pending) PSFUN_TRY(ar_join_est(c));
  if (ran_all) {
    hipLaunchKernelGGL(k_ar_ctrl, dim3(1), dim3(PSFUN_BS), 0, c->stream, P[cur], gcur, nl, ctl, H, ldh, model,
                       nl_k, istop_k, eps, fscale, c->d_flags);
    PSFUN_HIP_TRY(hipGetLastError());
    if (istop_k == 1) PSFUN_TRY(ar_launch_est(c, c->stream, ctl, H, ldh, model, nl, nl_k, eps, fscale, pade, S));
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
    hipLaunchKernelGGL(k_ar_final, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, H, ldh, model, fscale, pade, S);
    PSFUN_HIP_TRY(hipGetLastError());
  }
  const int Ga = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  hipLaunchKernelGGL(k_ar_assemble, dim3(Ga), dim3(PSFUN_BS), 0, c->stream, n, V, ldv, ctl, y);
  PSFUN_HIP_TRY(hipGetLastError());

  LzCtl *h = (LzCtl *)((char *)c->h_pinned + 256);
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(LzCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  const int itx = (h->done == LZ_DONE_CONVERGED) ? h->m : h->itx;
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

// yk = f(H_k^{-1} - tau I) e1 into W->y; false when H_k is singular.
__device__ bool sai_fproj(int k, const double *H, int ldh, int model, double *M, double *Hi, double tau,
                          double fscale, int pade, double *S, DenseLds *W) {
  if (!block_gj_inverse(k, H, ldh, M, Hi, W)) return false;
  for (int d = threadIdx.x; d < k; d += PSFUN_BS) Hi[d * k + d] -= tau;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W->y[i] = (i == 0) ? 1.0 : 0.0;
  __syncthreads();
  block_expmv_var(k, Hi, k, model, fscale, pade, S, W);
  return true;
}

// Reference control at loop index i (h(i,i-1) = ||w|| from the partials).
__global__ __launch_bounds__(PSFUN_BS) void k_sai_ctrl(const double *__restrict__ p_in, int g_in, int i, LzCtl *ctl,
                                                       double *H, int ldh, int model, double *M, double *Hi,
                                                       double tau, int nl, int istop, double eps, double fscale,
                                                       int pade, double *S) {
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
    if (!sai_fproj(k, H, ldh, model, M, Hi, tau, fscale, pade, S, &W)) {
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
                                                        double *Hi, double tau, double fscale, int pade, double *S) {
  __shared__ DenseLds W;
  const int done = ctl->done;
  if (done == LZ_DONE_CONVERGED || done == LZ_DONE_ZERO || done == LZ_DONE_BREAKDOWN) return;
  int k = ctl->itx - 1;
  if (k < 1) k = 1;
  if (!sai_fproj(k, H, ldh, model, M, Hi, tau, fscale, pade, S, &W)) {
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
  if (prec != PSFUN_PREC_NONE && prec != PSFUN_PREC_DIAG && prec != PSFUN_PREC_BJAC) return PSFUN_ERR_INVALID_ARG;
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
  // BJAC: the ILU(0) of the unshifted A is built once for all inner solves
  // (the reference builds its inner preconditioner on A, psfun_d_krylov_mod.F90:326-403)
  struct PrecGuard {
    psfun_prec_t P = nullptr;
    ~PrecGuard() { psfun_prec_destroy(P); }
  } pobj;
  if (prec == PSFUN_PREC_BJAC) PSFUN_TRY(psfun_prec_create(c, A, PSFUN_PREC_BJAC, &pobj.P));
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
  const size_t u_bytes = align_up(userf ? userf_scratch_bytes(nl) : 0, 256);
  const int pade = (!userf && variant == PSFUN_VAR_GENPADE && c->pade_degree != 6) ? c->pade_degree : 0;
  const size_t pade_bytes = pade ? 5 * m_bytes : 0;
  // the basis lives in the second workspace: the inner CG takes the first
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace2(
      c, ctl_bytes + h_bytes + 2 * m_bytes + vbytes + u_bytes + ar_ortho_bytes(c, n, nl) + pade_bytes, &ws));
  double *S = pade ? (double *)((char *)ws + ctl_bytes + h_bytes + 2 * m_bytes + vbytes + u_bytes +
                                ar_ortho_bytes(c, n, nl))
                   : nullptr;
  UserfScratch sc;
  if (userf) sc = userf_scratch_carve((char *)ws + ctl_bytes + h_bytes + 2 * m_bytes + vbytes, nl);
  const ArOrtho orth = ar_ortho_carve(c, (char *)ws + ctl_bytes + h_bytes + 2 * m_bytes + vbytes + u_bytes, n, nl);
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
      const int st = pobj.P ? psfun_cg_shifted_prec(c, A, pobj.P, V + (size_t)ldv * (j - 1), 1.0, tau, w, eps,
                                                    inner_itmax, inner_istop, &it_in, &err_in, nullptr)
                            : psfun_cg_shifted(c, A, prec, V + (size_t)ldv * (j - 1), 1.0, tau, w, eps,
                                               inner_itmax, inner_istop, &it_in, &err_in, nullptr);
      if (st != PSFUN_SUCCESS) return st;
      // modified Gram-Schmidt against v_1..v_j, then v_{j+1} = w / h(j+1, j)
      PSFUN_TRY(ar_orthogonalize(c, orth, n, j, V, ldv, w, P[0], ctl, H, ldh));
      hipLaunchKernelGGL(k_sai_normalize, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, w, V + (size_t)ldv * j,
                         P[0], orth.Gl, ctl, 0);
      hipLaunchKernelGGL(k_sai_ctrl, dim3(1), dim3(PSFUN_BS), 0, c->stream, P[0], orth.Gl, i, ctl, H, ldh, model, M,
                         Hi, tau, nl_k, istop_k, eps, fscale, pade, S);
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
    hipLaunchKernelGGL(k_sai_final, dim3(1), dim3(PSFUN_BS), 0, c->stream, ctl, H, ldh, model, M, Hi, tau, fscale,
                       pade, S);
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
