// Device-resident Lanczos f(A)x with the reference's semantics
// (krylov/polynomial/psfun_d_lanczos.f90:183-377).
//
// Reference loop (i = 2..nl):  itx = i-1; beta = ||w|| -> T(i-1,i) = T(i,i-1);
// exit if |beta| < 1e-16; v_i = (1/beta) w; w = A v_i; T(i,i) = w.v_i;
// w -= T(i,i) v_i; w -= beta v_{i-1}; if istop == 1: yk = f(T(1:i-1,1:i-1)) e1,
// rs(i) = rs(1) beta |yk(i-1)| / ||yk||, exit if rs(i) < eps.
// Assembly: y = rs(1) (yk(1) v_1 + sum_{l=2}^{itx-1} yk(l) v_l).
//
// Basis storage: slot l of V holds the unnormalised w_l (slot 0: x), and
// v_{l+1} = w_l * (1/beta_l) is formed wherever it is read, with the same
// 1/beta the SpMV step used, so every v value is the one a stored v_{l+1}
// would hold (beta_0 = rs1).
//
// Device schedule, per Krylov index j = 1, 2, ...:
//   k_lz_spmv(j):   every streaming block sums the partials of ||w||^2 left by
//                   the previous kernel (fixed order: identical in all blocks),
//                   beta = sqrt(.), computes u = opscale*A v_j gathering
//                   w_{j-1} and scaling by 1/beta on the fly, writes u into
//                   slot j, and leaves per-block partials of alpha_j = u.v_j.
//                   One extra control block runs the reference's loop control
//                   for index i = j (breakdown test and the a-posteriori
//                   estimate from T_{j-1}) concurrently with the streaming
//                   blocks.
//   k_lz_update(j): every block sums the alpha partials, w = u - alpha_j v_j -
//                   beta_{j-1} v_{j-1} in place in slot j, partials of ||w||^2.
// Scalars cross kernels only through kernel boundaries, so no grid-wide
// atomics or fences are needed.  When the control block ends the loop at
// index j, the SpMV/update it overlapped with feed nothing that is returned
// (the reference's own SpMV at the exit index is unused too).  Every kernel
// returns at once when ctl->done is set.  The control block also stores its
// loop index and stop state to coherent pinned host memory, on which the host
// paces its launches (no flag copies or events in the loop).
#include "common.hpp"
#include <type_traits>
#include "kernels.hpp"
#include "fexp.hpp"
#include "lanczos.hpp"
#include "userf.hpp"

// Reference loop control at index i = j (j >= 2): beta = T(j-1, j).
// Stop the loop: device state for the kernels, host flag for the launcher.
__device__ __forceinline__ void lz_stop(LzCtl *ctl, int *hf, int state) {
  ctl->done = state;
  host_flag_store(hf + 1, state);
}

// SMALLT: every T this loop can meet has order <= 64 (nl <= 65), so only the
// register f(T) routine is compiled in, which keeps the streaming kernel
// that hosts this control block within 64 VGPRs.
// (Lds: the coefficient table alone suffices for SMALLT, which keeps the
// streaming kernels that host the control block at 5.6 KB of LDS for it.)
template <bool SMALLT>
using LzCtlLds = typename std::conditional<SMALLT, ChebCoefLds, ChebLds>::type;

template <bool SMALLT>
__device__ void lz_control(LzCtl *ctl, int j, double beta, int nl, int istop, double eps, double fscale,
                           LzCtlLds<SMALLT> *W, int *hf) {
  const int lane = threadIdx.x;
  if (lane == 0) {
    host_flag_store(hf, j);
    ctl->beta[j - 1] = beta;
    ctl->itx = j - 1;
  }
  if (fabs(beta) < 1e-16) {  // psfun_d_lanczos.f90:225-230
    if (lane == 0) lz_stop(ctl, hf, LZ_DONE_BREAKDOWN);
    return;
  }
  if (istop == 1) {
    // yk = f(T(1:j-1,1:j-1)) e1; rs(j) = rs1 * beta * |yk(j-1)| / ||yk||  (lines 276-315)
    double nrm, last;
    if constexpr (SMALLT)
      wave_exp_tridiag_e1_small(j - 1, ctl->alpha + 1, ctl->beta + 1, fscale, W, ctl->yk, nrm, last);
    else
      wave_exp_tridiag_e1_lds(j - 1, ctl->alpha + 1, ctl->beta + 1, fscale, W, ctl->yk, nrm, last);
    const double errnum = ctl->rs1 * beta * fabs(last);
    const double rsj = errnum / nrm;
    if (lane == 0) {
      ctl->errnum = errnum;
      ctl->errden = nrm;
      ctl->rs[j] = rsj;
      if (rsj < eps) {
        ctl->m = j - 1;  // yk holds f(T_{j-1}) e1
        lz_stop(ctl, hf, LZ_DONE_CONVERGED);
      }
    }
    if (rsj < eps) return;
  }
  if (j >= nl && lane == 0) lz_stop(ctl, hf, LZ_DONE_MAXIT);
}

template <int L, bool SMALLT>
__global__ __launch_bounds__(PSFUN_BS) void k_lz_spmv(int n, const int *__restrict__ rowptr,
                                                      const int *__restrict__ col,
                                                      const double *__restrict__ val, double opscale,
                                                      const double *__restrict__ src, int j, int G,
                                                      const double *__restrict__ p_in, int g_in,
                                                      double *__restrict__ v_out, double *__restrict__ u_out,
                                                      double *__restrict__ p_out, LzCtl *ctl, int nl, int istop,
                                                      double eps, double fscale, int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  // block-uniform: the control block of this same grid may set the flag
  if (__syncthreads_or(ctl->done)) return;
  const double sigma = sqrt(block_sum_partials<PSFUN_BS>(p_in, g_in, lds));
  if (blockIdx.x == G) {  // control block
    if (threadIdx.x >= 64) return;
    if (j == 1) {
      if (threadIdx.x == 0) {
        host_flag_store(hf, 1);
        ctl->rs1 = sigma;
        ctl->rs[1] = sigma;
        if (sigma == 0.0) lz_stop(ctl, hf, LZ_DONE_ZERO);
      }
      return;
    }
    __shared__ LzCtlLds<SMALLT> W;
    lz_control<SMALLT>(ctl, j, sigma, nl, istop, eps, fscale, &W, hf);
    return;
  }
  if (sigma == 0.0 && j == 1) return;
  const double inv = 1.0 / sigma;
  constexpr int RPB = PSFUN_BS / L;
  const int64_t nchunks = (n + RPB - 1) / RPB;
  const XcdRange xr = xcd_range_g(nchunks, G);
  const int sub = threadIdx.x / L, lane = threadIdx.x % L;
  double dot = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * RPB + sub);
    double sum = 0.0;
    if (row < n) {
      const int s = rowptr[row], e = rowptr[row + 1];
      for (int k = s + lane; k < e; k += L) sum = fma(val[k], src[col[k]] * inv, sum);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
    if (row < n && lane == 0) {
      const double w = src[row];
      const double u = opscale * sum;
      if (j == 1) v_out[row] = w;  // the basis keeps w_0 = x (v_1 = x / rs1)
      u_out[row] = u;
      dot = fma(u, w * inv, dot);
    }
  }
  dot = block_sum<PSFUN_BS>(dot, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = dot;
}

// Same step on the CSR row blocks (A->csrs_blk, kinds as documented with the
// matrix): a block of short rows puts its entries' products a_k v(col_k) in
// LDS by coalesced loads of the whole block, then L lanes per row (the
// largest power of two with rows * L <= PSFUN_BS) sum each row from LDS; a
// long row alone is summed by the whole block; a segment of a split row is
// summed by the whole block into seg[b], and the block whose arrival
// completes the row (counter cnt[s]) adds the row's segments in order and
// puts the row's alpha term in partial slot G + s (the consumers fold G +
// nsplit partials), so the result does not depend on the arrival order.
// XT: the matrix has blocks of short rows with a staged x window (kind < 0);
// the kernel without it keeps the LDS and registers of a plain block.
template <bool SMALLT, bool XT>
__global__ __launch_bounds__(PSFUN_BS) void k_lz_spmv_csrs(int nblk, const int *__restrict__ blk,
                                                           const int *__restrict__ split, int *__restrict__ cnt,
                                                           double *__restrict__ seg, const int *__restrict__ rowptr,
                                                           const int *__restrict__ col,
                                                           const double *__restrict__ val, double opscale,
                                                           const double *__restrict__ src, int j, int G,
                                                           const double *__restrict__ p_in, int g_in,
                                                           double *__restrict__ v_out, double *__restrict__ u_out,
                                                           double *__restrict__ p_out, LzCtl *ctl, int nl, int istop,
                                                           double eps, double fscale, int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  // the streaming blocks' products and the control block's coefficient
  // table share one LDS buffer (a block is one or the other)
  constexpr size_t SH = sizeof(LzCtlLds<SMALLT>) > sizeof(double) * PSFUN_CSRS_CAP ? sizeof(LzCtlLds<SMALLT>)
                                                                                   : sizeof(double) * PSFUN_CSRS_CAP;
  __shared__ __align__(16) unsigned char shbuf[SH];
  __shared__ double xt[XT ? PSFUN_CSRS_XT : 1];  // x window of a block of short rows (kind < 0)
  double *prod = reinterpret_cast<double *>(shbuf);
  if (__syncthreads_or(ctl->done)) return;
  const double sigma = sqrt(block_sum_partials<PSFUN_BS>(p_in, g_in, lds));
  if (blockIdx.x == G) {  // control block
    if (threadIdx.x >= 64) return;
    if (j == 1) {
      if (threadIdx.x == 0) {
        host_flag_store(hf, 1);
        ctl->rs1 = sigma;
        ctl->rs[1] = sigma;
        if (sigma == 0.0) lz_stop(ctl, hf, LZ_DONE_ZERO);
      }
      return;
    }
    lz_control<SMALLT>(ctl, j, sigma, nl, istop, eps, fscale, reinterpret_cast<LzCtlLds<SMALLT> *>(shbuf), hf);
    return;
  }
  if (sigma == 0.0 && j == 1) return;
  const double inv = 1.0 / sigma;
  const XcdRange xr = xcd_range_g(nblk, G);
  double dot = 0.0;
  for (int64_t b = xr.begin; b < xr.end; b += xr.stride) {
    const int ba = blk[4 * b], bb = blk[4 * b + 1], bc = blk[4 * b + 2], kind = blk[4 * b + 3];
    int row = -1;
    double sum = 0.0;
    bool lead = false;
    if (kind <= 0) {  // short rows [ba, bb)
      const int e0 = rowptr[ba], ne = rowptr[bb] - e0;
      const int nr = bb - ba;
      int L = 1;
      while (L < 64 && nr * L * 2 <= PSFUN_BS) L *= 2;
      const int rr = threadIdx.x / L, ln = threadIdx.x % L;
      // this thread's row bounds, and all of its entries' column and value
      // loads (and its share of the x window), are in flight together, then
      // all of its gathers (one round trip each instead of one per entry)
      const int rs = (rr < nr) ? rowptr[ba + rr] : 0, re = (rr < nr) ? rowptr[ba + rr + 1] : 0;
      constexpr int NI = PSFUN_CSRS_CAP / PSFUN_BS;
      constexpr int NX = PSFUN_CSRS_XT / PSFUN_BS;
      int cq[NI];
      double aq[NI], xq[NI];
#pragma unroll
      for (int it = 0; it < NI; ++it) {
        const int t = threadIdx.x + it * PSFUN_BS;
        cq[it] = (t < ne) ? col[e0 + t] : 0;
        aq[it] = (t < ne) ? val[e0 + t] : 0.0;
      }
      if (XT && kind < 0) {
        // the block's columns lie in [bc, bc + span): that window of x is
        // staged in LDS with coalesced loads and the products read it there
        const int span = -kind;
        double xw[NX];
#pragma unroll
        for (int it = 0; it < NX; ++it) {
          const int t = threadIdx.x + it * PSFUN_BS;
          xw[it] = (t < span) ? src[bc + t] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < NX; ++it) xt[threadIdx.x + it * PSFUN_BS] = xw[it];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < NI; ++it) xq[it] = (threadIdx.x + it * PSFUN_BS < ne) ? xt[cq[it] - bc] : 0.0;
      } else {
#pragma unroll
        for (int it = 0; it < NI; ++it) xq[it] = (threadIdx.x + it * PSFUN_BS < ne) ? src[cq[it]] : 0.0;
      }
#pragma unroll
      for (int it = 0; it < NI; ++it) {
        const int t = threadIdx.x + it * PSFUN_BS;
        if (t < ne) prod[t] = aq[it] * (xq[it] * inv);
      }
      __syncthreads();
      if (rr < nr) {
        row = ba + rr;
        const int s = rs - e0, e = re - e0;
        for (int q = s + ln; q < e; q += L) sum += prod[q];
      }
      for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
      lead = rr < nr && ln == 0;
      __syncthreads();  // the next block of rows overwrites prod
    } else {  // one row, or a segment of one
      const int q0 = kind == 1 ? rowptr[ba] : bb, q1 = kind == 1 ? rowptr[ba + 1] : bc;
      // four strided entries per step, their loads in flight together; the
      // thread's sum keeps its entry order
      int q = q0 + (int)threadIdx.x;
      for (; q + 3 * PSFUN_BS < q1; q += 4 * PSFUN_BS) {
        int c4[4];
        double a4[4], x4[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          c4[m] = col[q + m * PSFUN_BS];
          a4[m] = val[q + m * PSFUN_BS];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) x4[m] = src[c4[m]];
#pragma unroll
        for (int m = 0; m < 4; ++m) sum = fma(a4[m], x4[m] * inv, sum);
      }
      for (; q < q1; q += PSFUN_BS) sum = fma(val[q], src[col[q]] * inv, sum);
      sum = block_sum<PSFUN_BS>(sum, lds);
      if (kind == 1) {
        row = ba;
        lead = threadIdx.x == 0;
      } else if (threadIdx.x == 0) {
        const int sp = kind - 2, first = split[2 * sp], nseg = split[2 * sp + 1];
        __hip_atomic_store(seg + b, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        if (atomicAdd(cnt + sp, 1) == nseg - 1) {  // the row's last segment to arrive
          __threadfence();
          double t = 0.0;
          for (int k = 0; k < nseg; ++k) t += __hip_atomic_load(seg + first + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const double w = src[ba];
          const double u = opscale * t;
          if (j == 1) v_out[ba] = w;
          u_out[ba] = u;
          p_out[G + sp] = u * (w * inv);
          __hip_atomic_store(cnt + sp, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    if (lead) {
      const double w = src[row];
      const double u = opscale * sum;
      if (j == 1) v_out[row] = w;  // the basis keeps w_0 = x (v_1 = x / rs1)
      u_out[row] = u;
      dot = fma(u, w * inv, dot);
    }
  }
  dot = block_sum<PSFUN_BS>(dot, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = dot;
}

// Same step on the sliced-ELL copy: block 0 is the control block (dispatched
// first), blocks 1..G stream one row per thread.  The stop flag, the partials
// and the first chunk's row sums are all loaded before anything waits on them,
// so the fold's latency hides behind the first chunk's loads; each row sum is
// taken unscaled and multiplied by 1/beta afterwards (u = opscale (1/beta)
// sum_k a_k w_k, equal to opscale sum_k a_k v_k up to rounding).
// D16: the columns come as 16-bit offsets from the row (A->sell_dcol).
template <bool SMALLT, bool D16>
__global__ __launch_bounds__(PSFUN_BS, 8) void k_lz_spmv_sell(int n, const int *__restrict__ sp,
                                                           const int *__restrict__ col,
                                                           const int16_t *__restrict__ dcol,
                                                           const double *__restrict__ val, double opscale,
                                                           const double *__restrict__ src, int j, int G,
                                                           const double *__restrict__ p_in, int g_in,
                                                           double *__restrict__ v_out, double *__restrict__ u_out,
                                                           double *__restrict__ p_out, LzCtl *ctl, int nl,
                                                           int istop, double eps, double fscale,
                                                           int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int stopped = ctl->done;
  double pacc = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < g_in; b += PSFUN_BS) pacc += p_in[b];
  const int64_t nchunks = (n + PSFUN_BS - 1) / PSFUN_BS;
  XcdRange xr = {0, 0, 1};
  if (blockIdx.x > 0) xr = xcd_range_c0(nchunks, G);
  double acc0 = 0.0, src0 = 0.0;
  if (xr.begin < xr.end) {
    const int row0 = (int)(xr.begin * PSFUN_BS + threadIdx.x);
    if (row0 < n) {
      acc0 = D16 ? sell_row_dot16(row0, sp, dcol, val, src, 1.0) : sell_row_dot(row0, sp, col, val, src, 1.0);
      src0 = src[row0];
    }
  }
  if (__syncthreads_or(stopped)) return;  // block-uniform (the control block may set it meanwhile)
  const double sigma = sqrt(block_sum<PSFUN_BS>(pacc, lds));
  if (blockIdx.x == 0) {  // control block
    if (threadIdx.x >= 64) return;
    if (j == 1) {
      if (threadIdx.x == 0) {
        host_flag_store(hf, 1);
        ctl->rs1 = sigma;
        ctl->rs[1] = sigma;
        if (sigma == 0.0) lz_stop(ctl, hf, LZ_DONE_ZERO);
      }
      return;
    }
    __shared__ LzCtlLds<SMALLT> W;
    lz_control<SMALLT>(ctl, j, sigma, nl, istop, eps, fscale, &W, hf);
    return;
  }
  if (sigma == 0.0 && j == 1) return;
  const double inv = 1.0 / sigma;
  const double uscale = opscale * inv;
  double dot = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * PSFUN_BS + threadIdx.x);
    if (row < n) {
      const bool first = (ch == xr.begin);
      const double sum = first ? acc0
                               : (D16 ? sell_row_dot16(row, sp, dcol, val, src, 1.0)
                                      : sell_row_dot(row, sp, col, val, src, 1.0));
      const double w = first ? src0 : src[row];
      const double u = uscale * sum;
      if (j == 1) v_out[row] = w;  // the basis keeps w_0 = x (v_1 = x / rs1)
      u_out[row] = u;
      dot = fma(u, w * inv, dot);
    }
  }
  dot = block_sum<PSFUN_BS>(dot, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x - 1] = dot;
}

// Same step on the diagonal slices (hdia_row_dot: no column indices, x read
// contiguously at each of the slice's offsets).
template <bool SMALLT>
__global__ __launch_bounds__(PSFUN_BS, 8) void k_lz_spmv_hdia(int n, const int *__restrict__ hp,
                                                           const int *__restrict__ off,
                                                           const double *__restrict__ hv, double opscale,
                                                           const double *__restrict__ src, int j, int G,
                                                           const double *__restrict__ p_in, int g_in,
                                                           double *__restrict__ v_out, double *__restrict__ u_out,
                                                           double *__restrict__ p_out, LzCtl *ctl, int nl,
                                                           int istop, double eps, double fscale,
                                                           int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int stopped = ctl->done;
  double pacc = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < g_in; b += PSFUN_BS) pacc += p_in[b];
  const int64_t nchunks = (n + PSFUN_BS - 1) / PSFUN_BS;
  XcdRange xr = {0, 0, 1};
  if (blockIdx.x > 0) xr = xcd_range_c0(nchunks, G);
  double acc0 = 0.0, src0 = 0.0;
  if (xr.begin < xr.end) {
    const int row0 = (int)(xr.begin * PSFUN_BS + threadIdx.x);
    if (row0 < n) {
      acc0 = hdia_row_dot(row0, n, hp, off, hv, src);
      src0 = src[row0];
    }
  }
  if (__syncthreads_or(stopped)) return;  // block-uniform (the control block may set it meanwhile)
  const double sigma = sqrt(block_sum<PSFUN_BS>(pacc, lds));
  if (blockIdx.x == 0) {  // control block
    if (threadIdx.x >= 64) return;
    if (j == 1) {
      if (threadIdx.x == 0) {
        host_flag_store(hf, 1);
        ctl->rs1 = sigma;
        ctl->rs[1] = sigma;
        if (sigma == 0.0) lz_stop(ctl, hf, LZ_DONE_ZERO);
      }
      return;
    }
    __shared__ LzCtlLds<SMALLT> W;
    lz_control<SMALLT>(ctl, j, sigma, nl, istop, eps, fscale, &W, hf);
    return;
  }
  if (sigma == 0.0 && j == 1) return;
  const double inv = 1.0 / sigma;
  const double uscale = opscale * inv;
  double dot = 0.0;
  for (int64_t ch = xr.begin; ch < xr.end; ch += xr.stride) {
    const int row = (int)(ch * PSFUN_BS + threadIdx.x);
    if (row < n) {
      const bool first = (ch == xr.begin);
      const double sum = first ? acc0 : hdia_row_dot(row, n, hp, off, hv, src);
      const double w = first ? src0 : src[row];
      const double u = uscale * sum;
      if (j == 1) v_out[row] = w;  // the basis keeps w_0 = x (v_1 = x / rs1)
      u_out[row] = u;
      dot = fma(u, w * inv, dot);
    }
  }
  dot = block_sum<PSFUN_BS>(dot, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x - 1] = dot;
}

// Loop control at the last index i = nl (no SpMV needed there).
__global__ __launch_bounds__(64) void k_lz_ctrl(const double *__restrict__ p_in, int g_in, int j, LzCtl *ctl,
                                                int nl, int istop, double eps, double fscale, int *hf) {
  __shared__ ChebLds W;
  if (ctl->done) return;
  double acc = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < g_in; b += 64) acc += p_in[b];
  const double beta = sqrt(wave_sum(acc));
  lz_control<false>(ctl, j, beta, nl, istop, eps, fscale, &W, hf);
}

// The basis slots hold the unnormalised w_{l} (slot 0: x), and v_{l+1} =
// w_l * (1/beta_l) is formed where it is read, with the same 1/beta the SpMV
// step used, so every v value equals the one a stored v would hold.
// beta_0 = rs1, beta_l = T(l, l+1).
__device__ __forceinline__ double lz_inv_beta(const LzCtl *ctl, int l) {
  const double b = (l == 0) ? ctl->rs1 : ctl->beta[l];
  return b != 0.0 ? 1.0 / b : 0.0;
}

// The stop flag, the partials and this thread's first LZ_UB (u, w_{j-1},
// w_{j-2}) triples are loaded before anything waits on them, so the fold's
// latency hides behind the first loads; later triples are also loaded LZ_UB
// at a time (a thread keeps 3*LZ_UB loads in flight, which the bandwidth
// needs).  Each thread still visits its rows in grid-stride order, so the
// ||w||^2 partials do not depend on LZ_UB.  No block of the grid writes the
// flag, so every thread reads the same value.
#define LZ_UB 4
__global__ __launch_bounds__(PSFUN_BS) void k_lz_update(int n, double *__restrict__ u,
                                                        const double *__restrict__ vj,
                                                        const double *__restrict__ vjm1, int j,
                                                        const double *__restrict__ p_in, int g_in,
                                                        double *__restrict__ p_out, LzCtl *ctl) {
  // u: slot j (u_j in, w_j out); vj: slot j-1 (w_{j-1}); vjm1: slot j-2 (w_{j-2})
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int stopped = ctl->done;
  double pacc = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < g_in; b += PSFUN_BS) pacc += p_in[b];
  const int64_t stride = (int64_t)gridDim.x * PSFUN_BS;
  const int64_t i0 = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x;
  double uu[LZ_UB], aa[LZ_UB], bb[LZ_UB];
#pragma unroll
  for (int q = 0; q < LZ_UB; ++q) {
    const int64_t i = i0 + q * stride;
    uu[q] = aa[q] = bb[q] = 0.0;
    if (i < n) {
      uu[q] = u[i];
      aa[q] = vj[i];
      if (j > 1) bb[q] = vjm1[i];
    }
  }
  if (stopped) return;
  const double be = (j > 1) ? ctl->beta[j - 1] : 0.0;
  const double ia = lz_inv_beta(ctl, j - 1);                 // v_j = w_{j-1} * ia
  const double ib = (j > 1) ? lz_inv_beta(ctl, j - 2) : 0.0;  // v_{j-1} = w_{j-2} * ib
  const double al = block_sum<PSFUN_BS>(pacc, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->alpha[j] = al;
  double acc = 0.0;
  for (int64_t base = i0; base < n; base += LZ_UB * stride) {
    if (base != i0) {
#pragma unroll
      for (int q = 0; q < LZ_UB; ++q) {
        const int64_t i = base + q * stride;
        if (i < n) {
          uu[q] = u[i];
          aa[q] = vj[i];
          if (j > 1) bb[q] = vjm1[i];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < LZ_UB; ++q) {
      const int64_t i = base + q * stride;
      if (i < n) {
        double w = fma(-al, aa[q] * ia, uu[q]);
        if (j > 1) w = fma(-be, bb[q] * ib, w);
        u[i] = w;
        acc = fma(w, w, acc);
      }
    }
  }
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = acc;
}

__global__ __launch_bounds__(PSFUN_BS) void k_sumsq_partials(int64_t n, const double *__restrict__ x,
                                                             double *__restrict__ p_out) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PSFUN_BS)
    acc = fma(x[i], x[i], acc);
  acc = block_sum<PSFUN_BS>(acc, lds);
  if (threadIdx.x == 0) p_out[blockIdx.x] = acc;
}

// Coefficients for the assembly when the loop did not end by convergence:
// yk = f(T(1:itx-1,1:itx-1)) e1 (psfun_d_lanczos.f90:330-362).  The reference
// leaves itx-1 = 0 undefined; here the 1x1 block T(1,1) is used.
__global__ __launch_bounds__(64) void k_lz_final(LzCtl *ctl, double fscale) {
  __shared__ ChebLds W;
  const int done = ctl->done;
  if (done == LZ_DONE_CONVERGED || done == LZ_DONE_ZERO) return;
  int k = ctl->itx - 1;
  if (k < 1) k = 1;
  double nrm, last;
  wave_exp_tridiag_e1(k, ctl->alpha + 1, ctl->beta + 1, fscale, &W, ctl->yk, nrm, last);
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
  __shared__ double yk[LZ_NL_MAX], ib[LZ_NL_MAX];
  for (int l = threadIdx.x; l < m; l += PSFUN_BS) {
    yk[l] = ctl->yk[l];
    ib[l] = lz_inv_beta(ctl, l);  // v_{l+1} = w_l * ib[l]
  }
  __syncthreads();
  basis_combine<true>(n, V, ldv, m, yk, ib, rs1, y);
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
static int lz_launch_spmv(psfun_ctx *c, const psfun_csr *A, double opscale, const double *src, int j, int G,
                          const double *p_in, int g_in, double *v_out, double *u_out, double *p_out, LzCtl *ctl,
                          int nl, int nl_ctl, int istop, double eps, double fscale, bool prof) {
  const int L = A->lanes;
  // algorithmic bytes: val+colind, rowptr, src rows read, u written (and x
  // copied into basis slot 0 at j = 1)
  const double bytes = 12.0 * A->nnz + 4.0 * (A->n + 1) + (j == 1 ? 24.0 : 16.0) * A->n;
  // The step is idempotent (it reads src and the partials and rewrites v_j, u
  // and its own partials and control fields with the same values), so timing
  // may repeat it back to back inside one event pair to amortise the events.
  const int reps = prof ? psfun_prof_reps(c) : 1;
  if (prof) PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_LZ_SPMV, bytes, j, reps));
  if (psfun_reads_hdia(A)) {
    for (int r = 0; r < reps; ++r) {
      if (nl <= 65)
        hipLaunchKernelGGL((k_lz_spmv_hdia<true>), dim3(G + 1), dim3(PSFUN_BS), 0, c->stream, A->n, A->hdia_ptr,
                           A->hdia_off, A->hdia_val, opscale, src, j, G, p_in, g_in, v_out, u_out, p_out, ctl, nl_ctl,
                           istop, eps, fscale, c->d_flags);
      else
        hipLaunchKernelGGL((k_lz_spmv_hdia<false>), dim3(G + 1), dim3(PSFUN_BS), 0, c->stream, A->n, A->hdia_ptr,
                           A->hdia_off, A->hdia_val, opscale, src, j, G, p_in, g_in, v_out, u_out, p_out, ctl, nl_ctl,
                           istop, eps, fscale, c->d_flags);
    }
    PSFUN_HIP_TRY(hipGetLastError());
    if (prof) PSFUN_TRY(psfun_prof_end(c));
    return PSFUN_SUCCESS;
  }
  if (A->format == PSFUN_FORMAT_CSRS && A->csrs_blk) {
#define LZ_CSRS(SM, XTV)                                                                                         \
  hipLaunchKernelGGL((k_lz_spmv_csrs<SM, XTV>), dim3(G + 1), dim3(PSFUN_BS), 0, c->stream, A->csrs_nblk,        \
                     A->csrs_blk, A->csrs_split, A->csrs_cnt, A->csrs_seg, A->rowptr, A->colind, A->val, opscale, \
                     src, j, G, p_in, g_in, v_out, u_out, p_out, ctl, nl_ctl, istop, eps, fscale, c->d_flags)
    for (int r = 0; r < reps; ++r) {
      if (nl <= 65) {
        if (A->csrs_nxt)
          LZ_CSRS(true, true);
        else
          LZ_CSRS(true, false);
      } else {
        if (A->csrs_nxt)
          LZ_CSRS(false, true);
        else
          LZ_CSRS(false, false);
      }
    }
#undef LZ_CSRS
    PSFUN_HIP_TRY(hipGetLastError());
    if (prof) PSFUN_TRY(psfun_prof_end(c));
    return PSFUN_SUCCESS;
  }
  if (psfun_reads_sell(A)) {
#define LZ_SELL(SM, D)                                                                                           \
  hipLaunchKernelGGL((k_lz_spmv_sell<SM, D>), dim3(G + 1), dim3(PSFUN_BS), 0, c->stream, A->n, A->sell_ptr,      \
                     A->sell_col, A->sell_dcol, A->sell_val, opscale, src, j, G, p_in, g_in, v_out, u_out, p_out, \
                     ctl, nl_ctl, istop, eps, fscale, c->d_flags)
    const bool d16 = A->sell_dcol != nullptr;
    for (int r = 0; r < reps; ++r) {
      if (nl <= 65) {
        if (d16) LZ_SELL(true, true);
        else LZ_SELL(true, false);
      } else {
        if (d16) LZ_SELL(false, true);
        else LZ_SELL(false, false);
      }
    }
#undef LZ_SELL
    PSFUN_HIP_TRY(hipGetLastError());
    if (prof) PSFUN_TRY(psfun_prof_end(c));
    return PSFUN_SUCCESS;
  }
#define LZ_LAUNCH(LL)                                                                                          \
  do {                                                                                                         \
    if (nl <= 65)                                                                                              \
      hipLaunchKernelGGL((k_lz_spmv<LL, true>), dim3(G + 1), dim3(PSFUN_BS), 0, c->stream, A->n, A->rowptr,   \
                         A->colind, A->val, opscale, src, j, G, p_in, g_in, v_out, u_out, p_out, ctl, nl_ctl, \
                         istop, eps, fscale, c->d_flags);                                                      \
    else                                                                                                       \
      hipLaunchKernelGGL((k_lz_spmv<LL, false>), dim3(G + 1), dim3(PSFUN_BS), 0, c->stream, A->n, A->rowptr,  \
                         A->colind, A->val, opscale, src, j, G, p_in, g_in, v_out, u_out, p_out, ctl, nl_ctl, \
                         istop, eps, fscale, c->d_flags);                                                      \
  } while (0)
  for (int r = 0; r < reps; ++r) switch (L) {
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
  if (prof) PSFUN_TRY(psfun_prof_end(c));
  return PSFUN_SUCCESS;
}

// Bytes of the loop workspace of one run: the control block, then the basis
// slots 0..nl-1 (slot 0 holds x, slot l the unnormalised w_l, u_l until the
// update of step l has run), each ldv = align64(n) doubles.
static size_t lz_ws_bytes(int64_t n, int nl) {
  return align_up(sizeof(LzCtl), 256) + sizeof(double) * align_up((size_t)n, 64) * (size_t)nl;
}

// The Lanczos run on a caller-provided workspace `ws` (lz_ws_bytes) and
// partial-sum lists Pw/Pa (PSFUN_MAX_GRID doubles each); `prof` enables the
// per-kernel event timing.
static int lz_apply(psfun_ctx *c, psfun_csr *A, bool userf, double fscale, double opscale, const double *x,
                    double *y, double eps, int nl, int istop, int *iter, double *err, double *res, int *nres,
                    void *ws, double *Pw, double *Pa, bool prof) {
  const int n = A->n;
  const int64_t ldv = (int64_t)align_up((size_t)n, 64);
  const size_t ctl_bytes = align_up(sizeof(LzCtl), 256);
  LzCtl *ctl = (LzCtl *)ws;
  double *V = (double *)((char *)ws + ctl_bytes);
  // USERF: the kernels' control blocks keep the loop bookkeeping and the
  // breakdown test; with istop = 1 the estimate (and with it the iteration
  // limit) moves to a host-synchronous step after each control index, since
  // the user function runs on the host.
  UserfScratch sc;
  if (userf) {
    void *ws2 = nullptr;
    PSFUN_TRY(psfun_workspace2(c, userf_scratch_bytes(nl), &ws2));
    sc = userf_scratch_carve(ws2, nl);
  }
  const bool fsync = userf && istop == 1;
  const int istop_k = userf ? 2 : istop;
  const int nl_k = fsync ? (1 << 30) : nl;
  // Pw: ||w||^2 (and ||x||^2) partials; Pa: alpha partials

  // sliced ELL and diagonal slices: one thread per row, chunks of PSFUN_BS rows
  const bool sell = psfun_row_per_thread(A);
  const int L = sell ? 1 : A->lanes;
  // CSR row blocks: one block of the grid per row block at a time
  const bool csrs = A->format == PSFUN_FORMAT_CSRS && A->csrs_blk;
  const int64_t chunks = csrs ? (int64_t)A->csrs_nblk : ((int64_t)n + PSFUN_BS / L - 1) / (PSFUN_BS / L);
  // alpha partials: one per streaming block, and with split rows one more per row
  const int nsplit = csrs ? A->csrs_nsplit : 0;
  // SpMV streaming blocks; with sliced ELL 8 blocks per CU are resident, and
  // the grid (control block + Gs) is kept within one residency round.  On
  // diagonal slices 6 per CU (5-6 row chunks per block at 128^3) measured
  // about 1% faster per f(A)x than 8 per CU.
  int Gs = psfun_grid(c, chunks, psfun_reads_hdia(A) ? 6 : (sell ? 8 : (csrs ? PSFUN_CSRS_PER_CU : 4)));
  if (sell && Gs + 1 > 8 * c->num_cu) Gs = (8 * c->num_cu - 1) / 8 * 8;
  if (Gs + nsplit > PSFUN_MAX_GRID) Gs = (PSFUN_MAX_GRID - nsplit) / 8 * 8;
  const int Gu = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 4);  // update blocks

  c->h_flags[0] = 0;
  c->h_flags[1] = 0;
  hipLaunchKernelGGL(k_lz_init, dim3(1), dim3(256), 0, c->stream, ctl);
  PSFUN_HIP_TRY(hipGetLastError());
  // rs(1) = ||x||; v_1 = (1/rs1) x; u = A v_1; partials of T(1,1) = u.v_1
  hipLaunchKernelGGL(k_sumsq_partials, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, (int64_t)n, x, Pw);
  PSFUN_HIP_TRY(hipGetLastError());
  PSFUN_TRY(lz_launch_spmv(c, A, opscale, x, 1, Gs, Pw, Gu, V, V + ldv, Pa, ctl, nl, nl_k, istop_k, eps, fscale,
                           prof));

  // Pacing: before queueing step j+2, wait until step j-1 has reported
  // (h_flags[0] >= j-1).  The stores from a running kernel may only reach the
  // host when that kernel ends, so this keeps two steps queued ahead of the
  // GPU; at most two steps run as no-ops after a stop.
  for (int j = 1; j <= nl - 1; ++j) {
    double *u = V + (size_t)ldv * j;
    const double *vj = V + (size_t)ldv * (j - 1);
    const double *vjm1 = (j > 1) ? V + (size_t)ldv * (j - 2) : nullptr;
    // algorithmic bytes: u read + w written, v_j (and v_{j-1}) read
    if (prof) PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_LZ_UPDATE, (j > 1 ? 32.0 : 24.0) * n, j));
    hipLaunchKernelGGL(k_lz_update, dim3(Gu), dim3(PSFUN_BS), 0, c->stream, n, u, vj, vjm1, j, Pa, Gs + nsplit, Pw,
                       ctl);
    PSFUN_HIP_TRY(hipGetLastError());
    if (prof) PSFUN_TRY(psfun_prof_end(c));
    if (j + 1 <= nl - 1) {
      PSFUN_TRY(lz_launch_spmv(c, A, opscale, u, j + 1, Gs, Pw, Gu, V, V + (size_t)ldv * (j + 1), Pa, ctl, nl,
                               nl_k, istop_k, eps, fscale, prof));
    } else {
      hipLaunchKernelGGL(k_lz_ctrl, dim3(1), dim3(64), 0, c->stream, Pw, Gu, nl, ctl, nl_k, istop_k, eps, fscale,
                         c->d_flags);
      PSFUN_HIP_TRY(hipGetLastError());
    }
    if (fsync) {
      // control index i = j+1: yk = f(T(1:j,1:j)) e1, rs(j+1) from beta = T(j, j+1)
      bool stopped = false;
      PSFUN_TRY(userf_eval(c, j, SYM_FROM_TRIDIAG, ctl->alpha + 1, ctl->beta + 1, 0, fscale, sc, &stopped));
      if (stopped) break;  // converged or at the limit at index j, or a breakdown at j+1
      PSFUN_TRY(userf_launch_est(c, ctl, j + 1, ctl->beta + j, sc, nl, eps));
      continue;
    }
    if (j < nl - 1 && j >= 2) {
      PSFUN_TRY(psfun_host_wait(c, j - 1));
      if (c->h_flags[1] != 0) break;
    }
  }
  if (userf) {
    // yk = f(T(1:itx-1,1:itx-1)) e1 unless the loop converged (or x = 0)
    LzCtl *hs = (LzCtl *)((char *)c->h_pinned + 256);
    PSFUN_HIP_TRY(hipMemcpyAsync(hs, ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    if (hs->done != LZ_DONE_CONVERGED && hs->done != LZ_DONE_ZERO) {
      const int k = hs->itx - 1 < 1 ? 1 : hs->itx - 1;
      PSFUN_TRY(userf_eval(c, k, SYM_FROM_TRIDIAG, ctl->alpha + 1, ctl->beta + 1, 0, fscale, sc));
      PSFUN_TRY(userf_launch_final(c, ctl, k, sc));
    }
  } else {
    hipLaunchKernelGGL(k_lz_final, dim3(1), dim3(64), 0, c->stream, ctl, fscale);
    PSFUN_HIP_TRY(hipGetLastError());
  }
  const int Ga = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 8);
  if (prof) PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_LZ_ASSEMBLE, 0.0));
  hipLaunchKernelGGL(k_lz_assemble, dim3(Ga), dim3(PSFUN_BS), 0, c->stream, n, V, ldv, ctl, y);
  PSFUN_HIP_TRY(hipGetLastError());
  if (prof) PSFUN_TRY(psfun_prof_end(c));

  // results back to the host
  LzCtl *h = (LzCtl *)((char *)c->h_pinned + 256);
  PSFUN_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(LzCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  const int itx = h->itx;
  // timing pairs of launches that came after the stop did no work: SpMV steps
  // 1..itx+1 and updates 1..itx ran (the last SpMV overlapped the stop test)
  // (with repeats, the repeats of the stopping step itself return at once)
  if (prof) {
    psfun_prof_discard_after(c, PSFUN_PROF_LZ_SPMV, psfun_prof_reps(c) > 1 ? itx : itx + 1);
    psfun_prof_discard_after(c, PSFUN_PROF_LZ_UPDATE, itx);
  }
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

extern "C" int psfun_lanczos_apply(psfun_ctx_t c, psfun_csr_t A, int fname, double fscale, double opscale,
                                   const double *x, double *y, double eps, int itmax, int istop, int *iter,
                                   double *err, double *res, int *nres) {
  if (!c || !A) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  const bool userf = fname == PSFUN_FN_USERF;
  if (userf)
    PSFUN_TRY(userf_check(c, PSFUN_VAR_SYM));
  else if (fname != PSFUN_FN_EXP)
    return PSFUN_ERR_UNSUPPORTED;
  if (itmax < 2 || itmax > LZ_NL_MAX) return PSFUN_ERR_INVALID_ARG;
  if (A->n == 0) {
    if (iter) *iter = 0;
    if (err) *err = 0.0;
    if (nres) *nres = 0;
    return PSFUN_SUCCESS;
  }
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, lz_ws_bytes(A->n, itmax), &ws));
  return lz_apply(c, A, userf, fscale, opscale, x, y, eps, itmax, istop, iter, err, res, nres, ws, c->d_partials,
                  c->d_partials + PSFUN_MAX_GRID, true);
}

// ---- allocation-free variant ------------------------------------------------
// Fixed device scratch of psfun_lanczos_apply_noalloc: the control block, the
// two partial-sum lists and the basis, carved as lz_apply expects them.  Each
// device the library runs on holds its own copy until the process ends.  The
// array's name is only a host-side handle: kernels get the device address
// looked up once per device by psfun_noalloc_setup.
#define PSFUN_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) double scratch[PSFUN_NOALLOC_SCRATCH_BYTES / sizeof(double)];

#define PSFUN_NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[PSFUN_NOALLOC_MAX_DEVICES];

// Called by psfun_ctx_create with `device` current: the device address of
// this device's copy of `scratch`.
int psfun_noalloc_setup(int device) {
  if (device < 0 || device >= PSFUN_NOALLOC_MAX_DEVICES) return PSFUN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return PSFUN_SUCCESS;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    psfun_record_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)", __FILE__,
                           __LINE__);
    return PSFUN_ERR_HIP;
  }
  g_scratch_ptr[device] = p;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_lanczos_apply_noalloc(psfun_ctx_t c, psfun_csr_t A, int fname, double fscale, double opscale,
                                           const double *x, double *y, double eps, int itmax, int istop, int *iter,
                                           double *err, double *res, int *nres) {
  if (!c || !A) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  if (fname != PSFUN_FN_EXP) return PSFUN_ERR_UNSUPPORTED;  // USERF needs host-visible scratch of its own
  if (itmax < 2 || itmax > LZ_NL_MAX) return PSFUN_ERR_INVALID_ARG;
  const size_t part_bytes = sizeof(double) * 2 * PSFUN_MAX_GRID;
  if (part_bytes + lz_ws_bytes(A->n, itmax) > PSFUN_NOALLOC_SCRATCH_BYTES) return PSFUN_ERR_TOO_LARGE;
  if (A->n == 0) {
    if (iter) *iter = 0;
    if (err) *err = 0.0;
    if (nres) *nres = 0;
    return PSFUN_SUCCESS;
  }
  if (c->device < 0 || c->device >= PSFUN_NOALLOC_MAX_DEVICES) return PSFUN_ERR_INVALID_ARG;
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[c->device];
  }
  if (!base) return PSFUN_ERR_HIP;  // psfun_ctx_create sets it up for the context's device
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  // layout: Pw | Pa | control block + basis (256-byte aligned throughout)
  double *Pw = (double *)base;
  double *Pa = Pw + PSFUN_MAX_GRID;
  void *ws = (char *)base + part_bytes;
  return lz_apply(c, A, false, fscale, opscale, x, y, eps, itmax, istop, iter, err, res, nres, ws, Pw, Pa, false);
}
