// The small f(H) step for a DENSE (general, Hessenberg or symmetric) H on the
// device: y = exp(s*H) x, evaluated by one block of PSFUN_BS threads.
//
// Method (the Taylor scheme of Al-Mohy & Higham, "Computing the action of the
// matrix exponential", SIAM J. Sci. Comput. 33 (2011)): shift by
// mu = s*trace(H)/k, split into p substeps with ||B||_1 / p <= 1 for
// B = s*H - mu*I, and on each substep sum the Taylor series of exp(B/p)
// applied to the current vector until two successive terms fall below
// 2^-53 of the partial sum (max norms); then multiply by exp(mu/p).  With
// substep norm <= 1 the terms never exceed the result by more than about e^2,
// so the cancellation error stays near the unit roundoff.  (A bound of 2
// saves a third of the terms but lets the error grow to ~e^4 roundoffs,
// which the a-posteriori estimate, a component near 1e-14 of ||yk|| at tight
// tolerances, can no longer resolve.)
//
// H is read from memory (column-major, leading dimension ldh) through the
// caches; how it is read is the variant's matrix model:
//   DENSE_FULL       every entry                        (GENPADE, TAYLOR, SASQ, CHBGEN)
//   DENSE_HESSENBERG entries with i <= j+1 only          (CHBHES: DNCHBV zeroes the rest)
//   DENSE_SYM_UPPER  H(i,j) read as H(min(i,j),max(i,j)) (CHBSYM: DSCHBV solves with ZSYSV 'U')
// The vectors live in LDS (struct DenseLds).  The matrix-vector product splits
// each row over nseg threads, whose partial sums meet in LDS in a fixed order.
#pragma once
#include "common.hpp"

#define PSFUN_DENSE_KMAX 512

enum { DENSE_FULL = 0, DENSE_HESSENBERG = 1, DENSE_SYM_UPPER = 2 };

struct DenseLds {
  double y[PSFUN_DENSE_KMAX];
  double t0[PSFUN_DENSE_KMAX];
  double t1[PSFUN_DENSE_KMAX];
  double part[PSFUN_BS];
  double red[PSFUN_BS / 64 + 1];
};

__device__ __forceinline__ double dense_h(const double *H, int ldh, int model, int i, int j) {
  if (model == DENSE_SYM_UPPER && i > j) {
    const int t = i;
    i = j;
    j = t;
  }
  if (model == DENSE_HESSENBERG && i > j + 1) return 0.0;
  return H[(int64_t)j * ldh + i];
}

// out = (s*H - mu*I) * in / div, with in/out in LDS.  All threads of the block.
__device__ inline void dense_matvec(int k, const double *H, int ldh, int model, double s, double mu, double div,
                                    const double *in, double *out, DenseLds *W) {
  const int tid = threadIdx.x;
  int nseg = PSFUN_BS / (k > 0 ? k : 1);
  if (nseg < 1) nseg = 1;
  if (nseg > 16) nseg = 16;
  if (nseg == 1) {
    for (int r = tid; r < k; r += PSFUN_BS) {
      double acc = 0.0;
#pragma unroll 2
      for (int c = 0; c < k; ++c) acc = fma(dense_h(H, ldh, model, r, c), in[c], acc);
      out[r] = fma(s, acc, -mu * in[r]) / div;
    }
    __syncthreads();
    return;
  }
  const int r = tid / nseg, sg = tid % nseg;
  if (r < k) {
    double acc = 0.0;
#pragma unroll 2
    for (int c = sg; c < k; c += nseg) acc = fma(dense_h(H, ldh, model, r, c), in[c], acc);
    W->part[tid] = acc;
  }
  __syncthreads();
  if (tid < k) {
    double acc = 0.0;
    for (int q = 0; q < nseg; ++q) acc += W->part[tid * nseg + q];
    out[tid] = fma(s, acc, -mu * in[tid]) / div;
  }
  __syncthreads();
}

__device__ inline double dense_block_max(double v, DenseLds *W) { return block_max<PSFUN_BS>(v, W->red); }
__device__ inline double dense_block_sum(double v, DenseLds *W) { return block_sum<PSFUN_BS>(v, W->red); }

// W->y <- exp(s*H) W->y.  Called by all PSFUN_BS threads of one block.
__device__ void block_expmv(int k, const double *H, int ldh, int model, double s, DenseLds *W) {
  const int tid = threadIdx.x;
  // shift mu = s*trace(H)/k and the 1-norm of B = s*H - mu*I
  double tr = 0.0;
  for (int i = tid; i < k; i += PSFUN_BS) tr += dense_h(H, ldh, model, i, i);
  const double mu = s * dense_block_sum(tr, W) / k;
  double cmax = 0.0;
  for (int c = tid; c < k; c += PSFUN_BS) {
    double cs = 0.0;
#pragma unroll 2
    for (int r = 0; r < k; ++r) {
      double b = s * dense_h(H, ldh, model, r, c);
      if (r == c) b -= mu;
      cs += fabs(b);
    }
    cmax = fmax(cmax, cs);
  }
  const double nrm = dense_block_max(cmax, W);
  int p = 1;
  if (nrm > 1.0) p = (nrm < 1e6) ? (int)ceil(nrm) : 1000000;
  const double emu = exp(mu / p);
  const double tiny = 0x1p-53;
  for (int step = 0; step < p; ++step) {
    for (int i = tid; i < k; i += PSFUN_BS) W->t0[i] = W->y[i];
    __syncthreads();
    double *t = W->t0, *tn = W->t1;
    double prev = INFINITY;
    for (int m = 1; m <= 64; ++m) {
      dense_matvec(k, H, ldh, model, s, mu, (double)p * m, t, tn, W);
      double tm = 0.0, ym = 0.0;
      for (int i = tid; i < k; i += PSFUN_BS) {
        const double v = W->y[i] + tn[i];
        W->y[i] = v;
        tm = fmax(tm, fabs(tn[i]));
        ym = fmax(ym, fabs(v));
      }
      tm = dense_block_max(tm, W);
      ym = dense_block_max(ym, W);
      double *sw = t;
      t = tn;
      tn = sw;
      if (prev + tm <= tiny * ym) break;
      prev = tm;
    }
    for (int i = tid; i < k; i += PSFUN_BS) W->y[i] *= emu;
    __syncthreads();
  }
}

// The same Taylor scheme by ONE wavefront for k <= 64: lane r owns row r, H
// sits in LDS (row stride 65, so a column read is conflict-free), and the
// norms are wavefront reductions, so no step waits on a block barrier.
#define PSFUN_WDENSE_K 64
struct WaveDenseLds {
  double h[PSFUN_WDENSE_K * (PSFUN_WDENSE_K + 1)];
  double t[2][PSFUN_WDENSE_K];
};

// y (lane r holds y_r, r < k) <- exp(s*H) y.  Called by the 64 lanes of one
// wavefront.
__device__ inline double wave_expmv(int k, const double *H, int ldh, int model, double s, double y,
                                    WaveDenseLds *L) {
  const int lane = threadIdx.x & 63;
  constexpr int LD = PSFUN_WDENSE_K + 1;
  for (int idx = lane; idx < k * k; idx += 64) {
    const int r = idx % k, c = idx / k;
    L->h[r * LD + c] = dense_h(H, ldh, model, r, c);
  }
  wave_lds_sync();
  const double mu = s * wave_sum(lane < k ? L->h[lane * LD + lane] : 0.0) / k;
  double cs = 0.0;
  if (lane < k) {
    for (int r = 0; r < k; ++r) {
      double b = s * L->h[r * LD + lane];
      if (r == lane) b -= mu;
      cs += fabs(b);
    }
  }
  const double nrm = wave_max_dpp(cs);
  int p = 1;
  if (nrm > 1.0) p = (nrm < 1e6) ? (int)ceil(nrm) : 1000000;
  const double emu = exp(mu / p);
  const double tiny = 0x1p-53;
  const double *hr = L->h + lane * LD;
  for (int step = 0; step < p; ++step) {
    int cur = 0;
    if (lane < k) L->t[0][lane] = y;
    wave_lds_sync();
    double prev = INFINITY;
    for (int m = 1; m <= 64; ++m) {
      const double *t = L->t[cur];
      double tn = 0.0;
      if (lane < k) {
        double acc = 0.0, acc2 = 0.0;  // two chains: half the dependent latency
        int c = 0;
        for (; c + 1 < k; c += 2) {
          acc = fma(hr[c], t[c], acc);
          acc2 = fma(hr[c + 1], t[c + 1], acc2);
        }
        if (c < k) acc = fma(hr[c], t[c], acc);
        acc += acc2;
        tn = fma(s, acc, -mu * t[lane]) / ((double)p * m);
        y += tn;
      }
      const double tm = wave_max_dpp(fabs(tn));
      const double ym = wave_max_dpp(lane < k ? fabs(y) : 0.0);
      if (lane < k) L->t[cur ^ 1][lane] = tn;
      wave_lds_sync();
      cur ^= 1;
      if (prev + tm <= tiny * ym) break;
      prev = tm;
    }
    y *= emu;
  }
  return y;
}

// Hi = H_k^{-1} (k x k, leading dimension k) by Gauss-Jordan elimination with
// partial pivoting, one block of PSFUN_BS threads (replaces the LAPACK
// dgetrf/dgetri pair of the shift-and-invert step).  H is read with leading
// dimension ldh; M (k*k) is the work copy.  Both M and Hi are global memory
// private to this block.  Returns false (in every thread) on a zero pivot.
__device__ bool block_gj_inverse(int k, const double *H, int ldh, double *M, double *Hi, DenseLds *W) {
  const int tid = threadIdx.x;
  const int kk = k * k;
  for (int idx = tid; idx < kk; idx += PSFUN_BS) {
    const int r = idx % k, c = idx / k;
    M[idx] = H[(int64_t)c * ldh + r];
    Hi[idx] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int c = 0; c < k; ++c) {
    // pivot row: largest |M(r,c)|, r >= c (ties: the smaller row)
    double best = -1.0;
    int bi = k;
    for (int r = c + tid; r < k; r += PSFUN_BS) {
      const double v = fabs(M[c * k + r]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if ((tid & 63) == 0) {
      W->part[tid >> 6] = best;
      W->part[8 + (tid >> 6)] = (double)bi;
    }
    __syncthreads();
    if (tid == 0) {
      double b = W->part[0];
      int p = (int)W->part[8];
      for (int w = 1; w < PSFUN_BS / 64; ++w) {
        const double ob = W->part[w];
        const int oi = (int)W->part[8 + w];
        if (ob > b || (ob == b && oi < p)) {
          b = ob;
          p = oi;
        }
      }
      W->part[16] = b;
      W->part[17] = (double)p;
    }
    __syncthreads();
    if (!(W->part[16] > 0.0)) return false;
    const int p = (int)W->part[17];
    if (p != c) {
      for (int j = tid; j < k; j += PSFUN_BS) {
        double t = M[j * k + c];
        M[j * k + c] = M[j * k + p];
        M[j * k + p] = t;
        t = Hi[j * k + c];
        Hi[j * k + c] = Hi[j * k + p];
        Hi[j * k + p] = t;
      }
    }
    __syncthreads();
    const double inv = 1.0 / M[c * k + c];
    __syncthreads();
    for (int j = tid; j < k; j += PSFUN_BS) {
      M[j * k + c] *= inv;
      Hi[j * k + c] *= inv;
    }
    __syncthreads();
    for (int r = tid; r < k; r += PSFUN_BS) W->t0[r] = (r == c) ? 0.0 : M[c * k + r];
    __syncthreads();
    for (int idx = tid; idx < kk; idx += PSFUN_BS) {
      const int r = idx % k, j = idx / k;
      const double f = W->t0[r];
      if (f != 0.0) {
        M[idx] = fma(-f, M[j * k + c], M[idx]);
        Hi[idx] = fma(-f, Hi[j * k + c], Hi[idx]);
      }
    }
    __syncthreads();
  }
  return true;
}

// C = alpha * A * B for k x k column-major matrices (A, B with leading
// dimensions lda, ldb; C with k), one block; C distinct from A and B.
__device__ inline void block_gemm(int k, double alpha, const double *A, int lda, const double *B, int ldb,
                                  double *C) {
  for (int idx = threadIdx.x; idx < k * k; idx += PSFUN_BS) {
    const int r = idx % k, c = idx / k;
    double acc = 0.0;
    for (int l = 0; l < k; ++l) acc = fma(A[(int64_t)l * lda + r], B[(int64_t)c * ldb + l], acc);
    C[idx] = alpha * acc;
  }
  __syncthreads();
}

// exp(t*H) (k x k) by the irreducible diagonal Pade approximant of degree
// 2 <= ideg <= PSFUN_PADE_MAX with scaling and squaring, the method of
// EXPOKIT's DGPADM (Sidje, ACM TOMS 24 (1998) 130-156; called by
// serial/psfun_d_serial_mod.F90:271 with fun%padedegree):
//   ns = max(0, int(log2(|t| ||H||_inf)) + 2),  x = (t / 2^ns) H
//   c_0 = 1, c_j = c_{j-1} (p + 1 - j) / (j (2p + 1 - j)):  N(x) = sum c_j x^j
//   and D(x) = N(-x), split into even and odd parts Ev(x^2), x Od(x^2)
//   (Horner in x^2 from c_p and c_{p-1}); r = N/D = 1 + 2 (Ev - x Od)^{-1} x Od,
//   and the same identity with the roles swapped gives -r when the odd part
//   lands in the other Horner polynomial; then r^(2^ns) by squaring (which
//   also removes that sign), the sign flipped explicitly when ns = 0.
// S holds 5 k*k doubles; returns the pointer to the result inside S, or
// nullptr when the denominator is singular (DGPADM stops there).  H = 0
// gives the identity (DGPADM stops there too).
__device__ const double *block_pade_expm(int k, const double *H, int ldh, double t, int ideg, double *S,
                                         DenseLds *W) {
  const int tid = threadIdx.x;
  const int kk = k * k;
  double rmax = 0.0;
  for (int r = tid; r < k; r += PSFUN_BS) {
    double rs = 0.0;
    for (int c = 0; c < k; ++c) rs += fabs(H[(int64_t)c * ldh + r]);
    rmax = fmax(rmax, rs);
  }
  const double hnorm = fabs(t * dense_block_max(rmax, W));
  double *buf[5] = {S, S + kk, S + 2 * (size_t)kk, S + 3 * (size_t)kk, S + 4 * (size_t)kk};
  if (hnorm == 0.0) {
    for (int idx = tid; idx < kk; idx += PSFUN_BS) buf[0][idx] = (idx % k == idx / k) ? 1.0 : 0.0;
    __syncthreads();
    return buf[0];
  }
  const int ns = max(0, (int)(log(hnorm) / log(2.0)) + 2);
  const double scale = ldexp(t, -ns);
  if (tid == 0) {  // Pade coefficients c_0..c_p in LDS
    W->part[0] = 1.0;
    for (int j = 1; j <= ideg; ++j)
      W->part[j] = (W->part[j - 1] * (double)(ideg + 1 - j)) / (double)(j * (2 * ideg + 1 - j));
  }
  __syncthreads();
  const double *cf = W->part;
  double *H2 = buf[0], *P = buf[1], *Q = buf[2], *F = buf[3], *G = buf[4];
  block_gemm(k, scale * scale, H, ldh, H, ldh, H2);
  for (int idx = tid; idx < kk; idx += PSFUN_BS) {
    const bool dg = idx % k == idx / k;
    P[idx] = dg ? cf[ideg - 1] : 0.0;
    Q[idx] = dg ? cf[ideg] : 0.0;
  }
  __syncthreads();
  int odd = 1;  // the polynomial the next Horner step extends: Q when odd, else P
  for (int j = ideg - 1; j > 0; --j) {
    double *used = odd ? Q : P;
    block_gemm(k, 1.0, used, k, H2, k, F);
    for (int d = tid; d < k; d += PSFUN_BS) F[d * k + d] += cf[j - 1];
    __syncthreads();
    if (odd)
      Q = F;
    else
      P = F;
    F = used;
    odd ^= 1;
  }
  // the polynomial whose turn would come next holds the odd part: multiply it by x
  if (odd) {
    block_gemm(k, scale, Q, k, H, ldh, F);
    double *tq = Q;
    Q = F;
    F = tq;
  } else {
    block_gemm(k, scale, P, k, H, ldh, F);
    double *tp = P;
    P = F;
    F = tp;
  }
  for (int idx = tid; idx < kk; idx += PSFUN_BS) Q[idx] -= P[idx];
  __syncthreads();
  // X = (Q - P)^{-1} P, then 2 X + I (F: elimination work, G: the inverse)
  if (!block_gj_inverse(k, Q, k, F, G, W)) return nullptr;
  block_gemm(k, 1.0, G, k, P, k, F);
  for (int idx = tid; idx < kk; idx += PSFUN_BS) F[idx] = 2.0 * F[idx] + ((idx % k == idx / k) ? 1.0 : 0.0);
  __syncthreads();
  if (ns == 0) {
    if (odd) {
      for (int idx = tid; idx < kk; idx += PSFUN_BS) F[idx] = -F[idx];
      __syncthreads();
    }
    return F;
  }
  double *cur = F, *nxt = G;
  for (int q = 0; q < ns; ++q) {
    block_gemm(k, 1.0, cur, k, cur, k, nxt);
    double *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  return cur;
}

// W->y <- exp(t*H) W->y by block_pade_expm; false on a singular denominator.
__device__ inline bool block_pade_expmv(int k, const double *H, int ldh, double t, int ideg, double *S,
                                        DenseLds *W) {
  const double *E = block_pade_expm(k, H, ldh, t, ideg, S, W);
  if (!E) return false;
  for (int i = threadIdx.x; i < k; i += PSFUN_BS) W->t0[i] = W->y[i];
  __syncthreads();
  for (int r = threadIdx.x; r < k; r += PSFUN_BS) {
    double acc = 0.0;
    for (int c = 0; c < k; ++c) acc = fma(E[(int64_t)c * k + r], W->t0[c], acc);
    W->y[r] = acc;
  }
  __syncthreads();
  return true;
}

// The dense exp step of the variants: the Taylor scheme, or with pade > 0 the
// degree-`pade` Pade approximant (S: 5 k*k doubles).  A singular Pade
// denominator leaves NaN in W->y.
__device__ inline void block_expmv_var(int k, const double *H, int ldh, int model, double s, int pade, double *S,
                                       DenseLds *W) {
  if (pade > 0) {
    if (!block_pade_expmv(k, H, ldh, s, pade, S, W)) {
      for (int i = threadIdx.x; i < k; i += PSFUN_BS) W->y[i] = NAN;
      __syncthreads();
    }
  } else {
    block_expmv(k, H, ldh, model, s, W);
  }
}
