// Preconditioned per-shift fallback of the multi-shift solver and of the
// quadrature apply.  No reference counterpart: the reference's quadrature
// object stores a preconditioner (psfun_z_quadrature%set(prec),
// quadratures/psfun_z_quadrature_mod.F90:156-166) for the shifted solves its
// poles need, and its shifted solver applies one (prec%apply,
// krylov/shifted/psfun_dcg.F90:283).
//
// A preconditioner built on A is not shift-invariant: M^{-1}(A + s I) spans
// a different Krylov space for every s, so the one-space multi-shift CG
// cannot use it.  The multi-shift solve therefore runs first, unpreconditioned;
// every shift it leaves above the tolerance after itmax iterations (the ill-
// conditioned poles near the spectrum) continues from its multi-shift iterate
// by preconditioned COCG (van der Vorst and Melissen 1990: CG with the
// bilinear form x^T y, valid for complex symmetric systems), all such shifts
// side by side in one launch sequence, one grid row (blockIdx.y) per shift:
//   r = b - (A + s I) x; z = M^{-1} r; p = z; rho = r^T z
//   repeat: q = (A + s I) p; mu = p^T q; alpha = rho/mu; x += alpha p;
//           r -= alpha q; stop when ||r||_2 <= eps ||b||_2 or after itmax
//           iterations; z = M^{-1} r; rho' = r^T z; p = z + (rho'/rho) p
// mu == 0 or rho == 0 stops that shift (a breakdown; its iterate is kept).
// M (complex symmetric, as COCG needs):
//   NONE: I;  DIAG: diag(A) + s I, the Jacobi matrix of the shifted system
//   (entries equal to 0 replaced by 1);  BJAC: L U, the ILU(0) of the real A
//   (L D L^T for a symmetric A), applied to the real and imaginary parts.
// Per iteration: k_pc_spmv (q, mu partials) -> k_pc_alpha (one block: alpha)
// -> k_pc_update (x, r, ||r||^2 partials) -> [BJAC: level-scheduled solves]
// -> k_pc_z (z for NONE/DIAG, r^T z partials) -> k_pc_check (one block: stop
// tests, beta, host flags) -> k_pc_dir (p).  Scalars are folded from per-block
// partials in a fixed order, so a run is deterministic.
#include "common.hpp"
#include "kernels.hpp"

#include <vector>

#define PC_MAX 64

struct PcCtl {
  int k;     // iterations done
  int stop;  // 1 once no shift runs
  int pad[2];
  double bn;  // ||b||_2
  int run[PC_MAX];
  int iters[PC_MAX];
  double resid[PC_MAX];
  double sig[PC_MAX][2];
  double rho[PC_MAX][2];
  double alpha[PC_MAX][2];
  double beta[PC_MAX][2];
};

struct PcPartials {
  double *rr, *zr, *zi, *qr, *qi, *bb;  // [s * G + block], bb: [block]
};

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cdiv2(double2 a, double2 b) {
  // Smith's algorithm: no overflow in |b|^2 for large entries
  if (fabs(b.x) >= fabs(b.y)) {
    const double t = b.y / b.x, d = b.x + b.y * t;
    return make_double2((a.x + a.y * t) / d, (a.y - a.x * t) / d);
  }
  const double t = b.x / b.y, d = b.y + b.x * t;
  return make_double2((a.x * t + a.y) / d, (a.y * t - a.x) / d);
}

// sum_k a_ik v_k + s v_i for row i of the CSR matrix
__device__ __forceinline__ double2 shifted_row(int i, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const double *__restrict__ va, const double2 *__restrict__ v,
                                               double2 sg) {
  double2 t = make_double2(0.0, 0.0);
  for (int q = rp[i]; q < rp[i + 1]; ++q) {
    const double a = va[q];
    const double2 u = v[ci[q]];
    t.x = fma(a, u.x, t.x);
    t.y = fma(a, u.y, t.y);
  }
  const double2 vi = v[i];
  return make_double2(t.x + (sg.x * vi.x - sg.y * vi.y), t.y + (sg.x * vi.y + sg.y * vi.x));
}

__global__ __launch_bounds__(PSFUN_BS) void k_pc_diag(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                      const double *__restrict__ va, double *__restrict__ d) {
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += gridDim.x * PSFUN_BS) {
    double t = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q)
      if (ci[q] == i) t += va[q];
    d[i] = t;
  }
}

// r_s = b - (A + s I) x_s and ||r_s||^2 partials for the shifts with on[s];
// the blocks of shift row 0 also take the ||b||^2 partials.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_resid(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                       const double *__restrict__ va, const double *__restrict__ b,
                                                       const double2 *__restrict__ X, int64_t ldx,
                                                       double2 *__restrict__ R, int64_t ld, const int *__restrict__ on,
                                                       const PcCtl *ctl, PcPartials P) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int s = blockIdx.y, G = gridDim.x;
  const bool act = on[s] != 0;
  double rr = 0.0, bb = 0.0;
  const double2 sg = make_double2(ctl->sig[s][0], ctl->sig[s][1]);
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += G * PSFUN_BS) {
    const double bi = b[i];
    bb = fma(bi, bi, bb);
    if (act) {
      const double2 ax = shifted_row(i, rp, ci, va, X + (size_t)s * ldx, sg);
      const double2 r = make_double2(bi - ax.x, -ax.y);
      R[(size_t)s * ld + i] = r;
      rr = fma(r.x, r.x, fma(r.y, r.y, rr));
    }
  }
  rr = block_sum<PSFUN_BS>(rr, lds);
  if (s == 0) bb = block_sum<PSFUN_BS>(bb, lds);
  if (threadIdx.x == 0) {
    P.rr[s * G + blockIdx.x] = rr;
    if (s == 0) P.bb[blockIdx.x] = bb;
  }
}

// z_s = M^{-1} r_s for NONE (z aliases r) and DIAG; BJAC's z is already
// formed.  Partials of r_s^T z_s.
template <bool DIAG>
__global__ __launch_bounds__(PSFUN_BS) void k_pc_z(int n, const double *__restrict__ d, const double2 *__restrict__ R,
                                                   double2 *__restrict__ Z, int64_t ld, const int *__restrict__ run,
                                                   const PcCtl *ctl, PcPartials P) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int s = blockIdx.y, G = gridDim.x;
  if (!run[s]) return;
  const double2 sg = make_double2(ctl->sig[s][0], ctl->sig[s][1]);
  double zr = 0.0, zi = 0.0;
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += G * PSFUN_BS) {
    const double2 r = R[(size_t)s * ld + i];
    double2 z;
    if (DIAG) {
      double2 m = make_double2(d[i] + sg.x, sg.y);
      if (m.x == 0.0 && m.y == 0.0) m.x = 1.0;
      z = cdiv2(r, m);
      Z[(size_t)s * ld + i] = z;
    } else {
      z = Z[(size_t)s * ld + i];
    }
    zr += r.x * z.x - r.y * z.y;
    zi += r.x * z.y + r.y * z.x;
  }
  zr = block_sum<PSFUN_BS>(zr, lds);
  zi = block_sum<PSFUN_BS>(zi, lds);
  if (threadIdx.x == 0) {
    P.zr[s * G + blockIdx.x] = zr;
    P.zi[s * G + blockIdx.x] = zi;
  }
}

// Start (one block): ||b||, the start residuals, rho; which shifts iterate.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_start(int S, int G, const int *__restrict__ on, PcCtl *ctl,
                                                       PcPartials P, double eps, int itmax) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const double bn = sqrt(block_sum_partials<PSFUN_BS>(P.bb, G, lds));
  for (int s = 0; s < S; ++s) {
    const double rn = sqrt(block_sum_partials<PSFUN_BS>(P.rr + s * G, G, lds));
    if (threadIdx.x == 0) {
      ctl->iters[s] = 0;
      if (!on[s]) {
        ctl->run[s] = 0;
      } else {
        ctl->resid[s] = (bn > 0.0) ? rn / bn : 0.0;
        ctl->run[s] = (bn > 0.0 && rn > eps * bn && itmax > 0) ? 1 : 0;
      }
    }
  }
  if (threadIdx.x == 0) {
    ctl->bn = bn;
    ctl->k = 0;
    ctl->stop = 0;
  }
}

// rho = r^T z of the start (one block, after k_pc_z): a zero rho stops the shift.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_rho0(int S, int G, PcCtl *ctl, PcPartials P, int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  int nrun = 0;
  for (int s = 0; s < S; ++s) {
    const int r = ctl->run[s];
    const double zr = block_sum_partials<PSFUN_BS>(P.zr + s * G, G, lds);
    const double zi = block_sum_partials<PSFUN_BS>(P.zi + s * G, G, lds);
    if (threadIdx.x == 0 && r) {
      ctl->rho[s][0] = zr;
      ctl->rho[s][1] = zi;
      ctl->beta[s][0] = 0.0;
      ctl->beta[s][1] = 0.0;
      if (zr == 0.0 && zi == 0.0) ctl->run[s] = 0;
    }
    nrun += (r && !(zr == 0.0 && zi == 0.0)) ? 1 : 0;
  }
  if (threadIdx.x == 0) {
    ctl->stop = nrun == 0;
    if (nrun == 0) host_flag_store(hf + 1, 1);
  }
}

// p_s = z_s + beta_s p_s (first: p_s = z_s); with b = 0 the first call sets
// every requested x_s to 0.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_dir(int n, const double2 *__restrict__ Z, double2 *__restrict__ Pv,
                                                     int64_t ld, double2 *__restrict__ X, int64_t ldx,
                                                     const int *__restrict__ on, const PcCtl *ctl, int first) {
  const int s = blockIdx.y, G = gridDim.x;
  if (first && on[s] && !(ctl->bn > 0.0)) {
    for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += G * PSFUN_BS)
      X[(size_t)s * ldx + i] = make_double2(0.0, 0.0);
    return;
  }
  if (ctl->stop || !ctl->run[s]) return;
  const double2 be = make_double2(ctl->beta[s][0], ctl->beta[s][1]);
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += G * PSFUN_BS) {
    const double2 z = Z[(size_t)s * ld + i];
    if (first) {
      Pv[(size_t)s * ld + i] = z;
    } else {
      const double2 bp = cmul2(be, Pv[(size_t)s * ld + i]);
      Pv[(size_t)s * ld + i] = make_double2(z.x + bp.x, z.y + bp.y);
    }
  }
}

// q_s = (A + s I) p_s and partials of mu_s = p_s^T q_s.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_spmv(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                      const double *__restrict__ va, const double2 *__restrict__ Pv,
                                                      double2 *__restrict__ Q, int64_t ld, const PcCtl *ctl,
                                                      PcPartials P) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int s = blockIdx.y, G = gridDim.x;
  if (ctl->stop || !ctl->run[s]) return;
  const double2 sg = make_double2(ctl->sig[s][0], ctl->sig[s][1]);
  const double2 *p = Pv + (size_t)s * ld;
  double qr = 0.0, qi = 0.0;
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += G * PSFUN_BS) {
    const double2 q = shifted_row(i, rp, ci, va, p, sg);
    Q[(size_t)s * ld + i] = q;
    const double2 pi = p[i];
    qr += pi.x * q.x - pi.y * q.y;
    qi += pi.x * q.y + pi.y * q.x;
  }
  qr = block_sum<PSFUN_BS>(qr, lds);
  qi = block_sum<PSFUN_BS>(qi, lds);
  if (threadIdx.x == 0) {
    P.qr[s * G + blockIdx.x] = qr;
    P.qi[s * G + blockIdx.x] = qi;
  }
}

// alpha_s = rho_s / mu_s (one block); mu_s == 0 stops the shift.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_alpha(int S, int G, PcCtl *ctl, PcPartials P) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->stop) return;
  for (int s = 0; s < S; ++s) {
    if (!ctl->run[s]) continue;  // uniform over the block
    const double mr = block_sum_partials<PSFUN_BS>(P.qr + s * G, G, lds);
    const double mi = block_sum_partials<PSFUN_BS>(P.qi + s * G, G, lds);
    if (threadIdx.x == 0) {
      if (mr == 0.0 && mi == 0.0) {
        ctl->run[s] = 0;
      } else {
        const double2 a = cdiv2(make_double2(ctl->rho[s][0], ctl->rho[s][1]), make_double2(mr, mi));
        ctl->alpha[s][0] = a.x;
        ctl->alpha[s][1] = a.y;
      }
    }
    __syncthreads();
  }
}

// x_s += alpha_s p_s; r_s -= alpha_s q_s; ||r_s||^2 partials.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_update(int n, const double2 *__restrict__ Pv,
                                                        const double2 *__restrict__ Q, double2 *__restrict__ R,
                                                        int64_t ld, double2 *__restrict__ X, int64_t ldx,
                                                        const PcCtl *ctl, PcPartials P) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  const int s = blockIdx.y, G = gridDim.x;
  if (ctl->stop || !ctl->run[s]) return;
  const double2 al = make_double2(ctl->alpha[s][0], ctl->alpha[s][1]);
  double rr = 0.0;
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += G * PSFUN_BS) {
    const double2 ap = cmul2(al, Pv[(size_t)s * ld + i]);
    const double2 aq = cmul2(al, Q[(size_t)s * ld + i]);
    double2 *xp = X + (size_t)s * ldx + i;
    const double2 x = *xp;
    *xp = make_double2(x.x + ap.x, x.y + ap.y);
    const double2 r0 = R[(size_t)s * ld + i];
    const double2 r = make_double2(r0.x - aq.x, r0.y - aq.y);
    R[(size_t)s * ld + i] = r;
    rr = fma(r.x, r.x, fma(r.y, r.y, rr));
  }
  rr = block_sum<PSFUN_BS>(rr, lds);
  if (threadIdx.x == 0) P.rr[s * G + blockIdx.x] = rr;
}

// Stop tests and beta (one block): resid_s = ||r_s|| / ||b||; a shift stops
// at ||r_s|| <= eps ||b||, at itmax, or when rho' = 0.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_check(int S, int G, PcCtl *ctl, PcPartials P, double eps, int itmax,
                                                       int *hf) {
  __shared__ double lds[PSFUN_BS / 64 + 1];
  if (ctl->stop) return;
  const int k = ctl->k + 1;
  const double bn = ctl->bn;
  int nrun = 0;
  for (int s = 0; s < S; ++s) {
    if (!ctl->run[s]) continue;  // uniform over the block
    const double rn = sqrt(block_sum_partials<PSFUN_BS>(P.rr + s * G, G, lds));
    const double zr = block_sum_partials<PSFUN_BS>(P.zr + s * G, G, lds);
    const double zi = block_sum_partials<PSFUN_BS>(P.zi + s * G, G, lds);
    bool go = !(rn <= eps * bn) && k < itmax && !(zr == 0.0 && zi == 0.0);
    if (threadIdx.x == 0) {
      ctl->iters[s] += 1;
      ctl->resid[s] = rn / bn;
      if (go) {
        const double2 be = cdiv2(make_double2(zr, zi), make_double2(ctl->rho[s][0], ctl->rho[s][1]));
        ctl->beta[s][0] = be.x;
        ctl->beta[s][1] = be.y;
        ctl->rho[s][0] = zr;
        ctl->rho[s][1] = zi;
      } else {
        ctl->run[s] = 0;
      }
    }
    nrun += go ? 1 : 0;
    __syncthreads();
  }
  __syncthreads();  // every wave has read ctl->k and ctl->stop
  if (threadIdx.x == 0) {
    ctl->k = k;
    host_flag_store(hf, k);
    if (nrun == 0) {
      ctl->stop = 1;
      host_flag_store(hf + 1, 1);
    }
  }
}

// Preconditioned COCG for the shifts s with on_host[s] != 0, continuing from
// X (shift-major, leading dimension ldx); iters[s] receives the COCG
// iterations and resid[s] the final ||r||/||b|| of those shifts.
static int pcocg_run(psfun_ctx *c, const psfun_csr *A, const psfun_prec *Pc, int S, const double *shifts,
                     const int *on_host, const double *b, double2 *X, int64_t ldx, double eps, int itmax, int *iters,
                     double *resid) {
  const int n = A->n;
  const int kind = psfun_prec_kind(Pc);
  const int64_t ld = (int64_t)align_up((size_t)n, 64);
  const int G = std::min(psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 2), 1024);
  const size_t ctl_bytes = align_up(sizeof(PcCtl), 256);
  const size_t vec = sizeof(double2) * (size_t)ld * S;
  const size_t nvec = (kind == PSFUN_PREC_BJAC) ? 5 : (kind == PSFUN_PREC_DIAG ? 4 : 3);
  const size_t part = sizeof(double) * (5 * (size_t)S * G + G);
  const size_t bytes = ctl_bytes + 256 + nvec * vec + sizeof(double) * ld + align_up(part, 256);
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, bytes, &ws));
  char *w = (char *)ws;
  PcCtl *ctl = (PcCtl *)w;
  int *on = (int *)(w + ctl_bytes);
  double2 *R = (double2 *)(w + ctl_bytes + 256);
  double2 *Pv = R + (size_t)ld * S;
  double2 *Q = Pv + (size_t)ld * S;
  double2 *Z = (kind == PSFUN_PREC_NONE) ? R : Q + (size_t)ld * S;
  double2 *Y = (kind == PSFUN_PREC_BJAC) ? Z + (size_t)ld * S : nullptr;
  double *d = (double *)(R + nvec * (size_t)ld * S);
  PcPartials Pp;
  Pp.rr = d + ld;
  Pp.zr = Pp.rr + (size_t)S * G;
  Pp.zi = Pp.zr + (size_t)S * G;
  Pp.qr = Pp.zi + (size_t)S * G;
  Pp.qi = Pp.qr + (size_t)S * G;
  Pp.bb = Pp.qi + (size_t)S * G;

  // control block and shift mask staged through pinned memory
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  PcCtl *hc = (PcCtl *)c->h_pinned;
  for (int s = 0; s < PC_MAX; ++s) {
    hc->sig[s][0] = s < S ? shifts[2 * s] : 0.0;
    hc->sig[s][1] = s < S ? shifts[2 * s + 1] : 0.0;
  }
  int *hon = (int *)((char *)c->h_pinned + align_up(sizeof(PcCtl), 256));
  for (int s = 0; s < S; ++s) hon[s] = on_host[s] ? 1 : 0;
  PSFUN_HIP_TRY(hipMemcpyAsync(ctl, hc, sizeof(PcCtl), hipMemcpyHostToDevice, c->stream));
  PSFUN_HIP_TRY(hipMemcpyAsync(on, hon, sizeof(int) * S, hipMemcpyHostToDevice, c->stream));
  c->h_flags[0] = 0;
  c->h_flags[1] = 0;

  const dim3 grid(G, S);
  if (kind == PSFUN_PREC_DIAG)
    hipLaunchKernelGGL(k_pc_diag, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, A->rowptr, A->colind, A->val, d);
  hipLaunchKernelGGL(k_pc_resid, grid, dim3(PSFUN_BS), 0, c->stream, n, A->rowptr, A->colind, A->val, b,
                     (const double2 *)X, ldx, R, ld, (const int *)on, (const PcCtl *)ctl, Pp);
  hipLaunchKernelGGL(k_pc_start, dim3(1), dim3(PSFUN_BS), 0, c->stream, S, G, (const int *)on, ctl, Pp, eps, itmax);
  auto precondition = [&]() -> int {
    if (kind == PSFUN_PREC_BJAC) {
      PSFUN_TRY(psfun_prec_launch_c(c, Pc, R, Z, Y, ld, S, ctl->run));
      hipLaunchKernelGGL(k_pc_z<false>, grid, dim3(PSFUN_BS), 0, c->stream, n, (const double *)d,
                         (const double2 *)R, Z, ld, (const int *)ctl->run, (const PcCtl *)ctl, Pp);
    } else if (kind == PSFUN_PREC_DIAG) {
      hipLaunchKernelGGL(k_pc_z<true>, grid, dim3(PSFUN_BS), 0, c->stream, n, (const double *)d, (const double2 *)R,
                         Z, ld, (const int *)ctl->run, (const PcCtl *)ctl, Pp);
    } else {
      hipLaunchKernelGGL(k_pc_z<false>, grid, dim3(PSFUN_BS), 0, c->stream, n, (const double *)d,
                         (const double2 *)R, Z, ld, (const int *)ctl->run, (const PcCtl *)ctl, Pp);
    }
    PSFUN_HIP_TRY(hipGetLastError());
    return PSFUN_SUCCESS;
  };
  PSFUN_TRY(precondition());
  hipLaunchKernelGGL(k_pc_rho0, dim3(1), dim3(PSFUN_BS), 0, c->stream, S, G, ctl, Pp, c->d_flags);
  hipLaunchKernelGGL(k_pc_dir, grid, dim3(PSFUN_BS), 0, c->stream, n, (const double2 *)Z, Pv, ld, X, ldx,
                     (const int *)on, (const PcCtl *)ctl, 1);
  PSFUN_HIP_TRY(hipGetLastError());
  for (int k = 0; k < itmax; ++k) {
    hipLaunchKernelGGL(k_pc_spmv, grid, dim3(PSFUN_BS), 0, c->stream, n, A->rowptr, A->colind, A->val,
                       (const double2 *)Pv, Q, ld, (const PcCtl *)ctl, Pp);
    hipLaunchKernelGGL(k_pc_alpha, dim3(1), dim3(PSFUN_BS), 0, c->stream, S, G, ctl, Pp);
    hipLaunchKernelGGL(k_pc_update, grid, dim3(PSFUN_BS), 0, c->stream, n, (const double2 *)Pv, (const double2 *)Q,
                       R, ld, X, ldx, (const PcCtl *)ctl, Pp);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_TRY(precondition());
    hipLaunchKernelGGL(k_pc_check, dim3(1), dim3(PSFUN_BS), 0, c->stream, S, G, ctl, Pp, eps, itmax, c->d_flags);
    hipLaunchKernelGGL(k_pc_dir, grid, dim3(PSFUN_BS), 0, c->stream, n, (const double2 *)Z, Pv, ld, X, ldx,
                       (const int *)on, (const PcCtl *)ctl, 0);
    PSFUN_HIP_TRY(hipGetLastError());
    // keep about two iterations queued ahead of the last one checked
    if (k >= 2 && k + 1 < itmax) {
      PSFUN_TRY(psfun_host_wait(c, k - 1));
      if (c->h_flags[1] != 0) break;
    }
  }
  PSFUN_HIP_TRY(hipMemcpyAsync(hc, ctl, sizeof(PcCtl), hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  for (int s = 0; s < S; ++s) {
    if (!on_host[s]) continue;
    iters[s] = hc->iters[s];
    resid[s] = hc->resid[s];
  }
  return PSFUN_SUCCESS;
}

// The multi-shift solve, then the preconditioned fallback for the shifts it
// left above eps; its/res are the caller's arrays or host scratch.
static int mscg_prec_run(psfun_ctx *c, const psfun_csr *A, const psfun_prec *P, int nshift, const double *shifts,
                         const double *b, double *X, int64_t ldx, double eps, int itmax, int *its, double *res) {
  PSFUN_TRY(psfun_mscg_run(c, A, nshift, shifts, b, X, ldx, eps, itmax, its, res));
  if (!P) return PSFUN_SUCCESS;
  std::vector<int> on(nshift, 0), it2(nshift, 0);
  std::vector<double> r2(nshift, 0.0);
  bool any = false;
  for (int s = 0; s < nshift; ++s) {
    on[s] = res[s] > eps ? 1 : 0;
    any = any || on[s];
  }
  if (!any) return PSFUN_SUCCESS;
  PSFUN_TRY(pcocg_run(c, A, P, nshift, shifts, on.data(), b, (double2 *)X, ldx, eps, itmax, it2.data(), r2.data()));
  for (int s = 0; s < nshift; ++s)
    if (on[s]) {
      its[s] += it2[s];
      res[s] = r2[s];
    }
  return PSFUN_SUCCESS;
}

static int prec_check(const psfun_csr *A, const psfun_prec *P) {
  if (!P) return PSFUN_SUCCESS;
  const int k = psfun_prec_kind(P);
  if (k != PSFUN_PREC_NONE && k != PSFUN_PREC_DIAG && k != PSFUN_PREC_BJAC) return PSFUN_ERR_INVALID_ARG;
  if (psfun_prec_size(P) != A->n) return PSFUN_ERR_INVALID_ARG;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_mscg_prec(psfun_ctx_t c, psfun_csr_t A, psfun_prec_t P, int nshift, const double *shifts,
                               const double *b, double *X, int64_t ldx, double eps, int itmax, int *iters,
                               double *resid) {
  if (!c || !A || nshift < 0 || nshift > PC_MAX || (nshift && !shifts)) return PSFUN_ERR_INVALID_ARG;
  if (!b || !X) return PSFUN_ERR_INVALID_VECT_STATE;
  if (itmax < 0 || ldx < A->n) return PSFUN_ERR_INVALID_ARG;
  PSFUN_TRY(prec_check(A, P));
  if (A->n == 0 || nshift == 0) return PSFUN_SUCCESS;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  if (itmax == 0) {
    PSFUN_HIP_TRY(hipMemsetAsync(X, 0, sizeof(double2) * (size_t)ldx * nshift, c->stream));
    return psfun_ctx_synchronize(c);
  }
  std::vector<int> its(nshift, 0);
  std::vector<double> res(nshift, 0.0);
  PSFUN_TRY(mscg_prec_run(c, A, P, nshift, shifts, b, X, ldx, eps, itmax, its.data(), res.data()));
  for (int s = 0; s < nshift; ++s) {
    if (iters) iters[s] = its[s];
    if (resid) resid[s] = res[s];
  }
  return PSFUN_SUCCESS;
}

// y = gamma Re sum_s w_s x_s (form 0) or acc = gamma Im sum_s w_s x_s (form 1,
// then y = A acc), shifts summed in order.
__global__ __launch_bounds__(PSFUN_BS) void k_pc_combine(int n, int S, const double2 *__restrict__ X, int64_t ldx,
                                                         const double *__restrict__ wg, int form,
                                                         double *__restrict__ out) {
  for (int i = blockIdx.x * PSFUN_BS + threadIdx.x; i < n; i += gridDim.x * PSFUN_BS) {
    double a = 0.0;
    for (int s = 0; s < S; ++s) {
      const double2 x = X[(size_t)s * ldx + i];
      const double wr = wg[2 * s], wi = wg[2 * s + 1];
      a += (form == 0) ? (wr * x.x - wi * x.y) : (wr * x.y + wi * x.x);
    }
    out[i] = a;
  }
}

extern "C" int psfun_quad_apply_prec(psfun_ctx_t c, psfun_csr_t A, psfun_prec_t P, int nshift, const double *shifts,
                                     const double *weights, double gamma, int form, const double *x, double *y,
                                     double eps, int itmax, int *iters, double *resid) {
  if (!c || !A || nshift < 0 || nshift > PC_MAX || (nshift && (!shifts || !weights))) return PSFUN_ERR_INVALID_ARG;
  if (form != PSFUN_QUAD_RE_SUM && form != PSFUN_QUAD_A_IM_SUM) return PSFUN_ERR_INVALID_ARG;
  if (!x || !y) return PSFUN_ERR_INVALID_VECT_STATE;
  if (itmax < 1) return PSFUN_ERR_INVALID_ARG;
  PSFUN_TRY(prec_check(A, P));
  const int n = A->n;
  if (n == 0) return PSFUN_SUCCESS;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  if (nshift == 0) {
    PSFUN_HIP_TRY(hipMemsetAsync(y, 0, sizeof(double) * n, c->stream));
    return psfun_ctx_synchronize(c);
  }
  // the shifted solutions live in the outer workspace (the solvers use the
  // inner one), followed by gamma * w_s and the form-1 accumulator
  const int64_t ld = (int64_t)align_up((size_t)n, 64);
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace2(c, sizeof(double2) * (size_t)ld * nshift + sizeof(double) * (2 * PC_MAX + ld), &ws));
  double2 *X = (double2 *)ws;
  double *wg = (double *)(X + (size_t)ld * nshift);
  double *acc = wg + 2 * PC_MAX;
  std::vector<int> its(nshift, 0);
  std::vector<double> res(nshift, 0.0);
  PSFUN_TRY(mscg_prec_run(c, A, P, nshift, shifts, x, (double *)X, ld, eps, itmax, its.data(), res.data()));
  double *hw = (double *)c->h_pinned;  // stream idle: mscg_prec_run ends synchronised
  for (int s = 0; s < 2 * nshift; ++s) hw[s] = gamma * weights[s];
  PSFUN_HIP_TRY(hipMemcpyAsync(wg, hw, sizeof(double) * 2 * nshift, hipMemcpyHostToDevice, c->stream));
  const int G = psfun_grid(c, (n + PSFUN_BS - 1) / PSFUN_BS, 4);
  hipLaunchKernelGGL(k_pc_combine, dim3(G), dim3(PSFUN_BS), 0, c->stream, n, nshift, (const double2 *)X, ld,
                     (const double *)wg, form, form == PSFUN_QUAD_RE_SUM ? y : acc);
  PSFUN_HIP_TRY(hipGetLastError());
  if (form == PSFUN_QUAD_A_IM_SUM) PSFUN_TRY(launch_spmv(c, A, 1.0, acc, 0.0, y));
  for (int s = 0; s < nshift; ++s) {
    if (iters) iters[s] = its[s];
    if (resid) resid[s] = res[s];
  }
  return psfun_ctx_synchronize(c);
}
