// Preconditioner objects of the shifted solver (psfun_prec_*): the block
// Jacobi preconditioner PSBLAS builds for prec%init(ctxt, 'BJAC', info)
// (examples/shiftedkrylov/shifted.inp, applied by krylov/shifted/psfun_dcg.F90:283).
// With one process the only diagonal block is the whole matrix, and PSBLAS's
// default block solver is ILU(0), so M = L U with L unit lower and U upper
// triangular on the sparsity pattern of A.
//
// Build (host, as PSBLAS builds its preconditioners): the matrix comes back
// from the device, each row sorted by column; ILU(0) in the IKJ order
// (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 10.4):
//   for i: for k < i in row i (ascending): a_ik /= a_kk;
//            for j > k in row i: if (k, j) is in the pattern: a_ij -= a_ik a_kj
// with explicit fma, so the factors are bitwise reproducible.  A zero pivot
// is PSFUN_ERR_BREAKDOWN.  Then the level sets of the two triangular solves:
// level(i) = 1 + max level of the rows i depends on (lower: j < i in row i;
// upper: j > i), rows of a level listed in ascending order.
//
// Apply z = U^{-1} L^{-1} r on the device, one level after another: every row
// of a level is independent, one thread per row,
//   y_i = r_i - sum_{j<i} l_ij y_j        (terms in column order, fma)
//   z_i = (y_i - sum_{j>i} u_ij z_j) / u_ii
// Runs of levels of at most PSFUN_BS rows go to one single-block launch that
// separates its levels by block barriers (every row of a level is then done,
// and visible to the block, before the next level starts); larger levels get a
// grid each.  Triangular solves are latency-bound: the cost is the chain of
// levels (3m - 2 per solve for an m^3 7-point grid in natural order).
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

struct psfun_prec {
  psfun_ctx *ctx = nullptr;
  int device = 0;
  int kind = PSFUN_PREC_NONE;
  int n = 0;
  // ILU(0) factors (device, 0-based CSR): L strictly lower, U strictly upper,
  // ud the diagonal of U
  int *Lp = nullptr, *Lc = nullptr, *Up = nullptr, *Uc = nullptr;
  double *Lv = nullptr, *Uv = nullptr, *ud = nullptr;
  // level schedules: rows of each level, launch groups [first level, end level)
  int *lrows = nullptr, *urows = nullptr;
  std::vector<int> lptr, uptr;                 // level boundaries into lrows / urows (host copy)
  int *d_lptr = nullptr, *d_uptr = nullptr;    // the same on the device
  std::vector<std::pair<int, int>> lgroups, ugroups;
  double *y = nullptr;  // intermediate L^{-1} r
};

static void prec_free(psfun_prec *P) {
  for (void *p : {(void *)P->Lp, (void *)P->Lc, (void *)P->Up, (void *)P->Uc, (void *)P->Lv, (void *)P->Uv,
                  (void *)P->ud, (void *)P->lrows, (void *)P->urows, (void *)P->d_lptr, (void *)P->d_uptr,
                  (void *)P->y})
    if (p) hipFree(p);
}

// Level sets of a triangular solve: dep(i) lists the rows i reads.
template <typename Dep>
static void level_sets(int n, bool forward, Dep dep, std::vector<int> &ptr, std::vector<int> &rows) {
  std::vector<int> lev(n, 0);
  int nlev = 0;
  for (int t = 0; t < n; ++t) {
    const int i = forward ? t : n - 1 - t;
    int l = 0;
    dep(i, [&](int j) { l = std::max(l, lev[j] + 1); });
    lev[i] = l;
    nlev = std::max(nlev, l + 1);
  }
  ptr.assign(nlev + 1, 0);
  for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
  rows.assign(n, 0);
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < n; ++i) rows[pos[lev[i]]++] = i;  // ascending within a level
}

static void level_groups(const std::vector<int> &ptr, std::vector<std::pair<int, int>> &groups) {
  groups.clear();
  const int nlev = (int)ptr.size() - 1;
  int l = 0;
  while (l < nlev) {
    if (ptr[l + 1] - ptr[l] > PSFUN_BS) {
      groups.push_back({l, l + 1});
      ++l;
      continue;
    }
    int e = l;
    while (e < nlev && ptr[e + 1] - ptr[e] <= PSFUN_BS) ++e;
    groups.push_back({l, e});
    l = e;
  }
}

// y_i = r_i - sum_{j<i} l_ij y_j (LOWER) or z_i = (y_i - sum_{j>i} u_ij z_j) / u_ii
// for the rows of levels [l0, l1); a multi-level launch is one block.
template <bool LOWER>
__global__ __launch_bounds__(PSFUN_BS) void k_trsv_levels(int l0, int l1, const int *__restrict__ lvl_ptr,
                                                          const int *__restrict__ rows, const int *__restrict__ tp,
                                                          const int *__restrict__ tc, const double *__restrict__ tv,
                                                          const double *__restrict__ diag,
                                                          const double *__restrict__ rhs, double *__restrict__ out) {
  for (int l = l0; l < l1; ++l) {
    const int b = lvl_ptr[l], e = lvl_ptr[l + 1];
    for (int idx = b + (int)(blockIdx.x * PSFUN_BS + threadIdx.x); idx < e; idx += (int)(gridDim.x * PSFUN_BS)) {
      const int i = rows[idx];
      double t = rhs[i];
      for (int q = tp[i]; q < tp[i + 1]; ++q) t = fma(-tv[q], out[tc[q]], t);
      out[i] = LOWER ? t : t / diag[i];
    }
    if (l + 1 < l1) __syncthreads();
  }
}

// The same solves for S complex right-hand sides stored shift-major with
// leading dimension ld, for the systems s with run[s] != 0 (read from device
// memory, the same value for every thread): the real and the imaginary part
// take the real solve's operations one by one, so each equals the real solve
// of that part bit for bit.
template <bool LOWER>
__global__ __launch_bounds__(PSFUN_BS) void k_trsv_levels_c(int l0, int l1, const int *__restrict__ lvl_ptr,
                                                            const int *__restrict__ rows, const int *__restrict__ tp,
                                                            const int *__restrict__ tc, const double *__restrict__ tv,
                                                            const double *__restrict__ diag,
                                                            const double2 *__restrict__ rhs, double2 *__restrict__ out,
                                                            int64_t ld, int S, const int *__restrict__ run) {
  for (int l = l0; l < l1; ++l) {
    const int b = lvl_ptr[l], e = lvl_ptr[l + 1];
    for (int idx = b + (int)(blockIdx.x * PSFUN_BS + threadIdx.x); idx < e; idx += (int)(gridDim.x * PSFUN_BS)) {
      const int i = rows[idx];
      const int q0 = tp[i], q1 = tp[i + 1];
      for (int s = 0; s < S; ++s) {
        if (!run[s]) continue;
        const double2 *o = out + (size_t)s * ld;
        double2 t = rhs[(size_t)s * ld + i];
        for (int q = q0; q < q1; ++q) {
          const double v = tv[q];
          const double2 u = o[tc[q]];
          t.x = fma(-v, u.x, t.x);
          t.y = fma(-v, u.y, t.y);
        }
        if (!LOWER) {
          const double d = diag[i];
          t.x = t.x / d;
          t.y = t.y / d;
        }
        out[(size_t)s * ld + i] = t;
      }
    }
    if (l + 1 < l1) __syncthreads();
  }
}

int psfun_prec_launch_c(psfun_ctx *c, const psfun_prec *P, const double2 *r, double2 *z, double2 *ybuf, int64_t ld,
                        int S, const int *run) {
  if (P->kind != PSFUN_PREC_BJAC) return PSFUN_ERR_INVALID_ARG;
  for (const auto &g : P->lgroups) {
    const int rows = P->lptr[g.second] - P->lptr[g.first];
    const int G = (g.second - g.first > 1) ? 1 : std::min((rows + PSFUN_BS - 1) / PSFUN_BS, 8 * c->num_cu);
    hipLaunchKernelGGL(k_trsv_levels_c<true>, dim3(G), dim3(PSFUN_BS), 0, c->stream, g.first, g.second, P->d_lptr,
                       P->lrows, P->Lp, P->Lc, P->Lv, (const double *)nullptr, r, ybuf, ld, S, run);
  }
  for (const auto &g : P->ugroups) {
    const int rows = P->uptr[g.second] - P->uptr[g.first];
    const int G = (g.second - g.first > 1) ? 1 : std::min((rows + PSFUN_BS - 1) / PSFUN_BS, 8 * c->num_cu);
    hipLaunchKernelGGL(k_trsv_levels_c<false>, dim3(G), dim3(PSFUN_BS), 0, c->stream, g.first, g.second, P->d_uptr,
                       P->urows, P->Up, P->Uc, P->Uv, P->ud, (const double2 *)ybuf, z, ld, S, run);
  }
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

// z = M^{-1} r on the context stream (r and z distinct device vectors).
int psfun_prec_launch(psfun_ctx *c, const psfun_prec *P, const double *r, double *z) {
  if (P->kind != PSFUN_PREC_BJAC) return PSFUN_ERR_INVALID_ARG;
  for (const auto &g : P->lgroups) {
    const int rows = P->lptr[g.second] - P->lptr[g.first];
    const int G = (g.second - g.first > 1) ? 1 : std::min((rows + PSFUN_BS - 1) / PSFUN_BS, 8 * c->num_cu);
    hipLaunchKernelGGL(k_trsv_levels<true>, dim3(G), dim3(PSFUN_BS), 0, c->stream, g.first, g.second, P->d_lptr,
                       P->lrows, P->Lp, P->Lc, P->Lv, (const double *)nullptr, r, P->y);
  }
  for (const auto &g : P->ugroups) {
    const int rows = P->uptr[g.second] - P->uptr[g.first];
    const int G = (g.second - g.first > 1) ? 1 : std::min((rows + PSFUN_BS - 1) / PSFUN_BS, 8 * c->num_cu);
    hipLaunchKernelGGL(k_trsv_levels<false>, dim3(G), dim3(PSFUN_BS), 0, c->stream, g.first, g.second, P->d_uptr,
                       P->urows, P->Up, P->Uc, P->Uv, P->ud, (const double *)P->y, z);
  }
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

int psfun_prec_kind(const psfun_prec *P) { return P ? P->kind : PSFUN_PREC_NONE; }
int psfun_prec_size(const psfun_prec *P) { return P ? P->n : -1; }
int psfun_prec_levels(const psfun_prec *P) { return P ? (int)(P->lptr.size() + P->uptr.size()) - 2 : 0; }

template <typename T>
static hipError_t upload(T **d, const std::vector<T> &h) {
  hipError_t e = hipMalloc((void **)d, sizeof(T) * std::max<size_t>(h.size(), 1));
  if (e == hipSuccess && !h.empty()) e = hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
  return e;
}

extern "C" int psfun_prec_create(psfun_ctx_t c, psfun_csr_t A, int kind, psfun_prec_t *out) {
  if (!c || !A || !out) return PSFUN_ERR_INVALID_ARG;
  *out = nullptr;
  if (kind != PSFUN_PREC_NONE && kind != PSFUN_PREC_DIAG && kind != PSFUN_PREC_BJAC) return PSFUN_ERR_INVALID_ARG;
  auto fresh = [&]() {
    psfun_prec *P = new psfun_prec();
    P->ctx = c;
    P->device = c->device;
    P->kind = kind;
    P->n = A->n;
    return P;
  };
  if (kind != PSFUN_PREC_BJAC || A->n == 0) {  // NONE and DIAG keep no state (the solver forms 1/diag itself)
    *out = fresh();
    return PSFUN_SUCCESS;
  }
  const int n = A->n;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<int> rp(n + 1), ci(A->nnz);
  std::vector<double> va(A->nnz);
  PSFUN_HIP_TRY(hipMemcpy(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  if (A->nnz) {
    PSFUN_HIP_TRY(hipMemcpy(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
    PSFUN_HIP_TRY(hipMemcpy(va.data(), A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost));
  }
  // rows sorted by column, repeated entries summed (the matrix the SpMV applies)
  std::vector<int> sp(n + 1, 0), sc;
  std::vector<double> sv;
  sc.reserve(A->nnz);
  sv.reserve(A->nnz);
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; ++i) {
    row.clear();
    for (int q = rp[i]; q < rp[i + 1]; ++q) row.push_back({ci[q], va[q]});
    std::stable_sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    for (size_t t = 0; t < row.size(); ++t) {
      if (!sc.empty() && (int)sc.size() > sp[i] && sc.back() == row[t].first) {
        sv.back() += row[t].second;
      } else {
        sc.push_back(row[t].first);
        sv.push_back(row[t].second);
      }
    }
    sp[i + 1] = (int)sc.size();
  }
  // ILU(0), IKJ order, in place
  std::vector<int> dpos(n, -1), where(n, -1);
  int st = PSFUN_SUCCESS;
  for (int i = 0; i < n && st == PSFUN_SUCCESS; ++i) {
    for (int q = sp[i]; q < sp[i + 1]; ++q) where[sc[q]] = q;
    for (int q = sp[i]; q < sp[i + 1]; ++q) {
      const int k = sc[q];
      if (k >= i) break;
      const int dk = dpos[k];
      if (dk < 0 || sv[dk] == 0.0) {
        st = PSFUN_ERR_BREAKDOWN;
        break;
      }
      const double lik = sv[q] / sv[dk];
      sv[q] = lik;
      for (int t = dk + 1; t < sp[k + 1]; ++t) {  // u_kj, j > k
        const int wj = where[sc[t]];
        if (wj >= 0) sv[wj] = fma(-lik, sv[t], sv[wj]);
      }
    }
    for (int q = sp[i]; q < sp[i + 1]; ++q) {
      if (sc[q] == i) dpos[i] = q;
      where[sc[q]] = -1;
    }
    if (st == PSFUN_SUCCESS && (dpos[i] < 0 || sv[dpos[i]] == 0.0)) st = PSFUN_ERR_BREAKDOWN;
  }
  if (st != PSFUN_SUCCESS) return st;
  psfun_prec *P = fresh();
  std::vector<int> lp(n + 1, 0), lc, up(n + 1, 0), uc;
  std::vector<double> lv, uv, ud(n);
  for (int i = 0; i < n; ++i) {
    for (int q = sp[i]; q < sp[i + 1]; ++q) {
      if (sc[q] < i) {
        lc.push_back(sc[q]);
        lv.push_back(sv[q]);
      } else if (sc[q] > i) {
        uc.push_back(sc[q]);
        uv.push_back(sv[q]);
      } else {
        ud[i] = sv[q];
      }
    }
    lp[i + 1] = (int)lc.size();
    up[i + 1] = (int)uc.size();
  }
  std::vector<int> lrows, urows;
  level_sets(n, true, [&](int i, auto use) { for (int q = lp[i]; q < lp[i + 1]; ++q) use(lc[q]); }, P->lptr, lrows);
  level_sets(n, false, [&](int i, auto use) { for (int q = up[i]; q < up[i + 1]; ++q) use(uc[q]); }, P->uptr, urows);
  level_groups(P->lptr, P->lgroups);
  level_groups(P->uptr, P->ugroups);
  hipError_t e = upload(&P->Lp, lp);
  if (e == hipSuccess) e = upload(&P->Lc, lc);
  if (e == hipSuccess) e = upload(&P->Lv, lv);
  if (e == hipSuccess) e = upload(&P->Up, up);
  if (e == hipSuccess) e = upload(&P->Uc, uc);
  if (e == hipSuccess) e = upload(&P->Uv, uv);
  if (e == hipSuccess) e = upload(&P->ud, ud);
  if (e == hipSuccess) e = upload(&P->lrows, lrows);
  if (e == hipSuccess) e = upload(&P->urows, urows);
  if (e == hipSuccess) e = upload(&P->d_lptr, P->lptr);
  if (e == hipSuccess) e = upload(&P->d_uptr, P->uptr);
  if (e == hipSuccess) e = hipMalloc((void **)&P->y, sizeof(double) * n);
  if (e != hipSuccess) {
    psfun_record_hip_error(e, "psfun_prec_create", __FILE__, __LINE__);
    prec_free(P);
    delete P;
    return PSFUN_ERR_ALLOC;
  }
  *out = P;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_prec_destroy(psfun_prec_t P) {
  if (!P) return PSFUN_SUCCESS;
  hipSetDevice(P->device);  // the context may already be gone
  hipDeviceSynchronize();
  prec_free(P);
  delete P;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_prec_info(psfun_prec_t P, int *kind, int *levels) {
  if (!P) return PSFUN_ERR_INVALID_ARG;
  if (kind) *kind = P->kind;
  if (levels) *levels = psfun_prec_levels(P);
  return PSFUN_SUCCESS;
}

extern "C" int psfun_prec_apply(psfun_ctx_t c, psfun_prec_t P, const double *r, double *z) {
  if (!c || !P) return PSFUN_ERR_INVALID_ARG;
  if (P->n == 0) return PSFUN_SUCCESS;
  if (!r || !z) return PSFUN_ERR_INVALID_VECT_STATE;
  if (P->kind != PSFUN_PREC_BJAC) return PSFUN_ERR_UNSUPPORTED;  // NONE/DIAG are applied inside the solver
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  return psfun_prec_launch(c, P, r, z);
}
