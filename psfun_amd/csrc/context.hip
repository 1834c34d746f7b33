// Context, workspace, CSR handles and device vectors of the C-ABI.
#include "common.hpp"
#include "kernels.hpp"

#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <chrono>

static thread_local char g_last_hip_error[512];

void psfun_record_hip_error(hipError_t e, const char *what, const char *file, int line) {
  snprintf(g_last_hip_error, sizeof(g_last_hip_error), "%s: %s (%s:%d)", hipGetErrorString(e), what,
           file, line);
}

extern "C" const char *psfun_strerror(int status) {
  switch (status) {
    case PSFUN_SUCCESS: return "success";
    case PSFUN_ERR_INVALID_ARG: return "invalid argument";
    case PSFUN_ERR_INVALID_ISTOP: return "invalid stopping criterion (istop must be 1 or 2)";
    case PSFUN_ERR_INVALID_VECT_STATE: return "vector not allocated";
    case PSFUN_ERR_INVALID_ARGS_COMBINATION: return "invalid combination of arguments";
    case PSFUN_ERR_NOT_SQUARE: return "matrix is not square";
    case PSFUN_ERR_UNSUPPORTED: return "function/variant not available on the device";
    case PSFUN_ERR_ALLOC: return "allocation failed";
    case PSFUN_ERR_HIP: return g_last_hip_error[0] ? g_last_hip_error : "HIP runtime error";
    case PSFUN_ERR_NO_DEVICE: return "no HIP device";
    case PSFUN_ERR_BREAKDOWN: return "solver breakdown";
    case PSFUN_ERR_TOO_LARGE: return "input exceeds the 64 MiB fixed scratch of the allocation-free entry point";
    default: return "unknown status";
  }
}

extern "C" int psfun_ctx_create(int device, psfun_ctx_t *out) {
  if (!out) return PSFUN_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PSFUN_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipSetDevice(device));
  psfun_ctx *c = new psfun_ctx();
  c->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->num_cu = cus;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->d_partials, sizeof(double) * PSFUN_MAX_GRID * 8);
  if (e == hipSuccess) e = hipMalloc(&c->d_scalar, sizeof(double) * 256);
  if (e == hipSuccess) e = hipHostMalloc(&c->h_pinned, PSFUN_PINNED_BYTES, hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&c->h_flags, 256, hipHostMallocCoherent | hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&c->d_flags, (void *)c->h_flags, 0);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    psfun_record_hip_error(e, "psfun_ctx_create", __FILE__, __LINE__);
    psfun_ctx_destroy(c);
    return PSFUN_ERR_HIP;
  }
  // the device address of this device's allocation-free scratch (once per device)
  const int st = psfun_noalloc_setup(device);
  if (st != PSFUN_SUCCESS) {
    psfun_ctx_destroy(c);
    return st;
  }
  c->stream = c->own_stream;
  *out = c;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_ctx_destroy(psfun_ctx_t c) {
  if (!c) return PSFUN_SUCCESS;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->own_stream) hipStreamSynchronize(c->own_stream);
  if (c->aux_stream) hipStreamSynchronize(c->aux_stream);
  if (c->ws) hipFree(c->ws);
  if (c->ws2) hipFree(c->ws2);
  if (c->d_partials) hipFree(c->d_partials);
  if (c->d_scalar) hipFree(c->d_scalar);
  if (c->h_pinned) hipHostFree(c->h_pinned);
  if (c->h_flags) hipHostFree((void *)c->h_flags);
  for (auto &e : c->ev)
    if (e) hipEventDestroy(e);
  for (auto &e : c->prof_ev) hipEventDestroy(e);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->aux_stream) hipStreamDestroy(c->aux_stream);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_ctx_set_stream(psfun_ctx_t c, void *s) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  c->stream = (hipStream_t)s;  // NULL selects the HIP null stream
  return PSFUN_SUCCESS;
}

extern "C" int psfun_set_userf(psfun_ctx_t c, psfun_userf_t f, void *user) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  c->userf = f;
  c->userf_data = user;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_set_pade_degree(psfun_ctx_t c, int degree) {
  if (!c || degree < 2 || degree > PSFUN_PADE_MAX) return PSFUN_ERR_INVALID_ARG;
  c->pade_degree = degree;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_ctx_use_own_stream(psfun_ctx_t c) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  c->stream = c->own_stream;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_ctx_synchronize(psfun_ctx_t c) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSFUN_SUCCESS;
}

extern "C" int psfun_ctx_num_cus(psfun_ctx_t c, int *n) {
  if (!c || !n) return PSFUN_ERR_INVALID_ARG;
  *n = c->num_cu;
  return PSFUN_SUCCESS;
}

int psfun_host_wait(psfun_ctx *c, int target) {
  volatile int *hf = c->h_flags;
  // Read the flags in a tight loop; ask the runtime whether the stream has
  // drained (after which every flag store is visible) only every 50 us, since
  // a stream query costs far more than a flag read.
  auto next = std::chrono::steady_clock::now() + std::chrono::microseconds(50);
  while (hf[0] < target && hf[1] == 0) {
    if (std::chrono::steady_clock::now() < next) continue;
    const hipError_t q = hipStreamQuery(c->stream);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) {
      psfun_record_hip_error(q, "hipStreamQuery", __FILE__, __LINE__);
      return PSFUN_ERR_HIP;
    }
    next = std::chrono::steady_clock::now() + std::chrono::microseconds(50);
  }
  return PSFUN_SUCCESS;
}

static int grow_workspace(psfun_ctx *c, void **buf, size_t *have, size_t bytes, void **ptr) {
  bytes = align_up(bytes < 256 ? 256 : bytes, 256);
  if (bytes > *have) {
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    if (*buf) PSFUN_HIP_TRY(hipFree(*buf));
    *buf = nullptr;
    *have = 0;
    size_t want = align_up(bytes + bytes / 8, 1 << 20);
    if (hipMalloc(buf, want) != hipSuccess) {
      (void)hipGetLastError();
      if (hipMalloc(buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *buf = nullptr;
        return PSFUN_ERR_ALLOC;
      }
      want = bytes;
    }
    *have = want;
  }
  *ptr = *buf;
  return PSFUN_SUCCESS;
}

int psfun_workspace(psfun_ctx *c, size_t bytes, void **ptr) { return grow_workspace(c, &c->ws, &c->ws_bytes, bytes, ptr); }
int psfun_workspace2(psfun_ctx *c, size_t bytes, void **ptr) {
  return grow_workspace(c, &c->ws2, &c->ws2_bytes, bytes, ptr);
}

// Lanes per row: the smallest power of two >= half the mean row length
// (each lane then handles about two nonzeros), between 2 and 64.
static int choose_lanes(int64_t n, int64_t nnz) {
  double avg = n > 0 ? (double)nnz / (double)n : 1.0;
  int l = 2;
  while (l < 0.5 * avg && l < 64) l <<= 1;
  return l;
}

// Sliced-ELL fill: thread r covers row r of slice r/64 (rows >= n pad the
// last slice).  Writes colind only when `cols` is set (value refresh skips it).
__global__ __launch_bounds__(PSFUN_BS) void k_sell_fill(int n, int nslice, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const double *__restrict__ va,
                                                        const int *__restrict__ sp, int *__restrict__ scol,
                                                        double *__restrict__ sval, int cols) {
  const int64_t r = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x;
  if (r >= (int64_t)nslice * PSFUN_SELL_C) return;
  const int s = (int)(r / PSFUN_SELL_C), l = (int)(r % PSFUN_SELL_C);
  const int b = sp[s];
  const int w = (sp[s + 1] - b) / PSFUN_SELL_C;
  const int start = r < n ? rp[r] : 0;
  const int len = r < n ? rp[r + 1] - start : 0;
  const int self = r < n ? (int)r : 0;
  for (int k = 0; k < w; ++k) {
    const int64_t idx = (int64_t)b + (int64_t)k * PSFUN_SELL_C + l;
    if (k < len) {
      if (cols) scol[idx] = ci[start + k];
      sval[idx] = va[start + k];
    } else {
      if (cols) scol[idx] = self;
      sval[idx] = 0.0;
    }
  }
}

// 16-bit offsets col - row of every sliced-ELL entry; *bad is set when an
// offset does not fit.  Padding entries have offset 0, and so do the rows
// beyond n that fill the last slice (no kernel computes them).
__global__ __launch_bounds__(PSFUN_BS) void k_sell_delta(int n, int nslice, const int *__restrict__ sp,
                                                         const int *__restrict__ scol, int16_t *__restrict__ dcol,
                                                         int *bad) {
  const int64_t r = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x;
  if (r >= (int64_t)nslice * PSFUN_SELL_C) return;
  const int s = (int)(r / PSFUN_SELL_C), l = (int)(r % PSFUN_SELL_C);
  const int b = sp[s];
  const int w = (sp[s + 1] - b) / PSFUN_SELL_C;
  int miss = 0;
  for (int k = 0; k < w; ++k) {
    const int64_t idx = (int64_t)b + (int64_t)k * PSFUN_SELL_C + l;
    const int64_t d = (r < n) ? (int64_t)scol[idx] - r : 0;
    const bool fits = d >= -32768 && d <= 32767;
    if (!fits) miss = 1;
    dcol[idx] = (int16_t)(fits ? d : 0);
  }
  if (miss) atomicOr(bad, 1);
}

static void sell_free(psfun_csr *A) {
  if (A->sell_ptr) hipFree(A->sell_ptr);
  if (A->sell_col) hipFree(A->sell_col);
  if (A->sell_val) hipFree(A->sell_val);
  if (A->sell_dcol) hipFree(A->sell_dcol);
  A->sell_ptr = A->sell_col = nullptr;
  A->sell_val = nullptr;
  A->sell_dcol = nullptr;
  A->sell_nnz = 0;
  A->nslice = 0;
}

static int sell_launch_fill(psfun_csr *A, int cols) {
  if (!A->sell_nnz || A->n == 0) return PSFUN_SUCCESS;
  const int64_t threads = (int64_t)A->nslice * PSFUN_SELL_C;
  hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)((threads + PSFUN_BS - 1) / PSFUN_BS)), dim3(PSFUN_BS), 0,
                     A->ctx->stream, A->n, A->nslice, A->rowptr, A->colind, A->val, A->sell_ptr, A->sell_col,
                     A->sell_val, cols);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

int psfun_sell_build(psfun_csr *A, const int *rp) {
  sell_free(A);
  const int n = A->n;
  if (n == 0) return PSFUN_SUCCESS;
  const int ns = (n + PSFUN_SELL_C - 1) / PSFUN_SELL_C;
  std::vector<int> sp((size_t)ns + 1);
  int64_t tot = 0;
  for (int s = 0; s < ns; ++s) {
    int w = 0;
    const int r1 = std::min(n, (s + 1) * PSFUN_SELL_C);
    for (int r = s * PSFUN_SELL_C; r < r1; ++r) w = std::max(w, rp[r + 1] - rp[r]);
    if (tot > INT32_MAX) return PSFUN_SUCCESS;  // offsets would not fit: CSR only
    sp[s] = (int)tot;
    tot += (int64_t)w * PSFUN_SELL_C;
  }
  if (tot > INT32_MAX) return PSFUN_SUCCESS;
  sp[ns] = (int)tot;
  hipError_t e = hipMalloc(&A->sell_ptr, sizeof(int) * ((size_t)ns + 1));
  if (e == hipSuccess) e = hipMalloc(&A->sell_col, sizeof(int) * (size_t)std::max<int64_t>(tot, 1));
  if (e == hipSuccess) e = hipMalloc(&A->sell_val, sizeof(double) * (size_t)std::max<int64_t>(tot, 1));
  if (e == hipSuccess) e = hipMemcpy(A->sell_ptr, sp.data(), sizeof(int) * ((size_t)ns + 1), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    sell_free(A);
    return PSFUN_SUCCESS;  // no room for the copy: CSR only
  }
  A->nslice = ns;
  A->sell_nnz = tot;
  PSFUN_TRY(sell_launch_fill(A, 1));
  // 16-bit column offsets when all of them fit (kept beside the 32-bit columns)
  if (hipMalloc(&A->sell_dcol, sizeof(int16_t) * (size_t)std::max<int64_t>(tot, 1)) == hipSuccess) {
    int *bad = (int *)A->ctx->d_scalar;
    int hbad = 0;
    PSFUN_HIP_TRY(hipMemcpyAsync(bad, &hbad, sizeof(int), hipMemcpyHostToDevice, A->ctx->stream));
    const int64_t threads = (int64_t)ns * PSFUN_SELL_C;
    hipLaunchKernelGGL(k_sell_delta, dim3((unsigned)((threads + PSFUN_BS - 1) / PSFUN_BS)), dim3(PSFUN_BS), 0,
                       A->ctx->stream, n, ns, A->sell_ptr, A->sell_col, A->sell_dcol, bad);
    PSFUN_HIP_TRY(hipGetLastError());
    PSFUN_HIP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, A->ctx->stream));
    PSFUN_HIP_TRY(hipStreamSynchronize(A->ctx->stream));
    if (hbad) {
      hipFree(A->sell_dcol);
      A->sell_dcol = nullptr;
    }
  } else {
    (void)hipGetLastError();
    A->sell_dcol = nullptr;
  }
  PSFUN_HIP_TRY(hipStreamSynchronize(A->ctx->stream));
  return PSFUN_SUCCESS;
}

// Values of the diagonal slices from the CSR arrays: thread r zeroes its
// slots, then adds its entries in CSR order at their offsets' slots (binary
// search in the slice's ascending offsets).  Threads of the rows beyond n that
// fill the last slice only zero theirs.
__global__ __launch_bounds__(PSFUN_BS) void k_hdia_fill(int n, int nslice, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const double *__restrict__ va,
                                                        const int *__restrict__ hp, const int *__restrict__ off,
                                                        double *__restrict__ hv) {
  const int64_t r = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x;
  if (r >= (int64_t)nslice * PSFUN_SELL_C) return;
  const int s = (int)(r / PSFUN_SELL_C), l = (int)(r % PSFUN_SELL_C);
  const int b = hp[s], e = hp[s + 1];
  for (int k = b; k < e; ++k) hv[(int64_t)k * PSFUN_SELL_C + l] = 0.0;
  if (r >= n) return;
  for (int q = rp[r]; q < rp[r + 1]; ++q) {
    const int d = ci[q] - (int)r;
    int lo = b, hi = e;  // first k with off[k] >= d (the list holds d)
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (off[mid] < d)
        lo = mid + 1;
      else
        hi = mid;
    }
    hv[(int64_t)lo * PSFUN_SELL_C + l] += va[q];
  }
}

static void hdia_free(psfun_csr *A) {
  if (A->hdia_ptr) hipFree(A->hdia_ptr);
  if (A->hdia_off) hipFree(A->hdia_off);
  if (A->hdia_val) hipFree(A->hdia_val);
  A->hdia_ptr = A->hdia_off = nullptr;
  A->hdia_val = nullptr;
  A->hdia_ndiag = 0;
}

static int hdia_launch_fill(psfun_csr *A) {
  if (!A->hdia_ptr || A->n == 0) return PSFUN_SUCCESS;
  const int ns = (A->n + PSFUN_SELL_C - 1) / PSFUN_SELL_C;
  const int64_t threads = (int64_t)ns * PSFUN_SELL_C;
  hipLaunchKernelGGL(k_hdia_fill, dim3((unsigned)((threads + PSFUN_BS - 1) / PSFUN_BS)), dim3(PSFUN_BS), 0,
                     A->ctx->stream, A->n, ns, A->rowptr, A->colind, A->val, A->hdia_ptr, A->hdia_off, A->hdia_val);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

// Diagonal slices: the offsets of each slice on the host (from the device
// CSR pattern), the values by k_hdia_fill.  PSFUN_ERR_UNSUPPORTED when the
// slices would hold more than 1.5 x nnz values.
static int hdia_build(psfun_csr *A) {
  hdia_free(A);
  const int n = A->n;
  const int ns = (n + PSFUN_SELL_C - 1) / PSFUN_SELL_C;
  const double limit = 1.5 * (double)std::max<int64_t>(A->nnz, 1);
  std::vector<int> rp((size_t)n + 1), ci((size_t)std::max<int64_t>(A->nnz, 1));
  PSFUN_HIP_TRY(hipStreamSynchronize(A->ctx->stream));
  PSFUN_HIP_TRY(hipMemcpy(rp.data(), A->rowptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost));
  if (A->nnz) PSFUN_HIP_TRY(hipMemcpy(ci.data(), A->colind, sizeof(int) * (size_t)A->nnz, hipMemcpyDeviceToHost));
  std::vector<int> hp((size_t)ns + 1), offs, tmp;
  double stored = 0.0;  // values of the rows that exist (the last slice may be partial)
  for (int s = 0; s < ns; ++s) {
    tmp.clear();
    const int r1 = std::min(n, (s + 1) * PSFUN_SELL_C);
    for (int r = s * PSFUN_SELL_C; r < r1; ++r)
      for (int q = rp[r]; q < rp[r + 1]; ++q) tmp.push_back(ci[q] - r);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    hp[s] = (int)offs.size();
    offs.insert(offs.end(), tmp.begin(), tmp.end());
    stored += (double)tmp.size() * (r1 - s * PSFUN_SELL_C);
    if (stored > limit || offs.size() > (size_t)INT32_MAX / PSFUN_SELL_C) return PSFUN_ERR_UNSUPPORTED;
  }
  hp[ns] = (int)offs.size();
  const size_t nd = offs.size();
  hipError_t e = hipMalloc(&A->hdia_ptr, sizeof(int) * ((size_t)ns + 1));
  if (e == hipSuccess) e = hipMalloc(&A->hdia_off, sizeof(int) * std::max<size_t>(nd, 1));
  if (e == hipSuccess) e = hipMalloc(&A->hdia_val, sizeof(double) * std::max<size_t>(nd * PSFUN_SELL_C, 1));
  if (e == hipSuccess) e = hipMemcpy(A->hdia_ptr, hp.data(), sizeof(int) * ((size_t)ns + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && nd) e = hipMemcpy(A->hdia_off, offs.data(), sizeof(int) * nd, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hdia_free(A);
    return PSFUN_ERR_ALLOC;
  }
  A->hdia_ndiag = (int64_t)nd;
  PSFUN_TRY(hdia_launch_fill(A));
  PSFUN_HIP_TRY(hipStreamSynchronize(A->ctx->stream));
  return PSFUN_SUCCESS;
}

int psfun_sell_refill(psfun_csr *A) {
  PSFUN_TRY(sell_launch_fill(A, 0));
  return hdia_launch_fill(A);
}

static void csrs_free(psfun_csr *A) {
  for (void *p : {(void *)A->csrs_blk, (void *)A->csrs_split, (void *)A->csrs_cnt, (void *)A->csrs_seg})
    if (p) hipFree(p);
  A->csrs_blk = A->csrs_split = A->csrs_cnt = nullptr;
  A->csrs_seg = nullptr;
  A->csrs_nblk = A->csrs_nsplit = A->csrs_nxt = 0;
}

// Blocks of the CSR-stream step, greedy over the rows: a block of short rows
// closes before it would pass PSFUN_BS rows or PSFUN_CSRS_CAP entries; a row
// with more entries is a block of its own, and a row with more than
// PSFUN_CSRS_SPLIT (the first PSFUN_CSRS_MAXSPLIT such rows) is cut into
// segments of PSFUN_CSRS_SEG entries, one block each.
// A block of short rows whose columns lie within PSFUN_CSRS_XT consecutive
// values is marked kind = -span with its first column in c: the step stages
// that window of x in LDS with coalesced loads instead of gathering it.
static int csrs_build(psfun_csr *A) {
  const int n = A->n;
  std::vector<int> rp(n + 1);
  PSFUN_HIP_TRY(hipMemcpy(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  std::vector<int> ci((size_t)std::max<int64_t>(A->nnz, 1));
  if (A->nnz) PSFUN_HIP_TRY(hipMemcpy(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
  std::vector<int> blk, split;
  blk.reserve(4 * (n / 64 + 16));
  auto add = [&](int a, int b, int c, int kind) {
    blk.push_back(a);
    blk.push_back(b);
    blk.push_back(c);
    blk.push_back(kind);
  };
  int r = 0;
  while (r < n) {
    const int e0 = rp[r], len = rp[r + 1] - e0;
    if (len <= PSFUN_CSRS_CAP) {
      const int start = r;
      while (r < n && r - start < PSFUN_BS && rp[r + 1] - e0 <= PSFUN_CSRS_CAP) ++r;
      int cmin = INT32_MAX, cmax = -1;
      for (int q = e0; q < rp[r]; ++q) {
        cmin = std::min(cmin, ci[q]);
        cmax = std::max(cmax, ci[q]);
      }
      const int64_t span = (cmax >= cmin) ? (int64_t)cmax - cmin + 1 : 0;
      if (span >= 1 && span <= PSFUN_CSRS_XT)
        add(start, r, cmin, -(int)span);
      else
        add(start, r, 0, 0);
    } else if (len <= PSFUN_CSRS_SPLIT || (int)split.size() / 2 >= PSFUN_CSRS_MAXSPLIT) {
      add(r, r + 1, 0, 1);
      ++r;
    } else {
      const int s = (int)split.size() / 2, first = (int)blk.size() / 4;
      int nseg = 0;
      for (int q = e0; q < rp[r + 1]; q += PSFUN_CSRS_SEG, ++nseg) add(r, q, std::min(q + PSFUN_CSRS_SEG, rp[r + 1]), 2 + s);
      split.push_back(first);
      split.push_back(nseg);
      ++r;
    }
  }
  const int nblk = (int)blk.size() / 4, nsplit = (int)split.size() / 2;
  hipError_t e = hipMalloc(&A->csrs_blk, sizeof(int) * std::max<size_t>(blk.size(), 4));
  if (e == hipSuccess && !blk.empty())
    e = hipMemcpy(A->csrs_blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&A->csrs_split, sizeof(int) * std::max<size_t>(split.size(), 2));
  if (e == hipSuccess && !split.empty())
    e = hipMemcpy(A->csrs_split, split.data(), sizeof(int) * split.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&A->csrs_cnt, sizeof(int) * std::max(nsplit, 1));
  if (e == hipSuccess) e = hipMemset(A->csrs_cnt, 0, sizeof(int) * std::max(nsplit, 1));
  if (e == hipSuccess) e = hipMalloc(&A->csrs_seg, sizeof(double) * std::max(nblk, 1));
  if (e != hipSuccess) {
    psfun_record_hip_error(e, "csrs_build", __FILE__, __LINE__);
    csrs_free(A);
    return PSFUN_ERR_ALLOC;
  }
  A->csrs_nblk = nblk;
  A->csrs_nsplit = nsplit;
  A->csrs_nxt = 0;
  for (int b = 0; b < nblk; ++b) A->csrs_nxt += (blk[4 * b + 3] < 0);
  return PSFUN_SUCCESS;
}

static int auto_format(const psfun_csr *A) {
  if (A->sell_nnz > 0 && (double)A->sell_nnz <= 1.25 * (double)std::max<int64_t>(A->nnz, 1))
    return PSFUN_FORMAT_SELL;
  return PSFUN_FORMAT_CSR;
}

extern "C" int psfun_csr_create(psfun_ctx_t c, int64_t n, int64_t nnz, const int32_t *rowptr,
                                const int32_t *colind, const double *val, int base, int mem,
                                psfun_csr_t *out) {
  if (!c || !out || n < 0 || nnz < 0 || n > INT32_MAX - 1 || nnz > INT32_MAX) return PSFUN_ERR_INVALID_ARG;
  if (base != 0 && base != 1) return PSFUN_ERR_INVALID_ARG;
  if ((n > 0 && !rowptr) || (nnz > 0 && (!colind || !val))) return PSFUN_ERR_INVALID_ARG;
  *out = nullptr;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  psfun_csr *A = new psfun_csr();
  A->ctx = c;
  A->device = c->device;
  A->n = (int)n;
  A->nnz = nnz;
  std::vector<int> host_rp;
  if (mem == PSFUN_MEM_DEVICE) {
    if (base != 0) {
      delete A;
      return PSFUN_ERR_INVALID_ARG;
    }
    A->rowptr = const_cast<int *>(rowptr);
    A->colind = const_cast<int *>(colind);
    A->val = const_cast<double *>(val);
    A->owns = false;
    A->max_row = -1;
  } else if (mem == PSFUN_MEM_HOST) {
    std::vector<int> rp(rowptr, rowptr + n + 1), ci(colind, colind + nnz);
    if (rp[0] != base || rp[n] - base != nnz) {
      delete A;
      return PSFUN_ERR_INVALID_ARG;
    }
    int mx = 0;
    for (int64_t i = 0; i <= n; ++i) rp[i] -= base;
    for (int64_t i = 0; i < n; ++i) {
      int len = rp[i + 1] - rp[i];
      if (len < 0) {
        delete A;
        return PSFUN_ERR_INVALID_ARG;
      }
      mx = std::max(mx, len);
    }
    for (int64_t j = 0; j < nnz; ++j) {
      ci[j] -= base;
      if (ci[j] < 0 || ci[j] >= n) {
        delete A;
        return PSFUN_ERR_INVALID_ARG;
      }
    }
    A->max_row = mx;
    hipError_t e = hipMalloc(&A->rowptr, sizeof(int) * (n + 1));
    if (e == hipSuccess) e = hipMalloc(&A->colind, sizeof(int) * std::max<int64_t>(nnz, 1));
    if (e == hipSuccess) e = hipMalloc(&A->val, sizeof(double) * std::max<int64_t>(nnz, 1));
    if (e == hipSuccess) e = hipMemcpy(A->rowptr, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(A->colind, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(A->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice);
    A->owns = true;
    host_rp.swap(rp);
    if (e != hipSuccess) {
      psfun_record_hip_error(e, "psfun_csr_create", __FILE__, __LINE__);
      psfun_csr_destroy(A);
      return PSFUN_ERR_HIP;
    }
  } else {
    delete A;
    return PSFUN_ERR_INVALID_ARG;
  }
  A->lanes = choose_lanes(n, nnz);
  {
    std::vector<int> rph;
    const int *rpp = nullptr;
    if (mem == PSFUN_MEM_DEVICE) {
      rph.resize((size_t)n + 1);
      PSFUN_HIP_TRY(hipMemcpy(rph.data(), A->rowptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost));
      rpp = rph.data();
    } else {
      rpp = host_rp.data();
    }
    const int st = psfun_sell_build(A, rpp);
    if (st != PSFUN_SUCCESS) {
      psfun_csr_destroy(A);
      return st;
    }
  }
  A->format = auto_format(A);
  *out = A;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_destroy(psfun_csr_t A) {
  if (!A) return PSFUN_SUCCESS;
  // The context may already be gone (a caller that destroys it first): only
  // the device recorded at creation is used here, and a device-wide wait
  // stands in for the context stream's.
  if (A->owns || A->sell_nnz || A->sell_ptr || A->hdia_ptr || A->csrs_blk) {
    hipSetDevice(A->device);
    hipDeviceSynchronize();
  }
  csrs_free(A);
  if (A->owns) {
    if (A->rowptr) hipFree(A->rowptr);
    if (A->colind) hipFree(A->colind);
    if (A->val) hipFree(A->val);
  }
  if (A->sell_nnz || A->sell_ptr || A->hdia_ptr) {
    sell_free(A);
    hdia_free(A);
  }
  delete A;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_set_values(psfun_csr_t A, const double *val) {
  if (!A || (!val && A->nnz)) return PSFUN_ERR_INVALID_ARG;
  if (A->owns) {
    if (A->nnz)
      PSFUN_HIP_TRY(hipMemcpyAsync(A->val, val, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice,
                                   A->ctx->stream));
  } else {
    A->val = const_cast<double *>(val);
  }
  return psfun_sell_refill(A);
}

extern "C" int psfun_csr_set_lanes(psfun_csr_t A, int lanes) {
  if (!A) return PSFUN_ERR_INVALID_ARG;
  if (lanes == 0) {
    A->lanes = choose_lanes(A->n, A->nnz);
    return PSFUN_SUCCESS;
  }
  if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1))) return PSFUN_ERR_INVALID_ARG;
  A->lanes = lanes;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_info(psfun_csr_t A, int64_t *n, int64_t *nnz, int *lanes) {
  if (!A) return PSFUN_ERR_INVALID_ARG;
  if (n) *n = A->n;
  if (nnz) *nnz = A->nnz;
  if (lanes) *lanes = A->lanes;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_set_format(psfun_csr_t A, int format) {
  if (!A) return PSFUN_ERR_INVALID_ARG;
  if (format == PSFUN_FORMAT_AUTO) {
    A->format = auto_format(A);
  } else if (format == PSFUN_FORMAT_CSR) {
    A->format = PSFUN_FORMAT_CSR;
  } else if (format == PSFUN_FORMAT_SELL) {
    if (A->sell_nnz == 0 && A->n > 0) return PSFUN_ERR_UNSUPPORTED;
    A->format = PSFUN_FORMAT_SELL;
  } else if (format == PSFUN_FORMAT_HDIA) {
    if (!A->hdia_ptr) {
      PSFUN_HIP_TRY(hipSetDevice(A->ctx->device));
      PSFUN_TRY(hdia_build(A));
    }
    A->format = PSFUN_FORMAT_HDIA;
  } else if (format == PSFUN_FORMAT_CSRS) {
    if (!A->csrs_blk && A->n > 0) {
      PSFUN_HIP_TRY(hipSetDevice(A->ctx->device));
      PSFUN_HIP_TRY(hipStreamSynchronize(A->ctx->stream));
      PSFUN_TRY(csrs_build(A));
    }
    A->format = PSFUN_FORMAT_CSRS;
  } else {
    return PSFUN_ERR_INVALID_ARG;
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_get_format(psfun_csr_t A, int *format, int64_t *sell_entries) {
  if (!A) return PSFUN_ERR_INVALID_ARG;
  if (format) *format = A->format;
  if (sell_entries) *sell_entries = A->sell_nnz;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_hdia_diagonals(psfun_csr_t A, int64_t *ndiag) {
  if (!A || !ndiag) return PSFUN_ERR_INVALID_ARG;
  *ndiag = A->hdia_ptr ? A->hdia_ndiag : 0;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_csr_sell_index_bits(psfun_csr_t A, int *bits) {
  if (!A || !bits) return PSFUN_ERR_INVALID_ARG;
  *bits = A->sell_nnz == 0 ? 0 : (A->sell_dcol ? 16 : 32);
  return PSFUN_SUCCESS;
}

extern "C" int psfun_vec_alloc(psfun_ctx_t c, int64_t n, double **d) {
  if (!c || !d || n < 0) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  if (hipMalloc((void **)d, sizeof(double) * (n > 0 ? n : 1)) != hipSuccess) {
    (void)hipGetLastError();
    *d = nullptr;
    return PSFUN_ERR_ALLOC;
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_vec_free(psfun_ctx_t c, double *d) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  if (d) {
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    PSFUN_HIP_TRY(hipFree(d));
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_vec_upload(psfun_ctx_t c, double *d, const double *h, int64_t n) {
  if (!c || n < 0 || (n && (!d || !h))) return PSFUN_ERR_INVALID_ARG;
  if (n) PSFUN_HIP_TRY(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSFUN_SUCCESS;
}

extern "C" int psfun_vec_download(psfun_ctx_t c, double *h, const double *d, int64_t n) {
  if (!c || n < 0 || (n && (!d || !h))) return PSFUN_ERR_INVALID_ARG;
  if (n) PSFUN_HIP_TRY(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSFUN_SUCCESS;
}

// ---------------------------------------------------------------------------
// Per-kernel timing
// ---------------------------------------------------------------------------
int psfun_prof_begin(psfun_ctx *c, int kid, double bytes, int tag, int reps) {
  if (!c->prof_on) return PSFUN_SUCCESS;
  if (c->prof_used + 1 >= PSFUN_PROF_MAXPAIRS) {  // pool full: fold finished pairs first
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    PSFUN_TRY(psfun_prof_collect(c));
  }
  const size_t need = 2 * (size_t)(c->prof_used + 1);
  while (c->prof_ev.size() < need) {
    hipEvent_t e;
    PSFUN_HIP_TRY(hipEventCreate(&e));
    c->prof_ev.push_back(e);
  }
  if (c->prof_kid.size() < (size_t)c->prof_used + 1) {
    c->prof_kid.resize(c->prof_used + 1);
    c->prof_tag.resize(c->prof_used + 1);
    c->prof_bytes_pending.resize(c->prof_used + 1);
    c->prof_reps.resize(c->prof_used + 1);
  }
  c->prof_kid[c->prof_used] = kid;
  c->prof_tag[c->prof_used] = tag;
  c->prof_bytes_pending[c->prof_used] = bytes * reps;
  c->prof_reps[c->prof_used] = reps;
  PSFUN_HIP_TRY(hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream));
  return PSFUN_SUCCESS;
}

void psfun_prof_discard_after(psfun_ctx *c, int kid, int tag_max) {
  for (int i = 0; i < c->prof_used; ++i)
    if (c->prof_kid[i] == kid && c->prof_tag[i] > tag_max) c->prof_kid[i] = -1;
}

int psfun_prof_end(psfun_ctx *c) {
  if (!c->prof_on) return PSFUN_SUCCESS;
  PSFUN_HIP_TRY(hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream));
  c->prof_used++;
  return PSFUN_SUCCESS;
}

int psfun_prof_collect(psfun_ctx *c) {
  for (int i = 0; i < c->prof_used; ++i) {
    float ms = 0.f;
    PSFUN_HIP_TRY(hipEventSynchronize(c->prof_ev[2 * i + 1]));
    PSFUN_HIP_TRY(hipEventElapsedTime(&ms, c->prof_ev[2 * i], c->prof_ev[2 * i + 1]));
    const int k = c->prof_kid[i];
    if (k >= 0 && k < PSFUN_PROF_NK) {
      c->prof_ms[k] += ms;
      c->prof_bytes[k] += c->prof_bytes_pending[i];
      c->prof_count[k] += c->prof_reps[i];
    }
  }
  c->prof_used = 0;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_prof_enable(psfun_ctx_t c, int on) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  if (!on && c->prof_used) {
    PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
    PSFUN_TRY(psfun_prof_collect(c));
  }
  c->prof_on = on <= 0 ? 0 : (on > 16 ? 16 : on);
  return PSFUN_SUCCESS;
}

extern "C" int psfun_prof_reset(psfun_ctx_t c) {
  if (!c) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  PSFUN_TRY(psfun_prof_collect(c));
  for (int k = 0; k < PSFUN_PROF_NK; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_bytes[k] = 0.0;
    c->prof_count[k] = 0;
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_prof_read(psfun_ctx_t c, int kid, double *ms, int64_t *count, double *bytes) {
  if (!c || kid < 0 || kid >= PSFUN_PROF_NK) return PSFUN_ERR_INVALID_ARG;
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  PSFUN_TRY(psfun_prof_collect(c));
  if (ms) *ms = c->prof_ms[kid];
  if (count) *count = c->prof_count[kid];
  if (bytes) *bytes = c->prof_bytes[kid];
  return PSFUN_SUCCESS;
}
