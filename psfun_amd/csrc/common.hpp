// Shared definitions for the gfx950 f(A)x library: context, workspace,
// status handling, wave/block reductions and the XCD-aware work split.
//
// Grid-wide reductions never use atomics or fences: a kernel leaves one
// partial per block, and every consumer (the next kernel in stream order)
// folds the partials itself in a fixed order, so results are deterministic
// and cross-XCD visibility comes from the kernel boundary alone.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <math.h>
#include <float.h>
#include <mutex>
#include <vector>

#include "../../include/psfun_hip.h"

#define PSFUN_BS 256           // threads per block for the streaming kernels
#define PSFUN_PADE_MAX 30       // highest degree of the GENPADE approximant
#define PSFUN_MAX_GRID 4096    // upper bound on reduction grids (partials slots)
#define PSFUN_PINNED_BYTES (1 << 20)
#define PSFUN_PROF_NK 16       // kernel ids for psfun_prof_*
#define PSFUN_PROF_MAXPAIRS 8192

#define PSFUN_HIP_TRY(call)                               \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) {                               \
      psfun_record_hip_error(e_, #call, __FILE__, __LINE__); \
      return PSFUN_ERR_HIP;                               \
    }                                                     \
  } while (0)

#define PSFUN_TRY(call)           \
  do {                            \
    int s_ = (call);              \
    if (s_ != PSFUN_SUCCESS) return s_; \
  } while (0)

void psfun_record_hip_error(hipError_t e, const char *what, const char *file, int line);

struct psfun_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  double *d_partials = nullptr;   // 8 slots of PSFUN_MAX_GRID doubles
  double *d_scalar = nullptr;     // 256 doubles of scalar results / small parameter tables
  void *h_pinned = nullptr;       // pinned host staging
  // Host-visible loop flags in fine-grained (coherent) pinned memory that the
  // control blocks of the iterative kernels store to directly: [0] progress
  // (last loop index started), [1] stop state (0 = running).  The host paces
  // its launches on them with no copies or events.
  volatile int *h_flags = nullptr;
  int *d_flags = nullptr;         // device address of h_flags
  hipEvent_t ev[2] = {nullptr, nullptr};
  // auxiliary stream (default priority) for work that overlaps the main
  // stream's inside one call, forked and joined by these events
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // grow-only workspaces: ws for each solver's own state; ws2 for an outer
  // method whose state must survive inner solver calls (shift-and-invert Arnoldi)
  void *ws = nullptr;
  size_t ws_bytes = 0;
  void *ws2 = nullptr;
  size_t ws2_bytes = 0;
  // degree of the Pade approximant of the GENPADE variant (psfun_set_pade_degree)
  int pade_degree = 6;
  // user function of the USERF f(H) step (psfun_set_userf), called on the host
  psfun_userf_t userf = nullptr;
  void *userf_data = nullptr;
  // optional per-launch timing of the hot kernels (psfun_prof_*)
  int prof_on = 0;                       // 0 off; R >= 1: timing on, idempotent steps run R times
  std::vector<hipEvent_t> prof_ev;       // event pool, 2 per timed launch
  std::vector<int> prof_kid;             // kernel id of each pending pair
  std::vector<int> prof_tag;             // caller tag (e.g. Krylov index) of each pending pair
  std::vector<double> prof_bytes_pending;
  std::vector<int> prof_reps;            // launches inside each pending pair
  int prof_used = 0;
  double prof_ms[PSFUN_PROF_NK] = {};
  double prof_bytes[PSFUN_PROF_NK] = {};
  int64_t prof_count[PSFUN_PROF_NK] = {};
};

// Timing hooks: bracket one launch of kernel `kid` (with `bytes` algorithmic
// bytes) by two events when profiling is on.  psfun_prof_collect() folds the
// finished pairs into the per-kernel totals (call after a synchronize).
int psfun_prof_begin(psfun_ctx *c, int kid, double bytes, int tag = 0, int reps = 1);
// Launches per timing pair for a kernel that may run repeatedly with the same
// result (1 unless profiling asked for repeats).
static inline int psfun_prof_reps(const psfun_ctx *c) { return c->prof_on > 1 ? c->prof_on : 1; }
// Drop pending pairs of kernel `kid` whose tag exceeds `tag_max` (launches
// that found their loop already finished and did no work).
void psfun_prof_discard_after(psfun_ctx *c, int kid, int tag_max);
int psfun_prof_end(psfun_ctx *c);
int psfun_prof_collect(psfun_ctx *c);

struct psfun_csr {
  psfun_ctx *ctx = nullptr;
  int device = 0;  // ctx->device, kept so destroy never reads the context
  int n = 0;
  int64_t nnz = 0;
  int *rowptr = nullptr;
  int *colind = nullptr;
  double *val = nullptr;
  bool owns = false;
  int lanes = 8;   // lanes per row for the vector-CSR kernels
  int max_row = 0;
  // Sliced ELL copy (slices of PSFUN_SELL_C consecutive rows, no row
  // reordering): slice s holds its width w_s = (sell_ptr[s+1]-sell_ptr[s])/C
  // entries per row, stored entry-major (entry k of row 64s+l at
  // sell_ptr[s] + k*C + l), each row's entries in CSR order, padding with
  // value 0 and the row's own column index.  One thread per row then reads
  // whole cache lines of colind/val with no intra-row reduction.
  int nslice = 0;
  int64_t sell_nnz = 0;  // padded entry count (0: not built)
  int *sell_ptr = nullptr;
  int *sell_col = nullptr;
  double *sell_val = nullptr;
  // optional 16-bit column offsets of the sliced-ELL entries (col - row),
  // built when every offset fits: half the column bytes for the kernels
  // that read them (nullptr otherwise)
  int16_t *sell_dcol = nullptr;
  // Optional diagonal slices (PSBLAS's "hacked DIA"), built when the format is
  // first chosen: slice s of PSFUN_SELL_C rows keeps the distinct offsets
  // col - row of its entries, ascending, in hdia_off[hdia_ptr[s] ..
  // hdia_ptr[s+1]), and for each one a value per row (0 where the row has no
  // such entry) at hdia_val[(hdia_ptr[s] + k) * C + l].  The SpMV reads no
  // column indices and x[row + offset] contiguously.
  int64_t hdia_ndiag = 0;  // slice diagonals stored
  int *hdia_ptr = nullptr;  // nullptr: not built
  int *hdia_off = nullptr;
  double *hdia_val = nullptr;
  // Optional CSR row blocks (built when PSFUN_FORMAT_CSRS is first chosen),
  // four ints per block {a, b, c, kind}:
  //   kind 0: rows [a, b), at most PSFUN_BS rows and PSFUN_CSRS_CAP entries;
  //           the Lanczos step streams their entries with coalesced loads
  //           into LDS and sums each row there (CSR-stream, Greathouse and
  //           Daga, SC14)
  //   kind -w: the same, with all of the block's columns in [c, c + w),
  //           w <= PSFUN_CSRS_XT: that window of x is staged in LDS too
  //   kind 1: row a alone (more than PSFUN_CSRS_CAP entries), summed by the
  //           whole block
  //   kind 2 + s: entries [b, c) of row a, one segment of split row s (rows
  //           longer than PSFUN_CSRS_SPLIT, at most PSFUN_CSRS_MAXSPLIT of
  //           them); the block that completes a row's segments (an arrival
  //           counter) adds them in segment order
  int *csrs_blk = nullptr;  // nullptr: not built
  int csrs_nblk = 0;
  int csrs_nsplit = 0;
  int csrs_nxt = 0;  // blocks of short rows with a staged x window
  int *csrs_split = nullptr;   // [2 * nsplit]: first block, segment count
  int *csrs_cnt = nullptr;     // [nsplit] arrival counters (0 between launches)
  double *csrs_seg = nullptr;  // [nblk] segment sums
  int format = PSFUN_FORMAT_CSR;  // format the kernels use
};
#define PSFUN_SELL_C 64
#ifndef PSFUN_CSRS_CAP
#define PSFUN_CSRS_CAP 2048      // entries per block of short rows
#endif
#ifndef PSFUN_CSRS_PER_CU
#define PSFUN_CSRS_PER_CU 12     // CSR-stream blocks per CU in the Lanczos grid (12 measured best of 4..16)
#endif
#define PSFUN_CSRS_XT 1024      // widest x window a block of short rows stages in LDS
#define PSFUN_CSRS_SPLIT 8192
#define PSFUN_CSRS_SEG 4096
#define PSFUN_CSRS_MAXSPLIT 1024
// lanes-per-row template value that selects the diagonal-slice row sum
#define PSFUN_L_HDIA (-1)

// (Re)build the sliced-ELL copy from the device CSR arrays; `rp_host` is the
// 0-based row pointer on the host.  Leaves sell_nnz = 0 when the padded size
// does not fit int32 offsets.
int psfun_sell_build(psfun_csr *A, const int *rp_host);
// Refresh the sliced-ELL values from A->val (same pattern).
int psfun_sell_refill(psfun_csr *A);

// Host side of the flag pacing: spin until h_flags[0] >= target, a stop is
// flagged, or the stream has drained (so the wait always ends).
int psfun_host_wait(psfun_ctx *c, int target);

// Device workspace: returns a pointer to at least `bytes` bytes (256-byte
// aligned).  Grows by reallocation; previous contents are not preserved.
int psfun_workspace(psfun_ctx *ctx, size_t bytes, void **ptr);
int psfun_workspace2(psfun_ctx *ctx, size_t bytes, void **ptr);

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Number of blocks for a grid-stride kernel over `chunks` work items.
// Always a multiple of 8 (one share per XCD) and at most PSFUN_MAX_GRID.
static inline int psfun_grid(const psfun_ctx *ctx, int64_t chunks, int per_cu = 16) {
  int64_t g = (int64_t)ctx->num_cu * per_cu;
  if (g > PSFUN_MAX_GRID) g = PSFUN_MAX_GRID;
  if (chunks < g) g = chunks;
  g = (g + 7) / 8 * 8;
  if (g < 8) g = 8;
  return (int)g;
}

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// LDS written by some lanes of a wavefront, then read by others: order the
// accesses within the wavefront (no block barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Krylov assembly y = rs1 * sum_{l<m} c[l] * (V_l * s[l]) over basis slots
// V_l = V + l*ldv (SCALED; otherwise the factor s[l] is left out), with c and
// s in LDS.  Each row's terms are folded in slot order, exactly as a one-row
// loop would.  The slots (16-byte aligned: ldv is even and V is) are read as
// pairs of rows with nontemporal loads, since nothing reads them again;
// y, which may have any 8-byte alignment, is written one double at a time.
// Measured for 24 slots of 2M rows: 64 us against 87 us for one row per
// thread.  Grid-stride over row pairs; the odd last row goes to thread 0 of
// the last block.
typedef double psfun_d2 __attribute__((ext_vector_type(2)));
template <bool SCALED>
__device__ __forceinline__ void basis_combine(int n, const double *__restrict__ V, int64_t ldv, int m,
                                              const double *c, const double *s, double rs1,
                                              double *__restrict__ y) {
  const int64_t n2 = n >> 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    psfun_d2 acc = {0.0, 0.0};
    if (m > 0) {
      const psfun_d2 v0 = __builtin_nontemporal_load((const psfun_d2 *)V + i);
      acc.x = SCALED ? c[0] * (v0.x * s[0]) : c[0] * v0.x;
      acc.y = SCALED ? c[0] * (v0.y * s[0]) : c[0] * v0.y;
#pragma unroll 8
      for (int l = 1; l < m; ++l) {
        const psfun_d2 v = __builtin_nontemporal_load((const psfun_d2 *)(V + (int64_t)l * ldv) + i);
        acc.x = fma(c[l], SCALED ? v.x * s[l] : v.x, acc.x);
        acc.y = fma(c[l], SCALED ? v.y * s[l] : v.y, acc.y);
      }
    }
    __builtin_nontemporal_store(rs1 * acc.x, y + 2 * i);
    __builtin_nontemporal_store(rs1 * acc.y, y + 2 * i + 1);
  }
  if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    const int64_t i = n - 1;
    double acc = 0.0;
    if (m > 0) {
      acc = SCALED ? c[0] * (V[i] * s[0]) : c[0] * V[i];
      for (int l = 1; l < m; ++l) acc = fma(c[l], SCALED ? V[(int64_t)l * ldv + i] * s[l] : V[(int64_t)l * ldv + i], acc);
    }
    y[i] = rs1 * acc;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// Max over the wavefront by DPP row moves (row_shr 1/2/4/8 within each
// 16-lane row, then row_bcast 15 and 31 across rows), the result read from
// lane 63: register-to-register steps instead of the cross-lane permutes of
// __shfl_xor.  Lanes whose DPP source lies outside the row keep their own
// value, which a max ignores.
template <int CTRL, int RM, int BM>
__device__ __forceinline__ double dpp_mov_f64(double old, double src) {
  const long long o = __double_as_longlong(old), s = __double_as_longlong(src);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, RM, BM, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, RM, BM, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_max_dpp(double v) {
  v = fmax(v, dpp_mov_f64<0x111, 0xf, 0xf>(v, v));  // row_shr:1
  v = fmax(v, dpp_mov_f64<0x112, 0xf, 0xf>(v, v));  // row_shr:2
  v = fmax(v, dpp_mov_f64<0x114, 0xf, 0xf>(v, v));  // row_shr:4
  v = fmax(v, dpp_mov_f64<0x118, 0xf, 0xf>(v, v));  // row_shr:8
  v = fmax(v, dpp_mov_f64<0x142, 0xa, 0xf>(v, v));  // row_bcast:15
  v = fmax(v, dpp_mov_f64<0x143, 0xc, 0xf>(v, v));  // row_bcast:31
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v of lane - 1 (0 in lane 0) and v of lane + 1 (0 in lane 63), by DPP
// wavefront shifts: register moves, no LDS permute.  Every lane of the
// wavefront must be active.
__device__ __forceinline__ double wave_from_prev(double v) { return dpp_mov_f64<0x138, 0xf, 0xf>(0.0, v); }  // wave_shr:1
__device__ __forceinline__ double wave_from_next(double v) { return dpp_mov_f64<0x130, 0xf, 0xf>(0.0, v); }  // wave_shl:1

// Sum over the wavefront by the same DPP row moves (lanes with no source
// add 0), read from lane 63: the result is wavefront-uniform.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_mov_f64<0x111, 0xf, 0xf>(0.0, v);  // row_shr:1
  v += dpp_mov_f64<0x112, 0xf, 0xf>(0.0, v);  // row_shr:2
  v += dpp_mov_f64<0x114, 0xf, 0xf>(0.0, v);  // row_shr:4
  v += dpp_mov_f64<0x118, 0xf, 0xf>(0.0, v);  // row_shr:8
  v += dpp_mov_f64<0x142, 0xa, 0xf>(0.0, v);  // row_bcast:15
  v += dpp_mov_f64<0x143, 0xc, 0xf>(0.0, v);  // row_bcast:31
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Value of v in lane `src` (wavefront-uniform index), to every lane.
__device__ __forceinline__ double wave_read_f64(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src), hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum over a block of BS threads; the result is valid in every thread.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *lds) {
  constexpr int NW = BS / 64;
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += lds[i];
  return t;
}
template <int BS>
__device__ __forceinline__ double block_max(double v, double *lds) {
  constexpr int NW = BS / 64;
  v = wave_max(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) t = fmax(t, lds[i]);
  return t;
}

// XCD-aware split of `nchunks` chunks over the grid: the 8 XCDs receive
// contiguous eighths of the chunk range (blocks are dispatched round-robin
// over XCDs, block b -> XCD b % 8), so neighbouring rows share an L2.
struct XcdRange {
  int64_t begin, end, stride;
};
// Sum of G partials in a fixed order (thread t takes t, t+BS, ..., then a
// fixed tree): every block that calls it gets the bit-identical total.
template <int BS>
__device__ __forceinline__ double block_sum_partials(const double *__restrict__ P, int G, double *lds) {
  double acc = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < G; b += BS) acc += P[b];
  return block_sum<BS>(acc, lds);
}

// Streaming blocks 0..G-1 of a grid that may carry extra control blocks.
__device__ __forceinline__ XcdRange xcd_range_g(int64_t nchunks, int G) {
  const int xcd = blockIdx.x & 7;
  const int lb = blockIdx.x >> 3;
  const int gp = G >> 3;
  XcdRange r;
  r.begin = nchunks * xcd / 8 + lb;
  r.end = nchunks * (xcd + 1) / 8;
  r.stride = gp;
  return r;
}

__device__ __forceinline__ XcdRange xcd_range(int64_t nchunks) {
  const int G = gridDim.x;          // multiple of 8
  const int xcd = blockIdx.x & 7;
  const int lb = blockIdx.x >> 3;
  const int gp = G >> 3;
  int64_t lo = nchunks * xcd / 8, hi = nchunks * (xcd + 1) / 8;
  XcdRange r;
  r.begin = lo + lb;
  r.end = hi;
  r.stride = gp;
  return r;
}

// Grid of 1 + G blocks whose block 0 is a loop-control block (dispatched
// first, so it starts at once however many streaming blocks are resident)
// and blocks 1..G stream (G a multiple of 8).  Block k runs on XCD k % 8; the
// blocks of one XCD get that XCD's contiguous eighth of the chunks.
__device__ __forceinline__ XcdRange xcd_range_c0(int64_t nchunks, int G) {
  const int k = blockIdx.x;  // 1..G
  const int xcd = k & 7;
  XcdRange r;
  r.begin = nchunks * xcd / 8 + ((k - 1) >> 3);
  r.end = nchunks * (xcd + 1) / 8;
  r.stride = G >> 3;
  return r;
}

// Operand of a padded-format product v * x: the padding slots of sliced ELL
// and diagonal slices hold v = 0 and point at some in-range x entry, which
// CSR would never read; taking x as 0 there keeps a non-finite x entry from
// turning unrelated rows into NaN (0 * Inf), so these formats agree with CSR
// on such input.  For finite x the product and the sum are bitwise unchanged.
// (An explicitly stored 0 of the matrix is treated like padding.)
__device__ __forceinline__ double pad_x(double v, double x) { return v != 0.0 ? x : 0.0; }

// Sum over the sliced-ELL entries of `row`, gathering x[col] * xs.
__device__ __forceinline__ double sell_row_dot(int row, const int *__restrict__ sp, const int *__restrict__ col,
                                               const double *__restrict__ val, const double *__restrict__ x,
                                               double xs) {
  const int s = row >> 6, l = row & 63;
  const int b = sp[s];
  const int w = (sp[s + 1] - b) >> 6;
  const int *cp = col + b + l;
  const double *vp = val + b + l;
  double sum = 0.0;
#pragma unroll 4
  for (int k = 0; k < w; ++k) {
    const double v = vp[k * 64];
    sum = fma(v, pad_x(v, x[cp[k * 64]] * xs), sum);
  }
  return sum;
}

// sell_row_dot with 16-bit column offsets: entry k of the row reads
// x[row + dcol_k].
__device__ __forceinline__ double sell_row_dot16(int row, const int *__restrict__ sp,
                                                 const int16_t *__restrict__ dcol, const double *__restrict__ val,
                                                 const double *__restrict__ x, double xs) {
  const int s = row >> 6, l = row & 63;
  const int b = sp[s];
  const int w = (sp[s + 1] - b) >> 6;
  const int16_t *cp = dcol + b + l;
  const double *vp = val + b + l;
  double sum = 0.0;
#pragma unroll 4
  for (int k = 0; k < w; ++k) {
    const double v = vp[k * 64];
    sum = fma(v, pad_x(v, x[row + (int)cp[k * 64]] * xs), sum);
  }
  return sum;
}

// Row sum of `row` in the diagonal slices: one term per diagonal of the row's
// slice, in ascending offset order (the row's CSR order for sorted columns;
// the slice's other diagonals add 0).  Every lane of the calling wavefront
// must hold a row of the same slice (64 consecutive rows from a multiple of
// 64), so the slice's offsets are scalar loads.  Out-of-range columns of the
// zero padding read a clamped, in-range x.
// With SCALED, each x entry is multiplied by xs before its product (as
// sell_row_dot does), for callers that form a scaled vector on the fly.
template <bool SCALED = false>
__device__ __forceinline__ double hdia_row_dot(int row, int n, const int *__restrict__ hp,
                                               const int *__restrict__ off, const double *__restrict__ val,
                                               const double *__restrict__ x, double xs = 1.0) {
  const int s = __builtin_amdgcn_readfirstlane(row >> 6), l = row & 63;
  const int b = hp[s], e = hp[s + 1];
  double sum = 0.0;
#pragma unroll 4
  for (int k = b; k < e; ++k) {
    const int c = min(max(row + off[k], 0), n - 1);
    const double v = val[(int64_t)k * PSFUN_SELL_C + l];
    sum = fma(v, pad_x(v, SCALED ? x[c] * xs : x[c]), sum);
  }
  return sum;
}

// Store to a host-visible flag word (coherent pinned memory): a relaxed
// system-scope atomic store, so the value leaves the GPU caches at once and
// the host sees it while the kernel still runs.
__device__ __forceinline__ void host_flag_store(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Row sum  sum_k val_k * g(col_k)  of `row` for the storage chosen at compile
// time: SELL (sliced ELL, one thread per row; rp = sell_ptr, entries in CSR
// order), diagonal slices (L = PSFUN_L_HDIA with SELL: rp/col/val = hdia_ptr,
// hdia_off, hdia_val; one thread per row, the wavefront's rows in one slice;
// padding columns off the matrix are clamped and meet a stored 0) or vector
// CSR with L lanes per row (every lane of the group calls it; the sum is
// valid in every lane).  g(c) yields the gathered operand entry.
template <int L, bool SELL, typename Gather>
__device__ __forceinline__ double row_dot(int row, int n, const int *__restrict__ rp, const int *__restrict__ col,
                                          const double *__restrict__ val, int lane, Gather g) {
  double sum = 0.0;
  if (L == PSFUN_L_HDIA) {
    if (row < n) {
      const int s = __builtin_amdgcn_readfirstlane(row >> 6), l = row & 63;
      const int b = rp[s], e = rp[s + 1];
#pragma unroll 4
      for (int k = b; k < e; ++k) {
        const double v = val[(int64_t)k * PSFUN_SELL_C + l];
        sum = fma(v, pad_x(v, g(min(max(row + col[k], 0), n - 1))), sum);
      }
    }
  } else if (SELL) {
    if (row < n) {
      const int s = row >> 6, l = row & 63;
      const int b = rp[s];
      const int w = (rp[s + 1] - b) >> 6;
      const int *cp = col + b + l;
      const double *vp = val + b + l;
#pragma unroll 4
      for (int k = 0; k < w; ++k) {
        const double v = vp[k * 64];
        sum = fma(v, pad_x(v, g(cp[k * 64])), sum);
      }
    }
  } else {
    if (row < n) {
      const int e = rp[row + 1];
      for (int q = rp[row] + lane; q < e; q += L) sum = fma(val[q], g(col[q]), sum);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
  }
  return sum;
}

// Kernels without a diagonal-slice variant read the sliced-ELL copy (always
// built beside it when it fits) when the diagonal-slice format is chosen.
static inline bool psfun_reads_sell(const psfun_csr *A) {
  return A->format == PSFUN_FORMAT_SELL || (A->format == PSFUN_FORMAT_HDIA && A->sell_nnz > 0);
}
static inline bool psfun_reads_hdia(const psfun_csr *A) {
  return A->format == PSFUN_FORMAT_HDIA && A->hdia_ptr != nullptr;
}
// Sliced ELL and diagonal slices: one thread per row, chunks of PSFUN_BS rows.
static inline bool psfun_row_per_thread(const psfun_csr *A) { return psfun_reads_sell(A) || psfun_reads_hdia(A); }

// Launch-side dispatch over the matrix storage: defines LL (lanes) and SS
// (sliced ELL) as constants and the matrix arrays MRP/MCOL/MVAL for __VA_ARGS__.
#define PSFUN_DISPATCH_STORAGE(A, ...)                                                           \
  do {                                                                                           \
    if (psfun_reads_hdia(A)) {                                                                   \
      constexpr int LL = PSFUN_L_HDIA;                                                           \
      constexpr bool SS = true;                                                                  \
      const int *MRP = (A)->hdia_ptr, *MCOL = (A)->hdia_off;                                     \
      const double *MVAL = (A)->hdia_val;                                                        \
      __VA_ARGS__;                                                                               \
    } else if (psfun_reads_sell(A)) {                                                            \
      constexpr int LL = 1;                                                                      \
      constexpr bool SS = true;                                                                  \
      const int *MRP = (A)->sell_ptr, *MCOL = (A)->sell_col;                                     \
      const double *MVAL = (A)->sell_val;                                                        \
      __VA_ARGS__;                                                                               \
    } else {                                                                                     \
      const int *MRP = (A)->rowptr, *MCOL = (A)->colind;                                         \
      const double *MVAL = (A)->val;                                                             \
      constexpr bool SS = false;                                                                 \
      switch ((A)->lanes) {                                                                      \
        case 1: { constexpr int LL = 1; __VA_ARGS__; } break;                                    \
        case 2: { constexpr int LL = 2; __VA_ARGS__; } break;                                    \
        case 4: { constexpr int LL = 4; __VA_ARGS__; } break;                                    \
        case 8: { constexpr int LL = 8; __VA_ARGS__; } break;                                    \
        case 16: { constexpr int LL = 16; __VA_ARGS__; } break;                                  \
        case 32: { constexpr int LL = 32; __VA_ARGS__; } break;                                  \
        default: { constexpr int LL = 64; __VA_ARGS__; } break;                                  \
      }                                                                                          \
    }                                                                                            \
  } while (0)

// Rows per chunk of a streaming SpMV kernel for the storage.
static inline int64_t psfun_spmv_chunks(const psfun_csr *A) {
  const int rpb = psfun_row_per_thread(A) ? PSFUN_BS : PSFUN_BS / A->lanes;
  return ((int64_t)A->n + rpb - 1) / rpb;
}
