// Batched Lanczos f(A)x: many independent square CSR matrices of uneven
// order ("graphs", concatenated block-diagonally) times a block of K vectors
// each, i.e. ngraph*K independent runs of the reference's Lanczos
// (krylov/polynomial/psfun_d_lanczos.f90:183-377), advanced together.
//
// Layout: vectors are row-major (total_rows x K) so that the K lanes that
// own one row gather K contiguous doubles per nonzero.  Rows are cut into
// chunks that never cross a graph; one block owns one chunk.  Every scalar of
// instance (g, k) is folded from the per-chunk partials of graph g in a fixed
// order by every block that needs it (deterministic, no atomics on data).
//
// Per Krylov index j (one launch sequence for all instances):
//   k_bt_spmm(j):   u = opscale*A v_j with v_j = (1/sigma) src formed on the
//                   fly (src = basis slot j-1; slot 0 is x), alpha partials
//   k_bt_fold(j):   one wavefront per instance folds its alpha partials
//   k_bt_update(j): w = u - alpha v_j - beta v_{j-1} into basis slot j (the
//                   basis keeps the unnormalised w: v_{l+1} = w_l / beta_l is
//                   formed where a slot is read); ||w||^2 partials
//   k_bt_beta/k_bt_est(j+1): one wavefront per instance: beta = ||w||, breakdown test,
//                   a-posteriori estimate from T_j, stop decision.
// Instances stop independently; a stopped instance's rows are skipped.  The
// whole sequence (itmax steps + assembly) can be captured once in a hipGraph
// and replayed: kernels of finished steps return at once.
#include "common.hpp"
#include "kernels.hpp"
#include "fexp.hpp"

#include <algorithm>
#include <map>
#include <type_traits>
#include <tuple>

#define BT_KMAX 64

enum { BT_RUN = 0, BT_CONV = 1, BT_BRK = 2, BT_MAXIT = 3, BT_ZERO = 4 };

struct BtCtl {
  int remaining;  // instances still running
  int pad[63];
};

struct psfun_batch {
  psfun_ctx *ctx = nullptr;
  int device = 0;  // ctx->device, kept so destroy never reads the context
  int ngraph = 0;
  int64_t nrows = 0, nnz = 0;
  int *rowptr = nullptr, *colind = nullptr;
  double *val = nullptr;
  std::vector<int64_t> row_off;
  int rows_per_chunk = 512;
  int nchunks = 0;
  int *d_chunk = nullptr;    // [3*nchunks]: row0, row1, graph
  int *d_gchunk = nullptr;   // [2*ngraph]: first chunk, end chunk
  double *d_part = nullptr;  // partials: 2 * nchunks * BT_KMAX
  size_t part_cap = 0;
  int use_graph = 0;
  // row order (PSFUN_ORDER_*): with RCM the device CSR is in the reordered
  // numbering, d_perm[r'] = original global row of row r', d_nzperm[q'] =
  // original position of stored entry q'
  int order = PSFUN_ORDER_NATURAL;
  int *d_perm = nullptr, *d_nzperm = nullptr;
  // per row: its place in its 64-row window (windows start at each graph's
  // first row) when the window's rows are taken longest first
  unsigned char *d_sched = nullptr;
  std::map<std::tuple<const void *, void *, const void *, int, int, int, double, double, double>, hipGraphExec_t>
      graphs;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

struct BtInst {
  double *alpha, *beta, *rs, *yk;  // [ninst][ld]
  // [ninst][ld]: 1/beta_l, the factor of basis slot l (1/rs1 for slot 0),
  // written by k_bt_final for the slots the assembly reads
  double *ib;
  double *rs1, *errnum, *errden;   // [ninst]
  int *done, *itx, *m;             // [ninst]
  int ld;
};

// ---------------------------------------------------------------------------
// Per-instance fold of per-chunk partials, one wavefront per instance (the
// 64 lanes stride the graph's chunks, then a fixed-order wave sum), so the
// streaming kernels read one scalar instead of folding a large graph's
// partials in every block.  MODE 0: alpha_j of running instances;
// MODE 1: ||x|| -> rs(1) (a zero start vector stops the instance).
template <int KP, int MODE>
__global__ __launch_bounds__(64) void k_bt_fold(const int *__restrict__ gchunk, int K, int j,
                                                const double *__restrict__ p_in, BtInst in, BtCtl *bctl) {
  const int I = blockIdx.x;
  if (MODE == 0 && in.done[I] != BT_RUN) return;
  const int g = I / K, k = I % K;
  const int c0 = gchunk[2 * g], c1 = gchunk[2 * g + 1];
  double acc = 0.0;
  for (int cc = c0 + threadIdx.x; cc < c1; cc += 64) acc += p_in[(int64_t)cc * KP + k];
  acc = wave_sum(acc);
  if (threadIdx.x != 0) return;
  const int64_t base = (int64_t)I * in.ld;
  if (MODE == 0) {
    in.alpha[base + j] = acc;
  } else {
    const double sg = sqrt(acc);
    in.rs1[I] = sg;
    in.rs[base + 1] = sg;
    if (sg == 0.0) {  // zero start vector: y = 0
      in.done[I] = BT_ZERO;
      atomicSub(&bctl->remaining, 1);
    }
  }
}

template <int KP>
__global__ __launch_bounds__(PSFUN_BS) void k_bt_norm(const int *__restrict__ chunk, int K,
                                                      const double *__restrict__ X, double *__restrict__ P) {
  constexpr int RP = PSFUN_BS / KP;
  __shared__ double red[PSFUN_BS];
  const int c = blockIdx.x;
  const int r0 = chunk[3 * c], r1 = chunk[3 * c + 1];
  const int k = threadIdx.x % KP, sub = threadIdx.x / KP;
  double acc = 0.0;
  if (k < K)
    for (int r = r0 + sub; r < r1; r += RP) {
      const double v = X[(int64_t)r * K + k];
      acc = fma(v, v, acc);
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < KP) {
    double s = 0.0;
    for (int q = 0; q < RP; ++q) s += red[q * KP + threadIdx.x];
    P[(int64_t)c * KP + threadIdx.x] = s;
  }
}

#ifndef BT_XCD
#define BT_XCD 1
#endif
// Chunk of this SpMM block.  Blocks are dispatched round-robin over the 8
// XCDs; block b runs on XCD b % 8 and takes the (b / 8)-th chunk of that
// XCD's contiguous eighth, so the blocks an XCD runs at a time cover
// consecutive rows and their neighbour gathers (rows close in the graph's
// ordering) hit that XCD's L2.  The grid is 8 * ceil(nc / 8); -1 for the
// padding blocks.
__device__ __forceinline__ int bt_block_chunk(int nc) {
  if (!BT_XCD) return blockIdx.x < (unsigned)nc ? (int)blockIdx.x : -1;
  const int per = (nc + 7) >> 3;
  const int c = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  return c < nc ? c : -1;
}
static inline int bt_spmm_grid(int nc) { return BT_XCD ? 8 * ((nc + 7) / 8) : nc; }

#define BT_UB 4     // rows per thread in flight in the 3-term update
#ifndef BT_AR
#define BT_AR 8     // rows per thread in flight in the assembly
#endif
#ifndef BT_EB
#define BT_EB 4     // entries per row whose gathers the staged SpMM issues together
#endif
#define BT_SR 64    // rows per LDS-staged sub-chunk of the SpMM
#define BT_CAP 768  // entries staged per sub-chunk (9 KiB with the row pointers)
// At least 7 waves per SIMD: the kernel waits on gather latency, and the
// bound holds it at 72 VGPRs (73 unbounded, 6 waves) with no spills; 8 waves
// would spill.  Measured on the 256-graph batch: +2% f(A)x/s.
template <int KP>
__global__ __launch_bounds__(PSFUN_BS, 7) void k_bt_spmm(const int *__restrict__ rowptr, const int *__restrict__ col,
                                                      const double *__restrict__ val, double opscale,
                                                      const int *__restrict__ chunk, const int *__restrict__ gchunk,
                                                      int K, int j, const double *__restrict__ src,
                                                      double *__restrict__ v_out, double *__restrict__ u_out,
                                                      const double *__restrict__ p_in, double *__restrict__ p_out,
                                                      BtInst in, BtCtl *bctl, int nc) {
  constexpr int RP = PSFUN_BS / KP;
  __shared__ double s_inv[KP];
  __shared__ int s_act[KP];
  __shared__ int s_any;
  __shared__ double red[PSFUN_BS];
  __shared__ int s_rp[BT_SR + 1];
  __shared__ int s_col[BT_CAP];
  __shared__ double s_val[BT_CAP];
  if (bctl->remaining == 0) return;
  const int c = bt_block_chunk(nc);
  if (c < 0) return;
  const int r0 = chunk[3 * c], r1 = chunk[3 * c + 1], g = chunk[3 * c + 2];
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  if (threadIdx.x < KP) {
    const int k = threadIdx.x;
    int act = 0;
    double inv = 0.0;
    if (k < K) {
      const int I = g * K + k;
      if (j == 1) {
        const double sg = in.rs1[I];  // folded by k_bt_fold<KP, 1>
        act = (sg != 0.0);
        inv = act ? 1.0 / sg : 0.0;
      } else {
        act = (in.done[I] == BT_RUN);
        if (act) inv = 1.0 / in.beta[(int64_t)I * in.ld + j - 1];
      }
    }
    s_inv[k] = inv;
    s_act[k] = act;
    if (act) s_any = 1;
  }
  __syncthreads();
  const int k = threadIdx.x % KP, sub = threadIdx.x / KP;
  const bool active = s_any && k < K && s_act[k];
  const double inv = active ? s_inv[k] : 0.0;
  double dot = 0.0;
  // The chunk is taken BT_SR rows at a time: their row pointers, and their
  // column indices and values when they fit in BT_CAP entries, are staged in
  // LDS by coalesced loads of the whole block, so a row's gathers issue after
  // one LDS read instead of after two dependent global loads.  A sub-chunk
  // with more entries reads them from global memory.  Each row's terms are
  // added in entry order either way.
  for (int sr0 = r0; sr0 < r1; sr0 += BT_SR) {
    const int sr1 = min(r1, sr0 + BT_SR);
    __syncthreads();  // the previous sub-chunk is done with the LDS
    for (int i = threadIdx.x; i <= sr1 - sr0; i += PSFUN_BS) s_rp[i] = rowptr[sr0 + i];
    __syncthreads();
    const int e0 = s_rp[0], ne = s_rp[sr1 - sr0] - e0;
    const bool staged = ne <= BT_CAP;
    if (staged && s_any) {
      for (int i = threadIdx.x; i < ne; i += PSFUN_BS) {
        s_col[i] = col[e0 + i];
        s_val[i] = val[e0 + i];
      }
    }
    __syncthreads();
    if (!active) continue;
    if (staged) {
      // two rows per thread at a time (r and r + RP): their gathers are in
      // flight together; each row's terms keep their entry order
      for (int r = sr0 + sub; r < sr1; r += 2 * RP) {
        const int r2 = r + RP;
        const bool two = r2 < sr1;
        const int s1 = s_rp[r - sr0], e1 = s_rp[r - sr0 + 1];
        const int s2 = two ? s_rp[r2 - sr0] : 0, e2 = two ? s_rp[r2 - sr0 + 1] : 0;
        const int lm = max(e1 - s1, e2 - s2);
        const double xr1 = src[(int64_t)r * K + k];
        const double xr2 = two ? src[(int64_t)r2 * K + k] : 0.0;
        double sum1 = 0.0, sum2 = 0.0;
        for (int t = 0; t < lm; t += BT_EB) {
          double a1[BT_EB], a2[BT_EB], g1[BT_EB], g2[BT_EB];
#pragma unroll
          for (int m = 0; m < BT_EB; ++m) {
            const int q1 = s1 + t + m, q2 = s2 + t + m;
            const bool ok1 = q1 < e1, ok2 = q2 < e2;
            a1[m] = ok1 ? s_val[q1 - e0] : 0.0;
            a2[m] = ok2 ? s_val[q2 - e0] : 0.0;
            g1[m] = ok1 ? src[(int64_t)s_col[q1 - e0] * K + k] : 0.0;
            g2[m] = ok2 ? src[(int64_t)s_col[q2 - e0] * K + k] : 0.0;
          }
#pragma unroll
          for (int m = 0; m < BT_EB; ++m) {
            if (s1 + t + m < e1) sum1 = fma(a1[m], g1[m] * inv, sum1);
            if (s2 + t + m < e2) sum2 = fma(a2[m], g2[m] * inv, sum2);
          }
        }
        const double v1 = xr1 * inv, u1 = opscale * sum1;
        if (j == 1) v_out[(int64_t)r * K + k] = xr1;  // basis slot 0 keeps x (v_1 = x / rs1)
        u_out[(int64_t)r * K + k] = u1;
        dot = fma(u1, v1, dot);
        if (two) {
          const double v2 = xr2 * inv, u2 = opscale * sum2;
          if (j == 1) v_out[(int64_t)r2 * K + k] = xr2;
          u_out[(int64_t)r2 * K + k] = u2;
          dot = fma(u2, v2, dot);
        }
      }
      continue;
    }
    for (int r = sr0 + sub; r < sr1; r += RP) {
      const int s = s_rp[r - sr0], e = s_rp[r - sr0 + 1];
      const double xr = src[(int64_t)r * K + k];
      double sum = 0.0;
      int q = s;
      if (staged) {
        for (; q + 4 <= e; q += 4) {
          const double x0 = src[(int64_t)s_col[q - e0] * K + k], x1 = src[(int64_t)s_col[q + 1 - e0] * K + k];
          const double x2 = src[(int64_t)s_col[q + 2 - e0] * K + k], x3 = src[(int64_t)s_col[q + 3 - e0] * K + k];
          sum = fma(s_val[q - e0], x0 * inv, sum);
          sum = fma(s_val[q + 1 - e0], x1 * inv, sum);
          sum = fma(s_val[q + 2 - e0], x2 * inv, sum);
          sum = fma(s_val[q + 3 - e0], x3 * inv, sum);
        }
        for (; q < e; ++q) sum = fma(s_val[q - e0], src[(int64_t)s_col[q - e0] * K + k] * inv, sum);
      } else {
        // entries four at a time: their column indices, then their gathers,
        // are issued together; the sum keeps the entry order
        for (; q + 4 <= e; q += 4) {
          const int c0 = col[q], c1 = col[q + 1], c2 = col[q + 2], c3 = col[q + 3];
          const double a0 = val[q], a1 = val[q + 1], a2 = val[q + 2], a3 = val[q + 3];
          const double x0 = src[(int64_t)c0 * K + k], x1 = src[(int64_t)c1 * K + k];
          const double x2 = src[(int64_t)c2 * K + k], x3 = src[(int64_t)c3 * K + k];
          sum = fma(a0, x0 * inv, sum);
          sum = fma(a1, x1 * inv, sum);
          sum = fma(a2, x2 * inv, sum);
          sum = fma(a3, x3 * inv, sum);
        }
        for (; q < e; ++q) sum = fma(val[q], src[(int64_t)col[q] * K + k] * inv, sum);
      }
      const double v = xr * inv;
      const double u = opscale * sum;
      if (j == 1) v_out[(int64_t)r * K + k] = xr;  // basis slot 0 keeps x (v_1 = x / rs1)
      u_out[(int64_t)r * K + k] = u;
      dot = fma(u, v, dot);
    }
  }
  red[threadIdx.x] = dot;
  __syncthreads();
  if (threadIdx.x < KP) {
    double s = 0.0;
    for (int q = 0; q < RP; ++q) s += red[q * KP + threadIdx.x];
    p_out[(int64_t)c * KP + threadIdx.x] = s;
  }
}

// Wide-load SpMM for even K (two columns per lane, H = KP / 2 lanes per row).
// Per 64-row window of a chunk the block stages in LDS the row pointers, the
// window's row schedule (its rows longest first, built at batch creation) and,
// for every entry, its value and the byte offset col * K * 8 of the row of src
// it gathers.  A wavefront takes W = 64 / H rows of the schedule at a time and
// walks their entries together up to the length of the group's first (longest)
// row, four entries per step: one 16-byte buffer load per lane and entry, with
// no per-entry branches.  An entry past its row's end loads from an offset
// beyond the buffer, which the hardware returns as 0, and adds 0 * 0.  Each
// column's sum runs over its row's entries in order with the same operations
// as k_bt_spmm, so u is the same to the bit; the alpha partials are summed in
// another order.  Against k_bt_spmm this issues about half the vector memory
// instructions and a third of the VALU work per nonzero, which is what bounds
// that kernel (its gathers wait on the texture-address path, not on HBM).
#define BTW_SENT 0x80000000u  // byte offset past every buffer this kernel reads
#define BTW_CAP 768           // entries staged per window
typedef double btw_d2 __attribute__((ext_vector_type(2)));
template <int KP>
__global__ __launch_bounds__(PSFUN_BS) void k_bt_spmm_wide(const int *__restrict__ rowptr, const int *__restrict__ col,
                                                           const double *__restrict__ val, double opscale,
                                                           const int *__restrict__ chunk,
                                                           const unsigned char *__restrict__ sched, int K, int j,
                                                           const double *__restrict__ src, unsigned int src_bytes,
                                                           unsigned int nnz, double *__restrict__ v_out,
                                                           double *__restrict__ u_out, double *__restrict__ p_out,
                                                           BtInst in, BtCtl *bctl, int nc) {
  constexpr int H = KP / 2, W = 64 / H;
  static_assert(2 * 2 * PSFUN_BS >= BTW_CAP + 1 && 4 * PSFUN_BS >= BTW_CAP + 3, "one staging pass per window");
  __shared__ double s_inv[KP];
  __shared__ int s_act[KP];
  __shared__ int s_any;
  __shared__ double red[2 * PSFUN_BS];
  __shared__ int s_rp[BT_SR + 1];
  __shared__ unsigned char s_sc[BT_SR];
  __shared__ unsigned int s_off[BTW_CAP + 64];
  __shared__ double s_val[BTW_CAP + 64];
  if (bctl->remaining == 0) return;
  const int c = bt_block_chunk(nc);
  if (c < 0) return;
  const int r0 = chunk[3 * c], r1 = chunk[3 * c + 1], g = chunk[3 * c + 2];
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  if (threadIdx.x < KP) {
    const int k = threadIdx.x;
    int act = 0;
    double inv = 0.0;
    if (k < K) {
      const int I = g * K + k;
      if (j == 1) {
        const double sg = in.rs1[I];  // folded by k_bt_fold<KP, 1>
        act = (sg != 0.0);
        inv = act ? 1.0 / sg : 0.0;
      } else {
        act = (in.done[I] == BT_RUN);
        if (act) inv = 1.0 / in.beta[(int64_t)I * in.ld + j - 1];
      }
    }
    s_inv[k] = inv;
    s_act[k] = act;
    if (act) s_any = 1;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int h = lane % H, jr = lane / H;
  const int K2 = K / 2;
  const bool act = s_any && h < K2 && (s_act[2 * h] || s_act[2 * h + 1]);
  const double inv0 = act ? s_inv[2 * h] : 0.0, inv1 = act ? s_inv[2 * h + 1] : 0.0;
  const unsigned int kb = (unsigned int)K * 8u, h16 = 16u * (unsigned int)h;
  const __amdgpu_buffer_rsrc_t xs = __builtin_amdgcn_make_buffer_rsrc((void *)src, (short)0, (int)src_bytes, 0x00020000);
  // the arrays are padded to whole 16-byte groups (bt_create)
  const unsigned int nnz4 = (nnz + 3u) & ~3u;
  const __amdgpu_buffer_rsrc_t cs = __builtin_amdgcn_make_buffer_rsrc((void *)col, (short)0, (int)(4u * nnz4), 0x00020000);
  const __amdgpu_buffer_rsrc_t vsr = __builtin_amdgcn_make_buffer_rsrc((void *)val, (short)0, (int)(8u * nnz4), 0x00020000);
  const double2 *src2 = reinterpret_cast<const double2 *>(src);
  double2 *u2 = reinterpret_cast<double2 *>(u_out);
  double2 *v2 = reinterpret_cast<double2 *>(v_out);
  double dot0 = 0.0, dot1 = 0.0;
  for (int sr0 = r0; s_any && sr0 < r1; sr0 += BT_SR) {
    const int nr = min(r1 - sr0, BT_SR);
    __syncthreads();  // the previous window is done with the LDS
    if (threadIdx.x <= nr) s_rp[threadIdx.x] = rowptr[sr0 + threadIdx.x];
    if (threadIdx.x < nr) s_sc[threadIdx.x] = sched[sr0 + threadIdx.x];
    __syncthreads();
    const int e0 = s_rp[0], ne = s_rp[nr] - e0;
    if (ne <= BTW_CAP) {
      // The window's entries in 16-byte buffer loads from the aligned entry
      // below e0 (four column indices, or two values, per lane; reads past
      // the arrays return 0), all in flight before the LDS stores.  Slots
      // past ne are never read as entries (the walk below masks entries past
      // a row's end).
      const int a0 = e0 & ~3, b0 = e0 & ~1;
      const int4 cw = __builtin_bit_cast(
          int4, __builtin_amdgcn_raw_buffer_load_b128(cs, (unsigned int)(a0 + 4 * (int)threadIdx.x) * 4u, 0, 0));
      btw_d2 vw[2];
#pragma unroll
      for (int it = 0; it < 2; ++it)
        vw[it] = __builtin_bit_cast(
            btw_d2, __builtin_amdgcn_raw_buffer_load_b128(
                        vsr, (unsigned int)(b0 + 2 * ((int)threadIdx.x + it * PSFUN_BS)) * 8u, 0, 0));
      {
        const int i0 = a0 + 4 * (int)threadIdx.x - e0;
        const int c4[4] = {cw.x, cw.y, cw.z, cw.w};
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (i0 + m >= 0 && i0 + m < ne) s_off[i0 + m] = (unsigned int)c4[m] * kb;
      }
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const int i0 = b0 + 2 * ((int)threadIdx.x + it * PSFUN_BS) - e0;
        if (i0 >= 0 && i0 < ne) s_val[i0] = vw[it].x;
        if (i0 + 1 >= 0 && i0 + 1 < ne) s_val[i0 + 1] = vw[it].y;
      }
      __syncthreads();
      for (int g0 = wv * W; g0 < nr; g0 += 4 * W) {
        const int slot = g0 + jr;
        const bool rok = act && slot < nr;
        const int lr = s_sc[slot < nr ? slot : g0];
        const int es = s_rp[lr] - e0, len = rok ? s_rp[lr + 1] - s_rp[lr] : 0;
        const int lf = s_sc[g0];
        const int M = __builtin_amdgcn_readfirstlane(s_rp[lf + 1] - s_rp[lf]);  // longest row of the group
        double sum0 = 0.0, sum1 = 0.0;
        for (int m = 0; m < M; m += 4) {
          btw_d2 x[4];
          double a[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const bool ok = m + q < len;
            const int ix = min(es + m + q, BTW_CAP + 63);
            x[q] = __builtin_bit_cast(
                btw_d2, __builtin_amdgcn_raw_buffer_load_b128(xs, ok ? s_off[ix] + h16 : BTW_SENT, 0, 0));
            a[q] = ok ? s_val[ix] : 0.0;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            sum0 = fma(a[q], x[q].x * inv0, sum0);
            sum1 = fma(a[q], x[q].y * inv1, sum1);
          }
        }
        if (rok) {
          const int64_t i2 = (int64_t)(sr0 + lr) * K2 + h;
          const double2 xv = src2[i2];
          const double2 uu = make_double2(opscale * sum0, opscale * sum1);
          if (j == 1) v2[i2] = xv;  // basis slot 0 keeps x (v_1 = x / rs1)
          u2[i2] = uu;
          dot0 = fma(uu.x, xv.x * inv0, dot0);
          dot1 = fma(uu.y, xv.y * inv1, dot1);
        }
      }
    } else {
      // more entries than the stage holds: the window's rows in order, their
      // entries straight from memory
      for (int lr = wv * W + jr; lr - jr < nr; lr += 4 * W) {
        if (!act || lr >= nr) continue;
        const int64_t i2 = (int64_t)(sr0 + lr) * K2 + h;
        const int s = s_rp[lr], e = s_rp[lr + 1];
        double sum0 = 0.0, sum1 = 0.0;
        for (int q = s; q < e; ++q) {
          const btw_d2 x = __builtin_bit_cast(
              btw_d2, __builtin_amdgcn_raw_buffer_load_b128(xs, (unsigned int)col[q] * kb + h16, 0, 0));
          const double a = val[q];
          sum0 = fma(a, x.x * inv0, sum0);
          sum1 = fma(a, x.y * inv1, sum1);
        }
        const double2 xv = src2[i2];
        const double2 uu = make_double2(opscale * sum0, opscale * sum1);
        if (j == 1) v2[i2] = xv;
        u2[i2] = uu;
        dot0 = fma(uu.x, xv.x * inv0, dot0);
        dot1 = fma(uu.y, xv.y * inv1, dot1);
      }
    }
  }
  red[2 * threadIdx.x] = dot0;
  red[2 * threadIdx.x + 1] = dot1;
  __syncthreads();
  if (threadIdx.x < KP) {
    const int hh = threadIdx.x >> 1, odd = threadIdx.x & 1;
    double s = 0.0;
    for (int q = 0; q < PSFUN_BS / H; ++q) s += red[2 * (q * H + hh) + odd];
    p_out[(int64_t)c * KP + threadIdx.x] = s;
  }
}

template <int KP>
__global__ __launch_bounds__(PSFUN_BS) void k_bt_update(const int *__restrict__ chunk, const int *__restrict__ gchunk,
                                                        int K, int j, const double *__restrict__ u,
                                                        double *__restrict__ w_out,
                                                        const double *__restrict__ vj, const double *__restrict__ vjm1,
                                                        const double *__restrict__ p_in, double *__restrict__ p_out,
                                                        BtInst in, const BtCtl *bctl) {
  constexpr int RP = PSFUN_BS / KP;
  __shared__ double s_al[KP], s_be[KP], s_ia[KP], s_ib[KP];
  __shared__ int s_act[KP];
  __shared__ double red[PSFUN_BS];
  if (bctl->remaining == 0) return;
  const int c = blockIdx.x;
  const int r0 = chunk[3 * c], r1 = chunk[3 * c + 1], g = chunk[3 * c + 2];
  if (threadIdx.x < KP) {
    const int k = threadIdx.x;
    int act = 0;
    double al = 0.0, be = 0.0, ia = 0.0, ib = 0.0;
    if (k < K) {
      const int I = g * K + k;
      const int64_t base = (int64_t)I * in.ld;
      act = (in.done[I] == BT_RUN);
      if (act) {
        al = in.alpha[base + j];  // folded by k_bt_fold<KP, 0>
        be = (j > 1) ? in.beta[base + j - 1] : 0.0;
        // v_j = w_{j-1} / beta_{j-1}, v_{j-1} = w_{j-2} / beta_{j-2} (beta_0 = rs1),
        // with the same reciprocals the SpMM steps used
        const double bj = (j == 1) ? in.rs1[I] : in.beta[base + j - 1];
        ia = 1.0 / bj;
        if (j > 1) ib = 1.0 / ((j == 2) ? in.rs1[I] : in.beta[base + j - 2]);
      }
    }
    s_al[k] = al;
    s_be[k] = be;
    s_ia[k] = ia;
    s_ib[k] = ib;
    s_act[k] = act;
  }
  __syncthreads();
  const int k = threadIdx.x % KP, sub = threadIdx.x / KP;
  double acc = 0.0;
  if (k < K && s_act[k]) {
    const double al = s_al[k], be = s_be[k], ia = s_ia[k], ib = s_ib[k];
    // BT_UB rows per thread at a time: their loads are issued together; the
    // ||w||^2 partial accumulates in row order as before
    int r = r0 + sub;
    for (; r + (BT_UB - 1) * RP < r1; r += BT_UB * RP) {
      double uu[BT_UB], a[BT_UB], b[BT_UB];
#pragma unroll
      for (int q = 0; q < BT_UB; ++q) {
        const int64_t i = (int64_t)(r + q * RP) * K + k;
        uu[q] = u[i];
        a[q] = vj[i];
        b[q] = (j > 1) ? vjm1[i] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < BT_UB; ++q) {
        double w = fma(-al, a[q] * ia, uu[q]);
        if (j > 1) w = fma(-be, b[q] * ib, w);
        w_out[(int64_t)(r + q * RP) * K + k] = w;
        acc = fma(w, w, acc);
      }
    }
    for (; r < r1; r += RP) {
      const int64_t i = (int64_t)r * K + k;
      double w = fma(-al, vj[i] * ia, u[i]);
      if (j > 1) w = fma(-be, vjm1[i] * ib, w);
      w_out[i] = w;
      acc = fma(w, w, acc);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < KP) {
    double s = 0.0;
    for (int q = 0; q < RP; ++q) s += red[q * KP + threadIdx.x];
    p_out[(int64_t)c * KP + threadIdx.x] = s;
  }
}

// Loop control at index j+1, split in two so that the estimate does not hold
// up the next SpMM (which needs only beta and the run state):
//   k_bt_beta (stream order): one wavefront per instance folds ||w||^2 into
//     beta = T(j, j+1), records it and itx, applies the breakdown test and,
//     when no estimate is taken (istop 2), the iteration limit;
//   k_bt_est (beside the SpMM, on the auxiliary stream): the reference's
//     estimate yk = f(T_j) e1, rs(j+1) = rs1 beta |yk(j)| / ||yk||, the
//     convergence test and the iteration limit.  An instance it stops may
//     have its SpMM work of the same index wasted, which feeds nothing.
template <int KP>
__global__ __launch_bounds__(64) void k_bt_beta(const int *__restrict__ gchunk, int K, int j,
                                                const double *__restrict__ p_in, BtInst in, BtCtl *bctl, int nl,
                                                int istop) {
  const int I = blockIdx.x;
  if (in.done[I] != BT_RUN) return;
  const int g = I / K, k = I % K;
  const int c0 = gchunk[2 * g], c1 = gchunk[2 * g + 1];
  double acc = 0.0;
  for (int cc = c0 + threadIdx.x; cc < c1; cc += 64) acc += p_in[(int64_t)cc * KP + k];
  const double beta = sqrt(wave_sum(acc));
  if (threadIdx.x != 0) return;
  const int64_t base = (int64_t)I * in.ld;
  in.beta[base + j - 1] = beta;
  in.itx[I] = j - 1;
  int fin = BT_RUN;
  if (fabs(beta) < 1e-16)
    fin = BT_BRK;
  else if (istop != 1 && j >= nl)
    fin = BT_MAXIT;
  if (fin != BT_RUN) {
    in.done[I] = fin;
    atomicSub(&bctl->remaining, 1);
  }
}

// ORD: the largest order of T this launch sequence reaches.  0: <= 64 (itmax
// <= 65), 1: <= 128 (itmax <= 129): only the register paths run (one or two
// rows of T per lane) and the LDS holds the coefficient table alone, so more
// waves fit per CU beside the SpMM this kernel overlaps.  2: any order.
template <int ORD>
__global__ __launch_bounds__(64) void k_bt_est(int j, BtInst in, BtCtl *bctl, int nl, double eps, double fscale) {
  __shared__ typename std::conditional<ORD < 2, ChebCoefLds, ChebLds>::type W;
  const int I = blockIdx.x;
  if (in.done[I] != BT_RUN) return;
  const int64_t base = (int64_t)I * in.ld;
  const double beta = in.beta[base + j - 1];
  const int lane = threadIdx.x;
  int fin = BT_RUN;
  double nrm, last;
  if constexpr (ORD == 0)
    wave_exp_tridiag_e1_small(j - 1, in.alpha + base + 1, in.beta + base + 1, fscale, &W, in.yk + base, nrm, last);
  else if constexpr (ORD == 1) {
    if (j - 1 <= 64)
      wave_exp_tridiag_e1_small(j - 1, in.alpha + base + 1, in.beta + base + 1, fscale, &W, in.yk + base, nrm, last);
    else
      wave_exp_tridiag_e1_mid(j - 1, in.alpha + base + 1, in.beta + base + 1, fscale, &W, in.yk + base, nrm, last);
  } else
    wave_exp_tridiag_e1(j - 1, in.alpha + base + 1, in.beta + base + 1, fscale, &W, in.yk + base, nrm, last);
  const double errnum = in.rs1[I] * beta * fabs(last);
  const double rsj = errnum / nrm;
  if (lane == 0) {
    in.errnum[I] = errnum;
    in.errden[I] = nrm;
    in.rs[base + j] = rsj;
  }
  if (rsj < eps) fin = BT_CONV;
  if (fin == BT_RUN && j >= nl) fin = BT_MAXIT;
  if (fin != BT_RUN && lane == 0) {
    if (fin == BT_CONV) in.m[I] = j - 1;
    in.done[I] = fin;
    atomicSub(&bctl->remaining, 1);
  }
}

__global__ __launch_bounds__(64) void k_bt_final(BtInst in, double fscale) {
  __shared__ ChebLds W;
  const int I = blockIdx.x;
  const int done = in.done[I];
  const int64_t base = (int64_t)I * in.ld;
  int m = 0;
  if (done == BT_CONV) {
    m = in.m[I];
  } else if (done != BT_ZERO) {
    int kk = in.itx[I] - 1;
    if (kk < 1) kk = 1;
    double nrm, last;
    wave_exp_tridiag_e1(kk, in.alpha + base + 1, in.beta + base + 1, fscale, &W, in.yk + base, nrm, last);
    if (threadIdx.x == 0) in.m[I] = kk;
    m = kk;
  }
  // factors of the basis slots the assembly reads
  for (int l = threadIdx.x; l < m; l += 64) {
    const double b = (l == 0) ? in.rs1[I] : in.beta[base + l];
    in.ib[base + l] = (b != 0.0) ? 1.0 / b : 0.0;
  }
}

// perm != nullptr: the batch is reordered and row r of the result goes to
// row perm[r] of Y (the caller's numbering).
template <int KP>
__global__ __launch_bounds__(PSFUN_BS) void k_bt_assemble(const int *__restrict__ chunk, int K,
                                                          const double *__restrict__ V, int64_t vstride, BtInst in,
                                                          const int *__restrict__ perm, double *__restrict__ Y) {
  constexpr int RP = PSFUN_BS / KP;
  __shared__ int s_m[KP];
  __shared__ double s_rs1[KP];
  const int c = blockIdx.x;
  const int r0 = chunk[3 * c], r1 = chunk[3 * c + 1], g = chunk[3 * c + 2];
  if (threadIdx.x < KP) {
    const int k = threadIdx.x;
    int m = 0;
    double rs1 = 0.0;
    if (k < K) {
      const int I = g * K + k;
      const int done = in.done[I];
      m = in.m[I];
      if (done == BT_CONV) m = (m - 1 < 1) ? 1 : m - 1;
      if (done == BT_ZERO || done == BT_RUN) m = 0;
      rs1 = in.rs1[I];
    }
    s_m[k] = m;
    s_rs1[k] = rs1;
  }
  __syncthreads();
  const int k = threadIdx.x % KP, sub = threadIdx.x / KP;
  if (k >= K) return;
  const int m = s_m[k];
  const double *yk = in.yk + (int64_t)(g * K + k) * in.ld;
  const double *ibk = in.ib + (int64_t)(g * K + k) * in.ld;  // v_{l+1} = slot l * ibk[l]
  // BT_AR rows per thread at a time, the slots in order for each (the basis is
  // not read again: nontemporal loads): the BT_AR loads of a slot are in
  // flight together and share one load of that slot's two coefficients
  int r = r0 + sub;
  for (; r + (BT_AR - 1) * RP < r1; r += BT_AR * RP) {
    double acc[BT_AR];
#pragma unroll
    for (int q = 0; q < BT_AR; ++q) acc[q] = 0.0;
    if (m > 0) {
      const double y0 = yk[0], b0 = ibk[0];
#pragma unroll
      for (int q = 0; q < BT_AR; ++q)
        acc[q] = y0 * (__builtin_nontemporal_load(V + (int64_t)(r + q * RP) * K + k) * b0);
      for (int l = 1; l < m; ++l) {
        const double yl = yk[l], bl = ibk[l];
        const double *Vl = V + (int64_t)l * vstride;
#pragma unroll
        for (int q = 0; q < BT_AR; ++q)
          acc[q] = fma(yl, __builtin_nontemporal_load(Vl + (int64_t)(r + q * RP) * K + k) * bl, acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < BT_AR; ++q) {
      const int rq = r + q * RP;
      Y[perm ? (int64_t)perm[rq] * K + k : (int64_t)rq * K + k] = s_rs1[k] * acc[q];
    }
  }
  for (; r < r1; r += RP) {
    const int64_t i = (int64_t)r * K + k;
    double acc = 0.0;
    if (m > 0) {
      acc = yk[0] * (__builtin_nontemporal_load(V + i) * ibk[0]);
#pragma unroll 8
      for (int l = 1; l < m; ++l)
        acc = fma(yk[l], __builtin_nontemporal_load(V + (int64_t)l * vstride + i) * ibk[l], acc);
    }
    Y[perm ? (int64_t)perm[r] * K + k : i] = s_rs1[k] * acc;
  }
}

// Xp[r] = X[perm[r]] row by row (K doubles per row): a reordered batch's
// input in its own numbering.
__global__ __launch_bounds__(PSFUN_BS) void k_bt_gather_rows(int64_t nrows, int K, const int *__restrict__ perm,
                                                            const double *__restrict__ X, double *__restrict__ Xp) {
  const int64_t tot = nrows * K;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < tot; i += (int64_t)gridDim.x * PSFUN_BS) {
    const int64_t r = i / K;
    Xp[i] = X[(int64_t)perm[r] * K + (i - r * K)];
  }
}

// Y[perm[r]] = Yp[r] + beta * Y[perm[r]] row by row.
__global__ __launch_bounds__(PSFUN_BS) void k_bt_scatter_rows(int64_t nrows, int K, const int *__restrict__ perm,
                                                             const double *__restrict__ Yp, double beta,
                                                             double *__restrict__ Y) {
  const int64_t tot = nrows * K;
  for (int64_t i = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; i < tot; i += (int64_t)gridDim.x * PSFUN_BS) {
    const int64_t r = i / K;
    const int64_t o = (int64_t)perm[r] * K + (i - r * K);
    Y[o] = (beta != 0.0) ? fma(beta, Y[o], Yp[i]) : Yp[i];
  }
}

// val[q] = src[nzperm[q]]: values given in the caller's entry order.
__global__ __launch_bounds__(PSFUN_BS) void k_bt_gather_vals(int64_t nnz, const int *__restrict__ nzperm,
                                                            const double *__restrict__ src, double *__restrict__ val) {
  for (int64_t q = (int64_t)blockIdx.x * PSFUN_BS + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * PSFUN_BS)
    val[q] = src[nzperm[q]];
}

__global__ void k_bt_init(BtInst in, int ninst, BtCtl *bctl) {
  const int64_t tot = (int64_t)ninst * in.ld;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x)
    in.rs[i] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) {
    in.done[i] = BT_RUN;
    in.itx[i] = 0;
    in.m[i] = 0;
    in.errnum[i] = 0.0;
    in.errden[i] = 1.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) bctl->remaining = ninst;
}

// ---------------------------------------------------------------------------
static int bt_build_chunks(psfun_batch *B) {
  std::vector<int> ch, gc(2 * B->ngraph);
  for (int g = 0; g < B->ngraph; ++g) {
    gc[2 * g] = (int)(ch.size() / 3);
    for (int64_t r = B->row_off[g]; r < B->row_off[g + 1]; r += B->rows_per_chunk) {
      ch.push_back((int)r);
      ch.push_back((int)std::min<int64_t>(r + B->rows_per_chunk, B->row_off[g + 1]));
      ch.push_back(g);
    }
    gc[2 * g + 1] = (int)(ch.size() / 3);
  }
  psfun_ctx *c = B->ctx;
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  if (B->d_chunk) PSFUN_HIP_TRY(hipFree(B->d_chunk));
  B->d_chunk = nullptr;
  B->nchunks = (int)(ch.size() / 3);
  PSFUN_HIP_TRY(hipMalloc(&B->d_chunk, sizeof(int) * (ch.size() + gc.size() + 3)));
  B->d_gchunk = B->d_chunk + ch.size();
  if (!ch.empty()) PSFUN_HIP_TRY(hipMemcpy(B->d_chunk, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice));
  PSFUN_HIP_TRY(hipMemcpy(B->d_gchunk, gc.data(), sizeof(int) * gc.size(), hipMemcpyHostToDevice));
  const size_t need = 2 * (size_t)std::max(B->nchunks, 1) * BT_KMAX;
  if (need > B->part_cap) {
    if (B->d_part) PSFUN_HIP_TRY(hipFree(B->d_part));
    PSFUN_HIP_TRY(hipMalloc(&B->d_part, sizeof(double) * need));
    B->part_cap = need;
  }
  for (auto &kv : B->graphs) hipGraphExecDestroy(kv.second);
  B->graphs.clear();
  return PSFUN_SUCCESS;
}

static int bt_validate(int ngraph, const int64_t *row_offsets, const int32_t *rowptr, const int32_t *colind,
                       const double *val) {
  if (ngraph < 0 || (ngraph && (!row_offsets || !rowptr))) return PSFUN_ERR_INVALID_ARG;
  const int64_t nrows = ngraph ? row_offsets[ngraph] : 0;
  if (nrows > INT32_MAX - 1 || (ngraph && row_offsets[0] != 0)) return PSFUN_ERR_INVALID_ARG;
  const int64_t nnz = nrows ? rowptr[nrows] : 0;
  if (rowptr && rowptr[0] != 0) return PSFUN_ERR_INVALID_ARG;
  if (nnz && (!colind || !val)) return PSFUN_ERR_INVALID_ARG;
  for (int g = 0; g < ngraph; ++g) {
    if (row_offsets[g + 1] < row_offsets[g]) return PSFUN_ERR_INVALID_ARG;
    for (int64_t r = row_offsets[g]; r < row_offsets[g + 1]; ++r) {
      if (rowptr[r + 1] < rowptr[r]) return PSFUN_ERR_INVALID_ARG;
      for (int q = rowptr[r]; q < rowptr[r + 1]; ++q)
        if (colind[q] < row_offsets[g] || colind[q] >= row_offsets[g + 1]) return PSFUN_ERR_INVALID_ARG;
    }
  }
  return PSFUN_SUCCESS;
}

// Device batch from validated host arrays in the order given.
static int bt_create(psfun_ctx_t c, int ngraph, const int64_t *row_offsets, const int32_t *rowptr,
                     const int32_t *colind, const double *val, psfun_batch_t *out) {
  const int64_t nrows = ngraph ? row_offsets[ngraph] : 0;
  const int64_t nnz = nrows ? rowptr[nrows] : 0;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  psfun_batch *B = new psfun_batch();
  B->ctx = c;
  B->device = c->device;
  B->ngraph = ngraph;
  B->nrows = nrows;
  B->nnz = nnz;
  B->row_off.assign(row_offsets, row_offsets + ngraph + 1);
  // column indices and values padded with zeros to whole 16-byte groups
  // (four indices, two values), so the wide SpMM's 16-byte staging loads are
  // either wholly inside the arrays or wholly past them
  const int64_t nnz4 = std::max<int64_t>((nnz + 3) / 4 * 4, 4);
  hipError_t e = hipMalloc(&B->rowptr, sizeof(int) * (nrows + 1));
  if (e == hipSuccess) e = hipMalloc(&B->colind, sizeof(int) * nnz4);
  if (e == hipSuccess) e = hipMalloc(&B->val, sizeof(double) * nnz4);
  if (e == hipSuccess) e = hipMemset(B->colind, 0, sizeof(int) * nnz4);
  if (e == hipSuccess) e = hipMemset(B->val, 0, sizeof(double) * nnz4);
  if (e == hipSuccess) e = hipMemcpy(B->rowptr, rowptr, sizeof(int) * (nrows + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(B->colind, colind, sizeof(int) * nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(B->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    psfun_record_hip_error(e, "psfun_batch_create", __FILE__, __LINE__);
    psfun_batch_destroy(B);
    return PSFUN_ERR_HIP;
  }
  // row schedule of every 64-row window: local rows by decreasing length
  // (ties in row order), read by the wide-load SpMM
  {
    std::vector<unsigned char> sc((size_t)std::max<int64_t>(nrows, 1));
    std::vector<int> idx(BT_SR);
    for (int g = 0; g < ngraph; ++g)
      for (int64_t a = row_offsets[g]; a < row_offsets[g + 1]; a += BT_SR) {
        const int nw = (int)std::min<int64_t>(BT_SR, row_offsets[g + 1] - a);
        for (int i = 0; i < nw; ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.begin() + nw, [&](int x, int y) {
          return rowptr[a + x + 1] - rowptr[a + x] > rowptr[a + y + 1] - rowptr[a + y];
        });
        for (int i = 0; i < nw; ++i) sc[(size_t)(a + i)] = (unsigned char)idx[i];
      }
    e = hipMalloc(&B->d_sched, sc.size());
    if (e == hipSuccess) e = hipMemcpy(B->d_sched, sc.data(), sc.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      psfun_record_hip_error(e, "psfun_batch_create", __FILE__, __LINE__);
      psfun_batch_destroy(B);
      return PSFUN_ERR_HIP;
    }
  }
  const int st = bt_build_chunks(B);
  if (st != PSFUN_SUCCESS) {
    psfun_batch_destroy(B);
    return st;
  }
  *out = B;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_batch_create(psfun_ctx_t c, int ngraph, const int64_t *row_offsets, const int32_t *rowptr,
                                  const int32_t *colind, const double *val, psfun_batch_t *out) {
  return psfun_batch_create_ordered(c, ngraph, row_offsets, rowptr, colind, val, PSFUN_ORDER_NATURAL, out);
}

extern "C" int psfun_batch_create_ordered(psfun_ctx_t c, int ngraph, const int64_t *row_offsets,
                                          const int32_t *rowptr, const int32_t *colind, const double *val, int order,
                                          psfun_batch_t *out) {
  if (!c || !out) return PSFUN_ERR_INVALID_ARG;
  *out = nullptr;
  if (order != PSFUN_ORDER_NATURAL && order != PSFUN_ORDER_RCM) return PSFUN_ERR_INVALID_ARG;
  PSFUN_TRY(bt_validate(ngraph, row_offsets, rowptr, colind, val));
  if (order == PSFUN_ORDER_NATURAL) return bt_create(c, ngraph, row_offsets, rowptr, colind, val, out);
  // reverse Cuthill-McKee inside every graph; each row keeps its entries in
  // the given order (so a row sum is the same sum as without reordering)
  const int64_t nrows = ngraph ? row_offsets[ngraph] : 0;
  const int64_t nnz = nrows ? rowptr[nrows] : 0;
  std::vector<int32_t> perm((size_t)nrows), iperm((size_t)nrows), rp2((size_t)nrows + 1), ci2((size_t)nnz),
      nzp((size_t)nnz);
  std::vector<double> va2((size_t)nnz);
  for (int g = 0; g < ngraph; ++g) {
    const int64_t r0 = row_offsets[g], ng = row_offsets[g + 1] - r0;
    PSFUN_TRY(psfun_rcm_order((int)ng, rowptr + r0, colind, (int)r0, perm.data() + r0));
    for (int64_t p = 0; p < ng; ++p) perm[r0 + p] += (int32_t)r0;
  }
  for (int64_t r = 0; r < nrows; ++r) iperm[perm[r]] = (int32_t)r;
  int64_t pos = 0;
  rp2[0] = 0;
  for (int64_t r = 0; r < nrows; ++r) {
    const int32_t o = perm[r];
    for (int q = rowptr[o]; q < rowptr[o + 1]; ++q) {
      ci2[pos] = iperm[colind[q]];
      va2[pos] = val[q];
      nzp[pos] = q;
      ++pos;
    }
    rp2[r + 1] = (int32_t)pos;
  }
  psfun_batch_t B = nullptr;
  PSFUN_TRY(bt_create(c, ngraph, row_offsets, rp2.data(), ci2.data(), va2.data(), &B));
  hipError_t e = hipMalloc(&B->d_perm, sizeof(int) * std::max<int64_t>(nrows, 1));
  if (e == hipSuccess) e = hipMalloc(&B->d_nzperm, sizeof(int) * std::max<int64_t>(nnz, 1));
  if (e == hipSuccess && nrows) e = hipMemcpy(B->d_perm, perm.data(), sizeof(int) * nrows, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(B->d_nzperm, nzp.data(), sizeof(int) * nnz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    psfun_record_hip_error(e, "psfun_batch_create_ordered", __FILE__, __LINE__);
    psfun_batch_destroy(B);
    return PSFUN_ERR_HIP;
  }
  B->order = PSFUN_ORDER_RCM;
  *out = B;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_batch_destroy(psfun_batch_t B) {
  if (!B) return PSFUN_SUCCESS;
  hipSetDevice(B->device);  // the context may already be gone
  hipDeviceSynchronize();
  for (auto &kv : B->graphs) hipGraphExecDestroy(kv.second);
  if (B->ev_in) hipEventDestroy(B->ev_in);
  if (B->ev_out) hipEventDestroy(B->ev_out);
  if (B->rowptr) hipFree(B->rowptr);
  if (B->colind) hipFree(B->colind);
  if (B->val) hipFree(B->val);
  if (B->d_chunk) hipFree(B->d_chunk);
  if (B->d_part) hipFree(B->d_part);
  if (B->d_perm) hipFree(B->d_perm);
  if (B->d_nzperm) hipFree(B->d_nzperm);
  if (B->d_sched) hipFree(B->d_sched);
  delete B;
  return PSFUN_SUCCESS;
}

extern "C" int psfun_batch_set_tuning(psfun_batch_t B, int rows_per_chunk, int use_graph) {
  if (!B || rows_per_chunk < 16 || rows_per_chunk > (1 << 20)) return PSFUN_ERR_INVALID_ARG;
  B->use_graph = use_graph ? 1 : 0;
  if (rows_per_chunk != B->rows_per_chunk) {
    B->rows_per_chunk = rows_per_chunk;
    return bt_build_chunks(B);
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_batch_info(psfun_batch_t B, int *ngraph, int64_t *nrows, int64_t *nnz, int *rows_per_chunk,
                                int *nchunks) {
  if (!B) return PSFUN_ERR_INVALID_ARG;
  if (ngraph) *ngraph = B->ngraph;
  if (nrows) *nrows = B->nrows;
  if (nnz) *nnz = B->nnz;
  if (rows_per_chunk) *rows_per_chunk = B->rows_per_chunk;
  if (nchunks) *nchunks = B->nchunks;
  return PSFUN_SUCCESS;
}

// The SpMM step u = opscale * A (src / sigma) with the alpha partials: the
// wide-load kernel where it applies (even K >= 4, chunks made of whole 64-row
// windows, the gathered operand and the entry arrays under 2 GiB so every
// byte offset and the out-of-range sentinel fit 32 bits), else k_bt_spmm.
static bool bt_wide(const psfun_batch *B, int K) {
  return K >= 4 && K % 2 == 0 && B->rows_per_chunk % BT_SR == 0 && B->d_sched &&
         (double)B->nrows * K * 8.0 < (double)BTW_SENT && (double)(B->nnz + 4) * 8.0 < (double)BTW_SENT;
}
template <int KP>
static void bt_spmm_launch(psfun_ctx *c, psfun_batch *B, double opscale, int K, int j, const double *src,
                           double *v_out, double *u_out, const double *p_in, double *p_out, const BtInst &in,
                           BtCtl *bctl) {
  const int NC = B->nchunks;
  if constexpr (KP >= 4) {
    if (bt_wide(B, K)) {
      hipLaunchKernelGGL((k_bt_spmm_wide<KP>), dim3(bt_spmm_grid(NC)), dim3(PSFUN_BS), 0, c->stream, B->rowptr,
                         B->colind, B->val, opscale, B->d_chunk, B->d_sched, K, j, src,
                         (unsigned int)(B->nrows * K * 8), (unsigned int)B->nnz, v_out, u_out, p_out, in, bctl, NC);
      return;
    }
  }
  hipLaunchKernelGGL((k_bt_spmm<KP>), dim3(bt_spmm_grid(NC)), dim3(PSFUN_BS), 0, c->stream, B->rowptr, B->colind,
                     B->val, opscale, B->d_chunk, B->d_gchunk, K, j, src, v_out, u_out, p_in, p_out, in, bctl, NC);
}

template <int KP>
static int bt_enqueue(psfun_ctx *c, psfun_batch *B, int K, double fscale, double opscale, const double *X, double *Y,
                      double eps, int itmax, int istop, const BtInst &in, BtCtl *bctl, double *V, double *U0,
                      double *U1, double *Xp, bool poll_host) {
  const int NC = B->nchunks;
  const int ninst = B->ngraph * K;
  const int64_t vs = B->nrows * K;
  if (B->d_perm) {  // reordered batch: X into its numbering first
    hipLaunchKernelGGL(k_bt_gather_rows, dim3(psfun_grid(c, (vs + PSFUN_BS - 1) / PSFUN_BS, 8)), dim3(PSFUN_BS), 0,
                       c->stream, (int64_t)B->nrows, K, (const int *)B->d_perm, X, Xp);
    X = Xp;
  }
  double *P0 = B->d_part, *P1 = B->d_part + (size_t)NC * BT_KMAX;
  hipLaunchKernelGGL(k_bt_init, dim3(std::min(1024, (ninst * in.ld + 255) / 256 + 1)), dim3(256), 0, c->stream, in,
                     ninst, bctl);
  hipLaunchKernelGGL((k_bt_norm<KP>), dim3(NC), dim3(PSFUN_BS), 0, c->stream, B->d_chunk, K, X, P0);
  hipLaunchKernelGGL((k_bt_fold<KP, 1>), dim3(ninst), dim3(64), 0, c->stream, B->d_gchunk, K, 0, P0, in, bctl);
  // src read and u written (the first step also copies x into basis slot 0)
  const double spmm_bytes = 12.0 * B->nnz + 4.0 * (B->nrows + 1) + 16.0 * B->nrows * K;
  PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_BATCH, spmm_bytes + 8.0 * B->nrows * K));
  bt_spmm_launch<KP>(c, B, opscale, K, 1, X, V, U0, P0, P1, in, bctl);
  PSFUN_TRY(psfun_prof_end(c));
  PSFUN_HIP_TRY(hipGetLastError());
  volatile int *hflag = (volatile int *)c->h_pinned;
  int pending = -1;
  for (int j = 1; j <= itmax - 1; ++j) {
    double *u = (j & 1) ? U0 : U1;
    const double *vj = V + (size_t)vs * (j - 1);
    const double *vjm1 = (j > 1) ? V + (size_t)vs * (j - 2) : nullptr;
    hipLaunchKernelGGL((k_bt_fold<KP, 0>), dim3(ninst), dim3(64), 0, c->stream, B->d_gchunk, K, j, P1, in, bctl);
    // w_j goes to basis slot j; the last step's w only feeds its norm
    double *w_out = (j + 1 <= itmax - 1) ? V + (size_t)vs * j : u;
    hipLaunchKernelGGL((k_bt_update<KP>), dim3(NC), dim3(PSFUN_BS), 0, c->stream, B->d_chunk, B->d_gchunk, K, j, u,
                       w_out, vj, vjm1, P1, P0, in, bctl);
    // the estimate of index j (forked in the previous step) is joined here,
    // before the beta step that reads the stop state it sets
    if (istop == 1 && j > 1) PSFUN_HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    hipLaunchKernelGGL((k_bt_beta<KP>), dim3(ninst), dim3(64), 0, c->stream, B->d_gchunk, K, j + 1, P0, in, bctl,
                       itmax, istop);
    if (istop == 1) {  // the estimate of index j+1 beside the SpMM of that index
      PSFUN_HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
      PSFUN_HIP_TRY(hipStreamWaitEvent(c->aux_stream, c->ev_fork, 0));
      if (itmax <= 65)
        hipLaunchKernelGGL(k_bt_est<0>, dim3(ninst), dim3(64), 0, c->aux_stream, j + 1, in, bctl, itmax, eps,
                           fscale);
      else if (itmax <= 129)
        hipLaunchKernelGGL(k_bt_est<1>, dim3(ninst), dim3(64), 0, c->aux_stream, j + 1, in, bctl, itmax, eps,
                           fscale);
      else
        hipLaunchKernelGGL(k_bt_est<2>, dim3(ninst), dim3(64), 0, c->aux_stream, j + 1, in, bctl, itmax, eps,
                           fscale);
      PSFUN_HIP_TRY(hipEventRecord(c->ev_join, c->aux_stream));
    }
    if (j + 1 <= itmax - 1) {
      PSFUN_TRY(psfun_prof_begin(c, PSFUN_PROF_BATCH, spmm_bytes));
      bt_spmm_launch<KP>(c, B, opscale, K, j + 1, V + (size_t)vs * j, nullptr, (j & 1) ? U1 : U0, P0, P1, in, bctl);
      PSFUN_TRY(psfun_prof_end(c));
    }
    PSFUN_HIP_TRY(hipGetLastError());
    if (poll_host && j % 2 == 0) {
      const int slot = (j / 2) & 1;
      PSFUN_HIP_TRY(hipMemcpyAsync((void *)(hflag + slot), &bctl->remaining, sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
      PSFUN_HIP_TRY(hipEventRecord(c->ev[slot], c->stream));
      if (pending >= 0) {
        PSFUN_HIP_TRY(hipEventSynchronize(c->ev[pending]));
        if (hflag[pending] == 0) break;
      }
      pending = slot;
    }
  }
  if (istop == 1 && itmax >= 2) PSFUN_HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));  // the last estimate
  hipLaunchKernelGGL(k_bt_final, dim3(ninst), dim3(64), 0, c->stream, in, fscale);
  hipLaunchKernelGGL((k_bt_assemble<KP>), dim3(NC), dim3(PSFUN_BS), 0, c->stream, B->d_chunk, K, V, vs, in,
                     (const int *)B->d_perm, Y);
  PSFUN_HIP_TRY(hipGetLastError());
  return PSFUN_SUCCESS;
}

__global__ void k_bt_export(BtInst in, int ninst, double *T_alpha, double *T_beta, int *nbasis, double *rs1) {
  const int64_t tot = (int64_t)ninst * in.ld;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    T_alpha[i] = in.alpha[i];
    T_beta[i] = in.beta[i];
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) {
    nbasis[i] = (in.done[i] == BT_ZERO) ? 0 : in.itx[i];
    rs1[i] = in.rs1[i];
  }
}

static int bt_lanczos(psfun_ctx_t c, psfun_batch_t B, int K, double fscale, double opscale, const double *X, double *Y,
                      double eps, int itmax, int istop, int *iters, double *errs, double *V_ext, double *T_alpha,
                      double *T_beta, int *nbasis, double *rs1_out);

extern "C" int psfun_batch_lanczos(psfun_ctx_t c, psfun_batch_t B, int K, double fscale, double opscale,
                                   const double *X, double *Y, double eps, int itmax, int istop, int *iters,
                                   double *errs) {
  return bt_lanczos(c, B, K, fscale, opscale, X, Y, eps, itmax, istop, iters, errs, nullptr, nullptr, nullptr,
                    nullptr, nullptr);
}

extern "C" int psfun_batch_lanczos_basis(psfun_ctx_t c, psfun_batch_t B, int K, double fscale, double opscale,
                                         const double *X, double *Y, double eps, int itmax, int istop, double *V,
                                         double *T_alpha, double *T_beta, int *nbasis, double *rs1) {
  if (!V || !T_alpha || !T_beta || !nbasis || !rs1) return PSFUN_ERR_INVALID_ARG;
  return bt_lanczos(c, B, K, fscale, opscale, X, Y, eps, itmax, istop, nullptr, nullptr, V, T_alpha, T_beta, nbasis,
                    rs1);
}

static int bt_lanczos(psfun_ctx_t c, psfun_batch_t B, int K, double fscale, double opscale, const double *X, double *Y,
                      double eps, int itmax, int istop, int *iters, double *errs, double *V_ext, double *T_alpha,
                      double *T_beta, int *nbasis, double *rs1_out) {
  if (!c || !B || K < 1 || K > BT_KMAX) return PSFUN_ERR_INVALID_ARG;
  if (istop < 1 || istop > 2) return PSFUN_ERR_INVALID_ISTOP;
  if (!X || !Y) return PSFUN_ERR_INVALID_VECT_STATE;
  if (itmax < 2 || itmax > PSFUN_FT_KMAX) return PSFUN_ERR_INVALID_ARG;
  if (B->nrows == 0 || B->ngraph == 0) return PSFUN_SUCCESS;
  // an exported basis would be in the batch's own row numbering
  if (V_ext && B->d_perm) return PSFUN_ERR_UNSUPPORTED;
  PSFUN_HIP_TRY(hipSetDevice(c->device));
  const int ninst = B->ngraph * K;
  const int ld = itmax + 2;
  const int64_t vs = B->nrows * K;
  // workspace: control, per-instance arrays, basis v_1..v_{itmax-1}, two u buffers
  const size_t inst_bytes = align_up(sizeof(double) * ((size_t)ninst * ld * 5 + (size_t)ninst * 3), 256) +
                            align_up(sizeof(int) * (size_t)ninst * 3, 256);
  // the basis lives in the caller's buffer when one is given
  // (a reordered batch also keeps X in its own numbering)
  const size_t bytes = 256 + inst_bytes + sizeof(double) * (size_t)vs * ((V_ext ? 2 : itmax + 1) + (B->d_perm ? 1 : 0));
  void *ws = nullptr;
  PSFUN_TRY(psfun_workspace(c, bytes, &ws));
  BtCtl *bctl = (BtCtl *)ws;
  BtInst in;
  double *d = (double *)((char *)ws + 256);
  in.ld = ld;
  in.alpha = d;
  in.beta = d + (size_t)ninst * ld;
  in.rs = d + 2 * (size_t)ninst * ld;
  in.yk = d + 3 * (size_t)ninst * ld;
  in.ib = d + 4 * (size_t)ninst * ld;
  in.rs1 = d + 5 * (size_t)ninst * ld;
  in.errnum = in.rs1 + ninst;
  in.errden = in.errnum + ninst;
  int *ip = (int *)((char *)d + align_up(sizeof(double) * ((size_t)ninst * ld * 5 + (size_t)ninst * 3), 256));
  in.done = ip;
  in.itx = ip + ninst;
  in.m = ip + 2 * ninst;
  double *V, *U0;
  if (V_ext) {
    V = V_ext;
    U0 = (double *)((char *)ws + 256 + inst_bytes);
  } else {
    V = (double *)((char *)ws + 256 + inst_bytes);
    U0 = V + (size_t)vs * (itmax - 1);
  }
  double *U1 = U0 + vs;
  double *Xp = B->d_perm ? U1 + vs : nullptr;

  int KP = 1;
  while (KP < K) KP <<= 1;
  // hipGraph path: captured and replayed on the context's own stream, joined
  // to the caller's stream by events (the null stream cannot be captured).
  const bool graph = B->use_graph && !c->prof_on && !V_ext;
  hipStream_t caller = c->stream;
  if (graph) {
    if (!B->ev_in) PSFUN_HIP_TRY(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
    if (!B->ev_out) PSFUN_HIP_TRY(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
    PSFUN_HIP_TRY(hipEventRecord(B->ev_in, caller));
    PSFUN_HIP_TRY(hipStreamWaitEvent(c->own_stream, B->ev_in, 0));
    c->stream = c->own_stream;
  }
  hipGraphExec_t exec = nullptr;
  auto key = std::make_tuple((const void *)X, (void *)Y, (const void *)ws, K, itmax, istop, eps, fscale, opscale);
  if (graph) {
    auto it = B->graphs.find(key);
    if (it != B->graphs.end()) exec = it->second;
  }
  int st = PSFUN_SUCCESS;
  if (!exec) {
    if (graph && hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) st = PSFUN_ERR_HIP;
    if (st == PSFUN_SUCCESS) {
#define BT_CASE(KPV) \
  case KPV: st = bt_enqueue<KPV>(c, B, K, fscale, opscale, X, Y, eps, itmax, istop, in, bctl, V, U0, U1, Xp, !graph); break;
      switch (KP) {
        BT_CASE(1) BT_CASE(2) BT_CASE(4) BT_CASE(8) BT_CASE(16) BT_CASE(32) BT_CASE(64)
        default: st = PSFUN_ERR_INVALID_ARG;
      }
#undef BT_CASE
    }
    if (graph) {
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (st == PSFUN_SUCCESS && e != hipSuccess) {
        psfun_record_hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
        st = PSFUN_ERR_HIP;
      }
      if (st == PSFUN_SUCCESS) {
        const hipError_t e2 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (e2 != hipSuccess) {
          psfun_record_hip_error(e2, "hipGraphInstantiate", __FILE__, __LINE__);
          st = PSFUN_ERR_HIP;
          exec = nullptr;
        } else {
          B->graphs[key] = exec;
        }
      }
      if (g) hipGraphDestroy(g);
    }
  }
  if (st == PSFUN_SUCCESS && exec && hipGraphLaunch(exec, c->stream) != hipSuccess) st = PSFUN_ERR_HIP;
  if (graph) {
    c->stream = caller;
    if (hipEventRecord(B->ev_out, c->own_stream) != hipSuccess || hipStreamWaitEvent(caller, B->ev_out, 0) != hipSuccess)
      st = (st == PSFUN_SUCCESS) ? PSFUN_ERR_HIP : st;
  }
  if (st != PSFUN_SUCCESS) return st;
  if (V_ext) {
    hipLaunchKernelGGL(k_bt_export, dim3(std::min(1024, (ninst * ld + 255) / 256 + 1)), dim3(256), 0, c->stream, in,
                       ninst, T_alpha, T_beta, nbasis, rs1_out);
    PSFUN_HIP_TRY(hipGetLastError());
  }
  // per-instance results
  PSFUN_HIP_TRY(hipStreamSynchronize(c->stream));
  if (iters || errs) {
    std::vector<int> itv(ninst);
    std::vector<double> en(ninst), ed(ninst);
    PSFUN_HIP_TRY(hipMemcpy(itv.data(), in.itx, sizeof(int) * ninst, hipMemcpyDeviceToHost));
    PSFUN_HIP_TRY(hipMemcpy(en.data(), in.errnum, sizeof(double) * ninst, hipMemcpyDeviceToHost));
    PSFUN_HIP_TRY(hipMemcpy(ed.data(), in.errden, sizeof(double) * ninst, hipMemcpyDeviceToHost));
    for (int i = 0; i < ninst; ++i) {
      if (iters) iters[i] = itv[i];
      if (errs) errs[i] = ed[i] == 0.0 ? en[i] : en[i] / ed[i];
    }
  }
  return PSFUN_SUCCESS;
}

// ---------------------------------------------------------------------------
// Row-major SpMM on the batch and value replacement (used by the adjoint).
template <int KP>
__global__ __launch_bounds__(PSFUN_BS) void k_bt_spmm_plain(const int *__restrict__ rowptr, const int *__restrict__ col,
                                                            const double *__restrict__ val,
                                                            const int *__restrict__ chunk, int K, double alpha,
                                                            const double *__restrict__ X, double beta,
                                                            double *__restrict__ Y) {
  constexpr int RP = PSFUN_BS / KP;
  const int c = blockIdx.x;
  const int r0 = chunk[3 * c], r1 = chunk[3 * c + 1];
  const int k = threadIdx.x % KP, sub = threadIdx.x / KP;
  if (k >= K) return;
  for (int r = r0 + sub; r < r1; r += RP) {
    double sum = 0.0;
    for (int q = rowptr[r]; q < rowptr[r + 1]; ++q) sum = fma(val[q], X[(int64_t)col[q] * K + k], sum);
    const int64_t i = (int64_t)r * K + k;
    double out = alpha * sum;
    if (beta != 0.0) out = fma(beta, Y[i], out);
    Y[i] = out;
  }
}

extern "C" int psfun_batch_spmm(psfun_ctx_t c, psfun_batch_t B, int K, double alpha, const double *X, double beta,
                                double *Y) {
  if (!c || !B || K < 1 || K > BT_KMAX || (B->nrows && (!X || !Y))) return PSFUN_ERR_INVALID_ARG;
  if (B->nrows == 0 || B->nchunks == 0) return PSFUN_SUCCESS;
  int KP = 1;
  while (KP < K) KP <<= 1;
  // reordered batch: X into its numbering, the product there, Y scattered back
  const int64_t vs = B->nrows * K;
  const double *Xs = X;
  double *Ys = Y;
  double beta_s = beta;
  if (B->d_perm) {
    void *ws = nullptr;
    PSFUN_TRY(psfun_workspace(c, sizeof(double) * 2 * (size_t)vs, &ws));
    double *Xp = (double *)ws, *Yp = Xp + vs;
    hipLaunchKernelGGL(k_bt_gather_rows, dim3(psfun_grid(c, (vs + PSFUN_BS - 1) / PSFUN_BS, 8)), dim3(PSFUN_BS), 0,
                       c->stream, (int64_t)B->nrows, K, (const int *)B->d_perm, X, Xp);
    Xs = Xp;
    Ys = Yp;
    beta_s = 0.0;
  }
#define SP_CASE(KPV)                                                                                       \
  case KPV:                                                                                                \
    hipLaunchKernelGGL((k_bt_spmm_plain<KPV>), dim3(B->nchunks), dim3(PSFUN_BS), 0, c->stream, B->rowptr, \
                       B->colind, B->val, B->d_chunk, K, alpha, Xs, beta_s, Ys);                          \
    break;
  switch (KP) {
    SP_CASE(1) SP_CASE(2) SP_CASE(4) SP_CASE(8) SP_CASE(16) SP_CASE(32) SP_CASE(64)
    default: return PSFUN_ERR_INVALID_ARG;
  }
#undef SP_CASE
  PSFUN_HIP_TRY(hipGetLastError());
  if (B->d_perm) {
    hipLaunchKernelGGL(k_bt_scatter_rows, dim3(psfun_grid(c, (vs + PSFUN_BS - 1) / PSFUN_BS, 8)), dim3(PSFUN_BS), 0,
                       c->stream, (int64_t)B->nrows, K, (const int *)B->d_perm, (const double *)Ys, beta, Y);
    PSFUN_HIP_TRY(hipGetLastError());
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_batch_set_values(psfun_batch_t B, const double *val) {
  if (!B || (B->nnz && !val)) return PSFUN_ERR_INVALID_ARG;
  if (!B->nnz) return PSFUN_SUCCESS;
  if (B->d_nzperm) {
    psfun_ctx *c = B->ctx;
    hipLaunchKernelGGL(k_bt_gather_vals, dim3(psfun_grid(c, (B->nnz + PSFUN_BS - 1) / PSFUN_BS, 8)), dim3(PSFUN_BS), 0,
                       c->stream, (int64_t)B->nnz, (const int *)B->d_nzperm, val, B->val);
    PSFUN_HIP_TRY(hipGetLastError());
  } else {
    PSFUN_HIP_TRY(hipMemcpyAsync(B->val, val, sizeof(double) * B->nnz, hipMemcpyDeviceToDevice, B->ctx->stream));
  }
  return PSFUN_SUCCESS;
}

extern "C" int psfun_batch_csr(psfun_batch_t B, const int32_t **rowptr, const int32_t **colind, const double **val) {
  if (!B) return PSFUN_ERR_INVALID_ARG;
  if (B->d_perm) return PSFUN_ERR_UNSUPPORTED;  // its arrays are in the batch's own numbering
  if (rowptr) *rowptr = B->rowptr;
  if (colind) *colind = B->colind;
  if (val) *val = B->val;
  return PSFUN_SUCCESS;
}
