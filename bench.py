"""Benchmark of the f(A)x hot path on one MI355X.

Headline (`value`): the plain GPU drop-in, y = exp(-tA)x with t = 1 by
Lanczos on the unscaled 3-D 7-point Dirichlet Laplacian of a 128^3 grid
(n = 2,097,152, nnz = 14,581,760, FP64 CSR), x ~ U[-1,1] (SplitMix64 seed),
Krylov dimension <= 50, eps = 1e-10.  One step = one f(A)x (one converged y);
value = f(A)x per second with the inputs resident in HBM.

The same JSON line also carries:
  roofline      -- dominant kernel of the headline (Lanczos SpMV + dot):
                   algorithmic bytes / HIP-event-timed launch duration
  cpu_baseline  -- the C oracle (the reference's algorithm, unfused) timed on
                   this host's cores on a bounded sample of the same workload
  batched       -- 256 random-geometric-graph Laplacians (n log-uniform in
                   [1k, 64k], mean degree ~8, weights U[0.5,1.5]) x 16 normal
                   vectors, exp(-0.5 L) X in one launch sequence, vs looping
                   the plain GPU path over the same instances
  adjoint       -- loss = <G, exp(-tA) X>, t = 0.1, on the 1024^2 anisotropic
                   operator with 32 vectors: forward + gradients w.r.t. X, t and
                   A's nonzeros, vs the same algorithm looped over the 32
                   columns one at a time (and, beside it, the plain GPU path
                   with a multi-shift CG gradient per column)
  irregular     -- Lanczos exp(A')x on a 2M-node power-law (Chung-Lu) graph
                   Laplacian, rows of 1 to ~150k entries: CSR row blocks
                   (CSRS) against the vector-CSR kernel and sliced ELL
  quadrature    -- exp(-A)x on the 96^3 Laplacian by the 16-pole Talbot rule
                   with one multi-shift CG over all poles (end-to-end error
                   <= 1e-8 against the closed form)
  blas1         -- the PSBLAS BLAS-1 primitives (dot, nrm2, amax, axpby) on
                   1 GiB vectors
  cg            -- the shifted PCG (A + 0.01 I) x = b on the 96^3 Laplacian,
                   DIAG preconditioner, eps 1e-8
  arnoldi       -- Arnoldi exp(-A)x on the headline's matrix (modified
                   Gram-Schmidt in its one-reduction form, GENPADE f(H) step)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "f(A)x/s at 1e-10 rel err: batched+adjoint vs plain GPU Lanczos; SpMV % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--grid", type=int, default=128, help="grid points per axis of the 3-D Laplacian")
    p.add_argument("--itmax", type=int, default=50)
    p.add_argument("--eps", type=float, default=1e-10)
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--prof-steps", type=int, default=3)
    p.add_argument("--copy-probe", action="store_true", help="also run the 1 GiB copy that scales PMC counters")
    p.add_argument("--prof-reps", type=int, default=4, help="back-to-back launches per timed SpMV step")
    p.add_argument("--no-extra", action="store_true", help="headline only (no blas1/batched/adjoint/quadrature)")
    p.add_argument("--sections", default="blas1,batched,adjoint,irregular,quadrature,cg,arnoldi",
                   help="comma-separated extra sections to run after the headline")
    p.add_argument("--format", default="", help="SELL, HDIA, CSRS, L1, L2, L4 or L8 (CSR lanes per row); default: autotune")
    p.add_argument("--bt-chunks", default="64,128,256,512,1024",
                   help="rows-per-chunk candidates of the batched autotuner")
    return p.parse_args()


def pmc_traffic(kernel, workload_file):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a 1 GiB copy of known size), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", workload_file)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch " + kernel + " "):
                    return float(line.split()[-1])
    except OSError:
        pass
    return None


def best_format(M, run, candidates=("SELL", "HDIA")):
    """Time one call of `run` on each storage format the matrix accepts and
    keep the fastest; returns (name, {format: ms})."""
    import torch
    from psfun_amd import _lib
    tune = {}
    for f in candidates:
        try:
            M.set_format(f)
        except _lib.PsfunError:
            continue
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        tune[f] = round((time.perf_counter() - t0) * 1e3, 3)
    best = min(tune, key=tune.get)
    M.set_format(best)
    return best, tune


def read_prof(L, _lib, ctx, name):
    import ctypes as C
    ms, cnt, by = C.c_double(), C.c_int64(), C.c_double()
    L.psfun_prof_read(ctx.handle, _lib.PROF_IDS[name], C.byref(ms), C.byref(cnt), C.byref(by))
    return dict(ms=ms.value, launches=cnt.value, bytes=by.value)


def headline(args):
    import torch
    from psfun_amd import _lib, matgen, psb
    from psfun_amd.krylov import psfun_d_krylov
    from psfun_amd.serial import psfun_d_serial

    m = args.grid
    dims = (m, m, m)
    n = m ** 3
    t = 1.0
    A = matgen.laplacian(dims, scale=-t)   # the reference Lanczos ignores fun%scaling: pass -tA
    nnz = len(A[1])
    x = matgen.uniform_pm1(20240601, n)
    ctx = psb.default_context(0)
    M = psb.psb_dspmat_type.from_csr(A, ctx)
    desc = psb.psb_desc_type(n, ctx)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    km = psfun_d_krylov()
    km.set("KNAME", "LANCZOS")
    xv = psb.psb_d_vect_type(torch.as_tensor(x, device="cuda:0"))
    yv = psb.psb_d_vect_type(torch.zeros(n, dtype=torch.float64, device="cuda:0"))

    def step():
        info, it, err, res = km.apply(fun, M, desc, yv, xv, args.eps, itmax=args.itmax, istop=1)
        if info != 0:
            raise RuntimeError(f"lanczos apply failed: info={info}")
        return it, err

    # storage autotune for this matrix: sliced ELL, or vector CSR with 1..8
    # lanes per row (timed separately, not in `value`)
    def use(choice):
        if choice in ("SELL", "HDIA", "CSRS"):
            M.set_format(choice)
        else:
            M.set_format("CSR")
            M.set_lanes(int(choice[1:]))

    tune = {}
    if args.format:
        use(args.format)
    else:
        for choice in ("SELL", "HDIA", "CSRS", "L1", "L2", "L4", "L8"):
            try:
                use(choice)
            except _lib.PsfunError:   # diagonal slices refused (too many diagonals per slice)
                continue
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                step()
            torch.cuda.synchronize()
            tune[choice] = (time.perf_counter() - t0) / 3
        use(min(tune, key=tune.get))
    fmt = M.get_format()[0]
    lanes = M.lanes() if fmt == "CSR" else None
    log(f"headline: n={n} nnz={nnz} format={fmt} lanes={lanes} "
        f"tune(ms)={ {k: round(v * 1e3, 3) for k, v in tune.items()} }")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it, err = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    y = yv.get_vect()
    yex = matgen.dst_expm_laplacian(dims, x, t)
    rel_err = float(np.linalg.norm(y - yex) / np.linalg.norm(yex))
    log(f"headline: {args.steps} f(A)x in {elapsed:.4f}s, iter={it}, rel_err={rel_err:.3e}")

    L = _lib.lib()
    L.psfun_prof_reset(ctx.handle)
    # the fused SpMV step is re-launched prof_reps times inside each event pair
    # (idempotent), so the events' own cost is spread over prof_reps launches
    L.psfun_prof_enable(ctx.handle, args.prof_reps)
    for _ in range(args.prof_steps):
        step()
    torch.cuda.synchronize()
    L.psfun_prof_enable(ctx.handle, 0)
    k = read_prof(L, _lib, ctx, "lz_spmv")
    ku = read_prof(L, _lib, ctx, "lz_update")
    # (launches that found the loop already stopped did no work and are not
    # counted by the library)
    per_launch_bytes = k["bytes"] / max(k["launches"], 1)
    avg_us = 1e3 * k["ms"] / max(k["launches"], 1)
    idx_bits = M.sell_index_bits() if fmt == "SELL" else (0 if fmt == "HDIA" else 32)
    n_, nnz_ = M.get_nrows(), M.get_nzeros()
    # bytes the kernel itself moves: sliced ELL does not read rowptr (its
    # padding, 0.2% here, is left out); diagonal slices read 64 values per
    # slice diagonal, one offset per slice diagonal and the slice pointers,
    # and no column indices
    if fmt == "SELL":
        kernel_bytes = per_launch_bytes - (4.0 - idx_bits / 8.0) * nnz_ - 4.0 * (n_ + 1)
    elif fmt == "HDIA":
        nd = M.hdia_diagonals()
        kernel_bytes = (per_launch_bytes - 12.0 * nnz_ - 4.0 * (n_ + 1)
                        + (8.0 * 64 + 4.0) * nd + 4.0 * ((n_ + 63) // 64 + 1))
    else:
        kernel_bytes = per_launch_bytes
    # `achieved`/`frac` count the bytes the timed kernel moves; the CSR-equivalent
    # count of the same step (int32 column indices and rowptr: 12*nnz + 4(n+1) + 16n)
    # is reported beside it under csr_equiv_*
    achieved = kernel_bytes / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0
    csr_equiv = per_launch_bytes / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0
    upd_us = 1e3 * ku["ms"] / max(ku["launches"], 1)
    upd_bytes = ku["bytes"] / max(ku["launches"], 1)
    kname = {"SELL": "k_lz_spmv_sell", "HDIA": "k_lz_spmv_hdia", "CSRS": "k_lz_spmv_csrs"}.get(
        fmt, f"k_lz_spmv<{lanes}>")
    per_unit = {"SELL": f"{idx_bits // 8 + 8}*nnz (values + {idx_bits}-bit column offsets) + 16n",
                "HDIA": "(8*64 + 4)*slice_diagonals + 4*(slices + 1) + 16n (no column indices)"}.get(
                    fmt, "12*nnz + 4(n+1) + 16n")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": (pmc_traffic(kname, f"lz_spmv_{fmt.lower()}_{m}cube_pmc.txt")
                            if fmt in ("SELL", "HDIA") else None),
                "kernel": kname + " (Lanczos SpMV step, alpha partials fused)",
                "bytes_per_launch": kernel_bytes, "bytes_per_unit": per_unit,
                "avg_launch_us": round(avg_us, 3), "launches": k["launches"],
                "launches_per_event_pair": args.prof_reps, "index_bits": idx_bits,
                "slice_diagonals": (M.hdia_diagonals() if fmt == "HDIA" else None),
                "csr_equiv_bytes_per_launch": per_launch_bytes,
                "csr_equiv_GBs": round(csr_equiv, 1),
                "csr_equiv_frac": round(csr_equiv / HBM_PEAK_GBS, 4),
                "note": ("achieved/frac: bytes the kernel moves per launch over its HIP-event-timed launch "
                         "duration (the first step of each apply also copies x: +8n, included); csr_equiv_*: "
                         "the same step charged the CSR count 12*nnz + 4(n+1) + 16n, which exceeds what a "
                         "format without column indices reads"),
                "update_kernel": {"name": "k_lz_update (3-term update + ||w||^2 partials, 32n bytes)",
                                  "avg_launch_us": round(upd_us, 3),
                                  "achieved_GBs": round(upd_bytes / (upd_us * 1e-6) / 1e9, 1) if upd_us else None}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, A, x, y, m)
    cfg = {"workload": f"lanczos exp(-tA)x, 3-D 7-point Dirichlet Laplacian {m}^3", "n": n, "nnz": nnz, "t": t,
           "itmax": args.itmax, "eps": args.eps, "format": fmt, "lanes_per_row": lanes, "global_batch": 1,
           "parallelism": "single-gpu"}
    return value, elapsed, it, rel_err, roofline, cpu, cfg


def cpu_baseline(args, A, x, y, m):
    """The oracle (the reference's Lanczos restated in C: unfused PSBLAS-style
    passes, OpenMP over rows, f(T) every iteration) on this host's cores: one
    warm-up run, then the median of at least 5 timed runs, on every core the
    process may use (bounded by OMP_NUM_THREADS when that is set) and on 1 core."""
    import statistics
    from oracle import oracle as O
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    threads = max(1, min(threads, avail))

    def timed(nthreads, min_runs, budget_s):
        O.set_threads(nthreads)
        ref = O.lanczos(A, x, args.eps, args.itmax, 1)           # warm-up
        times = []
        t_all = time.perf_counter()
        while len(times) < min_runs or (time.perf_counter() - t_all < budget_s and len(times) < 101):
            t0 = time.perf_counter()
            ref = O.lanczos(A, x, args.eps, args.itmax, 1)
            times.append(time.perf_counter() - t0)
        return ref, times

    ref, tm = timed(threads, 5, args.cpu_seconds * 2 / 3)
    _, t1 = timed(1, 5, args.cpu_seconds / 3)
    O.set_threads(threads)
    med, med1 = statistics.median(tm), statistics.median(t1)
    out = {"value": round(1.0 / med, 4), "unit": "f(A)x/s", "cores": threads, "kind": "port",
           "nproc": nproc, "cpus_available": avail, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
           "protocol": "1 warm-up, then the median of the timed runs (wall clock around the apply)",
           "runs": len(tm), "run_ms_min_median_max": [round(1e3 * min(tm), 2), round(1e3 * med, 2),
                                                      round(1e3 * max(tm), 2)],
           "value_1core": round(1.0 / med1, 4), "runs_1core": len(t1),
           "sample": f"{len(tm)} full f(A)x of the same {m}^3 workload after one warm-up "
                     f"(oracle/psfun_oracle.c: unfused PSBLAS-style passes, OpenMP over rows, f(T) by tridiagonal "
                     f"QL every iteration)",
           "gpu_vs_cpu_rel_diff": float(np.linalg.norm(y - ref["y"]) / np.linalg.norm(ref["y"])),
           "cpu_iter": ref["iter"]}
    log(f"cpu baseline: {out['value']} f(A)x/s on {threads} threads (nproc {nproc}, available {avail}), "
        f"{out['value_1core']} on 1 core")
    return out


def batched_section(args):
    import torch
    from psfun_amd import matgen, psb
    from psfun_amd.batched import psfun_batch
    from psfun_amd.krylov import lanczos_apply
    from psfun_amd.serial import psfun_d_serial

    t, K, eps, itmax = 0.5, 16, 1e-10, 60
    sizes = matgen.rgg_sizes(256, 3)
    tg = time.perf_counter()
    mats = [matgen.rgg_laplacian(int(nv), 1000 + i) for i, nv in enumerate(sizes)]
    log(f"batched: generated 256 graphs (sum n = {int(sizes.sum())}) in {time.perf_counter() - tg:.1f}s")
    # the batch in the graphs' own numbering and in reverse Cuthill-McKee order
    # (the generator numbers nodes in random spatial order); the faster counts
    timed = {}
    for order in ("natural", "rcm"):
        B = psfun_batch(mats, order=order)
        N = B.nrows
        X = torch.as_tensor(matgen.normal(77, N * K).reshape(N, K), device="cuda:0")
        cands = tuple(int(v) for v in args.bt_chunks.split(","))
        best, table = B.autotune(X, t=t, eps=eps, itmax=itmax, candidates=cands, reps=2)
        Y, iters, errs = B.expm(X, t=t, eps=eps, itmax=itmax)
        torch.cuda.synchronize()
        steps = 3
        t0 = time.perf_counter()
        for _ in range(steps):
            Y, _, _ = B.expm(X, t=t, eps=eps, itmax=itmax)
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t0) / steps
        timed[order] = (tb, B, best, Y, iters)
        log(f"batched: order {order}: tuning (rows_per_chunk, hipgraph) -> ms: "
            + ", ".join(f"{c[0]}/{int(c[1])}: {dt * 1e3:.2f}" for c, dt in table))
        log(f"batched: order {order}: {B.ngraph * K / tb:.1f} f(A)x/s")
    order = min(timed, key=lambda o: timed[o][0])
    tb, B, best, Y, iters = timed[order]
    rate = B.ngraph * K / tb
    # plain GPU path looped over every graph, one column each (the per-column
    # cost is the same for the other columns of the graph)
    ctx = psb.default_context(0)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    off = np.concatenate([[0], np.cumsum(sizes)])
    handles = [psb.psb_dspmat_type.from_csr(A, ctx) for A in mats]
    Xh = X.cpu().numpy()
    xs = [psb.psb_d_vect_type(torch.as_tensor(np.ascontiguousarray(Xh[off[g]:off[g + 1], 0]), device="cuda:0"))
          for g in range(len(mats))]
    ys = [psb.psb_d_vect_type(torch.zeros(int(sizes[g]), dtype=torch.float64, device="cuda:0"))
          for g in range(len(mats))]
    for g in range(4):  # warm
        lanczos_apply(fun, handles[g], None, ys[g], xs[g], eps, itmax=itmax, opscale=-t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in range(len(mats)):
        info, it, err, res = lanczos_apply(fun, handles[g], None, ys[g], xs[g], eps, itmax=itmax, opscale=-t)
    torch.cuda.synchronize()
    tl = time.perf_counter() - t0
    loop_rate = len(mats) / tl
    Yh = Y.cpu().numpy()
    maxrel = max(float(np.linalg.norm(Yh[off[g]:off[g + 1], 0] - ys[g].get_vect()) / np.linalg.norm(ys[g].get_vect()))
                 for g in range(len(mats)))
    log(f"batched: {rate:.1f} f(A)x/s vs loop {loop_rate:.1f} f(A)x/s, max rel diff {maxrel:.2e}")
    return {"workload": "256 RGG Laplacians (n in [1k,64k] log-uniform, degree ~8) x 16 vectors, exp(-0.5 L)X",
            "instances": int(B.ngraph * K), "rows": int(N), "nnz": int(B.nnz), "value": round(rate, 2),
            "unit": "f(A)x/s", "ms_per_batch": round(tb * 1e3, 3), "loop_plain_value": round(loop_rate, 2),
            "speedup_vs_loop": round(rate / loop_rate, 2), "max_rel_diff_vs_plain": maxrel,
            "tuned": {"rows_per_chunk": best[0], "hipgraph": bool(best[1]), "order": order},
            "value_by_order": {o: round(B.ngraph * K / v[0], 2) for o, v in timed.items()},
            "iters_min_max": [int(iters.min()), int(iters.max())]}


def adjoint_section(args):
    import ctypes as C
    import torch
    from psfun_amd import _lib, matgen, psb
    from psfun_amd.autograd import ExpmvOperator, expmv
    from psfun_amd.krylov import lanczos_apply
    from psfun_amd.serial import psfun_d_serial

    m, K, tt = 1024, 32, 0.1
    A = matgen.anisotropic_2d(m, seed=5)
    n = m * m
    dev = "cuda:0"
    torch.cuda.empty_cache()
    # the forward keeps its Krylov basis for the backward (2 x 51 GB here)
    op = ExpmvOperator(A[0], A[1], eps=1e-10, itmax=200, quad_poles=16, store_basis=True)
    vals = torch.tensor(A[2], device=dev, requires_grad=True)
    X = torch.tensor(matgen.normal(21, n * K).reshape(n, K), device=dev, requires_grad=True)
    t = torch.tensor(tt, dtype=torch.float64, device=dev, requires_grad=True)
    G = torch.tensor(matgen.normal(22, n * K).reshape(n, K), device=dev)

    def full_step():
        vals.grad = X.grad = t.grad = None
        Y = expmv(vals, X, t, op)
        (G * Y).sum().backward()
        return Y

    full_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Y = full_step()
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    iters = np.asarray(op.last_iters)
    rate = K / tb
    # like for like: the same operator and gradient algorithm (Galerkin
    # resolvents from the stored basis), one column at a time, all K columns
    def col_step(c):
        Yc = expmv(vals, X[:, c:c + 1].contiguous(), t, op)
        (G[:, c:c + 1] * Yc).sum().backward()

    vals.grad = X.grad = t.grad = None
    col_step(0)
    torch.cuda.synchronize()
    vals.grad = X.grad = t.grad = None
    t0 = time.perf_counter()
    for c in range(K):
        col_step(c)
    torch.cuda.synchronize()
    tcols = time.perf_counter() - t0
    loop_same_rate = K / tcols
    log(f"adjoint: loop of {K} single-column steps (same algorithm): {loop_same_rate:.2f}/s")
    # plain GPU path looped over columns: forward and dL/dX by plain Lanczos,
    # dL/dA by the per-column multi-shift CG + SDDMM (dL/dt reuses Y)
    ctx = psb.default_context(0)
    Mh = psb.psb_dspmat_type.from_csr(A, ctx)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    op2 = ExpmvOperator(A[0], A[1], eps=1e-10, itmax=200, quad_poles=16, grad_mode="mscg")
    op2.set_values(vals.detach())
    ncol = 2
    Xd = X.detach()
    xc = [psb.psb_d_vect_type(Xd[:, c].contiguous()) for c in range(ncol)]
    gc = [psb.psb_d_vect_type(G[:, c].contiguous()) for c in range(ncol)]
    yc = psb.psb_d_vect_type(torch.zeros(n, dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(ncol):
        lanczos_apply(fun, Mh, None, yc, xc[c], 1e-10, itmax=200, opscale=-tt)
        lanczos_apply(fun, Mh, None, yc, gc[c], 1e-10, itmax=200, opscale=-tt)
        op2.grad_values_mscg(Xd[:, c:c + 1].contiguous(), G[:, c:c + 1].contiguous(), tt)
    torch.cuda.synchronize()
    tl = (time.perf_counter() - t0) / ncol
    loop_rate = 1.0 / tl
    g2 = op2.grad_values_mscg(Xd[:, :ncol].contiguous(), G[:, :ncol].contiguous(), tt)
    # the batched gradient of the first two columns alone, for the cross-check
    vals.grad = X.grad = t.grad = None
    Y2 = expmv(vals, X[:, :ncol].contiguous(), t, op)
    (G[:, :ncol] * Y2).sum().backward()
    grel = float(torch.linalg.norm(vals.grad - g2) / torch.linalg.norm(g2))
    log(f"adjoint: {rate:.2f} (f(A)x+grad)/s vs loop {loop_rate:.2f}/s, grad rel diff krylov vs mscg {grel:.2e}")
    return {"workload": "loss=<G,exp(-tA)X>, t=0.1, 1024^2 anisotropic 5-point (log-normal sigma=1, x-edges x10), "
                        "32 vectors; forward + dL/dX, dL/dt, dL/dA(nnz)",
            "n": n, "nnz": int(len(A[1])), "vectors": K, "value": round(rate, 3), "unit": "(f(A)x + gradient)/s",
            "ms_per_step": round(tb * 1e3, 2),
            "loop_same_algorithm_value": round(loop_same_rate, 3),
            "speedup_vs_loop": round(rate / loop_same_rate, 2),
            "loop_same_algorithm": f"{K} single-column forward+backward steps of the same operator, all timed",
            "loop_plain_value": round(loop_rate, 3),
            "speedup_vs_plain_loop": round(rate / loop_rate, 2),
            "loop_plain": "per column: plain Lanczos forward and dL/dX, multi-shift CG + SDDMM for dL/dA "
                          f"(timed on {ncol} columns)",
            "grad_rel_diff_krylov_vs_mscg": grel,
            "lanczos_iters_min_max": [int(iters.min()), int(iters.max())], "quad_poles": 16}


def _lanczos_by_format(args, G, formats, tag):
    """Time the plain Lanczos exp(A)x on the CSR matrix G in each storage of
    `formats` ('L4'/'L8': vector CSR with that many lanes per row); returns
    the per-format figures and the results."""
    import torch
    from psfun_amd import _lib, matgen, psb
    from psfun_amd.krylov import psfun_d_krylov
    from psfun_amd.serial import psfun_d_serial

    n = len(G[0]) - 1
    ctx = psb.default_context(0)
    M = psb.psb_dspmat_type.from_csr(G, ctx)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    km = psfun_d_krylov()
    km.set("KNAME", "LANCZOS")
    x = matgen.uniform_pm1(5, n)
    xv = psb.psb_d_vect_type(torch.as_tensor(x, device="cuda:0"))
    yv = psb.psb_d_vect_type(torch.zeros(n, dtype=torch.float64, device="cuda:0"))
    L = _lib.lib()
    out, ys = {}, {}
    for f in formats:
        try:
            if f.startswith("L"):
                M.set_format("CSR")
                M.set_lanes(int(f[1:]))
            else:
                M.set_format(f)
        except _lib.PsfunError:
            continue
        km.apply(fun, M, None, yv, xv, args.eps, itmax=80, istop=1)
        torch.cuda.synchronize()
        L.psfun_prof_reset(ctx.handle)
        L.psfun_prof_enable(ctx.handle, 1)
        steps = 3
        t0 = time.perf_counter()
        for _ in range(steps):
            info, it, err, res = km.apply(fun, M, None, yv, xv, args.eps, itmax=80, istop=1)
        torch.cuda.synchronize()
        ta = (time.perf_counter() - t0) / steps
        k = read_prof(L, _lib, ctx, "lz_spmv")
        L.psfun_prof_enable(ctx.handle, 0)
        L.psfun_prof_reset(ctx.handle)
        us = 1e3 * k["ms"] / max(k["launches"], 1)
        by = k["bytes"] / max(k["launches"], 1)
        ys[f] = yv.get_vect()
        out[f] = {"f(A)x/s": round(1.0 / ta, 2), "ms_per_apply": round(ta * 1e3, 3), "iterations": it,
                  "spmv_us": round(us, 2), "spmv_GBs": round(by / (us * 1e-6) / 1e9, 1),
                  "spmv_frac": round(by / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"irregular: {tag}: {f}: {1.0 / ta:.2f} f(A)x/s, SpMV {us:.1f} us ({out[f]['spmv_GBs']} GB/s), iter {it}")
    ref = ys[formats[0]]
    diff = float(max(np.linalg.norm(y - ref) / np.linalg.norm(ref) for y in ys.values()))
    return out, diff


def irregular_section(args):
    """Lanczos exp(A')x on two general (non-stencil) matrices, each storage
    timed, the SpMV step charged the CSR count 12*nnz + 4(n+1) + 16n:
    - a power-law graph Laplacian (Chung-Lu, n = 2M, ~17M nonzeros, rows of 1
      to ~150k entries): the case no sliced or diagonal storage serves, its x
      gathers at random columns;
    - the 256 random-geometric-graph Laplacians of the batched section as one
      block-diagonal CSR, each graph in reverse Cuthill-McKee order (n = 3.7M,
      33M nonzeros, rows of ~9 entries): uneven rows with gather locality."""
    from psfun_amd import matgen
    from psfun_amd.batched import rcm_order

    n = 2_000_000
    tg = time.perf_counter()
    G = matgen.chung_lu_laplacian(n, 7)
    G = (G[0], G[1], (-10.0 / float(np.max(np.abs(G[2])))) * G[2])   # spectrum of the operator in [-20, 0]
    nnz = len(G[1])
    lens = np.diff(G[0])
    log(f"irregular: Chung-Lu graph n={n} nnz={nnz} row lengths {int(lens.min())}..{int(lens.max())} "
        f"(median {int(np.median(lens))}) in {time.perf_counter() - tg:.1f}s")
    out, diff = _lanczos_by_format(args, G, ("CSRS", "L4", "L8", "SELL"), "power law")
    vec_best = min((v for k_, v in out.items() if k_.startswith("L")), key=lambda v: v["ms_per_apply"])
    res = {"workload": "Lanczos exp(A')x, A' = -(10/max|a_ij|) L, L the Laplacian of a Chung-Lu power-law graph "
                       "(exponent 2.2, mean degree 8), n = 2M, eps 1e-10",
           "n": n, "nnz": nnz, "row_len_max": int(lens.max()), "by_format": out,
           "csrs_speedup_vs_vector_csr": round(vec_best["ms_per_apply"] / out["CSRS"]["ms_per_apply"], 2),
           "max_rel_diff_between_formats": diff,
           "bytes_per_unit": "12*nnz + 4(n+1) + 16n per SpMV step"}
    # the RGG batch as one CSR, each graph renumbered by RCM
    tg = time.perf_counter()
    sizes = matgen.rgg_sizes(256, 3)
    rps, cis, vas, off = [np.zeros(1, np.int64)], [], [], 0
    for i, nv in enumerate(sizes):
        rp, ci, va = matgen.rgg_laplacian(int(nv), 1000 + i)
        p = rcm_order(rp, ci)
        ip = np.empty_like(p)
        ip[p] = np.arange(len(p))
        A = matgen.to_scipy((rp, ci, va))[p][:, p].tocsr()
        A.sort_indices()
        rps.append(A.indptr[1:].astype(np.int64) + rps[-1][-1])
        cis.append(A.indices.astype(np.int64) + off)
        vas.append(A.data)
        off += int(nv)
    R = (np.concatenate(rps).astype(np.int32), np.concatenate(cis).astype(np.int32), -0.5 * np.concatenate(vas))
    lens = np.diff(R[0])
    log(f"irregular: RGG batch as one CSR n={off} nnz={len(R[1])} in {time.perf_counter() - tg:.1f}s")
    out2, diff2 = _lanczos_by_format(args, R, ("CSRS", "L4", "L8", "SELL"), "rgg")
    res["rgg_one_csr"] = {"workload": "Lanczos exp(-0.5 L)x, L the 256 RGG Laplacians of the batched section as one "
                                      "block-diagonal CSR (each graph in RCM order), eps 1e-10",
                          "n": off, "nnz": int(len(R[1])), "row_len_max": int(lens.max()), "by_format": out2,
                          "max_rel_diff_between_formats": diff2}
    return res


def quadrature_section(args):
    import torch
    from psfun_amd import _lib, matgen, psb
    from psfun_amd import quadrature as Q

    dims = (96, 96, 96)
    A = matgen.laplacian(dims)
    n = 96 ** 3
    x = matgen.uniform_pm1(4, n)
    ctx = psb.default_context(0)
    M = psb.psb_dspmat_type.from_csr(A, ctx)
    q = Q.psfun_z_quadrature()
    q.computepoles(Q.talbot, None, 16, rparams=[1.0])
    q.set(M)
    xv = psb.psb_d_vect_type(torch.as_tensor(x, device="cuda:0"))
    yv = psb.psb_d_vect_type(torch.zeros(n, dtype=torch.float64, device="cuda:0"))
    qeps = 1e-10  # solver tolerance; the 16-pole rule itself is accurate to ~2e-9 here
    fmt, tune = best_format(M, lambda: q.apply(xv, yv, eps=qeps, itmax=3000))
    steps = 3
    t0 = time.perf_counter()
    for _ in range(steps):
        info, iters, resid = q.apply(xv, yv, eps=qeps, itmax=3000)
    torch.cuda.synchronize()
    tq = (time.perf_counter() - t0) / steps
    L = _lib.lib()
    L.psfun_prof_reset(ctx.handle)
    L.psfun_prof_enable(ctx.handle, 1)
    q.apply(xv, yv, eps=qeps, itmax=3000)
    torch.cuda.synchronize()
    L.psfun_prof_enable(ctx.handle, 0)
    ku = read_prof(L, _lib, ctx, "ms_update")
    ks = read_prof(L, _lib, ctx, "ms_spmv")
    yex = matgen.dst_expm_laplacian(dims, x, 1.0)
    rel = float(np.linalg.norm(yv.get_vect() - yex) / np.linalg.norm(yex))
    log(f"quadrature: {1 / tq:.2f} f(A)x/s, rel err {rel:.2e}, seed iters {int(iters.max())}")
    return {"workload": "exp(-A)x, 3-D Laplacian 96^3, 16-pole Talbot rule, multi-shift CG over all poles "
                        "(conjugate pairs folded: 8 complex solves), shift residual tolerance 1e-10",
            "value": round(1.0 / tq, 3), "unit": "f(A)x/s", "ms_per_apply": round(tq * 1e3, 3), "rel_err_vs_closed_form": rel,
            "format": fmt, "tune_ms": tune,
            "seed_iters": int(iters.max()),
            "shift_kernel_GBs": round(ku["bytes"] / max(ku["ms"], 1e-9) / 1e6, 1),
            "seed_spmv_GBs": round(ks["bytes"] / max(ks["ms"], 1e-9) / 1e6, 1)}


def cg_section(args):
    """The shifted PCG of psfun_dcg_vect on the 96^3 Laplacian: (A + 0.01 I) x
    = b, DIAG preconditioner, istop 2, eps 1e-8, zero initial guess."""
    import torch
    from psfun_amd import _lib, matgen, psb
    from psfun_amd.shifted import psb_dprec_type, psb_krylov

    dims = (96, 96, 96)
    A = matgen.laplacian(dims)
    n = 96 ** 3
    ctx = psb.default_context(0)
    M = psb.psb_dspmat_type.from_csr(A, ctx)
    desc = psb.psb_desc_type(n, ctx)
    prec = psb_dprec_type()
    prec.init(ctx, "DIAG")
    prec.build(M, desc)
    b = psb.psb_d_vect_type(torch.as_tensor(matgen.uniform_pm1(9, n), device="cuda:0"))
    xv = psb.psb_d_vect_type(torch.zeros(n, dtype=torch.float64, device="cuda:0"))
    eta, zeta, eps = 1.0, 0.01, 1e-8

    def solve():
        xv.v.zero_()
        return psb_krylov("CG", M, prec, b, eta, zeta, xv, eps, desc, itmax=4000, istop=2)

    fmt, tune = best_format(M, solve)
    steps = 3
    t0 = time.perf_counter()
    for _ in range(steps):
        info, it, err, _ = solve()
    torch.cuda.synchronize()
    ts = (time.perf_counter() - t0) / steps
    L = _lib.lib()
    L.psfun_prof_reset(ctx.handle)
    L.psfun_prof_enable(ctx.handle, 1)
    solve()
    torch.cuda.synchronize()
    L.psfun_prof_enable(ctx.handle, 0)
    ka = read_prof(L, _lib, ctx, "cg")
    # residual check as shiftedtest.f90 does: ||b - (eta A + zeta I) x|| / ||b||
    xh = xv.get_vect()
    r = b.get_vect() - (eta * matgen.to_scipy(A) @ xh + zeta * xh)
    rel = float(np.linalg.norm(r) / np.linalg.norm(b.get_vect()))
    us_it = ts / max(it, 1) * 1e6
    log(f"cg: {1 / ts:.2f} solves/s, {it} iterations, {us_it:.1f} us/iteration, residual {rel:.2e}")
    return {"workload": "(A + 0.01 I) x = b, 3-D Laplacian 96^3, PCG with DIAG, istop 2, eps 1e-8",
            "value": round(1.0 / ts, 3), "unit": "solves/s", "ms_per_solve": round(ts * 1e3, 3),
            "iterations": int(it), "us_per_iteration": round(us_it, 2), "relative_residual": rel,
            "format": fmt, "tune_ms": tune,
            "spmv_kernel": "k_cg_a (p update + (eta A + zeta I) p + sigma partials; "
                           "12*nnz + 4(n+1) + 32n bytes: z and p_old read, p and q written)",
            "spmv_kernel_GBs": round(ka["bytes"] / max(ka["ms"], 1e-9) / 1e6, 1),
            "spmv_kernel_us": round(1e3 * ka["ms"] / max(ka["launches"], 1), 2)}


def arnoldi_section(args):
    """psfun_d_arnoldi (modified Gram-Schmidt, GENPADE-variant f(H) step) on
    the headline's 128^3 Laplacian: exp(-A)x, eps 1e-10."""
    import torch
    from psfun_amd import _lib, matgen, psb
    from psfun_amd.arnoldi import arnoldi_apply
    from psfun_amd.serial import psfun_d_serial

    m = args.grid
    dims = (m, m, m)
    n = m ** 3
    A = matgen.laplacian(dims)
    ctx = psb.default_context(0)
    M = psb.psb_dspmat_type.from_csr(A, ctx)
    x = matgen.uniform_pm1(1, n)
    xv = psb.psb_d_vect_type(torch.as_tensor(x, device="cuda:0"))
    yv = psb.psb_d_vect_type(torch.zeros(n, dtype=torch.float64, device="cuda:0"))
    fun = psfun_d_serial()
    fun.set("VARIANT", "GENPADE")
    fun.set("SCALING", -1.0)
    desc = psb.psb_desc_type(n, ctx)
    fmt, tune = best_format(M, lambda: arnoldi_apply(fun, M, desc, yv, xv, 1e-10, itmax=50))
    steps = 5
    t0 = time.perf_counter()
    for _ in range(steps):
        info, it, err, res = arnoldi_apply(fun, M, desc, yv, xv, 1e-10, itmax=50)
    torch.cuda.synchronize()
    ta = (time.perf_counter() - t0) / steps
    L = _lib.lib()
    L.psfun_prof_reset(ctx.handle)
    L.psfun_prof_enable(ctx.handle, 1)
    arnoldi_apply(fun, M, desc, yv, xv, 1e-10, itmax=50)
    torch.cuda.synchronize()
    L.psfun_prof_enable(ctx.handle, 0)
    ks = read_prof(L, _lib, ctx, "ar_spmv")
    km = read_prof(L, _lib, ctx, "ar_mgs")
    yex = matgen.dst_expm_laplacian(dims, x, 1.0)
    rel = float(np.linalg.norm(yv.get_vect() - yex) / np.linalg.norm(yex))
    log(f"arnoldi: {1 / ta:.2f} f(A)x/s, {it} iterations, rel err {rel:.2e}")
    return {"workload": f"Arnoldi exp(-A)x (SpMV scaled by -1, GENPADE f(H)), 3-D Laplacian {m}^3, eps 1e-10",
            "value": round(1.0 / ta, 3), "unit": "f(A)x/s", "ms_per_apply": round(ta * 1e3, 3),
            "iterations": int(it), "rel_err_vs_closed_form": rel, "format": fmt, "tune_ms": tune,
            "spmv_kernel_us": round(1e3 * ks["ms"] / max(ks["launches"], 1), 2),
            "spmv_kernel_GBs": round(ks["bytes"] / max(ks["ms"], 1e-9) / 1e6, 1),
            "ortho_ms_per_apply": round(km["ms"], 3),
            "ortho_columns": int(km["launches"]),
            "ortho_GBs": round(km["bytes"] / max(km["ms"], 1e-9) / 1e6, 1),
            "ortho": "one-reduction MGS per column: gram pass + fold/solve + update, (2j+3)*8n bytes"}


def blas1_section(args):
    """The PSBLAS BLAS-1 primitives behind the Krylov loops (psb_gedot,
    psb_genrm2, psb_geamax, psb_geaxpby) on 1 GiB vectors: per-kernel device
    time from HIP events on the library stream, against 8 TB/s."""
    import ctypes as C
    import torch
    from psfun_amd import _lib, psb
    ctx = psb.default_context(0)
    n = 1 << 27
    x = torch.rand(n, dtype=torch.float64, device="cuda:0")
    y = torch.rand(n, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    L = _lib.lib()
    ctx.bind_stream()
    r = C.c_double()
    xp, yp = C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())
    ops = {
        "dot": (lambda: L.psfun_dot(ctx.handle, n, xp, yp, C.byref(r)), "reduce"),
        "nrm2": (lambda: L.psfun_nrm2(ctx.handle, n, xp, C.byref(r)), "reduce"),
        "amax": (lambda: L.psfun_amax(ctx.handle, n, xp, C.byref(r)), "reduce"),
        "axpby": (lambda: L.psfun_axpby(ctx.handle, n, 0.5, xp, 0.25, yp), "axpby"),
    }
    out = {}
    for name, (fn, kid) in ops.items():
        for _ in range(2):
            _lib.check(fn())
        torch.cuda.synchronize()
        L.psfun_prof_reset(ctx.handle)
        L.psfun_prof_enable(ctx.handle, 1)
        for _ in range(5):
            _lib.check(fn())
        torch.cuda.synchronize()
        L.psfun_prof_enable(ctx.handle, 0)
        k = read_prof(L, _lib, ctx, kid)
        us = 1e3 * k["ms"] / max(k["launches"], 1)
        gbs = k["bytes"] / max(k["ms"], 1e-9) / 1e6
        out[name] = {"avg_launch_us": round(us, 2), "bytes_per_launch": k["bytes"] / max(k["launches"], 1),
                     "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    log("blas1: " + ", ".join(f"{k} {v['achieved_GBs']} GB/s" for k, v in out.items()))
    out["workload"] = "BLAS-1 over 2^27 doubles (1 GiB per vector), partial sums folded by a second launch"
    return out


def copy_probe():
    """One y = 1.0*x over 2^27 doubles (1 GiB read + 1 GiB written) through
    psfun_axpby: the known-size kernel that scales the rocprofv3 FETCH_SIZE /
    WRITE_SIZE counters into bytes."""
    import ctypes as C
    import torch
    from psfun_amd import _lib, psb
    ctx = psb.default_context(0)
    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    ctx.bind_stream()
    _lib.check(_lib.lib().psfun_axpby(ctx.handle, n, 1.0, C.c_void_p(x.data_ptr()), 0.0, C.c_void_p(y.data_ptr())))
    torch.cuda.synchronize()
    log(f"copy probe: {n} doubles, y == x: {bool(torch.equal(x, y))}")


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the f(A)x path here is single-GPU "
              "(multi-GPU row sharding is not implemented).", file=sys.stderr)
        sys.exit(2)
    value, elapsed, it, rel_err, roofline, cpu, cfg = headline(args)
    if args.copy_probe:
        copy_probe()
    extra = {}
    if not args.no_extra:
        wanted = set(args.sections.split(","))
        for name, fn in (("blas1", blas1_section), ("batched", batched_section), ("adjoint", adjoint_section),
                         ("irregular", irregular_section),
                         ("quadrature", quadrature_section), ("cg", cg_section), ("arnoldi", arnoldi_section)):
            if name not in wanted:
                continue
            try:
                extra[name] = fn(args)
            except Exception as e:  # report, do not hide, a failing section
                log(f"{name} section failed: {e!r}")
                extra[name] = {"error": repr(e)}
    out = {"metric": METRIC, "value": round(value, 3), "unit": "f(A)x/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": cfg,
           "iterations": it, "rel_err_vs_closed_form": rel_err, "roofline": roofline, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
