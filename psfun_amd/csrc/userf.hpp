// Host side of the USERF/SYM f(H) step (serial/psfun_d_serial_mod.F90:420-485)
// shared by the serial apply and the Krylov loops: the eigen-decomposition of
// the symmetric H runs on the device (one block, fsym.hpp), the user function
// on the host (it is a host procedure, psfun_d_serial%f_ptr), and the
// combination Q diag(f) Q^T back on the device.
#pragma once
#include "common.hpp"
#include "fsym.hpp"
#include "lanczos.hpp"

// Device scratch of one USERF step of order <= kmax.
struct UserfScratch {
  double *S = nullptr, *Q = nullptr, *w = nullptr, *fv = nullptr;
};
size_t userf_scratch_bytes(int kmax);
UserfScratch userf_scratch_carve(void *base, int kmax);

// Validates a USERF request: a function is set and the variant is SYM.
int userf_check(const psfun_ctx *c, int variant);

// S = the symmetric matrix (form, a, b, ld) of order k (device pointers);
// its eigenvalues w on the device, then fv = f(fscale*w) evaluated on the host
// and copied back.  Waits for the stream once.  With `stopped` given, a loop
// whose stop flag (host flag 1) is set by then is reported there and f is not
// called (the T or H of such a step is not the reference's).
int userf_eval(psfun_ctx *c, int k, int form, const double *a, const double *b, int ld, double fscale,
               const UserfScratch &sc, bool *stopped = nullptr);

// Krylov control at loop index j for istop = 1 (after userf_eval of order
// j-1): yk = Q diag(fv) Q^T e1, rs(j) = rs1 * sigma * |yk(j-1)| / ||yk|| with
// sigma read from *sig, then the convergence test (eps) and, at j >= nl, the
// iteration limit; stop states go to ctl->done and the host flags.
int userf_launch_est(psfun_ctx *c, LzCtl *ctl, int j, const double *sig, const UserfScratch &sc, int nl,
                     double eps);

// Assembly coefficients when the loop did not converge: yk = Q diag(fv) Q^T e1
// of order k (after userf_eval of order k), ctl->m = k.
int userf_launch_final(psfun_ctx *c, LzCtl *ctl, int k, const UserfScratch &sc);

// y = Q diag(fv) Q^T x (device vectors of length k), for the serial apply.
int userf_launch_apply(psfun_ctx *c, int k, const UserfScratch &sc, const double *x, double *y);
