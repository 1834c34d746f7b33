// Internal launch helpers shared between the translation units of the library.
#pragma once
#include "common.hpp"

int launch_spmv(psfun_ctx *c, const psfun_csr *A, double alpha, const double *x, double beta, double *y);
int launch_reduce(psfun_ctx *c, int op, int64_t n, const double *x, const double *y, double *d_out);
int launch_axpby(psfun_ctx *c, int64_t n, double alpha, const double *x, double beta, double *y);
// Looks up the device address of the allocation-free entry point's fixed
// scratch on `device` (current); defined beside that array in lanczos.hip.
int psfun_noalloc_setup(int device);
// Preconditioner objects (prec.hip): z = M^{-1} r on the context stream for a
// BJAC object; the kind and order of an object.
int psfun_prec_launch(psfun_ctx *c, const psfun_prec *P, const double *r, double *z);
// The same for S complex right-hand sides (shift-major, leading dimension
// ld; ybuf holds S*ld intermediates), for the systems with run[s] != 0.
int psfun_prec_launch_c(psfun_ctx *c, const psfun_prec *P, const double2 *r, double2 *z, double2 *ybuf, int64_t ld,
                        int S, const int *run);
int psfun_prec_kind(const psfun_prec *P);
int psfun_prec_size(const psfun_prec *P);
// Multi-shift CG (mscg.hip) with the solutions stored: arguments as
// psfun_mscg, already checked; iters and resid must be non-null.
int psfun_mscg_run(psfun_ctx *c, const psfun_csr *A, int nshift, const double *shifts, const double *b, double *X,
                   int64_t ldx, double eps, int itmax, int *iters, double *resid);
