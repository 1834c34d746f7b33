// Device control block of one Lanczos f(A)x run (1-based Krylov indices as in
// krylov/polynomial/psfun_d_lanczos.f90).
#pragma once
#include "common.hpp"

#define LZ_NL_MAX 512

enum {
  LZ_RUNNING = 0,
  LZ_DONE_CONVERGED = 1,
  LZ_DONE_BREAKDOWN = 2,
  LZ_DONE_MAXIT = 3,
  LZ_DONE_ZERO = 4,
};

struct LzCtl {
  int done;      // LZ_* state; every kernel returns at once when non-zero
  int itx;       // reference's itx
  int m;         // number of valid coefficients in yk
  int pad0;
  double rs1;    // ||x||
  double errnum, errden;
  double pad1;
  double alpha[LZ_NL_MAX + 2];  // T(i,i), i = 1..
  double beta[LZ_NL_MAX + 2];   // T(i,i+1), i = 1..; beta[0] = 0
  double rs[LZ_NL_MAX + 2];     // rs(i), i = 1..
  double yk[LZ_NL_MAX + 2];     // yk(1..m) stored 0-based
};
