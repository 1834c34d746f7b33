! psfun_hip_mod.f90 -- Fortran (iso_c_binding) interface to the C-ABI declared in
! psfun_hip.h.  This is the module a PSFUN Fortran layer uses to reach the
! device library: psfun_d_lanczos (krylov/polynomial/psfun_d_lanczos.f90),
! psfun_d_arnoldi (krylov/polynomial/psfun_d_arnoldi.f90), psfun_dcg_vect
! (krylov/shifted/psfun_dcg.F90), the f(H) step of psfun_d_serial
! (serial/psfun_d_serial_mod.F90) and the quadrature rules
! (quadratures/psfun_z_quadrules_mod.f90, utils/psfun_*_utils_mod.f90) each
! become one call below.  Status codes map onto PSBLAS info values in the
! caller (PSFUN_SUCCESS -> psb_success_, ...).
module psfun_hip_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PSFUN_SUCCESS = 0, PSFUN_ERR_INVALID_ARG = 1, PSFUN_ERR_INVALID_ISTOP = 2, &
       & PSFUN_ERR_INVALID_VECT_STATE = 3, PSFUN_ERR_INVALID_ARGS_COMBINATION = 4, PSFUN_ERR_NOT_SQUARE = 5, &
       & PSFUN_ERR_UNSUPPORTED = 6, PSFUN_ERR_ALLOC = 7, PSFUN_ERR_HIP = 8, PSFUN_ERR_NO_DEVICE = 9, &
       & PSFUN_ERR_BREAKDOWN = 10, PSFUN_ERR_TOO_LARGE = 11
  integer(c_int), parameter :: PSFUN_FN_EXP = 0, PSFUN_FN_USERF = 1
  integer(c_int), parameter :: PSFUN_VAR_CHBSYM = 0, PSFUN_VAR_CHBHES = 1, PSFUN_VAR_CHBGEN = 2, &
       & PSFUN_VAR_GENPADE = 3, PSFUN_VAR_TAYLOR = 4, PSFUN_VAR_SASQ = 5, PSFUN_VAR_SYM = 6
  integer(c_int), parameter :: PSFUN_MEM_HOST = 0, PSFUN_MEM_DEVICE = 1
  integer(c_int), parameter :: PSFUN_FORMAT_AUTO = 0, PSFUN_FORMAT_CSR = 1, PSFUN_FORMAT_SELL = 2, &
                                PSFUN_FORMAT_HDIA = 3
  integer(c_int), parameter :: PSFUN_PREC_NONE = 0, PSFUN_PREC_DIAG = 1, PSFUN_PREC_BJAC = 2
  integer(c_int), parameter :: PSFUN_QUAD_RE_SUM = 0, PSFUN_QUAD_A_IM_SUM = 1

  interface
    ! ---- context ---------------------------------------------------------
    function psfun_ctx_create(device, ctx) bind(C, name='psfun_ctx_create') result(st)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: st
    end function
    function psfun_ctx_destroy(ctx) bind(C, name='psfun_ctx_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: st
    end function
    function psfun_ctx_synchronize(ctx) bind(C, name='psfun_ctx_synchronize') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: st
    end function
    ! USERF: f is a bind(C) function  real(c_double) f(x, user)  with x and
    ! user passed by value; a PSFUN wrapper forwards to fun%f_ptr(x).
    function psfun_set_userf(ctx, f, user) bind(C, name='psfun_set_userf') result(st)
      import :: c_int, c_ptr, c_funptr
      type(c_ptr), value :: ctx
      type(c_funptr), value :: f
      type(c_ptr), value :: user
      integer(c_int) :: st
    end function
    function psfun_set_pade_degree(ctx, degree) bind(C, name='psfun_set_pade_degree') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: degree
      integer(c_int) :: st
    end function

    ! ---- psb_dspmat_type (CSR) ---------------------------------------------
    function psfun_csr_create(ctx, n, nnz, rowptr, colind, val, index_base, mem, a) &
         & bind(C, name='psfun_csr_create') result(st)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n, nnz
      integer(c_int32_t), intent(in) :: rowptr(*), colind(*)
      real(c_double), intent(in) :: val(*)
      integer(c_int), value :: index_base, mem
      type(c_ptr), intent(out) :: a
      integer(c_int) :: st
    end function
    function psfun_csr_destroy(a) bind(C, name='psfun_csr_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int) :: st
    end function
    function psfun_csr_set_format(a, format) bind(C, name='psfun_csr_set_format') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int), value :: format
      integer(c_int) :: st
    end function
    function psfun_csr_hdia_diagonals(a, ndiag) bind(C, name='psfun_csr_hdia_diagonals') result(st)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: a
      integer(c_int64_t), intent(out) :: ndiag
      integer(c_int) :: st
    end function

    ! ---- psb_d_vect_type storage (device FP64 arrays) ----------------------
    function psfun_vec_alloc(ctx, n, d) bind(C, name='psfun_vec_alloc') result(st)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), intent(out) :: d
      integer(c_int) :: st
    end function
    function psfun_vec_free(ctx, d) bind(C, name='psfun_vec_free') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d
      integer(c_int) :: st
    end function
    function psfun_vec_upload(ctx, d, host, n) bind(C, name='psfun_vec_upload') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, d
      real(c_double), intent(in) :: host(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: st
    end function
    function psfun_vec_download(ctx, host, d, n) bind(C, name='psfun_vec_download') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, d
      real(c_double), intent(out) :: host(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: st
    end function

    ! NUL-terminated message of a status code (c_f_pointer it to a character array)
    function psfun_strerror(status) bind(C, name='psfun_strerror') result(msg)
      import :: c_int, c_ptr
      integer(c_int), value :: status
      type(c_ptr) :: msg
    end function
    function psfun_csr_info(a, n, nnz, lanes) bind(C, name='psfun_csr_info') result(st)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: a
      integer(c_int64_t), intent(out) :: n, nnz
      integer(c_int), intent(out) :: lanes
      integer(c_int) :: st
    end function

    ! ---- PSBLAS primitives -------------------------------------------------
    function psfun_spmv(ctx, a, alpha, x, beta, y) bind(C, name='psfun_spmv') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, x, y
      real(c_double), value :: alpha, beta
      integer(c_int) :: st
    end function
    function psfun_dot(ctx, n, x, y, res) bind(C, name='psfun_dot') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x, y
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: res
      integer(c_int) :: st
    end function
    function psfun_nrm2(ctx, n, x, res) bind(C, name='psfun_nrm2') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: res
      integer(c_int) :: st
    end function
    function psfun_amax(ctx, n, x, res) bind(C, name='psfun_amax') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: res
      integer(c_int) :: st
    end function
    function psfun_scal(ctx, n, alpha, x) bind(C, name='psfun_scal') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x
      integer(c_int64_t), value :: n
      real(c_double), value :: alpha
      integer(c_int) :: st
    end function
    function psfun_axpby(ctx, n, alpha, x, beta, y) bind(C, name='psfun_axpby') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, x, y
      integer(c_int64_t), value :: n
      real(c_double), value :: alpha, beta
      integer(c_int) :: st
    end function

    ! ---- f(H) step (psfun_d_serial%apply) ----------------------------------
    function psfun_serial_apply_host(ctx, fname, variant, scaling, k, h, x, y) &
         & bind(C, name='psfun_serial_apply_host') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: fname, variant, k
      real(c_double), value :: scaling
      real(c_double), intent(in) :: h(*), x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int) :: st
    end function

    ! ---- Krylov f(A)x (psfun_d_krylov%apply) -------------------------------
    function psfun_lanczos_apply(ctx, a, fname, fscale, opscale, x, y, eps, itmax, istop, iter, err, res, nres) &
         & bind(C, name='psfun_lanczos_apply') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, x, y
      integer(c_int), value :: fname, itmax, istop
      real(c_double), value :: fscale, opscale, eps
      integer(c_int), intent(out) :: iter, nres
      real(c_double), intent(out) :: err
      real(c_double), intent(out) :: res(*)
      integer(c_int) :: st
    end function
    ! The same apply on the library's fixed device scratch (no allocation; EXP only)
    function psfun_lanczos_apply_noalloc(ctx, a, fname, fscale, opscale, x, y, eps, itmax, istop, iter, err, res, &
         & nres) bind(C, name='psfun_lanczos_apply_noalloc') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, x, y
      integer(c_int), value :: fname, itmax, istop
      real(c_double), value :: fscale, opscale, eps
      integer(c_int), intent(out) :: iter, nres
      real(c_double), intent(out) :: err
      real(c_double), intent(out) :: res(*)
      integer(c_int) :: st
    end function
    function psfun_arnoldi_apply(ctx, a, fname, variant, fscale, scaling, x, y, eps, itmax, istop, iter, err, &
         & res, nres) bind(C, name='psfun_arnoldi_apply') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, x, y
      integer(c_int), value :: fname, variant, itmax, istop
      real(c_double), value :: fscale, scaling, eps
      integer(c_int), intent(out) :: iter, nres
      real(c_double), intent(out) :: err
      real(c_double), intent(out) :: res(*)
      integer(c_int) :: st
    end function

    function psfun_sai_arnoldi_apply(ctx, a, fname, variant, fscale, tau, prec, inner_itmax, inner_istop, x, y, &
         & eps, itmax, istop, iter, err, res, nres) bind(C, name='psfun_sai_arnoldi_apply') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, x, y
      integer(c_int), value :: fname, variant, prec, inner_itmax, inner_istop, itmax, istop
      real(c_double), value :: fscale, tau, eps
      integer(c_int), intent(out) :: iter, nres
      real(c_double), intent(out) :: err
      real(c_double), intent(out) :: res(*)
      integer(c_int) :: st
    end function

    ! ---- shifted solves (psb_krylov) ---------------------------------------
    function psfun_cg_shifted(ctx, a, prec, b, eta, zeta, x, eps, itmax, istop, iter, err, cond) &
         & bind(C, name='psfun_cg_shifted') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, b, x
      integer(c_int), value :: prec, itmax, istop
      real(c_double), value :: eta, zeta, eps
      integer(c_int), intent(out) :: iter
      real(c_double), intent(out) :: err
      type(c_ptr), value :: cond     ! c_loc(a real(c_double)) or c_null_ptr
      integer(c_int) :: st
    end function
    function psfun_cg_shifted_prec(ctx, a, p, b, eta, zeta, x, eps, itmax, istop, iter, err, cond) &
         & bind(C, name='psfun_cg_shifted_prec') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, p, b, x
      integer(c_int), value :: itmax, istop
      real(c_double), value :: eta, zeta, eps
      integer(c_int), intent(out) :: iter
      real(c_double), intent(out) :: err
      type(c_ptr), value :: cond
      integer(c_int) :: st
    end function
    ! ---- preconditioner objects (prec%init / prec%build / prec%apply) -------
    function psfun_prec_create(ctx, a, kind, p) bind(C, name='psfun_prec_create') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a
      integer(c_int), value :: kind
      type(c_ptr), intent(out) :: p
      integer(c_int) :: st
    end function
    function psfun_prec_destroy(p) bind(C, name='psfun_prec_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int) :: st
    end function
    function psfun_prec_info(p, kind, levels) bind(C, name='psfun_prec_info') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int), intent(out) :: kind, levels
      integer(c_int) :: st
    end function
    function psfun_prec_apply(ctx, p, r, z) bind(C, name='psfun_prec_apply') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, p, r, z
      integer(c_int) :: st
    end function
    function psfun_mscg(ctx, a, nshift, shifts, b, x, ldx, eps, itmax, iters, resid) &
         & bind(C, name='psfun_mscg') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, a, b, x
      integer(c_int), value :: nshift, itmax
      real(c_double), intent(in) :: shifts(*)      ! (re, im) pairs
      integer(c_int64_t), value :: ldx
      real(c_double), value :: eps
      integer(c_int), intent(out) :: iters(*)
      real(c_double), intent(out) :: resid(*)
      integer(c_int) :: st
    end function
    function psfun_quad_apply(ctx, a, nshift, shifts, weights, gamma, form, x, y, eps, itmax, iters, resid) &
         & bind(C, name='psfun_quad_apply') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, x, y
      integer(c_int), value :: nshift, form, itmax
      real(c_double), intent(in) :: shifts(*), weights(*)
      real(c_double), value :: gamma, eps
      integer(c_int), intent(out) :: iters(*)
      real(c_double), intent(out) :: resid(*)
      integer(c_int) :: st
    end function

    function psfun_mscg_prec(ctx, a, p, nshift, shifts, b, x, ldx, eps, itmax, iters, resid) &
         & bind(C, name='psfun_mscg_prec') result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, a, p, b, x
      integer(c_int), value :: nshift, itmax
      real(c_double), intent(in) :: shifts(*)
      integer(c_int64_t), value :: ldx
      real(c_double), value :: eps
      integer(c_int), intent(out) :: iters(*)
      real(c_double), intent(out) :: resid(*)
      integer(c_int) :: st
    end function
    function psfun_quad_apply_prec(ctx, a, p, nshift, shifts, weights, gamma, form, x, y, eps, itmax, iters, resid) &
         & bind(C, name='psfun_quad_apply_prec') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, a, p, x, y
      integer(c_int), value :: nshift, form, itmax
      real(c_double), intent(in) :: shifts(*), weights(*)
      real(c_double), value :: gamma, eps
      integer(c_int), intent(out) :: iters(*)
      real(c_double), intent(out) :: resid(*)
      integer(c_int) :: st
    end function

    ! ---- quadrature rules and special functions (host) ---------------------
    function psfun_horner(nc, c, x) bind(C, name='psfun_horner') result(v)
      import :: c_int, c_double
      integer(c_int), value :: nc
      real(c_double), intent(in) :: c(*)
      real(c_double), value :: x
      real(c_double) :: v
    end function
    function psfun_ellipkkp(l, k) bind(C, name='psfun_ellipkkp') result(st)
      import :: c_int, c_double
      real(c_double), value :: l
      real(c_double), intent(out) :: k(2)
      integer(c_int) :: st
    end function
    function psfun_ellipj(u_re, u_im, l, out) bind(C, name='psfun_ellipj') result(st)
      import :: c_int, c_double
      real(c_double), value :: u_re, u_im, l
      real(c_double), intent(out) :: out(6)
      integer(c_int) :: st
    end function
    function psfun_hhtmethod1(zfun, user, n, rparams, xi, c, eta, sign) bind(C, name='psfun_hhtmethod1') result(st)
      import :: c_int, c_double, c_ptr, c_funptr
      type(c_funptr), value :: zfun    ! subroutine f(z, fz, user) bind(C): z(2), fz(2) real(c_double)
      type(c_ptr), value :: user
      integer(c_int), value :: n
      real(c_double), intent(in) :: rparams(2)
      real(c_double), intent(out) :: xi(*), c(*), eta, sign
      integer(c_int) :: st
    end function
    function psfun_talbot(n, t, xi, c, eta, sign) bind(C, name='psfun_talbot') result(st)
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), value :: t
      real(c_double), intent(out) :: xi(*), c(*), eta, sign
      integer(c_int) :: st
    end function
    ! batched Lanczos over block-diagonal graphs (host CSR, 0-based, copied)
    function psfun_batch_create(ctx, ngraph, row_offsets, rowptr, colind, val, b) &
         & bind(C, name='psfun_batch_create') result(st)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngraph
      integer(c_int64_t), intent(in) :: row_offsets(*)
      integer(c_int32_t), intent(in) :: rowptr(*), colind(*)
      real(c_double), intent(in) :: val(*)
      type(c_ptr), intent(out) :: b
      integer(c_int) :: st
    end function
    function psfun_batch_destroy(b) bind(C, name='psfun_batch_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: st
    end function
    function psfun_batch_set_values(b, val) bind(C, name='psfun_batch_set_values') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: b, val          ! val: device array of nnz doubles
      integer(c_int) :: st
    end function
    function psfun_batch_lanczos(ctx, b, k, fscale, opscale, x, y, eps, itmax, istop, iters, errs) &
         & bind(C, name='psfun_batch_lanczos') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, b, x, y
      integer(c_int), value :: k, itmax, istop
      real(c_double), value :: fscale, opscale, eps
      type(c_ptr), value :: iters, errs     ! host arrays of ngraph*k, or c_null_ptr
      integer(c_int) :: st
    end function
    ! vector-Jacobian product of Y = exp(-tA) X (one symmetric matrix in b)
    function psfun_expv_vjp(ctx, b, k, t, x, y, g, eps, itmax, npoles, gx, gt, gvals) &
         & bind(C, name='psfun_expv_vjp') result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, b, x, y, g, gx, gt, gvals   ! gt: host scalar or c_null_ptr
      integer(c_int), value :: k, itmax, npoles
      real(c_double), value :: t, eps
      integer(c_int) :: st
    end function
  end interface
end module psfun_hip_mod
