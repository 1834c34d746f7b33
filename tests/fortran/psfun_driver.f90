! Fortran caller of the device library through include/psfun_hip_mod.f90, in
! the shape of examples/polynomial/lanczostest.F90 and arnoldi.inp: the 50x50
! example matrix laplacian.mtx (1-D Laplacian, diagonal -2, off-diagonals 1),
! given as 1-based host CSR, x = (1,...,1), y = exp(A) x by
!   LANCZOS  (EXP, CHBSYM, scaling 1, eps 1e-6, itmax 100, istop 1)
!   ARNOLDI  (EXP, GENPADE, scaling 1, eps 1e-6, itmax 100, istop 1)
! and one shifted solve as in examples/shiftedkrylov/shifted.inp:
!   CG       ((-A + I) x = 1, DIAG, eps 1e-7, istop 2),
! and the LANCZOS deck again with fname USERF/SYM and a Fortran user function
! f(x) = exp(x) handed over with c_funloc (USERF).  Last, the batched and
! differentiable forms on B = -A (SPD) with K = 2 columns, t = 0.1:
!   BATCH    Y = exp(-tB) X                        (n x 2, row-major)
!   VJPGX    dL/dX = exp(-tB) G for L = <G, Y>     (err field: dL/dt)
!   VJPGV    dL/d(stored nonzeros of B)             (err field: dL/dt)
! Prints "<METHOD> <iter> <err>" followed by the entries of the result.
module driver_userf
  use iso_c_binding
  implicit none
contains
  function userf_exp(x, user) bind(C) result(fx)
    real(c_double), value :: x
    type(c_ptr), value :: user
    real(c_double) :: fx
    fx = exp(x)
  end function
end module driver_userf

program psfun_driver
  use iso_c_binding
  use psfun_hip_mod
  use driver_userf
  implicit none
  integer, parameter :: n = 50
  integer(c_int32_t) :: rowptr(n + 1), colind(3 * n - 2)
  real(c_double) :: val(3 * n - 2), x(n), y(n), res(102)
  real(c_double), target :: cond
  type(c_ptr) :: ctx, a, dx, dy
  integer(c_int) :: st, iter, nres
  integer :: i, k
  real(c_double) :: err

  k = 0
  do i = 1, n
    rowptr(i) = k + 1
    if (i > 1) then
      k = k + 1
      colind(k) = i - 1
      val(k) = 1.0d0
    end if
    k = k + 1
    colind(k) = i
    val(k) = -2.0d0
    if (i < n) then
      k = k + 1
      colind(k) = i + 1
      val(k) = 1.0d0
    end if
  end do
  rowptr(n + 1) = k + 1
  x = 1.0d0

  call check(psfun_ctx_create(0_c_int, ctx), 'psfun_ctx_create')
  call check(psfun_csr_create(ctx, int(n, c_int64_t), int(k, c_int64_t), rowptr, colind, val, 1_c_int, &
       & PSFUN_MEM_HOST, a), 'psfun_csr_create')
  call check(psfun_vec_alloc(ctx, int(n, c_int64_t), dx), 'psfun_vec_alloc')
  call check(psfun_vec_alloc(ctx, int(n, c_int64_t), dy), 'psfun_vec_alloc')
  call check(psfun_vec_upload(ctx, dx, x, int(n, c_int64_t)), 'psfun_vec_upload')

  st = psfun_lanczos_apply(ctx, a, PSFUN_FN_EXP, 1.0d0, 1.0d0, dx, dy, 1.0d-6, 100_c_int, 1_c_int, iter, err, &
       & res, nres)
  call check(st, 'psfun_lanczos_apply')
  call check(psfun_vec_download(ctx, y, dy, int(n, c_int64_t)), 'psfun_vec_download')
  call emit('LANCZOS', iter, err, y)

  st = psfun_arnoldi_apply(ctx, a, PSFUN_FN_EXP, PSFUN_VAR_GENPADE, 1.0d0, 1.0d0, dx, dy, 1.0d-6, 100_c_int, &
       & 1_c_int, iter, err, res, nres)
  call check(st, 'psfun_arnoldi_apply')
  call check(psfun_vec_download(ctx, y, dy, int(n, c_int64_t)), 'psfun_vec_download')
  call emit('ARNOLDI', iter, err, y)

  y = 0.0d0
  call check(psfun_vec_upload(ctx, dy, y, int(n, c_int64_t)), 'psfun_vec_upload')
  st = psfun_cg_shifted(ctx, a, PSFUN_PREC_DIAG, dx, -1.0d0, 1.0d0, dy, 1.0d-7, 1000_c_int, 2_c_int, iter, err, &
       & c_loc(cond))
  call check(st, 'psfun_cg_shifted')
  call check(psfun_vec_download(ctx, y, dy, int(n, c_int64_t)), 'psfun_vec_download')
  call emit('CG', iter, err, y)

  call check(psfun_set_userf(ctx, c_funloc(userf_exp), c_null_ptr), 'psfun_set_userf')
  st = psfun_lanczos_apply(ctx, a, PSFUN_FN_USERF, 1.0d0, 1.0d0, dx, dy, 1.0d-6, 100_c_int, 1_c_int, iter, err, &
       & res, nres)
  call check(st, 'psfun_lanczos_apply (USERF)')
  call check(psfun_vec_download(ctx, y, dy, int(n, c_int64_t)), 'psfun_vec_download')
  call emit('USERF', iter, err, y)

  call vjp_section(ctx, k)

  call check(psfun_vec_free(ctx, dx), 'psfun_vec_free')
  call check(psfun_vec_free(ctx, dy), 'psfun_vec_free')
  call check(psfun_csr_destroy(a), 'psfun_csr_destroy')
  call check(psfun_ctx_destroy(ctx), 'psfun_ctx_destroy')

contains

  subroutine vjp_section(ctx, nnz)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nnz
    integer, parameter :: kk = 2
    integer(c_int32_t) :: rp0(n + 1), ci0(nnz)
    integer(c_int64_t) :: offs(2)
    real(c_double) :: nv(nnz), xh(n * kk), gh(n * kk), out(n * kk), gv(nnz)
    real(c_double), target :: gt
    type(c_ptr) :: b, bx, by, bg, bgx, bgv
    integer :: i
    integer(c_int64_t) :: nk
    ! 0-based copy of the pattern, values of B = -A
    rp0 = rowptr - 1
    ci0 = colind(1:nnz) - 1
    nv = -val(1:nnz)
    offs = [0_c_int64_t, int(n, c_int64_t)]
    do i = 1, n
      xh(kk * (i - 1) + 1) = 1.0d0
      xh(kk * (i - 1) + 2) = dble(i) / dble(n)
      gh(kk * (i - 1) + 1) = sin(0.3d0 * i)
      gh(kk * (i - 1) + 2) = cos(0.7d0 * i)
    end do
    nk = int(n * kk, c_int64_t)
    call check(psfun_batch_create(ctx, 1_c_int, offs, rp0, ci0, nv, b), 'psfun_batch_create')
    call check(psfun_vec_alloc(ctx, nk, bx), 'psfun_vec_alloc')
    call check(psfun_vec_alloc(ctx, nk, by), 'psfun_vec_alloc')
    call check(psfun_vec_alloc(ctx, nk, bg), 'psfun_vec_alloc')
    call check(psfun_vec_alloc(ctx, nk, bgx), 'psfun_vec_alloc')
    call check(psfun_vec_alloc(ctx, int(nnz, c_int64_t), bgv), 'psfun_vec_alloc')
    call check(psfun_vec_upload(ctx, bx, xh, nk), 'psfun_vec_upload')
    call check(psfun_vec_upload(ctx, bg, gh, nk), 'psfun_vec_upload')
    call check(psfun_batch_lanczos(ctx, b, int(kk, c_int), 1.0d0, -0.1d0, bx, by, 1.0d-12, 50_c_int, 1_c_int, &
         & c_null_ptr, c_null_ptr), 'psfun_batch_lanczos')
    call check(psfun_expv_vjp(ctx, b, int(kk, c_int), 0.1d0, bx, by, bg, 1.0d-12, 50_c_int, 24_c_int, bgx, &
         & c_loc(gt), bgv), 'psfun_expv_vjp')
    call check(psfun_vec_download(ctx, out, by, nk), 'psfun_vec_download')
    call emit('BATCH', 0_c_int, 0.0d0, out)
    call check(psfun_vec_download(ctx, out, bgx, nk), 'psfun_vec_download')
    call emit('VJPGX', 0_c_int, gt, out)
    call check(psfun_vec_download(ctx, gv, bgv, int(nnz, c_int64_t)), 'psfun_vec_download')
    call emit('VJPGV', 0_c_int, gt, gv)
    call check(psfun_vec_free(ctx, bx), 'psfun_vec_free')
    call check(psfun_vec_free(ctx, by), 'psfun_vec_free')
    call check(psfun_vec_free(ctx, bg), 'psfun_vec_free')
    call check(psfun_vec_free(ctx, bgx), 'psfun_vec_free')
    call check(psfun_vec_free(ctx, bgv), 'psfun_vec_free')
    call check(psfun_batch_destroy(b), 'psfun_batch_destroy')
  end subroutine

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= PSFUN_SUCCESS) then
      write (*, '(a,a,i0)') what, ' failed with status ', status
      error stop 1
    end if
  end subroutine

  subroutine emit(name, it, e, v)
    character(len=*), intent(in) :: name
    integer(c_int), intent(in) :: it
    real(c_double), intent(in) :: e, v(:)
    integer :: q
    write (*, '(a,1x,i0,1x,es24.16)') name, it, e
    do q = 1, size(v)
      write (*, '(es24.16)') v(q)
    end do
  end subroutine
end program psfun_driver
