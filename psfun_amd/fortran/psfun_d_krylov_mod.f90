! psfun_d_krylov_mod.f90 -- PSFUN's Krylov f(A)x method object
! (krylov/psfun_d_krylov_mod.F90:36-269) over the gfx950 library.
!
! type(psfun_d_krylov) keeps the reference's components and defaults
! (kname 'ARNOLDI', kmethd 'CG', prec, tau 0, itmax 200, itrace 0, istopc 1,
! irst 10) and its type-bound procedures:
!   kmethd%set('KNAME', string, info), kmethd%set('ITMAX'|'ITRACE'|'ISTOPC'|'IRST', int, info)
!   kmethd%apply(fun, a, desc_a, y, x, eps, info, itmax, itrace, istop, iter, err, res)
!   kmethd%precinit(ctxt, ptype, info), kmethd%precbuild(a, desc_a, info)
!   kmethd%plot(fun, iter, res, info)
! apply dispatches on psb_toupper(kname) to psfun_d_lanczos, psfun_d_arnoldi
! or psfun_d_saiarnoldi (lines 254-269); each runs its whole Krylov loop, the
! f(H) step and the stopping test on the device and returns res(1:iter+1)
! allocated here, as the reference does (psfun_d_lanczos.f90:367-371).
! With itrace > 0 the convergence log is printed after the call, one line per
! itrace-th estimate (the values the reference's log_conv prints during it).
module psfun_d_krylov_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  use psb_prec_mod
  use psfun_d_serial_mod
  implicit none

  type, public :: psfun_d_krylov
    character(len=20) :: kname = 'ARNOLDI'
    character(len=20) :: kmethd = 'CG'
    type(psb_dprec_type) :: prec
    real(psb_dpk_) :: tau = dzero
    integer(psb_ipk_) :: itmax = 200
    integer(psb_ipk_) :: itrace = 0
    integer(psb_ipk_) :: istopc = 1
    integer(psb_ipk_) :: irst = 10
  contains
    procedure, pass(meth) :: setstring => psfun_d_setstring
    procedure, pass(meth) :: setinteger => psfun_d_setinteger
    generic, public :: set => setstring, setinteger
    procedure, pass(meth) :: apply => psfun_d_parallel_apply
    procedure, pass(meth) :: plot => psfun_d_plot_info
    procedure, pass(meth) :: precinit => psfun_d_prec_init
    procedure, pass(meth) :: precbuild => psfun_d_prec_build
  end type psfun_d_krylov

  private :: psfun_d_setstring, psfun_d_setinteger

contains

  subroutine psfun_d_setstring(meth, what, val, info)
    class(psfun_d_krylov), intent(inout) :: meth
    character(len=*), intent(in) :: what, val
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(what))
    case ('KNAME')
      meth%kname = val
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_setstring

  subroutine psfun_d_setinteger(meth, what, val, info)
    class(psfun_d_krylov), intent(inout) :: meth
    character(len=*), intent(in) :: what
    integer(psb_ipk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(what))
    case ('ITMAX')
      meth%itmax = val
    case ('ITRACE')
      meth%itrace = val
    case ('ISTOPC')
      meth%istopc = val
    case ('IRST')
      meth%irst = val
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_setinteger

  subroutine psfun_d_parallel_apply(meth, fun, a, desc_a, y, x, eps, info, itmax, itrace, istop, iter, err, res)
    class(psfun_d_krylov), intent(inout) :: meth
    type(psfun_d_serial), intent(inout) :: fun
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    type(psb_d_vect_type), intent(inout) :: y, x
    real(psb_dpk_), intent(in) :: eps
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: itmax, itrace, istop
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err
    real(psb_dpk_), optional, allocatable, intent(out) :: res(:)
    select case (psb_toupper(meth%kname))
    case ('ARNOLDI')
      call psfun_d_arnoldi(fun, a, desc_a, y, x, eps, info, itmax, itrace, istop, iter, err, res)
    case ('LANCZOS')
      call psfun_d_lanczos(fun, a, desc_a, y, x, eps, info, itmax, itrace, istop, iter, err, res)
    case ('SAIARNOLDI')
      call psfun_d_saiarnoldi(fun, a, meth%kmethd, meth%prec, meth%tau, desc_a, meth%itmax, meth%itrace, &
           & meth%istopc, y, x, eps, info, itmax, itrace, istop, iter, err, res)
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_parallel_apply

  ! The reference plots with gnuplot only when built WITHGNUPLOTFORTRAN and
  ! returns psb_err_from_subroutine_ otherwise; no plotting library here.
  subroutine psfun_d_plot_info(meth, fun, iter, res, info)
    class(psfun_d_krylov), intent(inout) :: meth
    type(psfun_d_serial), intent(inout) :: fun
    integer(psb_ipk_), intent(in) :: iter
    real(psb_dpk_), intent(in) :: res(:)
    integer(psb_ipk_), intent(out) :: info
    info = psb_err_from_subroutine_
  end subroutine psfun_d_plot_info

  subroutine psfun_d_prec_init(meth, ctxt, ptype, info)
    class(psfun_d_krylov), intent(inout) :: meth
    type(psb_ctxt_type), intent(in) :: ctxt
    character(len=*), intent(in) :: ptype
    integer(psb_ipk_) :: info
    call meth%prec%init(ctxt, ptype, info)
    if (info /= psb_success_) call psb_errpush(info, 'd_prec_init')
  end subroutine psfun_d_prec_init

  subroutine psfun_d_prec_build(meth, a, desc_a, info)
    class(psfun_d_krylov), intent(inout) :: meth
    type(psb_dspmat_type), intent(inout) :: a
    type(psb_desc_type), intent(inout) :: desc_a
    integer(psb_ipk_) :: info
    call meth%prec%build(a, desc_a, info)
    if (info /= psb_success_) call psb_errpush(info, 'd_prec_build')
  end subroutine psfun_d_prec_build

  ! ---- shared argument handling of the three methods ---------------------------
  ! istop (default 1, only 1 or 2), itmax (default 200), allocated vectors and
  ! matching sizes (psfun_d_lanczos.f90:97-142); ok = .false. sets info and err.
  subroutine krylov_args(a, desc_a, y, x, itmax, istop, nl, istop_, info, err, ok)
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    type(psb_d_vect_type), intent(in) :: y, x
    integer(psb_ipk_), optional, intent(in) :: itmax, istop
    integer(c_int), intent(out) :: nl, istop_
    integer(psb_ipk_), intent(out) :: info
    real(psb_dpk_), optional, intent(out) :: err
    logical, intent(out) :: ok
    ok = .false.
    info = psb_success_
    istop_ = 1
    if (present(istop)) istop_ = istop
    if (istop_ < 1 .or. istop_ > 2) then
      info = psb_err_invalid_istop_
      if (present(err)) err = real(info, psb_dpk_)   ! err = info, as lines 108-113 do
      call psb_errpush(info, 'psfun_krylov')
      return
    end if
    nl = 200
    if (present(itmax)) nl = itmax
    if (.not. c_associated(x%v) .or. .not. c_associated(y%v)) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psfun_krylov')
      return
    end if
    if (x%n /= desc_a%n .or. y%n /= desc_a%n .or. a%m /= desc_a%n) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psfun_krylov', a_err='sizes of a, x, y and desc_a differ')
      return
    end if
    ok = .true.
  end subroutine krylov_args

  ! Outputs of a device run: iter, err, res(1:nres) allocated here, the trace.
  subroutine krylov_outputs(methdname, st, it, derr, rbuf, nres, eps, itrace, estimates, info, iter, err, res)
    character(len=*), intent(in) :: methdname
    logical, intent(in) :: estimates
    integer(c_int), intent(in) :: st, it, nres
    real(c_double), intent(in) :: derr, rbuf(:), eps
    integer(psb_ipk_), optional, intent(in) :: itrace
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err
    real(psb_dpk_), optional, allocatable, intent(out) :: res(:)
    integer :: i, tr
    if (st /= PSFUN_SUCCESS) then
      info = psb_status_to_info(st)
      if (info == psb_err_unsupported_) info = psb_err_from_subroutine_non_
      call psb_errpush(info, methdname, a_err=trim(psb_status_text(st)))
      return
    end if
    info = psb_success_
    if (present(iter)) iter = it
    if (present(err)) err = derr
    if (present(res)) then
      allocate(res(nres))
      res(1:nres) = rbuf(1:nres)
    end if
    tr = 0
    if (present(itrace)) tr = itrace
    if (tr > 0) then
      ! log_header / log_conv / log_end of the reference's loop
      ! (psfun_d_lanczos.f90:180, 304-305, 365): the estimate at loop index i
      ! is rs(i) = res(i) for itx = i - 1 (res(it+1) is the final err); with
      ! istop 2 there are no estimates, and log_conv is not called.
      write(psb_out_unit, '(a,a)') trim(methdname), ':   Iteration   Error estimate   Tolerance'
      if (estimates) then
        do i = 2, min(it + 1, nres)
          if (mod(i - 1, tr) == 0) write(psb_out_unit, '(a,a,i6,2(2x,es15.8))') trim(methdname), ':', i - 1, &
               & rbuf(i), eps
        end do
      end if
      write(psb_out_unit, '(a,a,i6,a,es15.8)') trim(methdname), ': iterations ', it, ', error estimate ', derr
    end if
  end subroutine krylov_outputs

  ! ---- the methods ---------------------------------------------------------------
  ! psfun_d_lanczos (krylov/polynomial/psfun_d_lanczos.f90:37-400): the Krylov
  ! operator is A itself and f(T) is evaluated with scaling 1, as there
  ! (fun%scaling is ignored, lines 172-173, 198, 239).
  subroutine psfun_d_lanczos(fun, a, desc_a, y, x, eps, info, itmax, itrace, istop, iter, err, res)
    type(psfun_d_serial), intent(inout) :: fun
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    type(psb_d_vect_type), intent(inout) :: y, x
    real(psb_dpk_), intent(in) :: eps
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: itmax, itrace, istop
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err
    real(psb_dpk_), optional, allocatable, intent(out) :: res(:)
    integer(c_int) :: nl, istop_, fn, var, st, it, nres
    real(c_double) :: derr
    real(c_double), allocatable :: rbuf(:)
    type(psfun_userf_box), target :: box
    logical :: ok
    call krylov_args(a, desc_a, y, x, itmax, istop, nl, istop_, info, err, ok)
    if (.not. ok) return
    call fun%device_fn(fn, var, st)
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_non_
      call psb_errpush(info, 'psfun_d_lanczos', a_err=trim(fun%fname)//' '//trim(fun%variant))
      return
    end if
    allocate(rbuf(max(nl, 2) + 2))
    call psfun_bind_userf(fun, a%ctxt%ctx, box)
    st = psfun_lanczos_apply(a%ctxt%ctx, a%a, fn, 1.0_c_double, 1.0_c_double, x%v, y%v, eps, nl, istop_, it, &
         & derr, rbuf, nres)
    call psfun_unbind_userf(a%ctxt%ctx)
    call krylov_outputs('LANCZOS+'//trim(fun%fname)//'+'//trim(fun%variant), st, it, derr, rbuf, nres, eps, &
         & itrace, istop_ == 1, info, iter, err, res)
  end subroutine psfun_d_lanczos

  ! psfun_d_arnoldi (krylov/polynomial/psfun_d_arnoldi.f90:37-352): the SpMV
  ! is multiplied by fun%scaling and f(H) uses scaling 1 (lines 172-174, 202).
  subroutine psfun_d_arnoldi(fun, a, desc_a, y, x, eps, info, itmax, itrace, istop, iter, err, res)
    type(psfun_d_serial), intent(inout) :: fun
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    type(psb_d_vect_type), intent(inout) :: y, x
    real(psb_dpk_), intent(in) :: eps
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: itmax, itrace, istop
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err
    real(psb_dpk_), optional, allocatable, intent(out) :: res(:)
    integer(c_int) :: nl, istop_, fn, var, st, it, nres
    real(c_double) :: derr
    real(c_double), allocatable :: rbuf(:)
    type(psfun_userf_box), target :: box
    logical :: ok
    call krylov_args(a, desc_a, y, x, itmax, istop, nl, istop_, info, err, ok)
    if (.not. ok) return
    call fun%device_fn(fn, var, st)
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_non_
      call psb_errpush(info, 'psfun_d_arnoldi', a_err=trim(fun%fname)//' '//trim(fun%variant))
      return
    end if
    allocate(rbuf(max(nl, 2) + 2))
    call psfun_bind_userf(fun, a%ctxt%ctx, box)
    st = psfun_arnoldi_apply(a%ctxt%ctx, a%a, fn, var, 1.0_c_double, fun%scaling, x%v, y%v, eps, nl, istop_, &
         & it, derr, rbuf, nres)
    call psfun_unbind_userf(a%ctxt%ctx)
    call krylov_outputs('ARNOLDI+'//trim(fun%fname)//'+'//trim(fun%variant), st, it, derr, rbuf, nres, eps, &
         & itrace, istop_ == 1, info, iter, err, res)
  end subroutine psfun_d_arnoldi

  ! psfun_d_saiarnoldi (krylov/rational/psfun_d_sai.f90:37-419): Arnoldi on
  ! (A + tau I)^{-1} with every inner solve by the device CG ('CG'; any other
  ! kryl warns and runs CG, as psb_krylov does), preconditioned by prec.
  subroutine psfun_d_saiarnoldi(fun, a, kryl, prec, tau, desc_a, initmax, initrace, inistop, y, x, eps, info, &
       & itmax, itrace, istop, iter, err, res)
    type(psfun_d_serial), intent(inout) :: fun
    type(psb_dspmat_type), intent(in) :: a
    character(len=*), intent(in) :: kryl
    type(psb_dprec_type), intent(inout) :: prec
    real(psb_dpk_), intent(in) :: tau
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(in) :: initmax, initrace, inistop
    type(psb_d_vect_type), intent(inout) :: y, x
    real(psb_dpk_), intent(in) :: eps
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: itmax, itrace, istop
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err
    real(psb_dpk_), optional, allocatable, intent(out) :: res(:)
    integer(c_int) :: nl, istop_, fn, var, st, it, nres
    real(c_double) :: derr
    real(c_double), allocatable :: rbuf(:)
    type(psfun_userf_box), target :: box
    logical :: ok
    call krylov_args(a, desc_a, y, x, itmax, istop, nl, istop_, info, err, ok)
    if (.not. ok) return
    if (psb_toupper(trim(kryl)) /= 'CG') &
         & write(psb_err_unit, '(a,a,a)') 'psfun_krylov: Warning: Unknown method  ', trim(kryl), ', defaulting to CG'
    call fun%device_fn(fn, var, st)
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_non_
      call psb_errpush(info, 'psfun_d_saiarnoldi', a_err=trim(fun%fname)//' '//trim(fun%variant))
      return
    end if
    allocate(rbuf(max(nl, 2) + 2))
    call psfun_bind_userf(fun, a%ctxt%ctx, box)
    st = psfun_sai_arnoldi_apply(a%ctxt%ctx, a%a, fn, var, 1.0_c_double, tau, prec%get_code(), initmax, inistop, &
         & x%v, y%v, eps, nl, istop_, it, derr, rbuf, nres)
    call psfun_unbind_userf(a%ctxt%ctx)
    call krylov_outputs('SAIARNOLDI+'//trim(fun%fname)//'+'//trim(fun%variant), st, it, derr, rbuf, nres, eps, &
         & itrace, istop_ == 1, info, iter, err, res)
  end subroutine psfun_d_saiarnoldi

end module psfun_d_krylov_mod
