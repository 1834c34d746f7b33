! psfun_krylov_mod.f90 -- PSFUN's extension of PSBLAS's psb_krylov to shifted
! systems (eta A + zeta I) x = b (krylov/psfun_krylov_mod.f90:39-142), over
! the device CG of krylov/shifted/psfun_dcg.F90.
!
!   call psb_krylov(method, a, prec, b, eta, zeta, x, eps, desc_a, info,
!                   itmax, iter, err, itrace, irst, istop, cond)
!
! 'CG' is the only method; any other name prints the reference's warning and
! runs CG without the condition estimate (lines 116-126).  x holds the
! initial guess; istop defaults to 2 and itmax to 1000 (psfun_dcg.F90:193-229).
module psfun_krylov_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  use psb_prec_mod
  implicit none

  interface psb_krylov
    module procedure psfun_dkrylov_vect
  end interface psb_krylov

contains

  subroutine psfun_dkrylov_vect(method, a, prec, b, eta, zeta, x, eps, desc_a, info, itmax, iter, err, itrace, &
       & irst, istop, cond)
    character(len=*) :: method
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    class(psb_dprec_type), intent(inout) :: prec
    type(psb_d_vect_type), intent(inout) :: b
    real(psb_dpk_), intent(in) :: eta, zeta
    type(psb_d_vect_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: eps
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: itmax, itrace, irst, istop
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err, cond
    integer(c_int) :: st, it, itmax_, istop_
    real(c_double) :: derr
    real(c_double), target :: cnd
    type(c_ptr) :: pc
    logical :: want_cond
    info = psb_success_
    if (present(cond)) cond = dzero
    want_cond = present(cond)
    if (psb_toupper(trim(method)) /= 'CG') then
      write(psb_err_unit, '(a,a,a)') 'psfun_krylov: Warning: Unknown method  ', trim(method), ', defaulting to CG'
      want_cond = .false.
    end if
    if (.not. c_associated(b%v) .or. .not. c_associated(x%v)) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psfun_krylov')
      return
    end if
    istop_ = 2
    if (present(istop)) istop_ = istop
    itmax_ = 1000
    if (present(itmax)) itmax_ = itmax
    if (istop_ /= 1 .and. istop_ /= 2) then
      info = psb_err_invalid_istop_
      call psb_errpush(info, 'psfun_krylov')
      return
    end if
    cnd = dzero
    pc = c_null_ptr
    if (want_cond) pc = c_loc(cnd)
    if (c_associated(prec%p)) then      ! a built preconditioner object
      st = psfun_cg_shifted_prec(a%ctxt%ctx, a%a, prec%p, b%v, eta, zeta, x%v, eps, itmax_, istop_, it, derr, pc)
    else
      st = psfun_cg_shifted(a%ctxt%ctx, a%a, prec%get_code(), b%v, eta, zeta, x%v, eps, itmax_, istop_, it, &
           & derr, pc)
    end if
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psfun_krylov', a_err=trim(method))
      return
    end if
    if (present(iter)) iter = it
    if (present(err)) err = derr
    if (want_cond) cond = cnd
    if (present(itrace)) then
      if (itrace > 0) write(psb_out_unit, '(a,i6,a,es15.8)') 'CG: iterations ', it, ', error estimate ', derr
    end if
  end subroutine psfun_dkrylov_vect

end module psfun_krylov_mod
