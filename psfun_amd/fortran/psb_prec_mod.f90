! psb_prec_mod.f90 -- the preconditioner object PSFUN's shifted solver takes
! (class(psb_dprec_type), krylov/psfun_krylov_mod.f90:55 and
! krylov/shifted/psfun_dcg.F90:283), with the kinds the device CG applies:
! 'NONE' (or 'NOPREC'), 'DIAG' (the inverse diagonal of the matrix) and
! 'BJAC' (block Jacobi: with one process the whole matrix factored by ILU(0),
! held on the device by psfun_prec_create), each built on the unshifted A as
! the reference's drivers build it (examples/shiftedkrylov/shiftedtest.f90:219-223).
! init/build/apply keep PSBLAS's call shapes prec%init(ctxt, ptype, info),
! prec%build(a, desc_a, info) and prec%apply(x, y, desc_a, info).
module psb_prec_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  implicit none

  type :: psb_dprec_type
    character(len=20) :: ptype = ''
    integer(c_int) :: code = PSFUN_PREC_NONE
    logical :: built = .false.
    type(c_ptr) :: p = c_null_ptr       ! the device object once built
    integer(c_int) :: levels = 0
  contains
    procedure, pass(prec) :: init => psb_dprec_init
    procedure, pass(prec) :: build => psb_dprec_build
    procedure, pass(prec) :: free => psb_dprec_free
    procedure, pass(prec) :: apply => psb_dprec_apply_vect
    procedure, pass(prec) :: descr => psb_dprec_descr
    procedure, pass(prec) :: sizeof => psb_dprec_sizeof
    procedure, pass(prec) :: get_code => psb_dprec_get_code
    procedure, pass(prec) :: is_allocated_wrk => psb_dprec_is_allocated_wrk
    procedure, pass(prec) :: allocate_wrk => psb_dprec_allocate_wrk
    procedure, pass(prec) :: free_wrk => psb_dprec_free_wrk
  end type psb_dprec_type

contains

  subroutine psb_dprec_init(prec, ctxt, ptype, info)
    class(psb_dprec_type), intent(inout) :: prec
    type(psb_ctxt_type), intent(in) :: ctxt
    character(len=*), intent(in) :: ptype
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(trim(ptype)))
    case ('NONE', 'NOPREC')
      prec%ptype = 'NONE'
      prec%code = PSFUN_PREC_NONE
    case ('DIAG')
      prec%ptype = 'DIAG'
      prec%code = PSFUN_PREC_DIAG
    case ('BJAC')
      prec%ptype = 'BJAC'
      prec%code = PSFUN_PREC_BJAC
    case default
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_precinit', a_err=trim(ptype))
      return
    end select
    call release(prec)
    prec%built = .false.
  end subroutine psb_dprec_init

  subroutine release(prec)
    class(psb_dprec_type), intent(inout) :: prec
    integer(c_int) :: st
    if (c_associated(prec%p)) st = psfun_prec_destroy(prec%p)
    prec%p = c_null_ptr
    prec%levels = 0
  end subroutine release

  ! Creates the device object: NONE and DIAG keep no state (the solvers form
  ! the diagonal they need), BJAC factors A here (a zero pivot is an error).
  subroutine psb_dprec_build(prec, a, desc_a, info)
    class(psb_dprec_type), intent(inout) :: prec
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(c_int) :: st, kind
    info = psb_success_
    if (len_trim(prec%ptype) == 0) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_precbld', a_err='preconditioner not initialised')
      return
    end if
    call release(prec)
    st = psfun_prec_create(a%ctxt%ctx, a%a, prec%code, prec%p)
    if (st /= PSFUN_SUCCESS) then
      prec%p = c_null_ptr
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psb_precbld', a_err='ILU(0) factorization')
      return
    end if
    st = psfun_prec_info(prec%p, kind, prec%levels)
    prec%built = .true.
  end subroutine psb_dprec_build

  subroutine psb_dprec_free(prec, info)
    class(psb_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    call release(prec)
    prec%built = .false.
  end subroutine psb_dprec_free

  ! y = M^{-1} x for a built BJAC object (NONE and DIAG are applied inside
  ! the solver and report an error here).
  subroutine psb_dprec_apply_vect(prec, x, y, desc_a, info)
    class(psb_dprec_type), intent(inout) :: prec
    type(psb_d_vect_type), intent(inout) :: x, y
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    if (.not. c_associated(prec%p) .or. prec%code /= PSFUN_PREC_BJAC) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_precaply', a_err='no built BJAC object')
      return
    end if
    if (psfun_prec_apply(desc_a%ctxt%ctx, prec%p, x%v, y%v) /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psb_precaply')
    end if
  end subroutine psb_dprec_apply_vect

  subroutine psb_dprec_descr(prec, info, iout)
    class(psb_dprec_type), intent(in) :: prec
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iout
    integer :: u
    info = psb_success_
    u = psb_out_unit
    if (present(iout)) u = iout
    write(u, '(a,a)') 'Preconditioner: ', trim(prec%ptype)
    if (prec%code == PSFUN_PREC_BJAC .and. c_associated(prec%p)) &
         & write(u, '(a,i8)') 'ILU(0) triangular-solve levels: ', prec%levels
  end subroutine psb_dprec_descr

  function psb_dprec_sizeof(prec) result(s)
    class(psb_dprec_type), intent(in) :: prec
    integer(psb_epk_) :: s
    s = 0
  end function psb_dprec_sizeof

  function psb_dprec_get_code(prec) result(c)
    class(psb_dprec_type), intent(in) :: prec
    integer(c_int) :: c
    c = prec%code
  end function psb_dprec_get_code

  ! PSFUN's psb_krylov asks for per-call work space (psfun_krylov_mod.f90:113-114);
  ! the device solver keeps its own, so these report it as present.
  function psb_dprec_is_allocated_wrk(prec) result(r)
    class(psb_dprec_type), intent(in) :: prec
    logical :: r
    r = .true.
  end function psb_dprec_is_allocated_wrk

  subroutine psb_dprec_allocate_wrk(prec, info, vmold, desc)
    class(psb_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    class(*), intent(in), optional :: vmold
    type(psb_desc_type), intent(in), optional :: desc
    info = psb_success_
  end subroutine psb_dprec_allocate_wrk

  subroutine psb_dprec_free_wrk(prec, info)
    class(psb_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
  end subroutine psb_dprec_free_wrk

end module psb_prec_mod
