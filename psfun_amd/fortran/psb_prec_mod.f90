! psb_prec_mod.f90 -- the preconditioner object PSFUN's shifted solver takes
! (class(psb_dprec_type), krylov/psfun_krylov_mod.f90:55 and
! krylov/shifted/psfun_dcg.F90:283), with the kinds the device CG applies:
! 'NONE' (or 'NOPREC') and 'DIAG' (the inverse diagonal of the matrix, built
! on the unshifted A as the reference's drivers build it,
! examples/shiftedkrylov/shiftedtest.f90:219-223).  init/build keep PSBLAS's
! call shapes prec%init(ctxt, ptype, info) and prec%build(a, desc_a, info).
module psb_prec_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  implicit none

  type :: psb_dprec_type
    character(len=20) :: ptype = ''
    integer(c_int) :: code = PSFUN_PREC_NONE
    logical :: built = .false.
  contains
    procedure, pass(prec) :: init => psb_dprec_init
    procedure, pass(prec) :: build => psb_dprec_build
    procedure, pass(prec) :: free => psb_dprec_free
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
    case default
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_precinit', a_err=trim(ptype))
      return
    end select
    prec%built = .false.
  end subroutine psb_dprec_init

  ! The device CG forms the inverse diagonal of the matrix it is given (the
  ! unshifted A) itself; building only checks the object is initialised.
  subroutine psb_dprec_build(prec, a, desc_a, info)
    class(psb_dprec_type), intent(inout) :: prec
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    if (len_trim(prec%ptype) == 0) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_precbld', a_err='preconditioner not initialised')
      return
    end if
    prec%built = .true.
  end subroutine psb_dprec_build

  subroutine psb_dprec_free(prec, info)
    class(psb_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    prec%built = .false.
  end subroutine psb_dprec_free

  subroutine psb_dprec_descr(prec, info, iout)
    class(psb_dprec_type), intent(in) :: prec
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iout
    integer :: u
    info = psb_success_
    u = psb_out_unit
    if (present(iout)) u = iout
    write(u, '(a,a)') 'Preconditioner: ', trim(prec%ptype)
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
