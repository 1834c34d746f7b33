! psfun_d_serial_mod.f90 -- the function object of PSFUN's f(H) step
! (serial/psfun_d_serial_mod.F90:43-170) over the gfx950 library.
!
! type(psfun_d_serial) keeps the reference's components and defaults (fname
! 'EXP', variant 'EXPOKIT', scaling 1, padedegree 6, phiorder 1, f_ptr) and
! its type-bound set/apply generics:
!   fun%set('FNAME'|'VARIANT', string, info), fun%set('SCALING', real, info),
!   fun%set('PADE_DEGREE', integer, info), fun%set('FPOINTER', f, info)
!   fun%apply(a(:,:), y, x, info)            (psfun_d_serial_apply_array)
!   fun%apply(a_sparse, y, x, info)          (psfun_d_serial_apply_sparse)
! Option names go through psb_toupper; values are matched case-sensitively
! (lines 214-216), so the default variant 'EXPOKIT' is rejected by apply
! (lines 387-390).  The EXP variants TAYLOR, SASQ, GENPADE, CHBHES, CHBGEN,
! CHBSYM and USERF/SYM run on the device (psfun_serial_apply_host); PHI
! needs an external library the reference links only WITHPHILIBRARY and
! reports psb_err_from_subroutine_ as the reference does without it.
module psfun_d_serial_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  implicit none

  type, public :: psfun_d_serial
    character(len=20) :: fname = 'EXP'
    character(len=20) :: variant = 'EXPOKIT'
    real(psb_dpk_) :: scaling = 1.0_psb_dpk_
    integer(psb_ipk_) :: padedegree = 6_psb_ipk_
    integer(psb_ipk_) :: phiorder = 1_psb_ipk_
    procedure(func), pointer, nopass :: f_ptr => null()
  contains
    procedure, pass(fun) :: setstring => psfun_d_setstring
    procedure, pass(fun) :: setreal => psfun_d_setreal
    procedure, pass(fun) :: setinteger => psfun_d_setinteger
    procedure, pass(fun) :: setfunction => psfun_d_setpointer
    generic, public :: set => setstring, setreal, setinteger, setfunction
    procedure, pass(fun) :: applya => psfun_d_serial_apply_array
    procedure, pass(fun) :: applys => psfun_d_serial_apply_sparse
    generic, public :: apply => applya, applys
    procedure, pass(fun) :: device_fn => psfun_d_device_fn
  end type psfun_d_serial

  abstract interface
    function func(x, k)
      import :: psb_dpk_, psb_ipk_
      real(psb_dpk_) :: func
      real(psb_dpk_), intent(in) :: x
      integer(psb_ipk_), intent(in), optional :: k
    end function func
  end interface

  ! Carrier of a USERF function across the C boundary: the library calls
  ! psfun_userf_trampoline(x, c_loc(box)), which calls box%f(x).
  type :: psfun_userf_box
    procedure(func), pointer, nopass :: f => null()
  end type psfun_userf_box

  private :: psfun_d_setstring, psfun_d_setreal, psfun_d_setinteger, psfun_d_setpointer

contains

  function psfun_userf_trampoline(x, user) bind(C) result(r)
    real(c_double), value :: x
    type(c_ptr), value :: user
    real(c_double) :: r
    type(psfun_userf_box), pointer :: box
    call c_f_pointer(user, box)
    r = box%f(x)
  end function psfun_userf_trampoline

  ! Make fun%f_ptr the library context's USERF function (the box must stay
  ! alive, at a fixed address, during the calls that may use it), and
  ! fun%padedegree its GENPADE degree (DGPADM's ideg, serial_mod:271).
  subroutine psfun_bind_userf(fun, ctx, box)
    type(psfun_d_serial), intent(in) :: fun
    type(c_ptr), intent(in) :: ctx
    type(psfun_userf_box), intent(inout), target :: box
    integer(c_int) :: st
    if (trim(fun%fname) == 'EXP' .and. trim(fun%variant) == 'GENPADE') &
         & st = psfun_set_pade_degree(ctx, int(fun%padedegree, c_int))
    if (trim(fun%fname) /= 'USERF' .or. .not. associated(fun%f_ptr)) return
    box%f => fun%f_ptr
    st = psfun_set_userf(ctx, c_funloc(psfun_userf_trampoline), c_loc(box))
  end subroutine psfun_bind_userf

  ! Undo psfun_bind_userf: no USERF function, GENPADE degree 6.
  subroutine psfun_unbind_userf(ctx)
    type(c_ptr), intent(in) :: ctx
    integer(c_int) :: st
    st = psfun_set_userf(ctx, c_null_funptr, c_null_ptr)
    st = psfun_set_pade_degree(ctx, 6_c_int)
  end subroutine psfun_unbind_userf

  subroutine psfun_d_setstring(fun, what, val, info)
    class(psfun_d_serial), intent(inout) :: fun
    character(len=*), intent(in) :: what, val
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(what))
    case ('FNAME')
      fun%fname = val
    case ('VARIANT')
      fun%variant = val
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_setstring

  subroutine psfun_d_setreal(fun, what, val, info)
    class(psfun_d_serial), intent(inout) :: fun
    character(len=*), intent(in) :: what
    real(psb_dpk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(what))
    case ('SCALING')
      fun%scaling = val
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_setreal

  subroutine psfun_d_setinteger(fun, what, val, info)
    class(psfun_d_serial), intent(inout) :: fun
    character(len=*), intent(in) :: what
    integer(psb_ipk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(what))
    case ('PADE_DEGREE')
      fun%padedegree = val
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_setinteger

  subroutine psfun_d_setpointer(fun, what, val, info)
    class(psfun_d_serial), intent(inout) :: fun
    character(len=*), intent(in) :: what
    procedure(func) :: val
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    select case (psb_toupper(what))
    case ('FPOINTER')
      fun%f_ptr => val
    case default
      info = psb_err_invalid_args_combination_
    end select
  end subroutine psfun_d_setpointer

  ! Device function and variant codes of this object (the select cases of
  ! serial_mod:214-505); st /= PSFUN_SUCCESS when the apply must fail.
  subroutine psfun_d_device_fn(fun, fn, var, st)
    class(psfun_d_serial), intent(in) :: fun
    integer(c_int), intent(out) :: fn, var, st
    st = PSFUN_SUCCESS
    fn = PSFUN_FN_EXP
    var = -1
    select case (trim(fun%fname))
    case ('EXP')
      select case (trim(fun%variant))
      case ('CHBSYM')
        var = PSFUN_VAR_CHBSYM
      case ('CHBHES')
        var = PSFUN_VAR_CHBHES
      case ('CHBGEN')
        var = PSFUN_VAR_CHBGEN
      case ('GENPADE')
        var = PSFUN_VAR_GENPADE
        if (fun%padedegree < 2 .or. fun%padedegree > 30) st = PSFUN_ERR_INVALID_ARG
      case ('TAYLOR')
        var = PSFUN_VAR_TAYLOR
      case ('SASQ')
        var = PSFUN_VAR_SASQ
      case default
        st = PSFUN_ERR_INVALID_ARG
      end select
    case ('USERF')
      fn = PSFUN_FN_USERF
      var = PSFUN_VAR_SYM
      if (trim(fun%variant) /= 'SYM' .or. .not. associated(fun%f_ptr)) st = PSFUN_ERR_INVALID_ARG
    case default
      st = PSFUN_ERR_UNSUPPORTED
    end select
  end subroutine psfun_d_device_fn

  ! y = f(scaling * a) x for a small dense square a, on the device of the
  ! context psb_init created.
  subroutine psfun_d_serial_apply_array(fun, a, y, x, info)
    class(psfun_d_serial), intent(inout) :: fun
    real(psb_dpk_), intent(in) :: a(:, :)
    real(psb_dpk_), intent(out) :: y(:)
    real(psb_dpk_), intent(in) :: x(:)
    integer(psb_ipk_), intent(out) :: info
    integer(c_int) :: fn, var, st, k
    real(c_double), allocatable :: h(:, :), xx(:), yy(:)
    type(psfun_userf_box), target :: box
    info = psb_success_
    if (size(a, 1) /= size(a, 2)) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psb_krylov', a_err=trim(fun%fname))
      return
    end if
    call fun%device_fn(fn, var, st)
    if (st /= PSFUN_SUCCESS .or. .not. c_associated(psb_default_ctxt%ctx)) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psb_krylov', a_err=trim(fun%fname)//' '//trim(fun%variant))
      return
    end if
    k = size(a, 1)
    h = a
    xx = x(1:k)
    allocate(yy(k))
    call psfun_bind_userf(fun, psb_default_ctxt%ctx, box)
    st = psfun_serial_apply_host(psb_default_ctxt%ctx, fn, var, fun%scaling, k, h, xx, yy)
    call psfun_unbind_userf(psb_default_ctxt%ctx)
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'psb_krylov', a_err=trim(psb_status_text(st)))
      return
    end if
    y(1:k) = yy
  end subroutine psfun_d_serial_apply_array

  ! The sparse matrix densified, then the array apply (serial_mod:526-573);
  ! repeated entries add up.
  subroutine psfun_d_serial_apply_sparse(fun, a, y, x, info)
    class(psfun_d_serial), intent(inout) :: fun
    type(psb_dspmat_type), intent(inout) :: a
    real(psb_dpk_), intent(in) :: x(:)
    real(psb_dpk_), intent(out) :: y(:)
    integer(psb_ipk_), intent(out) :: info
    real(psb_dpk_), allocatable :: amat(:, :)
    integer :: i, q, n
    if (.not. allocated(a%hrp)) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psfun_d_serial_apply_sparse', a_err='matrix not assembled')
      return
    end if
    n = int(a%m)
    allocate(amat(n, n))
    amat = dzero
    do i = 1, n
      do q = a%hrp(i) + 1, a%hrp(i + 1)
        amat(i, a%hci(q) + 1) = amat(i, a%hci(q) + 1) + a%hva(q)
      end do
    end do
    call fun%apply(amat, y, x, info)
  end subroutine psfun_d_serial_apply_sparse

end module psfun_d_serial_mod
