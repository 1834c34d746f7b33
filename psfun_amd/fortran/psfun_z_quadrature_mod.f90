! psfun_z_quadrature_mod.f90 -- PSFUN's rational-quadrature object
! (quadratures/psfun_z_quadrature_mod.F90:42-166, psfun_base_quadrature_mod.f90:39-43)
! over the gfx950 library, with the apply the reference documents
! (docs/pages/algorithms.md:106-125) but does not implement.
!
!   call quad%computepoles(quadformula, zfun, N, info, cparams, rparams)
!   call quad%set(a)          ! setmatrix
!   call quad%set(prec)       ! setpreconditioner
!   call quad%apply(x, y, info, eps, itmax, iters, resid)
!
! The rules are procedures of the reference's abstract interface zquadrule:
! hhtmethod1 (HHT 2008 method 1, quadratures/psfun_z_quadrules_mod.f90:40-85)
! and talbot (the modified Talbot rule for exp(-t z), rparams(1) = t; no
! reference counterpart).  computepoles tests cparams before rparams, as the
! reference does (lines 130-138), so its "both present" branch is unreachable
! there and here.
!
! The stored rule (xi, c, eta, sign) means the shifted systems
! (xi_j I + sign A) v_j = x.  apply solves all of them with one multi-shift
! CG on the device (psfun_quad_apply; with a built preconditioner set, the
! poles left above eps continue by preconditioned COCG, psfun_quad_apply_prec)
! and assembles
!   hhtmethod1:  y = eta * A * Im(sum_j c_j v_j)
!   other rules: y = eta * Re(sum_j c_j v_j)     (complex-conjugate pole pairs
!                                                 folded into one solve)
module psfun_base_quadrature_mod
  use psb_base_mod
  implicit none
  type, public :: psfun_quadrature
    type(psb_desc_type) :: desc_a
  end type psfun_quadrature
end module psfun_base_quadrature_mod

module psfun_z_quadrature_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  use psb_prec_mod
  use psfun_base_quadrature_mod
  implicit none

  integer, parameter :: psfun_form_resolvent = 0, psfun_form_hht1 = 1

  type, extends(psfun_quadrature) :: psfun_z_quadrature
    complex(psb_dpk_), allocatable, dimension(:) :: xi
    complex(psb_dpk_), allocatable, dimension(:) :: c
    real(psb_dpk_) :: eta = dzero
    real(psb_dpk_) :: sign = dzero
    type(psb_dspmat_type), pointer :: a => null()
    type(psb_dprec_type), pointer :: prec => null()
    integer :: form = psfun_form_resolvent
  contains
    procedure, pass(quad) :: computepoles => psfun_z_computepoles
    procedure, pass(quad) :: setmatrix => psfun_z_setmatrix
    procedure, pass(quad) :: setpreconditioner => psfun_z_setpreconditioner
    generic, public :: set => setmatrix, setpreconditioner
    procedure, pass(quad) :: plot => psfun_z_quadratureplot
    procedure, pass(quad) :: apply => psfun_z_quadrature_apply
  end type psfun_z_quadrature

  abstract interface
    function zquadfun(z) result(res)
      import :: psb_dpk_
      complex(psb_dpk_), intent(in) :: z
      complex(psb_dpk_) :: res
    end function zquadfun
  end interface

  abstract interface
    subroutine zquadrule(zfun, xi, c, eta, sign, N, info, cparams, rparams)
      import :: psb_dpk_, psb_ipk_, zquadfun
      complex(psb_dpk_), allocatable, dimension(:), intent(out) :: xi
      complex(psb_dpk_), allocatable, dimension(:), intent(out) :: c
      real(psb_dpk_), intent(out) :: eta
      real(psb_dpk_), intent(out) :: sign
      procedure(zquadfun), pointer, intent(in) :: zfun
      integer(psb_ipk_), intent(in) :: N
      integer(psb_ipk_), intent(out) :: info
      complex(psb_dpk_), dimension(:), optional, intent(in) :: cparams
      real(psb_dpk_), dimension(:), optional, intent(in) :: rparams
    end subroutine zquadrule
  end interface

  ! Carrier of the function to integrate across the C boundary.
  type :: zfun_box
    procedure(zquadfun), pointer, nopass :: f => null()
  end type zfun_box

contains

  subroutine zfun_trampoline(z, fz, user) bind(C)
    real(c_double), intent(in) :: z(2)
    real(c_double), intent(out) :: fz(2)
    type(c_ptr), value :: user
    type(zfun_box), pointer :: box
    complex(psb_dpk_) :: v
    call c_f_pointer(user, box)
    v = box%f(cmplx(z(1), z(2), psb_dpk_))
    fz(1) = real(v, c_double)
    fz(2) = aimag(v)
  end subroutine zfun_trampoline

  ! ---- rules ------------------------------------------------------------------
  ! HHT 2008 method 1 for f = zfun on the spectrum interval rparams(1:2).
  subroutine hhtmethod1(zfun, xi, c, eta, sign, N, info, cparams, rparams)
    procedure(zquadfun), pointer, intent(in) :: zfun
    complex(psb_dpk_), allocatable, dimension(:), intent(out) :: xi
    complex(psb_dpk_), allocatable, dimension(:), intent(out) :: c
    real(psb_dpk_), intent(out) :: eta
    real(psb_dpk_), intent(out) :: sign
    integer(psb_ipk_), intent(in) :: N
    integer(psb_ipk_), intent(out) :: info
    complex(psb_dpk_), dimension(:), optional, intent(in) :: cparams
    real(psb_dpk_), dimension(:), optional, intent(in) :: rparams
    type(zfun_box), target :: box
    real(c_double) :: rp(2)
    real(c_double), allocatable :: xr(:), cr(:)
    integer(c_int) :: st
    integer :: j
    info = psb_success_
    eta = dzero
    sign = dzero
    if (.not. present(rparams)) then
      info = psb_err_invalid_args_combination_
      return
    end if
    if (size(rparams) < 2 .or. .not. associated(zfun)) then
      info = psb_err_invalid_args_combination_
      return
    end if
    box%f => zfun
    rp = rparams(1:2)
    allocate(xr(2 * N), cr(2 * N))
    st = psfun_hhtmethod1(c_funloc(zfun_trampoline), c_loc(box), N, rp, xr, cr, eta, sign)
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_
      return
    end if
    allocate(xi(N), c(N))
    do j = 1, N
      xi(j) = cmplx(xr(2 * j - 1), xr(2 * j), psb_dpk_)
      c(j) = cmplx(cr(2 * j - 1), cr(2 * j), psb_dpk_)
    end do
  end subroutine hhtmethod1

  ! Modified Talbot rule for exp(-t z), t = rparams(1) (default 1); zfun is
  ! not evaluated (the rule is specific to exp).
  subroutine talbot(zfun, xi, c, eta, sign, N, info, cparams, rparams)
    procedure(zquadfun), pointer, intent(in) :: zfun
    complex(psb_dpk_), allocatable, dimension(:), intent(out) :: xi
    complex(psb_dpk_), allocatable, dimension(:), intent(out) :: c
    real(psb_dpk_), intent(out) :: eta
    real(psb_dpk_), intent(out) :: sign
    integer(psb_ipk_), intent(in) :: N
    integer(psb_ipk_), intent(out) :: info
    complex(psb_dpk_), dimension(:), optional, intent(in) :: cparams
    real(psb_dpk_), dimension(:), optional, intent(in) :: rparams
    real(c_double) :: t
    real(c_double), allocatable :: xr(:), cr(:)
    integer(c_int) :: st
    integer :: j
    info = psb_success_
    t = done
    if (present(rparams)) then
      if (size(rparams) >= 1) t = rparams(1)
    end if
    allocate(xr(2 * N), cr(2 * N))
    st = psfun_talbot(N, t, xr, cr, eta, sign)
    if (st /= PSFUN_SUCCESS) then
      info = psb_err_from_subroutine_
      return
    end if
    allocate(xi(N), c(N))
    do j = 1, N
      xi(j) = cmplx(xr(2 * j - 1), xr(2 * j), psb_dpk_)
      c(j) = cmplx(cr(2 * j - 1), cr(2 * j), psb_dpk_)
    end do
  end subroutine talbot

  ! ---- the object ----------------------------------------------------------------
  subroutine psfun_z_computepoles(quad, quadformula, zfun, N, info, cparams, rparams)
    class(psfun_z_quadrature), intent(inout) :: quad
    procedure(zquadrule), pointer, intent(in) :: quadformula
    procedure(zquadfun), pointer, intent(in) :: zfun
    integer(psb_ipk_), intent(in) :: N
    integer(psb_ipk_), intent(out) :: info
    complex(psb_dpk_), dimension(:), optional, intent(in) :: cparams
    real(psb_dpk_), dimension(:), optional, intent(in) :: rparams
    info = 0
    if (present(cparams)) then
      call quadformula(zfun, quad%xi, quad%c, quad%eta, quad%sign, N, info, cparams=cparams)
    else if (present(rparams)) then
      call quadformula(zfun, quad%xi, quad%c, quad%eta, quad%sign, N, info, rparams=rparams)
    else if ((present(cparams)) .and. (present(rparams))) then
      call quadformula(zfun, quad%xi, quad%c, quad%eta, quad%sign, N, info, cparams, rparams)
    else
      call quadformula(zfun, quad%xi, quad%c, quad%eta, quad%sign, N, info)
    end if
    quad%form = psfun_form_resolvent
    if (associated(quadformula, hhtmethod1)) quad%form = psfun_form_hht1
  end subroutine psfun_z_computepoles

  subroutine psfun_z_setmatrix(quad, a)
    class(psfun_z_quadrature), intent(inout) :: quad
    type(psb_dspmat_type), target :: a
    quad%a => a
  end subroutine psfun_z_setmatrix

  subroutine psfun_z_setpreconditioner(quad, prec)
    class(psfun_z_quadrature), intent(inout) :: quad
    type(psb_dprec_type), target :: prec
    quad%prec => prec
  end subroutine psfun_z_setpreconditioner

  ! The reference plots with gnuplot only when built WITHGNUPLOTFORTRAN; here
  ! the poles and weights are listed on the output unit instead.
  subroutine psfun_z_quadratureplot(quad, zfun, info, filename)
    class(psfun_z_quadrature), intent(inout) :: quad
    procedure(zquadfun), pointer, intent(in) :: zfun
    integer(psb_ipk_), intent(out) :: info
    character(len=*), optional, intent(in) :: filename
    integer :: j
    info = psb_success_
    if (.not. allocated(quad%xi)) then
      info = psb_err_invalid_args_combination_
      return
    end if
    do j = 1, size(quad%xi)
      write(psb_out_unit, '(i4,4(1x,es24.16))') j, real(quad%xi(j)), aimag(quad%xi(j)), real(quad%c(j)), &
           & aimag(quad%c(j))
    end do
  end subroutine psfun_z_quadratureplot

  ! y = f(A) x with the stored rule (see the module header).  iters/resid
  ! (optional, allocated here) hold each solved shift's iteration count and
  ! final relative residual.
  subroutine psfun_z_quadrature_apply(quad, x, y, info, eps, itmax, iters, resid)
    class(psfun_z_quadrature), intent(inout) :: quad
    type(psb_d_vect_type), intent(inout) :: x, y
    integer(psb_ipk_), intent(out) :: info
    real(psb_dpk_), optional, intent(in) :: eps
    integer(psb_ipk_), optional, intent(in) :: itmax
    integer(psb_ipk_), optional, allocatable, intent(out) :: iters(:)
    real(psb_dpk_), optional, allocatable, intent(out) :: resid(:)
    complex(psb_dpk_), allocatable :: s(:), w(:)
    real(c_double), allocatable :: sh(:), ww(:), rr(:)
    integer(c_int), allocatable :: it(:)
    logical, allocatable :: used(:)
    real(c_double) :: eps_, gamma, scale, tol, d, best
    integer(c_int) :: itmax_, st, form, ns
    integer :: i, j, jb, n
    info = psb_success_
    if (.not. associated(quad%a) .or. .not. allocated(quad%xi)) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psfun_z_quadrature_apply')
      return
    end if
    if (.not. c_associated(x%v) .or. .not. c_associated(y%v)) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psfun_z_quadrature_apply')
      return
    end if
    eps_ = 1.0e-8_c_double
    if (present(eps)) eps_ = eps
    itmax_ = 1000
    if (present(itmax)) itmax_ = itmax
    n = size(quad%xi)
    ! (xi I + sign A) = sign (A + sign xi I)
    allocate(s(n), w(n))
    s = quad%sign * quad%xi
    w = quad%c
    gamma = quad%eta * quad%sign
    ns = 0
    if (quad%form == psfun_form_hht1) then
      form = PSFUN_QUAD_A_IM_SUM
      ns = n
    else
      ! fold (s, w), (conj s, conj w) pairs into (s, 2 w) for the Re(...) form
      form = PSFUN_QUAD_RE_SUM
      allocate(used(n))
      used = .false.
      tol = 1.0e-13_c_double
      do i = 1, n
        if (used(i)) cycle
        used(i) = .true.
        ns = ns + 1
        s(ns) = quad%sign * quad%xi(i)
        w(ns) = quad%c(i)
        scale = max(abs(s(ns)), 1.0e-300_c_double)
        if (abs(aimag(s(ns))) > tol * scale) then
          jb = 0
          best = huge(best)
          do j = 1, n
            if (used(j)) cycle
            d = abs(quad%sign * quad%xi(j) - conjg(s(ns))) + &
                 & abs(quad%c(j) - conjg(w(ns))) / max(abs(w(ns)), 1.0e-300_c_double) * scale
            if (d < best) then
              best = d
              jb = j
            end if
          end do
          if (jb > 0 .and. best <= 10 * tol * scale) then
            used(jb) = .true.
            w(ns) = 2 * w(ns)
          end if
        end if
      end do
    end if
    allocate(sh(2 * ns), ww(2 * ns), it(ns), rr(ns))
    do i = 1, ns
      sh(2 * i - 1) = real(s(i), c_double)
      sh(2 * i) = aimag(s(i))
      ww(2 * i - 1) = real(w(i), c_double)
      ww(2 * i) = aimag(w(i))
    end do
    if (associated(quad%prec)) then
      if (.not. c_associated(quad%prec%p)) then
        info = psb_err_invalid_args_combination_
        call psb_errpush(info, 'psfun_z_quadrature_apply', a_err='preconditioner not built')
        return
      end if
      st = psfun_quad_apply_prec(quad%a%ctxt%ctx, quad%a%a, quad%prec%p, ns, sh, ww, gamma, form, x%v, y%v, &
           & eps_, itmax_, it, rr)
    else
      st = psfun_quad_apply(quad%a%ctxt%ctx, quad%a%a, ns, sh, ww, gamma, form, x%v, y%v, eps_, itmax_, it, rr)
    end if
    if (st /= PSFUN_SUCCESS) then
      info = psb_status_to_info(st)
      call psb_errpush(info, 'psfun_z_quadrature_apply', a_err=trim(psb_status_text(st)))
      return
    end if
    if (present(iters)) iters = it
    if (present(resid)) resid = rr
  end subroutine psfun_z_quadrature_apply

end module psfun_z_quadrature_mod

! The reference's umbrella module of the quadrature objects.
module psfun_quadrature_mod
  use psfun_base_quadrature_mod
  use psfun_z_quadrature_mod
end module psfun_quadrature_mod
