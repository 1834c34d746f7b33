! psfun_utils_mod.f90 -- PSFUN's utility functions (utils/psfun_utils_mod.f90:
! ellipkkp, ellipj, horner) with the reference's generic names and
! arguments, computed by the library's host routines (psfun_ellipkkp,
! psfun_ellipj, psfun_horner; utils/psfun_d_utils_mod.f90:37-138,
! utils/psfun_z_utils_mod.f90:39-107).
module psfun_utils_mod
  use iso_c_binding
  use psfun_hip_mod
  use psb_base_mod
  implicit none

  interface ellipkkp
    module procedure psfun_ellipkkp_f
  end interface ellipkkp

  interface ellipj
    module procedure d_ellipj, z_ellipj
  end interface ellipj

  interface horner
    module procedure psfun_horner_f
  end interface horner

contains

  ! K(m) and K(1 - m) at m = exp(-2 pi L)
  function psfun_ellipkkp_f(L) result(K)
    real(psb_dpk_), intent(in) :: L
    real(psb_dpk_) :: K(2)
    integer(c_int) :: st
    st = psfun_ellipkkp(L, K)
  end function psfun_ellipkkp_f

  ! Jacobi sn, cn, dn at real u, m = exp(-2 pi L)
  subroutine d_ellipj(u, L, sn, cn, dn)
    real(psb_dpk_), intent(in) :: u, L
    real(psb_dpk_), intent(out) :: sn, cn, dn
    real(c_double) :: out(6)
    integer(c_int) :: st
    st = psfun_ellipj(u, 0.0_c_double, L, out)
    sn = out(1)
    cn = out(3)
    dn = out(5)
  end subroutine d_ellipj

  ! Jacobi sn, cn, dn at complex u (flag is accepted for the reference's
  ! interface; the descending Landen recursion needs no caller state)
  subroutine z_ellipj(u, L, sn, cn, dn, flag)
    complex(psb_dpk_), intent(in) :: u
    real(psb_dpk_), intent(in) :: L
    complex(psb_dpk_), intent(out) :: sn, cn, dn
    logical, optional, intent(in) :: flag
    real(c_double) :: out(6)
    integer(c_int) :: st
    st = psfun_ellipj(real(u, c_double), aimag(u), L, out)
    sn = cmplx(out(1), out(2), psb_dpk_)
    cn = cmplx(out(3), out(4), psb_dpk_)
    dn = cmplx(out(5), out(6), psb_dpk_)
  end subroutine z_ellipj

  ! sum_i coeffs(i) x^(i-1)
  function psfun_horner_f(coeffs, x) result(res)
    real(psb_dpk_), intent(in) :: coeffs(:)
    real(psb_dpk_), intent(in) :: x
    real(psb_dpk_) :: res
    real(c_double), allocatable :: c(:)
    c = coeffs
    res = psfun_horner(int(size(c), c_int), c, x)
  end function psfun_horner_f

end module psfun_utils_mod
