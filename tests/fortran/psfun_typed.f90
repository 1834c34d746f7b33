! psfun_typed.f90 -- PSFUN's example drivers written against the typed
! Fortran layer (psfun_amd/fortran: psb_base_mod, psfun_d_krylov_mod, ...),
! in the call shapes of the reference's examples, so that the parity tests
! exercise exactly what a PSFUN program calls:
!   psfun_typed krylov  <deck> [KNAME] [keep]  examples/polynomial/lanczostest.F90:76-255
!                                              (KNAME default LANCZOS; without
!                                              "keep" eps/itmax/itrace/istop are
!                                              overridden as lines 211-215 do)
!   psfun_typed shifted <deck>                 examples/shiftedkrylov/shiftedtest.f90
!   psfun_typed serial  <deck>                 examples/serial/serialtest.f90
!   psfun_typed quad <matrix> HHT|TALBOT <N> <p1> [p2] [PREC=<ptype>] [ITMAX=<n>]
!                                              psfun_z_quadrature: HHT sqrt(z) on
!                                              [p1, p2], or Talbot exp(-p1 z),
!                                              applied to x = 1; with PREC, a
!                                              preconditioner built on A and
!                                              given by quad%set(prec)
! Decks are read as the reference's get_parms reads them (list-directed, the
! line count first).  Machine-readable results follow lines starting '@'.
! Functions handed to the quadrature rules as procedure pointers (module
! procedures, so no internal-procedure trampoline is needed).
module psfun_typed_funcs
  use psb_base_mod
  implicit none
contains
  function zsqrt(z) result(r)
    complex(psb_dpk_), intent(in) :: z
    complex(psb_dpk_) :: r
    r = sqrt(z)
  end function zsqrt

  function zexpm(z) result(r)
    complex(psb_dpk_), intent(in) :: z
    complex(psb_dpk_) :: r
    r = exp(-z)
  end function zexpm

end module psfun_typed_funcs

program psfun_typed
  use psfun_typed_funcs
  use psb_base_mod
  use psb_util_mod
  use psb_prec_mod
  use psfun_d_serial_mod
  use psfun_d_krylov_mod
  use psfun_krylov_mod
  use psfun_quadrature_mod
  implicit none
  character(len=256) :: mode
  type(psb_ctxt_type) :: ctxt
  integer(psb_ipk_) :: iam, np

  call psb_init(ctxt)
  call psb_info(ctxt, iam, np)
  if (iam < 0) then
    call psb_exit(ctxt)
    error stop 2
  end if
  call get_command_argument(1, mode)
  select case (trim(mode))
  case ('krylov')
    call run_krylov()
  case ('shifted')
    call run_shifted()
  case ('serial')
    call run_serial()
  case ('quad')
    call run_quad()
  case default
    write(psb_err_unit, *) 'usage: psfun_typed krylov|shifted|serial|quad ...'
    call psb_exit(ctxt)
    error stop 2
  end select
  call psb_exit(ctxt)

contains

  subroutine print_vec(tag, v)
    character(len=*), intent(in) :: tag
    real(psb_dpk_), intent(in) :: v(:)
    integer :: i
    do i = 1, size(v)
      write(psb_out_unit, '(a,1x,es24.16)') tag, v(i)
    end do
  end subroutine print_vec

  subroutine run_krylov()
    character(len=256) :: deck, kname, keep
    character(len=20) :: mname, rhs_file, filefmt, part, afmt, fname, variant
    real(psb_dpk_) :: scaling, eps, err
    integer(psb_ipk_) :: itmax, itrace, istop, iter, info, ip, m_problem
    type(psb_ldspmat_type) :: aux_a
    type(psb_dspmat_type) :: a
    type(psb_desc_type) :: desc_a
    type(psb_d_vect_type) :: x_col, y_col
    type(psfun_d_serial) :: fun
    type(psfun_d_krylov) :: kmethd
    real(psb_dpk_), allocatable :: aux_x(:, :), x_glob(:), res(:), y(:)
    real(psb_dpk_) :: t1, t2

    call get_command_argument(2, deck)
    kname = 'LANCZOS'
    if (command_argument_count() >= 3) call get_command_argument(3, kname)
    keep = ''
    if (command_argument_count() >= 4) call get_command_argument(4, keep)
    open(30, file=trim(deck), action='read')
    read(30, *) ip
    read(30, *) mname
    read(30, *) rhs_file
    read(30, *) filefmt
    read(30, *) part
    read(30, *) afmt
    read(30, *) fname
    read(30, *) variant
    read(30, *) scaling
    eps = 1e-6
    itmax = 100
    itrace = 1
    istop = 1
    if (ip == 12) then
      read(30, *) eps
      read(30, *) itmax
      read(30, *) itrace
      read(30, *) istop
    end if
    close(30)
    if (trim(keep) /= 'keep') then   ! lanczostest.F90:211-215
      eps = 1e-6
      itmax = 100
      itrace = 1
      istop = 1
    end if

    if (psb_toupper(trim(filefmt)) == 'HB') then
      call hb_read(aux_a, info, iunit=12, b=aux_x, filename=mname)
    else
      call mm_mat_read(aux_a, info, iunit=12, filename=mname)
    end if
    if (info /= psb_success_) then
      call psb_error(ctxt)
      error stop 3
    end if
    m_problem = int(aux_a%get_nrows())
    if (allocated(aux_x)) then          ! the HB file's own right-hand side
      x_glob = aux_x(:, 1)
    else if (trim(rhs_file) /= 'NONE') then
      call mm_array_read(aux_x, info, iunit=12, filename=rhs_file)
      x_glob = aux_x(:, 1)
    else
      allocate(x_glob(m_problem))
      x_glob = done
    end if
    call psb_matdist(aux_a, a, ctxt, desc_a, info, fmt=afmt, parts=part_block)
    call psb_scatter(x_glob, x_col, desc_a, info, root=psb_root_)
    call psb_geall(y_col, desc_a, info)
    call y_col%zero()
    call psb_geasb(y_col, desc_a, info)

    call fun%set("FNAME", fname, info)
    call fun%set("VARIANT", variant, info)
    call fun%set("SCALING", scaling, info)
    call kmethd%set("KNAME", trim(kname), info)

    t1 = psb_wtime()
    call kmethd%apply(fun, a, desc_a, y_col, x_col, eps, info, itmax, itrace, istop, iter, err, res)
    t2 = psb_wtime() - t1
    write(psb_out_unit, '("Time to solve system          : ",es12.5)') t2
    write(psb_out_unit, '("Number of iterations          : ",i12)') iter
    write(psb_out_unit, '("Convergence indicator on exit : ",es12.5)') err
    write(psb_out_unit, '("Info  on exit                 : ",i12)') info
    write(psb_out_unit, '("Storage format for               A: ",a)') a%get_fmt()
    write(psb_out_unit, '(a,1x,i0)') '@INFO', info
    if (info == psb_success_) then
      write(psb_out_unit, '(a,1x,i0)') '@ITER', iter
      write(psb_out_unit, '(a,1x,es24.16)') '@ERR', err
      call print_vec('@RES', res)
      y = y_col%get_vect()
      call print_vec('@Y', y)
    end if
    call psb_gefree(x_col, desc_a, info)
    call psb_gefree(y_col, desc_a, info)
    call psb_spfree(a, desc_a, info)
    call psb_cdfree(desc_a, info)
  end subroutine run_krylov

  subroutine run_shifted()
    character(len=256) :: deck
    character(len=40) :: mtrx_file, rhs_file, filefmt, kmethd, ptype, afmt, part
    real(psb_dpk_) :: eta, zeta, eps, err, cond, resmx, resmxp
    integer(psb_ipk_) :: ip, istopc, itmax, itrace, irst, iter, info, m_problem
    type(psb_ldspmat_type) :: aux_a
    type(psb_dspmat_type) :: a
    type(psb_desc_type) :: desc_a
    type(psb_d_vect_type) :: b_col, x_col, r_col
    type(psb_dprec_type) :: prec
    real(psb_dpk_), allocatable :: b_glob(:), aux_b(:, :), xg(:)

    call get_command_argument(2, deck)
    open(30, file=trim(deck), action='read')
    read(30, *) ip
    read(30, *) mtrx_file
    read(30, *) rhs_file
    read(30, *) eta
    read(30, *) zeta
    read(30, *) filefmt
    read(30, *) kmethd
    read(30, *) ptype
    read(30, *) afmt
    read(30, *) part
    read(30, *) istopc
    read(30, *) itmax
    read(30, *) itrace
    read(30, *) irst
    read(30, *) eps
    close(30)

    call mm_mat_read(aux_a, info, iunit=12, filename=mtrx_file)
    if (info /= psb_success_) then
      call psb_error(ctxt)
      error stop 3
    end if
    m_problem = int(aux_a%get_nrows())
    if (trim(rhs_file) /= 'NONE') then
      call mm_array_read(aux_b, info, iunit=12, filename=rhs_file)
      b_glob = aux_b(:, 1)
    else
      allocate(b_glob(m_problem))
      b_glob = done
    end if
    call psb_matdist(aux_a, a, ctxt, desc_a, info, fmt=afmt, parts=part_block)
    call psb_scatter(b_glob, b_col, desc_a, info, root=psb_root_)
    call psb_geall(x_col, desc_a, info)
    call x_col%zero()
    call psb_geasb(x_col, desc_a, info)
    call psb_geall(r_col, desc_a, info)
    call r_col%zero()
    call psb_geasb(r_col, desc_a, info)

    call prec%init(ctxt, ptype, info)
    call prec%build(a, desc_a, info)
    if (info /= psb_success_) then
      call psb_error(ctxt)
      error stop 3
    end if
    cond = dzero
    call psb_krylov(kmethd, a, prec, b_col, eta, zeta, x_col, eps, desc_a, info, itmax=itmax, iter=iter, err=err, &
         & itrace=itrace, istop=istopc, irst=irst, cond=cond)
    ! the driver's residual uses the unshifted A (shiftedtest.f90:248-251)
    call psb_geaxpby(done, b_col, dzero, r_col, desc_a, info)
    call psb_spmm(-done, a, x_col, done, r_col, desc_a, info)
    resmx = psb_genrm2(r_col, desc_a, info)
    resmxp = psb_geamax(r_col, desc_a, info)
    write(psb_out_unit, '("Iterations to convergence: ",i6)') iter
    write(psb_out_unit, '("Error estimate on exit   : ",es12.5)') err
    write(psb_out_unit, '("Residual norm 2          : ",es12.5)') resmx
    write(psb_out_unit, '("Residual norm inf        : ",es12.5)') resmxp
    write(psb_out_unit, '("Condition number         : ",es12.5)') cond
    write(psb_out_unit, '(a,1x,i0)') '@INFO', info
    write(psb_out_unit, '(a,1x,i0)') '@ITER', iter
    write(psb_out_unit, '(a,1x,es24.16)') '@ERR', err
    write(psb_out_unit, '(a,1x,es24.16)') '@COND', cond
    write(psb_out_unit, '(a,1x,es24.16)') '@RESN2', resmx
    write(psb_out_unit, '(a,1x,es24.16)') '@RESNI', resmxp
    call psb_gather(xg, x_col, desc_a, info, root=psb_root_)
    call print_vec('@Y', xg)
    call prec%free(info)
    call psb_gefree(b_col, desc_a, info)
    call psb_gefree(x_col, desc_a, info)
    call psb_gefree(r_col, desc_a, info)
    call psb_spfree(a, desc_a, info)
    call psb_cdfree(desc_a, info)
  end subroutine run_shifted

  subroutine run_serial()
    character(len=256) :: deck
    character(len=20) :: mname, fname, variant
    character(len=100) :: dumpchar
    real(psb_dpk_) :: scaling
    integer(psb_ipk_) :: ip, info, n
    type(psb_dspmat_type) :: a
    type(psfun_d_serial) :: fun
    real(psb_dpk_), allocatable :: x(:), y(:)

    call get_command_argument(2, deck)
    open(30, file=trim(deck), action='read')
    read(30, *) ip
    read(30, *) mname
    read(30, *) fname
    read(30, *) variant
    read(30, *) scaling
    read(30, *) dumpchar
    close(30)
    call mm_mat_read(a, info, iunit=12, filename=mname)
    if (info /= psb_success_) then
      call psb_error(ctxt)
      error stop 3
    end if
    n = a%get_nrows()
    write(psb_out_unit, '("Matrix of size matrix ",i6,"x",i6," nnz =",i6)') n, n, a%get_nzeros()
    call fun%set("FNAME", fname, info)
    call fun%set("VARIANT", variant, info)
    call fun%set("SCALING", scaling, info)
    allocate(x(n), y(n))
    x = done
    call fun%apply(a, y, x, info)
    write(psb_out_unit, '(a,1x,i0)') '@INFO', info
    if (info == psb_success_) call print_vec('@Y', y)
  end subroutine run_serial

  subroutine run_quad()
    character(len=256) :: mname, rule, arg
    integer(psb_ipk_) :: N, info, j
    real(psb_dpk_) :: p1, p2
    type(psb_ldspmat_type) :: aux_a
    type(psb_dspmat_type), target :: a
    type(psb_desc_type) :: desc_a
    type(psb_d_vect_type) :: x_col, y_col
    type(psfun_z_quadrature) :: quad
    procedure(zquadfun), pointer :: zfun
    procedure(zquadrule), pointer :: quadformula
    real(psb_dpk_), allocatable :: xg(:), y(:), resid(:)
    integer(psb_ipk_), allocatable :: iters(:)
    type(psb_dprec_type), target :: prec
    character(len=20) :: ptype
    integer(psb_ipk_) :: itmax, k

    call get_command_argument(2, mname)
    call get_command_argument(3, rule)
    call get_command_argument(4, arg)
    read(arg, *) N
    call get_command_argument(5, arg)
    read(arg, *) p1
    p2 = p1
    ptype = ''
    itmax = 2000
    do k = 6, command_argument_count()
      call get_command_argument(k, arg)
      if (arg(1:5) == 'PREC=') then
        ptype = arg(6:)
      else if (arg(1:6) == 'ITMAX=') then
        read(arg(7:), *) itmax
      else
        read(arg, *) p2
      end if
    end do
    call mm_mat_read(aux_a, info, iunit=12, filename=mname)
    call psb_matdist(aux_a, a, ctxt, desc_a, info, fmt='CSR', parts=part_block)
    allocate(xg(desc_a%get_local_rows()))
    xg = done
    call psb_scatter(xg, x_col, desc_a, info)
    call psb_geall(y_col, desc_a, info)
    call y_col%zero()
    call psb_geasb(y_col, desc_a, info)
    if (trim(rule) == 'HHT') then
      zfun => zsqrt
      quadformula => hhtmethod1
      call quad%computepoles(quadformula, zfun, N, info, rparams=[p1, p2])
    else
      zfun => zexpm
      quadformula => talbot
      call quad%computepoles(quadformula, zfun, N, info, rparams=[p1])
    end if
    write(psb_out_unit, '(a,1x,i0)') '@POLES_INFO', info
    do j = 1, N
      write(psb_out_unit, '(a,4(1x,es24.16))') '@POLE', real(quad%xi(j)), aimag(quad%xi(j)), real(quad%c(j)), &
           & aimag(quad%c(j))
    end do
    write(psb_out_unit, '(a,2(1x,es24.16))') '@ETA_SIGN', quad%eta, quad%sign
    call quad%set(a)
    if (len_trim(ptype) > 0) then
      call prec%init(ctxt, ptype, info)
      call prec%build(a, desc_a, info)
      call quad%set(prec)
    end if
    call quad%apply(x_col, y_col, info, eps=1.0e-12_psb_dpk_, itmax=itmax, iters=iters, resid=resid)
    write(psb_out_unit, '(a,1x,i0)') '@INFO', info
    write(psb_out_unit, '(a,1x,i0)') '@NSOLVE', size(iters)
    if (info == 0) then
      write(psb_out_unit, '(a,1x,i0)') '@MAXITER', maxval(iters)
      write(psb_out_unit, '(a,1x,es24.16)') '@MAXRESID', maxval(resid)
    end if
    if (len_trim(ptype) > 0) call prec%free(info)
    y = y_col%get_vect()
    call print_vec('@Y', y)
    call psb_gefree(x_col, desc_a, info)
    call psb_gefree(y_col, desc_a, info)
    call psb_spfree(a, desc_a, info)
  end subroutine run_quad

end program psfun_typed
