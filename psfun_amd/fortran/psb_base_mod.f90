! psb_base_mod.f90 -- the part of PSBLAS's base module that PSFUN's f(A)x path
! and its example drivers use, written for one process on one GPU of the
! gfx950 library (libpsfun_hip.so, through psfun_hip_mod).
!
! PSBLAS itself is not part of the PSFUN tree.  This module keeps the names,
! argument order and info-code conventions PSFUN's sources call
! (krylov/polynomial/psfun_d_lanczos.f90:97-386, krylov/shifted/psfun_dcg.F90,
! examples/polynomial/lanczostest.F90:76-255) so that code written against
! PSBLAS compiles against it:
!   kinds          psb_dpk_, psb_ipk_, psb_lpk_, psb_mpk_, psb_epk_
!   environment    psb_ctxt_type, psb_init, psb_info, psb_exit, psb_abort,
!                  psb_barrier, psb_wtime, psb_bcast, psb_amx, psb_sum
!   descriptor     psb_desc_type, psb_cdall, psb_cdasb, psb_cdfree
!   sparse matrix  psb_dspmat_type (a device CSR handle), psb_spall,
!                  psb_spins, psb_spasb, psb_spfree
!   dense vector   psb_d_vect_type (a device pointer), psb_geall,
!                  psb_geins, psb_geasb, psb_gefree, psb_scatter, psb_gather
!   computation    psb_spmm, psb_gedot, psb_genrm2, psb_geamax, psb_geaxpby
!   errors         psb_errpush, psb_erractionsave/restore, psb_error_handler,
!                  psb_get_errstatus, psb_error, psb_toupper
! With one process every rank-local size equals the global one; collectives
! are identities.  All vector and matrix data live in device memory; the
! host keeps only what a build step stages (psb_spins/psb_geins triplets).
module psb_base_mod
  use iso_c_binding
  use psfun_hip_mod
  implicit none

  ! ---- kinds ---------------------------------------------------------------
  integer, parameter :: psb_dpk_ = c_double
  integer, parameter :: psb_spk_ = c_float
  integer, parameter :: psb_ipk_ = c_int32_t
  integer, parameter :: psb_mpk_ = c_int32_t
  integer, parameter :: psb_lpk_ = c_int64_t
  integer, parameter :: psb_epk_ = c_int64_t

  real(psb_dpk_), parameter :: dzero = 0.0_psb_dpk_, done = 1.0_psb_dpk_
  complex(psb_dpk_), parameter :: zzero = (0.0_psb_dpk_, 0.0_psb_dpk_), zone = (1.0_psb_dpk_, 0.0_psb_dpk_)
  integer(psb_ipk_), parameter :: izero = 0, ione = 1
  integer(psb_ipk_), parameter :: psb_root_ = 0
  integer, parameter :: psb_inp_unit = 5, psb_out_unit = 6, psb_err_unit = 0

  ! ---- info codes (the names PSFUN compares against; the values are this
  ! module's own, equal to the Python mirror's psfun_amd/psb.py) ------------
  integer(psb_ipk_), parameter :: psb_success_ = 0
  integer(psb_ipk_), parameter :: psb_err_alloc_dealloc_ = 4000
  integer(psb_ipk_), parameter :: psb_err_from_subroutine_ = 4010
  integer(psb_ipk_), parameter :: psb_err_from_subroutine_non_ = 4011
  integer(psb_ipk_), parameter :: psb_err_unsupported_ = 4050
  integer(psb_ipk_), parameter :: psb_err_invalid_vect_state_ = 1124
  integer(psb_ipk_), parameter :: psb_err_invalid_args_combination_ = 1126
  integer(psb_ipk_), parameter :: psb_err_invalid_istop_ = 5001
  integer(psb_ipk_), parameter :: psb_act_ret_ = 0, psb_act_abort_ = 1

  ! ---- environment ---------------------------------------------------------
  type :: psb_ctxt_type
    type(c_ptr) :: ctx = c_null_ptr       ! libpsfun_hip context (psfun_ctx_t)
    integer(psb_ipk_) :: device = -1
  end type psb_ctxt_type

  ! ---- descriptor: one process owns every row -------------------------------
  type :: psb_desc_type
    type(psb_ctxt_type) :: ctxt
    integer(psb_lpk_) :: n = -1
    logical :: asb = .false.
  contains
    procedure, pass(desc) :: get_context => psb_cd_get_context
    procedure, pass(desc) :: get_global_rows => psb_cd_get_global_rows
    procedure, pass(desc) :: get_local_rows => psb_cd_get_local_rows
    procedure, pass(desc) :: get_local_cols => psb_cd_get_local_cols
    procedure, pass(desc) :: is_asb => psb_cd_is_asb
    procedure, pass(desc) :: sizeof => psb_cd_sizeof
    procedure, pass(desc) :: get_fmt => psb_cd_get_fmt
  end type psb_desc_type

  ! ---- sparse matrix: the device CSR of libpsfun_hip ------------------------
  type :: psb_dspmat_type
    type(c_ptr) :: a = c_null_ptr          ! psfun_csr_t
    type(psb_ctxt_type) :: ctxt
    integer(psb_lpk_) :: m = 0, nz = 0
    character(len=5) :: fmt = 'CSR'
    ! build state between psb_spall and psb_spasb: 1-based triplets
    integer(psb_lpk_), allocatable :: bia(:), bja(:)
    real(psb_dpk_), allocatable :: bval(:)
    integer(psb_lpk_) :: nins = 0
    ! host copy of the assembled CSR (0-based), as PSBLAS keeps the matrix on
    ! the host; read by the dense f(H) step's sparse apply
    integer(c_int32_t), allocatable :: hrp(:), hci(:)
    real(c_double), allocatable :: hva(:)
  contains
    procedure, pass(a) :: get_nrows => psb_d_get_nrows
    procedure, pass(a) :: get_ncols => psb_d_get_nrows
    procedure, pass(a) :: get_nzeros => psb_d_get_nzeros
    procedure, pass(a) :: get_fmt => psb_d_get_fmt
    procedure, pass(a) :: sizeof => psb_d_sizeof
    procedure, pass(a) :: is_asb => psb_d_is_asb
    procedure, pass(a) :: set_dev_format => psb_d_set_dev_format
  end type psb_dspmat_type

  ! ---- dense vector: a device array of n FP64 -------------------------------
  type :: psb_d_vect_type
    type(c_ptr) :: v = c_null_ptr          ! device pointer
    type(psb_ctxt_type) :: ctxt
    integer(psb_lpk_) :: n = 0
    real(psb_dpk_), allocatable :: h(:)    ! staged entries between psb_geall and psb_geasb
  contains
    procedure, pass(x) :: get_nrows => psb_d_vect_get_nrows
    procedure, pass(x) :: is_allocated => psb_d_vect_is_allocated
    procedure, pass(x) :: zero => psb_d_vect_zero
    procedure, pass(x) :: set_scal => psb_d_vect_set_scal
    procedure, pass(x) :: set_vect => psb_d_vect_set_vect
    generic, public :: set => set_scal, set_vect
    procedure, pass(x) :: get_vect => psb_d_vect_get_vect
    procedure, pass(x) :: free => psb_d_vect_free
  end type psb_d_vect_type

  interface psb_error_handler
    module procedure psb_error_handler_act, psb_error_handler_ctxt
  end interface
  interface psb_bcast
    module procedure psb_bcast_i, psb_bcast_l, psb_bcast_d, psb_bcast_c, psb_bcast_lg
  end interface
  interface psb_amx
    module procedure psb_amx_d
  end interface
  interface psb_sum
    module procedure psb_sum_i, psb_sum_l, psb_sum_d
  end interface
  interface psb_cdall
    module procedure psb_cdall_n
  end interface
  interface psb_geins
    module procedure psb_geins_v, psb_geins_s
  end interface

  ! error stack: the last pushed error (one process, one thread)
  integer(psb_ipk_), save :: err_code = 0
  character(len=64), save :: err_name = ''
  character(len=128), save :: err_text = ''
  integer(psb_ipk_), save :: err_status = 0
  ! the context psb_init created (used by routines without a descriptor)
  type(psb_ctxt_type), save :: psb_default_ctxt

contains

  ! ---- helpers ---------------------------------------------------------------
  ! Map a libpsfun_hip status onto the info names PSFUN uses.
  function psb_status_to_info(st) result(info)
    integer(c_int), intent(in) :: st
    integer(psb_ipk_) :: info
    select case (st)
    case (PSFUN_SUCCESS)
      info = psb_success_
    case (PSFUN_ERR_INVALID_ISTOP)
      info = psb_err_invalid_istop_
    case (PSFUN_ERR_INVALID_VECT_STATE)
      info = psb_err_invalid_vect_state_
    case (PSFUN_ERR_INVALID_ARGS_COMBINATION)
      info = psb_err_invalid_args_combination_
    case (PSFUN_ERR_ALLOC)
      info = psb_err_alloc_dealloc_
    case (PSFUN_ERR_UNSUPPORTED)
      info = psb_err_unsupported_
    case default
      info = psb_err_from_subroutine_non_
    end select
  end function psb_status_to_info

  ! Text of a libpsfun_hip status.
  function psb_status_text(st) result(msg)
    integer(c_int), intent(in) :: st
    character(len=128) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ''
    call c_f_pointer(psfun_strerror(st), p, [128])
    do i = 1, 128
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function psb_status_text

  function psb_toupper(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i, k
    u = s
    do i = 1, len(s)
      k = iachar(s(i:i))
      if (k >= iachar('a') .and. k <= iachar('z')) u(i:i) = achar(k - 32)
    end do
  end function psb_toupper

  function psb_tolower(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i, k
    u = s
    do i = 1, len(s)
      k = iachar(s(i:i))
      if (k >= iachar('A') .and. k <= iachar('Z')) u(i:i) = achar(k + 32)
    end do
  end function psb_tolower

  ! ---- errors ----------------------------------------------------------------
  subroutine psb_errpush(err, name, i_err, a_err)
    integer(psb_ipk_), intent(in) :: err
    character(len=*), intent(in) :: name
    integer(psb_ipk_), intent(in), optional :: i_err(:)
    character(len=*), intent(in), optional :: a_err
    err_code = err
    err_name = name
    err_text = ''
    if (present(a_err)) err_text = a_err
    if (present(i_err)) then
      if (size(i_err) > 0) write(err_text, '(a,1x,i0)') trim(err_text), i_err(1)
    end if
    err_status = 1
  end subroutine psb_errpush

  subroutine psb_erractionsave(err_act)
    integer(psb_ipk_), intent(out) :: err_act
    err_act = psb_act_ret_
  end subroutine psb_erractionsave

  subroutine psb_erractionrestore(err_act)
    integer(psb_ipk_), intent(in) :: err_act
  end subroutine psb_erractionrestore

  subroutine psb_error_handler_act(err_act)
    integer(psb_ipk_), intent(in) :: err_act
    if (err_act == psb_act_abort_) call psb_error()
  end subroutine psb_error_handler_act

  subroutine psb_error_handler_ctxt(ctxt, err_act)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_ipk_), intent(in) :: err_act
    if (err_act == psb_act_abort_) call psb_error(ctxt)
  end subroutine psb_error_handler_ctxt

  function psb_get_errstatus() result(st)
    integer(psb_ipk_) :: st
    st = err_status
  end function psb_get_errstatus

  ! Print the last pushed error (PSBLAS prints its error stack).
  subroutine psb_error(ctxt)
    type(psb_ctxt_type), intent(in), optional :: ctxt
    if (err_status /= 0) then
      write(psb_err_unit, '(a,i0,a,a,a,a)') 'PSBLAS error ', err_code, ' in ', trim(err_name), ': ', trim(err_text)
    end if
    err_status = 0
  end subroutine psb_error

  ! ---- environment ---------------------------------------------------------------
  ! Create the device context (device 0 unless given) that the objects built
  ! on this ctxt use.
  subroutine psb_init(ctxt, np, basectxt, ids, device)
    type(psb_ctxt_type), intent(out) :: ctxt
    integer(psb_ipk_), intent(in), optional :: np
    type(psb_ctxt_type), intent(in), optional :: basectxt
    integer(psb_ipk_), intent(in), optional :: ids(:)
    integer(psb_ipk_), intent(in), optional :: device
    integer(c_int) :: st, dev
    dev = 0
    if (present(device)) dev = device
    st = psfun_ctx_create(dev, ctxt%ctx)
    if (st /= PSFUN_SUCCESS) then
      write(psb_err_unit, '(a,a)') 'psb_init: ', trim(psb_status_text(st))
      ctxt%ctx = c_null_ptr
      ctxt%device = -1
      err_status = 1
      return
    end if
    ctxt%device = dev
    psb_default_ctxt = ctxt
  end subroutine psb_init

  subroutine psb_info(ctxt, iam, np)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_ipk_), intent(out) :: iam, np
    if (c_associated(ctxt%ctx)) then
      iam = 0
      np = 1
    else
      iam = -1
      np = -1
    end if
  end subroutine psb_info

  subroutine psb_exit(ctxt, close)
    type(psb_ctxt_type), intent(inout) :: ctxt
    logical, intent(in), optional :: close
    integer(c_int) :: st
    if (c_associated(ctxt%ctx)) then
      st = psfun_ctx_destroy(ctxt%ctx)
      if (c_associated(psb_default_ctxt%ctx, ctxt%ctx)) psb_default_ctxt%ctx = c_null_ptr
    end if
    ctxt%ctx = c_null_ptr
    ctxt%device = -1
  end subroutine psb_exit

  subroutine psb_abort(ctxt, errc)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_ipk_), intent(in), optional :: errc
    call psb_error(ctxt)
    error stop 1
  end subroutine psb_abort

  subroutine psb_barrier(ctxt)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(c_int) :: st
    if (c_associated(ctxt%ctx)) st = psfun_ctx_synchronize(ctxt%ctx)
  end subroutine psb_barrier

  function psb_wtime() result(t)
    real(psb_dpk_) :: t
    integer(psb_epk_) :: cnt, rate
    call system_clock(cnt, rate)
    t = real(cnt, psb_dpk_) / real(rate, psb_dpk_)
  end function psb_wtime

  ! collectives over one process are identities
  subroutine psb_bcast_i(ctxt, v, root)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_ipk_), intent(inout) :: v
    integer(psb_ipk_), intent(in), optional :: root
  end subroutine psb_bcast_i
  subroutine psb_bcast_l(ctxt, v, root)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_lpk_), intent(inout) :: v
    integer(psb_ipk_), intent(in), optional :: root
  end subroutine psb_bcast_l
  subroutine psb_bcast_d(ctxt, v, root)
    type(psb_ctxt_type), intent(in) :: ctxt
    real(psb_dpk_), intent(inout) :: v
    integer(psb_ipk_), intent(in), optional :: root
  end subroutine psb_bcast_d
  subroutine psb_bcast_c(ctxt, v, root)
    type(psb_ctxt_type), intent(in) :: ctxt
    character(len=*), intent(inout) :: v
    integer(psb_ipk_), intent(in), optional :: root
  end subroutine psb_bcast_c
  subroutine psb_bcast_lg(ctxt, v, root)
    type(psb_ctxt_type), intent(in) :: ctxt
    logical, intent(inout) :: v
    integer(psb_ipk_), intent(in), optional :: root
  end subroutine psb_bcast_lg
  subroutine psb_amx_d(ctxt, v)
    type(psb_ctxt_type), intent(in) :: ctxt
    real(psb_dpk_), intent(inout) :: v
  end subroutine psb_amx_d
  subroutine psb_sum_i(ctxt, v)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_ipk_), intent(inout) :: v
  end subroutine psb_sum_i
  subroutine psb_sum_l(ctxt, v)
    type(psb_ctxt_type), intent(in) :: ctxt
    integer(psb_lpk_), intent(inout) :: v
  end subroutine psb_sum_l
  subroutine psb_sum_d(ctxt, v)
    type(psb_ctxt_type), intent(in) :: ctxt
    real(psb_dpk_), intent(inout) :: v
  end subroutine psb_sum_d

  ! ---- descriptor ----------------------------------------------------------------
  ! psb_cdall(ctxt, desc, info, nl=n) or mg=n: n rows, all owned here.
  subroutine psb_cdall_n(ctxt, desc, info, nl, mg, vg, parts)
    type(psb_ctxt_type), intent(in) :: ctxt
    type(psb_desc_type), intent(out) :: desc
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: nl
    integer(psb_lpk_), intent(in), optional :: mg
    integer(psb_ipk_), intent(in), optional :: vg(:)
    external :: parts
    optional :: parts
    info = psb_success_
    desc%ctxt = ctxt
    if (present(nl)) then
      desc%n = nl
    else if (present(mg)) then
      desc%n = mg
    else if (present(vg)) then
      desc%n = size(vg)
    else
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_cdall')
      return
    end if
    if (.not. c_associated(ctxt%ctx)) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_cdall', a_err='context not initialised (psb_init)')
    end if
  end subroutine psb_cdall_n

  subroutine psb_cdasb(desc, info)
    type(psb_desc_type), intent(inout) :: desc
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    desc%asb = .true.
  end subroutine psb_cdasb

  subroutine psb_cdfree(desc, info)
    type(psb_desc_type), intent(inout) :: desc
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    desc%n = -1
    desc%asb = .false.
  end subroutine psb_cdfree

  function psb_cd_get_context(desc) result(ctxt)
    class(psb_desc_type), intent(in) :: desc
    type(psb_ctxt_type) :: ctxt
    ctxt = desc%ctxt
  end function psb_cd_get_context
  function psb_cd_get_global_rows(desc) result(n)
    class(psb_desc_type), intent(in) :: desc
    integer(psb_lpk_) :: n
    n = desc%n
  end function psb_cd_get_global_rows
  function psb_cd_get_local_rows(desc) result(n)
    class(psb_desc_type), intent(in) :: desc
    integer(psb_ipk_) :: n
    n = int(desc%n, psb_ipk_)
  end function psb_cd_get_local_rows
  function psb_cd_get_local_cols(desc) result(n)
    class(psb_desc_type), intent(in) :: desc
    integer(psb_ipk_) :: n
    n = int(desc%n, psb_ipk_)
  end function psb_cd_get_local_cols
  function psb_cd_is_asb(desc) result(r)
    class(psb_desc_type), intent(in) :: desc
    logical :: r
    r = desc%asb
  end function psb_cd_is_asb
  function psb_cd_sizeof(desc) result(s)
    class(psb_desc_type), intent(in) :: desc
    integer(psb_epk_) :: s
    s = 32
  end function psb_cd_sizeof
  function psb_cd_get_fmt(desc) result(f)
    class(psb_desc_type), intent(in) :: desc
    character(len=5) :: f
    f = 'REPL'
  end function psb_cd_get_fmt

  ! ---- sparse matrix -----------------------------------------------------------------
  subroutine psb_spall(a, desc_a, info, nnz)
    type(psb_dspmat_type), intent(inout) :: a
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: nnz
    integer(psb_lpk_) :: cap
    info = psb_success_
    call psb_spfree(a, desc_a, info)
    cap = 16
    if (present(nnz)) cap = max(int(nnz, psb_lpk_), cap)
    allocate(a%bia(cap), a%bja(cap), a%bval(cap), stat=info)
    if (info /= 0) then
      info = psb_err_alloc_dealloc_
      call psb_errpush(info, 'psb_spall')
      return
    end if
    a%nins = 0
    a%m = desc_a%n
    a%ctxt = desc_a%ctxt
  end subroutine psb_spall

  ! Insert nz entries (ia(k), ja(k), val(k)), 1-based global indices.
  subroutine psb_spins(nz, ia, ja, val, a, desc_a, info)
    integer(psb_ipk_), intent(in) :: nz
    integer(psb_ipk_), intent(in) :: ia(:), ja(:)
    real(psb_dpk_), intent(in) :: val(:)
    type(psb_dspmat_type), intent(inout) :: a
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_lpk_), allocatable :: ti(:), tj(:)
    real(psb_dpk_), allocatable :: tv(:)
    integer(psb_lpk_) :: need, cap
    integer :: k
    info = psb_success_
    if (.not. allocated(a%bia)) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_spins', a_err='matrix not in the build state (psb_spall)')
      return
    end if
    do k = 1, nz
      if (ia(k) < 1 .or. ia(k) > desc_a%n .or. ja(k) < 1 .or. ja(k) > desc_a%n) then
        info = psb_err_invalid_args_combination_
        call psb_errpush(info, 'psb_spins', a_err='index out of range')
        return
      end if
    end do
    need = a%nins + nz
    cap = size(a%bia, kind=psb_lpk_)
    if (need > cap) then
      cap = max(need, 2 * cap)
      allocate(ti(cap), tj(cap), tv(cap))
      ti(1:a%nins) = a%bia(1:a%nins)
      tj(1:a%nins) = a%bja(1:a%nins)
      tv(1:a%nins) = a%bval(1:a%nins)
      call move_alloc(ti, a%bia)
      call move_alloc(tj, a%bja)
      call move_alloc(tv, a%bval)
    end if
    a%bia(a%nins + 1:need) = ia(1:nz)
    a%bja(a%nins + 1:need) = ja(1:nz)
    a%bval(a%nins + 1:need) = val(1:nz)
    a%nins = need
  end subroutine psb_spins

  ! Assemble: sort the triplets by row (stable, so each row keeps the
  ! insertion order of its columns after a sort by column), sum repeated
  ! entries, and create the device CSR.  afmt: 'CSR' lets the library pick
  ! its storage (sliced ELL when it pads little, else CSR), 'HLL'/'ELL' ask
  ! for sliced ELL, 'HDIA'/'DIA' for diagonal slices.
  subroutine psb_spasb(a, desc_a, info, afmt, dupl, upd, mold)
    type(psb_dspmat_type), intent(inout) :: a
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    character(len=*), intent(in), optional :: afmt
    integer(psb_ipk_), intent(in), optional :: dupl, upd
    class(*), intent(in), optional :: mold
    integer(c_int32_t), allocatable :: rp(:), ci(:)
    real(c_double), allocatable :: va(:)
    integer(psb_lpk_), allocatable :: cnt(:), pos(:), ord(:)
    integer(psb_lpk_) :: n, k, i, q, p, nnz, r0, r1
    integer(c_int) :: st
    character(len=5) :: f
    info = psb_success_
    if (.not. allocated(a%bia)) then
      info = psb_err_invalid_args_combination_
      call psb_errpush(info, 'psb_spasb', a_err='matrix not in the build state (psb_spall)')
      return
    end if
    n = desc_a%n
    ! counting sort by row
    allocate(cnt(n + 1), pos(n + 1), ord(max(a%nins, 1_psb_lpk_)))
    cnt = 0
    do k = 1, a%nins
      cnt(a%bia(k) + 1) = cnt(a%bia(k) + 1) + 1
    end do
    cnt(1) = 1
    do i = 2, n + 1
      cnt(i) = cnt(i) + cnt(i - 1)
    end do
    pos = cnt
    do k = 1, a%nins
      ord(pos(a%bia(k))) = k
      pos(a%bia(k)) = pos(a%bia(k)) + 1
    end do
    ! inside each row: insertion sort by column, then merge repeats
    allocate(rp(n + 1), ci(max(a%nins, 1_psb_lpk_)), va(max(a%nins, 1_psb_lpk_)))
    nnz = 0
    rp(1) = 0
    do i = 1, n
      r0 = cnt(i)
      r1 = cnt(i + 1) - 1
      do q = r0 + 1, r1
        k = ord(q)
        p = q - 1
        do while (p >= r0)
          if (a%bja(ord(p)) <= a%bja(k)) exit
          ord(p + 1) = ord(p)
          p = p - 1
        end do
        ord(p + 1) = k
      end do
      do q = r0, r1
        k = ord(q)
        if (nnz > rp(i) .and. q > r0) then
          if (ci(nnz) == int(a%bja(k) - 1, c_int32_t)) then
            va(nnz) = va(nnz) + a%bval(k)
            cycle
          end if
        end if
        nnz = nnz + 1
        ci(nnz) = int(a%bja(k) - 1, c_int32_t)
        va(nnz) = a%bval(k)
      end do
      rp(i + 1) = int(nnz, c_int32_t)
    end do
    st = psfun_csr_create(desc_a%ctxt%ctx, n, nnz, rp, ci, va, 0_c_int, PSFUN_MEM_HOST, a%a)
    if (st /= PSFUN_SUCCESS) then
      info = psb_status_to_info(st)
      call psb_errpush(info, 'psb_spasb', a_err=trim(psb_status_text(st)))
      return
    end if
    a%m = n
    a%nz = nnz
    a%ctxt = desc_a%ctxt
    deallocate(a%bia, a%bja, a%bval)
    call move_alloc(rp, a%hrp)
    call move_alloc(ci, a%hci)
    call move_alloc(va, a%hva)
    a%nins = 0
    f = 'CSR'
    if (present(afmt)) f = psb_toupper(afmt)
    call a%set_dev_format(f, info)
  end subroutine psb_spasb

  ! Device storage for an assembled matrix ('CSR', 'HLL'/'ELL', 'HDIA'/'DIA').
  subroutine psb_d_set_dev_format(a, afmt, info)
    class(psb_dspmat_type), intent(inout) :: a
    character(len=*), intent(in) :: afmt
    integer(psb_ipk_), intent(out) :: info
    integer(c_int) :: st, code
    info = psb_success_
    select case (psb_toupper(trim(afmt)))
    case ('HLL', 'ELL')
      code = PSFUN_FORMAT_SELL
    case ('HDIA', 'DIA')
      code = PSFUN_FORMAT_HDIA
    case default
      code = PSFUN_FORMAT_AUTO
    end select
    st = psfun_csr_set_format(a%a, code)
    if (st /= PSFUN_SUCCESS) then
      info = psb_status_to_info(st)
      call psb_errpush(info, 'set_dev_format', a_err=trim(psb_status_text(st)))
      return
    end if
    a%fmt = psb_toupper(afmt)
  end subroutine psb_d_set_dev_format

  subroutine psb_spfree(a, desc_a, info)
    type(psb_dspmat_type), intent(inout) :: a
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(c_int) :: st
    info = psb_success_
    if (c_associated(a%a)) st = psfun_csr_destroy(a%a)
    a%a = c_null_ptr
    if (allocated(a%bia)) deallocate(a%bia)
    if (allocated(a%bja)) deallocate(a%bja)
    if (allocated(a%bval)) deallocate(a%bval)
    if (allocated(a%hrp)) deallocate(a%hrp)
    if (allocated(a%hci)) deallocate(a%hci)
    if (allocated(a%hva)) deallocate(a%hva)
    a%nins = 0
    a%m = 0
    a%nz = 0
  end subroutine psb_spfree

  function psb_d_get_nrows(a) result(n)
    class(psb_dspmat_type), intent(in) :: a
    integer(psb_ipk_) :: n
    n = int(a%m, psb_ipk_)
  end function psb_d_get_nrows
  function psb_d_get_nzeros(a) result(n)
    class(psb_dspmat_type), intent(in) :: a
    integer(psb_lpk_) :: n
    n = a%nz
  end function psb_d_get_nzeros
  function psb_d_get_fmt(a) result(f)
    class(psb_dspmat_type), intent(in) :: a
    character(len=5) :: f
    f = a%fmt
  end function psb_d_get_fmt
  function psb_d_sizeof(a) result(s)
    class(psb_dspmat_type), intent(in) :: a
    integer(psb_epk_) :: s
    s = 12 * a%nz + 4 * (a%m + 1)
  end function psb_d_sizeof
  function psb_d_is_asb(a) result(r)
    class(psb_dspmat_type), intent(in) :: a
    logical :: r
    r = c_associated(a%a)
  end function psb_d_is_asb

  ! ---- dense vectors -------------------------------------------------------------
  subroutine psb_geall(x, desc_a, info)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    call x%free(info)
    x%n = desc_a%n
    x%ctxt = desc_a%ctxt
    allocate(x%h(x%n), stat=info)
    if (info /= 0) then
      info = psb_err_alloc_dealloc_
      call psb_errpush(info, 'psb_geall')
      return
    end if
    x%h = dzero
  end subroutine psb_geall

  ! Entries irw(1:m) (1-based global) take val(1:m) (added to what is there, as
  ! PSBLAS's default psb_dupl_add_ does during the build).
  subroutine psb_geins_v(m, irw, val, x, desc_a, info, dupl)
    integer(psb_ipk_), intent(in) :: m
    integer(psb_ipk_), intent(in) :: irw(:)
    real(psb_dpk_), intent(in) :: val(:)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: dupl
    integer :: k
    info = psb_success_
    if (.not. allocated(x%h)) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psb_geins')
      return
    end if
    do k = 1, m
      if (irw(k) < 1 .or. irw(k) > x%n) then
        info = psb_err_invalid_args_combination_
        call psb_errpush(info, 'psb_geins', a_err='index out of range')
        return
      end if
      x%h(irw(k)) = x%h(irw(k)) + val(k)
    end do
  end subroutine psb_geins_v

  subroutine psb_geins_s(m, irw, val, x, desc_a, info, dupl)
    integer(psb_ipk_), intent(in) :: m
    integer(psb_ipk_), intent(in) :: irw(:)
    real(psb_dpk_), intent(in) :: val
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: dupl
    real(psb_dpk_), allocatable :: v(:)
    allocate(v(m))
    v = val
    call psb_geins_v(m, irw, v, x, desc_a, info, dupl)
  end subroutine psb_geins_s

  ! Assemble: the staged entries go to a new device array.  mold is accepted
  ! for PSBLAS source compatibility: every vector here is a device vector
  ! (krylov/polynomial/psfun_d_lanczos.f90:159-164 allocates its work vectors
  ! with mold=x%v).
  subroutine psb_geasb(x, desc_a, info, mold, scratch)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    class(*), intent(in), optional :: mold
    logical, intent(in), optional :: scratch
    integer(c_int) :: st
    info = psb_success_
    if (.not. allocated(x%h)) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psb_geasb')
      return
    end if
    if (.not. c_associated(x%v)) then
      st = psfun_vec_alloc(x%ctxt%ctx, max(x%n, 1_psb_lpk_), x%v)
      if (st /= PSFUN_SUCCESS) then
        info = psb_status_to_info(st)
        call psb_errpush(info, 'psb_geasb', a_err=trim(psb_status_text(st)))
        return
      end if
    end if
    if (x%n > 0) then
      st = psfun_vec_upload(x%ctxt%ctx, x%v, x%h, x%n)
      if (st /= PSFUN_SUCCESS) then
        info = psb_status_to_info(st)
        call psb_errpush(info, 'psb_geasb', a_err=trim(psb_status_text(st)))
        return
      end if
    end if
    deallocate(x%h)
  end subroutine psb_geasb

  subroutine psb_gefree(x, desc_a, info)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    call x%free(info)
  end subroutine psb_gefree

  ! x = glob (one process: the whole vector), as an assembled device vector.
  subroutine psb_scatter(glob, x, desc_a, info, root, mold)
    real(psb_dpk_), intent(in) :: glob(:)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: root
    class(*), intent(in), optional :: mold
    call psb_geall(x, desc_a, info)
    if (info /= psb_success_) return
    x%h(1:x%n) = glob(1:x%n)
    call psb_geasb(x, desc_a, info)
  end subroutine psb_scatter

  ! glob = x copied to the host (allocated here).
  subroutine psb_gather(glob, x, desc_a, info, root)
    real(psb_dpk_), allocatable, intent(out) :: glob(:)
    type(psb_d_vect_type), intent(in) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: root
    glob = x%get_vect()
    info = psb_success_
  end subroutine psb_gather

  function psb_d_vect_get_nrows(x) result(n)
    class(psb_d_vect_type), intent(in) :: x
    integer(psb_ipk_) :: n
    n = int(x%n, psb_ipk_)
  end function psb_d_vect_get_nrows

  function psb_d_vect_is_allocated(x) result(r)
    class(psb_d_vect_type), intent(in) :: x
    logical :: r
    r = c_associated(x%v) .or. allocated(x%h)
  end function psb_d_vect_is_allocated

  subroutine psb_d_vect_zero(x)
    class(psb_d_vect_type), intent(inout) :: x
    call x%set_scal(dzero)
  end subroutine psb_d_vect_zero

  ! every entry = val (on the device once assembled, else in the staging)
  subroutine psb_d_vect_set_scal(x, val)
    class(psb_d_vect_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: val
    real(psb_dpk_), allocatable :: t(:)
    integer(c_int) :: st
    if (allocated(x%h)) then
      x%h = val
    else if (c_associated(x%v) .and. x%n > 0) then
      allocate(t(x%n))
      t = val
      st = psfun_vec_upload(x%ctxt%ctx, x%v, t, x%n)
    end if
  end subroutine psb_d_vect_set_scal

  subroutine psb_d_vect_set_vect(x, val)
    class(psb_d_vect_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: val(:)
    integer(c_int) :: st
    if (allocated(x%h)) then
      x%h(1:min(size(val, kind=psb_lpk_), x%n)) = val(1:min(size(val, kind=psb_lpk_), x%n))
    else if (c_associated(x%v) .and. x%n > 0) then
      st = psfun_vec_upload(x%ctxt%ctx, x%v, val, min(size(val, kind=psb_lpk_), x%n))
    end if
  end subroutine psb_d_vect_set_vect

  function psb_d_vect_get_vect(x) result(res)
    class(psb_d_vect_type), intent(in) :: x
    real(psb_dpk_), allocatable :: res(:)
    integer(c_int) :: st
    allocate(res(x%n))
    if (allocated(x%h)) then
      res = x%h
    else if (c_associated(x%v) .and. x%n > 0) then
      st = psfun_vec_download(x%ctxt%ctx, res, x%v, x%n)
    end if
  end function psb_d_vect_get_vect

  subroutine psb_d_vect_free(x, info)
    class(psb_d_vect_type), intent(inout) :: x
    integer(psb_ipk_), intent(out) :: info
    integer(c_int) :: st
    info = psb_success_
    if (c_associated(x%v)) st = psfun_vec_free(x%ctxt%ctx, x%v)
    x%v = c_null_ptr
    if (allocated(x%h)) deallocate(x%h)
    x%n = 0
  end subroutine psb_d_vect_free

  ! ---- computational routines (one device kernel each) ---------------------------
  logical function vect_ready(x)
    type(psb_d_vect_type), intent(in) :: x
    vect_ready = c_associated(x%v)
  end function vect_ready

  ! y = alpha A x + beta y
  subroutine psb_spmm(alpha, a, x, beta, y, desc_a, info, trans, work)
    real(psb_dpk_), intent(in) :: alpha, beta
    type(psb_dspmat_type), intent(in) :: a
    type(psb_d_vect_type), intent(inout) :: x, y
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    character, intent(in), optional :: trans
    real(psb_dpk_), intent(inout), optional, target :: work(:)
    integer(c_int) :: st
    info = psb_success_
    if (.not. (vect_ready(x) .and. vect_ready(y))) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psb_spmm')
      return
    end if
    st = psfun_spmv(a%ctxt%ctx, a%a, alpha, x%v, beta, y%v)
    if (st /= PSFUN_SUCCESS) then
      info = psb_status_to_info(st)
      call psb_errpush(info, 'psb_spmm', a_err=trim(psb_status_text(st)))
    end if
  end subroutine psb_spmm

  function psb_gedot(x, y, desc_a, info, global) result(res)
    type(psb_d_vect_type), intent(inout) :: x, y
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    logical, intent(in), optional :: global
    real(psb_dpk_) :: res
    integer(c_int) :: st
    res = dzero
    info = psb_success_
    if (.not. (vect_ready(x) .and. vect_ready(y))) then
      info = psb_err_invalid_vect_state_
      return
    end if
    st = psfun_dot(x%ctxt%ctx, x%n, x%v, y%v, res)
    info = psb_status_to_info(st)
  end function psb_gedot

  function psb_genrm2(x, desc_a, info, global) result(res)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    logical, intent(in), optional :: global
    real(psb_dpk_) :: res
    integer(c_int) :: st
    res = dzero
    info = psb_success_
    if (.not. vect_ready(x)) then
      info = psb_err_invalid_vect_state_
      return
    end if
    st = psfun_nrm2(x%ctxt%ctx, x%n, x%v, res)
    info = psb_status_to_info(st)
  end function psb_genrm2

  function psb_geamax(x, desc_a, info, global) result(res)
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    logical, intent(in), optional :: global
    real(psb_dpk_) :: res
    integer(c_int) :: st
    res = dzero
    info = psb_success_
    if (.not. vect_ready(x)) then
      info = psb_err_invalid_vect_state_
      return
    end if
    st = psfun_amax(x%ctxt%ctx, x%n, x%v, res)
    info = psb_status_to_info(st)
  end function psb_geamax

  ! y = alpha x + beta y
  subroutine psb_geaxpby(alpha, x, beta, y, desc_a, info)
    real(psb_dpk_), intent(in) :: alpha, beta
    type(psb_d_vect_type), intent(inout) :: x, y
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(c_int) :: st
    info = psb_success_
    if (.not. (vect_ready(x) .and. vect_ready(y))) then
      info = psb_err_invalid_vect_state_
      call psb_errpush(info, 'psb_geaxpby')
      return
    end if
    st = psfun_axpby(y%ctxt%ctx, y%n, alpha, x%v, beta, y%v)
    info = psb_status_to_info(st)
  end subroutine psb_geaxpby

end module psb_base_mod
