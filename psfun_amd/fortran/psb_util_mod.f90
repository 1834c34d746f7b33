! psb_util_mod.f90 -- matrix input and distribution as PSFUN's example
! drivers call them (examples/polynomial/lanczostest.F90:107-178,
! examples/shiftedkrylov/shiftedtest.f90:150-200), for one process:
!   mm_mat_read(a, info, iunit, filename)    MatrixMarket coordinate file, into
!                                            the host global matrix or straight
!                                            into a device matrix
!   hb_read(a, info, iunit, b, filename, mtitle)
!                                            Harwell-Boeing file and the
!                                            right-hand sides it holds
!   mm_array_read(b, info, iunit, filename)  MatrixMarket array file
!   mm_array_write(y, title, info, iunit, filename)
!   psb_matdist(aux_a, a, ctxt, desc_a, info, fmt, parts)
!   part_block                               the block partition (one block)
! psb_ldspmat_type is the global matrix on the host (1-based triplets).
module psb_util_mod
  use iso_c_binding
  use psb_base_mod
  implicit none

  type :: psb_ldspmat_type
    integer(psb_lpk_) :: m = 0, nz = 0
    integer(psb_lpk_), allocatable :: ia(:), ja(:)
    real(psb_dpk_), allocatable :: val(:)
  contains
    procedure, pass(a) :: get_nrows => psb_ld_get_nrows
    procedure, pass(a) :: get_ncols => psb_ld_get_nrows
    procedure, pass(a) :: get_nzeros => psb_ld_get_nzeros
  end type psb_ldspmat_type

  interface mm_mat_read
    module procedure mm_lmat_read, mm_dmat_read
  end interface mm_mat_read

contains

  function psb_ld_get_nrows(a) result(n)
    class(psb_ldspmat_type), intent(in) :: a
    integer(psb_lpk_) :: n
    n = a%m
  end function psb_ld_get_nrows
  function psb_ld_get_nzeros(a) result(n)
    class(psb_ldspmat_type), intent(in) :: a
    integer(psb_lpk_) :: n
    n = a%nz
  end function psb_ld_get_nzeros

  ! The block partition of PSBLAS's examples; with one process every row is
  ! in the only block.
  subroutine part_block(global_indx, n, np, pv, nv)
    integer(psb_ipk_), intent(in) :: global_indx, n, np
    integer(psb_ipk_), intent(out) :: nv
    integer(psb_ipk_), intent(out) :: pv(*)
    nv = 1
    pv(1) = 0
  end subroutine part_block

  subroutine lower_words(line, w, nw)
    character(len=*), intent(in) :: line
    character(len=32), intent(out) :: w(8)
    integer, intent(out) :: nw
    integer :: i, s
    character(len=len(line)) :: l
    l = psb_tolower(line)
    nw = 0
    i = 1
    do while (i <= len_trim(l) .and. nw < 8)
      do while (i <= len_trim(l))
        if (l(i:i) /= ' ') exit
        i = i + 1
      end do
      if (i > len_trim(l)) exit
      s = i
      do while (i <= len_trim(l))
        if (l(i:i) == ' ') exit
        i = i + 1
      end do
      nw = nw + 1
      w(nw) = l(s:i - 1)
    end do
  end subroutine lower_words

  ! MatrixMarket coordinate matrix: real/integer/pattern entries, general,
  ! symmetric or skew-symmetric storage (only one triangle stored: the other
  ! is generated), repeated entries kept (psb_spasb sums them).
  subroutine mm_lmat_read(a, info, iunit, filename)
    type(psb_ldspmat_type), intent(out) :: a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iunit
    character(len=*), intent(in), optional :: filename
    character(len=1024) :: line
    character(len=32) :: w(8)
    integer :: u, nw, ios, k
    integer(psb_lpk_) :: nr, nc, nnz, i, j, q, nsym
    real(psb_dpk_) :: v
    logical :: pattern, sym, skew
    info = psb_success_
    u = 12
    if (present(iunit)) u = iunit
    if (.not. present(filename)) then
      info = psb_err_invalid_args_combination_
      return
    end if
    open(u, file=filename, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_mat_read', a_err=trim(filename))
      return
    end if
    read(u, '(a)', iostat=ios) line
    call lower_words(line, w, nw)
    if (ios /= 0 .or. nw < 5 .or. w(1) /= '%%matrixmarket' .or. w(2) /= 'matrix' .or. w(3) /= 'coordinate') then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_mat_read', a_err='not a MatrixMarket coordinate matrix')
      close(u)
      return
    end if
    pattern = w(4) == 'pattern'
    if (w(4) /= 'real' .and. w(4) /= 'integer' .and. .not. pattern) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_mat_read', a_err='unsupported field '//trim(w(4)))
      close(u)
      return
    end if
    sym = w(5) == 'symmetric'
    skew = w(5) == 'skew-symmetric'
    if (w(5) /= 'general' .and. .not. sym .and. .not. skew) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_mat_read', a_err='unsupported symmetry '//trim(w(5)))
      close(u)
      return
    end if
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      if (line(1:1) /= '%') exit
    end do
    read(line, *, iostat=ios) nr, nc, nnz
    if (ios /= 0 .or. nr /= nc) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_mat_read', a_err='bad size line or non-square matrix')
      close(u)
      return
    end if
    allocate(a%ia(2 * nnz), a%ja(2 * nnz), a%val(2 * nnz))
    q = 0
    nsym = 0
    do k = 1, int(nnz)
      if (pattern) then
        read(u, *, iostat=ios) i, j
        v = done
      else
        read(u, *, iostat=ios) i, j, v
      end if
      if (ios /= 0) then
        info = psb_err_from_subroutine_
        call psb_errpush(info, 'mm_mat_read', a_err='short file')
        close(u)
        return
      end if
      q = q + 1
      a%ia(q) = i
      a%ja(q) = j
      a%val(q) = v
      if ((sym .or. skew) .and. i /= j) then
        q = q + 1
        a%ia(q) = j
        a%ja(q) = i
        a%val(q) = merge(-v, v, skew)
      end if
    end do
    close(u)
    a%m = nr
    a%nz = q
  end subroutine mm_lmat_read

  ! Harwell-Boeing file (the drivers' filefmt 'HB', lanczostest.F90:119-122;
  ! format of Duff, Grimes and Lewis, 1992): a 4-line header (5 with
  ! right-hand sides), then the column pointers, row indices, values and
  ! right-hand sides, each block on its own cards read with the format the
  ! header gives.  Types R (real) or P (pattern: values 1) with U, S or Z
  ! (one triangle stored, the other generated) and A (assembled).  b receives
  ! the full ('F') right-hand sides, when the file holds any.
  subroutine hb_read(a, info, iunit, b, filename, mtitle)
    type(psb_ldspmat_type), intent(out) :: a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iunit
    real(psb_dpk_), allocatable, intent(inout), optional :: b(:, :)
    character(len=*), intent(in), optional :: filename
    character(len=72), intent(out), optional :: mtitle
    character(len=72) :: title
    character(len=8) :: key
    character(len=3) :: mxtype, rhstyp
    character(len=16) :: ptrfmt, indfmt
    character(len=20) :: valfmt, rhsfmt
    integer :: u, ios
    integer(psb_lpk_) :: totcrd, ptrcrd, indcrd, valcrd, rhscrd, nrow, ncol, nnz, neltvl, nrhs, nrhsix
    integer(psb_lpk_) :: j, k, q
    integer(psb_lpk_), allocatable :: colptr(:), rowind(:)
    real(psb_dpk_), allocatable :: vals(:)
    logical :: sym, skew
    info = psb_success_
    u = 12
    if (present(iunit)) u = iunit
    if (.not. present(filename)) then
      info = psb_err_invalid_args_combination_
      return
    end if
    open(u, file=filename, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'hb_read', a_err=trim(filename))
      return
    end if
    rhscrd = 0
    read(u, '(a72,a8)', iostat=ios) title, key
    if (ios == 0) read(u, '(5i14)', iostat=ios) totcrd, ptrcrd, indcrd, valcrd, rhscrd
    if (ios == 0) read(u, '(a3,11x,4i14)', iostat=ios) mxtype, nrow, ncol, nnz, neltvl
    if (ios == 0) read(u, '(2a16,2a20)', iostat=ios) ptrfmt, indfmt, valfmt, rhsfmt
    nrhs = 0
    rhstyp = ' '
    if (ios == 0 .and. rhscrd > 0) read(u, '(a3,11x,2i14)', iostat=ios) rhstyp, nrhs, nrhsix
    mxtype = psb_toupper(mxtype)
    rhstyp = psb_toupper(rhstyp)
    if (ios /= 0 .or. scan(mxtype(1:1), 'RP') == 0 .or. scan(mxtype(2:2), 'USZR') == 0 .or. mxtype(3:3) /= 'A' &
         & .or. nrow /= ncol) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'hb_read', a_err='bad header or unsupported type '//mxtype)
      close(u)
      return
    end if
    allocate(colptr(ncol + 1), rowind(nnz), vals(nnz))
    read(u, ptrfmt, iostat=ios) colptr
    if (ios == 0) read(u, indfmt, iostat=ios) rowind
    if (ios == 0) then
      if (mxtype(1:1) == 'R') then
        read(u, valfmt, iostat=ios) vals
      else
        vals = done
      end if
    end if
    if (ios == 0 .and. nrhs > 0 .and. rhstyp(1:1) == 'F' .and. present(b)) then
      if (allocated(b)) deallocate(b)
      allocate(b(nrow, nrhs))
      if (len_trim(rhsfmt) == 0) rhsfmt = valfmt
      read(u, rhsfmt, iostat=ios) b
    end if
    close(u)
    if (ios /= 0) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'hb_read', a_err='short or malformed data cards')
      return
    end if
    sym = mxtype(2:2) == 'S'
    skew = mxtype(2:2) == 'Z'
    allocate(a%ia(2 * nnz), a%ja(2 * nnz), a%val(2 * nnz))
    q = 0
    do j = 1, ncol
      do k = colptr(j), colptr(j + 1) - 1
        q = q + 1
        a%ia(q) = rowind(k)
        a%ja(q) = j
        a%val(q) = vals(k)
        if ((sym .or. skew) .and. rowind(k) /= j) then
          q = q + 1
          a%ia(q) = j
          a%ja(q) = rowind(k)
          a%val(q) = merge(-vals(k), vals(k), skew)
        end if
      end do
    end do
    a%m = nrow
    a%nz = q
    if (present(mtitle)) mtitle = title
  end subroutine hb_read

  ! The same file assembled as a device matrix on the context psb_init
  ! created (serial drivers read straight into psb_dspmat_type,
  ! examples/serial/serialtest.f90:77).
  subroutine mm_dmat_read(a, info, iunit, filename)
    type(psb_dspmat_type), intent(inout) :: a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iunit
    character(len=*), intent(in), optional :: filename
    type(psb_ldspmat_type) :: g
    type(psb_desc_type) :: desc
    call mm_lmat_read(g, info, iunit, filename)
    if (info /= psb_success_) return
    call psb_matdist(g, a, psb_default_ctxt, desc, info)
  end subroutine mm_dmat_read

  ! Dense vector as a MatrixMarket array file (one column).
  subroutine mm_array_write(y, title, info, iunit, filename)
    real(psb_dpk_), intent(in) :: y(:)
    character(len=*), intent(in) :: title
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iunit
    character(len=*), intent(in), optional :: filename
    integer :: u, ios, i
    info = psb_success_
    u = 12
    if (present(iunit)) u = iunit
    if (present(filename)) then
      open(u, file=filename, status='replace', action='write', iostat=ios)
      if (ios /= 0) then
        info = psb_err_from_subroutine_
        return
      end if
    end if
    write(u, '(a)') '%%MatrixMarket matrix array real general'
    write(u, '(a,a)') '% ', trim(title)
    write(u, '(i0,1x,i0)') size(y), 1
    do i = 1, size(y)
      write(u, '(es24.16)') y(i)
    end do
    if (present(filename)) close(u)
  end subroutine mm_array_write

  ! MatrixMarket dense array (column-major, general): b(nrows, ncols).
  subroutine mm_array_read(b, info, iunit, filename)
    real(psb_dpk_), allocatable, intent(out) :: b(:, :)
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iunit
    character(len=*), intent(in), optional :: filename
    character(len=1024) :: line
    character(len=32) :: w(8)
    integer :: u, nw, ios
    integer(psb_lpk_) :: nr, nc, i, j
    info = psb_success_
    u = 12
    if (present(iunit)) u = iunit
    open(u, file=filename, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_array_read', a_err=trim(filename))
      return
    end if
    read(u, '(a)', iostat=ios) line
    call lower_words(line, w, nw)
    if (ios /= 0 .or. nw < 4 .or. w(1) /= '%%matrixmarket' .or. w(3) /= 'array') then
      info = psb_err_from_subroutine_
      call psb_errpush(info, 'mm_array_read', a_err='not a MatrixMarket array')
      close(u)
      return
    end if
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      if (line(1:1) /= '%') exit
    end do
    read(line, *) nr, nc
    allocate(b(nr, nc))
    do j = 1, nc
      do i = 1, nr
        read(u, *, iostat=ios) b(i, j)
        if (ios /= 0) then
          info = psb_err_from_subroutine_
          close(u)
          return
        end if
      end do
    end do
    close(u)
  end subroutine mm_array_read

  ! Distribute the global matrix: with one process the descriptor owns every
  ! row and the local matrix is the whole matrix, assembled on the device in
  ! storage `fmt`.
  subroutine psb_matdist(aux_a, a, ctxt, desc_a, info, fmt, parts, vg)
    type(psb_ldspmat_type), intent(inout) :: aux_a
    type(psb_dspmat_type), intent(inout) :: a
    type(psb_ctxt_type), intent(in) :: ctxt
    type(psb_desc_type), intent(out) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    character(len=*), intent(in), optional :: fmt
    external :: parts
    optional :: parts
    integer(psb_ipk_), intent(in), optional :: vg(:)
    integer(psb_ipk_), allocatable :: ia(:), ja(:)
    call psb_cdall(ctxt, desc_a, info, nl=int(aux_a%m, psb_ipk_))
    if (info /= psb_success_) return
    call psb_spall(a, desc_a, info, nnz=int(aux_a%nz, psb_ipk_))
    if (info /= psb_success_) return
    ia = int(aux_a%ia(1:aux_a%nz), psb_ipk_)
    ja = int(aux_a%ja(1:aux_a%nz), psb_ipk_)
    call psb_spins(int(aux_a%nz, psb_ipk_), ia, ja, aux_a%val(1:aux_a%nz), a, desc_a, info)
    if (info /= psb_success_) return
    call psb_cdasb(desc_a, info)
    if (present(fmt)) then
      call psb_spasb(a, desc_a, info, afmt=fmt)
    else
      call psb_spasb(a, desc_a, info)
    end if
  end subroutine psb_matdist

end module psb_util_mod
