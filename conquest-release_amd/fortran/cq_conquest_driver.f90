! Fortran host of the drop-in at CONQUEST's own data structures: reads a
! one-process index set (the arrays a CONQUEST run holds in its matrix_halo,
! matrix_trans, matrix(:) and bindex, plus the operand values) from a stream
! file, fills the index types of module cqmm_conquest from them, and calls
! cq_gpu_mat_mult_x for one mult(x) entry twice: with B, then with 2B, which
! reuses the cached plan. Writes both results to a raw file.
!
! File layout (native-endian stream): int32 mult_type, n_prim, groups_on_node,
! ncovery, ncoverz; per halo (A then C): int32 np_in_halo, ni_in_halo, int64
! len_i_halo, len_i_hbeg, then int32 nh_part, j_beg, lab_hcell, lab_hcover
! (np_in_halo each), j_seq, ndimj (ni_in_halo each), i_h2d (n_prim*ni_in_halo),
! i_halo, i_hbeg, ndimi (n_prim); ltrans: int64 len_i_prim, int32 n_hnab,
! i_beg, i_nd_beg (A halo atoms each), i_prim; bmat: int32 array_posn,
! n_atoms, part_nabs, nd_offset (groups_on_node each), int64 len_bindex, int32
! bindex; parts: int32 n_cell_parts, i_cc2seq; operands: int64 lena, lenb,
! lenc, real64 a, b, c.
!
! usage: cq_conquest_driver infile outfile
program cq_conquest_driver
  use iso_c_binding
  use cqmm_f
  use cqmm_conquest
  implicit none
  integer, parameter :: mult_id = 7
  type(cq_matrix_mult) :: abc
  type(cq_matrix_halo), target :: ah, ch
  type(cq_matrix_trans), target :: lt
  type(cq_group_set), target :: parts
  type(cq_primary_set), target :: prim
  type(cq_cover_set), target :: gcs
  type(cq_matrix), pointer :: bmat(:)
  integer(c_int32_t), pointer :: bindex(:), i_cc2seq(:)
  integer(c_int32_t), allocatable :: scal(:)
  integer(c_int32_t) :: mult_type, n_prim, ngroups, n_cell
  integer(c_int64_t) :: len_i_prim, len_bindex, lena, lenb, lenc
  real(c_double), allocatable, target :: a(:), b(:), c(:), r1(:)
  type(c_ptr) :: p1, p2
  character(len=512) :: infile, outfile
  integer :: u, nn

  call get_command_argument(1, infile)
  call get_command_argument(2, outfile)
  open(newunit=u, file=trim(infile), access='stream', form='unformatted', status='old')
  read (u) mult_type, n_prim, ngroups, gcs%ncovery, gcs%ncoverz
  call read_halo(u, n_prim, ah)
  call read_halo(u, n_prim, ch)
  read (u) len_i_prim
  allocate(lt%n_hnab(ah%ni_in_halo), lt%i_beg(ah%ni_in_halo), lt%i_nd_beg(ah%ni_in_halo), lt%i_prim(len_i_prim))
  read (u) lt%n_hnab, lt%i_beg, lt%i_nd_beg, lt%i_prim
  allocate(bmat(ngroups), scal(ngroups))
  read (u) scal
  do nn = 1, ngroups
     bmat(nn)%array_posn = scal(nn)
  end do
  read (u) scal
  do nn = 1, ngroups
     bmat(nn)%n_atoms = scal(nn)
  end do
  read (u) scal
  do nn = 1, ngroups
     bmat(nn)%part_nabs = scal(nn)
  end do
  read (u) scal
  do nn = 1, ngroups
     bmat(nn)%nd_offset = scal(nn)
  end do
  read (u) len_bindex
  allocate(bindex(len_bindex))
  read (u) bindex
  read (u) n_cell
  allocate(i_cc2seq(n_cell))
  read (u) i_cc2seq
  read (u) lena, lenb, lenc
  allocate(a(lena), b(lenb), c(lenc), r1(merge(lenc, lena, mult_type == 1)))
  read (u) a, b, c
  close(u)

  ! The matrix_mult entry, as immi fills mult(x) (src/mult_module.f90:271-1102).
  parts%i_cc2seq => i_cc2seq
  prim%n_prim = n_prim
  prim%groups_on_node = ngroups
  abc%mult_type = mult_type
  abc%parts => parts
  abc%prim => prim
  abc%gcs => gcs
  abc%bmat => bmat
  abc%ahalo => ah
  abc%chalo => ch
  abc%ltrans => lt
  abc%bindex => bindex

  call cq_gpu_check(cqmm_init(0), 'init')
  call cq_gpu_mat_mult_x(0, a, int(lena), b, int(lenb), c, int(lenc), abc, mult_id)
  p1 = cq_gpu_plan_of(mult_id, abc, int(lena), int(lenb), int(lenc))
  if (mult_type == 1) then
     r1 = c
  else
     r1 = a
  end if
  b = 2.0_c_double * b
  call cq_gpu_mat_mult_x(0, a, int(lena), b, int(lenb), c, int(lenc), abc, mult_id)
  p2 = cq_gpu_plan_of(mult_id, abc, int(lena), int(lenb), int(lenc))
  if (.not. c_associated(p1, p2)) error stop 'plan cache: mult(x) rebuilt'
  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write (u) r1
  if (mult_type == 1) then
     write (u) c
  else
     write (u) a
  end if
  close(u)
  write (*, '(a,i0,a)') 'cq_conquest_driver: mult ', mult_id, ' plan built once, reused'
  call cq_gpu_forget_plans()

contains

  subroutine read_halo(u, n_prim, h)
    integer, intent(in) :: u
    integer(c_int32_t), intent(in) :: n_prim
    type(cq_matrix_halo), intent(inout) :: h
    integer(c_int64_t) :: len_i_halo, len_i_hbeg
    read (u) h%np_in_halo, h%ni_in_halo, len_i_halo, len_i_hbeg
    allocate(h%nh_part(h%np_in_halo), h%j_beg(h%np_in_halo), h%lab_hcell(h%np_in_halo), &
             h%lab_hcover(h%np_in_halo), h%j_seq(h%ni_in_halo), h%ndimj(h%ni_in_halo), &
             h%i_h2d(int(n_prim, c_int64_t) * h%ni_in_halo), h%i_halo(len_i_halo), h%i_hbeg(len_i_hbeg), &
             h%ndimi(n_prim))
    read (u) h%nh_part, h%j_beg, h%lab_hcell, h%lab_hcover, h%j_seq, h%ndimj, h%i_h2d, h%i_halo, h%i_hbeg, h%ndimi
  end subroutine read_halo

end program cq_conquest_driver
