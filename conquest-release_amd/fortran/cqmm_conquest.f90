! cqmm_conquest — the Fortran side of the drop-in for mat_mult: index types
! with the components CONQUEST's own types carry, their flattening into the C
! descriptor of include/cqmm.h, and a mat_mult wrapper that keeps one plan per
! mult(x) entry.
!
! The derived types below are this library's own. Their components have the
! names and meaning of the reference's types (matrix, matrix_trans, matrix_halo,
! matrix_mult: src/matrix_module.f90:63-203; the parts of group_set,
! primary_set and cover_set that mat_mult reads), so that a CONQUEST build can
! pass its own arrays component by component (or replace these types by its
! own, with the same component names, and keep the routines unchanged). No
! array is copied except the per-partition scalars of bmat(:), which the C
! descriptor wants as arrays (set_matrix_pointers2 keeps them per partition).
module cqmm_conquest
  use iso_c_binding
  use cqmm_f
  implicit none
  private
  public :: cq_matrix, cq_matrix_trans, cq_matrix_halo, cq_group_set, cq_primary_set, cq_cover_set, cq_matrix_mult
  public :: halo_desc, trans_desc, bmat_desc, mult_desc, cq_gpu_mat_mult_x, cq_gpu_forget_plans, cq_gpu_plan_of

  ! type(matrix), src/matrix_module.f90:63-82 (one primary-set partition)
  type cq_matrix
     integer :: length = 0
     integer(c_int32_t) :: array_posn = 0, offset = 0, nd_offset = 0
     integer(c_int32_t) :: n_atoms = 0, part_nabs = 0, part_nd_nabs = 0
     integer(c_int32_t), pointer :: n_nab(:) => null(), ndimi(:) => null(), ndimj(:) => null()
     integer(c_int32_t), pointer :: i_seq(:) => null(), i_acc(:) => null(), i_nd_acc(:) => null()
     integer(c_int32_t), pointer :: npxyz(:) => null(), i_part(:) => null(), onsite(:) => null()
  end type cq_matrix

  ! type(matrix_trans), src/matrix_module.f90:97-104
  type cq_matrix_trans
     integer(c_int32_t), pointer :: i_prim(:) => null(), i_beg(:) => null(), i_nd_beg(:) => null()
     integer(c_int32_t), pointer :: n_hnab(:) => null()
  end type cq_matrix_trans

  ! type(matrix_halo), src/matrix_module.f90:121-136
  type cq_matrix_halo
     integer(c_int32_t) :: np_in_halo = 0, ni_in_halo = 0
     integer(c_int32_t), allocatable :: nh_part(:), j_beg(:), lab_hcell(:), lab_hcover(:), j_seq(:)
     integer(c_int32_t), allocatable :: i_h2d(:), i_halo(:), i_hbeg(:), ndimi(:), ndimj(:)
  end type cq_matrix_halo

  ! The components of group_set, primary_set and cover_set that mat_mult reads.
  type cq_group_set
     integer(c_int32_t), pointer :: i_cc2seq(:) => null()
  end type cq_group_set
  type cq_primary_set
     integer(c_int32_t) :: n_prim = 0, groups_on_node = 0
  end type cq_primary_set
  type cq_cover_set
     integer(c_int32_t) :: ncovery = 0, ncoverz = 0
  end type cq_cover_set

  ! type(matrix_mult), src/matrix_module.f90:193-203, for one process (no comms)
  type cq_matrix_mult
     integer :: mult_type = 0
     type(cq_group_set), pointer :: parts => null()
     type(cq_primary_set), pointer :: prim => null()
     type(cq_cover_set), pointer :: gcs => null()
     type(cq_matrix), pointer :: bmat(:) => null()
     type(cq_matrix_halo), pointer :: ahalo => null(), chalo => null()
     type(cq_matrix_trans), pointer :: ltrans => null()
     integer(c_int32_t), pointer :: bindex(:) => null()
  end type cq_matrix_mult

  ! bmat(:)'s per-partition scalars as contiguous arrays, kept alive while a
  ! plan is built from the descriptor that points at them.
  type bmat_arrays
     integer(c_int32_t), allocatable :: array_posn(:), n_atoms(:), part_nabs(:), nd_offset(:)
  end type bmat_arrays

  ! One plan per mult(x) entry (the reference's mult(:) table, src/mult_module.f90:271-1102).
  integer, parameter :: max_mults = 256
  type(c_ptr), save :: plans(max_mults) = c_null_ptr

contains

  ! matrix_halo -> cqmm_halo (pointers into h; h must stay allocated).
  subroutine halo_desc(h, d)
    type(cq_matrix_halo), target, intent(in) :: h
    type(cqmm_halo), intent(out) :: d
    d%np_in_halo = h%np_in_halo
    d%ni_in_halo = h%ni_in_halo
    d%nh_part = c_loc(h%nh_part)
    d%j_beg = c_loc(h%j_beg)
    d%lab_hcell = c_loc(h%lab_hcell)
    d%lab_hcover = c_loc(h%lab_hcover)
    d%j_seq = c_loc(h%j_seq)
    d%ndimj = c_loc(h%ndimj)
    d%i_h2d = c_loc(h%i_h2d)
    d%i_halo = c_loc(h%i_halo)
    d%i_hbeg = c_loc(h%i_hbeg)
    d%ndimi = c_loc(h%ndimi)
    d%len_i_halo = size(h%i_halo, kind=c_int64_t)
    d%len_i_hbeg = size(h%i_hbeg, kind=c_int64_t)
  end subroutine halo_desc

  ! matrix_trans -> cqmm_trans.
  subroutine trans_desc(t, d)
    type(cq_matrix_trans), intent(in) :: t
    type(cqmm_trans), intent(out) :: d
    d%n_hnab = c_loc(t%n_hnab)
    d%i_beg = c_loc(t%i_beg)
    d%i_nd_beg = c_loc(t%i_nd_beg)
    d%i_prim = c_loc(t%i_prim)
    d%len_i_prim = size(t%i_prim, kind=c_int64_t)
  end subroutine trans_desc

  ! bmat(1:n_parts) and the packed bindex -> cqmm_bindex; the scalars of each
  ! partition are gathered into w.
  subroutine bmat_desc(bmat, n_parts, bindex, w, d)
    type(cq_matrix), intent(in) :: bmat(:)
    integer(c_int32_t), intent(in) :: n_parts
    integer(c_int32_t), pointer, intent(in) :: bindex(:)
    type(bmat_arrays), target, intent(inout) :: w
    type(cqmm_bindex), intent(out) :: d
    integer :: nn
    if (n_parts > size(bmat)) error stop 'bmat_desc: more partitions than bmat(:) holds'
    allocate(w%array_posn(max(1, n_parts)), w%n_atoms(max(1, n_parts)), w%part_nabs(max(1, n_parts)), &
             w%nd_offset(max(1, n_parts)))
    do nn = 1, n_parts
       w%array_posn(nn) = bmat(nn)%array_posn
       w%n_atoms(nn) = bmat(nn)%n_atoms
       w%part_nabs(nn) = bmat(nn)%part_nabs
       w%nd_offset(nn) = bmat(nn)%nd_offset
    end do
    d%n_parts = n_parts
    d%array_posn = c_loc(w%array_posn)
    d%n_atoms = c_loc(w%n_atoms)
    d%part_nabs = c_loc(w%part_nabs)
    d%nd_offset = c_loc(w%nd_offset)
    d%bindex = c_loc(bindex)
    d%len_bindex = size(bindex, kind=c_int64_t)
  end subroutine bmat_desc

  ! type(matrix_mult) + the array lengths mat_mult receives -> cqmm_mult_desc.
  subroutine mult_desc(abc, lena, lenb, lenc, w, d)
    type(cq_matrix_mult), intent(in) :: abc
    integer, intent(in) :: lena, lenb, lenc
    type(bmat_arrays), intent(inout) :: w
    type(cqmm_mult_desc), intent(out) :: d
    if (.not. (associated(abc%parts) .and. associated(abc%prim) .and. associated(abc%gcs) .and. &
               associated(abc%bmat) .and. associated(abc%ahalo) .and. associated(abc%chalo) .and. &
               associated(abc%ltrans) .and. associated(abc%bindex))) &
       error stop 'mult_desc: incomplete matrix_mult'
    d%mult_type = abc%mult_type
    d%n_prim = abc%prim%n_prim
    d%ncovery = abc%gcs%ncovery
    d%ncoverz = abc%gcs%ncoverz
    call halo_desc(abc%ahalo, d%ahalo)
    call halo_desc(abc%chalo, d%chalo)
    call trans_desc(abc%ltrans, d%ltrans)
    call bmat_desc(abc%bmat, abc%prim%groups_on_node, abc%bindex, w, d%bmat)
    d%i_cc2seq = c_loc(abc%parts%i_cc2seq)
    d%n_cell_parts = size(abc%parts%i_cc2seq, kind=c_int32_t)
    d%lena = lena
    d%lenb = lenb
    d%lenc = lenc
  end subroutine mult_desc

  ! The plan of mult(mult_id), built from abc on first use.
  function cq_gpu_plan_of(mult_id, abc, lena, lenb, lenc) result(plan)
    integer, intent(in) :: mult_id, lena, lenb, lenc
    type(cq_matrix_mult), intent(in) :: abc
    type(c_ptr) :: plan
    type(cqmm_mult_desc) :: d
    type(bmat_arrays) :: w
    if (mult_id < 1 .or. mult_id > max_mults) error stop 'cq_gpu_plan_of: mult id out of range'
    if (.not. c_associated(plans(mult_id))) then
       call mult_desc(abc, lena, lenb, lenc, w, d)
       call cq_gpu_check(cqmm_plan_create(d, plans(mult_id)), 'plan_create')
    end if
    plan = plans(mult_id)
  end function cq_gpu_plan_of

  ! mat_mult(myid, a, lena, b, lenb, c, lenc, a_b_c) (src/multiply_module.f90:115)
  ! on the GPU for the product mult(mult_id) = a_b_c: its plan is built on the
  ! first call and reused after. Host arrays; type 1 overwrites c, type 2 a.
  subroutine cq_gpu_mat_mult_x(myid, a, lena, b, lenb, c, lenc, abc, mult_id)
    integer, intent(in) :: myid, lena, lenb, lenc, mult_id
    real(c_double), target :: a(lena), b(lenb), c(lenc)
    type(cq_matrix_mult), intent(in) :: abc
    call cq_gpu_mat_mult(myid, a, lena, b, lenb, c, lenc, cq_gpu_plan_of(mult_id, abc, lena, lenb, lenc))
  end subroutine cq_gpu_mat_mult_x

  ! Drop every cached plan (atoms moved: the index set is rebuilt).
  subroutine cq_gpu_forget_plans()
    integer :: k
    do k = 1, max_mults
       if (c_associated(plans(k))) call cq_gpu_check(cqmm_plan_destroy(plans(k)), 'plan_destroy')
       plans(k) = c_null_ptr
    end do
  end subroutine cq_gpu_forget_plans

end module cqmm_conquest
