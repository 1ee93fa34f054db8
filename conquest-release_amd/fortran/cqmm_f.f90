! cqmm_f — iso_c_binding interface to libcqmm.so for a Fortran (CONQUEST) host.
!
! Interoperable copies of the C descriptor types of include/cqmm.h and
! include/cqmm_matrix.h, the bind(C) interfaces, and two Fortran-level
! wrappers a CONQUEST build would call:
!   cq_gpu_mat_mult  — same argument list as mat_mult
!                      (src/multiply_module.f90:115) with a prebuilt plan in
!                      place of the type(matrix_mult) descriptor; a nonzero
!                      status becomes an abort with the library's message, as
!                      cq_abort does (src/generic_comms.f90:95-100).
!   cq_gpu_check     — status -> abort.
module cqmm_f
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: CQMM_DEVICE_PTRS = 1, CQMM_FP32 = 2, CQMM_NO_SYNC = 4

  ! type(matrix_halo), src/matrix_module.f90:121-136
  type, bind(c) :: cqmm_halo
     integer(c_int32_t) :: np_in_halo, ni_in_halo
     type(c_ptr) :: nh_part, j_beg, lab_hcell, lab_hcover, j_seq, ndimj, i_h2d, i_halo, i_hbeg, ndimi
     integer(c_int64_t) :: len_i_halo, len_i_hbeg
  end type cqmm_halo

  ! type(matrix_trans), src/matrix_module.f90:97-104
  type, bind(c) :: cqmm_trans
     type(c_ptr) :: n_hnab, i_beg, i_nd_beg, i_prim
     integer(c_int64_t) :: len_i_prim
  end type cqmm_trans

  ! B index: type(matrix) per partition + packed bindex (src/matrix_module.f90:350-421)
  type, bind(c) :: cqmm_bindex
     integer(c_int32_t) :: n_parts
     type(c_ptr) :: array_posn, n_atoms, part_nabs, nd_offset, bindex
     integer(c_int64_t) :: len_bindex
  end type cqmm_bindex

  ! type(matrix_mult), src/matrix_module.f90:193-203, for one process
  type, bind(c) :: cqmm_mult_desc
     integer(c_int32_t) :: mult_type, n_prim, ncovery, ncoverz
     type(cqmm_halo) :: ahalo, chalo
     type(cqmm_trans) :: ltrans
     type(cqmm_bindex) :: bmat
     type(c_ptr) :: i_cc2seq
     integer(c_int32_t) :: n_cell_parts
     integer(c_int64_t) :: lena, lenb, lenc
  end type cqmm_mult_desc

  type, bind(c) :: cqmm_plan_info
     integer(c_int64_t) :: n_triplets, n_out_blocks, n_out_total
     real(c_double) :: flops
     integer(c_int32_t) :: uniform_nsf, mult_type
     integer(c_int32_t) :: kernel, n_tiles
     integer(c_int64_t) :: n_entries, n_loads
  end type cqmm_plan_info

  type, bind(c) :: cqmm_mcw_setup
     integer(c_int32_t) :: range_L, range_S, range_H, range_LS, range_LSL
     integer(c_int32_t) :: mult_L_S_LS, mult_LS_L_LSL, mult_LSL_SL_L
     integer(c_int32_t) :: max_iter, precision, n_final_fp64
     real(c_double) :: fp32_tol
  end type cqmm_mcw_setup

  type, bind(c) :: cqmm_mcw_result
     integer(c_int32_t) :: iterations, fp32_iterations, stop_reason, restored
     real(c_double) :: energy, electrons, c, seconds
     real(c_double) :: energy_trace(256)
  end type cqmm_mcw_result

  interface
     function cqmm_last_error() bind(c, name="cqmm_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function cqmm_last_error
     function cqmm_init(device) bind(c, name="cqmm_init") result(st)
       import :: c_int
       integer(c_int), value :: device
       integer(c_int) :: st
     end function cqmm_init
     function cqmm_plan_create(desc, plan) bind(c, name="cqmm_plan_create") result(st)
       import :: c_int, c_ptr, cqmm_mult_desc
       type(cqmm_mult_desc), intent(in) :: desc
       type(c_ptr), intent(out) :: plan
       integer(c_int) :: st
     end function cqmm_plan_create
     function cqmm_plan_destroy(plan) bind(c, name="cqmm_plan_destroy") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: plan
       integer(c_int) :: st
     end function cqmm_plan_destroy
     function cqmm_plan_get_info(plan, info) bind(c, name="cqmm_plan_get_info") result(st)
       import :: c_int, c_ptr, cqmm_plan_info
       type(c_ptr), value :: plan
       type(cqmm_plan_info), intent(out) :: info
       integer(c_int) :: st
     end function cqmm_plan_get_info
     function cqmm_mat_mult(plan, a, lena, b, lenb, c, lenc, flags) bind(c, name="cqmm_mat_mult") result(st)
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: plan, a, b, c
       integer(c_int64_t), value :: lena, lenb, lenc
       integer(c_int32_t), value :: flags
       integer(c_int) :: st
     end function cqmm_mat_mult
     ! ---- host index mirror (include/cqmm_system.h) -------------------------
     function cqmm_random_structure(natoms, density, dmin, seed, xyz, edge) &
          bind(c, name="cqmm_random_structure") result(st)
       import :: c_int, c_int32_t, c_double, c_int64_t
       integer(c_int32_t), value :: natoms
       real(c_double), value :: density, dmin
       integer(c_int64_t), value :: seed
       real(c_double), intent(out) :: xyz(*), edge
       integer(c_int) :: st
     end function cqmm_random_structure
     function cqmm_sys_create(natoms, xyz, box, nsf, ngrid, rcut_cover, sys) &
          bind(c, name="cqmm_sys_create") result(st)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: natoms
       real(c_double), intent(in) :: xyz(*), box(3)
       integer(c_int32_t), intent(in) :: nsf(*), ngrid(3)
       real(c_double), value :: rcut_cover
       type(c_ptr), intent(out) :: sys
       integer(c_int) :: st
     end function cqmm_sys_create
     function cqmm_sys_destroy(sys) bind(c, name="cqmm_sys_destroy") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: sys
       integer(c_int) :: st
     end function cqmm_sys_destroy
     function cqmm_sys_add_range(sys, rcut, range_id) bind(c, name="cqmm_sys_add_range") result(st)
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: sys
       real(c_double), value :: rcut
       integer(c_int32_t), intent(out) :: range_id
       integer(c_int) :: st
     end function cqmm_sys_add_range
     function cqmm_sys_mult_desc(sys, mult_type, ra, rb, rc, desc) bind(c, name="cqmm_sys_mult_desc") result(st)
       import :: c_int, c_ptr, c_int32_t, cqmm_mult_desc
       type(c_ptr), value :: sys
       integer(c_int32_t), value :: mult_type, ra, rb, rc
       type(cqmm_mult_desc), intent(out) :: desc
       integer(c_int) :: st
     end function cqmm_sys_mult_desc
     function cqmm_sys_fill(sys, range_id, kind, seed, tag, p, np, values) bind(c, name="cqmm_sys_fill") result(st)
       import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
       type(c_ptr), value :: sys
       integer(c_int32_t), value :: range_id, kind, np
       integer(c_int64_t), value :: seed, tag
       real(c_double), intent(in) :: p(*)
       real(c_double), intent(out) :: values(*)
       integer(c_int) :: st
     end function cqmm_sys_fill
  end interface

contains

  ! Abort with the library's last message when status /= 0.
  subroutine cq_gpu_check(st, where)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (st == 0) return
    call c_f_pointer(cqmm_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write (0, '(a,a,i0,a,a)') trim(where), ': status ', st, ': ', trim(text)
    error stop 1
  end subroutine cq_gpu_check

  ! mat_mult(myid, a, lena, b, lenb, c, lenc, a_b_c) on the GPU: the plan stands
  ! for a_b_c (built once per mult(x) with cqmm_plan_create). Host arrays; type 1
  ! overwrites c, type 2 overwrites a, as the reference does.
  subroutine cq_gpu_mat_mult(myid, a, lena, b, lenb, c, lenc, plan)
    integer, intent(in) :: myid, lena, lenb, lenc
    real(c_double), target :: a(lena), b(lenb), c(lenc)
    type(c_ptr), intent(in) :: plan
    integer(c_int) :: st
    if (myid < 0) error stop 'cq_gpu_mat_mult: bad process id'
    st = cqmm_mat_mult(plan, c_loc(a), int(lena, c_int64_t), c_loc(b), int(lenb, c_int64_t), &
                       c_loc(c), int(lenc, c_int64_t), 0_c_int32_t)
    call cq_gpu_check(st, 'mat_mult')
  end subroutine cq_gpu_mat_mult

end module cqmm_f
