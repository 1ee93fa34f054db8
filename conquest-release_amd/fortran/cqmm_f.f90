! cqmm_f — iso_c_binding interface to libcqmm.so for a Fortran (CONQUEST) host.
!
! Interoperable copies of the C descriptor types of include/cqmm.h and
! include/cqmm_matrix.h, the bind(C) interfaces (the drop-in multiply, the
! one-process index builder, and the mult_module-style handle API with the
! McWeeny, InitMcW, Hotelling, LNV and FindMinDM loops), and two Fortran-level
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
     integer(c_int32_t) :: tile_family, tile_wgs
     integer(c_int64_t) :: n_mirror
  end type cqmm_plan_info

  ! type(matrix) index of one range, src/matrix_module.f90:63-82 (include/cqmm_matrix.h)
  type, bind(c) :: cqmm_matrix_index
     integer(c_int32_t) :: n_parts
     type(c_ptr) :: n_atoms, nab_offset, nd_offset
     integer(c_int32_t) :: n_prim
     type(c_ptr) :: n_nab, i_acc, i_nd_acc, ndimi, onsite, i_part, i_seq, ndimj
     integer(c_int64_t) :: n_nabs, length
  end type cqmm_matrix_index

  type, bind(c) :: cqmm_hotelling_setup
     integer(c_int32_t) :: range_T, range_S, range_TS
     integer(c_int32_t) :: mult_T_S_TS, mult_TS_T_T
     integer(c_int32_t) :: max_iter
     real(c_double) :: tol
  end type cqmm_hotelling_setup

  type, bind(c) :: cqmm_initmcw_setup
     integer(c_int32_t) :: range_T, range_S, range_H, range_L, range_TS, range_TH
     integer(c_int32_t) :: mult_T_S_TS, mult_T_H_TH, mult_TH_T_L
     real(c_double) :: n_electrons
  end type cqmm_initmcw_setup

  ! LNV_matrix_multiply's ranges and mult entries (src/mult_module.f90:1478-1823)
  type, bind(c) :: cqmm_lnv_setup
     integer(c_int32_t) :: range_L, range_S, range_H, range_LS, range_LH, range_LSL, range_SLS
     integer(c_int32_t) :: mult_L_S_LS, mult_L_H_LH, mult_HL_S_LSL, mult_HLS_LS_L, mult_SL_S_SLS, &
                           mult_SLS_LS_L, mult_LS_L_LSL, mult_LSL_SL_H
  end type cqmm_lnv_setup

  ! FindMinDM's parameters (src/DMMinModule.f90:172-312)
  type, bind(c) :: cqmm_dmm_setup
     type(cqmm_lnv_setup) :: lnv
     integer(c_int32_t) :: range_T, range_TL, mult_T_L_TL, mult_TL_T_L
     integer(c_int32_t) :: n_L_iterations, vary_mu, maxpulay, global_tolerance, n_atoms
     real(c_double) :: tolerance, min_pulay_step, max_pulay_step, lin_tol, electrons_target, delta_e
  end type cqmm_dmm_setup

  type, bind(c) :: cqmm_dmm_result
     integer(c_int32_t) :: done, early_iterations, late_iterations, inflexions
     real(c_double) :: energy, electrons, residual, seconds
  end type cqmm_dmm_result

  integer(c_int32_t), parameter :: CQMM_LNV_K = 1, CQMM_LNV_E = 2, CQMM_LNV_M3 = 4, CQMM_LNV_PHI = 8

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

  ! InfoGlobal file contents (dump_InfoMatGlobal / grab_InfoMatGlobal).
  type, bind(c) :: cqmm_info_global
     integer(c_int32_t) :: natoms, numprocs
     integer(c_int32_t) :: npcell(3)
     real(c_double) :: rcell(3)
     integer(c_int32_t) :: md_step, has_velocity, has_md_step, index
  end type cqmm_info_global

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
     function cqmm_plan_create_symmetric(desc, npairs, p, q, n1, n2, plan) &
         bind(c, name="cqmm_plan_create_symmetric") result(st)
       import :: c_int, c_ptr, c_int32_t, c_int64_t, cqmm_mult_desc
       type(cqmm_mult_desc), intent(in) :: desc
       integer(c_int64_t), value :: npairs
       integer(c_int64_t), intent(in) :: p(*), q(*)
       integer(c_int32_t), intent(in) :: n1(*), n2(*)
       type(c_ptr), intent(out) :: plan
       integer(c_int) :: st
     end function cqmm_plan_create_symmetric
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
     ! ---- mult_module handle API (include/cqmm_matrix.h) ----
     function cqmm_sys_plan_create(sys, mult_type, ra, rb, rc, plan) bind(c, name="cqmm_sys_plan_create") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: sys
       integer(c_int32_t), value :: mult_type, ra, rb, rc
       type(c_ptr) :: plan
       integer(c_int) :: st
     end function cqmm_sys_plan_create
     function cqmm_sys_plan_create_symmetric(sys, mult_type, ra, rb, rc, plan) &
         bind(c, name="cqmm_sys_plan_create_symmetric") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: sys
       integer(c_int32_t), value :: mult_type, ra, rb, rc
       type(c_ptr) :: plan
       integer(c_int) :: st
     end function cqmm_sys_plan_create_symmetric
     function cqmm_sys_matrix_index(sys, range_id, idx) bind(c, name="cqmm_sys_matrix_index") result(st)
       import :: c_ptr, c_int, c_int32_t, cqmm_matrix_index
       type(c_ptr), value :: sys
       integer(c_int32_t), value :: range_id
       type(cqmm_matrix_index) :: idx
       integer(c_int) :: st
     end function cqmm_sys_matrix_index
     function cqmm_sys_transpose_pairs(sys, range_id, npairs, p, q, n1, n2) &
         bind(c, name="cqmm_sys_transpose_pairs") result(st)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: sys
       integer(c_int32_t), value :: range_id
       integer(c_int64_t) :: npairs
       type(c_ptr), value :: p, q, n1, n2
       integer(c_int) :: st
     end function cqmm_sys_transpose_pairs
     function cqmm_ctx_create(ctx) bind(c, name="cqmm_ctx_create") result(st)
       import :: c_ptr, c_int
       type(c_ptr) :: ctx
       integer(c_int) :: st
     end function cqmm_ctx_create
     function cqmm_ctx_destroy(ctx) bind(c, name="cqmm_ctx_destroy") result(st)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: st
     end function cqmm_ctx_destroy
     function cqmm_ctx_add_range(ctx, range_id, idx) bind(c, name="cqmm_ctx_add_range") result(st)
       import :: c_ptr, c_int, c_int32_t, cqmm_matrix_index
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: range_id
       type(cqmm_matrix_index), intent(in) :: idx
       integer(c_int) :: st
     end function cqmm_ctx_add_range
     function cqmm_ctx_set_transpose(ctx, range_id, npairs, p, q, n1, n2) bind(c, name="cqmm_ctx_set_transpose") result(st)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: range_id
       integer(c_int64_t), value :: npairs
       type(c_ptr), value :: p, q, n1, n2
       integer(c_int) :: st
     end function cqmm_ctx_set_transpose
     function cqmm_ctx_add_mult(ctx, mult_id, ra, rb, rc, plan) bind(c, name="cqmm_ctx_add_mult") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mult_id, ra, rb, rc
       type(c_ptr), value :: plan
       integer(c_int) :: st
     end function cqmm_ctx_add_mult
     function cqmm_matrix_allocate(ctx, range_id, handle) bind(c, name="cqmm_matrix_allocate") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: range_id
       integer(c_int32_t) :: handle
       integer(c_int) :: st
     end function cqmm_matrix_allocate
     function cqmm_matrix_free(ctx, handle) bind(c, name="cqmm_matrix_free") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: handle
       integer(c_int) :: st
     end function cqmm_matrix_free
     function cqmm_matrix_set(ctx, handle, host) bind(c, name="cqmm_matrix_set") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: handle
       real(c_double), intent(in) :: host(*)
       integer(c_int) :: st
     end function cqmm_matrix_set
     function cqmm_matrix_get(ctx, handle, host) bind(c, name="cqmm_matrix_get") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: handle
       real(c_double) :: host(*)
       integer(c_int) :: st
     end function cqmm_matrix_get
     function cqmm_matrix_product(ctx, a, b, c, mult_id) bind(c, name="cqmm_matrix_product") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: a, b, c, mult_id
       integer(c_int) :: st
     end function cqmm_matrix_product
     function cqmm_matrix_sum(ctx, alpha, a, beta, b) bind(c, name="cqmm_matrix_sum") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: alpha, beta
       integer(c_int32_t), value :: a, b
       integer(c_int) :: st
     end function cqmm_matrix_sum
     function cqmm_matrix_product_trace(ctx, a, b, res) bind(c, name="cqmm_matrix_product_trace") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: a, b
       real(c_double) :: res
       integer(c_int) :: st
     end function cqmm_matrix_product_trace
     function cqmm_hotelling(ctx, setup, mats, matt, iterations, omega) bind(c, name="cqmm_hotelling") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double, cqmm_hotelling_setup
       type(c_ptr), value :: ctx
       type(cqmm_hotelling_setup), intent(in) :: setup
       integer(c_int32_t), value :: mats, matt
       integer(c_int32_t) :: iterations
       real(c_double) :: omega
       integer(c_int) :: st
     end function cqmm_hotelling
     function cqmm_init_mcweeny(ctx, setup, math, mats, matt, matl, mu_lambda) bind(c, name="cqmm_init_mcweeny") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double, cqmm_initmcw_setup
       type(c_ptr), value :: ctx
       type(cqmm_initmcw_setup), intent(in) :: setup
       integer(c_int32_t), value :: math, mats, matt, matl
       real(c_double) :: mu_lambda(2)
       integer(c_int) :: st
     end function cqmm_init_mcweeny
     function cqmm_mcweeny(ctx, setup, math, mats, matl, res) bind(c, name="cqmm_mcweeny") result(st)
       import :: c_ptr, c_int, c_int32_t, cqmm_mcw_setup, cqmm_mcw_result
       type(c_ptr), value :: ctx
       type(cqmm_mcw_setup), intent(in) :: setup
       integer(c_int32_t), value :: math, mats, matl
       type(cqmm_mcw_result) :: res
       integer(c_int) :: st
     end function cqmm_mcweeny
     function cqmm_mcweeny_spin(ctx, setup, fixed_populations, math, mats, matl, res) &
         bind(c, name="cqmm_mcweeny_spin") result(st)
       import :: c_ptr, c_int, c_int32_t, cqmm_mcw_setup, cqmm_mcw_result
       type(c_ptr), value :: ctx
       type(cqmm_mcw_setup), intent(in) :: setup
       integer(c_int32_t), value :: fixed_populations
       integer(c_int32_t), intent(in) :: math(2), mats(2), matl(2)
       type(cqmm_mcw_result) :: res(2)
       integer(c_int) :: st
     end function cqmm_mcweeny_spin
     function cqmm_init_mcweeny_spin(ctx, setup, fixed_populations, math, mats, matt, matl, n_electrons, &
         mu_lambda) bind(c, name="cqmm_init_mcweeny_spin") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double, cqmm_initmcw_setup
       type(c_ptr), value :: ctx
       type(cqmm_initmcw_setup), intent(in) :: setup
       integer(c_int32_t), value :: fixed_populations
       integer(c_int32_t), intent(in) :: math(2), mats(2), matt(2), matl(2)
       real(c_double), intent(in) :: n_electrons(2)
       real(c_double) :: mu_lambda(4)
       integer(c_int) :: st
     end function cqmm_init_mcweeny_spin
     function cqmm_lnv(ctx, setup, flags, math, mats, matl, matk, matm3, matphi, electrons, energy) &
         bind(c, name="cqmm_lnv") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double, cqmm_lnv_setup
       type(c_ptr), value :: ctx
       type(cqmm_lnv_setup), intent(in) :: setup
       integer(c_int32_t), value :: flags, math, mats, matl, matk, matm3, matphi
       real(c_double) :: electrons, energy
       integer(c_int) :: st
     end function cqmm_lnv
     function cqmm_correct_electrons(ctx, setup, math, mats, matt, matl, electrons) &
         bind(c, name="cqmm_correct_electrons") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double, cqmm_dmm_setup
       type(c_ptr), value :: ctx
       type(cqmm_dmm_setup), intent(in) :: setup
       integer(c_int32_t), value :: math, mats, matt, matl
       real(c_double) :: electrons
       integer(c_int) :: st
     end function cqmm_correct_electrons
     function cqmm_dmm(ctx, setup, math, mats, matt, matl, matk, res) bind(c, name="cqmm_dmm") result(st)
       import :: c_ptr, c_int, c_int32_t, cqmm_dmm_setup, cqmm_dmm_result
       type(c_ptr), value :: ctx
       type(cqmm_dmm_setup), intent(in) :: setup
       integer(c_int32_t), value :: math, mats, matt, matl, matk
       type(cqmm_dmm_result) :: res
       integer(c_int) :: st
     end function cqmm_dmm
     ! Matrix operations of mult_module (matrix_scale, matrix_trace, symmetrise_L,
     ! matrix_transpose) on context handles.
     function cqmm_matrix_length(ctx, a, length) bind(c, name="cqmm_matrix_length") result(st)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: a
       integer(c_int64_t) :: length
       integer(c_int) :: st
     end function cqmm_matrix_length
     function cqmm_matrix_scale(ctx, alpha, a) bind(c, name="cqmm_matrix_scale") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: alpha
       integer(c_int32_t), value :: a
       integer(c_int) :: st
     end function cqmm_matrix_scale
     function cqmm_matrix_trace(ctx, a, res) bind(c, name="cqmm_matrix_trace") result(st)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: a
       real(c_double) :: res
       integer(c_int) :: st
     end function cqmm_matrix_trace
     function cqmm_matrix_symmetrise(ctx, a) bind(c, name="cqmm_matrix_symmetrise") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: a
       integer(c_int) :: st
     end function cqmm_matrix_symmetrise
     function cqmm_matrix_transpose(ctx, a, b) bind(c, name="cqmm_matrix_transpose") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: a, b
       integer(c_int) :: st
     end function cqmm_matrix_transpose
     ! Kernel choice for a plan: fixed, or timed on the caller's operands.
     function cqmm_plan_set_kernel(plan, kernel) bind(c, name="cqmm_plan_set_kernel") result(st)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: plan
       integer(c_int32_t), value :: kernel
       integer(c_int) :: st
     end function cqmm_plan_set_kernel
     function cqmm_plan_tune(plan, a, lena, b, lenb, c, lenc, flags, reps, ms_tile, ms_triplet) &
          bind(c, name="cqmm_plan_tune") result(st)
       import :: c_ptr, c_int, c_int64_t, c_int32_t, c_float
       type(c_ptr), value :: plan, a, b, c
       integer(c_int64_t), value :: lena, lenb, lenc
       integer(c_int32_t), value :: flags, reps
       real(c_float) :: ms_tile, ms_triplet
       integer(c_int) :: st
     end function cqmm_plan_tune
     ! Matrix files (store_matrix_module / io_module formats). Strings are
     ! NUL-terminated: pass trim(name)//c_null_char.
     function cqmm_matrix2_file_name(root, index, inode, out, out_len) bind(c, name="cqmm_matrix2_file_name") result(st)
       import :: c_int, c_int32_t, c_char
       character(kind=c_char), dimension(*), intent(in) :: root
       integer(c_int32_t), value :: index, inode, out_len
       character(kind=c_char), dimension(*) :: out
       integer(c_int) :: st
     end function cqmm_matrix2_file_name
     function cqmm_sys_dump_matrix2(sys, range_id, nspin, values, path, binary) &
          bind(c, name="cqmm_sys_dump_matrix2") result(st)
       import :: c_ptr, c_int, c_int32_t, c_char
       type(c_ptr), value :: sys, values
       integer(c_int32_t), value :: range_id, nspin, binary
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int) :: st
     end function cqmm_sys_dump_matrix2
     function cqmm_sys_load_matrix2(sys, range_id, path, binary, nspin_max, outp, nspin, blocks_loaded, &
          blocks_skipped) bind(c, name="cqmm_sys_load_matrix2") result(st)
       import :: c_ptr, c_int, c_int32_t, c_int64_t, c_char
       type(c_ptr), value :: sys, outp
       integer(c_int32_t), value :: range_id, binary, nspin_max
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int32_t) :: nspin
       integer(c_int64_t) :: blocks_loaded, blocks_skipped
       integer(c_int) :: st
     end function cqmm_sys_load_matrix2
     function cqmm_info_global_file_name(index, out, out_len) bind(c, name="cqmm_info_global_file_name") result(st)
       import :: c_int, c_int32_t, c_char
       integer(c_int32_t), value :: index, out_len
       character(kind=c_char), dimension(*) :: out
       integer(c_int) :: st
     end function cqmm_info_global_file_name
     function cqmm_sys_dump_info_global(sys, path, index, has_md_step, md_step, velocity) &
          bind(c, name="cqmm_sys_dump_info_global") result(st)
       import :: c_ptr, c_int, c_int32_t, c_char
       type(c_ptr), value :: sys, velocity
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int32_t), value :: index, has_md_step, md_step
       integer(c_int) :: st
     end function cqmm_sys_dump_info_global
     function cqmm_load_info_global(path, index, natoms_max, info, glob_to_node, coords, velocity) &
          bind(c, name="cqmm_load_info_global") result(st)
       import :: c_ptr, c_int, c_int32_t, c_char, cqmm_info_global
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int32_t), value :: index, natoms_max
       type(cqmm_info_global) :: info
       type(c_ptr), value :: glob_to_node, coords, velocity
       integer(c_int) :: st
     end function cqmm_load_info_global
     function cqmm_matrix_file_name(stub, numprocs, inode, out, out_len) bind(c, name="cqmm_matrix_file_name") result(st)
       import :: c_int, c_int32_t, c_char
       character(kind=c_char), dimension(*), intent(in) :: stub
       integer(c_int32_t), value :: numprocs, inode, out_len
       character(kind=c_char), dimension(*) :: out
       integer(c_int) :: st
     end function cqmm_matrix_file_name
     function cqmm_dump_matrix_text(path, values, length) bind(c, name="cqmm_dump_matrix_text") result(st)
       import :: c_int, c_int64_t, c_double, c_char
       character(kind=c_char), dimension(*), intent(in) :: path
       real(c_double), dimension(*), intent(in) :: values
       integer(c_int64_t), value :: length
       integer(c_int) :: st
     end function cqmm_dump_matrix_text
     function cqmm_grab_matrix_text(path, values, length) bind(c, name="cqmm_grab_matrix_text") result(st)
       import :: c_int, c_int64_t, c_double, c_char
       character(kind=c_char), dimension(*), intent(in) :: path
       real(c_double), dimension(*) :: values
       integer(c_int64_t), value :: length
       integer(c_int) :: st
     end function cqmm_grab_matrix_text
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
