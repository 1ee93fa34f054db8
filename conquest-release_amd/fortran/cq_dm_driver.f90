! Fortran host for the density-matrix loops through the mult_module-style
! handle API of libcqmm: registers the ranges of a seeded periodic system,
! creates mult(x) plans as CONQUEST's immi does, and runs the O(N) sequence
! Hotelling S^-1 (src/S_matrix_module.f90:746-1034) -> InitMcW
! (src/McWeeny.f90:340-570) -> McWMin (src/McWeeny.f90:107-272) -> FindMinDM
! (src/DMMinModule.f90:172-312, LNV functional src/mult_module.f90:1478-1823).
! S and H follow the synthetic recipe (include/cqmm_system.h kinds 1 and 2)
! plus 3 S. Writes the scalar results to a text file.
!
! usage: cq_dm_driver natoms nsf seed outfile   (ranges S = H = T 4.5 A, L 6 A)
program cq_dm_driver
  use iso_c_binding
  use cqmm_f
  implicit none
  integer(c_int32_t) :: natoms, nsf1, ngrid(3), st_iter
  integer(c_int64_t) :: seed
  real(c_double) :: edge, box(3), rS, rL, omega, mu_lambda(2)
  real(c_double), allocatable :: xyz(:), svals(:), hvals(:)
  integer(c_int32_t), allocatable :: nsf(:)
  type(c_ptr) :: sys, ctx
  character(len=256) :: arg, outfile
  ! Range ids (S = H = T share one range, as do LS = LH = TL and TS = TH).
  integer(c_int32_t) :: rgS, rgL, rgLS, rgLSL, rgTS, rgSLS
  integer(c_int32_t) :: matS, matH, matT, matL, matK
  type(cqmm_hotelling_setup) :: hs
  type(cqmm_initmcw_setup) :: is
  type(cqmm_mcw_setup) :: ms
  type(cqmm_mcw_result) :: mr
  type(cqmm_dmm_setup) :: ds
  type(cqmm_dmm_result) :: dr
  real(c_double) :: prm_s(2), prm_h(4), ne
  integer :: u
  ! mult ids (names as in src/mult_module.f90:128-160)
  integer(c_int32_t), parameter :: L_S_LS = 1, LS_L_LSL = 2, LSL_SL_L = 3, T_S_TS = 4, TS_T_T = 5, T_H_TH = 6, &
       TH_T_L = 7, L_H_LH = 8, HL_S_LSL = 9, HLS_LS_L = 10, SL_S_SLS = 11, SLS_LS_L = 12, LSL_SL_H = 13, &
       T_L_TL = 14, TL_T_L = 15

  call get_command_argument(1, arg); read (arg, *) natoms
  call get_command_argument(2, arg); read (arg, *) nsf1
  call get_command_argument(3, arg); read (arg, *) seed
  call get_command_argument(4, outfile)
  rS = 4.5_c_double
  rL = 6.0_c_double

  allocate(xyz(3 * natoms), nsf(natoms))
  nsf = nsf1
  call cq_gpu_check(cqmm_random_structure(natoms, 0.05_c_double, 2.0_c_double, seed, xyz, edge), 'structure')
  box = edge
  ngrid = 2
  ! Cover-set radius: the largest range any product builds (LSL = 2L + S).
  call cq_gpu_check(cqmm_sys_create(natoms, xyz, box, nsf, ngrid, 2 * rL + rS, sys), 'sys_create')
  call cq_gpu_check(cqmm_sys_add_range(sys, rS, rgS), 'range S')
  call cq_gpu_check(cqmm_sys_add_range(sys, rL, rgL), 'range L')
  call cq_gpu_check(cqmm_sys_add_range(sys, rL + rS, rgLS), 'range LS')
  call cq_gpu_check(cqmm_sys_add_range(sys, 2 * rL + rS, rgLSL), 'range LSL')
  call cq_gpu_check(cqmm_sys_add_range(sys, 2 * rS, rgTS), 'range TS')
  call cq_gpu_check(cqmm_sys_add_range(sys, 2 * rS + rL, rgSLS), 'range SLS')
  call cq_gpu_check(cqmm_init(0), 'init')
  call cq_gpu_check(cqmm_ctx_create(ctx), 'ctx_create')
  call add_range(rgS, .true.)
  call add_range(rgL, .true.)
  call add_range(rgLS, .true.)
  call add_range(rgLSL, .true.)
  call add_range(rgTS, .false.)
  call add_range(rgSLS, .false.)
  call add_mult(L_S_LS, 1, rgL, rgS, rgLS)
  call add_mult(LS_L_LSL, 1, rgLS, rgL, rgLSL)
  call add_mult(LSL_SL_L, 2, rgL, rgLS, rgLSL)
  call add_mult(T_S_TS, 1, rgS, rgS, rgTS)
  call add_mult(TS_T_T, 2, rgS, rgS, rgTS)
  call add_mult(T_H_TH, 1, rgS, rgS, rgTS)
  call add_mult(TH_T_L, 1, rgTS, rgS, rgL)
  call add_mult(L_H_LH, 1, rgL, rgS, rgLS)
  call add_mult(HL_S_LSL, 1, rgLS, rgS, rgLSL)
  call add_mult(HLS_LS_L, 2, rgL, rgLS, rgLSL)
  call add_mult(SL_S_SLS, 1, rgLS, rgS, rgSLS)
  call add_mult(SLS_LS_L, 2, rgL, rgLS, rgSLS)
  call add_mult(LSL_SL_H, 2, rgS, rgLS, rgLSL)
  call add_mult(T_L_TL, 1, rgS, rgL, rgLS)
  call add_mult(TL_T_L, 2, rgL, rgS, rgLS)

  call cq_gpu_check(cqmm_matrix_allocate(ctx, rgS, matS), 'allocate S')
  call cq_gpu_check(cqmm_matrix_allocate(ctx, rgS, matH), 'allocate H')
  call cq_gpu_check(cqmm_matrix_allocate(ctx, rgS, matT), 'allocate T')
  call cq_gpu_check(cqmm_matrix_allocate(ctx, rgL, matL), 'allocate L')
  block
    type(cqmm_matrix_index) :: idx
    call cq_gpu_check(cqmm_sys_matrix_index(sys, rgS, idx), 'index S')
    allocate(svals(idx%length), hvals(idx%length))
  end block
  prm_s = [0.1_c_double, 1.5_c_double]
  prm_h = [-0.2_c_double, 2.0_c_double, 0.5_c_double, 1.5_c_double]
  call cq_gpu_check(cqmm_sys_fill(sys, rgS, 1, seed, 11_c_int64_t, prm_s, 2, svals), 'fill S')
  call cq_gpu_check(cqmm_sys_fill(sys, rgS, 2, seed, 12_c_int64_t, prm_h, 4, hvals), 'fill H')
  hvals = hvals + 3.0_c_double * svals
  call cq_gpu_check(cqmm_matrix_set(ctx, matS, svals), 'set S')
  call cq_gpu_check(cqmm_matrix_set(ctx, matH, hvals), 'set H')

  ne = real((natoms * nsf1) / 2, c_double)
  hs = cqmm_hotelling_setup(rgS, rgS, rgTS, T_S_TS, TS_T_T, 50, 1.0e-10_c_double)
  call cq_gpu_check(cqmm_hotelling(ctx, hs, matS, matT, st_iter, omega), 'hotelling')
  is = cqmm_initmcw_setup(rgS, rgS, rgS, rgL, rgTS, rgTS, T_S_TS, T_H_TH, TH_T_L, ne)
  call cq_gpu_check(cqmm_init_mcweeny(ctx, is, matH, matS, matT, matL, mu_lambda), 'init_mcweeny')
  ms = cqmm_mcw_setup(rgL, rgS, rgS, rgLS, rgLSL, L_S_LS, LS_L_LSL, LSL_SL_L, 100, 0, 2, 1.0e-6_c_double)
  call cq_gpu_check(cqmm_mcweeny(ctx, ms, matH, matS, matL, mr), 'mcweeny')
  ds%lnv = cqmm_lnv_setup(rgL, rgS, rgS, rgLS, rgLS, rgLSL, rgSLS, L_S_LS, L_H_LH, HL_S_LSL, HLS_LS_L, &
                          SL_S_SLS, SLS_LS_L, LS_L_LSL, LSL_SL_H)
  ds%range_T = rgS
  ds%range_TL = rgLS
  ds%mult_T_L_TL = T_L_TL
  ds%mult_TL_T_L = TL_T_L
  ds%n_L_iterations = 100
  ds%vary_mu = 1
  ds%maxpulay = 5
  ds%global_tolerance = 0
  ds%n_atoms = natoms
  ds%tolerance = 1.0e-7_c_double
  ds%min_pulay_step = 1.0e-3_c_double
  ds%max_pulay_step = 0.1_c_double
  ds%lin_tol = 0.1_c_double
  ds%electrons_target = ne
  ds%delta_e = 0.0_c_double
  call cq_gpu_check(cqmm_matrix_allocate(ctx, rgS, matK), 'allocate K')
  call cq_gpu_check(cqmm_dmm(ctx, ds, matH, matS, matT, matL, matK, dr), 'dmm')

  open(newunit=u, file=trim(outfile), status='replace')
  write (u, '(i0,1x,i0,1x,es25.17,1x,es25.17)') st_iter, mr%iterations, mr%energy, mr%electrons
  write (u, '(i0,1x,i0,1x,i0,1x,es25.17,1x,es25.17,1x,es25.17)') dr%done, dr%early_iterations, dr%late_iterations, &
       dr%energy, dr%electrons, dr%residual
  close(u)
  write (*, '(a,i0,a,i0,a,i0,a,es20.12)') 'cq_dm_driver: McWeeny ', mr%iterations, ' steps; FindMinDM early ', &
       dr%early_iterations, ' late ', dr%late_iterations, ' band energy ', dr%energy
  call cq_gpu_check(cqmm_matrix_free(ctx, matK), 'free K')
  call cq_gpu_check(cqmm_ctx_destroy(ctx), 'ctx_destroy')
  call cq_gpu_check(cqmm_sys_destroy(sys), 'sys_destroy')

contains

  ! Register a range with the context, and its transpose pairs when asked.
  subroutine add_range(rg, trans)
    integer(c_int32_t), intent(in) :: rg
    logical, intent(in) :: trans
    type(cqmm_matrix_index) :: idx
    integer(c_int64_t) :: npairs
    integer(c_int64_t), allocatable, target :: p(:), q(:)
    integer(c_int32_t), allocatable, target :: n1(:), n2(:)
    call cq_gpu_check(cqmm_sys_matrix_index(sys, rg, idx), 'matrix_index')
    call cq_gpu_check(cqmm_ctx_add_range(ctx, rg, idx), 'ctx_add_range')
    if (.not. trans) return
    call cq_gpu_check(cqmm_sys_transpose_pairs(sys, rg, npairs, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), &
                      'transpose_pairs count')
    allocate(p(max(npairs, 1_c_int64_t)), q(max(npairs, 1_c_int64_t)), n1(max(npairs, 1_c_int64_t)), &
             n2(max(npairs, 1_c_int64_t)))
    call cq_gpu_check(cqmm_sys_transpose_pairs(sys, rg, npairs, c_loc(p), c_loc(q), c_loc(n1), c_loc(n2)), &
                      'transpose_pairs')
    call cq_gpu_check(cqmm_ctx_set_transpose(ctx, rg, npairs, c_loc(p), c_loc(q), c_loc(n1), c_loc(n2)), &
                      'set_transpose')
  end subroutine add_range

  ! mult(id): a plan from the system's index set, handed to the context.
  subroutine add_mult(id, mult_type, ra, rb, rc)
    integer(c_int32_t), intent(in) :: id, mult_type, ra, rb, rc
    type(c_ptr) :: plan
    call cq_gpu_check(cqmm_sys_plan_create(sys, mult_type, ra, rb, rc, plan), 'plan_create')
    call cq_gpu_check(cqmm_ctx_add_mult(ctx, id, ra, rb, rc, plan), 'ctx_add_mult')
  end subroutine add_mult

end program cq_dm_driver
