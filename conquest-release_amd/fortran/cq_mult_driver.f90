! Fortran host exercising the drop-in: builds a one-process index set for a
! seeded random structure, creates the plan for one product and calls
! cq_gpu_mat_mult exactly where CONQUEST's matrix_product calls mat_mult
! (src/mult_module.f90:2591-2600). Writes the result array to a raw file.
!
! usage: cq_mult_driver natoms nsf ra rb rc mult_type seed outfile
program cq_mult_driver
  use iso_c_binding
  use cqmm_f
  implicit none
  integer(c_int32_t) :: natoms, nsf1, mult_type, ida, idb, idc, ngrid(3)
  real(c_double) :: ra, rb, rc, edge, box(3), rcover, prm(1)
  integer(c_int64_t) :: seed
  real(c_double), allocatable :: xyz(:)
  integer(c_int32_t), allocatable :: nsf(:)
  real(c_double), allocatable, target :: a(:), b(:), c(:)
  type(c_ptr) :: sys, plan
  type(cqmm_mult_desc) :: desc
  type(cqmm_plan_info) :: info
  character(len=256) :: arg, outfile
  integer :: u

  call get_command_argument(1, arg); read (arg, *) natoms
  call get_command_argument(2, arg); read (arg, *) nsf1
  call get_command_argument(3, arg); read (arg, *) ra
  call get_command_argument(4, arg); read (arg, *) rb
  call get_command_argument(5, arg); read (arg, *) rc
  call get_command_argument(6, arg); read (arg, *) mult_type
  call get_command_argument(7, arg); read (arg, *) seed
  call get_command_argument(8, outfile)

  allocate(xyz(3 * natoms), nsf(natoms))
  nsf = nsf1
  call cq_gpu_check(cqmm_random_structure(natoms, 0.05_c_double, 2.0_c_double, seed, xyz, edge), 'structure')
  box = edge
  ngrid = max(1, nint(edge / 4.0_c_double))
  rcover = max(ra + rb, rc)
  call cq_gpu_check(cqmm_sys_create(natoms, xyz, box, nsf, ngrid, rcover, sys), 'sys_create')
  call cq_gpu_check(cqmm_sys_add_range(sys, ra, ida), 'add_range a')
  call cq_gpu_check(cqmm_sys_add_range(sys, rb, idb), 'add_range b')
  call cq_gpu_check(cqmm_sys_add_range(sys, rc, idc), 'add_range c')
  call cq_gpu_check(cqmm_sys_mult_desc(sys, mult_type, ida, idb, idc, desc), 'mult_desc')
  call cq_gpu_check(cqmm_init(0), 'init')
  call cq_gpu_check(cqmm_plan_create(desc, plan), 'plan_create')
  call cq_gpu_check(cqmm_plan_get_info(plan, info), 'plan_info')

  allocate(a(desc%lena), b(desc%lenb), c(desc%lenc))
  prm(1) = 2.0_c_double
  call cq_gpu_check(cqmm_sys_fill(sys, idb, 0, seed, 2_c_int64_t, prm, 1, b), 'fill b')
  if (mult_type == 1) then
     call cq_gpu_check(cqmm_sys_fill(sys, ida, 0, seed, 1_c_int64_t, prm, 1, a), 'fill a')
     c = 0.0_c_double
  else
     call cq_gpu_check(cqmm_sys_fill(sys, idc, 0, seed, 3_c_int64_t, prm, 1, c), 'fill c')
     a = 0.0_c_double
  end if
  ! matrix_product(A,B,C): type 1 -> mat_mult(A,B,C); type 2 -> mat_mult(C,B,A)
  call cq_gpu_mat_mult(0, a, int(desc%lena), b, int(desc%lenb), c, int(desc%lenc), plan)

  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write (u) int(info%n_triplets, c_int64_t)
  if (mult_type == 1) then
     write (u) c
  else
     write (u) a
  end if
  close(u)
  write (*, '(a,i0,a,es12.5)') 'cq_mult_driver: products ', info%n_triplets, ' flops ', info%flops
  call cq_gpu_check(cqmm_plan_destroy(plan), 'plan_destroy')
  call cq_gpu_check(cqmm_sys_destroy(sys), 'sys_destroy')
end program cq_mult_driver
