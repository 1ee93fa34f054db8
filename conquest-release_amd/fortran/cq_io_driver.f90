! Fortran host for the matrix files of libcqmm (host-only calls, no device):
! writes a seeded system's matrix in CONQUEST's dump_matrix2 layout (binary
! and text, src/store_matrix_module.f90:340-470), its InfoGlobal file
! (dump_InfoMatGlobal, :700-781) and the older dump_matrix text dump
! (src/io_module.f90:1960-2012), reads each back, and writes the file names,
! block counts and largest differences to a text file.
!
! usage: cq_io_driver natoms nsf seed dir outfile   (range 6 A)
program cq_io_driver
  use iso_c_binding
  use cqmm_f
  implicit none
  integer(c_int32_t) :: natoms, nsf1, ngrid(3), rg, nspin
  integer(c_int64_t) :: seed, loaded, skipped
  real(c_double) :: edge, box(3), prm(2)
  real(c_double), allocatable, target :: xyz(:), vals(:), back(:), vel(:), coords(:), vel2(:)
  integer(c_int32_t), allocatable, target :: nsf(:), nodes(:)
  type(c_ptr), target :: ptrs(1)
  type(c_ptr) :: sys
  type(cqmm_matrix_index) :: idx
  type(cqmm_info_global) :: info
  character(len=256) :: arg, dir, outfile
  character(kind=c_char, len=512) :: name2, nameg, namem
  character(len=512) :: p2b, p2t, pg, pm
  integer :: u, k

  call get_command_argument(1, arg); read (arg, *) natoms
  call get_command_argument(2, arg); read (arg, *) nsf1
  call get_command_argument(3, arg); read (arg, *) seed
  call get_command_argument(4, dir)
  call get_command_argument(5, outfile)

  allocate(xyz(3 * natoms), nsf(natoms), vel(3 * natoms), coords(3 * natoms), vel2(3 * natoms), nodes(natoms))
  nsf = nsf1
  call cq_gpu_check(cqmm_random_structure(natoms, 0.05_c_double, 2.0_c_double, seed, xyz, edge), 'structure')
  box = edge
  ngrid = 2
  call cq_gpu_check(cqmm_sys_create(natoms, xyz, box, nsf, ngrid, 6.0_c_double, sys), 'sys_create')
  call cq_gpu_check(cqmm_sys_add_range(sys, 6.0_c_double, rg), 'range')
  call cq_gpu_check(cqmm_sys_matrix_index(sys, rg, idx), 'matrix_index')
  allocate(vals(idx%length), back(idx%length))
  prm = [0.1_c_double, 1.5_c_double]
  call cq_gpu_check(cqmm_sys_fill(sys, rg, 1, seed, 11_c_int64_t, prm, 2, vals), 'fill')

  ! dump_matrix2 / grab: binary and text, read back into the same range.
  call cq_gpu_check(cqmm_matrix2_file_name('L'//c_null_char, 0, 1, name2, len(name2)), 'matrix2 name')
  p2b = trim(dir)//'/'//c_str(name2)
  p2t = trim(dir)//'/text_'//c_str(name2)
  ptrs(1) = c_loc(vals)
  call cq_gpu_check(cqmm_sys_dump_matrix2(sys, rg, 1, c_loc(ptrs), trim(p2b)//c_null_char, 1), 'dump_matrix2 binary')
  call cq_gpu_check(cqmm_sys_dump_matrix2(sys, rg, 1, c_loc(ptrs), trim(p2t)//c_null_char, 0), 'dump_matrix2 text')
  open(newunit=u, file=trim(outfile), status='replace')
  write (u, '(a)') c_str(name2)
  back = 0
  ptrs(1) = c_loc(back)
  call cq_gpu_check(cqmm_sys_load_matrix2(sys, rg, trim(p2b)//c_null_char, 1, 1, c_loc(ptrs), nspin, loaded, skipped), &
                    'load_matrix2 binary')
  write (u, '(i0,1x,i0,1x,i0,1x,es25.17)') nspin, loaded, skipped, maxval(abs(back - vals))
  back = 0
  call cq_gpu_check(cqmm_sys_load_matrix2(sys, rg, trim(p2t)//c_null_char, 0, 1, c_loc(ptrs), nspin, loaded, skipped), &
                    'load_matrix2 text')
  write (u, '(i0,1x,i0,1x,i0,1x,es25.17)') nspin, loaded, skipped, maxval(abs(back - vals))

  ! InfoGlobal with velocities and an MD step.
  do k = 1, 3 * natoms
    vel(k) = 1.0e-3_c_double * real(mod(k * 37, 101) - 50, c_double)
  end do
  call cq_gpu_check(cqmm_info_global_file_name(4, nameg, len(nameg)), 'info name')
  pg = trim(dir)//'/'//c_str(nameg)
  call cq_gpu_check(cqmm_sys_dump_info_global(sys, trim(pg)//c_null_char, 4, 1, 17, c_loc(vel)), 'dump_info_global')
  call cq_gpu_check(cqmm_load_info_global(trim(pg)//c_null_char, 4, natoms, info, c_loc(nodes), c_loc(coords), &
                    c_loc(vel2)), 'load_info_global')
  write (u, '(a)') c_str(nameg)
  write (u, '(i0,1x,i0,1x,3(i0,1x),i0,1x,i0,1x,i0,1x,i0)') info%natoms, info%numprocs, info%npcell, info%md_step, &
       info%has_velocity, info%has_md_step, info%index
  write (u, '(es25.17,1x,es25.17,1x,i0)') maxval(abs(vel2 - vel)), info%rcell(1), sum(nodes)
  write (u, '(es25.17)') (coords(k), k = 1, 3 * natoms)

  ! dump_matrix / grab_matrix.
  call cq_gpu_check(cqmm_matrix_file_name('K'//c_null_char, 1, 1, namem, len(namem)), 'matrix name')
  pm = trim(dir)//'/'//c_str(namem)
  call cq_gpu_check(cqmm_dump_matrix_text(trim(pm)//c_null_char, vals, idx%length), 'dump_matrix')
  back = 0
  call cq_gpu_check(cqmm_grab_matrix_text(trim(pm)//c_null_char, back, idx%length), 'grab_matrix')
  write (u, '(a)') c_str(namem)
  write (u, '(es25.17)') maxval(abs(back - vals))
  close(u)
  call cq_gpu_check(cqmm_sys_destroy(sys), 'sys_destroy')

contains

  ! The characters of a NUL-terminated C string.
  function c_str(s) result(r)
    character(kind=c_char, len=*), intent(in) :: s
    character(len=:), allocatable :: r
    integer :: n
    n = index(s, c_null_char) - 1
    if (n < 0) n = len_trim(s)
    r = s(1:n)
  end function c_str

end program cq_io_driver
