! ppe_driver.f90 -- Fortran host driving the GPU pressure-Poisson solve
! through moons_amd_PCG_mod, as a MOONS host would at clean_div/solve: the
! solver is built with the reference's init(PCG, operator, prec, m, ISP, MFP,
! x, x_BC, k, var_dir, var_name, testSymmetry, exportOperator) argument order
! (code/handwritten/governing_equations/momentum_extend.f90:166-167) and solved
! with its solve(PCG, x, b, m, compute_norms).
!
! usage: ppe_driver N beta prec iter_max tol b.bin x.bin [rawdir]
!   prec: MOONS_PREC_* of the PCG, or 0 for the cosine-transform direct
!         solver of the x-planes (init_FFT / solve; needs beta <= 1)
!   mesh: x uniform on [0,1], y and z Roberts beta (uniform when beta <= 1)
!   BCs:  homogeneous Neumann on all faces, all_Neumann
!   b.bin: the right-hand side, real(8), ghost-padded (N+2)^3, column-major
!   x.bin: the solution written back (x0 = 0)
!   rawdir: (PCG only) the device solution is also written as the raw
!         Tecplot file rawdir//'xc.dat' and the restart primitives
!         rawdir//'x_restart.txt'
! prints: iterations and the relative residual.
program ppe_driver
  use iso_c_binding
  use moons_amd_bindings
  use moons_amd_PCG_mod
  implicit none
  integer :: N, iter_max, prec_i, u
  real(c_double) :: beta, tol
  character(len=512) :: arg, bpath, xpath, rawdir
  real(c_double), allocatable :: hx(:), hy(:), x(:, :, :), b(:, :, :)
  type(mesh_gpu) :: m
  type(PCG_solver_SF_gpu) :: PCG
  type(FFT_solver_SF_gpu) :: FFT
  type(iter_solver_params_gpu) :: ISP
  type(matrix_free_params_gpu) :: MFP
  type(SF_gpu) :: p_gpu, k_none(3)
  integer(c_int) :: types(6)
  real(c_double) :: vals(6)

  call get_command_argument(1, arg); read(arg, *) N
  call get_command_argument(2, arg); read(arg, *) beta
  call get_command_argument(3, arg); read(arg, *) prec_i
  call get_command_argument(4, arg); read(arg, *) iter_max
  call get_command_argument(5, arg); read(arg, *) tol
  call get_command_argument(6, bpath)
  call get_command_argument(7, xpath)
  rawdir = ''
  if (command_argument_count() .ge. 8) call get_command_argument(8, rawdir)

  allocate(hx(N + 3), hy(N + 3))
  call moons_check(moons_nodes_uniform(0.0_c_double, 1.0_c_double, int(N, c_int), hx), 'nodes')
  if (beta .gt. 1.0_c_double) then
    call moons_check(moons_nodes_roberts_both(0.0_c_double, 1.0_c_double, int(N, c_int), beta, hy), 'nodes')
  else
    hy = hx
  endif
  call init(m, hx, hy, hy)

  allocate(x(N + 2, N + 2, N + 2), b(N + 2, N + 2, N + 2))
  open(newunit=u, file=trim(bpath), access='stream', form='unformatted', status='old')
  read(u) b
  close(u)
  x = 0.0_c_double

  types = MOONS_BC_NEUMANN
  vals = 0.0_c_double
  if (prec_i .eq. 0) then
    call init(FFT, m, 1, types, vals)
    call solve(FFT, x, b)
    write(*, '(a)') ' FFT solve done'
    call delete(FFT)
  else
    ! p with its BCs, as mom%p carries them into init
    call init_CC(p_gpu, m)
    call set_BCs(p_gpu, types, vals, .true.)
    call init(ISP, iter_max, tol, 1.0d-300, 1, .false., .false.)
    call init(PCG, MOONS_OP_LAP_UNIFORM_SF, int(prec_i, c_int), m, ISP, MFP, p_gpu, p_gpu, k_none, &
              'out/', 'p', .false., .false.)
    call solve(PCG, x, b, m, .true.)
    if (len_trim(rawdir) .gt. 0) then
      ! the device copy of the solution as a raw Tecplot file and as restart primitives
      call export_raw(m, PCG%x, trim(rawdir), 'x', 0)
      call export(PCG%x, trim(rawdir)//'x_restart.txt')
    endif
    call delete(PCG)
    call delete(p_gpu)
  endif

  open(newunit=u, file=trim(xpath), access='stream', form='unformatted', status='replace')
  write(u) x
  close(u)
  call delete(m)
end program
