! nrg_driver.f90 -- Fortran host driving the GPU energy step through
! moons_amd_PCG_mod, as solve_MOONS does each step (add_all_energy_sources,
! then solve(nrg, ...); code/handwritten/MOONS/MOONS_solver.f90:80-83).
!
! usage: nrg_driver N beta method rk_stages nsteps in.bin out.bin
!   mesh: x uniform on [0,1], y and z Roberts beta (uniform when beta <= 1)
!   T: Dirichlet 1 / 0 on the x walls, Neumann elsewhere; dt 2e-3, Pe 50
!   (coeff_implicit -1/Pe), advection -1, diffusion 1/Pe, KE diffusion 0.02,
!   viscous dissipation 0.5, Joule heating 0.25, volumetric heating 0.7;
!   the T solver to 1e-12 (50 iterations at most); k = 1, 1.5, 2 on the x, y, z faces.
!   in.bin: T0, U (three faces), J (three edges), sigmaInv_CC one after the
!   other, real(8), ghost-padded, column-major; out.bin receives T then F.
program nrg_driver
  use iso_c_binding
  use moons_amd_bindings
  use moons_amd_PCG_mod
  implicit none
  integer :: N, method, rk, nsteps, u, step
  real(c_double) :: beta
  character(len=512) :: arg, inpath, outpath
  real(c_double), allocatable :: hx(:), hy(:)
  real(c_double), allocatable :: T(:, :, :), F(:, :, :), sig(:, :, :)
  real(c_double), allocatable :: U1(:, :, :), U2(:, :, :), U3(:, :, :), J1(:, :, :), J2(:, :, :), J3(:, :, :)
  real(c_double), allocatable :: k1(:, :, :), k2(:, :, :), k3(:, :, :)
  type(mesh_gpu) :: m
  type(energy_gpu) :: nrg
  type(moons_nrg_params) :: P
  type(c_ptr) :: hF
  integer(c_int) :: types(6)
  real(c_double) :: vals(6)

  call get_command_argument(1, arg); read(arg, *) N
  call get_command_argument(2, arg); read(arg, *) beta
  call get_command_argument(3, arg); read(arg, *) method
  call get_command_argument(4, arg); read(arg, *) rk
  call get_command_argument(5, arg); read(arg, *) nsteps
  call get_command_argument(6, inpath)
  call get_command_argument(7, outpath)

  allocate(hx(N + 3), hy(N + 3))
  call moons_check(moons_nodes_uniform(0.0_c_double, 1.0_c_double, int(N, c_int), hx), 'nodes')
  if (beta .gt. 1.0_c_double) then
    call moons_check(moons_nodes_roberts_both(0.0_c_double, 1.0_c_double, int(N, c_int), beta, hy), 'nodes')
  else
    hy = hx
  endif
  call init(m, hx, hy, hy)

  allocate(T(N + 2, N + 2, N + 2), F(N + 2, N + 2, N + 2), sig(N + 2, N + 2, N + 2))
  allocate(U1(N + 3, N + 2, N + 2), U2(N + 2, N + 3, N + 2), U3(N + 2, N + 2, N + 3))
  allocate(k1(N + 3, N + 2, N + 2), k2(N + 2, N + 3, N + 2), k3(N + 2, N + 2, N + 3))
  allocate(J1(N + 2, N + 3, N + 3), J2(N + 3, N + 2, N + 3), J3(N + 3, N + 3, N + 2))
  open(newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(u) T, U1, U2, U3, J1, J2, J3, sig
  close(u)
  k1 = 1.0_c_double; k2 = 1.5_c_double; k3 = 2.0_c_double

  types = MOONS_BC_NEUMANN
  types(1:2) = MOONS_BC_DIRICHLET
  vals = 0.0_c_double
  vals(1) = 1.0_c_double
  P%dt = 2.0d-3
  P%method = int(method, c_int)
  P%coeff_implicit = -1.0_c_double / 50.0_c_double
  P%advection = 1; P%advection_scale = -1.0_c_double
  P%diffusion = 1; P%diffusion_scale = 1.0_c_double / 50.0_c_double
  P%iter_max = 50; P%tol_rel = 1.0d-12; P%tol_abs = 1.0d-15; P%n_skip = 1
  P%KE_diffusion = 1; P%KE_diffusion_scale = 0.02_c_double
  P%viscous_dissipation = 1; P%viscous_dissipation_scale = 0.5_c_double
  P%joule_heating = 1; P%joule_heating_scale = 0.25_c_double
  P%volumetric_heating = 1; P%volumetric_heating_scale = 0.7_c_double
  P%rk_stages = int(rk, c_int)
  call init(nrg, m, types, vals, P, k1, k2, k3)
  call set_joule(nrg, m, J1, J2, J3, sig)
  do step = 1, nsteps
    call solve(nrg, T, U1, U2, U3, .true.)
  enddo
  call moons_check(moons_nrg_field(nrg%h, MOONS_NRG_F, hF), 'field F')
  call moons_check(moons_field_download(hF, F, int(size(F), c_size_t)), 'download F')

  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(u) T, F
  close(u)
  call delete(nrg)
  call delete(m)
end program
