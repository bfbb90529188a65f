! vf_driver.f90 -- Fortran host driving the GPU face-centred implicit
! diffusion solve through moons_amd_PCG_mod, as solve_momentum /
! solve_induction do (init_PCG_VF, update_MFP, solve): the momentum solve with
! the BC-array init, the induction solve with the reference's argument order.
!
! usage: vf_driver N beta op c iter_max tol x0.bin b.bin x.bin [sig.bin]
!   mesh: x uniform on [0,1], y and z Roberts beta (uniform when beta <= 1)
!   op 1: mom_diffusion, lid-driven-cavity BCs (u_x = 1 on the y-max lid, no slip)
!   op 2: ind_diffusion, pseudo-vacuum BCs, sigmaInv on the x, y, z edges from sig.bin
!   x0.bin, b.bin: the three face components one after the other, real(8),
!   ghost-padded, column-major; x.bin receives the solution the same way.
program vf_driver
  use iso_c_binding
  use moons_amd_bindings
  use moons_amd_PCG_mod
  implicit none
  integer :: N, iter_max, op, u, d
  real(c_double) :: beta, c, tol
  character(len=512) :: arg, x0path, bpath, xpath, spath
  real(c_double), allocatable :: hx(:), hy(:)
  real(c_double), allocatable :: x1(:, :, :), x2(:, :, :), x3(:, :, :), b1(:, :, :), b2(:, :, :), b3(:, :, :)
  real(c_double), allocatable :: s1(:, :, :), s2(:, :, :), s3(:, :, :)
  type(mesh_gpu) :: m
  type(PCG_solver_VF_gpu) :: PCG
  type(iter_solver_params_gpu) :: ISP
  type(matrix_free_params_gpu) :: MFP
  type(SF_gpu) :: xbc(3), sig(3)
  integer(c_int) :: types(6, 3)
  real(c_double) :: vals(6, 3)

  call get_command_argument(1, arg); read(arg, *) N
  call get_command_argument(2, arg); read(arg, *) beta
  call get_command_argument(3, arg); read(arg, *) op
  call get_command_argument(4, arg); read(arg, *) c
  call get_command_argument(5, arg); read(arg, *) iter_max
  call get_command_argument(6, arg); read(arg, *) tol
  call get_command_argument(7, x0path)
  call get_command_argument(8, bpath)
  call get_command_argument(9, xpath)

  allocate(hx(N + 3), hy(N + 3))
  call moons_check(moons_nodes_uniform(0.0_c_double, 1.0_c_double, int(N, c_int), hx), 'nodes')
  if (beta .gt. 1.0_c_double) then
    call moons_check(moons_nodes_roberts_both(0.0_c_double, 1.0_c_double, int(N, c_int), beta, hy), 'nodes')
  else
    hy = hx
  endif
  call init(m, hx, hy, hy)

  ! face d: N + 1 + 2 points along d, N + 2 otherwise
  allocate(x1(N + 3, N + 2, N + 2), x2(N + 2, N + 3, N + 2), x3(N + 2, N + 2, N + 3))
  allocate(b1(N + 3, N + 2, N + 2), b2(N + 2, N + 3, N + 2), b3(N + 2, N + 2, N + 3))
  open(newunit=u, file=trim(x0path), access='stream', form='unformatted', status='old')
  read(u) x1, x2, x3
  close(u)
  open(newunit=u, file=trim(bpath), access='stream', form='unformatted', status='old')
  read(u) b1, b2, b3
  close(u)

  vals = 0.0_c_double
  if (op .eq. 1) then
    types = MOONS_BC_DIRICHLET
    vals(4, 1) = 1.0_c_double
    call init(PCG, m, MOONS_OP_MOM_DIFFUSION, types, vals, iter_max, tol, 1.0d-13, 1)
  else
    types = MOONS_BC_DIRICHLET
    do d = 1, 3
      types(2 * d - 1, d) = MOONS_BC_NEUMANN
      types(2 * d, d) = MOONS_BC_NEUMANN
    enddo
    call get_command_argument(10, spath)
    ! edge-x: N + 2 along x, N + 3 along y, z; edge-y, edge-z likewise
    allocate(s1(N + 2, N + 3, N + 3), s2(N + 3, N + 2, N + 3), s3(N + 3, N + 3, N + 2))
    open(newunit=u, file=trim(spath), access='stream', form='unformatted', status='old')
    read(u) s1, s2, s3
    close(u)
    ! the reference's init(PCG_B, ind_diffusion, ..., ISP, MFP, Bstar, B,
    ! sigmaInv_edge, ...) order (induction_extend.f90:206-207), with B's BCs
    ! on device fields and sigmaInv on the edges
    do d = 1, 3
      call init_Face(xbc(d), m, d)
      call set_BCs(xbc(d), types(:, d), vals(:, d), .false.)
      call init_Edge(sig(d), m, d)
    enddo
    call moons_check(moons_field_upload(sig(1)%h, s1, int(size(s1), c_size_t)), 'upload sigmaInv')
    call moons_check(moons_field_upload(sig(2)%h, s2, int(size(s2), c_size_t)), 'upload sigmaInv')
    call moons_check(moons_field_upload(sig(3)%h, s3, int(size(s3), c_size_t)), 'upload sigmaInv')
    call init(ISP, iter_max, tol, 1.0d-13, 1, .false., .false.)
    MFP%coeff_implicit_time_split = c
    call init(PCG, MOONS_OP_IND_DIFFUSION, m, ISP, MFP, xbc, xbc, sig, 'out/', 'B', .false., .false.)
  endif
  if (op .eq. 1) call update_MFP(PCG, c, .true.)
  call solve(PCG, x1, x2, x3, b1, b2, b3, .true.)

  open(newunit=u, file=trim(xpath), access='stream', form='unformatted', status='replace')
  write(u) x1, x2, x3
  close(u)
  call delete(PCG)
  do d = 1, 3
    call delete(xbc(d))
    call delete(sig(d))
  enddo
  call delete(m)
end program
