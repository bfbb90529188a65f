! moons_amd_PCG.f90 -- Fortran host surface of the GPU pressure-Poisson path.
!
! Mirrors the reference's call surface so a MOONS host swaps in the GPU path
! at the same call sites:
!   init(PCG, ...)          init_PCG_SF   (code/handwritten/PCG/PCG_solver_extend.f90:51-107)
!   solve(PCG, x, b, ...)   solve_PCG_SF  (code/handwritten/PCG/PCG_solver_extend.f90:182-193)
!   clean_div(PCG, ...)     clean_div_PCG (code/handwritten/solvers/clean_divergence.f90:37-59)
!   init(FFT, m, dir), solve(FFT, u, f), clean_div(FFT, ...)
!                           init_FFT / solve (code/handwritten/FFT_solver/FFT_solver_extend.f90:43-137)
!                           and clean_div_FFT (clean_divergence.f90:61-83)
!   init(nrg, ...), set_joule(nrg, ...), solve(nrg, T, U, ...)
!                           the energy equation's step: add_all_energy_sources then
!                           solve_energy (code/handwritten/solvers/energy/add_all_energy_sources.f90:21-80,
!                           code/handwritten/governing_equations/energy_extend.f90:236-262)
!   export(f, path), import(f, path)
!                           grid_field restart primitives of a device field
!                           (code/generated_code/grid_field/grid_field.f90:108-156;
!                           a file path in place of an open unit)
!   export_raw(m, f, dir, name, pad), import_raw(...)
!                           Tecplot raw files (code/handwritten/IO/export_raw_processed.f90:44-60,
!                           code/handwritten/IO/import_raw.f90:23-31)
! Fields are the host's ghost-padded, column-major real(8) arrays (the layout
! of grid_field%f, code/generated_code/grid_field/grid_field.f90:39-43); they
! are copied to the device at solve entry and back at exit (the PCIe-inclusive
! boundary).  Device-resident callers use the C ABI handles directly.
module moons_amd_PCG_mod
  use iso_c_binding
  use moons_amd_bindings
  implicit none
  private

  public :: mesh_gpu, SF_gpu, PCG_solver_SF_gpu, PCG_solver_VF_gpu, FFT_solver_SF_gpu, energy_gpu
  public :: exit_criteria_gpu, iter_solver_params_gpu, matrix_free_params_gpu
  public :: init, solve, delete, clean_div, update_MFP, set_joule
  public :: init_CC, init_Face, init_Edge, set_BCs
  public :: export, import, export_raw, import_raw

  type mesh_gpu
    type(c_ptr) :: h = c_null_ptr
    integer :: sn(3) = 0
  end type

  type SF_gpu
    type(c_ptr) :: h = c_null_ptr
    integer :: s(3) = 0
    integer(c_int) :: dl = 0
  end type

  ! The reference's solver parameters with its field names:
  ! iter_solver_params (code/generated_code/var_set/iter_solver_params.f90:44-55)
  ! with its exit_criteria (code/generated_code/var_set/exit_criteria.f90:44-50),
  ! and matrix_free_params (code/generated_code/PCG/matrix_free_params.f90:39-47).
  type exit_criteria_gpu
    integer :: iter_max = 0
    real(c_double) :: tol_abs = 0.0d0, tol_rel = 0.0d0
  end type

  type iter_solver_params_gpu
    type(exit_criteria_gpu) :: EC
    logical :: exit_loop(3) = .false.
    integer :: iter_total = 0, iter_per_call = 0, n_skip_check_res = 0
    logical :: export_convergence = .false., export_heavy = .false.
  end type

  type matrix_free_params_gpu
    logical :: suppress_warning = .false.
    real(c_double) :: alpha = 0.0d0, beta = 0.0d0, coeff_natural = 0.0d0, coeff_explicit = 0.0d0
    real(c_double) :: coeff_implicit = 0.0d0, coeff_implicit_time_split = 0.0d0
  end type

  type PCG_solver_SF_gpu
    type(c_ptr) :: h = c_null_ptr
    integer :: iter_max = 1000, n_skip_check_res = 1
    real(c_double) :: tol_rel = 1.0d-6, tol_abs = 1.0d-13
    type(moons_pcg_info) :: info
    type(SF_gpu) :: x, b ! device mirrors of the solve's x and b
  end type

  ! init_PCG_VF / solve_PCG_VF / update_MFP_PCG_VF over three face arrays
  ! (code/handwritten/PCG/PCG_solver_extend.f90:117-180,195-206,281-291)
  type PCG_solver_VF_gpu
    type(c_ptr) :: h = c_null_ptr
    type(moons_pcg_info) :: info
    type(SF_gpu) :: x(3), b(3), k(3)
  end type

  ! init_FFT / solve / clean_div_FFT: the cosine-transform direct solver of
  ! the planes normal to dir (its u carries the BCs)
  type FFT_solver_SF_gpu
    type(c_ptr) :: h = c_null_ptr
    type(SF_gpu) :: u, f
  end type

  ! The energy step over host arrays: T (cell-centred, with its BCs) and the
  ! momentum's face velocity U are uploaded at each solve, T downloaded after;
  ! F, Fnm1, L and Tnm1 stay on the device between steps.
  type energy_gpu
    type(c_ptr) :: h = c_null_ptr
    type(moons_nrg_info) :: info
    type(SF_gpu) :: T, U(3), k(3), J(3), sig
  end type

  interface init;   module procedure init_mesh_gpu; module procedure init_PCG_SF_gpu
                    module procedure init_PCG_VF_gpu; module procedure init_FFT_gpu
                    module procedure init_energy_gpu; module procedure init_ISP_gpu
                    module procedure init_PCG_SF_ref; module procedure init_PCG_VF_ref; end interface
  interface solve;  module procedure solve_PCG_SF_gpu; module procedure solve_PCG_VF_gpu
                    module procedure solve_FFT_gpu; module procedure solve_energy_gpu
                    module procedure solve_PCG_SF_ref; end interface
  interface set_joule; module procedure set_joule_energy_gpu; end interface
  interface clean_div; module procedure clean_div_PCG_gpu; module procedure clean_div_FFT_gpu; end interface
  interface update_MFP; module procedure update_MFP_PCG_VF_gpu; end interface
  interface export; module procedure export_SF_gpu; end interface
  interface import; module procedure import_SF_gpu; end interface
  interface export_raw; module procedure export_raw_SF_gpu; end interface
  interface import_raw; module procedure import_raw_SF_gpu; end interface
  interface delete; module procedure delete_mesh_gpu; module procedure delete_SF_gpu;
                    module procedure delete_PCG_SF_gpu; module procedure delete_PCG_VF_gpu
                    module procedure delete_FFT_gpu; module procedure delete_energy_gpu; end interface

contains

  subroutine init_mesh_gpu(m, hn_x, hn_y, hn_z)
    type(mesh_gpu), intent(inout) :: m
    real(c_double), intent(in) :: hn_x(:), hn_y(:), hn_z(:)
    call moons_check(moons_init(0_c_int), 'moons_init')
    call moons_check(moons_mesh_create(hn_x, size(hn_x), hn_y, size(hn_y), hn_z, size(hn_z), m%h), 'mesh')
    m%sn = [size(hn_x), size(hn_y), size(hn_z)]
  end subroutine

  subroutine init_field(f, m, dl)
    type(SF_gpu), intent(inout) :: f
    type(mesh_gpu), intent(in) :: m
    integer(c_int), intent(in) :: dl
    integer :: d
    call moons_check(moons_field_create(m%h, dl, f%h), 'field')
    f%dl = dl
    do d = 1, 3
      if (btest(dl, d - 1)) then
        f%s(d) = m%sn(d)
      else
        f%s(d) = m%sn(d) - 1
      endif
    enddo
  end subroutine

  subroutine init_CC(f, m)
    type(SF_gpu), intent(inout) :: f
    type(mesh_gpu), intent(in) :: m
    call init_field(f, m, MOONS_DL_CC)
  end subroutine

  subroutine init_Face(f, m, dir)
    type(SF_gpu), intent(inout) :: f
    type(mesh_gpu), intent(in) :: m
    integer, intent(in) :: dir
    call init_field(f, m, int(ishft(1, dir - 1), c_int))
  end subroutine

  ! edge field along dir (1, 2, 3): node-centred off dir
  subroutine init_Edge(f, m, dir)
    type(SF_gpu), intent(inout) :: f
    type(mesh_gpu), intent(in) :: m
    integer, intent(in) :: dir
    integer(c_int) :: edl(3)
    edl = [MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z]
    call init_field(f, m, edl(dir))
  end subroutine

  ! BC type per face (MOONS_BC_*), uniform value per face, all_Neumann flag.
  subroutine set_BCs(f, types, vals, all_Neumann)
    type(SF_gpu), intent(inout) :: f
    integer(c_int), intent(in) :: types(6)
    real(c_double), intent(in) :: vals(6)
    logical, intent(in) :: all_Neumann
    integer(c_int) :: an
    an = 0
    if (all_Neumann) an = 1
    call moons_check(moons_field_set_bcs(f%h, types, vals, an), 'set_BCs')
  end subroutine

  ! init_ISP (code/handwritten/var_set/iter_solver_params_extend.f90:44-59)
  subroutine init_ISP_gpu(ISP, iter_max, tol_rel, tol_abs, n_skip_check_res, export_convergence, export_heavy)
    type(iter_solver_params_gpu), intent(inout) :: ISP
    integer, intent(in) :: iter_max, n_skip_check_res
    real(c_double), intent(in) :: tol_rel, tol_abs
    logical, intent(in) :: export_convergence, export_heavy
    ISP%iter_total = 0
    ISP%iter_per_call = 0
    ISP%EC%iter_max = iter_max
    ISP%EC%tol_rel = tol_rel
    ISP%EC%tol_abs = tol_abs
    ISP%export_convergence = export_convergence
    ISP%export_heavy = export_heavy
    ISP%n_skip_check_res = n_skip_check_res
    ISP%exit_loop = .false.
  end subroutine

  subroutine refuse_debug_options(testSymmetry, exportOperator, where)
    logical, intent(in) :: testSymmetry, exportOperator
    character(len=*), intent(in) :: where
    if (testSymmetry .or. exportOperator) then
      write(*, '(a)') 'Error: '//where//': testSymmetry and exportOperator have no device path'
      error stop 'moons_amd: unsupported solver option'
    endif
  end subroutine

  ! init_PCG_SF in the reference's argument order (code/handwritten/PCG/PCG_solver_extend.f90:51-107).
  ! operator (with operator_explicit) and prec are the codes MOONS_OP_* and
  ! MOONS_PREC_*, since a Fortran procedure cannot run on the device; x_BC
  ! carries the BCs (a device field, set with set_BCs); k: the three face
  ! fields of nrg_diffusion's coefficient (their handles may be null for the
  ! Laplacian, whose k the reference does not read); MFP%coeff_implicit_time_split
  ! is the operator's coefficient (nrg_diffusion); var_dir and var_name name the
  ! solve in its messages.  testSymmetry and exportOperator must be .false.
  subroutine init_PCG_SF_ref(PCG, operator, prec, m, ISP, MFP, x, x_BC, k, var_dir, var_name, testSymmetry, &
                             exportOperator)
    type(PCG_solver_SF_gpu), intent(inout) :: PCG
    integer(c_int), intent(in) :: operator, prec
    type(mesh_gpu), intent(in) :: m
    type(iter_solver_params_gpu), intent(in) :: ISP
    type(matrix_free_params_gpu), intent(in) :: MFP
    type(SF_gpu), intent(in) :: x, x_BC
    type(SF_gpu), intent(in) :: k(3)
    character(len=*), intent(in) :: var_dir, var_name
    logical, intent(in) :: testSymmetry, exportOperator
    type(c_ptr) :: hk(3)
    integer :: d
    call refuse_debug_options(testSymmetry, exportOperator, 'init_PCG_SF '//var_dir//var_name)
    if (x%dl .ne. MOONS_DL_CC) error stop 'Error: init_PCG_SF_gpu: x must be cell-centred'
    call init_CC(PCG%x, m)
    call init_CC(PCG%b, m)
    call moons_check(moons_field_copy_bcs(PCG%x%h, x_BC%h), 'init_PCG_SF '//var_name)
    call moons_check(moons_pcg_sf_create(m%h, operator, prec, PCG%x%h, PCG%h), 'init_PCG_SF '//var_name)
    PCG%iter_max = ISP%EC%iter_max
    PCG%tol_rel = ISP%EC%tol_rel
    PCG%tol_abs = ISP%EC%tol_abs
    PCG%n_skip_check_res = max(ISP%n_skip_check_res, 1)
    call moons_check(moons_pcg_set_isp(PCG%h, int(PCG%iter_max, c_int), PCG%tol_rel, PCG%tol_abs, &
                                       int(PCG%n_skip_check_res, c_int)), 'ISP '//var_name)
    if (operator .eq. MOONS_OP_NRG_DIFFUSION) then
      do d = 1, 3
        hk(d) = k(d)%h
      enddo
      call moons_check(moons_pcg_sf_set_k(PCG%h, hk), 'init_PCG_SF k '//var_name)
      call moons_check(moons_pcg_sf_update_mfp(PCG%h, MFP%coeff_implicit_time_split, 1_c_int), 'MFP '//var_name)
    endif
  end subroutine

  ! solve_PCG_SF in the reference's argument order (PCG_solver_extend.f90:182-193):
  ! x, b host arrays (ghosts included); m names the mesh as there.
  subroutine solve_PCG_SF_ref(PCG, x, b, m, compute_norms)
    type(PCG_solver_SF_gpu), intent(inout) :: PCG
    real(c_double), intent(inout), contiguous, target :: x(:, :, :)
    real(c_double), intent(in), contiguous, target :: b(:, :, :)
    type(mesh_gpu), intent(in) :: m
    logical, intent(in) :: compute_norms
    if (.not. c_associated(m%h)) error stop 'Error: solve_PCG_SF_gpu: mesh not initialised'
    call solve_PCG_SF_gpu(PCG, x, b, compute_norms)
  end subroutine

  ! init_PCG_SF: operator Lap_uniform_SF, prec prec_Lap_SF (prec = MOONS_PREC_LAP_SF)
  ! or the FP32 multigrid (MOONS_PREC_MG_FP32); x carries the BCs.
  subroutine init_PCG_SF_gpu(PCG, m, x_types, x_vals, all_Neumann, iter_max, tol_rel, tol_abs, &
                             n_skip_check_res, prec)
    type(PCG_solver_SF_gpu), intent(inout) :: PCG
    type(mesh_gpu), intent(in) :: m
    integer(c_int), intent(in) :: x_types(6)
    real(c_double), intent(in) :: x_vals(6)
    logical, intent(in) :: all_Neumann
    integer, intent(in) :: iter_max, n_skip_check_res
    real(c_double), intent(in) :: tol_rel, tol_abs
    integer(c_int), intent(in) :: prec
    call init_CC(PCG%x, m)
    call init_CC(PCG%b, m)
    call set_BCs(PCG%x, x_types, x_vals, all_Neumann)
    call moons_check(moons_pcg_sf_create(m%h, MOONS_OP_LAP_UNIFORM_SF, prec, PCG%x%h, PCG%h), 'init_PCG_SF')
    PCG%iter_max = iter_max
    PCG%tol_rel = tol_rel
    PCG%tol_abs = tol_abs
    PCG%n_skip_check_res = n_skip_check_res
    call moons_check(moons_pcg_set_isp(PCG%h, int(iter_max, c_int), tol_rel, tol_abs, &
                                       int(n_skip_check_res, c_int)), 'ISP')
  end subroutine

  ! solve_PCG_SF on host arrays: x (in/out, ghosts included) and b.
  subroutine solve_PCG_SF_gpu(PCG, x, b, compute_norms)
    type(PCG_solver_SF_gpu), intent(inout) :: PCG
    real(c_double), intent(inout), contiguous, target :: x(:, :, :)
    real(c_double), intent(in), contiguous, target :: b(:, :, :)
    logical, intent(in) :: compute_norms
    integer(c_size_t) :: n
    n = int(size(x), c_size_t)
    if (any(shape(x) .ne. PCG%x%s) .or. any(shape(b) .ne. PCG%x%s)) then
      error stop 'Error: x and b must have the solver field shape in solve_PCG_SF_gpu'
    endif
    call moons_check(moons_field_upload(PCG%x%h, x, n), 'upload x')
    call moons_check(moons_field_upload(PCG%b%h, b, n), 'upload b')
    call moons_check(moons_pcg_sf_solve(PCG%h, PCG%x%h, PCG%b%h, PCG%info), 'solve_PCG_SF')
    call moons_check(moons_field_download(PCG%x%h, x, n), 'download x')
    if (compute_norms) then
      write(*, '(a,i8,a,es12.4)') ' PCG iterations = ', PCG%info%iterations, &
        '   rel. residual = ', PCG%info%res_norm / PCG%info%res_norm0
    endif
  end subroutine

  ! clean_div_PCG on host face arrays: X = Xstar - grad(phi)/scale with
  ! lap(phi) = scale*div(Xstar); X carries its BCs (types, vals).
  subroutine clean_div_PCG_gpu(PCG, m, X1, X2, X3, Xs1, Xs2, Xs3, X_types, X_vals, phi, scale)
    type(PCG_solver_SF_gpu), intent(inout) :: PCG
    type(mesh_gpu), intent(in) :: m
    real(c_double), intent(inout), contiguous :: X1(:, :, :), X2(:, :, :), X3(:, :, :)
    real(c_double), intent(in), contiguous :: Xs1(:, :, :), Xs2(:, :, :), Xs3(:, :, :)
    integer(c_int), intent(in) :: X_types(6)
    real(c_double), intent(in) :: X_vals(6)
    real(c_double), intent(inout), contiguous :: phi(:, :, :)
    real(c_double), intent(in) :: scale
    type(SF_gpu) :: X(3), Xs(3), T(3), tcc
    type(c_ptr) :: hX(3), hXs(3), hT(3)
    integer :: d
    do d = 1, 3
      call init_Face(X(d), m, d)
      call init_Face(Xs(d), m, d)
      call init_Face(T(d), m, d)
      call set_BCs(X(d), X_types, X_vals, .false.)
      hX(d) = X(d)%h; hXs(d) = Xs(d)%h; hT(d) = T(d)%h
    enddo
    call init_CC(tcc, m)
    call moons_check(moons_field_upload(Xs(1)%h, Xs1, int(size(Xs1), c_size_t)), 'upload')
    call moons_check(moons_field_upload(Xs(2)%h, Xs2, int(size(Xs2), c_size_t)), 'upload')
    call moons_check(moons_field_upload(Xs(3)%h, Xs3, int(size(Xs3), c_size_t)), 'upload')
    call moons_check(moons_field_upload(PCG%x%h, phi, int(size(phi), c_size_t)), 'upload')
    call moons_check(moons_clean_div(PCG%h, hX, hXs, PCG%x%h, scale, hT, tcc%h, PCG%info), 'clean_div')
    call moons_check(moons_field_download(X(1)%h, X1, int(size(X1), c_size_t)), 'download')
    call moons_check(moons_field_download(X(2)%h, X2, int(size(X2), c_size_t)), 'download')
    call moons_check(moons_field_download(X(3)%h, X3, int(size(X3), c_size_t)), 'download')
    call moons_check(moons_field_download(PCG%x%h, phi, int(size(phi), c_size_t)), 'download')
    do d = 1, 3
      call delete(X(d)); call delete(Xs(d)); call delete(T(d))
    enddo
    call delete(tcc)
  end subroutine

  ! init_FFT: direction dir (1, 2, 3); u's BCs (types, vals)
  subroutine init_FFT_gpu(FFT, m, dir, u_types, u_vals)
    type(FFT_solver_SF_gpu), intent(inout) :: FFT
    type(mesh_gpu), intent(in) :: m
    integer, intent(in) :: dir
    integer(c_int), intent(in) :: u_types(6)
    real(c_double), intent(in) :: u_vals(6)
    call init_CC(FFT%u, m)
    call init_CC(FFT%f, m)
    call set_BCs(FFT%u, u_types, u_vals, .false.)
    call moons_check(moons_fft_create(m%h, int(dir, c_int), FFT%h), 'init_FFT')
  end subroutine

  ! solve(FFT, u, f) on host arrays (ghosts included)
  subroutine solve_FFT_gpu(FFT, u, f)
    type(FFT_solver_SF_gpu), intent(inout) :: FFT
    real(c_double), intent(inout), contiguous :: u(:, :, :)
    real(c_double), intent(in), contiguous :: f(:, :, :)
    integer(c_size_t) :: n
    n = int(size(u), c_size_t)
    if (any(shape(u) .ne. FFT%u%s) .or. any(shape(f) .ne. FFT%u%s)) then
      error stop 'Error: u and f must have the cell-centred shape in solve_FFT_gpu'
    endif
    call moons_check(moons_field_upload(FFT%f%h, f, n), 'upload f')
    call moons_check(moons_fft_solve(FFT%h, FFT%u%h, FFT%f%h), 'solve_FFT')
    call moons_check(moons_field_download(FFT%u%h, u, n), 'download u')
  end subroutine

  ! clean_div_FFT on host face arrays: X = Xstar - grad(phi)/scale with
  ! phi from the cosine-transform solve of scale*div(Xstar)
  subroutine clean_div_FFT_gpu(FFT, m, X1, X2, X3, Xs1, Xs2, Xs3, X_types, X_vals, phi, scale)
    type(FFT_solver_SF_gpu), intent(inout) :: FFT
    type(mesh_gpu), intent(in) :: m
    real(c_double), intent(inout), contiguous :: X1(:, :, :), X2(:, :, :), X3(:, :, :)
    real(c_double), intent(in), contiguous :: Xs1(:, :, :), Xs2(:, :, :), Xs3(:, :, :)
    integer(c_int), intent(in) :: X_types(6)
    real(c_double), intent(in) :: X_vals(6)
    real(c_double), intent(inout), contiguous :: phi(:, :, :)
    real(c_double), intent(in) :: scale
    type(SF_gpu) :: X(3), Xs(3), T(3)
    type(c_ptr) :: hX(3), hXs(3), hT(3)
    integer :: d
    do d = 1, 3
      call init_Face(X(d), m, d)
      call init_Face(Xs(d), m, d)
      call init_Face(T(d), m, d)
      call set_BCs(X(d), X_types, X_vals, .false.)
      hX(d) = X(d)%h; hXs(d) = Xs(d)%h; hT(d) = T(d)%h
    enddo
    call moons_check(moons_field_upload(Xs(1)%h, Xs1, int(size(Xs1), c_size_t)), 'upload')
    call moons_check(moons_field_upload(Xs(2)%h, Xs2, int(size(Xs2), c_size_t)), 'upload')
    call moons_check(moons_field_upload(Xs(3)%h, Xs3, int(size(Xs3), c_size_t)), 'upload')
    call moons_check(moons_clean_div_fft(FFT%h, hX, hXs, FFT%u%h, scale, hT, FFT%f%h), 'clean_div_FFT')
    call moons_check(moons_field_download(X(1)%h, X1, int(size(X1), c_size_t)), 'download')
    call moons_check(moons_field_download(X(2)%h, X2, int(size(X2), c_size_t)), 'download')
    call moons_check(moons_field_download(X(3)%h, X3, int(size(X3), c_size_t)), 'download')
    call moons_check(moons_field_download(FFT%u%h, phi, int(size(phi), c_size_t)), 'download')
    do d = 1, 3
      call delete(X(d)); call delete(Xs(d)); call delete(T(d))
    enddo
  end subroutine

  subroutine delete_FFT_gpu(FFT)
    type(FFT_solver_SF_gpu), intent(inout) :: FFT
    if (c_associated(FFT%h)) call moons_check(moons_fft_destroy(FFT%h), 'fft destroy')
    FFT%h = c_null_ptr
    call delete(FFT%u)
    call delete(FFT%f)
  end subroutine

  ! init_PCG_VF: op = MOONS_OP_MOM_DIFFUSION (prec_mom_VF) or
  ! MOONS_OP_IND_DIFFUSION (prec_ind_VF, with sigmaInv on the x, y, z edges).
  ! X_types(6, 3), X_vals(6, 3): the BCs of the three face components.
  subroutine init_PCG_VF_gpu(PCG, m, op, X_types, X_vals, iter_max, tol_rel, tol_abs, n_skip_check_res, &
                             sig_x, sig_y, sig_z)
    type(PCG_solver_VF_gpu), intent(inout) :: PCG
    type(mesh_gpu), intent(in) :: m
    integer(c_int), intent(in) :: op
    integer(c_int), intent(in) :: X_types(6, 3)
    real(c_double), intent(in) :: X_vals(6, 3)
    integer, intent(in) :: iter_max, n_skip_check_res
    real(c_double), intent(in) :: tol_rel, tol_abs
    real(c_double), intent(in), optional, contiguous :: sig_x(:, :, :), sig_y(:, :, :), sig_z(:, :, :)
    type(c_ptr) :: hx(3), hk(3)
    integer(c_int) :: edl(3)
    integer :: d
    edl = [MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z]
    do d = 1, 3
      call init_Face(PCG%x(d), m, d)
      call init_Face(PCG%b(d), m, d)
      call set_BCs(PCG%x(d), X_types(:, d), X_vals(:, d), .false.)
      hx(d) = PCG%x(d)%h
      hk(d) = c_null_ptr
    enddo
    if (op .eq. MOONS_OP_IND_DIFFUSION) then
      if (.not. (present(sig_x) .and. present(sig_y) .and. present(sig_z))) &
        error stop 'Error: ind_diffusion needs sigmaInv on the edges in init_PCG_VF_gpu'
      do d = 1, 3
        call init_field(PCG%k(d), m, edl(d))
        hk(d) = PCG%k(d)%h
      enddo
      call moons_check(moons_field_upload(PCG%k(1)%h, sig_x, int(size(sig_x), c_size_t)), 'upload sigmaInv')
      call moons_check(moons_field_upload(PCG%k(2)%h, sig_y, int(size(sig_y), c_size_t)), 'upload sigmaInv')
      call moons_check(moons_field_upload(PCG%k(3)%h, sig_z, int(size(sig_z), c_size_t)), 'upload sigmaInv')
    endif
    call moons_check(moons_pcg_vf_create(m%h, op, hx, hk, PCG%h), 'init_PCG_VF')
    call moons_check(moons_pcg_vf_set_isp(PCG%h, int(iter_max, c_int), tol_rel, tol_abs, &
                                          int(n_skip_check_res, c_int)), 'ISP')
  end subroutine

  ! init_PCG_VF in the reference's argument order (PCG_solver_extend.f90:117-180):
  ! operator MOONS_OP_MOM_DIFFUSION (prec_mom_VF) or MOONS_OP_IND_DIFFUSION
  ! (prec_ind_VF); x_BC(3) carries the BCs of the three face components; k(3):
  ! sigmaInv on the x, y, z edges (ind_diffusion; null handles for momentum).
  subroutine init_PCG_VF_ref(PCG, operator, m, ISP, MFP, x, x_BC, k, var_dir, var_name, testSymmetry, exportOperator)
    type(PCG_solver_VF_gpu), intent(inout) :: PCG
    integer(c_int), intent(in) :: operator
    type(mesh_gpu), intent(in) :: m
    type(iter_solver_params_gpu), intent(in) :: ISP
    type(matrix_free_params_gpu), intent(in) :: MFP
    type(SF_gpu), intent(in) :: x(3), x_BC(3), k(3)
    character(len=*), intent(in) :: var_dir, var_name
    logical, intent(in) :: testSymmetry, exportOperator
    type(c_ptr) :: hx(3), hk(3)
    integer :: d
    call refuse_debug_options(testSymmetry, exportOperator, 'init_PCG_VF '//var_dir//var_name)
    do d = 1, 3
      call init_Face(PCG%x(d), m, d)
      call init_Face(PCG%b(d), m, d)
      call moons_check(moons_field_copy_bcs(PCG%x(d)%h, x_BC(d)%h), 'init_PCG_VF '//var_name)
      hx(d) = PCG%x(d)%h
      hk(d) = c_null_ptr
      if (operator .eq. MOONS_OP_IND_DIFFUSION) hk(d) = k(d)%h
    enddo
    if (x(1)%dl .ne. MOONS_DL_FACE_X) error stop 'Error: init_PCG_VF_gpu: x must be the three face components'
    call moons_check(moons_pcg_vf_create(m%h, operator, hx, hk, PCG%h), 'init_PCG_VF '//var_name)
    call moons_check(moons_pcg_vf_set_isp(PCG%h, int(ISP%EC%iter_max, c_int), ISP%EC%tol_rel, ISP%EC%tol_abs, &
                                          int(max(ISP%n_skip_check_res, 1), c_int)), 'ISP '//var_name)
    if (MFP%coeff_implicit_time_split .ne. 0.0d0) &
      call moons_check(moons_pcg_vf_update_mfp(PCG%h, MFP%coeff_implicit_time_split, 1_c_int), 'MFP '//var_name)
  end subroutine

  subroutine update_MFP_PCG_VF_gpu(PCG, coeff_implicit_time_split, update_preconditioner)
    type(PCG_solver_VF_gpu), intent(inout) :: PCG
    real(c_double), intent(in) :: coeff_implicit_time_split
    logical, intent(in) :: update_preconditioner
    integer(c_int) :: up
    up = 0
    if (update_preconditioner) up = 1
    call moons_check(moons_pcg_vf_update_mfp(PCG%h, coeff_implicit_time_split, up), 'update_MFP_PCG_VF')
  end subroutine

  ! solve_PCG_VF on host face arrays (ghosts included), x in/out
  subroutine solve_PCG_VF_gpu(PCG, x1, x2, x3, b1, b2, b3, compute_norms)
    type(PCG_solver_VF_gpu), intent(inout) :: PCG
    real(c_double), intent(inout), contiguous :: x1(:, :, :), x2(:, :, :), x3(:, :, :)
    real(c_double), intent(in), contiguous :: b1(:, :, :), b2(:, :, :), b3(:, :, :)
    logical, intent(in) :: compute_norms
    type(c_ptr) :: hx(3), hb(3)
    integer :: d
    call moons_check(moons_field_upload(PCG%x(1)%h, x1, int(size(x1), c_size_t)), 'upload x')
    call moons_check(moons_field_upload(PCG%x(2)%h, x2, int(size(x2), c_size_t)), 'upload x')
    call moons_check(moons_field_upload(PCG%x(3)%h, x3, int(size(x3), c_size_t)), 'upload x')
    call moons_check(moons_field_upload(PCG%b(1)%h, b1, int(size(b1), c_size_t)), 'upload b')
    call moons_check(moons_field_upload(PCG%b(2)%h, b2, int(size(b2), c_size_t)), 'upload b')
    call moons_check(moons_field_upload(PCG%b(3)%h, b3, int(size(b3), c_size_t)), 'upload b')
    do d = 1, 3
      hx(d) = PCG%x(d)%h
      hb(d) = PCG%b(d)%h
    enddo
    call moons_check(moons_pcg_vf_solve(PCG%h, hx, hb, PCG%info), 'solve_PCG_VF')
    call moons_check(moons_field_download(PCG%x(1)%h, x1, int(size(x1), c_size_t)), 'download x')
    call moons_check(moons_field_download(PCG%x(2)%h, x2, int(size(x2), c_size_t)), 'download x')
    call moons_check(moons_field_download(PCG%x(3)%h, x3, int(size(x3), c_size_t)), 'download x')
    if (compute_norms) then
      write(*, '(a,i8,a,es12.4)') ' PCG_VF iterations = ', PCG%info%iterations, &
        '   rel. residual = ', PCG%info%res_norm / PCG%info%res_norm0
    endif
  end subroutine

  subroutine delete_PCG_VF_gpu(PCG)
    type(PCG_solver_VF_gpu), intent(inout) :: PCG
    integer :: d
    if (c_associated(PCG%h)) call moons_check(moons_pcg_vf_destroy(PCG%h), 'delete PCG_VF')
    PCG%h = c_null_ptr
    do d = 1, 3
      call delete(PCG%x(d)); call delete(PCG%b(d))
      if (c_associated(PCG%k(d)%h)) call delete(PCG%k(d))
    enddo
  end subroutine

  ! init_energy: T's BCs (types, vals), the step's parameters (moons_nrg_params:
  ! method, dt, coeff_implicit, the sources' flags and scales, the T solver's
  ! ISP, rk_stages) and, for methods 3-5, the face arrays k1, k2, k3 bound to
  ! the implicit operator (absent: zeros, as init_energy binds,
  ! energy_extend.f90:103,129-130).
  subroutine init_energy_gpu(nrg, m, T_types, T_vals, params, k1, k2, k3)
    type(energy_gpu), intent(inout) :: nrg
    type(mesh_gpu), intent(in) :: m
    integer(c_int), intent(in) :: T_types(6)
    real(c_double), intent(in) :: T_vals(6)
    type(moons_nrg_params), intent(in) :: params
    real(c_double), intent(in), optional, contiguous :: k1(:, :, :), k2(:, :, :), k3(:, :, :)
    type(c_ptr) :: hU(3), hk(3)
    integer :: d
    call init_CC(nrg%T, m)
    call set_BCs(nrg%T, T_types, T_vals, .false.)
    call moons_check(moons_field_fill(nrg%T%h, 0.0_c_double), 'fill')
    do d = 1, 3
      call init_Face(nrg%U(d), m, d)
      call init_Face(nrg%k(d), m, d)
      call moons_check(moons_field_fill(nrg%U(d)%h, 0.0_c_double), 'fill')
      call moons_check(moons_field_fill(nrg%k(d)%h, 0.0_c_double), 'fill')
      hU(d) = nrg%U(d)%h
      hk(d) = nrg%k(d)%h
    enddo
    if (present(k1) .and. present(k2) .and. present(k3)) then
      call moons_check(moons_field_upload(nrg%k(1)%h, k1, int(size(k1), c_size_t)), 'upload k')
      call moons_check(moons_field_upload(nrg%k(2)%h, k2, int(size(k2), c_size_t)), 'upload k')
      call moons_check(moons_field_upload(nrg%k(3)%h, k3, int(size(k3), c_size_t)), 'upload k')
    endif
    call moons_check(moons_nrg_create(m%h, params, nrg%T%h, hU, hk, nrg%h), 'init_energy')
  end subroutine

  ! The induction's J (x, y, z edge arrays) and sigmaInv_CC that Joule heating
  ! reads (ind%J, ind%sigmaInv_CC), copied to the device.
  subroutine set_joule_energy_gpu(nrg, m, J1, J2, J3, sigmaInv_CC)
    type(energy_gpu), intent(inout) :: nrg
    type(mesh_gpu), intent(in) :: m
    real(c_double), intent(in), contiguous :: J1(:, :, :), J2(:, :, :), J3(:, :, :), sigmaInv_CC(:, :, :)
    type(c_ptr) :: hJ(3)
    integer(c_int) :: edl(3)
    integer :: d
    edl = [MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z]
    if (.not. c_associated(nrg%J(1)%h)) then
      do d = 1, 3
        call init_field(nrg%J(d), m, edl(d))
      enddo
      call init_CC(nrg%sig, m)
    endif
    call moons_check(moons_field_upload(nrg%J(1)%h, J1, int(size(J1), c_size_t)), 'upload J')
    call moons_check(moons_field_upload(nrg%J(2)%h, J2, int(size(J2), c_size_t)), 'upload J')
    call moons_check(moons_field_upload(nrg%J(3)%h, J3, int(size(J3), c_size_t)), 'upload J')
    call moons_check(moons_field_upload(nrg%sig%h, sigmaInv_CC, int(size(sigmaInv_CC), c_size_t)), 'upload sigmaInv')
    do d = 1, 3
      hJ(d) = nrg%J(d)%h
    enddo
    call moons_check(moons_nrg_set_joule(nrg%h, hJ, nrg%sig%h), 'set_joule')
  end subroutine

  ! One energy step (the RK stages when rk_stages = 3) on host arrays:
  ! T in/out (ghosts included), U1, U2, U3 the momentum's face velocity.
  subroutine solve_energy_gpu(nrg, T, U1, U2, U3, compute_norms)
    type(energy_gpu), intent(inout) :: nrg
    real(c_double), intent(inout), contiguous :: T(:, :, :)
    real(c_double), intent(in), contiguous :: U1(:, :, :), U2(:, :, :), U3(:, :, :)
    logical, intent(in) :: compute_norms
    if (any(shape(T) .ne. nrg%T%s)) error stop 'Error: T must have the cell-centred shape in solve_energy_gpu'
    call moons_check(moons_field_upload(nrg%T%h, T, int(size(T), c_size_t)), 'upload T')
    call moons_check(moons_field_upload(nrg%U(1)%h, U1, int(size(U1), c_size_t)), 'upload U')
    call moons_check(moons_field_upload(nrg%U(2)%h, U2, int(size(U2), c_size_t)), 'upload U')
    call moons_check(moons_field_upload(nrg%U(3)%h, U3, int(size(U3), c_size_t)), 'upload U')
    call moons_check(moons_nrg_step(nrg%h, 1_c_int, nrg%info), 'solve_energy')
    call moons_check(moons_field_download(nrg%T%h, T, int(size(T), c_size_t)), 'download T')
    if (compute_norms) write(*, '(a,i8)') ' energy T-solve iterations = ', nrg%info%iterations
  end subroutine

  subroutine delete_energy_gpu(nrg)
    type(energy_gpu), intent(inout) :: nrg
    integer :: d
    if (c_associated(nrg%h)) call moons_check(moons_nrg_destroy(nrg%h), 'delete energy')
    nrg%h = c_null_ptr
    call delete(nrg%T)
    do d = 1, 3
      call delete(nrg%U(d)); call delete(nrg%k(d))
      if (c_associated(nrg%J(d)%h)) call delete(nrg%J(d))
    enddo
    if (c_associated(nrg%sig%h)) call delete(nrg%sig)
  end subroutine

  subroutine export_SF_gpu(f, path)
    type(SF_gpu), intent(in) :: f
    character(len=*), intent(in) :: path
    call moons_check(moons_field_export(f%h, trim(path)//c_null_char), 'export')
  end subroutine

  subroutine import_SF_gpu(f, path)
    type(SF_gpu), intent(inout) :: f
    character(len=*), intent(in) :: path
    call moons_check(moons_field_import(f%h, trim(path)//c_null_char), 'import')
  end subroutine

  ! export_raw(m, x, dir, name, pad): file dir//name//<c|n|f_x|...>//'.dat'
  subroutine export_raw_SF_gpu(m, f, dir, name, pad)
    type(mesh_gpu), intent(in) :: m
    type(SF_gpu), intent(in) :: f
    character(len=*), intent(in) :: dir, name
    integer, intent(in) :: pad
    if (.not. c_associated(m%h)) stop 'Error: export_raw on a mesh that was not initialised'
    call moons_check(moons_export_raw(f%h, dir//c_null_char, name//c_null_char, int(pad, c_int), c_null_ptr), &
                     'export_raw')
  end subroutine

  subroutine import_raw_SF_gpu(m, f, dir, name, pad)
    type(mesh_gpu), intent(in) :: m
    type(SF_gpu), intent(inout) :: f
    character(len=*), intent(in) :: dir, name
    integer, intent(in) :: pad
    if (.not. c_associated(m%h)) stop 'Error: import_raw on a mesh that was not initialised'
    call moons_check(moons_import_raw(f%h, dir//c_null_char, name//c_null_char, int(pad, c_int)), 'import_raw')
  end subroutine

  subroutine delete_mesh_gpu(m)
    type(mesh_gpu), intent(inout) :: m
    if (c_associated(m%h)) call moons_check(moons_mesh_destroy(m%h), 'mesh destroy')
    m%h = c_null_ptr
  end subroutine

  subroutine delete_SF_gpu(f)
    type(SF_gpu), intent(inout) :: f
    if (c_associated(f%h)) call moons_check(moons_field_destroy(f%h), 'field destroy')
    f%h = c_null_ptr
  end subroutine

  subroutine delete_PCG_SF_gpu(PCG)
    type(PCG_solver_SF_gpu), intent(inout) :: PCG
    if (c_associated(PCG%h)) call moons_check(moons_pcg_destroy(PCG%h), 'pcg destroy')
    PCG%h = c_null_ptr
    call delete(PCG%x)
    call delete(PCG%b)
  end subroutine

end module
