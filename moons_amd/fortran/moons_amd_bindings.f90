! moons_amd_bindings.f90 -- iso_c_binding interfaces of libmoons_amd
! (include/moons_amd.h), one per C entry point the Fortran host uses, and
! the status check that turns a non-zero status into `error stop` with the
! library's message, as the reference stops on errors
! (code/handwritten/ops/ops_del.f90:267-288).
module moons_amd_bindings
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MOONS_DL_CC = 0
  integer(c_int), parameter, public :: MOONS_DL_FACE_X = 1, MOONS_DL_FACE_Y = 2, MOONS_DL_FACE_Z = 4
  integer(c_int), parameter, public :: MOONS_BC_NONE = 0, MOONS_BC_DIRICHLET = 1, MOONS_BC_NEUMANN = 2
  integer(c_int), parameter, public :: MOONS_BC_PERIODIC = 3, MOONS_BC_SYMMETRIC = 4, MOONS_BC_ANTISYMMETRIC = 5
  integer(c_int), parameter, public :: MOONS_BC_ROBIN = 6
  integer(c_int), parameter, public :: MOONS_OP_LAP_UNIFORM_SF = 1, MOONS_OP_NRG_DIFFUSION = 3
  integer(c_int), parameter, public :: MOONS_PREC_LAP_SF = 1, MOONS_PREC_MG_FP32 = 2
  integer(c_int), parameter, public :: MOONS_OP_MOM_DIFFUSION = 1, MOONS_OP_IND_DIFFUSION = 2
  integer(c_int), parameter, public :: MOONS_DL_EDGE_X = 6, MOONS_DL_EDGE_Y = 5, MOONS_DL_EDGE_Z = 3
  ! grid_continue extension kinds of moons_nodes_ext
  integer(c_int), parameter, public :: MOONS_EXT_UNIFORM = 0, MOONS_EXT_ROBERTS_NEAR = 1, MOONS_EXT_ROBERTS_FAR = 2
  integer(c_int), parameter, public :: MOONS_EXT_ROBERTS_B = 3, MOONS_EXT_PREP_UNIFORM = 4
  integer(c_int), parameter, public :: MOONS_EXT_PREP_ROBERTS_L = 5, MOONS_EXT_PREP_ROBERTS_R = 6
  integer(c_int), parameter, public :: MOONS_EXT_PREP_ROBERTS_B = 7, MOONS_EXT_PREP_ROBERTS_C2F = 8
  integer(c_int), parameter, public :: MOONS_EXT_APP_UNIFORM = 9, MOONS_EXT_APP_ROBERTS_L = 10
  integer(c_int), parameter, public :: MOONS_EXT_APP_ROBERTS_R = 11, MOONS_EXT_APP_ROBERTS_B = 12
  integer(c_int), parameter, public :: MOONS_EXT_APP_ROBERTS_C2F = 13

  ! moons_mhd_params / moons_mhd_info of include/moons_amd.h
  type, bind(C), public :: moons_mhd_params
    real(c_double) :: Re, Ha, Rem, dt
    real(c_double) :: alpha_U
    integer(c_int) :: induction
    integer(c_int) :: prec_p
    real(c_double) :: B0(3)
    integer(c_int) :: iter_max(4)
    real(c_double) :: tol_rel(4)
    real(c_double) :: tol_abs(4)
    integer(c_int) :: n_skip(4)
    integer(c_int) :: mpg_dir
    real(c_double) :: mpg_scale
    integer(c_int) :: finite_Rem
    integer(c_int) :: prec_phi
  end type

  type, bind(C), public :: moons_mhd_info
    integer(c_int) :: iterations(4)
    real(c_double) :: ms
  end type

  ! moons_nrg_params / moons_nrg_info of include/moons_amd.h
  integer(c_int), parameter, public :: MOONS_NRG_F = 1, MOONS_NRG_FNM1 = 2, MOONS_NRG_TNM1 = 3, MOONS_NRG_L = 4

  type, bind(C), public :: moons_nrg_params
    real(c_double) :: dt
    integer(c_int) :: method
    real(c_double) :: coeff_implicit
    integer(c_int) :: advection
    real(c_double) :: advection_scale
    integer(c_int) :: diffusion
    real(c_double) :: diffusion_scale
    integer(c_int) :: iter_max
    real(c_double) :: tol_rel, tol_abs
    integer(c_int) :: n_skip
    integer(c_int) :: KE_diffusion
    real(c_double) :: KE_diffusion_scale
    integer(c_int) :: viscous_dissipation
    real(c_double) :: viscous_dissipation_scale
    integer(c_int) :: joule_heating
    real(c_double) :: joule_heating_scale
    integer(c_int) :: volumetric_heating
    real(c_double) :: volumetric_heating_scale
    integer(c_int) :: rk_stages
  end type

  type, bind(C), public :: moons_nrg_info
    integer(c_int) :: iterations
    real(c_double) :: ms
  end type

  type, bind(C), public :: moons_pcg_info
    integer(c_int) :: iterations
    integer(c_int) :: exit_loop
    real(c_double) :: res_norm0
    real(c_double) :: res_norm
    real(c_double) :: ms
  end type

  public :: moons_check
  public :: moons_init, moons_finalize, moons_synchronize
  public :: moons_nodes_uniform, moons_nodes_roberts_both, moons_nodes_ext
  public :: moons_mesh_create, moons_mesh_destroy
  public :: moons_field_create, moons_field_destroy, moons_field_set_bcs, moons_field_set_bc_props
  public :: moons_field_copy_bcs, moons_pcg_sf_set_k, moons_pcg_sf_update_mfp
  public :: moons_field_upload, moons_field_download, moons_field_fill
  public :: moons_apply_bcs, moons_grad, moons_div
  public :: moons_pcg_sf_create, moons_pcg_destroy, moons_pcg_set_isp, moons_pcg_sf_solve
  public :: moons_clean_div
  public :: moons_fft_create, moons_fft_destroy, moons_fft_solve, moons_clean_div_fft, moons_fft_set_coefficients
  public :: moons_field_copy, moons_apply_bcs_implicit, moons_curl, moons_interp, moons_advect_B
  public :: moons_pcg_vf_create, moons_pcg_vf_destroy, moons_pcg_vf_set_isp, moons_pcg_vf_update_mfp
  public :: moons_pcg_vf_solve
  public :: moons_mhd_create, moons_mhd_destroy, moons_mhd_step, moons_mhd_graph_capture, moons_mhd_graph_launch
  public :: moons_mhd_set_buoyancy
  public :: moons_field_export, moons_field_import, moons_export_raw, moons_import_raw
  public :: moons_nrg_create, moons_nrg_destroy, moons_nrg_step, moons_nrg_set_joule, moons_nrg_field

  interface
    integer(c_int) function moons_init(device) bind(C, name="moons_init")
      import c_int
      integer(c_int), value :: device
    end function
    integer(c_int) function moons_finalize() bind(C, name="moons_finalize")
      import c_int
    end function
    integer(c_int) function moons_synchronize() bind(C, name="moons_synchronize")
      import c_int
    end function
    type(c_ptr) function moons_last_error() bind(C, name="moons_last_error")
      import c_ptr
    end function
    integer(c_int) function moons_nodes_uniform(hmin, hmax, N, hn) bind(C, name="moons_nodes_uniform")
      import c_int, c_double
      real(c_double), value :: hmin, hmax
      integer(c_int), value :: N
      real(c_double), intent(out) :: hn(*)
    end function
    integer(c_int) function moons_nodes_roberts_both(hmin, hmax, N, beta, hn) bind(C, name="moons_nodes_roberts_both")
      import c_int, c_double
      real(c_double), value :: hmin, hmax, beta
      integer(c_int), value :: N
      real(c_double), intent(out) :: hn(*)
    end function
    integer(c_int) function moons_nodes_ext(hn, sn, kind, L, N, n_iter, max_stretch_ratio, out, cap, n_out) &
        bind(C, name="moons_nodes_ext")
      import c_int, c_double
      real(c_double), intent(in) :: hn(*)
      integer(c_int), value :: sn, kind, N, n_iter, cap
      real(c_double), value :: L, max_stretch_ratio
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: n_out
    end function
    integer(c_int) function moons_mesh_create(hx, sx, hy, sy, hz, sz, m) bind(C, name="moons_mesh_create")
      import c_int, c_double, c_ptr
      real(c_double), intent(in) :: hx(*), hy(*), hz(*)
      integer(c_int), value :: sx, sy, sz
      type(c_ptr), intent(out) :: m
    end function
    integer(c_int) function moons_mesh_destroy(m) bind(C, name="moons_mesh_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function moons_field_create(m, dl, f) bind(C, name="moons_field_create")
      import c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: dl
      type(c_ptr), intent(out) :: f
    end function
    integer(c_int) function moons_field_destroy(f) bind(C, name="moons_field_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: f
    end function
    integer(c_int) function moons_field_set_bcs(f, types, vals, all_neumann) bind(C, name="moons_field_set_bcs")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: f
      integer(c_int), intent(in) :: types(6)
      real(c_double), intent(in) :: vals(6)
      integer(c_int), value :: all_neumann
    end function
    integer(c_int) function moons_field_set_bc_props(f, c_w, robin_coeff) bind(C, name="moons_field_set_bc_props")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: f
      real(c_double), intent(in) :: c_w(6), robin_coeff(6)
    end function
    integer(c_int) function moons_field_copy_bcs(dst, src) bind(C, name="moons_field_copy_bcs")
      import c_int, c_ptr
      type(c_ptr), value :: dst, src
    end function
    integer(c_int) function moons_pcg_sf_set_k(pcg, k) bind(C, name="moons_pcg_sf_set_k")
      import c_int, c_ptr
      type(c_ptr), value :: pcg
      type(c_ptr), intent(in) :: k(3)
    end function
    integer(c_int) function moons_pcg_sf_update_mfp(pcg, coeff, update_prec) bind(C, name="moons_pcg_sf_update_mfp")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: pcg
      real(c_double), value :: coeff
      integer(c_int), value :: update_prec
    end function
    integer(c_int) function moons_field_upload(f, host, n) bind(C, name="moons_field_upload")
      import c_int, c_double, c_ptr, c_size_t
      type(c_ptr), value :: f
      real(c_double), intent(in) :: host(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function moons_field_download(f, host, n) bind(C, name="moons_field_download")
      import c_int, c_double, c_ptr, c_size_t
      type(c_ptr), value :: f
      real(c_double), intent(out) :: host(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function moons_field_export(f, path) bind(C, name="moons_field_export")
      import c_int, c_ptr, c_char
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function moons_field_import(f, path) bind(C, name="moons_field_import")
      import c_int, c_ptr, c_char
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function moons_export_raw(f, dir, name, pad, t) bind(C, name="moons_export_raw")
      import c_int, c_ptr, c_char
      type(c_ptr), value :: f, t
      character(kind=c_char), intent(in) :: dir(*), name(*)
      integer(c_int), value :: pad
    end function
    integer(c_int) function moons_import_raw(f, dir, name, pad) bind(C, name="moons_import_raw")
      import c_int, c_ptr, c_char
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: dir(*), name(*)
      integer(c_int), value :: pad
    end function
    integer(c_int) function moons_field_fill(f, v) bind(C, name="moons_field_fill")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: f
      real(c_double), value :: v
    end function
    integer(c_int) function moons_apply_bcs(f) bind(C, name="moons_apply_bcs")
      import c_int, c_ptr
      type(c_ptr), value :: f
    end function
    integer(c_int) function moons_grad(gx, gy, gz, u) bind(C, name="moons_grad")
      import c_int, c_ptr
      type(c_ptr), value :: gx, gy, gz, u
    end function
    integer(c_int) function moons_div(d, u, v, w) bind(C, name="moons_div")
      import c_int, c_ptr
      type(c_ptr), value :: d, u, v, w
    end function
    integer(c_int) function moons_pcg_sf_create(m, op, prec, x, pcg) bind(C, name="moons_pcg_sf_create")
      import c_int, c_ptr
      type(c_ptr), value :: m, x
      integer(c_int), value :: op, prec
      type(c_ptr), intent(out) :: pcg
    end function
    integer(c_int) function moons_pcg_destroy(pcg) bind(C, name="moons_pcg_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: pcg
    end function
    integer(c_int) function moons_pcg_set_isp(pcg, iter_max, tol_rel, tol_abs, n_skip) bind(C, name="moons_pcg_set_isp")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: pcg
      integer(c_int), value :: iter_max, n_skip
      real(c_double), value :: tol_rel, tol_abs
    end function
    integer(c_int) function moons_pcg_sf_solve(pcg, x, b, info) bind(C, name="moons_pcg_sf_solve")
      import c_int, c_ptr, moons_pcg_info
      type(c_ptr), value :: pcg, x, b
      type(moons_pcg_info), intent(out) :: info
    end function
    integer(c_int) function moons_clean_div(pcg, X, Xstar, phi, scale, temp_F, temp_CC, info) &
        bind(C, name="moons_clean_div")
      import c_int, c_double, c_ptr, moons_pcg_info
      type(c_ptr), value :: pcg, phi, temp_CC
      type(c_ptr), intent(in) :: X(3), Xstar(3), temp_F(3)
      real(c_double), value :: scale
      type(moons_pcg_info), intent(out) :: info
    end function
    integer(c_int) function moons_fft_create(m, dir, fft) bind(C, name="moons_fft_create")
      import c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: dir
      type(c_ptr), intent(out) :: fft
    end function
    integer(c_int) function moons_fft_set_coefficients(fft, kind) bind(C, name="moons_fft_set_coefficients")
      import c_int, c_ptr
      type(c_ptr), value :: fft
      integer(c_int), value :: kind
    end function
    integer(c_int) function moons_fft_destroy(fft) bind(C, name="moons_fft_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: fft
    end function
    integer(c_int) function moons_fft_solve(fft, u, f) bind(C, name="moons_fft_solve")
      import c_int, c_ptr
      type(c_ptr), value :: fft, u, f
    end function
    integer(c_int) function moons_clean_div_fft(fft, X, Xstar, phi, scale, temp_F, temp_CC) &
        bind(C, name="moons_clean_div_fft")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: fft, phi, temp_CC
      type(c_ptr), intent(in) :: X(3), Xstar(3), temp_F(3)
      real(c_double), value :: scale
    end function
    integer(c_int) function moons_field_copy(dst, src) bind(C, name="moons_field_copy")
      import c_int, c_ptr
      type(c_ptr), value :: dst, src
    end function
    integer(c_int) function moons_apply_bcs_implicit(f) bind(C, name="moons_apply_bcs_implicit")
      import c_int, c_ptr
      type(c_ptr), value :: f
    end function
    integer(c_int) function moons_curl(c, u, v, w, dir) bind(C, name="moons_curl")
      import c_int, c_ptr
      type(c_ptr), value :: c, u, v, w
      integer(c_int), value :: dir
    end function
    integer(c_int) function moons_interp(out, f, dir) bind(C, name="moons_interp")
      import c_int, c_ptr
      type(c_ptr), value :: out, f
      integer(c_int), value :: dir
    end function
    integer(c_int) function moons_advect_B(adv, U_E, B, temp_E_TF, temp_E) bind(C, name="moons_advect_B")
      import c_int, c_ptr
      type(c_ptr), intent(in) :: adv(3), U_E(6), B(3), temp_E_TF(6), temp_E(3)
    end function
    integer(c_int) function moons_pcg_vf_create(m, op, x, k, pcg) bind(C, name="moons_pcg_vf_create")
      import c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: op
      type(c_ptr), intent(in) :: x(3), k(3)
      type(c_ptr), intent(out) :: pcg
    end function
    integer(c_int) function moons_pcg_vf_destroy(pcg) bind(C, name="moons_pcg_vf_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: pcg
    end function
    integer(c_int) function moons_pcg_vf_set_isp(pcg, iter_max, tol_rel, tol_abs, n_skip) &
        bind(C, name="moons_pcg_vf_set_isp")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: pcg
      integer(c_int), value :: iter_max, n_skip
      real(c_double), value :: tol_rel, tol_abs
    end function
    integer(c_int) function moons_pcg_vf_update_mfp(pcg, coeff, update_prec) bind(C, name="moons_pcg_vf_update_mfp")
      import c_int, c_double, c_ptr
      type(c_ptr), value :: pcg
      real(c_double), value :: coeff
      integer(c_int), value :: update_prec
    end function
    integer(c_int) function moons_pcg_vf_solve(pcg, x, b, info) bind(C, name="moons_pcg_vf_solve")
      import c_int, c_ptr, moons_pcg_info
      type(c_ptr), value :: pcg
      type(c_ptr), intent(in) :: x(3), b(3)
      type(moons_pcg_info), intent(out) :: info
    end function
    integer(c_int) function moons_mhd_create(m, params, U, p, B, phi, s) bind(C, name="moons_mhd_create")
      import c_int, c_ptr, moons_mhd_params
      type(c_ptr), value :: m, p, phi
      type(moons_mhd_params), intent(in) :: params
      type(c_ptr), intent(in) :: U(3), B(3)
      type(c_ptr), intent(out) :: s
    end function
    integer(c_int) function moons_mhd_destroy(s) bind(C, name="moons_mhd_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function moons_mhd_step(s, n, info) bind(C, name="moons_mhd_step")
      import c_int, c_ptr, moons_mhd_info
      type(c_ptr), value :: s
      integer(c_int), value :: n
      type(moons_mhd_info), intent(out) :: info
    end function
    integer(c_int) function moons_mhd_graph_capture(s) bind(C, name="moons_mhd_graph_capture")
      import c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function moons_mhd_graph_launch(s, n, stream) bind(C, name="moons_mhd_graph_launch")
      import c_int, c_ptr
      type(c_ptr), value :: s, stream
      integer(c_int), value :: n
    end function
    ! buoyancy F += cellCenter2Face(T g) * buoyancy_scale, gravity F += cellCenter2Face(g) * gravity_scale
    ! (compute_add_buoyancy / compute_add_gravity, momentum_sources.f90:133-167); gravity(3) CC fields
    integer(c_int) function moons_mhd_set_buoyancy(s, T, gravity, add_buoyancy, buoyancy_scale, add_gravity, &
                                                   gravity_scale) bind(C, name="moons_mhd_set_buoyancy")
      import c_int, c_ptr, c_double
      type(c_ptr), value :: s, T
      type(c_ptr), intent(in) :: gravity(3)
      integer(c_int), value :: add_buoyancy, add_gravity
      real(c_double), value :: buoyancy_scale, gravity_scale
    end function
    integer(c_int) function moons_nrg_create(m, params, T, U, k, s) bind(C, name="moons_nrg_create")
      import c_int, c_ptr, moons_nrg_params
      type(c_ptr), value :: m, T
      type(moons_nrg_params), intent(in) :: params
      type(c_ptr), intent(in) :: U(3), k(3)
      type(c_ptr), intent(out) :: s
    end function
    integer(c_int) function moons_nrg_destroy(s) bind(C, name="moons_nrg_destroy")
      import c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function moons_nrg_step(s, n, info) bind(C, name="moons_nrg_step")
      import c_int, c_ptr, moons_nrg_info
      type(c_ptr), value :: s
      integer(c_int), value :: n
      type(moons_nrg_info), intent(out) :: info
    end function
    integer(c_int) function moons_nrg_set_joule(s, J, sigma_inv_cc) bind(C, name="moons_nrg_set_joule")
      import c_int, c_ptr
      type(c_ptr), value :: s, sigma_inv_cc
      type(c_ptr), intent(in) :: J(3)
    end function
    integer(c_int) function moons_nrg_field(s, which, f) bind(C, name="moons_nrg_field")
      import c_int, c_ptr
      type(c_ptr), value :: s
      integer(c_int), value :: which
      type(c_ptr), intent(out) :: f
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Non-zero status: print the library's message and stop, the reference's
  ! convention for errors.
  subroutine moons_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    type(c_ptr) :: p
    character(kind=c_char), pointer :: msg(:)
    integer(c_size_t) :: n
    character(len=:), allocatable :: text
    integer :: i
    if (status .eq. 0) return
    p = moons_last_error()
    n = c_strlen(p)
    call c_f_pointer(p, msg, [n])
    allocate(character(len=n) :: text)
    do i = 1, int(n)
      text(i:i) = msg(i)
    enddo
    write(*, '(a)') 'Error: '//where//': '//text
    error stop 'libmoons_amd call failed'
  end subroutine

end module
