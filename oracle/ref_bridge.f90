! ref_bridge.f90 -- bind(C) entry points over routines of the reference
! compiled from their own sources (oracle/Makefile, target _ref).
!
! TEST INFRASTRUCTURE ONLY: lets tests/ check the C restatement in
! oracle/moons_oracle.c against the reference's own compiled code for the
! parts that compile stand-alone: 1D grid distributions, derivative and
! interpolation stencil coefficients, the pencil kernels of stencils.f90,
! the raw ghost-plane BC kernels, the Laplacian and curl-curl diagonals of
! the Jacobi preconditioners, the staggered interpolations with their ghost
! extrapolation, and the grid_field restart primitives and Tecplot raw files
! written and read by the reference's own routines.
! Nothing here is used by the product.
module moons_ref_bridge
  use iso_c_binding
  use current_precision_mod
  use sparse_mod
  use coordinate_distribution_funcs_mod
  use derivative_stencils_mod
  use interpolation_stencils_mod
  use stencils_mod
  use apply_BCs_faces_raw_mod
  use apply_BCs_faces_raw_implicit_mod
  use string_mod, only: string, str, delete
  use IO_tools_mod, only: new_and_open
  use grid_mod, only: grid
  use grid_extend_mod, only: init, get_shape, get_face_GI
  use data_location_mod, only: data_location
  use data_location_extend_mod, only: init_CC, init_Node, init_Face, init_Edge
  use grid_field_mod, only: grid_field, export, import
  use exp_Tecplot_Header_mod, only: exp_Header_3D_1C
  use construct_suffix_mod, only: construct_suffix
  use GF_export_mod, only: exp_3D_1C_GF
  use GF_import_mod, only: imp_3D_1C_GF
  use grid_continue_mod
  use mesh_quality_params_mod, only: mesh_quality_params
  use coordinates_mod, only: coordinates
  use coordinates_extend_mod, only: coord_init => init
  use overlap_mod, only: overlap
  use overlap_extend_mod, only: get_C_overlap, get_N_overlap
  use GF_diagonals_mod, only: laplacian_diagonal, curl_curl_diagonal
  use GF_interp_mod, only: interp_C2N, interp_N2C
  implicit none
contains

  ! ext_uniform_IO(g, N, dir) (handwritten/grid/grid_continue.f90:41-47) on a
  ! grid whose three directions all have the node array hn; out gets the
  ! extended nodes of direction 1 (sn + 2N entries)
  subroutine ref_ext_uniform(hn, sn, N, out) bind(C, name="ref_ext_uniform")
    integer(c_int), value :: sn, N
    real(c_double), intent(in) :: hn(sn)
    real(c_double), intent(out) :: out(sn + 2 * N)
    type(grid) :: g
    call init(g, hn, hn, hn)
    call ext_uniform_IO(g, int(N), 1)
    out = g%c(1)%hn%f
  end subroutine

  ! The grid_continue extension of kind (0 ext_uniform_IO, 1 ext_Roberts_near_IO,
  ! 2 ext_Roberts_far_IO, 3 ext_Roberts_B_IO, 4 ext_prep_uniform_IO,
  ! 5 ext_prep_Roberts_L_IO, 6 _R, 7 _B, 8 ext_prep_Roberts_C2F_IO,
  ! 9 ext_app_uniform_IO, 10 ext_app_Roberts_L_IO, 11 _R, 12 _B,
  ! 13 ext_app_Roberts_C2F_IO; handwritten/grid/grid_continue.f90) of direction 1
  ! of a grid with node array hn in every direction, segment length L, N cells,
  ! mesh quality N_iter and max_mesh_stretch_ratio; out (cap entries) gets
  ! the extended nodes, n_out their number.
  subroutine ref_ext(kind, hn, sn, L, N, n_iter, max_ratio, out, cap, n_out) bind(C, name="ref_ext")
    integer(c_int), value :: kind, sn, N, n_iter, cap
    real(c_double), value :: L, max_ratio
    real(c_double), intent(in) :: hn(sn)
    real(c_double), intent(out) :: out(cap)
    integer(c_int), intent(out) :: n_out
    type(grid) :: g
    type(mesh_quality_params) :: MQP
    integer :: Ni
    MQP%N_iter = n_iter
    MQP%max_mesh_stretch_ratio = max_ratio
    MQP%N_max_points_add = max(n_iter - 1, 0)
    MQP%auto_find_N = n_iter > 1
    Ni = int(N)
    call init(g, hn, hn, hn)
    select case (kind)
    case (0); call ext_uniform_IO(g, Ni, 1)
    case (1); call ext_Roberts_near_IO(g, L, Ni, 1, MQP)
    case (2); call ext_Roberts_far_IO(g, L, Ni, 1, MQP)
    case (3); call ext_Roberts_B_IO(g, L, Ni, 1, MQP)
    case (4); call ext_prep_uniform_IO(g, Ni, 1)
    case (5); call ext_prep_Roberts_L_IO(g, L, Ni, 1, MQP)
    case (6); call ext_prep_Roberts_R_IO(g, L, Ni, 1, MQP)
    case (7); call ext_prep_Roberts_B_IO(g, L, Ni, 1, MQP)
    case (8); call ext_prep_Roberts_C2F_IO(g, L, Ni, 1, MQP)
    case (9); call ext_app_uniform_IO(g, Ni, 1)
    case (10); call ext_app_Roberts_L_IO(g, L, Ni, 1, MQP)
    case (11); call ext_app_Roberts_R_IO(g, L, Ni, 1, MQP)
    case (12); call ext_app_Roberts_B_IO(g, L, Ni, 1, MQP)
    case (13); call ext_app_Roberts_C2F_IO(g, L, Ni, 1, MQP)
    end select
    n_out = int(g%c(1)%sn, c_int)
    if (n_out <= cap) out(1:n_out) = g%c(1)%hn%f
  end subroutine

  ! get_C_overlap (kind 0) / get_N_overlap (kind 1) of the coordinates hn1
  ! (R1) and hn2 (R2), tol 1e-12, padding p (handwritten/sub_domain/overlap_extend.f90:204-278);
  ! r = (i1(1), i1(2), i2(1), i2(2)), 1-based, 0 when none
  subroutine ref_overlap(kind, hn1, sn1, hn2, sn2, p, r) bind(C, name="ref_overlap")
    integer(c_int), value :: kind, sn1, sn2, p
    real(c_double), intent(in) :: hn1(sn1), hn2(sn2)
    integer(c_int), intent(out) :: r(4)
    type(coordinates), dimension(2) :: c
    type(overlap) :: OL
    call coord_init(c(1), hn1, sn1)
    call coord_init(c(2), hn2, sn2)
    if (kind == 0) then
      OL = get_C_overlap(c, 1.0d-12, int(p))
    else
      OL = get_N_overlap(c, 1.0d-12, int(p))
    end if
    r = (/ OL%i1(1), OL%i1(2), OL%i2(1), OL%i2(2) /)
  end subroutine

  function c2f(c, n) result(f)
    integer(c_int), intent(in) :: n
    character(kind=c_char), intent(in) :: c(n)
    character(len=n) :: f
    integer :: i
    do i = 1, n
      f(i:i) = c(i)
    end do
  end function

  ! Data location by the library's code: bit d set = node-centred along d+1.
  subroutine dl_from_code(DL, code)
    type(data_location), intent(inout) :: DL
    integer(c_int), intent(in) :: code
    select case (code)
    case (0); call init_CC(DL)
    case (7); call init_Node(DL)
    case (1); call init_Face(DL, 1)
    case (2); call init_Face(DL, 2)
    case (4); call init_Face(DL, 3)
    case (6); call init_Edge(DL, 1)
    case (5); call init_Edge(DL, 2)
    case (3); call init_Edge(DL, 3)
    end select
  end subroutine

  ! export(grid_field, un) of generated_code/grid_field/grid_field.f90 on a
  ! grid_field holding f(s(1),s(2),s(3))
  subroutine ref_gf_export(f, s, path, plen) bind(C, name="ref_gf_export")
    integer(c_int), intent(in) :: s(3)
    real(c_double), intent(in) :: f(s(1), s(2), s(3))
    integer(c_int), value :: plen
    character(kind=c_char), intent(in) :: path(plen)
    type(grid_field) :: GF
    integer :: un
    allocate(GF%f(s(1), s(2), s(3)))
    GF%f = f
    GF%s = s
    GF%s_1D = s(1) * s(2) * s(3)
    open(newunit=un, file=c2f(path, plen), status='replace', action='write')
    call export(GF, un)
    close(un)
  end subroutine

  ! import(grid_field, un): s gets the shape read, f (n entries) the values
  subroutine ref_gf_import(path, plen, f, n, s) bind(C, name="ref_gf_import")
    integer(c_int), value :: plen, n
    character(kind=c_char), intent(in) :: path(plen)
    real(c_double), intent(out) :: f(n)
    integer(c_int), intent(out) :: s(3)
    type(grid_field) :: GF
    integer :: un
    open(newunit=un, file=c2f(path, plen), status='old', action='read')
    call import(GF, un)
    close(un)
    s = 0
    if (allocated(GF%f)) then
      s = shape(GF%f)
      if (size(GF%f) == n) f = reshape(GF%f, (/ n /))
    end if
  end subroutine

  ! export_3D_1C_steady (handwritten/IO/IO_export.f90:265-277) with
  ! exp_3D_1C (handwritten/IO/base_export.f90:150-159) on a one-block grid
  subroutine ref_export_raw(hx, sx, hy, sy, hz, sz, code, f, nf, dir, dlen, name, nlen, pad) &
      bind(C, name="ref_export_raw")
    integer(c_int), value :: sx, sy, sz, code, nf, dlen, nlen, pad
    real(c_double), intent(in) :: hx(sx), hy(sy), hz(sz), f(nf)
    character(kind=c_char), intent(in) :: dir(dlen), name(nlen)
    type(grid) :: g
    type(data_location) :: DL
    type(grid_field) :: GF
    type(string) :: s
    integer :: sh(3), un
    call init(g, hx, hy, hz)
    call dl_from_code(DL, code)
    sh = get_shape(g, DL)
    allocate(GF%f(sh(1), sh(2), sh(3)))
    GF%f = reshape(f, sh)
    call construct_suffix(s, c2f(name, nlen), DL)
    un = new_and_open(c2f(dir, dlen), str(s))
    call exp_Header_3D_1C(un, str(s))
    call exp_3D_1C_GF(g, DL, 1, pad, un, GF)
    close(un)
    call delete(s)
  end subroutine

  ! import_3D_1C (handwritten/IO/IO_import.f90:20-33) with imp_3D_1C
  ! (handwritten/IO/base_import.f90:142-150) on a one-block grid; f holds the
  ! field before the call (its pad cells are kept)
  subroutine ref_import_raw(hx, sx, hy, sy, hz, sz, code, f, nf, dir, dlen, name, nlen, pad) &
      bind(C, name="ref_import_raw")
    integer(c_int), value :: sx, sy, sz, code, nf, dlen, nlen, pad
    real(c_double), intent(in) :: hx(sx), hy(sy), hz(sz)
    real(c_double), intent(inout) :: f(nf)
    character(kind=c_char), intent(in) :: dir(dlen), name(nlen)
    type(grid) :: g
    type(data_location) :: DL
    type(grid_field) :: GF
    type(string) :: s
    integer :: sh(3), un
    call init(g, hx, hy, hz)
    call dl_from_code(DL, code)
    sh = get_shape(g, DL)
    allocate(GF%f(sh(1), sh(2), sh(3)))
    GF%f = reshape(f, sh)
    call construct_suffix(s, c2f(name, nlen), DL)
    un = new_and_open(c2f(dir, dlen), str(s))
    read(un, *); read(un, *)
    call imp_3D_1C_GF(g, DL, pad, un, GF)
    close(un)
    call delete(s)
    f = reshape(GF%f, (/ nf /))
  end subroutine

  ! hn(1:N+1) = uniform(hmin,hmax,N) / robertsBoth / robertsLeft / robertsRight
  subroutine ref_grid(kind, hmin, hmax, N, beta, hn) bind(C, name="ref_grid")
    integer(c_int), value :: kind, N
    real(c_double), value :: hmin, hmax, beta
    real(c_double), intent(out) :: hn(N+1)
    select case (kind)
    case (0); hn = uniform(hmin, hmax, N)
    case (1); hn = robertsBoth(hmin, hmax, N, beta)
    case (2); hn = robertsLeft(hmin, hmax, N, beta)
    case (3); hn = robertsRight(hmin, hmax, N, beta)
    end select
  end subroutine

  ! kind: 0 staggered_CC2N, 1 staggered_N2CC, 2 collocated_CC_1, 3 collocated_CC_2,
  !       4 collocated_Node_1, 5 collocated_Node_2.  dh has s-1 entries.
  ! L,D,U receive up to s entries each; nL,nD,nU their lengths.
  subroutine ref_stencil(kind, dh, s, L, D, U, nL, nD, nU) bind(C, name="ref_stencil")
    integer(c_int), value :: kind, s
    real(c_double), intent(in) :: dh(s-1)
    real(c_double), intent(out) :: L(s), D(s), U(s)
    integer(c_int), intent(out) :: nL, nD, nU
    type(sparse) :: SP
    select case (kind)
    case (0); SP = staggered_CC2N(dh, s)
    case (1); SP = staggered_N2CC(dh, s)
    case (2); SP = collocated_CC_1(dh, s)
    case (3); SP = collocated_CC_2(dh, s)
    case (4); SP = collocated_Node_1(dh, s)
    case (5); SP = collocated_Node_2(dh, s)
    end select
    nL = SP%L%N; nD = SP%D%N; nU = SP%U%N
    L = 0.0_cp; D = 0.0_cp; U = 0.0_cp
    L(1:nL) = SP%L%f(1:nL); D(1:nD) = SP%D%f(1:nD); U(1:nU) = SP%U%f(1:nU)
  end subroutine

  ! interpolation_stencil(hc,hn,sc,sn): theta D and U (sc-1 entries)
  subroutine ref_interp_stencil(hc, sc, hn, sn, D, U, n) bind(C, name="ref_interp_stencil")
    integer(c_int), value :: sc, sn
    real(c_double), intent(in) :: hc(sc), hn(sn)
    real(c_double), intent(out) :: D(sc), U(sc)
    integer(c_int), intent(out) :: n
    type(sparse) :: SP
    SP = interpolation_stencil(hc, hn, sc, sn)
    n = SP%D%N
    D = 0.0_cp; U = 0.0_cp
    D(1:n) = SP%D%f(1:n); U(1:n) = SP%U%f(1:n)
  end subroutine

  ! Pencil kernels of stencils.f90 on one 1D pencil.
  ! kind: 0 stag_assign, 1 stag_add, 2 stag_subtract, 3 col_CC_assign,
  !       4 col_CC_add, 5 col_CC_subtract, 6 col_N_assign, 7 col_N_add, 8 col_N_subtract
  subroutine ref_pencil(kind, dfdh, sd, f, s, L, D, U, nc, gt) bind(C, name="ref_pencil")
    integer(c_int), value :: kind, sd, s, nc, gt
    real(c_double), intent(inout) :: dfdh(sd)
    real(c_double), intent(in) :: f(s), L(nc), D(nc), U(nc)
    type(sparse) :: SP
    allocate(SP%L%f(nc), SP%D%f(nc), SP%U%f(nc))
    SP%L%f = L; SP%D%f = D; SP%U%f = U
    SP%L%N = nc; SP%D%N = nc; SP%U%N = nc
    select case (kind)
    case (0); call stag_assign(dfdh, f, SP, s, sd, gt)
    case (1); call stag_add(dfdh, f, SP, s, sd, gt)
    case (2); call stag_subtract(dfdh, f, SP, s, sd, gt)
    case (3); call col_CC_assign(dfdh, f, SP, s, 0, 0)
    case (4); call col_CC_add(dfdh, f, SP, s, 0, 0)
    case (5); call col_CC_subtract(dfdh, f, SP, s, 0, 0)
    case (6); call col_N_assign(dfdh, f, SP, s, 0, 0)
    case (7); call col_N_add(dfdh, f, SP, s, 0, 0)
    case (8); call col_N_subtract(dfdh, f, SP, s, 0, 0)
    end select
  end subroutine

  ! Raw ghost-plane BC kernels on an x-by-y plane, p = 0.
  ! kind: 0 Dirichlet_C, 1 Dirichlet_N, 2 Neumann_C, 3 Neumann_N, 4 Periodic_C,
  !       5 Dirichlet_C_implicit, 6 Dirichlet_N_implicit, 7 Neumann_C_implicit
  subroutine ref_bc_plane(kind, ug, ub, ui, bvals, dh_nhat, x, y) bind(C, name="ref_bc_plane")
    integer(c_int), value :: kind, x, y
    real(c_double), intent(inout) :: ug(x, y), ub(x, y)
    real(c_double), intent(in) :: ui(x, y), bvals(x, y)
    real(c_double), value :: dh_nhat
    select case (kind)
    case (0); call apply_Dirichlet_C(ug, ui, bvals, x, y, 0)
    case (1); call apply_Dirichlet_N(ug, ub, ui, bvals, x, y, 0)
    case (2); call apply_Neumann_C(ug, ui, bvals, dh_nhat, x, y, 0)
    case (3); call apply_Neumann_N(ug, ui, bvals, dh_nhat, x, y, 0)
    case (4); call apply_Periodic_C(ug, ui, x, y, 0)
    case (5); call apply_Dirichlet_C_implicit(ug, ui, x, y, 0)
    case (6); call apply_Dirichlet_N_implicit(ug, ub, ui, x, y, 0)
    case (7); call apply_Neumann_C_implicit(ug, ui, x, y, 0)
    end select
  end subroutine

  ! Raw Robin kernels on an x-by-y plane, p = 0 (u + c_w du/dh = 0).
  ! kind: 0 Robin_C, 1 Robin_N, 2 Robin_C_implicit, 3 Robin_N_implicit
  subroutine ref_robin_plane(kind, ug, ub, ui, bvals, dh, nhat, c_w, x, y) bind(C, name="ref_robin_plane")
    integer(c_int), value :: kind, x, y
    real(c_double), intent(inout) :: ug(x, y)
    real(c_double), intent(in) :: ub(x, y), ui(x, y), bvals(x, y)
    real(c_double), value :: dh, nhat, c_w
    select case (kind)
    case (0); call apply_Robin_C(ug, ui, bvals, dh, nhat, c_w, x, y, 0)
    case (1); call apply_Robin_N(ug, ui, ub, bvals, dh, nhat, c_w, x, y, 0)
    case (2); call apply_Robin_C_implicit(ug, ui, dh, nhat, c_w, x, y, 0)
    case (3); call apply_Robin_N_implicit(ug, ub, ui, dh, nhat, c_w, x, y, 0)
    end select
  end subroutine

  ! dh of face (1..6) on a grid whose three directions have the node array
  ! hn (ghosts included): dhn(1) of the face grid get_face_GI keeps
  ! (handwritten/grid/grid_extend.f90:260-268), which get_dh_boundary returns
  ! for a grid of two or more cells (handwritten/sub_domain/sub_domain_extend.f90:125-134)
  ! and face_SD stores as FSD%dh (handwritten/sub_domain/face_SD_extend.f90:63).
  subroutine ref_face_dh(hn, sn, face, dh) bind(C, name="ref_face_dh")
    integer(c_int), value :: sn, face
    real(c_double), intent(in) :: hn(sn)
    real(c_double), intent(out) :: dh
    type(grid) :: g, gf
    call init(g, hn, hn, hn)
    call get_face_GI(gf, g, int(face))
    dh = gf%c((face + 1) / 2)%dhn%f(1)
  end subroutine

  ! A grid_field of the shape of DL (code) on g, zero-filled.
  subroutine gf_alloc(GF, g, code)
    type(grid_field), intent(inout) :: GF
    type(grid), intent(in) :: g
    integer(c_int), intent(in) :: code
    type(data_location) :: DL
    integer :: sh(3)
    call dl_from_code(DL, code)
    sh = get_shape(g, DL)
    allocate(GF%f(sh(1), sh(2), sh(3)))
    GF%f = 0.0_cp
    GF%s = sh
    GF%s_1D = sh(1) * sh(2) * sh(3)
  end subroutine

  ! laplacian_diagonal (handwritten/grid_field/GF_diagonals.f90:15-35) of a
  ! field at data location code on the grid (hx, hy, hz); out gets n values.
  subroutine ref_laplacian_diagonal(hx, sx, hy, sy, hz, sz, code, out, n) bind(C, name="ref_laplacian_diagonal")
    integer(c_int), value :: sx, sy, sz, code, n
    real(c_double), intent(in) :: hx(sx), hy(sy), hz(sz)
    real(c_double), intent(out) :: out(n)
    type(grid) :: g
    type(data_location) :: DL
    type(grid_field) :: GF
    call init(g, hx, hy, hz)
    call dl_from_code(DL, code)
    call gf_alloc(GF, g, code)
    call laplacian_diagonal(GF, g, DL)
    out = reshape(GF%f, (/ n /))
  end subroutine

  ! curl_curl_diagonal (GF_diagonals.f90:37-71): sig_x, sig_y, sig_z on the
  ! x, y, z edges (nsx, nsy, nsz values), dx, dy, dz on the x, y, z faces.
  subroutine ref_curl_curl_diagonal(hx, sx, hy, sy, hz, sz, sgx, nsx, sgy, nsy, sgz, nsz, dx, ndx, dy, ndy, dz, ndz) &
      bind(C, name="ref_curl_curl_diagonal")
    integer(c_int), value :: sx, sy, sz, nsx, nsy, nsz, ndx, ndy, ndz
    real(c_double), intent(in) :: hx(sx), hy(sy), hz(sz), sgx(nsx), sgy(nsy), sgz(nsz)
    real(c_double), intent(out) :: dx(ndx), dy(ndy), dz(ndz)
    type(grid) :: g
    type(grid_field) :: S1, S2, S3, D1, D2, D3
    call init(g, hx, hy, hz)
    call gf_alloc(S1, g, 6_c_int); call gf_alloc(S2, g, 5_c_int); call gf_alloc(S3, g, 3_c_int)
    call gf_alloc(D1, g, 1_c_int); call gf_alloc(D2, g, 2_c_int); call gf_alloc(D3, g, 4_c_int)
    S1%f = reshape(sgx, S1%s); S2%f = reshape(sgy, S2%s); S3%f = reshape(sgz, S3%s)
    call curl_curl_diagonal(D1, D2, D3, g, S1, S2, S3)
    dx = reshape(D1%f, (/ ndx /)); dy = reshape(D2%f, (/ ndy /)); dz = reshape(D3%f, (/ ndz /))
  end subroutine

  ! interpO2_SF's per-block step (handwritten/ops/ops_interp.f90:190-219):
  ! interp_C2N (with extrap_staggered of the two end planes) or interp_N2C
  ! (handwritten/grid_field/GF_interp/GF_interp.f90:19-76,
  ! GF_interp/GF_extrap.f90:17-44) along dir (1..3), from data location
  ! code_in to code_out.
  subroutine ref_interp(hx, sx, hy, sy, hz, sz, code_out, code_in, dir, fin, nin, fout, nout) &
      bind(C, name="ref_interp")
    integer(c_int), value :: sx, sy, sz, code_out, code_in, dir, nin, nout
    real(c_double), intent(in) :: hx(sx), hy(sy), hz(sz), fin(nin)
    real(c_double), intent(out) :: fout(nout)
    type(grid) :: g
    type(grid_field) :: GI, GO
    integer :: e(3)
    call init(g, hx, hy, hz)
    call gf_alloc(GI, g, code_in)
    call gf_alloc(GO, g, code_out)
    GI%f = reshape(fin, GI%s)
    e = 0
    e(dir) = 1
    if (btest(code_out, dir - 1)) then
      call interp_C2N(GO, GI, g, int(dir), e(1), e(2), e(3))
    else
      call interp_N2C(GO, GI, e(1), e(2), e(3))
    end if
    fout = reshape(GO%f, (/ nout /))
  end subroutine
end module
