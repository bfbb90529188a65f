! ref_bridge.f90 -- bind(C) entry points over routines of the reference
! compiled from their own sources (oracle/Makefile, target _ref).
!
! TEST INFRASTRUCTURE ONLY: lets tests/ check the C restatement in
! oracle/moons_oracle.c against the reference's own compiled code for the
! parts that compile stand-alone: 1D grid distributions, derivative and
! interpolation stencil coefficients, the pencil kernels of stencils.f90 and
! the raw ghost-plane BC kernels.  Nothing here is used by the product.
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
  implicit none
contains

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
end module
