/*
 * moons_oracle.h -- CPU FP64 restatement of the MOONS per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * moons_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The product never links, loads or calls it.
 *
 * Layout conventions follow the reference (Fortran column-major, i fastest,
 * one ghost layer): a field of shape s[0] x s[1] x s[2] is stored as
 * f[i + s[0]*(j + s[1]*k)], 0-based here where the reference is 1-based.
 * Along direction d a field is node-centred (N) or cell-centred (CC):
 *   N along d  -> s[d] = sn_d = N_d + 3   (reference hw/grid/grid_extend.f90:220-234)
 *   CC along d -> s[d] = sc_d = N_d + 2
 * A data location is a 3-bit mask, bit d set when the field is N along d:
 *   CC = 0, Face-x = 1, Face-y = 2, Face-z = 4, Edge-z = 3, Edge-y = 5,
 *   Edge-x = 6, Node = 7  (reference gen/fields/data_location.f90:38-51).
 *
 * Parity status of each function is stated in the header of moons_oracle.c.
 */
#ifndef MOONS_ORACLE_H
#define MOONS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DL_CC 0
#define ORC_DL_FX 1
#define ORC_DL_FY 2
#define ORC_DL_FZ 4
#define ORC_DL_EX 6
#define ORC_DL_EY 5
#define ORC_DL_EZ 3
#define ORC_DL_N 7

/* BC types (reference hw/boundary_conditions; values are this library's own) */
#define ORC_BC_NONE 0
#define ORC_BC_DIRICHLET 1
#define ORC_BC_NEUMANN 2
#define ORC_BC_PERIODIC 3
#define ORC_BC_SYMMETRIC 4
#define ORC_BC_ANTISYMMETRIC 5
#define ORC_BC_ROBIN 6

typedef struct {
  int sn, sc;
  double *hn, *hc, *dhn, *dhc;        /* sn, sc, sn-1, sc-1 */
  double *cc2n_D, *cc2n_U;            /* sc-1 : staggered CC->N   */
  double *n2cc_D, *n2cc_U;            /* sn-1 : staggered N->CC   */
  double *theta_D, *theta_U;          /* sc-1 : C->N interpolation */
  double *colCC_L[2], *colCC_D[2], *colCC_U[2]; /* sc-2, n = 1,2 */
  double *colN_L[2], *colN_D[2], *colN_U[2];    /* sn-2, n = 1,2 */
} orc_coords;

typedef struct {
  orc_coords c[3];
} orc_grid;

typedef struct {
  int type[6];   /* faces xmin,xmax,ymin,ymax,zmin,zmax */
  double val[6]; /* uniform BC value per face */
  double cw[6];  /* wall conductance ratio c_w per face (Robin faces only) */
} orc_bcs;

/* ---- 1D grids ---- */
int orc_uniform_nodes(double hmin, double hmax, int N, double *hn_out /* N+3 */);
int orc_roberts_both_nodes(double hmin, double hmax, int N, double beta, double *hn_out /* N+3 */);
int orc_roberts_left_nodes(double hmin, double hmax, int N, double beta, double *hn_out);
int orc_roberts_right_nodes(double hmin, double hmax, int N, double beta, double *hn_out);

/* ---- grid ---- */
orc_grid *orc_grid_create(const double *hn_x, int sn_x, const double *hn_y, int sn_y,
                          const double *hn_z, int sn_z);
void orc_grid_destroy(orc_grid *g);
void orc_field_shape(const orc_grid *g, int dl, int s[3]);
size_t orc_field_size(const orc_grid *g, int dl);
/* copy out one coordinate array: which = 0 hn,1 hc,2 dhn,3 dhc,4 cc2n_D,5 cc2n_U,
   6 n2cc_D,7 n2cc_U,8 theta_D,9 theta_U,10..15 colCC(1)L,D,U,(2)L,D,U,
   16..21 colN(1)L,D,U,(2)L,D,U.  Returns length. */
int orc_coord_array(const orc_grid *g, int dir, int which, double *out);

/* ---- derivative dispatch (reference hw/ops/ops_del.f90) ----
   gen: 1 assign, 2 add, 3 subtract; n = 1 or 2; dir = 0,1,2 */
void orc_del(double *dfdh, int dl_out, const double *f, int dl_in, const orc_grid *g,
             int n, int dir, int pad, int gen);

/* ---- discrete operators (reference hw/ops/ops_discrete.f90) ---- */
void orc_grad(double *gx, double *gy, double *gz, const double *u, int dl_u, const orc_grid *g);
void orc_div(double *divU, int dl_out, const double *u, const double *v, const double *w,
             int dl_u, const orc_grid *g, const orc_bcs *bc_out);
void orc_curl_component(double *curlU, int dl_out, const double *u, const double *v,
                        const double *w, int dl_in, const orc_grid *g, int dir);
void orc_lap_centered_sf(double *lapU, const double *U, int dl, const orc_grid *g,
                         const orc_bcs *bc_out);

/* ---- interpolation (reference hw/ops/ops_interp.f90:190-221,
   hw/grid_field/GF_interp/GF_interp.f90:19-76, GF_extrap.f90:17-46) ----
   C->N along dir (theta weights, ghost nodes extrapolated) or N->C (half sum). */
void orc_interp(double *out, int dl_out, const double *in, int dl_in, const orc_grid *g, int dir);

/* ---- constrained transport (reference hw/ops/ops_advect.f90:186-218) ----
   U_E: the six off-diagonal edge components of velocity, ordered
   U%y%x, U%z%x (on x-edges), U%x%y, U%z%y (y-edges), U%x%z, U%y%z (z-edges).
   B: the three face components.  E = U_E x face2Edge_no_diag(B) on edges;
   advect_B: adv = curl(E) on faces. */
void orc_edge_cross_face_E(double *Ex, double *Ey, double *Ez, double *const U_E[6], double *const B[3],
                           const orc_grid *g);
void orc_advect_B(double *adv[3], double *const U_E[6], double *const B[3], const orc_grid *g);

/* ---- BCs (reference hw/BCs/direct and hw/BCs/implicit raw kernels) ---- */
void orc_apply_bcs(double *u, int dl, const orc_grid *g, const orc_bcs *bc, int implicit_);
/* dh of the face sub-domain at the min (is_max 0) or max face of a direction */
double orc_face_dh(const orc_coords *c, int is_max);
double orc_grid_face_dh(const orc_grid *g, int d, int is_max);
void orc_assign_ghost_xperiodic(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val);
void orc_assign_wall_dirichlet(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val);
void orc_multiply_wall_neumann(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val);

/* ---- auxiliaries ---- */
void orc_volume(double *vol, int dl, const orc_grid *g);
void orc_lap_diagonal(double *diag, int dl, const orc_grid *g);
/* curl_curl_diagonal: hw/grid_field/GF_diagonals.f90:37-71; sig on the x, y, z edges, dg on the faces */
void orc_curl_curl_diagonal(double *dg[3], double *const sig[3], const orc_grid *g);
double orc_sum_pad(const double *u, const int s[3], int pad);
/* Threads of the OpenMP loops (1, the default: sequential, the exact order). */
void orc_set_threads(int n);
double orc_sum_all(const double *u, size_t n);

/* ---- PCG for the scalar Laplacian (PPE / phi cleaning) ----
   Restates hw/solvers/PCG/PCG_solver_algorithm.f90:37-107 with op = Lap_uniform_SF
   (hw/solvers/matrix_free_operators.f90:35-62) and prec = prec_Lap_SF
   (hw/solvers/preconditioners/preconditioners.f90:55-71).
   x is in/out (ghosts included); b is the right-hand side.  res_hist (optional, length
   iter_max+1) receives res_norm%L2 after each check; out[0] = iterations,
   out[1] = res_norm0%L2, out[2] = final res_norm%L2.  */
int orc_pcg_lap_sf(const orc_grid *g, int dl, const orc_bcs *bc, int all_neumann,
                   double *x, const double *b, int iter_max, double tol_rel, double tol_abs,
                   int n_skip_check_res, double *out, double *res_hist);

/* nrg_diffusion (matrix_free_operators.f90:148-174): Ax = c div(k grad x) + x
   on a cell-centred x (implicit BCs applied to x first); k: x-, y-, z-face arrays. */
void orc_nrg_op(double *Ax, double *x, double *const k[3], double c, const orc_grid *g, const orc_bcs *bc);
/* The same PCG with op = nrg_diffusion and prec = prec_lap_SF (the energy
   equation's solver, energy_extend.f90:129-130). */
int orc_pcg_nrg_sf(const orc_grid *g, const orc_bcs *bc, int all_neumann, double *const k[3], double c,
                   double *x, const double *b, int iter_max, double tol_rel, double tol_abs,
                   int n_skip_check_res, double *out, double *res_hist);

/* ---- vector-field PCG (momentum and induction implicit diffusion) ----
   Restates solve_PCG_VF (hw/solvers/PCG/PCG_solver_algorithm.f90:109-179) and
   modify_RHS_VF (PCG_aux.f90:56-79) with
     kind 1: mom_diffusion  A = I + c lap_centered_VF  (matrix_free_operators.f90:176-202),
             prec_mom_VF    (preconditioners.f90:86-113)
     kind 2: ind_diffusion  A = I + c curl(k curl)     (matrix_free_operators.f90:120-146),
             prec_ind_VF    (preconditioners.f90:115-129), k = sigmaInv on edges.
   x, b: the three face components (ghosts included); bc: BCs of each component. */
#define ORC_OP_MOM_DIFFUSION 1
#define ORC_OP_IND_DIFFUSION 2
void orc_op_vf(double *Ax[3], double *x[3], int kind, double c, double *const k[3], const orc_grid *g,
               const orc_bcs bc[3], int implicit_);
int orc_pcg_vf(const orc_grid *g, int kind, double c, double *const k[3], const orc_bcs bc[3], double *x[3],
               double *const b[3], int iter_max, double tol_rel, double tol_abs, int n_skip, double *out,
               double *res_hist);
void orc_minv_vf(double *Minv[3], int kind, double c, double *const k[3], const orc_grid *g);

/* seeded SplitMix64 -> uniform [-1,1) */
void orc_splitmix_fill(double *out, size_t n, unsigned long long seed, double amplitude);

#ifdef __cplusplus
}
#endif
#endif
