/*
 * moons_amd.h -- C ABI of the MI355X-native MOONS hot path (libmoons_amd.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  Every
 * entry point returns MOONS_OK (0) or a non-zero status; moons_last_error()
 * returns the message of the last failure on the calling thread.  The
 * reference stops the program on error (`stop 'Error: ...'`,
 * code/handwritten/ops/ops_del.f90:267-288) -- the Fortran shim
 * (moons_amd/fortran/moons_amd_bindings.f90) turns a non-zero status into
 * `error stop` with this message, so callers see the reference behaviour.
 *
 * Data layout (device): one allocation per scalar field, Fortran column-major
 * with one ghost layer, exactly the layout of the reference's
 * grid_field%f(:,:,:) (code/generated_code/grid_field/grid_field.f90:39-43), so
 * host<->device transfers of a Fortran allocatable array are one memcpy.
 *
 * Threading: one host thread per library; the library owns one HIP stream
 * (default priority) and every call is stream-ordered on it.  Not re-entrant,
 * as the reference (single host thread + OpenMP inside each routine).
 * One block per mesh (the reference's m%s > 1 multi-block meshes are rejected).
 */
#ifndef MOONS_AMD_H
#define MOONS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define MOONS_OK 0
#define MOONS_ERR_ARG 1
#define MOONS_ERR_HIP 2
#define MOONS_ERR_STATE 3
#define MOONS_ERR_NAN 4
#define MOONS_ERR_IO 5 /* a file could not be opened, written or parsed */

/* --------------------------------------------------------- data location
 * 3-bit mask, bit d set when the field is node-centred along direction d
 * (code/generated_code/fields/data_location.f90:38-51; the shape along d is
 * sn = N+3 when node-centred, sc = N+2 when cell-centred,
 * code/handwritten/grid/grid_extend.f90:220-234). */
#define MOONS_DL_CC 0
#define MOONS_DL_FACE_X 1
#define MOONS_DL_FACE_Y 2
#define MOONS_DL_FACE_Z 4
#define MOONS_DL_EDGE_X 6
#define MOONS_DL_EDGE_Y 5
#define MOONS_DL_EDGE_Z 3
#define MOONS_DL_NODE 7

/* ------------------------------------------------------------- BC types
 * Face order xmin,xmax,ymin,ymax,zmin,zmax (code/handwritten/BCs/
 * face_edge_corner_indexing.f90:75-85,326-339). */
#define MOONS_BC_NONE 0
#define MOONS_BC_DIRICHLET 1
#define MOONS_BC_NEUMANN 2
#define MOONS_BC_PERIODIC 3
#define MOONS_BC_SYMMETRIC 4
#define MOONS_BC_ANTISYMMETRIC 5
/* u + c_w du/dh = 0 with the face's c_w (moons_field_set_bc_props): the
 * thin-wall conducting BC, apply_Robin_C/N
 * (code/handwritten/BCs/direct/apply_BCs_faces_raw.f90:236-282) and
 * apply_Robin_C/N_implicit
 * (code/handwritten/BCs/implicit/apply_BCs_faces_raw_implicit.f90:213-257).
 * As there, the cell-located rule's ghost factor is
 * (2 c_w nhat/dh - 1)/(2 c_w nhat/dh + 1) with nhat = -1 on min faces and
 * dh > 0, i.e. (2c + 1)/(2c - 1), c = c_w/dh, on min faces: the condition
 * u - c_w du/dn = 0 there (singular at c = 1/2).  Kept for parity. */
#define MOONS_BC_ROBIN 6

/* ----------------------------------------------------------- operators */
#define MOONS_OP_LAP_UNIFORM_SF 1 /* code/handwritten/solvers/matrix_free_operators.f90:35-62 */
/* nrg_diffusion: Ax = c div(k grad x) + x on a cell-centred x with k on the
 * faces (code/handwritten/solvers/matrix_free_operators.f90:148-174), the
 * energy equation's implicit operator; with MOONS_PREC_LAP_SF only (the
 * reference binds prec_lap_SF, code/handwritten/governing_equations/energy_extend.f90:129-130). */
#define MOONS_OP_NRG_DIFFUSION 3
#define MOONS_PREC_LAP_SF 1       /* code/handwritten/solvers/preconditioners/preconditioners.f90:55-71 */
/* FP32 geometric multigrid V-cycle inside an FP64 flexible PCG.  No compiled
 * counterpart in the reference (its V-cycle is dead code,
 * code_not_compiled/9.6.2017/MG/MG.f90:241-396); same solve/exit interface. */
#define MOONS_PREC_MG_FP32 2
/* Learned preconditioner: z = s N(r/vol * vol^(2/3) / s) with N the BF16
 * U-net bound by moons_pcg_set_unet, inside the FP64 flexible PCG.  No
 * reference counterpart. */
#define MOONS_PREC_UNET 3

/* Face-centred vector operators of the implicit diffusion solves:
 *   mom_diffusion: Ax = x + c lap(x)             (matrix_free_operators.f90:176-202)
 *   ind_diffusion: Ax = x + c curl(k curl(x))    (matrix_free_operators.f90:120-146)
 * with c = coeff_implicit_time_split (negative for momentum, where the
 * reference passes -theta dt/Re) and k = sigmaInv on the edges. */
#define MOONS_OP_MOM_DIFFUSION 1
#define MOONS_OP_IND_DIFFUSION 2

typedef struct moons_mesh_s *moons_mesh_t;
typedef struct moons_field_s *moons_field_t;
typedef struct moons_pcg_s *moons_pcg_t;
typedef struct moons_pcgvf_s *moons_pcgvf_t;
typedef struct moons_mhd_s *moons_mhd_t;
typedef struct moons_unet_s *moons_unet_t;
typedef struct moons_fft_s *moons_fft_t;
typedef struct moons_nrg_s *moons_nrg_t;
typedef struct moons_domain_s *moons_domain_t;

/* Exit information of one solve, the reference's ISP and norms after
 * solve_PCG_SF (code/handwritten/solvers/PCG/PCG_solver_algorithm.f90:37-107):
 * iterations = ISP%iter_per_call, res_norm0 = res_norm0%L2 (sqrt of the sum of
 * r0^2 over all cells), res_norm = last checked res_norm%L2, exit_loop =
 * ISP%exit_loop(1:3) as bits 0 (tol_rel), 1 (tol_abs), 2 (iter_max); bit 3
 * marks a MOONS_PREC_UNET solve that broke down or reached iter_max and was
 * redone with the Jacobi preconditioner (the figures are then the redo's). */
typedef struct {
  int iterations;
  int exit_loop;
  double res_norm0;
  double res_norm;
  double ms; /* device time of the solve (HIP events on the library stream) */
} moons_pcg_info;

/* ------------------------------------------------------------ library */
/* No reference counterpart (the reference has no device): selects the GPU and
 * creates the library stream.  Must precede every other call. */
int moons_init(int device);
/* Extra HIP streams (default priority) for replaying captured steps of
 * independent states concurrently; synchronize(NULL) waits for the library stream. */
/* In-situ kernel timer (no reference counterpart; instrumentation): while
 * armed, event pairs are recorded on the library stream around the launches
 * of the kernel class `which` (each launch of the FP64 PCG kernels; each
 * group of consecutive launches of the line half-sweeps, one line sweep), up
 * to max_launches pairs; stop returns the number of launches timed and their
 * mean duration (the group's time over its launches). */
#define MOONS_KTIMER_MG_LINE_FINE 1 /* zebra line half-sweeps of the finest multigrid level */
#define MOONS_KTIMER_PCG_MATVEC 2   /* FP64 PCG kernels of the PPE */
#define MOONS_KTIMER_PCG_UPDATE 3
#define MOONS_KTIMER_PCG_PUPDATE 4
#define MOONS_KTIMER_UNET_CONV16 5  /* the U-net's forward 16 -> 16 convolutions of its finest level */
int moons_ktimer_start(int which, int max_launches);
int moons_ktimer_stop(int *launches, double *avg_ms);
int moons_stream_create(void **out);
int moons_stream_destroy(void *hip_stream);
int moons_stream_synchronize(void *hip_stream);
int moons_finalize(void);
const char *moons_last_error(void);
int moons_synchronize(void);
/* The library's hipStream_t, as void*, for callers that time or order work on it. */
void *moons_stream(void);

/* ------------------------------------------------------------ 1D grids */
/* uniform_func + add_ghost_nodes: code/handwritten/grid/coordinate_distribution_funcs.f90:34-61,
 * code/handwritten/grid/coordinates_extend.f90:529-538.  hn has N+3 entries. */
int moons_nodes_uniform(double hmin, double hmax, int N, double *hn);
/* robertsBoth_func (Tannehill transformation 2, alpha = 1/2) + ghosts:
 * code/handwritten/grid/coordinate_distribution_funcs.f90:192-248,325-331,
 * code/handwritten/grid/grid_init.f90:60-68.  hn has N+3 entries. */
int moons_nodes_roberts_both(double hmin, double hmax, int N, double beta, double *hn);

/* ext_uniform_IO (code/handwritten/grid/grid_continue.f90:41-47,94-107,197-222):
 * hn (with ghosts, sn >= 4 entries) extended by N uniform cells at both ends,
 * each with its end's ghost spacing, continued from hmin down and hmax up;
 * out gets sn + 2N nodes, with ghosts. */
int moons_nodes_ext_uniform(const double *hn, int sn, int N, double *out);

/* The grid_continue extensions (code/handwritten/grid/grid_continue.f90) the
 * reference's mesh bank builds wall and vacuum meshes with, by kind:
 *   ext_uniform_IO (:41-47), ext_Roberts_near_IO (:49-57: prep R + app L),
 *   ext_Roberts_far_IO (:59-67: prep L + app R), ext_Roberts_B_IO (:69-77),
 *   ext_prep_uniform_IO / _Roberts_L / _R / _B / _C2F_IO (:94-193),
 *   ext_app_uniform_IO / _Roberts_L / _R / _B / _C2F_IO (:211-306).
 * A Roberts segment of N cells and length L continues the mesh from hmin
 * down (prep) or hmax up (app); its stretching beta is matched by Newton
 * iteration (coordinate_stretch_param_match.f90:176-233) so that its cell at
 * the matched end has the mesh's end spacing.  n_iter, max_stretch_ratio:
 * the mesh-quality parameters N_iter and max_mesh_stretch_ratio (1 and any
 * value for a manual mesh; with N_iter > 1 a segment whose neighbouring
 * spacings differ by more than the ratio is redone with more cells).
 * hn: nodes with ghosts, sn >= 4; out: room for cap nodes, *n_out gets the
 * count.  A match that does not converge is MOONS_ERR_ARG (the reference
 * stops at init on mesh_generation_error). */
#define MOONS_EXT_UNIFORM 0
#define MOONS_EXT_ROBERTS_NEAR 1
#define MOONS_EXT_ROBERTS_FAR 2
#define MOONS_EXT_ROBERTS_B 3
#define MOONS_EXT_PREP_UNIFORM 4
#define MOONS_EXT_PREP_ROBERTS_L 5
#define MOONS_EXT_PREP_ROBERTS_R 6
#define MOONS_EXT_PREP_ROBERTS_B 7
#define MOONS_EXT_PREP_ROBERTS_C2F 8
#define MOONS_EXT_APP_UNIFORM 9
#define MOONS_EXT_APP_ROBERTS_L 10
#define MOONS_EXT_APP_ROBERTS_R 11
#define MOONS_EXT_APP_ROBERTS_B 12
#define MOONS_EXT_APP_ROBERTS_C2F 13
int moons_nodes_ext(const double *hn, int sn, int kind, double L, int N, int n_iter, double max_stretch_ratio,
                    double *out, int cap, int *n_out);

/* ---------------------------------------------------------------- mesh */
/* init_c per direction (code/handwritten/grid/coordinates_extend.f90:78-173):
 * node coordinates including the two ghost nodes, sn = N+3 >= 5 entries. */
int moons_mesh_create(const double *hn_x, int sn_x, const double *hn_y, int sn_y,
                      const double *hn_z, int sn_z, moons_mesh_t *out);
int moons_mesh_destroy(moons_mesh_t m);
/* Copies one 1D coordinate array (which: 0 hn, 1 hc, 2 dhn, 3 dhc) of direction
 * dir (0..2) into out; returns its length through *n. */
int moons_mesh_coords(moons_mesh_t m, int dir, int which, double *out, int *n);

/* --------------------------------------------------------------- fields */
/* A scalar field (SF, code/generated_code/fields/SF.f90:43-51) of one block. */
int moons_field_create(moons_mesh_t m, int dl, moons_field_t *out);
int moons_field_destroy(moons_field_t f);
int moons_field_shape(moons_field_t f, int s[3]);
/* BCs of the field (code/handwritten/boundary_conditions): per-face type and
 * uniform value; all_neumann is the reference's SF%all_Neumann flag, which the
 * caller sets (code/handwritten/solvers/momentum/init_P_BCs.f90:26). */
int moons_field_set_bcs(moons_field_t f, const int type[6], const double val[6], int all_neumann);
/* init_BC_props (code/handwritten/block_field/block_field_extend.f90:395-400,
 * code/handwritten/boundary_conditions/boundary_conditions_extend.f90:504-511):
 * the wall conductance ratio c_w of each face, which Robin faces use, and the
 * Robin coefficients kept beside it (robin_coeff may be NULL). */
int moons_field_set_bc_props(moons_field_t f, const double c_w[6], const double robin_coeff[6]);
/* dst takes src's BCs (types, values, all_Neumann, c_w, Robin coefficients),
 * as the reference's init(SF, SF) copies them for a solver's fields
 * (init(PCG%p, x) and init(PCG%x_BC, x_BC), code/handwritten/PCG/PCG_solver_extend.f90:72-73);
 * the two fields share a data location. */
int moons_field_copy_bcs(moons_field_t dst, moons_field_t src);
/* Host <-> device copies of the whole ghost-padded array (n = element count). */
int moons_field_upload(moons_field_t f, const double *host, size_t n);
int moons_field_download(moons_field_t f, double *host, size_t n);
int moons_field_fill(moons_field_t f, double val);
int moons_field_copy(moons_field_t dst, moons_field_t src);
/* Device address of the field's array, for callers that share memory with
 * the library on its stream (e.g. torch tensors over the same buffers). */
int moons_field_device_ptr(moons_field_t f, void **dptr);

/* ---------------------------------------------------- multi-domain meshes */
/* mesh_domain (init_MD, code/handwritten/mesh_domain/mesh_domain_extend.f90:31-38):
 * where the interior mesh (e.g. the fluid) lies in the total mesh (e.g. the
 * induction mesh with its walls and vacuum), per direction the range of
 * coinciding interior cell centres and nodes (physical_sub_domain with
 * padding 1 and tolerance 1e-12,
 * code/handwritten/physical_domain/physical_sub_domain_extend.f90:36-45,
 * code/handwritten/sub_domain/overlap_extend.f90:204-278).  MOONS_ERR_ARG
 * when a direction has no common range. */
int moons_domain_create(moons_mesh_t interior, moons_mesh_t total, moons_domain_t *out);
int moons_domain_destroy(moons_domain_t d);
/* r = interior lo, interior hi, total lo, total hi (0-based, inclusive) of
 * direction dir (0..2): node = 0 the cell centres, 1 the nodes. */
int moons_domain_overlap(moons_domain_t d, int dir, int node, int r[4]);
/* embed / extract (code/handwritten/ops/ops_embedExtract.f90:169-278,
 * code/handwritten/grid_field/GF_embed_extract/GF_embed_extract.f90:14-60):
 * copy the overlap of two fields of one location from the interior field
 * into the total one, or back; every other value is left alone. */
int moons_embed(moons_field_t total, moons_field_t interior, moons_domain_t d);
int moons_extract(moons_field_t interior, moons_field_t total, moons_domain_t d);
/* set_sigma_inv_SF/VF (code/handwritten/solvers/induction/induction_aux.f90:162-194)
 * on one field of the total mesh: 1/sig_local_over_sig_f everywhere, 1 on
 * the overlap of the conducting mesh d->interior. */
int moons_set_sigma_inv(moons_field_t sig, moons_domain_t d, double sig_local_over_sig_f);

/* ------------------------------------------ restart primitives, raw export */
/* Text files in the reference's list-directed layout (reals with the fewest
 * digits that read back exactly; F form for 0.095 <= |x| < 9.5e14, E form
 * otherwise; records of at most 79 columns).  They read back, in the
 * reference or here, to the same doubles bit for bit, and the readers accept
 * any list-directed text.  Host arrays are column-major, ghost cells included.
 *
 * grid_field restart primitives (export_primitives_grid_field /
 * import_primitives_grid_field, code/generated_code/grid_field/grid_field.f90:123-156):
 * shape, "f     = ", the values, "s     = ", shape, "s_1D  = ", size. */
int moons_gf_write(const char *path, const double *f, const int s[3]);
/* Reads the shape into s_out; with f non-NULL also the values (n must equal
 * the size on file).  A file of an unallocated array (shape 0) leaves f alone. */
int moons_gf_read(const char *path, double *f, size_t n, int s_out[3]);
/* Tecplot raw export of one component on a one-block mesh (export_raw ->
 * export_3D_1C_steady/_unsteady, code/handwritten/IO/export_raw_processed.f90:44-60,
 * code/handwritten/IO/IO_export.f90:265-292; exp_3D_1C_GF,
 * code/handwritten/grid_field/GF_export_import/GF_export.f90:70-89): file
 * dir name <c|n|f_x|...|e_z> [_t=<t as F40.20>] .dat, header, zone, one
 * "x y z value" record per point, pad cells left out at each end.  hn_* are
 * the node coordinates with ghosts (sn = N+3); t NULL: the steady name. */
int moons_raw_write(const double *hn_x, int sn_x, const double *hn_y, int sn_y, const double *hn_z, int sn_z,
                    int dl, const double *f, const char *dir, const char *name, int pad, const double *t);
/* import_raw -> import_3D_1C (code/handwritten/IO/import_raw.f90:23-31,
 * code/handwritten/IO/IO_import.f90:20-33, imp_3D_1C_GF
 * code/handwritten/grid_field/GF_export_import/GF_import.f90:69-89): reads
 * the values of the padded range into f (shape s); pad cells keep theirs. */
int moons_raw_read(const int s[3], int dl, double *f, const char *dir, const char *name, int pad);
/* The same on device fields: the array is copied between HBM and the host
 * once per call (export: download then write; import: read then upload). */
int moons_field_export(moons_field_t f, const char *path);
int moons_field_import(moons_field_t f, const char *path);
int moons_export_raw(moons_field_t f, const char *dir, const char *name, int pad, const double *t);
int moons_import_raw(moons_field_t f, const char *dir, const char *name, int pad);

/* --------------------------------------------------------------- BCs */
/* apply_BCs / apply_BCs_implicit (code/handwritten/BCs/apply/apply_BCs.f90:42-64,
 * code/handwritten/BCs/direct/apply_BCs_faces_raw.f90:29-283,
 * code/handwritten/BCs/implicit/apply_BCs_faces_raw_implicit.f90:27-257).
 * Faces are applied in order 1..6. */
int moons_apply_bcs(moons_field_t u);
int moons_apply_bcs_implicit(moons_field_t u);

/* ------------------------------------------------------ discrete operators */
/* del%assign/add/subtract (code/handwritten/ops/ops_del.f90:236-261):
 * gen 1 assign, 2 add, 3 subtract; n = 1,2; dir = 0..2; pad = 0,1. */
int moons_del(moons_field_t dfdh, moons_field_t f, int n, int dir, int pad, int gen);
/* grad_SF (code/handwritten/ops/ops_discrete.f90:173-182) */
int moons_grad(moons_field_t gx, moons_field_t gy, moons_field_t gz, moons_field_t u);
/* div_SF (code/handwritten/ops/ops_discrete.f90:158-171) */
int moons_div(moons_field_t divU, moons_field_t u, moons_field_t v, moons_field_t w);
/* curl_SF component dir = 0..2 (code/handwritten/ops/ops_discrete.f90:184-202) */
int moons_curl(moons_field_t curlU, moons_field_t u, moons_field_t v, moons_field_t w, int dir);
/* Lap_uniform_SF (code/handwritten/solvers/matrix_free_operators.f90:51-62):
 * applies the implicit BCs to x (its ghosts change) then Ax = div(grad(x)),
 * one fused kernel for cell-centred x. */
int moons_lap_uniform_sf(moons_field_t Ax, moons_field_t x);
/* interpO2_SF along dir (code/handwritten/ops/ops_interp.f90:190-221): C->N
 * with theta weights and extrapolated end nodes
 * (code/handwritten/grid_field/GF_interp/GF_interp.f90:19-47, GF_extrap.f90:17-46)
 * or N->C half sums (GF_interp.f90:49-74), whichever the locations imply. */
int moons_interp(moons_field_t out, moons_field_t in, int dir);
/* edgeCrossFace_E (code/handwritten/ops/ops_advect.f90:202-218): E = U_E x
 * face2Edge_no_diag(B) on edges.  U_E: Uy@x, Uz@x, Ux@y, Uz@y, Ux@z, Uy@z;
 * tmp: six edge work fields Bx@y, Bx@z, By@x, By@z, Bz@x, Bz@y. */
int moons_edge_cross_face_E(moons_field_t E[3], moons_field_t U_E[6], moons_field_t B[3], moons_field_t tmp[6]);
/* advect_B (code/handwritten/ops/ops_advect.f90:186-200): the constrained-
 * transport source adv = curl(U_E x B) on faces; div(adv) = 0 discretely. */
int moons_advect_B(moons_field_t adv[3], moons_field_t U_E[6], moons_field_t B[3], moons_field_t temp_E_TF[6],
                   moons_field_t temp_E[3]);
/* volume_SF (code/handwritten/fields/SF_extend.f90:571-581) */
int moons_volume(moons_field_t vol);

/* ------------------------------------------------------------------ PCG */
/* init_PCG_SF (code/handwritten/PCG/PCG_solver_extend.f90:51-107): binds the
 * operator and preconditioner, computes vol and Minv; x carries the BCs used
 * by the solver (its copy `p` in the reference). */
int moons_pcg_sf_create(moons_mesh_t m, int op, int prec, moons_field_t x, moons_pcg_t *out);
int moons_pcg_destroy(moons_pcg_t pcg);
/* init(ISP, iter_max, tol_rel, tol_abs, n_skip_check_res, ...)
 * (code/handwritten/var_set/iter_solver_params_extend.f90:44-58). */
int moons_pcg_set_isp(moons_pcg_t pcg, int iter_max, double tol_rel, double tol_abs,
                      int n_skip_check_res);
/* The k of an nrg_diffusion solver: init(PCG%k, k) of init_PCG_SF
 * (code/handwritten/PCG/PCG_solver_extend.f90:77); k[0..2] are the x-, y- and
 * z-face fields of the mesh, copied.  Required before its first solve. */
int moons_pcg_sf_set_k(moons_pcg_t pcg, moons_field_t k[3]);
/* update_MFP_PCG_SF (code/handwritten/PCG/PCG_solver_extend.f90:269-279): the
 * coefficient c of an nrg_diffusion solver; prec_Lap_SF does not depend on it. */
int moons_pcg_sf_update_mfp(moons_pcg_t pcg, double coeff_implicit_time_split, int update_preconditioner);
/* solve_PCG_SF (code/handwritten/PCG/PCG_solver_extend.f90:182-193): x in/out,
 * b the right-hand side, both device fields of the solver's location.  The
 * whole loop runs on the device; info may be NULL. */
int moons_pcg_sf_solve(moons_pcg_t pcg, moons_field_t x, moons_field_t b, moons_pcg_info *info);
/* The PPE solve of moons_pcg_sf_solve (Lap_uniform_SF with prec_Lap_SF, FP64
 * Jacobi-PCG, ISP as moons_pcg_set_isp) without a solver object and without
 * allocating: no counterpart in the reference; it is for callers that cannot
 * allocate device memory (tight time-step loops, harnesses that forbid
 * hipMalloc).  Its work vectors, block
 * partial sums and control block live in one static device array of about
 * 15.7 MB per GPU (held until the process ends): x may have at most 2^19
 * padded cells (e.g. 78^3 interior), else MOONS_ERR_ARG.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (several host
 * threads, or streams other than the library's) race; each call clears what
 * it reads before use.  Not inside a graph capture (MOONS_ERR_STATE). */
int moons_pcg_sf_solve_noalloc(moons_field_t x, moons_field_t b, int iter_max, double tol_rel, double tol_abs,
                               int n_skip_check_res, moons_pcg_info *info);
/* clean_div_PCG (code/handwritten/solvers/clean_divergence.f90:37-59):
 *   temp_CC = scale*div(Xstar); solve(PCG, phi, temp_CC);
 *   X = Xstar - grad(phi)/scale; apply_BCs(X).
 * X, Xstar, temp_F are the three face components. */
int moons_clean_div(moons_pcg_t pcg, moons_field_t X[3], moons_field_t Xstar[3], moons_field_t phi,
                    double scale, moons_field_t temp_F[3], moons_field_t temp_CC,
                    moons_pcg_info *info);

/* ------------------------------------------ cosine-transform direct solver */

/* init(FFT, u, m, direction, var_name)
 * (code/handwritten/FFT_solver/FFT_solver_extend.f90:43-126): a direct solver
 * of the cell-centred Poisson problem in which every plane normal to dir
 * (1, 2, 3) is an independent 2D problem in the transverse directions, solved
 * by staggered DCT-II transforms (code/handwritten/ops/ops_dct.f90:36-70).
 * The transverse spacings must be uniform (MOONS_ERR_ARG otherwise).  The
 * mode factors are the reference's, 0.5 ha hb / (cos + cos - 2), with its
 * dir = 1 guard of the zero mode by the indices (i, j) (the modes (0, q) of
 * the first x-plane keep their value), unless moons_fft_set_coefficients
 * selects the exact eigenvalues. */
int moons_fft_create(moons_mesh_t m, int dir, moons_fft_t *fft);
#define MOONS_FFT_COEF_REFERENCE 0 /* init_FFT's factors (FFT_solver_extend.f90:62-112) */
#define MOONS_FFT_COEF_EXACT 1     /* 1/eigenvalue of the 5-point Neumann Laplacian for any spacings */
/* Selects the mode factors (no reference counterpart: the reference has only
 * its own; the exact eigenvalues differ from them for unequal transverse
 * spacings and in the first x-plane for dir = 1). */
int moons_fft_set_coefficients(moons_fft_t fft, int kind);
int moons_fft_destroy(moons_fft_t fft);
/* solve(FFT, u, f, m, compute_norms) (FFT_solver_extend.f90:128-137;
 * solve_FFT_SF_x/y/z, code/handwritten/FFT_solver/FFT_algorithm.f90:27-210):
 * u's interior from the transforms of f's interior, the zero transverse mode
 * of every plane set to 0 ("pin down pressure"), then apply_BCs(u).  u and f
 * are cell-centred fields of the solver's mesh. */
int moons_fft_solve(moons_fft_t fft, moons_field_t u, moons_field_t f);
/* clean_div_FFT (code/handwritten/solvers/clean_divergence.f90:61-83):
 *   temp_CC = scale*div(Xstar); solve(FFT, phi, temp_CC);
 *   X = Xstar - grad(phi)/scale; apply_BCs(X). */
int moons_clean_div_fft(moons_fft_t fft, moons_field_t X[3], moons_field_t Xstar[3], moons_field_t phi,
                        double scale, moons_field_t temp_F[3], moons_field_t temp_CC);

/* Multigrid parameters of a MOONS_PREC_MG_FP32 solver (rebuilds the level
 * hierarchy): at most max_levels levels, nu1/nu2 smoothing sweeps before/after
 * the coarse correction, coarse_sweeps sweeps on the coarsest level.  A sweep
 * is red-black Gauss-Seidel, or on a level of more than 32^3 cells whose
 * non-periodic y or z cell widths vary by more than 1.5x (lines of at most 128
 * cells), zebra line Gauss-Seidel along those directions: y-lines of both
 * colours, then z-lines (reversed after the correction).  Defaults 16, 2, 2,
 * 16.  No reference counterpart (see MOONS_PREC_MG_FP32). */
int moons_pcg_set_mg(moons_pcg_t pcg, int max_levels, int nu1, int nu2, int coarse_sweeps);
/* Number of multigrid levels and the coarsest level's cells per direction. */
int moons_pcg_mg_info(moons_pcg_t pcg, int *levels, int n_coarsest[3]);
/* Applies the solver's preconditioner to a volume-weighted residual r:
 * z = Minv r (prec_Lap_SF) or z = V(r/vol) (one FP32 V-cycle).  The
 * preconditioner interface of the reference is preconditioner_SF
 * (code/handwritten/abstract_interfaces/preconditioner_interfaces.f90:13-23),
 * which builds Minv; here the application itself is exposed. */
int moons_pcg_apply_prec(moons_pcg_t pcg, moons_field_t r, moons_field_t z);

/* Times one application of each kernel of the solver's iteration (matvec +
 * (p,Ap); x/r update + (r,r),(z,r); p update) with HIP events on the library
 * stream, averaged over reps runs on the solver's current vectors.  ms[3]
 * receives the per-kernel averages.  No reference counterpart (profiling). */
int moons_pcg_time_kernels(moons_pcg_t pcg, int reps, double ms[3]);

/* Times the multigrid kernels of a MOONS_PREC_MG_FP32 solver with HIP events
 * on the library stream, reps launches back to back each: ms[0] one
 * half-sweep (one colour) of the finest level, ms[1] restriction of its
 * residual, ms[2] prolongation into it, ms[3] the coarse end, ms[4] a whole
 * V-cycle, ms[5] one fused whole sweep (both colours, one launch) of the
 * finest level, ms[6]/ms[7] a half-sweep / fused sweep of level 1 and
 * ms[8]/ms[9] of level 2 (0 where the level does not exist).  On a level
 * smoothed by zebra lines the half-sweep entry is one line half-sweep (one
 * colour of the first line direction) and the fused-sweep entry is 0.
 * Overwrites the multigrid arrays.  No reference counterpart (profiling). */
int moons_pcg_time_mg(moons_pcg_t pcg, int reps, double ms[10]);

/* --------------------------------------------- vector (VF) Jacobi-PCG */
/* init_PCG_VF (code/handwritten/PCG/PCG_solver_extend.f90:117-180) for
 * op = MOONS_OP_MOM_DIFFUSION with prec_mom_VF, or MOONS_OP_IND_DIFFUSION
 * with prec_ind_VF (code/handwritten/solvers/preconditioners/preconditioners.f90:104-129).
 * x: the three face components (their BCs are the solver's); k: sigmaInv on
 * the x,y,z edges (copied; NULL for momentum).  The coefficient starts at 0:
 * set it with moons_pcg_vf_update_mfp before the first solve. */
int moons_pcg_vf_create(moons_mesh_t m, int op, moons_field_t x[3], moons_field_t k[3], moons_pcgvf_t *out);
int moons_pcg_vf_destroy(moons_pcgvf_t pcg);
/* ISP of the VF solver (iter_max, tol_rel, tol_abs, n_skip_check_res). */
int moons_pcg_vf_set_isp(moons_pcgvf_t pcg, int iter_max, double tol_rel, double tol_abs,
                         int n_skip_check_res);
/* update_MFP_PCG_VF (code/handwritten/PCG/PCG_solver_extend.f90:281-291):
 * sets the operator coefficient; rebuilds Minv when update_preconditioner. */
int moons_pcg_vf_update_mfp(moons_pcgvf_t pcg, double coeff_implicit_time_split, int update_preconditioner);
/* solve_PCG_VF (code/handwritten/PCG/PCG_solver_extend.f90:195-206 ->
 * code/handwritten/solvers/PCG/PCG_solver_algorithm.f90:109-179): x in/out,
 * b the right-hand side, three face components each. */
int moons_pcg_vf_solve(moons_pcgvf_t pcg, moons_field_t x[3], moons_field_t b[3], moons_pcg_info *info);

/* ------------------------------------------------------- the time step */
/* Parameters of one MOONS MHD time step (solve_MOONS,
 * code/handwritten/MOONS/MOONS_solver.f90:47-144, one RK stage):
 *   momentum sources F = -u.grad(u) + (1-alpha_U)/Re lap(U) + J x B
 *     (add_all_momentum_sources, solvers/momentum/add_all_momentum_sources.f90:22-86);
 *   alpha_U = 0: Euler_time_no_diff_AB2_sources (method 4), alpha_U = 1:
 *     Euler_time_AB2_sources (method 6) with A = I - dt/Re lap
 *     (solvers/time_marching_methods.f90:117-150,266-296);
 *   induction (when induction != 0): CT source curl(U_E x Btot),
 *     backward-Euler curl-curl diffusion with sigma = 1,
 *     phi cleaning; J = curl(B)/Rem, J x B scaled by N = Ha^2/Re
 *     (induction_extend.f90:475-536, induction_sources.f90:28-84).
 * Solver settings index 0..3 = U, p, B, phi (iter_max, tol_rel, tol_abs,
 * n_skip_check_res); prec_p selects the PPE preconditioner
 * (MOONS_PREC_LAP_SF or MOONS_PREC_MG_FP32).
 * mpg_dir 1..3 (0: off) drives the flow by a mean pressure gradient: before
 * the sources, every value of U's mpg_dir component gains mpg_scale*dt, then
 * U's BCs are applied (compute_add_MPG,
 * solvers/momentum/momentum_sources.f90:97-108, called at
 * add_all_momentum_sources.f90:31-33).
 * finite_Rem != 0: Btot = B0 + B after each induction solve, else Btot = B0
 * (compute_Btot, governing_equations/induction_extend.f90:527-536); Btot
 * enters the CT source and J x B. */
typedef struct {
  double Re, Ha, Rem, dt;
  double alpha_U;
  int induction;
  int prec_p;
  double B0[3];
  int iter_max[4];
  double tol_rel[4];
  double tol_abs[4];
  int n_skip[4];
  int mpg_dir;
  double mpg_scale;
  int finite_Rem;
  /* the phi-cleaning preconditioner: 0 or MOONS_PREC_LAP_SF (the reference's
   * Jacobi), or MOONS_PREC_MG_FP32 (the mixed-precision multigrid, as for the PPE) */
  int prec_phi;
} moons_mhd_params;

typedef struct {
  int iterations[4]; /* of the last step's U, p, B, phi solves */
  double ms;         /* device time of the call */
} moons_mhd_info;

#define MOONS_MHD_F 1       /* momentum sources F (3 faces) */
#define MOONS_MHD_FNM1 2    /* previous step's F */
#define MOONS_MHD_USTAR 3   /* intermediate velocity */
#define MOONS_MHD_U_E 4     /* U on edges, 6 entries: Uy@x, Uz@x, Ux@y, Uz@y, Ux@z, Uy@z */
#define MOONS_MHD_J 5       /* current density on edges */
#define MOONS_MHD_JCROSSB 6 /* N J x B on faces */
#define MOONS_MHD_FB 7      /* induction sources */
#define MOONS_MHD_PPE_RHS 8 /* right-hand side of the last pressure solve (1 CC entry) */

/* Creates the step state over the caller's U (faces, with BCs), p (CC, with
 * BCs), and for induction B (faces, with BCs) and phi (CC, with BCs); these
 * stay the caller's and are advanced in place.  F and Fnm1 start at zero
 * (the reference's start-up), U_E, J and J x B are computed from U and B. */
int moons_mhd_create(moons_mesh_t m, const moons_mhd_params *params, moons_field_t U[3], moons_field_t p,
                     moons_field_t B[3], moons_field_t phi, moons_mhd_t *out);
int moons_mhd_destroy(moons_mhd_t s);
/* The same step with the induction on its own mesh (the fluid mesh m
 * extended by walls and vacuum, MOONS' multi-domain MHD): B, phi and
 * sigma_inv live on m_ind, d is the fluid mesh inside m_ind
 * (moons_domain_create(m, m_ind)).  Each step embeds the fluid's edge
 * velocity into m_ind before the CT source (embedVelocity_E,
 * code/handwritten/solvers/induction/add_all_induction_sources.f90:35-36,
 * code/handwritten/solvers/induction/induction_aux.f90:135-143; zero outside
 * the fluid) and extracts J x B to the fluid faces (extract_add_JCrossB,
 * code/handwritten/solvers/momentum/momentum_sources.f90:111-121).
 * sigma_inv: the three edge fields of 1/sigma on m_ind (copied; NULL: 1),
 * e.g. from moons_set_sigma_inv.  m_ind == m with d == NULL is
 * moons_mhd_create. */
int moons_mhd_create_md(moons_mesh_t m, moons_mesh_t m_ind, moons_domain_t d, const moons_mhd_params *params,
                        moons_field_t U[3], moons_field_t p, moons_field_t B[3], moons_field_t phi,
                        moons_field_t sigma_inv[3], moons_mhd_t *out);
/* Advances n_steps time steps; info (may be NULL) gets the last step's solver
 * iterations and the device time of the call. */
int moons_mhd_step(moons_mhd_t s, int n_steps, moons_mhd_info *info);
/* Captures one time step into a HIP graph (no reference counterpart).  Inside
 * the graph every PCG solve runs exactly its iter_max iterations, each
 * iteration returning at once on the device after convergence, and nothing is
 * read back to the host; the implicit-diffusion preconditioners (which
 * depend on dt only) are built before the capture.  Not with the learned PPE
 * solver. */
int moons_mhd_graph_capture(moons_mhd_t s);
/* Replays the captured step n_steps times on hip_stream (NULL: the library
 * stream), asynchronously. */
int moons_mhd_graph_launch(moons_mhd_t s, int n_steps, void *hip_stream);
/* Handles of internal fields (MOONS_MHD_*), 3 or 6 entries, for inspection. */
int moons_mhd_field(moons_mhd_t s, int which, moons_field_t out[6]);
/* Buoyancy and gravity in the momentum sources, after J x B
 * (add_all_momentum_sources.f90:70-78):
 *   buoyancy (compute_add_buoyancy, momentum_sources.f90:133-150):
 *     F_d += cellCenter2Face_d(T * g_d) * buoyancy_scale   (Gr/Re^2)
 *   gravity (compute_add_gravity, momentum_sources.f90:152-167):
 *     F_d += cellCenter2Face_d(g_d) * gravity_scale        (1/Fr^2)
 * T is the energy's temperature (cell-centred, on the fluid mesh) and g the
 * three cell-centred components of the gravity field (nrg%gravity, e.g. the
 * reference's uniform_gravity_field: 1 in one component everywhere, ghosts
 * included, 0 in the others, init_gravity_field.f90:41-50).  The faces are
 * taken onto the fluid mesh as J x B is (extractFace with the energy on the
 * fluid mesh: every face but the non-periodic ghost planes, which are 0).
 * The fields stay the caller's and are read at every step (T as the energy
 * step left it).  add_buoyancy / add_gravity 0 switch a term off; T may be
 * NULL when buoyancy is off; both off unbinds. */
int moons_mhd_set_buoyancy(moons_mhd_t s, moons_field_t T, moons_field_t gravity[3], int add_buoyancy,
                           double buoyancy_scale, int add_gravity, double gravity_scale);

/* ---------------------------------------------------- energy equation */
/* One energy step (solve_energy after add_all_energy_sources,
 * code/handwritten/governing_equations/energy_extend.f90:236-262,
 * code/handwritten/solvers/energy/add_all_energy_sources.f90:21-80), the
 * first equation of each time step of solve_MOONS
 * (code/handwritten/MOONS/MOONS_solver.f90:80-83):
 *   Fnm1 = F; F = 0;
 *   advection (add_advection, code/handwritten/solvers/energy/energy_sources.f90:26-38):
 *     F += advection_scale * (-div(cellCenter2Face(T) * U));
 *   diffusion (add_diffusion, energy_sources.f90:40-48):
 *     F += diffusion_scale * lapUniformCoeff(T);
 *   KE diffusion, viscous dissipation, Joule and volumetric heating (below);
 *   method 1: T += dt F; method 2: T += dt AB2(F, Fnm1); method 3: Tnm1 = T,
 *     solve (I + c div(k grad)) T = T + dt AB2(F, Fnm1) with c = dt coeff_implicit
 *     (code/handwritten/solvers/time_marching_methods_SF.f90:71-98,135-170);
 *   method 4 (O2_BDF_time_AB2_sources_SF, :42-69): rhs = 2/3 dt AB2(F, Fnm1)
 *     + 4/3 T - 1/3 Tnm1, Tnm1 = T, c = dt 2/3 coeff_implicit;
 *   method 5 (Euler_time_RK_sources_SF, :100-129): rhs = dt (gamma F + zeta Fnm1
 *     + alpha L) + T, Tnm1 = T, c = dt beta coeff_implicit.
 * U_CC = face2CellCenter(U) (momentum_extend.f90:370) is formed from U at
 * each step's sources.
 * coeff_implicit is the MFP's (-alpha/Pe; code/handwritten/var_set/var_set_extend.f90:82).
 * The reference's presets scale advection by -1 and diffusion by 1/Pe
 * (e.g. code/handwritten/sim_params/sim_params_default_MHD_Shercliff/sim_params_init_flow_config.f90:71-72). */
typedef struct {
  double dt;
  int method;
  double coeff_implicit;
  int advection;
  double advection_scale;
  int diffusion;
  double diffusion_scale;
  int iter_max;
  double tol_rel, tol_abs;
  int n_skip;
  /* add_KE_diffusion (energy_sources.f90:50-63): F += scale lap(|U_CC|^2) */
  int KE_diffusion;
  double KE_diffusion_scale;
  /* add_viscous_dissipation (:65-76): F += scale grad(U_CC) : grad(U_CC) */
  int viscous_dissipation;
  double viscous_dissipation_scale;
  /* add_joule_heating (:78-93): F += scale |edge2CellCenter(J)|^2 sigmaInv_CC
   * (J and sigmaInv_CC bound by moons_nrg_set_joule) */
  int joule_heating;
  double joule_heating_scale;
  /* add_volumetric_heating_Nuclear (:95-113): F += scale q(y), q = scale 2/(1 - e^-2) e^-(y_c + 1) */
  int volumetric_heating;
  double volumetric_heating_scale;
  /* 1: RK inactive (one stage, coefficients one); 3: the 3-stage RK of init_RKP
   * (code/handwritten/var_set/RK_params_extend.f90:29-62) is active: a step runs
   * three stages, diffusion goes to L, and method 5 reads the stage's coefficients.
   * 0 is taken as 1. */
  int rk_stages;
} moons_nrg_params;

typedef struct {
  int iterations; /* of the last step's T solve (methods 3-5) */
  double ms;      /* device time of the call */
} moons_nrg_info;

#define MOONS_NRG_F 1    /* energy sources F */
#define MOONS_NRG_FNM1 2 /* previous step's F */
#define MOONS_NRG_TNM1 3 /* T before the last implicit solve */
#define MOONS_NRG_L 4    /* the RK stage's explicit diffusion */

/* Creates the energy state over the caller's T (cell-centred, with BCs) and
 * face velocity U (read at each step: the momentum state's U); k are the
 * face fields bound to the implicit operator (copied).  The reference binds
 * a zero-filled temporary there (energy_extend.f90:103,129-130), which makes
 * its implicit operator the identity; pass zero fields to reproduce that, or
 * the face conductivity (cellCenter2Face of k_cc, energy_extend.f90:121-124)
 * for I + c div(k grad).  F and Fnm1 start at zero, Tnm1 = T. */
int moons_nrg_create(moons_mesh_t m, const moons_nrg_params *params, moons_field_t T, moons_field_t U[3],
                     moons_field_t k[3], moons_nrg_t *out);
int moons_nrg_destroy(moons_nrg_t s);
/* Binds the induction's current density J (three edge fields: x edges,
 * y edges, z edges) and sigmaInv_CC (cell-centred) that Joule heating reads
 * (ind%J, ind%sigmaInv_CC of add_all_energy_sources.f90:59-62), on the
 * energy mesh.  Read at each step. */
int moons_nrg_set_joule(moons_nrg_t s, moons_field_t J[3], moons_field_t sigma_inv_cc);
/* Advances n_steps energy steps with the current U (each step runs the RK
 * stages when rk_stages = 3). */
int moons_nrg_step(moons_nrg_t s, int n_steps, moons_nrg_info *info);
/* Handle of an internal field (MOONS_NRG_*). */
int moons_nrg_field(moons_nrg_t s, int which, moons_field_t *out);

/* ------------------------------------------- learned preconditioner */
/* A 3-level 3D U-net (channels 16/32/64, 3x3x3 convolutions, ReLU, average
 * pooling, nearest upsampling with skip concatenation) with BF16 weights and
 * activations and FP32 accumulation on the MFMA units; FP32 master weights
 * trained with Adam by the library's own forward and backward kernels.
 * Input: 4 channels per cell (f vol^(2/3) / s and log(dh_d / vol^(1/3)));
 * output: 1 channel.  None of these has a reference counterpart. */
int moons_unet_create(unsigned long long seed, moons_unet_t *out);
int moons_unet_destroy(moons_unet_t net);
size_t moons_unet_num_params(moons_unet_t net);
/* FP32 master weights, layer by layer: W[tap][c_in][c_out] then bias[c_out]. */
int moons_unet_get_weights(moons_unet_t net, float *host, size_t n);
int moons_unet_set_weights(moons_unet_t net, const float *host, size_t n);
/* in4: 4 channel arrays of (n0+2)(n1+2)(n2+2) FP32 (column-major, halo
 * ignored); out: one such array (interior written).  n0, n1, n2 multiples of 4. */
int moons_unet_forward_host(moons_unet_t net, int n0, int n1, int n2, const float *in4, float *out);
/* loss = mean over the interior of (N(in4) - target)^2 and d loss / d param
 * (same order as the weights). */
int moons_unet_grad_host(moons_unet_t net, int n0, int n1, int n2, const float *in4, const float *target,
                         double *loss, float *grad);
/* Adam training on npair samples (in4: npair x 4 arrays, target: npair
 * arrays), batch samples per step, epochs passes in a seeded order;
 * loss_hist[epoch] = mean loss.  The data set is copied to the device once. */
int moons_unet_train(moons_unet_t net, int n0, int n1, int n2, int npair, const float *in4, const float *target,
                     int epochs, int batch, double lr, unsigned long long seed, double *loss_hist);
/* D = A B for 16x16 A and B rounded to BF16, through one MFMA with the lane
 * layout the U-net kernels use (a self-check of that layout). */
int moons_mfma_check(const float *A, const float *B, float *D);
/* D = A B for a 16x32 A and a 32x16 B rounded to BF16, through one
 * v_mfma_f32_16x16x32_bf16 with the lane layout of the finest-level
 * convolutions (a self-check of that layout).  No reference counterpart. */
int moons_mfma_check32(const float *A, const float *B, float *D);
/* Learned + mixed: inside a MOONS_PREC_MG_FP32 solver's V-cycle, the network
 * gives the initial guess of multigrid level `level` (0 = finest; a level
 * launched over the grid, with cell counts that are multiples of 4): before
 * that level's pre-smoothing x = c N(f), c = (N, f) / (N, L N) (0 when not
 * positive).  net = NULL removes it.  No reference counterpart. */
int moons_pcg_set_mg_unet(moons_pcg_t pcg, moons_unet_t net, int level);
/* The learned initial guess in the first `cycles` V-cycles of each solve only
 * (0, the default: in every V-cycle).  No reference counterpart. */
int moons_pcg_set_mg_unet_cycles(moons_pcg_t pcg, int cycles);
/* Inside a MOONS_PREC_MG_FP32 solver's V-cycle, the part from multigrid
 * level `level` down (its smoothing, the coarser levels and the coarse end)
 * is run `cycles` times per visit, each from the previous result: a
 * gamma-cycle at one level.  Many cycles make that level's correction
 * nearly exact, the best any initial guess of that level (such as the
 * learned one) can give.  level < 0 or cycles = 1: the plain V-cycle.  The
 * level must be launched over the grid (not in the single-workgroup coarse
 * end; MOONS_ERR_ARG otherwise).  Holds until moons_pcg_set_mg rebuilds the
 * hierarchy.  No reference counterpart. */
int moons_pcg_set_mg_gamma(moons_pcg_t pcg, int level, int cycles);
/* Enables (default) or disables the Jacobi redo of failed MOONS_PREC_UNET solves. */
int moons_pcg_set_fallback(moons_pcg_t pcg, int enable);
/* Binds the network to a MOONS_PREC_UNET solver (required before its first solve). */
int moons_pcg_set_unet(moons_pcg_t pcg, moons_unet_t net);

#ifdef __cplusplus
}
#endif
#endif
