// step.hip -- one MOONS MHD time step on the device: momentum (sources, AB2
// time march with backward-Euler or explicit diffusion, pressure projection)
// and low magnetic Reynolds number induction (CT source curl(U_E x Btot),
// implicit curl-curl diffusion, phi cleaning, J and J x B).
//
// Sequence of code/handwritten/MOONS/MOONS_solver.f90:47-144 for one RK stage:
//   add_all_momentum_sources   solvers/momentum/add_all_momentum_sources.f90:22-86
//   solve_momentum             governing_equations/momentum_extend.f90:326-374
//     method 4 Euler_time_no_diff_AB2_sources  solvers/time_marching_methods.f90:266-296
//     method 6 Euler_time_AB2_sources          solvers/time_marching_methods.f90:117-150
//   add_all_induction_sources  solvers/induction/add_all_induction_sources.f90:24-60
//   solve_induction (method 6) governing_equations/induction_extend.f90:475-525
//   compute_J, compute_Btot (finite_Rem = F: Btot = B0), compute_JCrossB
//                              solvers/induction/induction_aux.f90:80-88,
//                              solvers/induction/induction_sources.f90:70-84
// Every operator is a device kernel of this library; element-wise chains are
// fused into one kernel each, keeping the reference's operation order (AB2 =
// 1.5 F, then + (-0.5) Fnm1, + L, * dt; Dirichlet wall planes zeroed; + X), so
// each value is rounded exactly as in the sequential passes.  The fluid and
// induction meshes are the same single block, so embedding and extraction
// are identities.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace moons;

namespace moons {

constexpr int kST = 256;

// y = y + x*s
__global__ void __launch_bounds__(kST) k_axpy(double *y, const double *x, double s, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t < n) y[t] = y[t] + x[t] * s;
}

// y = y + s over every value, ghosts included (add of a scalar to a grid field)
__global__ void __launch_bounds__(kST) k_adds(double *y, double s, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t < n) y[t] = y[t] + s;
}

// y = a*b
__global__ void __launch_bounds__(kST) k_mul3(double *y, const double *a, const double *b, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t < n) y[t] = a[t] * b[t];
}

// y = y + a*b (the product rounded first, as multiply then add)
__global__ void __launch_bounds__(kST) k_addmul(double *y, const double *a, const double *b, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t < n) {
    const double p = a[t] * b[t];
    y[t] = y[t] + p;
  }
}

// out = ((((F*1.5) + Fnm1*(-0.5)) + L) * dt), zero on the Dirichlet wall
// planes of X (normal direction d, planes 1 and s_d-2 by mask), then + X.
// has_L = 0 skips the + L (method 4).
__global__ void __launch_bounds__(kST) k_ab2(double *out, const double *F, const double *Fnm1, const double *L,
                                            const double *X, double dt, int s0, int s1, int s2, int d, int wall_mask,
                                            int has_L) {
  const size_t n = (size_t)s0 * s1 * s2;
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t >= n) return;
  double v = F[t] * 1.5;
  v = v + Fnm1[t] * (-0.5);
  if (has_L) v = v + L[t];
  v = v * dt;
  const int i = (int)(t % s0), j = (int)((t / s0) % s1), k = (int)(t / ((size_t)s0 * s1));
  const int ii = d == 0 ? i : (d == 1 ? j : k);
  const int sd = d == 0 ? s0 : (d == 1 ? s1 : s2);
  if (((wall_mask & 1) && ii == 1) || ((wall_mask & 2) && ii == sd - 2)) v = 0.0;
  out[t] = v + X[t];
}

// acc = x*x (first) or acc + x*x: a left-to-right sum of squares over
// successive launches, as square() then add() of the reference's fields
__global__ void __launch_bounds__(kST) k_sqacc(double *acc, const double *x, int first, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t >= n) return;
  const double q = x[t] * x[t];
  acc[t] = first ? q : acc[t] + q;
}

// O2_BDF_time_AB2_sources_SF's right-hand side (time_marching_methods_SF.f90:61-65):
// ((F*1.5 + Fnm1*(-0.5)) * (2/3 dt) + X*(4/3)) + Xnm1*(-1/3)
__global__ void __launch_bounds__(kST) k_bdf2_rhs(double *out, const double *F, const double *Fnm1, const double *X,
                                                 const double *Xnm1, double c23dt, double c43, double cm13, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t >= n) return;
  double v = F[t] * 1.5;
  v = v + Fnm1[t] * (-0.5);
  v = v * c23dt;
  v = v + X[t] * c43;
  out[t] = v + Xnm1[t] * cm13;
}

// Euler_time_RK_sources_SF's right-hand side (time_marching_methods_SF.f90:121-125):
// ((F*(dt gamma) + Fnm1*(dt zeta)) + L*(dt alpha)) + X
__global__ void __launch_bounds__(kST) k_rk_rhs(double *out, const double *F, const double *Fnm1, const double *L,
                                               const double *X, double cg, double cz, double ca, size_t n) {
  const size_t t = (size_t)blockIdx.x * kST + threadIdx.x;
  if (t >= n) return;
  double v = F[t] * cg;
  v = v + Fnm1[t] * cz;
  v = v + L[t] * ca;
  out[t] = v + X[t];
}

static inline unsigned sblk(size_t n) { return (unsigned)((n + kST - 1) / kST); }

} // namespace moons

struct moons_mhd_s {
  moons_mesh_s *m = nullptr;
  moons_mesh_s *mi = nullptr;          // mesh of the induction fields (m, or the extended mesh)
  const moons_domain_s *dom = nullptr; // the fluid mesh inside mi (multi-domain only)
  moons_mhd_params P{};
  moons_field_s *U[3] = {}, *p = nullptr, *B[3] = {}, *phi = nullptr; // caller's
  moons_field_s *Ustar[3] = {}, *F[3] = {}, *Fnm1[3] = {}, *L[3] = {}, *T1[3] = {}, *T2[3] = {}, *T3[3] = {};
  moons_field_s *UE[6] = {}, *cc = nullptr, *tcc = nullptr;
  moons_field_s *Bstar[3] = {}, *FB[3] = {}, *FBnm1[3] = {}, *LB[3] = {}, *B0[3] = {}, *J[3] = {}, *JxB[3] = {};
  moons_field_s *Btot[3] = {}; // B0 (low Rem) or a field of its own, B0 + B (finite Rem)
  moons_field_s *sig[3] = {}, *E[3] = {}, *ETF[6] = {}, *A6[6] = {}, *B6[6] = {};
  // induction-mesh temporaries: U_E embedded (multi-domain only; else UE),
  // face temporaries (multi-domain only; else T1, T2) and the phi right-hand side
  moons_field_s *UEi[6] = {}, *T1i[3] = {}, *T2i[3] = {}, *cci = nullptr, *tcci = nullptr;
  bool own_i = false; // UEi, T1i, T2i, cci are fields of their own
  moons_pcg_t pcg_p = nullptr, pcg_phi = nullptr;
  moons_pcgvf_t pcg_U = nullptr, pcg_B = nullptr;
  // buoyancy and gravity sources (caller's fields: T cell-centred, g three cell-centred components)
  const moons_field_s *Tb = nullptr, *grav[3] = {};
  int add_buoy = 0, add_grav = 0;
  double buoy_scale = 0.0, grav_scale = 0.0;
  int iters[4] = {0, 0, 0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

const int kEdgeDL[3] = {MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z};
// (component, location) of the six off-diagonal tensor entries: U_y@x, U_z@x,
// U_x@y, U_z@y, U_x@z, U_y@z (edges for U_E, faces for the J x B operands)
const int kNoDiag[6][2] = {{1, 0}, {2, 0}, {0, 1}, {2, 1}, {0, 2}, {1, 2}};

int fcreate(moons_mesh_s *m, int dl, moons_field_s **f) { return moons_field_create(m, dl, f); }

void mhd_free(moons_mhd_s *S) {
  if (S->exec) (void)hipGraphExecDestroy(S->exec);
  if (S->graph) (void)hipGraphDestroy(S->graph);
  S->exec = nullptr;
  S->graph = nullptr;
  if (S->P.finite_Rem)
    for (int d = 0; d < 3; ++d)
      if (S->Btot[d]) moons_field_destroy(S->Btot[d]);
  moons_field_s **groups3[] = {S->Ustar, S->F, S->Fnm1, S->L, S->T1, S->T2, S->T3, S->Bstar, S->FB,
                               S->FBnm1, S->LB, S->B0, S->J, S->JxB, S->sig, S->E};
  for (auto g : groups3)
    for (int d = 0; d < 3; ++d)
      if (g[d]) moons_field_destroy(g[d]);
  moons_field_s **groups6[] = {S->UE, S->ETF, S->A6, S->B6};
  for (auto g : groups6)
    for (int q = 0; q < 6; ++q)
      if (g[q]) moons_field_destroy(g[q]);
  if (S->own_i) {
    for (int q = 0; q < 6; ++q)
      if (S->UEi[q]) moons_field_destroy(S->UEi[q]);
    for (int d = 0; d < 3; ++d) {
      if (S->T1i[d]) moons_field_destroy(S->T1i[d]);
      if (S->T2i[d]) moons_field_destroy(S->T2i[d]);
    }
    if (S->cci) moons_field_destroy(S->cci);
  }
  if (S->tcci) moons_field_destroy(S->tcci);
  if (S->cc) moons_field_destroy(S->cc);
  if (S->tcc) moons_field_destroy(S->tcc);
  if (S->pcg_p) moons_pcg_destroy(S->pcg_p);
  if (S->pcg_phi) moons_pcg_destroy(S->pcg_phi);
  if (S->pcg_U) moons_pcg_vf_destroy(S->pcg_U);
  if (S->pcg_B) moons_pcg_vf_destroy(S->pcg_B);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  delete S;
}

int axpy(moons_field_s *y, const moons_field_s *x, double s) {
  hipLaunchKernelGGL(k_axpy, dim3(sblk(y->n)), dim3(kST), 0, stream(), y->d, x->d, s, y->n);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int mul3(moons_field_s *y, const moons_field_s *a, const moons_field_s *b) {
  hipLaunchKernelGGL(k_mul3, dim3(sblk(y->n)), dim3(kST), 0, stream(), y->d, a->d, b->d, y->n);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int addmul(moons_field_s *y, const moons_field_s *a, const moons_field_s *b) {
  hipLaunchKernelGGL(k_addmul, dim3(sblk(y->n)), dim3(kST), 0, stream(), y->d, a->d, b->d, y->n);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// face i -> CC (along i) -> face j (along j)
// face i -> CC (cc, a cell-centred temporary of u's mesh) -> face j
int face2face(moons_field_s *cc, moons_field_s *out, const moons_field_s *u, int i, int j) {
  int st;
  if ((st = launch_interp(cc, u, i))) return st;
  return launch_interp(out, cc, j);
}

// U_E = face2Edge_no_diag(U) (ops/ops_interp.f90:531-544)
int face2edge_no_diag(moons_mhd_s *S) {
  int st;
  for (int q = 0; q < 6; ++q) {
    const int i = kNoDiag[q][0], j = kNoDiag[q][1];
    if ((st = launch_interp(S->UE[q], S->U[i], 3 - i - j))) return st;
  }
  return MOONS_OK;
}

// T1 = u_j d/dx_j u_i (advect_U_convection1, ops/ops_advect.f90:71-127)
int convection(moons_mhd_s *S, moons_field_s *div[3]) {
  int st;
  for (int d = 0; d < 3; ++d) {
    if ((st = launch_interp(S->cc, S->U[d], d))) return st;
    if ((st = launch_del(S->T2[d], S->cc, 1, d, 0, 1))) return st;
    if ((st = mul3(div[d], S->T2[d], S->U[d]))) return st;
  }
  // (U_E entry, derivative direction, velocity component), reference order
  const int terms[3][2][3] = {{{4, 1, 1}, {2, 2, 2}}, {{5, 0, 0}, {0, 2, 2}}, {{3, 0, 0}, {1, 1, 1}}};
  for (int d = 0; d < 3; ++d)
    for (int t = 0; t < 2; ++t) {
      const int q = terms[d][t][0], ddir = terms[d][t][1], vel = terms[d][t][2];
      if ((st = launch_del(S->T2[d], S->UE[q], 1, ddir, 0, 1))) return st;
      if ((st = face2face(S->cc, S->T3[d], S->U[vel], vel, d))) return st;
      if ((st = addmul(div[d], S->T2[d], S->T3[d]))) return st;
    }
  return MOONS_OK;
}

int wall_mask(const moons_field_s *X, int d) {
  if (!X->has_bcs) return 0;
  return (X->bct[2 * d] == MOONS_BC_DIRICHLET ? 1 : 0) | (X->bct[2 * d + 1] == MOONS_BC_DIRICHLET ? 2 : 0);
}

int ab2_rhs(moons_field_s *out[3], moons_field_s *F[3], moons_field_s *Fnm1[3], moons_field_s *L[3],
            moons_field_s *X[3], double dt, bool has_L) {
  for (int d = 0; d < 3; ++d) {
    const moons_field_s *x = X[d];
    hipLaunchKernelGGL(k_ab2, dim3(sblk(x->n)), dim3(kST), 0, stream(), out[d]->d, F[d]->d, Fnm1[d]->d,
                       has_L ? L[d]->d : F[d]->d, x->d, dt, x->s[0], x->s[1], x->s[2], d, wall_mask(x, d),
                       has_L ? 1 : 0);
    MOONS_LAUNCH_CHECK();
  }
  return MOONS_OK;
}

// J = curl(B)/Rem; Btot = B0 (+ B at finite Rem); JxB = N edge2Face(J) x face2Face(Btot)
int post_induction(moons_mhd_s *S) {
  int st;
  for (int d = 0; d < 3; ++d) {
    if ((st = moons_curl(S->J[d], S->B[0], S->B[1], S->B[2], d))) return st;
    if ((st = launch_scale(S->J[d], 1.0 / S->P.Rem))) return st;
  }
  if (S->P.finite_Rem)
    for (int d = 0; d < 3; ++d) {
      if ((st = moons_field_copy(S->Btot[d], S->B0[d]))) return st;
      if ((st = axpy(S->Btot[d], S->B[d], 1.0))) return st; // B0 + B*1 = B0 + B exactly
    }
  for (int q = 0; q < 6; ++q) {
    const int i = kNoDiag[q][0], j = kNoDiag[q][1];
    if ((st = launch_interp(S->A6[q], S->J[i], 3 - i - j))) return st;  // edge2Face
    if ((st = face2face(S->cci, S->B6[q], S->Btot[i], i, j))) return st;   // face2Face(Btot)
  }
  // (J x B)_x = J_y Bz - J_z By on x-faces; y negated form; z = J_x By - J_y Bx
  if ((st = launch_cross(S->JxB[0], S->A6[0], S->B6[1], S->A6[1], S->B6[0], 1))) return st;
  if ((st = launch_cross(S->JxB[1], S->A6[2], S->B6[3], S->A6[3], S->B6[2], -1))) return st;
  if ((st = launch_cross(S->JxB[2], S->A6[4], S->B6[5], S->A6[5], S->B6[4], 1))) return st;
  const double N = S->P.Ha * S->P.Ha / S->P.Re;
  for (int d = 0; d < 3; ++d)
    if ((st = launch_scale(S->JxB[d], N))) return st;
  return MOONS_OK;
}

int momentum(moons_mhd_s *S) {
  int st;
  moons_pcg_info info;
  if (S->P.mpg_dir) {
    // compute_add_MPG (solvers/momentum/momentum_sources.f90:97-108): the
    // component gains scale*dt everywhere, then apply_BCs(U)
    moons_field_s *u = S->U[S->P.mpg_dir - 1];
    hipLaunchKernelGGL(k_adds, dim3(sblk(u->n)), dim3(kST), 0, stream(), u->d, S->P.mpg_scale * S->P.dt, u->n);
    MOONS_LAUNCH_CHECK();
    if ((st = launch_apply_bcs3(S->U, false))) return st;
  }
  for (int d = 0; d < 3; ++d) {
    if ((st = moons_field_copy(S->Fnm1[d], S->F[d]))) return st;
    if ((st = moons_field_fill(S->L[d], 0.0))) return st;
    if ((st = moons_field_fill(S->F[d], 0.0))) return st;
  }
  if ((st = convection(S, S->T1))) return st;
  for (int d = 0; d < 3; ++d)
    if ((st = axpy(S->F[d], S->T1[d], -1.0))) return st;
  const double ce = (1.0 / S->P.Re) * (1.0 - S->P.alpha_U);
  for (int d = 0; d < 3; ++d) {
    if ((st = launch_del(S->T1[d], S->U[d], 2, 0, 0, 1))) return st;
    if ((st = launch_del(S->T1[d], S->U[d], 2, 1, 0, 2))) return st;
    if ((st = launch_del(S->T1[d], S->U[d], 2, 2, 0, 2))) return st;
    if ((st = axpy(S->F[d], S->T1[d], ce))) return st;
  }
  if (S->P.induction)
    for (int d = 0; d < 3; ++d) {
      // extract_add_JCrossB: the overlap of the induction mesh's J x B onto the
      // fluid faces (every fluid face but the ghosts), ghosts zeroed
      if ((st = S->dom ? launch_extract(S->T1[d], S->JxB[d], S->dom) : moons_field_copy(S->T1[d], S->JxB[d])))
        return st;
      if ((st = launch_assign_ghost_xperiodic(S->T1[d], 0.0))) return st;
      if ((st = axpy(S->F[d], S->T1[d], 1.0))) return st;
    }
  if (S->add_buoy)
    for (int d = 0; d < 3; ++d) {
      // compute_add_buoyancy: temp_CC = T, times gravity, cellCenter2Face,
      // extractFace (as J x B), F += buoyancy * scale
      if ((st = mul3(S->tcc, S->Tb, S->grav[d]))) return st;
      if ((st = launch_interp(S->T1[d], S->tcc, d))) return st;
      if ((st = launch_assign_ghost_xperiodic(S->T1[d], 0.0))) return st;
      if ((st = axpy(S->F[d], S->T1[d], S->buoy_scale))) return st;
    }
  if (S->add_grav)
    for (int d = 0; d < 3; ++d) {
      // compute_add_gravity: cellCenter2Face(g), extractFace, F += gravity * scale
      if ((st = launch_interp(S->T1[d], S->grav[d], d))) return st;
      if ((st = launch_assign_ghost_xperiodic(S->T1[d], 0.0))) return st;
      if ((st = axpy(S->F[d], S->T1[d], S->grav_scale))) return st;
    }
  double scale;
  if (S->P.alpha_U == 0.0) {
    if ((st = ab2_rhs(S->Ustar, S->F, S->Fnm1, S->L, S->U, S->P.dt, false))) return st;
    scale = 1.0;
  } else {
    if ((st = ab2_rhs(S->T1, S->F, S->Fnm1, S->L, S->U, S->P.dt, true))) return st;
    for (int d = 0; d < 3; ++d)
      if ((st = moons_field_copy(S->Ustar[d], S->U[d]))) return st;
    if ((st = moons_pcg_vf_update_mfp(S->pcg_U, S->P.dt * 1.0 * (-(1.0 / S->P.Re) * S->P.alpha_U), 1))) return st;
    if ((st = moons_pcg_vf_solve(S->pcg_U, S->Ustar, S->T1, &info))) return st;
    S->iters[0] = info.iterations;
    scale = 1.0 / S->P.dt;
  }
  if ((st = moons_clean_div(S->pcg_p, S->U, S->Ustar, S->p, scale, S->T2, S->tcc, &info))) return st;
  S->iters[1] = info.iterations;
  return face2edge_no_diag(S);
}

int induction(moons_mhd_s *S) {
  int st;
  moons_pcg_info info;
  for (int d = 0; d < 3; ++d) {
    if ((st = moons_field_copy(S->FBnm1[d], S->FB[d]))) return st;
    if ((st = moons_field_fill(S->LB[d], 0.0))) return st;
    if ((st = moons_field_fill(S->FB[d], 0.0))) return st;
  }
  if (S->dom) // embedVelocity_E: the fluid's edge velocity into the induction mesh
    for (int q = 0; q < 6; ++q)
      if ((st = launch_embed(S->UEi[q], S->UE[q], S->dom))) return st;
  if ((st = moons_advect_B(S->T1i, S->UEi, S->Btot, S->ETF, S->E))) return st;
  for (int d = 0; d < 3; ++d)
    if ((st = axpy(S->FB[d], S->T1i[d], 1.0))) return st;
  for (int d = 0; d < 3; ++d)
    if ((st = mul3(S->E[d], S->J[d], S->sig[d]))) return st;
  for (int d = 0; d < 3; ++d) {
    if ((st = moons_curl(S->T2i[d], S->E[0], S->E[1], S->E[2], d))) return st;
    if ((st = axpy(S->FB[d], S->T2i[d], -0.0))) return st; // -beta, beta = 1 - alpha = 0
  }
  if ((st = ab2_rhs(S->T1i, S->FB, S->FBnm1, S->LB, S->B, S->P.dt, true))) return st;
  for (int d = 0; d < 3; ++d)
    if ((st = moons_field_copy(S->Bstar[d], S->B[d]))) return st;
  if ((st = moons_pcg_vf_update_mfp(S->pcg_B, S->P.dt * 1.0 * (1.0 / S->P.Rem), 1))) return st;
  if ((st = moons_pcg_vf_solve(S->pcg_B, S->Bstar, S->T1i, &info))) return st;
  S->iters[2] = info.iterations;
  if ((st = moons_clean_div(S->pcg_phi, S->B, S->Bstar, S->phi, 1.0 / S->P.dt, S->T2i, S->tcci, &info))) return st;
  S->iters[3] = info.iterations;
  return post_induction(S);
}

} // namespace

extern "C" {

int moons_mhd_create(moons_mesh_t m, const moons_mhd_params *P, moons_field_t U[3], moons_field_t p,
                     moons_field_t B[3], moons_field_t phi, moons_mhd_t *out) {
  return moons_mhd_create_md(m, m, nullptr, P, U, p, B, phi, nullptr, out);
}

int moons_mhd_create_md(moons_mesh_t m, moons_mesh_t mi, moons_domain_t dom, const moons_mhd_params *P,
                        moons_field_t U[3], moons_field_t p, moons_field_t B[3], moons_field_t phi,
                        moons_field_t sigma_inv[3], moons_mhd_t *out) {
  if (!stream()) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!m || !P || !U || !p || !out) return fail(MOONS_ERR_ARG, "moons_mhd_create: null argument");
  if (!mi) mi = m;
  if (!(P->dt > 0.0) || !(P->Re > 0.0) || (P->alpha_U != 0.0 && P->alpha_U != 1.0))
    return fail(MOONS_ERR_ARG, "moons_mhd_create: need dt > 0, Re > 0, alpha_U 0 (method 4) or 1 (method 6)");
  if (P->mpg_dir < 0 || P->mpg_dir > 3) return fail(MOONS_ERR_ARG, "moons_mhd_create: mpg_dir must be 0..3");
  if (P->prec_phi != 0 && P->prec_phi != MOONS_PREC_LAP_SF && P->prec_phi != MOONS_PREC_MG_FP32)
    return fail(MOONS_ERR_ARG, "moons_mhd_create: prec_phi must be MOONS_PREC_LAP_SF or MOONS_PREC_MG_FP32");
  if (P->induction && (!B || !phi || !(P->Rem > 0.0)))
    return fail(MOONS_ERR_ARG, "moons_mhd_create: induction needs B, phi and Rem > 0");
  if (mi != m && (!dom || dom->mi != m || dom->mt != mi || !P->induction))
    return fail(MOONS_ERR_ARG, "moons_mhd_create_md: a separate induction mesh needs induction and the domain of "
                               "the fluid mesh inside it");
  if (mi == m && dom) return fail(MOONS_ERR_ARG, "moons_mhd_create_md: a domain needs a separate induction mesh");
  for (int d = 0; d < 3; ++d)
    if (!U[d] || U[d]->dl != (1 << d) || U[d]->m != m ||
        (P->induction && (!B[d] || B[d]->dl != (1 << d) || B[d]->m != mi)))
      return fail(MOONS_ERR_ARG, "moons_mhd_create: U and B must be face components on the fluid and induction meshes");
  if (p->dl != MOONS_DL_CC || (P->induction && (phi->dl != MOONS_DL_CC || phi->m != mi)))
    return fail(MOONS_ERR_ARG, "moons_mhd_create: p and phi must be cell-centred on their meshes");
  if (sigma_inv)
    for (int d = 0; d < 3; ++d)
      if (!sigma_inv[d] || sigma_inv[d]->m != mi || sigma_inv[d]->dl != kEdgeDL[d])
        return fail(MOONS_ERR_ARG, "moons_mhd_create_md: sigma_inv must be the three edge fields of the induction mesh");
  moons_mhd_s *S = new moons_mhd_s();
  S->m = m;
  S->mi = mi;
  S->dom = dom;
  S->P = *P;
  S->p = p;
  S->phi = phi;
  int st = MOONS_OK;
  for (int d = 0; d < 3 && !st; ++d) {
    S->U[d] = U[d];
    if (P->induction) S->B[d] = B[d];
    moons_field_s **vecs[] = {S->Ustar, S->F, S->Fnm1, S->L, S->T1, S->T2, S->T3};
    for (auto v : vecs)
      if (!st) st = fcreate(m, 1 << d, &v[d]);
    if (!st) copy_bcs_of(S->Ustar[d], U[d]);
  }
  for (int q = 0; q < 6 && !st; ++q) st = fcreate(m, kEdgeDL[kNoDiag[q][1]], &S->UE[q]);
  if (!st) st = fcreate(m, MOONS_DL_CC, &S->cc);
  if (!st) st = fcreate(m, MOONS_DL_CC, &S->tcc);
  if (!st) {
    for (int d = 0; d < 3; ++d) st = st ? st : moons_field_fill(S->F[d], 0.0);
  }
  if (!st) st = moons_pcg_sf_create(m, MOONS_OP_LAP_UNIFORM_SF, P->prec_p, p, &S->pcg_p);
  if (!st) st = moons_pcg_set_isp(S->pcg_p, P->iter_max[1], P->tol_rel[1], P->tol_abs[1], P->n_skip[1]);
  if (!st && P->alpha_U != 0.0) {
    st = moons_pcg_vf_create(m, MOONS_OP_MOM_DIFFUSION, S->Ustar, nullptr, &S->pcg_U);
    if (!st) st = moons_pcg_vf_set_isp(S->pcg_U, P->iter_max[0], P->tol_rel[0], P->tol_abs[0], P->n_skip[0]);
  }
  if (!st && P->induction) {
    for (int d = 0; d < 3 && !st; ++d) {
      moons_field_s **vecs[] = {S->Bstar, S->FB, S->FBnm1, S->LB, S->B0, S->JxB};
      for (auto v : vecs)
        if (!st) st = fcreate(mi, 1 << d, &v[d]);
      if (!st) copy_bcs_of(S->Bstar[d], B[d]);
      if (!st) st = moons_field_fill(S->B0[d], P->B0[d]);
      if (P->finite_Rem) {
        if (!st) st = fcreate(mi, 1 << d, &S->Btot[d]);
      } else {
        S->Btot[d] = S->B0[d];
      }
      if (!st) st = moons_field_fill(S->FB[d], 0.0);
      if (!st) st = fcreate(mi, kEdgeDL[d], &S->J[d]);
      if (!st) st = fcreate(mi, kEdgeDL[d], &S->sig[d]);
      if (!st) st = sigma_inv ? moons_field_copy(S->sig[d], sigma_inv[d]) : moons_field_fill(S->sig[d], 1.0);
      if (!st) st = fcreate(mi, kEdgeDL[d], &S->E[d]);
    }
    // temp_E_TF of advect_B: Bx@y, Bx@z, By@x, By@z, Bz@x, Bz@y edges
    const int etf_edge[6] = {1, 2, 0, 2, 0, 1};
    for (int q = 0; q < 6 && !st; ++q) st = fcreate(mi, kEdgeDL[etf_edge[q]], &S->ETF[q]);
    for (int q = 0; q < 6 && !st; ++q) {
      st = fcreate(mi, 1 << kNoDiag[q][1], &S->A6[q]);
      if (!st) st = fcreate(mi, 1 << kNoDiag[q][1], &S->B6[q]);
    }
    if (!st) st = fcreate(mi, MOONS_DL_CC, &S->tcci);
    S->own_i = mi != m;
    for (int q = 0; q < 6 && !st; ++q) {
      if (!S->own_i) S->UEi[q] = S->UE[q];
      else if (!(st = fcreate(mi, kEdgeDL[kNoDiag[q][1]], &S->UEi[q]))) st = moons_field_fill(S->UEi[q], 0.0);
    }
    if (!S->own_i) S->cci = S->cc;
    else if (!st) st = fcreate(mi, MOONS_DL_CC, &S->cci);
    for (int d = 0; d < 3 && !st; ++d) {
      if (!S->own_i) {
        S->T1i[d] = S->T1[d];
        S->T2i[d] = S->T2[d];
      } else {
        st = fcreate(mi, 1 << d, &S->T1i[d]);
        if (!st) st = fcreate(mi, 1 << d, &S->T2i[d]);
      }
    }
    if (!st) st = moons_pcg_vf_create(mi, MOONS_OP_IND_DIFFUSION, S->Bstar, S->sig, &S->pcg_B);
    if (!st) st = moons_pcg_vf_set_isp(S->pcg_B, P->iter_max[2], P->tol_rel[2], P->tol_abs[2], P->n_skip[2]);
    if (!st) st = moons_pcg_sf_create(mi, MOONS_OP_LAP_UNIFORM_SF, P->prec_phi ? P->prec_phi : MOONS_PREC_LAP_SF, phi,
                                      &S->pcg_phi);
    if (!st) st = moons_pcg_set_isp(S->pcg_phi, P->iter_max[3], P->tol_rel[3], P->tol_abs[3], P->n_skip[3]);
    if (!st) st = post_induction(S);
  }
  if (!st) st = face2edge_no_diag(S);
  if (!st && (hipEventCreate(&S->ev0) != hipSuccess || hipEventCreate(&S->ev1) != hipSuccess))
    st = fail(MOONS_ERR_HIP, "moons_mhd_create: event creation failed");
  if (st) {
    mhd_free(S);
    return st;
  }
  *out = S;
  return MOONS_OK;
}

int moons_mhd_destroy(moons_mhd_t S) {
  if (!S) return MOONS_OK;
  // a captured step may have been replayed on a caller's stream: wait for the device
  if (S->exec) (void)hipDeviceSynchronize();
  else if (stream()) (void)hipStreamSynchronize(stream());
  mhd_free(S);
  return MOONS_OK;
}

int moons_mhd_step(moons_mhd_t S, int n_steps, moons_mhd_info *info) {
  if (!S || n_steps < 0) return fail(MOONS_ERR_ARG, "moons_mhd_step: bad argument");
  MOONS_HIP(hipEventRecord(S->ev0, stream()));
  int st;
  for (int n = 0; n < n_steps; ++n) {
    if ((st = momentum(S))) return st;
    if (S->P.induction && (st = induction(S))) return st;
  }
  MOONS_HIP(hipEventRecord(S->ev1, stream()));
  MOONS_HIP(hipEventSynchronize(S->ev1));
  if (info) {
    for (int q = 0; q < 4; ++q) info->iterations[q] = S->iters[q];
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, S->ev0, S->ev1);
    info->ms = ms;
  }
  return MOONS_OK;
}

int moons_mhd_graph_capture(moons_mhd_t S) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_mhd_graph_capture: null state");
  if (S->P.prec_p == MOONS_PREC_UNET) return fail(MOONS_ERR_STATE, "moons_mhd_graph_capture: not with the learned PPE solver");
  int st;
  // the implicit-diffusion preconditioners depend on dt only: built here, outside the graph
  if (S->pcg_U && (st = moons_pcg_vf_update_mfp(S->pcg_U, S->P.dt * 1.0 * (-(1.0 / S->P.Re) * S->P.alpha_U), 1)))
    return st;
  if (S->pcg_B && (st = moons_pcg_vf_update_mfp(S->pcg_B, S->P.dt * 1.0 * (1.0 / S->P.Rem), 1))) return st;
  MOONS_HIP(hipStreamSynchronize(stream()));
  if (S->exec) (void)hipGraphExecDestroy(S->exec);
  if (S->graph) (void)hipGraphDestroy(S->graph);
  S->exec = nullptr;
  S->graph = nullptr;
  MOONS_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
  set_sync_free(true);
  st = momentum(S);
  if (!st && S->P.induction) st = induction(S);
  set_sync_free(false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream(), &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return fail(MOONS_ERR_HIP, std::string("moons_mhd_graph_capture: ") + hipGetErrorString(e));
  S->graph = g;
  MOONS_HIP(hipGraphInstantiate(&S->exec, S->graph, nullptr, nullptr, 0));
  return MOONS_OK;
}

int moons_mhd_graph_launch(moons_mhd_t S, int n_steps, void *hip_stream) {
  if (!S || n_steps < 0) return fail(MOONS_ERR_ARG, "moons_mhd_graph_launch: bad argument");
  if (!S->exec) return fail(MOONS_ERR_STATE, "moons_mhd_graph_launch: capture the step first");
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : stream();
  for (int n = 0; n < n_steps; ++n) MOONS_HIP(hipGraphLaunch(S->exec, st));
  return MOONS_OK;
}

int moons_mhd_set_buoyancy(moons_mhd_t S, moons_field_t T, moons_field_t gravity[3], int add_buoyancy,
                           double buoyancy_scale, int add_gravity, double gravity_scale) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_mhd_set_buoyancy: null state");
  if (S->exec) return fail(MOONS_ERR_STATE, "moons_mhd_set_buoyancy: set the sources before capturing the step");
  if (!add_buoyancy && !add_gravity) {
    S->add_buoy = S->add_grav = 0;
    S->Tb = nullptr;
    for (int d = 0; d < 3; ++d) S->grav[d] = nullptr;
    return MOONS_OK;
  }
  if (!gravity) return fail(MOONS_ERR_ARG, "moons_mhd_set_buoyancy: the gravity field is needed");
  for (int d = 0; d < 3; ++d)
    if (!gravity[d] || gravity[d]->m != S->m || gravity[d]->dl != MOONS_DL_CC)
      return fail(MOONS_ERR_ARG, "moons_mhd_set_buoyancy: gravity must be three cell-centred fields on the fluid mesh");
  if (add_buoyancy && (!T || T->m != S->m || T->dl != MOONS_DL_CC))
    return fail(MOONS_ERR_ARG, "moons_mhd_set_buoyancy: T must be a cell-centred field on the fluid mesh");
  S->Tb = add_buoyancy ? T : nullptr;
  for (int d = 0; d < 3; ++d) S->grav[d] = gravity[d];
  S->add_buoy = add_buoyancy ? 1 : 0;
  S->add_grav = add_gravity ? 1 : 0;
  S->buoy_scale = buoyancy_scale;
  S->grav_scale = gravity_scale;
  return MOONS_OK;
}

int moons_mhd_field(moons_mhd_t S, int which, moons_field_t out[6]) {
  if (!S || !out) return fail(MOONS_ERR_ARG, "moons_mhd_field: null argument");
  moons_field_s **src = nullptr;
  int n = 3;
  switch (which) {
    case MOONS_MHD_F: src = S->F; break;
    case MOONS_MHD_FNM1: src = S->Fnm1; break;
    case MOONS_MHD_USTAR: src = S->Ustar; break;
    case MOONS_MHD_U_E: src = S->UE; n = 6; break;
    case MOONS_MHD_J: src = S->J; break;
    case MOONS_MHD_JCROSSB: src = S->JxB; break;
    case MOONS_MHD_FB: src = S->FB; break;
    case MOONS_MHD_PPE_RHS: out[0] = S->tcc; return MOONS_OK;
    default: return fail(MOONS_ERR_ARG, "moons_mhd_field: unknown field");
  }
  if (!src[0]) return fail(MOONS_ERR_STATE, "moons_mhd_field: field not present (no induction)");
  for (int q = 0; q < n; ++q) out[q] = src[q];
  return MOONS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// The energy equation: sources (advection, diffusion) and one of the three
// time-marching methods of solve_energy, composed from the kernels above.
// ---------------------------------------------------------------------------

struct moons_nrg_s {
  moons_mesh_s *m = nullptr;
  moons_nrg_params P{};
  moons_field_s *T = nullptr, *U[3] = {}; // caller's
  moons_field_s *J[3] = {}, *sig = nullptr; // caller's (Joule heating)
  moons_field_s *F = nullptr, *Fnm1 = nullptr, *L = nullptr, *Tnm1 = nullptr, *rhs = nullptr, *tcc = nullptr,
                *tcc2 = nullptr, *tF[3] = {}, *UCC[3] = {}, *qv = nullptr;
  moons_pcg_t pcg = nullptr;
  int iters = 0;
  int stage = 0; // RK stage (0-based) of the running step
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

// init_RKP (code/handwritten/var_set/RK_params_extend.f90:55-58): gamma, zeta,
// alpha of the three stages; beta = alpha
constexpr double kRKGamma[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
constexpr double kRKZeta[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
constexpr double kRKAlpha[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};

bool rk_active(const moons_nrg_s *S) { return S->P.rk_stages == 3; }

void nrg_free(moons_nrg_s *S) {
  moons_field_s *fs[] = {S->F,     S->Fnm1,   S->L,      S->Tnm1,   S->rhs,    S->tcc,    S->tcc2,
                         S->tF[0], S->tF[1],  S->tF[2],  S->UCC[0], S->UCC[1], S->UCC[2], S->qv};
  for (auto f : fs)
    if (f) moons_field_destroy(f);
  if (S->pcg) moons_pcg_destroy(S->pcg);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  delete S;
}

int sqacc(moons_field_s *acc, const moons_field_s *x, bool first) {
  hipLaunchKernelGGL(k_sqacc, dim3(sblk(acc->n)), dim3(kST), 0, stream(), acc->d, x->d, first ? 1 : 0, acc->n);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// lapUniformCoeff_SF (ops/ops_discrete.f90:101-110): collocated second
// derivatives along x (assign), y and z (add)
int lap_uniform(moons_field_s *out, const moons_field_s *f) {
  int st;
  if ((st = launch_del(out, f, 2, 0, 0, 1))) return st;
  if ((st = launch_del(out, f, 2, 1, 0, 2))) return st;
  return launch_del(out, f, 2, 2, 0, 2);
}

// add_all_energy_sources (solvers/energy/add_all_energy_sources.f90:21-80)
int nrg_sources(moons_nrg_s *S) {
  int st;
  if ((st = moons_field_copy(S->Fnm1, S->F))) return st;
  if (rk_active(S) && (st = moons_field_fill(S->L, 0.0))) return st;
  if ((st = moons_field_fill(S->F, 0.0))) return st;
  if (S->P.advection) {
    // temp_F = cellCenter2Face(T); temp_F = temp_F * U; temp_CC = div(temp_F);
    // temp_CC = -temp_CC; F = F + temp_CC * scale (the negation is exact, so
    // it folds into the scale: (-t) s == t (-s) bit for bit)
    for (int d = 0; d < 3; ++d) {
      if ((st = launch_interp(S->tF[d], S->T, d))) return st;
      hipLaunchKernelGGL(k_mul3, dim3(sblk(S->tF[d]->n)), dim3(kST), 0, stream(), S->tF[d]->d, S->tF[d]->d,
                         S->U[d]->d, S->tF[d]->n);
      MOONS_LAUNCH_CHECK();
    }
    if ((st = moons_div(S->tcc, S->tF[0], S->tF[1], S->tF[2]))) return st;
    if ((st = axpy(S->F, S->tcc, -S->P.advection_scale))) return st;
  }
  if (S->P.diffusion) {
    if ((st = lap_uniform(S->tcc, S->T))) return st;
    if ((st = axpy(rk_active(S) ? S->L : S->F, S->tcc, S->P.diffusion_scale))) return st;
  }
  if (S->P.KE_diffusion || S->P.viscous_dissipation) // U_CC = face2CellCenter(U)
    for (int d = 0; d < 3; ++d)
      if ((st = launch_interp(S->UCC[d], S->U[d], d))) return st;
  if (S->P.KE_diffusion) {
    // temp_CC1 = Ux^2 + Uy^2 + Uz^2; temp_CC2 = lap(temp_CC1); F += temp_CC2 * scale
    for (int d = 0; d < 3; ++d)
      if ((st = sqacc(S->tcc, S->UCC[d], d == 0))) return st;
    if ((st = lap_uniform(S->tcc2, S->tcc))) return st;
    if ((st = axpy(S->F, S->tcc2, S->P.KE_diffusion_scale))) return st;
  }
  if (S->P.viscous_dissipation) {
    // grad_TF (ops_discrete.f90:313-321): G[c][e] = d U_c / d x_e collocated
    // on the cell centres; summed squared in add_SF_TF's order (TF_extend.f90:248-255):
    // G[x][x], G[y][x], G[z][x], G[x][y], ..., G[z][z]
    for (int e = 0; e < 3; ++e)
      for (int c = 0; c < 3; ++c) {
        if ((st = launch_del(S->tcc2, S->UCC[c], 1, e, 0, 1))) return st;
        if ((st = sqacc(S->tcc, S->tcc2, e == 0 && c == 0))) return st;
      }
    if ((st = axpy(S->F, S->tcc, S->P.viscous_dissipation_scale))) return st;
  }
  if (S->P.joule_heating) {
    if (!S->J[0] || !S->sig) return fail(MOONS_ERR_STATE, "moons_nrg_step: Joule heating needs moons_nrg_set_joule");
    // edge2CellCenter_VF (ops_interp.f90:688-696): component d goes edge ->
    // face fd = d+1 mod 3 along the third direction, then -> cell centre along fd
    for (int d = 0; d < 3; ++d) {
      const int fd = (d + 1) % 3, od = 3 - d - fd;
      if ((st = launch_interp(S->tF[fd], S->J[d], od))) return st;
      if ((st = launch_interp(S->tcc2, S->tF[fd], fd))) return st;
      if ((st = sqacc(S->tcc, S->tcc2, d == 0))) return st;
    }
    if ((st = mul3(S->tcc, S->tcc, S->sig))) return st; // multiply(temp_CC, sigmaInv_CC)
    if ((st = axpy(S->F, S->tcc, S->P.joule_heating_scale))) return st;
  }
  if (S->P.volumetric_heating && (st = axpy(S->F, S->qv, S->P.volumetric_heating_scale))) return st;
  return MOONS_OK;
}

// solve_energy (governing_equations/energy_extend.f90:236-259), methods 1-5.
// A cell-centred field has no wall planes, so assign_wall_Dirichlet is a no-op.
int nrg_march(moons_nrg_s *S) {
  const moons_field_s *T = S->T;
  const unsigned nb = sblk(T->n);
  const double dt = S->P.dt;
  if (S->P.method == 1) { // T = T + F*dt
    hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(kST), 0, stream(), S->T->d, S->F->d, dt, T->n);
    MOONS_LAUNCH_CHECK();
    return MOONS_OK;
  }
  if (S->P.method == 2) { // T = ((F*1.5 + Fnm1*(-0.5)) * dt) + T
    hipLaunchKernelGGL(k_ab2, dim3(nb), dim3(kST), 0, stream(), S->T->d, S->F->d, S->Fnm1->d, (const double *)nullptr,
                       S->T->d, dt, T->s[0], T->s[1], T->s[2], 0, 0, 0);
    MOONS_LAUNCH_CHECK();
    return MOONS_OK;
  }
  double c;
  if (S->P.method == 3) {
    hipLaunchKernelGGL(k_ab2, dim3(nb), dim3(kST), 0, stream(), S->rhs->d, S->F->d, S->Fnm1->d,
                       (const double *)nullptr, S->T->d, dt, T->s[0], T->s[1], T->s[2], 0, 0, 0);
    c = dt * 1.0 * S->P.coeff_implicit;
  } else if (S->P.method == 4) {
    const double two_thirds = 2.0 / 3.0, four_thirds = 4.0 / 3.0, neg_one_third = -1.0 / 3.0;
    hipLaunchKernelGGL(k_bdf2_rhs, dim3(nb), dim3(kST), 0, stream(), S->rhs->d, S->F->d, S->Fnm1->d, S->T->d,
                       S->Tnm1->d, two_thirds * dt, four_thirds, neg_one_third, T->n);
    c = dt * two_thirds * S->P.coeff_implicit;
  } else {
    const bool rk = rk_active(S);
    const double g = rk ? kRKGamma[S->stage] : 1.0, z = rk ? kRKZeta[S->stage] : 1.0,
                 a = rk ? kRKAlpha[S->stage] : 1.0;
    hipLaunchKernelGGL(k_rk_rhs, dim3(nb), dim3(kST), 0, stream(), S->rhs->d, S->F->d, S->Fnm1->d, S->L->d, S->T->d,
                       dt * g, dt * z, dt * a, T->n);
    c = dt * a * S->P.coeff_implicit; // beta = alpha
  }
  MOONS_LAUNCH_CHECK();
  int st;
  if ((st = moons_field_copy(S->Tnm1, S->T))) return st;
  if ((st = moons_pcg_sf_update_mfp(S->pcg, c, 1))) return st;
  moons_pcg_info info;
  if ((st = moons_pcg_sf_solve(S->pcg, S->T, S->rhs, &info))) return st;
  S->iters = info.iterations;
  return MOONS_OK;
}

bool is_face_set(moons_field_t U[3], const moons_mesh_s *m) {
  for (int d = 0; d < 3; ++d)
    if (!U || !U[d] || U[d]->dl != (1 << d) || U[d]->m != m) return false;
  return true;
}

} // namespace

extern "C" {

int moons_nrg_create(moons_mesh_t m, const moons_nrg_params *P, moons_field_t T, moons_field_t U[3],
                     moons_field_t k[3], moons_nrg_t *out) {
  if (!stream()) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!m || !P || !T || !out) return fail(MOONS_ERR_ARG, "moons_nrg_create: null argument");
  if (P->method < 1 || P->method > 5 || !(P->dt > 0.0))
    return fail(MOONS_ERR_ARG, "moons_nrg_create: need dt > 0 and method 1 to 5");
  if (P->rk_stages != 0 && P->rk_stages != 1 && P->rk_stages != 3)
    return fail(MOONS_ERR_ARG, "moons_nrg_create: rk_stages must be 1 or 3");
  if (T->dl != MOONS_DL_CC || T->m != m) return fail(MOONS_ERR_ARG, "moons_nrg_create: T must be cell-centred on the mesh");
  if ((P->advection || P->KE_diffusion || P->viscous_dissipation) && !is_face_set(U, m))
    return fail(MOONS_ERR_ARG, "moons_nrg_create: advection, KE diffusion and viscous dissipation need U as the three "
                               "face components on the mesh");
  if (P->method >= 3 && !k) return fail(MOONS_ERR_ARG, "moons_nrg_create: methods 3-5 need k");
  moons_nrg_s *S = new moons_nrg_s();
  S->m = m;
  S->P = *P;
  if (S->P.rk_stages == 0) S->P.rk_stages = 1;
  S->T = T;
  for (int d = 0; d < 3; ++d) S->U[d] = U ? U[d] : nullptr;
  int st = MOONS_OK;
  moons_field_s **cc[] = {&S->F, &S->Fnm1, &S->L, &S->Tnm1, &S->rhs, &S->tcc, &S->tcc2};
  for (auto f : cc)
    if (!st) st = fcreate(m, MOONS_DL_CC, f);
  for (int d = 0; d < 3 && !st; ++d) st = fcreate(m, 1 << d, &S->tF[d]);
  if (P->KE_diffusion || P->viscous_dissipation)
    for (int d = 0; d < 3 && !st; ++d) st = fcreate(m, MOONS_DL_CC, &S->UCC[d]);
  if (!st) st = moons_field_fill(S->F, 0.0);
  if (!st) st = moons_field_fill(S->L, 0.0);
  if (!st) st = moons_field_copy(S->Tnm1, T);
  if (!st && P->volumetric_heating) {
    // add_volumetric_heating_Nuclear (energy_sources.f90:95-113): a = L = 1,
    // m_ = a/L, F = scale 2 m_/(1 - exp(-2 m_)), q(j) = F exp(-m_ (y_c(j) + 1)),
    // on every cell (ghosts included); formed once on the host
    const double a = 1.0, Lh = 1.0, mm = a / Lh;
    const double Fc = P->volumetric_heating_scale * 2.0 * mm / (1.0 - std::exp(-2.0 * mm));
    const moons::HostCoords &cy = m->hc[1];
    std::vector<double> q((size_t)T->n);
    for (int kk = 0; kk < T->s[2]; ++kk)
      for (int j = 0; j < T->s[1]; ++j) {
        const double v = Fc * std::exp(-mm * (cy.hc[j] + 1.0));
        for (int i = 0; i < T->s[0]; ++i) q[((size_t)kk * T->s[1] + j) * T->s[0] + i] = v;
      }
    st = fcreate(m, MOONS_DL_CC, &S->qv);
    if (!st) st = moons_field_upload(S->qv, q.data(), q.size());
  }
  if (!st && P->method >= 3) {
    st = moons_pcg_sf_create(m, MOONS_OP_NRG_DIFFUSION, MOONS_PREC_LAP_SF, T, &S->pcg);
    if (!st) st = moons_pcg_set_isp(S->pcg, P->iter_max, P->tol_rel, P->tol_abs, P->n_skip);
    if (!st) st = moons_pcg_sf_set_k(S->pcg, k);
  }
  if (!st && (hipEventCreate(&S->ev0) != hipSuccess || hipEventCreate(&S->ev1) != hipSuccess))
    st = fail(MOONS_ERR_HIP, "moons_nrg_create: event creation failed");
  if (st) {
    nrg_free(S);
    return st;
  }
  *out = S;
  return MOONS_OK;
}

int moons_nrg_destroy(moons_nrg_t S) {
  if (!S) return MOONS_OK;
  if (stream()) (void)hipStreamSynchronize(stream());
  nrg_free(S);
  return MOONS_OK;
}

int moons_nrg_set_joule(moons_nrg_t S, moons_field_t J[3], moons_field_t sigma_inv_cc) {
  if (!S || !J || !sigma_inv_cc) return fail(MOONS_ERR_ARG, "moons_nrg_set_joule: null argument");
  const int edl[3] = {6, 5, 3}; // x, y, z edges
  for (int d = 0; d < 3; ++d)
    if (!J[d] || J[d]->dl != edl[d] || J[d]->m != S->m)
      return fail(MOONS_ERR_ARG, "moons_nrg_set_joule: J must be the x, y and z edge fields on the energy mesh");
  if (sigma_inv_cc->dl != MOONS_DL_CC || sigma_inv_cc->m != S->m)
    return fail(MOONS_ERR_ARG, "moons_nrg_set_joule: sigmaInv_CC must be cell-centred on the energy mesh");
  for (int d = 0; d < 3; ++d) S->J[d] = J[d];
  S->sig = sigma_inv_cc;
  return MOONS_OK;
}

int moons_nrg_step(moons_nrg_t S, int n_steps, moons_nrg_info *info) {
  if (!S || n_steps < 0) return fail(MOONS_ERR_ARG, "moons_nrg_step: bad argument");
  if (S->P.joule_heating && (!S->J[0] || !S->sig))
    return fail(MOONS_ERR_STATE, "moons_nrg_step: Joule heating needs moons_nrg_set_joule");
  MOONS_HIP(hipEventRecord(S->ev0, stream()));
  int st;
  for (int n = 0; n < n_steps; ++n)
    for (S->stage = 0; S->stage < S->P.rk_stages; ++S->stage) {
      if ((st = nrg_sources(S))) return st;
      if ((st = nrg_march(S))) return st;
    }
  S->stage = 0;
  MOONS_HIP(hipEventRecord(S->ev1, stream()));
  MOONS_HIP(hipEventSynchronize(S->ev1));
  if (info) {
    info->iterations = S->iters;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, S->ev0, S->ev1);
    info->ms = ms;
  }
  return MOONS_OK;
}

int moons_nrg_field(moons_nrg_t S, int which, moons_field_t *out) {
  if (!S || !out) return fail(MOONS_ERR_ARG, "moons_nrg_field: null argument");
  switch (which) {
    case MOONS_NRG_F: *out = S->F; return MOONS_OK;
    case MOONS_NRG_FNM1: *out = S->Fnm1; return MOONS_OK;
    case MOONS_NRG_TNM1: *out = S->Tnm1; return MOONS_OK;
    case MOONS_NRG_L: *out = S->L; return MOONS_OK;
    default: return fail(MOONS_ERR_ARG, "moons_nrg_field: unknown field");
  }
}

} // extern "C"
