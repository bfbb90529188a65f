// vfpcg.hip -- FP64 Jacobi-PCG for the face-centred vector solves: the
// momentum implicit diffusion (mom_diffusion, A = I + c lap_centered_VF) and
// the induction implicit diffusion (ind_diffusion, A = I + c curl(k curl)).
//
// Restates solve_PCG_VF (code/handwritten/solvers/PCG/PCG_solver_algorithm.f90:109-179),
// modify_RHS_VF (PCG_aux.f90:56-79), the operators (matrix_free_operators.f90:120-146,
// 176-202) and prec_mom_VF / prec_ind_VF (preconditioners.f90:86-129).  The
// operator is composed from the general staggered kernels (implicit BCs, del,
// curl); vector updates and dot products are three-component kernels (one
// grid row per component) reading alpha/beta from a device control block,
// with the reductions finished by one-block scalar kernels, so the loop needs
// no host round trip; the host reads the exit flag between chunks.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace moons {

struct VF3 {
  double *p[3];
  int s[3][3];
  size_t n[3];
};

struct VfCtrl {
  int iter, done, exit_loop, iter_max, n_skip, pad_;
  double tol_rel, tol_abs, rho, alpha, beta, res0, res, pq, rr, rz;
};

static VF3 vf3(moons_field_s *const f[3]) {
  VF3 v;
  for (int d = 0; d < 3; ++d) {
    v.p[d] = f[d]->d;
    for (int e = 0; e < 3; ++e) v.s[d][e] = f[d]->s[e];
    v.n[d] = f[d]->n;
  }
  return v;
}

// ------------------------------------------------------------------ kernels
constexpr int kVT = 256;

__global__ void __launch_bounds__(kVT) k_vf_mul(VF3 y, VF3 a, const int *done) {
  if (*done) return;
  const int d = blockIdx.y;
  const size_t t = (size_t)blockIdx.x * kVT + threadIdx.x;
  if (t < y.n[d]) y.p[d][t] = y.p[d][t] * a.p[d][t];
}

// The staggered 1D stencils of the three directions (cc2n: CC -> N, n2cc: N -> CC).
struct StagCoefs {
  const double *cc2n_D[3], *cc2n_U[3], *n2cc_D[3], *n2cc_U[3];
};

// The wall-plane rules of the operator's output in the PCG loop
// (multiply_wall_Neumann 0.5, assign_wall_Dirichlet 0, assign_wall_Periodic_single
// 0 on the boundary-normal planes 1 and s-2 of each component; vf_wall_fix),
// then the product with the cell volume: applied by the operator kernels to
// each element they write when `on` is set, with the same arithmetic.
struct WallVol {
  int on;
  const double *vol[3];
  int lo[3], hi[3]; // 0 none, 1 Dirichlet (set 0), 2 Neumann (times 0.5)
  int per[3];       // periodic pair: plane s-2 set to 0
};

// (volv: the cell volume W.vol[d][t], loaded by the caller ahead of the
// arithmetic so that its latency overlaps the operator's loads)
__device__ __forceinline__ double wall_vol(const WallVol &W, int d, int pd, int sd, double volv, double v) {
  if (pd == 1) {
    if (W.lo[d] == 1) v = 0.0;
    else if (W.lo[d] == 2) v = v * 0.5;
  }
  if (pd == sd - 2) {
    if (W.hi[d] == 1) v = 0.0;
    else if (W.hi[d] == 2) v = v * 0.5;
    if (W.per[d]) v = 0.0;
  }
  return v * volv;
}

// The value k_del (staggered kinds, gen 1, pad 0) stores at position q along
// a pencil of the first pass: kind 0 CC -> N reads input q-1, q; kind 1 N -> CC
// reads q, q+1; entries without both inputs are the zeroed ends.
// Entries without both inputs load from position 0 and are discarded, so no
// branch separates the loads of the calls.
__device__ __forceinline__ double stag_first(const double *__restrict__ fp, int fs, int q, int s_in, int kind,
                                             const double *__restrict__ D, const double *__restrict__ U) {
  const int i = q - (kind == 0 ? 1 : 0);
  const bool ok = i >= 0 && i <= s_in - 2;
  const int ic = ok ? i : 0;
  const double x0 = fp[ic * fs] * D[ic];
  const double x1 = fp[(ic + 1) * fs] * U[ic];
  return ok ? x0 + x1 : 0.0;
}

// Ax = x + c lap(x) for the three face components: mom_diffusion
// (code/handwritten/solvers/matrix_free_operators.f90:176-202) over
// lap_centered_SF_given_both (code/handwritten/ops/ops_discrete.f90:112-122),
// one thread per element of Ax_d.  For each direction e in turn, the staggered
// derivative out and back (two del passes through a temporary) is evaluated
// from x: the second pass's two inputs are the first pass's values recomputed
// here.  The arithmetic, the zeroed end entries and the accumulation of the
// two passes per direction (assign for x, add for y and z, the add starting
// from zero at the pencil ends) and the final Ax = lap*c + x (multiply then
// add, as mom_diffusion_explicit) are those of the separate launches, so the
// result is identical.
__global__ void __launch_bounds__(kVT) k_vf_mom_op(VF3 Ax, VF3 x, StagCoefs C, double c, WallVol W,
                                                   const int *gate) {
  if (gate && *gate) return;
  const int d = blockIdx.y;
  // fields stay below 2^31 elements; XCD-contiguous block numbering keeps the
  // neighbouring planes a derivative reads in one L2
  const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x) * kVT + (int)threadIdx.x;
  if (t >= (int)Ax.n[d]) return;
  const int s0 = x.s[d][0], s1 = x.s[d][1];
  const int s[3] = {s0, s1, x.s[d][2]};
  const int io[3] = {t % s0, (t / s0) % s1, t / (s0 * s1)};
  const int str[3] = {1, s0, s0 * s1};
  const double *__restrict__ xp = x.p[d];
  const double volv = W.on ? W.vol[d][t] : 0.0;
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const bool n_along = (e == d); // face-d data is node-located along d
    const int kind1 = n_along ? 1 : 0, kind2 = 1 - kind1;
    const double *D1 = kind1 == 0 ? C.cc2n_D[e] : C.n2cc_D[e];
    const double *U1 = kind1 == 0 ? C.cc2n_U[e] : C.n2cc_U[e];
    const double *D2 = kind2 == 0 ? C.cc2n_D[e] : C.n2cc_D[e];
    const double *U2 = kind2 == 0 ? C.cc2n_U[e] : C.n2cc_U[e];
    const int se = s[e];
    const int nt = kind1 == 0 ? se + 1 : se - 1; // temporary's length along e
    const int o = io[e];
    const double *fp = xp + (t - o * str[e]); // start of the pencil along e
    const int i2 = o - (kind2 == 0 ? 1 : 0);
    const bool written = (i2 >= 0 && i2 <= nt - 2);
    const bool zero_end = (o == 0 || o == se - 1);
    // both passes' loads without a branch (positions clamped in range where
    // the entry is not written), then the entry selected
    const int i2c = written ? i2 : 0;
    const double t0 = stag_first(fp, str[e], i2c, se, kind1, D1, U1);
    const double t1 = stag_first(fp, str[e], i2c + 1, se, kind1, D1, U1);
    const double x0 = t0 * D2[i2c];
    const double x1 = t1 * U2[i2c];
    double v;
    if (written) {
      if (e == 0) {
        v = x0 + x1;
      } else {
        const double cur = zero_end ? 0.0 : acc;
        v = cur + x0 + x1;
      }
    } else {
      v = (e == 0 || zero_end) ? 0.0 : acc;
    }
    acc = v;
  }
  const double w = acc * c;
  double res = w + xp[t];
  if (W.on) res = wall_vol(W, d, io[d], s[d], volv, res);
  Ax.p[d][t] = res;
}

// out_d = curl(in)_d, then (mode 0) out_d = curl * aux_d (as k_vf_mul) or
// (mode 1) out_d = curl * c + aux_d (multiply then add, as
// ind_diffusion_explicit): curl_SF component d
// (code/handwritten/ops/ops_discrete.f90:184-202) assigns the derivative of
// in[(d+2)%3] along (d+1)%3 and subtracts that of in[(d+1)%3] along (d+2)%3,
// both staggered (mode 0: faces to edges, CC -> N; mode 1: edges to faces,
// N -> CC).  One thread per element of out_d with the arithmetic and zeroed
// end entries of the two del launches and of the following element-wise
// kernel, so the result is identical.
__global__ void __launch_bounds__(kVT) k_vf_curl_op(VF3 out, VF3 in, StagCoefs C, VF3 aux, double c, int mode,
                                                    WallVol W, const int *gate) {
  if (gate && *gate) return;
  const int d = blockIdx.y;
  // fields stay below 2^31 elements; XCD-contiguous block numbering keeps the
  // neighbouring planes a derivative reads in one L2
  const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x) * kVT + (int)threadIdx.x;
  if (t >= (int)out.n[d]) return;
  const int s0 = out.s[d][0], s1 = out.s[d][1];
  const int io[3] = {t % s0, (t / s0) % s1, t / (s0 * s1)};
  const int kind = mode == 0 ? 0 : 1;
  const double volv = (mode != 0 && W.on) ? W.vol[d][t] : 0.0;
  const double auxv = aux.p[d][t];
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int ci = q == 0 ? (d + 2) % 3 : (d + 1) % 3; // input component
    const int e = q == 0 ? (d + 1) % 3 : (d + 2) % 3;  // derivative direction
    const int *si = in.s[ci];
    const int stri[3] = {1, si[0], si[0] * si[1]};
    const double *D = kind == 0 ? C.cc2n_D[e] : C.n2cc_D[e];
    const double *U = kind == 0 ? C.cc2n_U[e] : C.n2cc_U[e];
    const int o = io[e];
    const int nd = out.s[d][e], s_in = si[e];
    int base = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
      if (r != e) base += io[r] * stri[r];
    const double *fp = in.p[ci] + base;
    const int i = o - (kind == 0 ? 1 : 0);
    const bool written = (i >= 0 && i <= s_in - 2);
    const bool zero_end = (o == 0 || o == nd - 1);
    // loads at a clamped position where the entry is not written (discarded)
    const int ic = written ? i : 0;
    const double x0 = fp[ic * stri[e]] * D[ic];
    const double x1 = fp[(ic + 1) * stri[e]] * U[ic];
    if (q == 0) {
      v = written ? x0 + x1 : 0.0;
    } else {
      const double cur = zero_end ? 0.0 : v;
      v = written ? cur - (x0 + x1) : cur;
    }
  }
  if (mode == 0) {
    out.p[d][t] = v * auxv;
  } else {
    const double w = v * c;
    double res = w + auxv;
    if (W.on) res = wall_vol(W, d, io[d], out.s[d][d], volv, res);
    out.p[d][t] = res;
  }
}

// y = y - x
__global__ void __launch_bounds__(kVT) k_vf_sub(VF3 y, VF3 x, const int *done) {
  if (*done) return;
  const int d = blockIdx.y;
  const size_t t = (size_t)blockIdx.x * kVT + threadIdx.x;
  if (t < y.n[d]) y.p[d][t] = y.p[d][t] - x.p[d][t];
}

// x = x + p*alpha and r = r + Ax*(-alpha) in one pass (two k_vf_axpy)
__global__ void __launch_bounds__(kVT) k_vf_axpy2(VF3 x, VF3 p, VF3 r, VF3 Ax, const VfCtrl *ctrl) {
  if (ctrl->done) return;
  const int d = blockIdx.y;
  const size_t t = (size_t)blockIdx.x * kVT + threadIdx.x;
  const double a = ctrl->alpha, ma = -ctrl->alpha;
  if (t < x.n[d]) {
    x.p[d][t] = x.p[d][t] + p.p[d][t] * a;
    r.p[d][t] = r.p[d][t] + Ax.p[d][t] * ma;
  }
}

// z = Minv*r (all cells)
__global__ void __launch_bounds__(kVT) k_vf_minv(VF3 z, VF3 M, VF3 r, const int *done) {
  if (*done) return;
  const int d = blockIdx.y;
  const size_t t = (size_t)blockIdx.x * kVT + threadIdx.x;
  if (t < z.n[d]) z.p[d][t] = M.p[d][t] * r.p[d][t];
}

// p = p*beta + z
__global__ void __launch_bounds__(kVT) k_vf_pupdate(VF3 p, VF3 z, const VfCtrl *ctrl) {
  if (ctrl->done) return;
  const int d = blockIdx.y;
  const size_t t = (size_t)blockIdx.x * kVT + threadIdx.x;
  const double beta = ctrl->beta;
  if (t < p.n[d]) p.p[d][t] = p.p[d][t] * beta + z.p[d][t];
}

// Partial sums of a*b over the interior (pad 1) of each component, the
// Dirichlet wall-node planes excluded (dot_product_VF: multiply,
// assign_wall_Dirichlet, sum pad 1; code/handwritten/ops/ops_aux.f90:199-207).
// excl: bit 2d = exclude plane 1 of component d, bit 2d+1 = plane s-2.
__global__ void __launch_bounds__(kVT) k_vf_dot(VF3 a, VF3 b, int excl, double *part, int nbmax, const int *done) {
  if (*done) return;
  __shared__ double sh[kVT / kWave];
  const int d = blockIdx.y;
  const int s0 = a.s[d][0], s1 = a.s[d][1], s2 = a.s[d][2];
  const int n0 = s0 - 2, n01 = (s0 - 2) * (s1 - 2);
  const int ni = n01 * (s2 - 2); // fields of the VF solver stay below 2^31 elements (checked at creation)
  double acc = 0.0;
  for (int t = blockIdx.x * kVT + threadIdx.x; t < ni; t += gridDim.x * kVT) {
    const int k = 1 + t / n01;
    const int r = t - (k - 1) * n01;
    const int j = 1 + r / n0;
    const int i = 1 + (r - (j - 1) * n0);
    const int ii[3] = {i, j, k};
    const int sd = a.s[d][d];
    if (((excl >> (2 * d)) & 1) && ii[d] == 1) continue;
    if (((excl >> (2 * d + 1)) & 1) && ii[d] == sd - 2) continue;
    const size_t o = (size_t)i + (size_t)s0 * ((size_t)j + (size_t)s1 * k);
    acc += a.p[d][o] * b.p[d][o];
  }
  const double s = block_sum<kVT>(acc, sh);
  if (threadIdx.x == 0) part[(size_t)d * nbmax + blockIdx.x] = s;
}

enum { OP_RES0 = 0, OP_RHO0 = 1, OP_ALPHA = 2, OP_CHECK = 3, OP_BETA = 4 };

// Finishes a dot product (fixed order) and advances the PCG scalars.
__global__ void __launch_bounds__(kVT) k_vf_scalar(VfCtrl *ctrl, const double *part, int nbmax, int op, int it) {
  if (op >= OP_ALPHA && ctrl->done) return;
  __shared__ double sh[kVT / kWave];
  double v = 0.0;
  // components summed in order x, y, z as sum(temp%x,1) + sum(temp%y,1) + sum(temp%z,1)
  double tot = 0.0;
  for (int d = 0; d < 3; ++d) {
    v = 0.0;
    for (int q = threadIdx.x; q < nbmax; q += kVT) v += part[(size_t)d * nbmax + q];
    const double s = block_sum<kVT>(v, sh);
    if (threadIdx.x == 0) tot = tot + s;
  }
  if (threadIdx.x != 0) return;
  switch (op) {
    case OP_RES0: ctrl->res0 = sqrt(tot); break;
    case OP_RHO0: {
      ctrl->rho = tot;
      const double res = sqrt(fabs(tot));
      ctrl->res = res;
      const bool e1 = res / ctrl->res0 < ctrl->tol_rel, e2 = res < ctrl->tol_abs, e3 = 0 >= ctrl->iter_max;
      ctrl->exit_loop = (e1 ? 1 : 0) | (e2 ? 2 : 0) | (e3 ? 4 : 0);
      ctrl->iter = e2 ? 1 : 0;
      ctrl->done = e2 ? 1 : 0;
    } break;
    case OP_ALPHA:
      ctrl->pq = tot;
      ctrl->alpha = ctrl->rho / tot;
      break;
    case OP_CHECK: {
      ctrl->iter = it;
      ctrl->rr = tot;
      if (it % ctrl->n_skip == 0) {
        const double res = sqrt(fabs(tot));
        ctrl->res = res;
        const bool e1 = res / ctrl->res0 < ctrl->tol_rel, e2 = res < ctrl->tol_abs, e3 = it >= ctrl->iter_max;
        ctrl->exit_loop = (e1 ? 1 : 0) | (e2 ? 2 : 0) | (e3 ? 4 : 0);
        if (e1 || e2 || e3) ctrl->done = 1;
      }
    } break;
    case OP_BETA:
      ctrl->rz = tot;
      ctrl->beta = tot / ctrl->rho;
      ctrl->rho = tot;
      if (it >= ctrl->iter_max) {
        ctrl->exit_loop |= 4;
        ctrl->done = 1;
      }
      break;
  }
}

} // namespace moons

using namespace moons;

struct moons_pcgvf_s {
  moons_mesh_s *m = nullptr;
  int kind = 0;
  double c = 0.0;
  int iter_max = 1000, n_skip = 1;
  double tol_rel = 1e-6, tol_abs = 1e-13;
  moons_field_s bcs[3];                     // locations and BCs of x (no storage)
  moons_field_s *r[3] = {}, *p[3] = {}, *Ax[3] = {}, *z[3] = {}, *xbc[3] = {}, *vol[3] = {}, *Minv[3] = {};
  moons_field_s *k[3] = {};                 // sigmaInv on edges (induction)
  moons_field_s *E[3] = {};                 // edge work (induction)
  double *part = nullptr;
  int nbmax = 0;
  int excl = 0;
  VfCtrl *ctrl = nullptr, *host = nullptr, *host_init = nullptr; // pinned: read-backs, initial block
  bool minv_ready = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static void vf_free(moons_pcgvf_s *S) {
  moons_field_s **all[] = {S->r, S->p, S->Ax, S->z, S->xbc, S->vol, S->Minv, S->k, S->E};
  for (auto arr : all)
    for (int d = 0; d < 3; ++d)
      if (arr[d]) moons_field_destroy(arr[d]);
  if (S->part) (void)hipFree(S->part);
  if (S->ctrl) (void)hipFree(S->ctrl);
  if (S->host) (void)hipHostFree(S->host);
  if (S->host_init) (void)hipHostFree(S->host_init);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  delete S;
}

// Fields of the solver carry x's BCs (p is x's copy in the reference).
static void copy_bcs(moons_field_s *dst, const moons_field_s &src) { copy_bcs_of(dst, &src); }

// Minv of prec_mom_VF / prec_ind_VF computed on the host in the reference's
// operation order (laplacian_diagonal / curl_curl_diagonal, *c, +1, *vol, 1/x,
// ghosts zeroed), then uploaded.
static int vf_minv(moons_pcgvf_s *S) {
  const HostCoords *C = S->m->hc;
  std::vector<double> sig[3];
  int ss[3][3];
  const int edl[3] = {MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z};
  if (S->kind == 2) {
    for (int d = 0; d < 3; ++d) {
      sig[d].resize(S->k[d]->n);
      int st = moons_field_download(S->k[d], sig[d].data(), S->k[d]->n);
      if (st) return st;
      for (int e = 0; e < 3; ++e) ss[d][e] = ((edl[d] >> e) & 1) ? C[e].sn : C[e].sc;
    }
  }
  for (int d = 0; d < 3; ++d) {
    const int dl = 1 << d;
    int s[3];
    for (int e = 0; e < 3; ++e) s[e] = ((dl >> e) & 1) ? C[e].sn : C[e].sc;
    const size_t n = (size_t)s[0] * s[1] * s[2];
    std::vector<double> M(n, 0.0);
    auto I3 = [&](const int *sh, int i, int j, int k) { return (size_t)i + (size_t)sh[0] * ((size_t)j + (size_t)sh[1] * k); };
    const int p0 = (dl >> 0) & 1, p1 = (dl >> 1) & 1, p2 = (dl >> 2) & 1;
    for (int k = 1; k < s[2] - 1; ++k)
      for (int j = 1; j < s[1] - 1; ++j)
        for (int i = 1; i < s[0] - 1; ++i) {
          double dg;
          if (S->kind == 1) {
            dg = C[0].n2cc_D[i] * C[0].cc2n_U[i - 1] + C[0].n2cc_U[i - p0] * C[0].cc2n_D[i - p0] +
                 C[1].n2cc_D[j] * C[1].cc2n_U[j - 1] + C[1].n2cc_U[j - p1] * C[1].cc2n_D[j - p1] +
                 C[2].n2cc_D[k] * C[2].cc2n_U[k - 1] + C[2].n2cc_U[k - p2] * C[2].cc2n_D[k - p2];
          } else if (d == 0) {
            dg = -C[1].n2cc_D[j] * sig[2][I3(ss[2], i, j, k)] * C[1].cc2n_U[j - 1] -
                 C[1].n2cc_U[j] * sig[2][I3(ss[2], i, j + 1, k)] * C[1].cc2n_D[j] -
                 C[2].n2cc_D[k] * sig[1][I3(ss[1], i, j, k)] * C[2].cc2n_U[k - 1] -
                 C[2].n2cc_U[k] * sig[1][I3(ss[1], i, j, k + 1)] * C[2].cc2n_D[k];
          } else if (d == 1) {
            dg = -C[0].n2cc_D[i] * sig[2][I3(ss[2], i, j, k)] * C[0].cc2n_U[i - 1] -
                 C[0].n2cc_U[i] * sig[2][I3(ss[2], i + 1, j, k)] * C[0].cc2n_D[i] -
                 C[2].n2cc_D[k] * sig[0][I3(ss[0], i, j, k)] * C[2].cc2n_U[k - 1] -
                 C[2].n2cc_U[k] * sig[0][I3(ss[0], i, j, k + 1)] * C[2].cc2n_D[k];
          } else {
            dg = -C[0].n2cc_D[i] * sig[1][I3(ss[1], i, j, k)] * C[0].cc2n_U[i - 1] -
                 C[0].n2cc_U[i] * sig[1][I3(ss[1], i + 1, j, k)] * C[0].cc2n_D[i] -
                 C[1].n2cc_D[j] * sig[0][I3(ss[0], i, j, k)] * C[1].cc2n_U[j - 1] -
                 C[1].n2cc_U[j] * sig[0][I3(ss[0], i, j + 1, k)] * C[1].cc2n_D[j];
          }
          const double *dh0 = p0 ? C[0].dhc.data() : C[0].dhn.data();
          const double *dh1 = p1 ? C[1].dhc.data() : C[1].dhn.data();
          const double *dh2 = p2 ? C[2].dhc.data() : C[2].dhn.data();
          const double vol = dh0[i - p0] * dh1[j - p1] * dh2[k - p2];
          double v = dg * S->c;
          v = v + 1.0;
          v = v * vol;
          M[I3(s, i, j, k)] = 1.0 / v;
        }
    int st = moons_field_upload(S->Minv[d], M.data(), n);
    if (st) return st;
  }
  return MOONS_OK;
}

static inline unsigned vblk(size_t n) { return (unsigned)((n + kVT - 1) / kVT); }
static inline size_t vmax(moons_field_s *const f[3]) { return std::max(f[0]->n, std::max(f[1]->n, f[2]->n)); }

// Ax = op(x): implicit BCs on x when implicit_, then lap_centered_VF or
// curl(k curl x), then Ax = Ax*c + x.
// The operator (implicit BCs on x first when implicit_); with wall_vol the
// loop's wall-plane rules and volume product are applied to Ax as well.
static int vf_op(moons_pcgvf_s *S, moons_field_s *Ax[3], moons_field_s *x[3], bool implicit_, bool wall_vol_ = false) {
  int st;
  if (implicit_ && (st = launch_apply_bcs3(x, true))) return st;
  StagCoefs C;
  for (int e = 0; e < 3; ++e) {
    const DirCoefDev &cd = S->m->dev.c[e];
    C.cc2n_D[e] = cd.cc2n_D; C.cc2n_U[e] = cd.cc2n_U;
    C.n2cc_D[e] = cd.n2cc_D; C.n2cc_U[e] = cd.n2cc_U;
  }
  WallVol W{};
  if (wall_vol_) {
    W.on = 1;
    for (int d = 0; d < 3; ++d) {
      const moons_field_s &b = S->bcs[d];
      W.vol[d] = S->vol[d]->d;
      const int s = x[d]->s[d];
      auto rule = [](int t) { return t == MOONS_BC_DIRICHLET ? 1 : (t == MOONS_BC_NEUMANN ? 2 : 0); };
      W.lo[d] = rule(b.bct[2 * d]);
      W.hi[d] = rule(b.bct[2 * d + 1]);
      W.per[d] = b.bct[2 * d] == MOONS_BC_PERIODIC && b.bct[2 * d + 1] == MOONS_BC_PERIODIC && s > 4;
    }
  }
  const int *gate = current_gate();
  if (S->kind == 1) {
    // lap and x + c lap in one launch (the del passes fused; same arithmetic)
    hipLaunchKernelGGL(k_vf_mom_op, dim3(vblk(vmax(Ax)), 3), dim3(kVT), 0, stream(), vf3(Ax), vf3(x), C, S->c, W,
                       gate);
  } else {
    // E = k curl(x), then Ax = x + c curl(E): two launches (the curls'
    // del passes and the element-wise products fused; same arithmetic)
    hipLaunchKernelGGL(k_vf_curl_op, dim3(vblk(vmax(S->E)), 3), dim3(kVT), 0, stream(), vf3(S->E), vf3(x), C,
                       vf3(S->k), 0.0, 0, WallVol{}, gate);
    hipLaunchKernelGGL(k_vf_curl_op, dim3(vblk(vmax(Ax)), 3), dim3(kVT), 0, stream(), vf3(Ax), vf3(S->E), C, vf3(x),
                       S->c, 1, W, gate);
  }
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// Wall planes of the boundary-normal component: multiply_wall_Neumann(0.5),
// assign_wall_Dirichlet(0), assign_wall_Periodic_single(0)
// (code/handwritten/block_field/block_field_extend.f90:273-377).
static int vf_wall_fix(moons_pcgvf_s *S, moons_field_s *u[3], bool periodic_first) {
  int st;
  for (int d = 0; d < 3; ++d) {
    const moons_field_s &b = S->bcs[d];
    const int s = u[d]->s[d];
    bool per = b.bct[2 * d] == MOONS_BC_PERIODIC && b.bct[2 * d + 1] == MOONS_BC_PERIODIC && s > 4;
    auto plane_scale = [&](int idx, double v) -> int { return launch_plane_scale(u[d], d, idx, v); };
    auto plane_set = [&](int idx, double v) -> int { return launch_plane_set(u[d], d, idx, v); };
    if (periodic_first && per && (st = plane_set(s - 2, 0.0))) return st;
    for (int side = 0; side < 2; ++side) {
      const int t = b.bct[2 * d + side];
      const int idx = side ? s - 2 : 1;
      if (t == MOONS_BC_DIRICHLET && (st = plane_set(idx, 0.0))) return st;
    }
    for (int side = 0; side < 2; ++side) {
      const int t = b.bct[2 * d + side];
      const int idx = side ? s - 2 : 1;
      if (t == MOONS_BC_NEUMANN && (st = plane_scale(idx, 0.5))) return st;
    }
    if (!periodic_first && per && (st = plane_set(s - 2, 0.0))) return st;
  }
  return MOONS_OK;
}

static int vf_dot(moons_pcgvf_s *S, moons_field_s *a[3], moons_field_s *b[3], int op, int it) {
  hipLaunchKernelGGL(k_vf_dot, dim3(S->nbmax, 3), dim3(kVT), 0, stream(), vf3(a), vf3(b), S->excl, S->part, S->nbmax,
                     &S->ctrl->done);
  hipLaunchKernelGGL(k_vf_scalar, dim3(1), dim3(kVT), 0, stream(), S->ctrl, S->part, S->nbmax, op, it);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

extern "C" {

int moons_pcg_vf_create(moons_mesh_t m, int op, moons_field_t x[3], moons_field_t k[3], moons_pcgvf_t *out) {
  if (!stream()) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!m || !x || !out) return fail(MOONS_ERR_ARG, "moons_pcg_vf_create: null argument");
  if (op != MOONS_OP_MOM_DIFFUSION && op != MOONS_OP_IND_DIFFUSION)
    return fail(MOONS_ERR_ARG, "moons_pcg_vf_create: operator must be mom_diffusion or ind_diffusion");
  for (int d = 0; d < 3; ++d)
    if (!x[d] || x[d]->dl != (1 << d) || x[d]->m != m)
      return fail(MOONS_ERR_ARG, "moons_pcg_vf_create: x must be the three face components on this mesh");
  for (int d = 0; d < 3; ++d)
    if (x[d]->n >= ((size_t)1 << 31))
      return fail(MOONS_ERR_ARG, "moons_pcg_vf_create: fields of 2^31 or more elements are not supported");
  if (op == MOONS_OP_IND_DIFFUSION) {
    const int edl[3] = {MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z};
    if (!k) return fail(MOONS_ERR_ARG, "moons_pcg_vf_create: ind_diffusion needs sigmaInv on edges");
    for (int d = 0; d < 3; ++d)
      if (!k[d] || k[d]->dl != edl[d]) return fail(MOONS_ERR_ARG, "moons_pcg_vf_create: k must be on edges x,y,z");
  }
  moons_pcgvf_s *S = new moons_pcgvf_s();
  S->m = m;
  S->kind = op;
  int st = MOONS_OK;
  for (int d = 0; d < 3 && !st; ++d) {
    S->bcs[d] = *x[d];
    S->bcs[d].d = nullptr;
    moons_field_s **vecs[] = {S->r, S->p, S->Ax, S->z, S->xbc, S->vol, S->Minv};
    for (auto v : vecs)
      if (!st) st = moons_field_create(m, 1 << d, &v[d]);
    if (!st) {
      copy_bcs(S->p[d], *x[d]);
      copy_bcs(S->xbc[d], *x[d]);
      copy_bcs(S->Ax[d], *x[d]);
      st = moons_volume(S->vol[d]);
    }
  }
  if (!st && op == MOONS_OP_IND_DIFFUSION) {
    const int edl[3] = {MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z};
    for (int d = 0; d < 3 && !st; ++d) {
      st = moons_field_create(m, edl[d], &S->k[d]);
      if (!st) st = moons_field_copy(S->k[d], k[d]);
      if (!st) st = moons_field_create(m, edl[d], &S->E[d]);
    }
  }
  if (st) {
    vf_free(S);
    return st;
  }
  S->nbmax = (int)std::min<size_t>(1024, vblk(std::max(S->r[0]->n, std::max(S->r[1]->n, S->r[2]->n))));
  for (int d = 0; d < 3; ++d)
    for (int side = 0; side < 2; ++side)
      if (x[d]->has_bcs && x[d]->bct[2 * d + side] == MOONS_BC_DIRICHLET) S->excl |= 1 << (2 * d + side);
  bool ok = hipMalloc(&S->part, 3 * (size_t)S->nbmax * sizeof(double)) == hipSuccess &&
            hipMalloc(&S->ctrl, sizeof(VfCtrl)) == hipSuccess &&
            hipHostMalloc(&S->host, sizeof(VfCtrl), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&S->host_init, sizeof(VfCtrl), hipHostMallocDefault) == hipSuccess &&
            hipEventCreate(&S->ev0) == hipSuccess && hipEventCreate(&S->ev1) == hipSuccess;
  if (!ok) {
    vf_free(S);
    return fail(MOONS_ERR_HIP, "moons_pcg_vf_create: allocation failed");
  }
  (void)hipMemsetAsync(S->ctrl, 0, sizeof(VfCtrl), stream());
  std::memset(S->host_init, 0xff, sizeof(VfCtrl));
  *out = S;
  return MOONS_OK;
}

int moons_pcg_vf_destroy(moons_pcgvf_t S) {
  if (!S) return MOONS_OK;
  if (stream()) (void)hipStreamSynchronize(stream());
  vf_free(S);
  return MOONS_OK;
}

int moons_pcg_vf_set_isp(moons_pcgvf_t S, int iter_max, double tol_rel, double tol_abs, int n_skip_check_res) {
  if (!S || iter_max < 1 || n_skip_check_res < 1) return fail(MOONS_ERR_ARG, "moons_pcg_vf_set_isp: bad argument");
  S->iter_max = iter_max;
  S->tol_rel = tol_rel;
  S->tol_abs = tol_abs;
  S->n_skip = n_skip_check_res;
  return MOONS_OK;
}

int moons_pcg_vf_update_mfp(moons_pcgvf_t S, double coeff_implicit_time_split, int update_preconditioner) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_pcg_vf_update_mfp: null solver");
  const bool changed = S->c != coeff_implicit_time_split;
  S->c = coeff_implicit_time_split;
  // Minv depends on c and the mesh only: rebuilding it for an unchanged c
  // would reproduce the same array
  if (update_preconditioner && (changed || !S->minv_ready)) {
    if (sync_free()) return fail(MOONS_ERR_STATE, "moons_pcg_vf_update_mfp: a new coefficient cannot be set inside a graph");
    const int st = vf_minv(S);
    if (st) return st;
    S->minv_ready = true;
  }
  return MOONS_OK;
}

int moons_pcg_vf_solve(moons_pcgvf_t S, moons_field_t x[3], moons_field_t b[3], moons_pcg_info *info) {
  if (!S || !x || !b) return fail(MOONS_ERR_ARG, "moons_pcg_vf_solve: null argument");
  for (int d = 0; d < 3; ++d)
    if (!x[d] || !b[d] || x[d]->dl != (1 << d) || b[d]->dl != (1 << d) || x[d]->m != S->m || b[d]->m != S->m)
      return fail(MOONS_ERR_ARG, "moons_pcg_vf_solve: x and b must be face components on the solver's mesh");
  VfCtrl h;
  std::memset(&h, 0, sizeof(VfCtrl));
  h.iter_max = S->iter_max;
  h.n_skip = S->n_skip;
  h.tol_rel = S->tol_rel;
  h.tol_abs = S->tol_abs;
  if (std::memcmp(&h, S->host_init, sizeof(VfCtrl)) != 0) {
    if (!sync_free()) MOONS_HIP(hipStreamSynchronize(stream()));
    *S->host_init = h;
  }
  MOONS_HIP(hipMemcpyAsync(S->ctrl, S->host_init, sizeof(VfCtrl), hipMemcpyHostToDevice, stream()));
  const bool nosync = sync_free();
  if (!nosync) MOONS_HIP(hipEventRecord(S->ev0, stream()));
  // the solve's x carries its own BCs; keep a copy with x's BCs for the implicit application
  moons_field_s xs[3];
  moons_field_s *xp[3];
  for (int d = 0; d < 3; ++d) {
    xs[d] = *x[d];
    xp[d] = &xs[d];
  }
  int st;
  const int *done = &S->ctrl->done;
  const size_t nmax = vmax(S->r);
  const dim3 g3(vblk(nmax), 3), blk(kVT);
  set_gate(done);
  // ---- modify_RHS_VF
  for (int d = 0; d < 3; ++d)
    if ((st = moons_field_copy(S->r[d], b[d]))) goto out;
  for (int d = 0; d < 3; ++d) {
    if ((st = moons_field_fill(S->xbc[d], 0.0))) goto out;
    if ((st = launch_apply_bcs(S->xbc[d], false))) goto out;
    for (int f = 0; f < 6; ++f) {
      const int e = f / 2;
      if (!((S->xbc[d]->dl >> e) & 1) || S->xbc[d]->bct[f] == MOONS_BC_PERIODIC) continue;
      if ((st = launch_plane_set(S->xbc[d], e, (f % 2) ? S->xbc[d]->s[e] - 1 : 0, 0.0))) goto out;
    }
  }
  if ((st = vf_op(S, S->Ax, S->xbc, false))) goto out;
  for (int d = 0; d < 3; ++d)
    if ((st = launch_assign_ghost_xperiodic(S->Ax[d], 0.0))) goto out;
  hipLaunchKernelGGL(k_vf_sub, g3, blk, 0, stream(), vf3(S->r), vf3(S->Ax), done);
  if ((st = vf_op(S, S->Ax, xp, true))) goto out;
  hipLaunchKernelGGL(k_vf_sub, g3, blk, 0, stream(), vf3(S->r), vf3(S->Ax), done);
  hipLaunchKernelGGL(k_vf_mul, g3, blk, 0, stream(), vf3(S->r), vf3(S->vol), done);
  if ((st = vf_wall_fix(S, S->r, true))) goto out;
  // ---- set-up: res0, z = Minv r, p = z, rho = (r,z)
  if ((st = vf_dot(S, S->r, S->r, OP_RES0, 0))) goto out;
  hipLaunchKernelGGL(k_vf_minv, g3, blk, 0, stream(), vf3(S->z), vf3(S->Minv), vf3(S->r), done);
  for (int d = 0; d < 3; ++d)
    if ((st = moons_field_copy(S->p[d], S->z[d]))) goto out;
  if ((st = vf_dot(S, S->r, S->z, OP_RHO0, 0))) goto out;
  {
    int launched = 0;
    for (;;) {
      if (nosync) {
        if (launched >= S->iter_max) break;
      } else {
        MOONS_HIP(hipMemcpyAsync(S->host, S->ctrl, sizeof(VfCtrl), hipMemcpyDeviceToHost, stream()));
        MOONS_HIP(hipStreamSynchronize(stream()));
        if (S->host->done || launched >= S->iter_max) break;
      }
      int chunk = nosync ? S->iter_max : std::min(std::max(2, std::min(16, launched / 4)), S->iter_max - launched);
      for (int c = 0; c < chunk; ++c) {
        const int it = launched + c + 1;
        if ((st = vf_op(S, S->Ax, S->p, true, true))) goto out; // with the wall rules and vol
        if ((st = vf_dot(S, S->p, S->Ax, OP_ALPHA, it))) goto out;
        hipLaunchKernelGGL(k_vf_axpy2, g3, blk, 0, stream(), vf3(xp), vf3(S->p), vf3(S->r), vf3(S->Ax), S->ctrl);
        if ((st = vf_dot(S, S->r, S->r, OP_CHECK, it))) goto out;
        hipLaunchKernelGGL(k_vf_minv, g3, blk, 0, stream(), vf3(S->z), vf3(S->Minv), vf3(S->r), done);
        if ((st = vf_dot(S, S->z, S->r, OP_BETA, it))) goto out;
        hipLaunchKernelGGL(k_vf_pupdate, g3, blk, 0, stream(), vf3(S->p), vf3(S->z), S->ctrl);
        MOONS_LAUNCH_CHECK();
      }
      launched += chunk;
    }
  }
  set_gate(nullptr);
  // apply_BCs(x) after the last update (PCG_solver_algorithm.f90:149,170)
  for (int d = 0; d < 3; ++d)
    if ((st = launch_apply_bcs(x[d], false))) return st;
  if (nosync) {
    if (info) {
      info->iterations = -1;
      info->exit_loop = 0;
      info->res_norm0 = info->res_norm = info->ms = 0.0;
    }
    return MOONS_OK;
  }
  MOONS_HIP(hipEventRecord(S->ev1, stream()));
  MOONS_HIP(hipMemcpyAsync(S->host, S->ctrl, sizeof(VfCtrl), hipMemcpyDeviceToHost, stream()));
  MOONS_HIP(hipStreamSynchronize(stream()));
  if (std::isnan(S->host->res) || std::isnan(S->host->res0))
    return fail(MOONS_ERR_NAN, "Error: NaN in PCG residual (check_nans, PCG_solver_algorithm.f90:223-248)");
  if (info) {
    info->iterations = S->host->iter;
    info->exit_loop = S->host->exit_loop;
    info->res_norm0 = S->host->res0;
    info->res_norm = S->host->res;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, S->ev0, S->ev1);
    info->ms = ms;
  }
  return MOONS_OK;
out:
  set_gate(nullptr);
  return st;
}

} // extern "C"
