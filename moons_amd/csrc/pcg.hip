// pcg.hip -- FP64 Jacobi-preconditioned CG for the cell-centred Laplacian
// (pressure-Poisson and phi-cleaning solves), the whole loop on the device.
//
// Restates solve_PCG_SF (code/handwritten/solvers/PCG/PCG_solver_algorithm.f90:37-107)
// with op = Lap_uniform_SF (code/handwritten/solvers/matrix_free_operators.f90:35-62)
// or nrg_diffusion (the energy equation's I + c div(k grad),
// matrix_free_operators.f90:148-174), prec = prec_Lap_SF
// (code/handwritten/solvers/preconditioners/preconditioners.f90:55-71) and
// modify_RHS_SF (code/handwritten/solvers/PCG/PCG_aux.f90:30-54).
//
// One iteration is three HBM-bound kernels over the interior cells:
//   k_pcg_matvec   q = vol*lap(p), partial (p,q)              read p, write q        16 B/cell
//   k_pcg_update   x += a p, r -= a q, partials (r,r),(Minv r,r) read x,p,r,q write x,r 48 B/cell
//   k_pcg_pupdate  exit test; p = b p + Minv r (+ ghosts of p)  read r,p write p        24 B/cell
// Minv = 1/(vol*diag) is rebuilt in registers from 1D arrays (no 3D array).
//
// Scalars live in a device control block.  A kernel producing a dot product
// writes one partial per block; the next kernel finishes that sum in its
// prologue: every block sums the same partials in the same fixed order, so
// all blocks derive identical alpha, beta and exit decisions with no atomics
// and no cross-block fences (the kernel boundary makes the partials visible).
// rho is double-buffered by iteration parity, since block 0 publishes the new
// value while the other blocks of that kernel still read the old one.  Every
// kernel returns at once after the exit decision; the host launches chunks of
// iterations (passing the iteration number) and reads the exit flag between
// chunks, sized by the residual's observed rate of decrease.
#include "cc_stencil.hpp"
#include "mg.hpp"
#include "unet.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

namespace moons {

struct PcgCtrl {
  int iter, done, exit_loop, iter_max, n_skip, pad_;
  double tol_rel, tol_abs;
  double rho[2]; // rho of the last iteration, indexed by its parity
  double alpha, beta, res0, res, mean, volsum, pq, rr, rz;
};

// ------------------------------------------------------------------ kernels

// Where an FP32 preconditioner vector keeps cell (i,j,k): at the padded index
// o of the field (split == 0) or in the multigrid's colour-split layout.
struct F32Map {
  int split, h0, n1, half;
};
__device__ __forceinline__ size_t f32_at(const F32Map &M, size_t o, int i, int j, int k) {
  return M.split ? (size_t)(((i + j + k) & 1) * M.half + ((i - 1) >> 1) + M.h0 * ((j - 1) + M.n1 * (k - 1))) : o;
}

// The nrg_diffusion operator's face coefficients k (x-, y-, z-face arrays of
// the mesh) and c = coeff_implicit_time_split.
struct NrgK {
  const double *k[3];
  double c;
};

// c div(k grad x) + x at an interior cell, in nrg_diffusion_explicit's order
// (matrix_free_operators.f90:148-161): the face gradients of lap7, each times
// its face k, the divergence summed as in lap7, times c, plus x.
__device__ __forceinline__ double nrg7(const NrgK &K, const CCGeom &G, double c0, double xm, double xp, double ym,
                                       double yp, double zm, double zp, int i, int j, int k) {
  const size_t s0 = G.s0, s1 = G.s1;
  const size_t ox = (size_t)i + (s0 + 1) * ((size_t)j + s1 * (size_t)k); // x-faces: (s0+1) x s1 x s2
  const size_t oy = (size_t)i + s0 * ((size_t)j + (s1 + 1) * (size_t)k); // y-faces: s0 x (s1+1) x s2
  const size_t oz = (size_t)i + s0 * ((size_t)j + s1 * (size_t)k);       // z-faces: s0 x s1 x (s2+1)
  const Coef1 &X = G.c[0], &Y = G.c[1], &Z = G.c[2];
  const double gxm = (xm * X.cD[i - 1] + c0 * X.cU[i - 1]) * K.k[0][ox];
  const double gxp = (c0 * X.cD[i] + xp * X.cU[i]) * K.k[0][ox + 1];
  double v = gxm * X.nD[i] + gxp * X.nU[i];
  const double gym = (ym * Y.cD[j - 1] + c0 * Y.cU[j - 1]) * K.k[1][oy];
  const double gyp = (c0 * Y.cD[j] + yp * Y.cU[j]) * K.k[1][oy + s0];
  v = v + gym * Y.nD[j] + gyp * Y.nU[j];
  const double gzm = (zm * Z.cD[k - 1] + c0 * Z.cU[k - 1]) * K.k[2][oz];
  const double gzp = (c0 * Z.cD[k] + zp * Z.cU[k]) * K.k[2][oz + s0 * s1];
  v = v + gzm * Z.nD[k] + gzp * Z.nU[k];
  v = v * K.c;
  return v + c0;
}

// Partial sums of b*vol over interior cells (subtract_phys_mean_vol_SF,
// code/handwritten/ops/ops_aux.f90:146-167).
__global__ void __launch_bounds__(kNT) k_pcg_mean(const double *__restrict__ b, CCGeom G, double *partM) {
  __shared__ double sh[kNT / kWave];
  const TileIdx t = tile_of(G);
  double acc = 0.0;
  if (t.active) {
    const double dxy = G.c[0].dh[t.i] * G.c[1].dh[t.j];
    const size_t sxy = (size_t)G.s0 * G.s1;
    const double *bp = b + t.i + (size_t)t.j * G.s0;
    for (int k = t.k0; k < t.k1; ++k) acc += bp[(size_t)k * sxy] * (dxy * G.c[2].dh[k]);
  }
  const double s = block_sum<kNT>(acc, sh);
  if (threadIdx.x == 0) partM[blockIdx.x] = s;
}

// modify_RHS_SF + PCG set-up (PCG_aux.f90:30-54; PCG_solver_algorithm.f90:57-66):
// r = vol*((b - mean) - A x_BC - A x), z = Minv r, p = z; partials of (r,r), (r,z).
template <bool MG, bool NRG>
__global__ void __launch_bounds__(kNT) k_pcg_rhs(const double *__restrict__ b, const double *__restrict__ x,
                                                double *__restrict__ r, double *__restrict__ p, CCGeom G,
                                                const PcgCtrl *ctrl, const double *partM, double *partB,
                                                int all_neumann, float *__restrict__ f0, F32Map M, NrgK K) {
  __shared__ double sh[kNT / kWave];
  double mean = 0.0;
  if (all_neumann) mean = prologue_sum<kNT>(partM, gridDim.x, 1, 0, sh) / ctrl->volsum;
  const TileIdx t = tile_of(G);
  double a_rr = 0.0, a_rz = 0.0;
  if (t.active) {
    const int i = t.i, j = t.j;
    const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
    const double dxy = G.c[0].dh[i] * G.c[1].dh[j];
    const double axy = G.c[0].lA[i] + G.c[0].lB[i] + G.c[1].lA[j] + G.c[1].lB[j];
    for (int k = t.k0; k < t.k1; ++k) {
      const size_t o = (size_t)i + j * sx + k * sxy;
      const double c = x[o];
      // neighbours of x with the implicit ghost rules, and of x_BC (zero interior,
      // ghosts from the BC values: compute_x_BC, PCG_aux.f90:109-115)
      double nb[6], nbc[6];
      const size_t off[6] = {o - 1, o + 1, o - sx, o + sx, o - sxy, o + sxy};
      const bool ghost[6] = {i == 1, i == G.n0, j == 1, j == G.n1, k == 1, k == G.n2};
      const size_t opp[6] = {o + (size_t)(G.n0 - 1), o - (size_t)(G.n0 - 1), o + (size_t)(G.n1 - 1) * sx,
                             o - (size_t)(G.n1 - 1) * sx, o + (size_t)(G.n2 - 1) * sxy, o - (size_t)(G.n2 - 1) * sxy};
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        if (!ghost[f]) {
          nb[f] = x[off[f]];
          nbc[f] = 0.0;
        } else {
          const int ty = G.bct[f];
          const double vmem = x[off[f]];
          const double vio = (ty == MOONS_BC_PERIODIC) ? x[opp[f]] : 0.0;
          nb[f] = ghost_cc(ty, true, 0.0, 0.0, G.robc[f], c, vio, vmem);
          nbc[f] = (ty == MOONS_BC_NONE) ? 0.0
                                         : ghost_cc(ty, false, G.bval[f], G.dh_nhat[f], G.robc[f], 0.0, 0.0, 0.0);
        }
      }
      const double axbc =
          NRG ? nrg7(K, G, 0.0, nbc[0], nbc[1], nbc[2], nbc[3], nbc[4], nbc[5], i, j, k)
              : lap7(0.0, nbc[0], nbc[1], nbc[2], nbc[3], nbc[4], nbc[5], G.c[0], i, G.c[1], j, G.c[2], k);
      const double ax = NRG ? nrg7(K, G, c, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], i, j, k)
                            : lap7(c, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], G.c[0], i, G.c[1], j, G.c[2], k);
      double rv = b[o];
      if (all_neumann) rv = rv - mean;
      rv = rv - axbc;
      rv = rv - ax;
      const double vol = dxy * G.c[2].dh[k];
      rv = rv * vol;
      r[o] = rv;
      a_rr += rv * rv;
      if (MG) {
        f0[f32_at(M, o, i, j, k)] = (float)(rv / vol);
      } else {
        const double z = minv_at<double>(G, dxy, axy, k) * rv;
        p[o] = z;
        fill_ghosts_implicit(p, G, i, j, k, z);
        a_rz += rv * z;
      }
    }
  }
  const double srr = block_sum<kNT>(a_rr, sh);
  const double srz = block_sum<kNT>(a_rz, sh);
  if (threadIdx.x == 0) {
    partB[2 * blockIdx.x] = srr;
    partB[2 * blockIdx.x + 1] = srz;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) const_cast<PcgCtrl *>(ctrl)->mean = mean;
}

// Set-up exit test (PCG_solver_algorithm.f90:60-68), one block:
// res0 = |r0|, rho = (r,z), res = sqrt|rho|; exit_loop(2) skips the loop.
__global__ void __launch_bounds__(kNT) k_pcg_init(PcgCtrl *ctrl, const double *partB, int nb) {
  __shared__ double sh[2 * (kNT / kWave) + 2];
  double rr, rz;
  prologue_sum2<kNT>(partB, 0, partB, 1, nb, 2, sh, rr, rz);
  if (threadIdx.x == 0) {
    const double res0 = sqrt(rr);
    const double res = sqrt(fabs(rz));
    ctrl->res0 = res0;
    ctrl->rho[0] = rz;
    ctrl->res = res;
    ctrl->iter = 0;
    const bool e1 = res / res0 < ctrl->tol_rel, e2 = res < ctrl->tol_abs, e3 = 0 >= ctrl->iter_max;
    ctrl->exit_loop = (e1 ? 1 : 0) | (e2 ? 2 : 0) | (e3 ? 4 : 0);
    ctrl->done = e2 ? 1 : 0;
    if (e2) ctrl->iter = 1; // the reference's else-branch counts one iteration
  }
}

// q = vol*A(p) over the interior (A = lap, or nrg_diffusion when NRG; p's
// face ghosts hold the implicit BC values) and partial (p,q).
template <int KZ, bool NRG>
__global__ void __launch_bounds__(kNT) k_pcg_matvec(const double *__restrict__ p, double *__restrict__ q, CCGeom G,
                                                   const PcgCtrl *ctrl, double *partA, NrgK K) {
  if (ctrl->done) return;
  __shared__ double sh[kNT / kWave];
  const TileIdx t = tile_of(G);
  double acc = 0.0;
  if (t.active) {
    const int i = t.i, j = t.j;
    const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
    const double dxy = G.c[0].dh[i] * G.c[1].dh[j];
    const double *pc = p + i + (size_t)j * sx;
    const int nk = t.k1 - t.k0; // 1 .. KZ planes, the same in the whole block
    // every load of the thread's planes first (planes past the slab read the
    // last one's and are not used), so that they are in flight together
    double pz[KZ + 2], nb[KZ][4];
    const St6<double> SX = st6(G.c[0], i), SY = st6(G.c[1], j);
#pragma unroll
    for (int u = 0; u < KZ + 2; ++u) pz[u] = pc[(size_t)min(t.k0 - 1 + u, t.k1) * sxy];
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      const size_t ok = (size_t)min(t.k0 + u, t.k1 - 1) * sxy;
      nb[u][0] = pc[ok - 1];
      nb[u][1] = pc[ok + 1];
      nb[u][2] = pc[ok - sx];
      nb[u][3] = pc[ok + sx];
    }
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      if (u < nk) {
        const int k = t.k0 + u;
        const double p0 = pz[u + 1];
        const double v = NRG ? nrg7(K, G, p0, nb[u][0], nb[u][1], nb[u][2], nb[u][3], pz[u], pz[u + 2], i, j, k)
                             : lap7s(p0, nb[u][0], nb[u][1], nb[u][2], nb[u][3], pz[u], pz[u + 2], SX, SY,
                                     st6(G.c[2], k));
        const double qv = v * (dxy * G.c[2].dh[k]);
        q[i + (size_t)j * sx + (size_t)k * sxy] = qv;
        acc += p0 * qv;
      }
    }
  }
  const double s = block_sum<kNT>(acc, sh);
  if (threadIdx.x == 0) partA[blockIdx.x] = s;
}

// alpha = rho/(p,q); x = x + p*alpha; r = r + q*(-alpha); partials (r,r), (Minv r, r).
template <int KZ, bool MG>
__global__ void __launch_bounds__(kNT) k_pcg_update(double *__restrict__ x, const double *__restrict__ p,
                                                   double *__restrict__ r, const double *__restrict__ q, CCGeom G,
                                                   PcgCtrl *ctrl, const double *partA, double *partB, int it,
                                                   float *__restrict__ f0, F32Map M) {
  if (ctrl->done) return;
  __shared__ double sh[kNT / kWave];
  const TileIdx t = tile_of(G);
  const int i = t.i, j = t.j;
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  const int nk = t.active ? t.k1 - t.k0 : 0;
  const size_t base = (size_t)i + (size_t)j * sx + (size_t)t.k0 * sxy;
  // stream loads first, so they are in flight during the prologue reduction
  double xv[KZ], pv[KZ], rv[KZ], qv[KZ];
#pragma unroll
  for (int u = 0; u < KZ; ++u) {
    if (u < nk) {
      const size_t o = base + (size_t)u * sxy;
      xv[u] = x[o];
      pv[u] = p[o];
      rv[u] = r[o];
      qv[u] = q[o];
    }
  }
  const double pq = prologue_sum<kNT>(partA, gridDim.x, 1, 0, sh);
  const double alpha = ctrl->rho[(it - 1) & 1] / pq;
  const double malpha = -alpha;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctrl->pq = pq;
    ctrl->alpha = alpha;
  }
  double a_rr = 0.0, a_rz = 0.0;
  if (t.active) {
    const double dxy = G.c[0].dh[i] * G.c[1].dh[j];
    const double axy = G.c[0].lA[i] + G.c[0].lB[i] + G.c[1].lA[j] + G.c[1].lB[j];
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      if (u < nk) {
        const size_t o = base + (size_t)u * sxy;
        x[o] = xv[u] + pv[u] * alpha;
        const double rn = rv[u] + qv[u] * malpha;
        r[o] = rn;
        a_rr += rn * rn;
        if (MG) {
          f0[f32_at(M, o, i, j, t.k0 + u)] = (float)(rn / (dxy * G.c[2].dh[t.k0 + u]));
        } else {
          const double z = minv_at<double>(G, dxy, axy, t.k0 + u) * rn;
          a_rz += z * rn;
        }
      }
    }
  }
  const double srr = block_sum<kNT>(a_rr, sh);
  const double srz = block_sum<kNT>(a_rz, sh);
  if (threadIdx.x == 0) {
    partB[2 * blockIdx.x] = srr;
    partB[2 * blockIdx.x + 1] = srz;
  }
}

// Exit decision after the x/r update of iteration `it` (check_res every
// n_skip, update_exit_loop: hw/var_set/iter_solver_params_extend.f90:86-113).
struct ExitDecision {
  bool stop, checked;
  int flags;
  double res;
};

__device__ __forceinline__ ExitDecision exit_decision(const PcgCtrl *ctrl, int it, double rr) {
  ExitDecision e{false, false, 0, 0.0};
  if (it % ctrl->n_skip == 0) {
    e.checked = true;
    e.res = sqrt(fabs(rr));
    const bool e1 = e.res / ctrl->res0 < ctrl->tol_rel, e2 = e.res < ctrl->tol_abs, e3 = it >= ctrl->iter_max;
    e.flags = (e1 ? 1 : 0) | (e2 ? 2 : 0) | (e3 ? 4 : 0);
    e.stop = e1 || e2 || e3;
  }
  if (!e.stop && it >= ctrl->iter_max) {
    e.flags |= 4;
    e.stop = true;
  }
  return e;
}

__device__ __forceinline__ void publish_exit(PcgCtrl *ctrl, int it, double rr, const ExitDecision &e) {
  ctrl->iter = it;
  ctrl->rr = rr;
  if (e.checked) {
    ctrl->res = e.res;
    ctrl->exit_loop = e.flags;
  } else if (e.stop) {
    ctrl->exit_loop |= e.flags;
  }
  if (e.stop) ctrl->done = 1;
}

// Exit test of iteration `it` (check_res every n_skip, update_exit_loop,
// hw/var_set/iter_solver_params_extend.f90:86-113), then p = p*beta + Minv r
// and the face ghosts of p.  Block 0 publishes the new state.
template <int KZ>
__global__ void __launch_bounds__(kNT) k_pcg_pupdate(double *__restrict__ p, const double *__restrict__ r, CCGeom G,
                                                    PcgCtrl *ctrl, const double *partB, int it) {
  if (ctrl->done) return;
  __shared__ double sh[2 * (kNT / kWave) + 2];
  const TileIdx t = tile_of(G);
  const int i = t.i, j = t.j;
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  const int nk = t.active ? t.k1 - t.k0 : 0;
  const size_t base = (size_t)i + (size_t)j * sx + (size_t)t.k0 * sxy;
  double rv[KZ], pv[KZ];
#pragma unroll
  for (int u = 0; u < KZ; ++u) {
    if (u < nk) {
      rv[u] = r[base + (size_t)u * sxy];
      pv[u] = p[base + (size_t)u * sxy];
    }
  }
  double rr, rz;
  prologue_sum2<kNT>(partB, 0, partB, 1, gridDim.x, 2, sh, rr, rz);
  const double rho_old = ctrl->rho[(it - 1) & 1];
  const ExitDecision e = exit_decision(ctrl, it, rr);
  const double beta = rz / rho_old;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    publish_exit(ctrl, it, rr, e);
    ctrl->rz = rz;
    ctrl->beta = beta;
    ctrl->rho[it & 1] = rz;
  }
  if (e.stop || !t.active) return;
  const double dxy = G.c[0].dh[i] * G.c[1].dh[j];
  const double axy = G.c[0].lA[i] + G.c[0].lB[i] + G.c[1].lA[j] + G.c[1].lB[j];
  const bool edge_ij = (i == 1) || (i == G.n0) || (j == 1) || (j == G.n1);
#pragma unroll
  for (int u = 0; u < KZ; ++u) {
    if (u < nk) {
      const int k = t.k0 + u;
      const double z = minv_at<double>(G, dxy, axy, k) * rv[u];
      const double pn = pv[u] * beta + z;
      p[base + (size_t)u * sxy] = pn;
      if (edge_ij || k == 1 || k == G.n2) fill_ghosts_implicit(p, G, i, j, k, pn);
    }
  }
}

// ----------------------------------------------------- multigrid-mode kernels
// The FP64 flexible PCG around the FP32 V-cycle: z = V(r/vol), rho = (z,r),
// beta = (z_new, r_new - r_old)/rho_old = -alpha (z_new, q)/rho_old
// (Polak-Ribiere form, exact for a fixed preconditioner and robust to the
// FP32 V-cycle's rounding).

// Exit test after the x/r update (one block; rr partials of k_pcg_update).
__global__ void __launch_bounds__(kNT) k_pcg_check(PcgCtrl *ctrl, const double *partB, int nb, int it) {
  if (ctrl->done) return;
  __shared__ double sh[kNT / kWave];
  const double rr = prologue_sum<kNT>(partB, nb, 2, 0, sh);
  const ExitDecision e = exit_decision(ctrl, it, rr);
  if (threadIdx.x == 0) publish_exit(ctrl, it, rr, e);
}

// Partials of (z,r) and (z,q) with z the V-cycle output.
template <int KZ>
__global__ void __launch_bounds__(kNT) k_mg_zdots(const float *__restrict__ z32, const double *__restrict__ r,
                                                 const double *__restrict__ q, CCGeom G, const PcgCtrl *ctrl,
                                                 double *partC, int use_q, F32Map M) {
  if (ctrl->done) return;
  __shared__ double sh[kNT / kWave];
  const TileIdx t = tile_of(G);
  double a_zr = 0.0, a_zq = 0.0;
  if (t.active) {
    const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
    const size_t base = (size_t)t.i + (size_t)t.j * sx + (size_t)t.k0 * sxy;
    const int nk = t.k1 - t.k0;
#pragma unroll
    for (int u = 0; u < KZ; ++u) {
      if (u < nk) {
        const size_t o = base + (size_t)u * sxy;
        const double z = (double)z32[f32_at(M, o, t.i, t.j, t.k0 + u)];
        a_zr += z * r[o];
        if (use_q) a_zq += z * q[o];
      }
    }
  }
  const double s0 = block_sum<kNT>(a_zr, sh);
  const double s1 = block_sum<kNT>(a_zq, sh);
  if (threadIdx.x == 0) {
    partC[2 * blockIdx.x] = s0;
    partC[2 * blockIdx.x + 1] = s1;
  }
}

// Set-up exit test with z = V(r0): res0 = |r0|, rho0 = (z,r0), res = sqrt|rho0|.
__global__ void __launch_bounds__(kNT) k_mg_init(PcgCtrl *ctrl, const double *partB, const double *partC, int nb) {
  __shared__ double sh[2 * (kNT / kWave) + 2];
  double rr, zr;
  prologue_sum2<kNT>(partB, 0, partC, 0, nb, 2, sh, rr, zr);
  if (threadIdx.x == 0) {
    const double res0 = sqrt(rr);
    const double res = sqrt(fabs(zr));
    ctrl->res0 = res0;
    ctrl->rho[0] = zr;
    ctrl->res = res;
    ctrl->iter = 0;
    const bool e1 = res / res0 < ctrl->tol_rel, e2 = res < ctrl->tol_abs, e3 = 0 >= ctrl->iter_max;
    ctrl->exit_loop = (e1 ? 1 : 0) | (e2 ? 2 : 0) | (e3 ? 4 : 0);
    ctrl->done = e2 ? 1 : 0;
    if (e2) ctrl->iter = 1;
  }
}

// p = p*beta + z (it >= 1) or p = z (it == 0), then the face ghosts of p.
template <int KZ>
__global__ void __launch_bounds__(kNT) k_mg_pupdate(double *__restrict__ p, const float *__restrict__ z32, CCGeom G,
                                                   PcgCtrl *ctrl, const double *partC, int it, F32Map M) {
  if (ctrl->done) return;
  __shared__ double sh[2 * (kNT / kWave) + 2];
  const TileIdx t = tile_of(G);
  const int i = t.i, j = t.j;
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  const int nk = t.active ? t.k1 - t.k0 : 0;
  const size_t base = (size_t)i + (size_t)j * sx + (size_t)t.k0 * sxy;
  float zv[KZ];
  double pv[KZ];
#pragma unroll
  for (int u = 0; u < KZ; ++u) {
    if (u < nk) {
      zv[u] = z32[f32_at(M, base + (size_t)u * sxy, i, j, t.k0 + u)];
      pv[u] = it > 0 ? p[base + (size_t)u * sxy] : 0.0;
    }
  }
  double beta = 0.0;
  if (it > 0) {
    double zr, zq;
    prologue_sum2<kNT>(partC, 0, partC, 1, gridDim.x, 2, sh, zr, zq);
    const double rho_old = ctrl->rho[(it - 1) & 1];
    beta = -ctrl->alpha * zq / rho_old;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctrl->rz = zr;
      ctrl->beta = beta;
      ctrl->rho[it & 1] = zr;
    }
  }
  if (!t.active) return;
  const bool edge_ij = (i == 1) || (i == G.n0) || (j == 1) || (j == G.n1);
#pragma unroll
  for (int u = 0; u < KZ; ++u) {
    if (u < nk) {
      const int k = t.k0 + u;
      const double z = (double)zv[u];
      const double pn = it > 0 ? pv[u] * beta + z : z;
      p[base + (size_t)u * sxy] = pn;
      if (edge_ij || k == 1 || k == G.n2) fill_ghosts_implicit(p, G, i, j, k, pn);
    }
  }
}

// f0 = r/vol (FP32) for one preconditioner application; z = x0 back to FP64.
__global__ void __launch_bounds__(kNT) k_mg_r2f(const double *__restrict__ r, float *__restrict__ f0, CCGeom G,
                                               F32Map M) {
  const TileIdx t = tile_of(G);
  if (!t.active) return;
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  const double dxy = G.c[0].dh[t.i] * G.c[1].dh[t.j];
  for (int k = t.k0; k < t.k1; ++k) {
    const size_t o = (size_t)t.i + (size_t)t.j * sx + (size_t)k * sxy;
    f0[f32_at(M, o, t.i, t.j, k)] = (float)(r[o] / (dxy * G.c[2].dh[k]));
  }
}
__global__ void __launch_bounds__(kNT) k_mg_x2z(const float *__restrict__ x0, double *__restrict__ z, CCGeom G,
                                               F32Map M) {
  const TileIdx t = tile_of(G);
  if (!t.active) return;
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  for (int k = t.k0; k < t.k1; ++k) {
    const size_t o = (size_t)t.i + (size_t)t.j * sx + (size_t)k * sxy;
    z[o] = (double)x0[f32_at(M, o, t.i, t.j, k)];
  }
}

// Fused Lap_uniform_SF for a cell-centred x whose ghosts already hold the
// implicit BC values: Ax = lap(x) in the interior.
__global__ void __launch_bounds__(kNT) k_lap_cc(const double *__restrict__ x, double *__restrict__ Ax, CCGeom G) {
  const TileIdx t = tile_of(G);
  if (!t.active) return;
  const int i = t.i, j = t.j;
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  const double *pc = x + i + (size_t)j * sx;
  double pm = pc[(size_t)(t.k0 - 1) * sxy];
  double p0 = pc[(size_t)t.k0 * sxy];
  for (int k = t.k0; k < t.k1; ++k) {
    const size_t ok = (size_t)k * sxy;
    const double pp = pc[ok + sxy];
    Ax[i + (size_t)j * sx + ok] = lap7(p0, pc[ok - 1], pc[ok + 1], pc[ok - sx], pc[ok + sx], pm, pp, G.c[0], i, G.c[1], j, G.c[2], k);
    pm = p0;
    p0 = pp;
  }
}

CCGeom make_geom(const moons_field_s *x, int kz) {
  CCTiles T = make_cc_tiles(x->s[0], x->s[1], x->s[2], kz);
  CCGeom G;
  G.s0 = T.s0; G.s1 = T.s1; G.s2 = T.s2;
  G.n0 = T.n0; G.n1 = T.n1; G.n2 = T.n2;
  G.ntx = T.ntx; G.nty = T.nty; G.ntz = T.ntz; G.kz = T.kz; G.nblocks = T.nblocks;
  const MeshDev &M = x->m->dev;
  for (int d = 0; d < 3; ++d) {
    G.c[d].cD = M.c[d].cc2n_D; G.c[d].cU = M.c[d].cc2n_U;
    G.c[d].nD = M.c[d].n2cc_D; G.c[d].nU = M.c[d].n2cc_U;
    G.c[d].dh = M.c[d].dhn;
    G.c[d].lA = M.c[d].lapA; G.c[d].lB = M.c[d].lapB;
  }
  BCDev b = make_bcdev(x);
  for (int f = 0; f < 6; ++f) {
    G.bct[f] = b.type[f];
    G.bval[f] = b.val[f];
    G.dh_nhat[f] = b.dh_nhat[f];
    G.robc[f] = b.robi[f];
  }
  return G;
}

int pick_kz(int n2) {
  // z cells marched per thread: enough blocks to fill 256 CUs several times
  // over.  MOONS_KZ (2, 4, 8 or 16) overrides the choice, for tuning sweeps.
  if (const char *e = getenv("MOONS_KZ")) {
    int v = atoi(e);
    if (v == 2 || v == 4 || v == 8 || v == 16) return v;
  }
  if (n2 >= 128) return 8;
  if (n2 >= 32) return 4;
  return 2;
}

} // namespace moons

using namespace moons;

struct moons_pcg_s {
  moons_mesh_s *m = nullptr;
  moons_field_s xbc; // copy of x's location and BCs (no storage)
  int op = 0, prec = 0;
  int iter_max = 1000, n_skip = 1;
  double tol_rel = 1e-6, tol_abs = 1e-13;
  double volsum = 0.0;
  CCGeom G;
  double *r = nullptr, *p = nullptr, *q = nullptr; // device work vectors
  double *partA = nullptr, *partB = nullptr, *partM = nullptr, *partC = nullptr;
  MGHier *mg = nullptr; // MOONS_PREC_MG_FP32 only
  UnetPrec *unet = nullptr;                   // MOONS_PREC_UNET: bound network
  bool jacobi_only = false;                   // set while a learned solve falls back to Jacobi
  bool fallback = true;                       // learned solves redo failures with Jacobi
  moons_unet_s *mg_net = nullptr;             // learned initial guess inside the V-cycle (MG solvers)
  int mg_net_level = -1;
  double *xbak = nullptr;                     // x at entry of a learned solve (for the fallback)
  float *uf0 = nullptr, *uz0 = nullptr;       // its FP32 input r/vol and output z (padded)
  int mg_max_levels = 16, mg_nu1 = 2, mg_nu2 = 2, mg_coarse = 16;
  double *kf[3] = {nullptr, nullptr, nullptr}; // nrg_diffusion: the solver's copy of k (PCG%k)
  double cn = 0.0;                             // nrg_diffusion: coeff_implicit_time_split
  PcgCtrl *ctrl = nullptr;
  PcgCtrl *host_ctrl = nullptr; // pinned, read-backs
  PcgCtrl *host_init = nullptr; // pinned, initial control block (the source of every upload)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool noalloc = false;         // a one-shot solver over the static scratch (no host_init)
};

// Iterations to launch before the next read-back of the control block: the
// count at which the relative residual, decreasing at its average rate so
// far, reaches tol_rel, within [lo, hi] (lo without a usable rate).  Each
// read-back idles the GPU for a host round trip, and each iteration launched
// after convergence still costs its kernels' launches, so the estimate aims
// at the iteration the exit test will stop at.
static int next_chunk(const PcgCtrl &h, double tol_rel, int lo, int hi) {
  int chunk = lo;
  if (h.iter >= 2 && h.res0 > 0.0 && tol_rel > 0.0) {
    const double rel = h.res / h.res0;
    if (rel > tol_rel && rel < 1.0) {
      const double rate = std::pow(rel, 1.0 / h.iter);
      if (rate < 0.999) chunk = (int)std::ceil(std::log(tol_rel / rel) / std::log(rate));
    }
  }
  return std::max(lo, std::min(hi, chunk));
}

// The initial control block written by a kernel (the allocation-free solve).
__global__ void k_pcg_ctrl_set(PcgCtrl *ctrl, int iter_max, int n_skip, double tol_rel, double tol_abs,
                               double volsum) {
  if (threadIdx.x != 0) return;
  PcgCtrl h;
  h.iter = h.done = h.exit_loop = h.pad_ = 0;
  h.iter_max = iter_max;
  h.n_skip = n_skip;
  h.tol_rel = tol_rel;
  h.tol_abs = tol_abs;
  h.rho[0] = h.rho[1] = 0.0;
  h.alpha = h.beta = h.res0 = h.res = h.mean = h.pq = h.rr = h.rz = 0.0;
  h.volsum = volsum;
  *ctrl = h;
}

// The initial control block lives in its own pinned buffer, written only
// here, so a captured copy of it replays the same values.
static int pcg_upload_ctrl(moons_pcg_s *S) {
  if (S->noalloc) {
    hipLaunchKernelGGL(k_pcg_ctrl_set, dim3(1), dim3(64), 0, stream(), S->ctrl, S->iter_max, S->n_skip, S->tol_rel,
                       S->tol_abs, S->volsum);
    MOONS_LAUNCH_CHECK();
    return MOONS_OK;
  }
  PcgCtrl h;
  std::memset(&h, 0, sizeof(PcgCtrl));
  h.iter_max = S->iter_max;
  h.n_skip = S->n_skip;
  h.tol_rel = S->tol_rel;
  h.tol_abs = S->tol_abs;
  h.volsum = S->volsum;
  if (std::memcmp(&h, S->host_init, sizeof(PcgCtrl)) != 0) {
    // a copy of the old contents may still be pending
    if (!sync_free()) MOONS_HIP(hipStreamSynchronize(stream()));
    *S->host_init = h;
  }
  MOONS_HIP(hipMemcpyAsync(S->ctrl, S->host_init, sizeof(PcgCtrl), hipMemcpyHostToDevice, stream()));
  return MOONS_OK;
}

// flexible (non-Jacobi) preconditioners: FP32 r/vol in, FP32 z out
static bool flex(const moons_pcg_s *S) {
  return !S->jacobi_only && (S->mg != nullptr || S->prec == MOONS_PREC_UNET);
}
static float *flex_f0(moons_pcg_s *S) { return S->mg ? mg_f0(S->mg) : S->uf0; }
static float *flex_z0(moons_pcg_s *S) { return S->mg ? mg_x0(S->mg) : S->uz0; }
static F32Map flex_map(const moons_pcg_s *S) {
  F32Map M{0, 0, 0, 0};
  if (S->mg) {
    int m[3];
    mg_f32_map(S->mg, m);
    M = F32Map{1, m[0], m[1], m[2]};
  }
  return M;
}
static int flex_apply(moons_pcg_s *S) {
  if (S->mg) return mg_vcycle(S->mg, &S->ctrl->done);
  return unet_prec_apply(S->unet, S->uf0, S->uz0, &S->ctrl->done);
}

static NrgK nrg_k(const moons_pcg_s *S) {
  NrgK K;
  for (int d = 0; d < 3; ++d) K.k[d] = S->kf[d];
  K.c = S->cn;
  return K;
}

template <int KZ>
static void pcg_launch_kz(moons_pcg_s *S, double *x, int it, int which) {
  const dim3 grid(S->G.nblocks), block(kNT);
  const bool mg = flex(S);
  float *f0 = mg ? flex_f0(S) : nullptr;
  const F32Map M = flex_map(S);
  const NrgK K = nrg_k(S);
  // the kernel timer times the plain (Jacobi) PPE's kernels only
  const int tm = (!mg && S->op == MOONS_OP_LAP_UNIFORM_SF) ? 1 : 0;
  if (which & 1) {
    ktimer_mark(tm * MOONS_KTIMER_PCG_MATVEC);
    if (S->op == MOONS_OP_NRG_DIFFUSION)
      hipLaunchKernelGGL((k_pcg_matvec<KZ, true>), grid, block, 0, stream(), S->p, S->q, S->G, S->ctrl, S->partA, K);
    else
      hipLaunchKernelGGL((k_pcg_matvec<KZ, false>), grid, block, 0, stream(), S->p, S->q, S->G, S->ctrl, S->partA, K);
    ktimer_mark(tm * MOONS_KTIMER_PCG_MATVEC);
  }
  if (which & 2) {
    ktimer_mark(tm * MOONS_KTIMER_PCG_UPDATE);
    if (mg)
      hipLaunchKernelGGL((k_pcg_update<KZ, true>), grid, block, 0, stream(), x, S->p, S->r, S->q, S->G, S->ctrl,
                         S->partA, S->partB, it, f0, M);
    else
      hipLaunchKernelGGL((k_pcg_update<KZ, false>), grid, block, 0, stream(), x, S->p, S->r, S->q, S->G, S->ctrl,
                         S->partA, S->partB, it, f0, M);
    ktimer_mark(tm * MOONS_KTIMER_PCG_UPDATE);
  }
  if (which & 4) {
    ktimer_mark(tm * MOONS_KTIMER_PCG_PUPDATE);
    hipLaunchKernelGGL(k_pcg_pupdate<KZ>, grid, block, 0, stream(), S->p, S->r, S->G, S->ctrl, S->partB, it);
    ktimer_mark(tm * MOONS_KTIMER_PCG_PUPDATE);
  }
  if (which & 8) hipLaunchKernelGGL(k_mg_zdots<KZ>, grid, block, 0, stream(), flex_z0(S), S->r, S->q, S->G, S->ctrl,
                                    S->partC, it > 0 ? 1 : 0, M);
  if (which & 16) hipLaunchKernelGGL(k_mg_pupdate<KZ>, grid, block, 0, stream(), S->p, flex_z0(S), S->G, S->ctrl,
                                     S->partC, it, M);
}

// which: bit 0 matvec, 1 update, 2 p-update (Jacobi), 3 (z,r),(z,q) dots, 4 p-update (multigrid)
static int pcg_launch(moons_pcg_s *S, double *x, int it, int which) {
  switch (S->G.kz) {
    case 16: pcg_launch_kz<16>(S, x, it, which); break;
    case 8: pcg_launch_kz<8>(S, x, it, which); break;
    case 4: pcg_launch_kz<4>(S, x, it, which); break;
    default: pcg_launch_kz<2>(S, x, it, which); break;
  }
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// One multigrid-preconditioned iteration: matvec, update (+ f0 = r/vol),
// exit test (the head), then V-cycle, (z,r),(z,q), p update (the tail).
static int pcg_launch_iteration_mg_head(moons_pcg_s *S, double *x, int it) {
  int st = pcg_launch(S, x, it, 1 | 2);
  if (st) return st;
  hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(kNT), 0, stream(), S->ctrl, S->partB, S->G.nblocks, it);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}
static int pcg_launch_iteration_mg_tail(moons_pcg_s *S, double *x, int it) {
  int st = flex_apply(S);
  if (st) return st;
  return pcg_launch(S, x, it, 8 | 16);
}
static int pcg_launch_iteration_mg(moons_pcg_s *S, double *x, int it) {
  int st = pcg_launch_iteration_mg_head(S, x, it);
  return st ? st : pcg_launch_iteration_mg_tail(S, x, it);
}

static int pcg_launch_iteration(moons_pcg_s *S, double *x, int it) { return pcg_launch(S, x, it, 7); }

extern "C" {

int moons_pcg_sf_create(moons_mesh_t m, int op, int prec, moons_field_t x, moons_pcg_t *out) {
  if (!stream()) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!m || !x || !out) return fail(MOONS_ERR_ARG, "moons_pcg_sf_create: null argument");
  const bool lap_ok = op == MOONS_OP_LAP_UNIFORM_SF &&
                      (prec == MOONS_PREC_LAP_SF || prec == MOONS_PREC_MG_FP32 || prec == MOONS_PREC_UNET);
  const bool nrg_ok = op == MOONS_OP_NRG_DIFFUSION && prec == MOONS_PREC_LAP_SF;
  if (!lap_ok && !nrg_ok)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_create: Lap_uniform_SF with prec_Lap_SF, the FP32 multigrid or the "
                               "U-net, or nrg_diffusion with prec_Lap_SF");
  if (x->dl != MOONS_DL_CC)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_create: the solver needs a cell-centred field");
  if (x->m != m) return fail(MOONS_ERR_ARG, "moons_pcg_sf_create: field is on another mesh");
  moons_pcg_s *S = new moons_pcg_s();
  S->m = m;
  S->xbc = *x;
  S->xbc.d = nullptr;
  S->op = op;
  S->prec = prec;
  S->G = make_geom(x, pick_kz(x->s[2] - 2));
  {
    // vol%vol (volume_SF, SF_extend.f90:571-581), sequential over interior cells
    const HostCoords *c = m->hc;
    double vs = 0.0;
    for (int k = 1; k < x->s[2] - 1; ++k)
      for (int j = 1; j < x->s[1] - 1; ++j)
        for (int i = 1; i < x->s[0] - 1; ++i) vs += c[0].dhn[i] * c[1].dhn[j] * c[2].dhn[k];
    S->volsum = vs;
  }
  const size_t bytes = x->n * sizeof(double);
  const size_t nb = (size_t)S->G.nblocks;
  bool ok = hipMalloc(&S->r, bytes) == hipSuccess && hipMalloc(&S->p, bytes) == hipSuccess &&
            hipMalloc(&S->q, bytes) == hipSuccess && hipMalloc(&S->partA, nb * sizeof(double)) == hipSuccess &&
            hipMalloc(&S->partB, 2 * nb * sizeof(double)) == hipSuccess &&
            hipMalloc(&S->partM, nb * sizeof(double)) == hipSuccess &&
            hipMalloc(&S->partC, 2 * nb * sizeof(double)) == hipSuccess &&
            hipMalloc(&S->ctrl, sizeof(PcgCtrl)) == hipSuccess &&
            hipHostMalloc(&S->host_ctrl, sizeof(PcgCtrl), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&S->host_init, sizeof(PcgCtrl), hipHostMallocDefault) == hipSuccess &&
            hipEventCreate(&S->ev0) == hipSuccess && hipEventCreate(&S->ev1) == hipSuccess;
  if (!ok) {
    moons_pcg_destroy(S);
    return fail(MOONS_ERR_HIP, "moons_pcg_sf_create: allocation failed");
  }
  std::memset(S->host_init, 0xff, sizeof(PcgCtrl)); // differs from any real block: first upload writes it
  (void)hipMemsetAsync(S->r, 0, bytes, stream());
  (void)hipMemsetAsync(S->p, 0, bytes, stream());
  (void)hipMemsetAsync(S->q, 0, bytes, stream());
  if (prec == MOONS_PREC_MG_FP32) {
    std::string err;
    S->mg = mg_create(x, S->mg_max_levels, S->mg_nu1, S->mg_nu2, S->mg_coarse, &err);
    if (!S->mg) {
      moons_pcg_destroy(S);
      return fail(MOONS_ERR_HIP, "moons_pcg_sf_create: " + err);
    }
  }
  if (prec == MOONS_PREC_UNET) {
    const size_t nf = (size_t)x->n;
    if (hipMalloc(&S->uf0, nf * sizeof(float)) != hipSuccess || hipMalloc(&S->uz0, nf * sizeof(float)) != hipSuccess) {
      moons_pcg_destroy(S);
      return fail(MOONS_ERR_HIP, "moons_pcg_sf_create: allocation failed");
    }
    (void)hipMemsetAsync(S->uf0, 0, nf * sizeof(float), stream());
    (void)hipMemsetAsync(S->uz0, 0, nf * sizeof(float), stream());
  }
  int st = pcg_upload_ctrl(S);
  if (st) {
    moons_pcg_destroy(S);
    return st;
  }
  *out = S;
  return MOONS_OK;
}

int moons_pcg_set_mg_unet(moons_pcg_t S, moons_unet_t net, int level) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_pcg_set_mg_unet: null solver");
  if (!S->mg) return fail(MOONS_ERR_STATE, "moons_pcg_set_mg_unet: solver was not created with MOONS_PREC_MG_FP32");
  std::string err;
  const int st = mg_set_learned(S->mg, net, level, &err);
  if (st) return fail(st, "moons_pcg_set_mg_unet: " + err);
  S->mg_net = net;
  S->mg_net_level = level;
  return MOONS_OK;
}

int moons_pcg_set_mg_unet_cycles(moons_pcg_t S, int cycles) {
  if (!S || cycles < 0) return fail(MOONS_ERR_ARG, "moons_pcg_set_mg_unet_cycles: bad argument");
  if (!S->mg) return fail(MOONS_ERR_STATE, "moons_pcg_set_mg_unet_cycles: solver was not created with MOONS_PREC_MG_FP32");
  mg_set_learned_cycles(S->mg, cycles);
  return MOONS_OK;
}

int moons_pcg_set_mg_gamma(moons_pcg_t S, int level, int cycles) {
  if (!S || cycles < 1) return fail(MOONS_ERR_ARG, "moons_pcg_set_mg_gamma: bad argument");
  if (!S->mg) return fail(MOONS_ERR_STATE, "moons_pcg_set_mg_gamma: solver was not created with MOONS_PREC_MG_FP32");
  if (mg_set_gamma(S->mg, level, cycles))
    return fail(MOONS_ERR_ARG, "moons_pcg_set_mg_gamma: level is not one of the grid-launched multigrid levels");
  return MOONS_OK;
}

int moons_pcg_set_fallback(moons_pcg_t S, int enable) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_pcg_set_fallback: null solver");
  S->fallback = enable != 0;
  return MOONS_OK;
}

int moons_pcg_set_unet(moons_pcg_t S, moons_unet_t net) {
  if (!S || !net) return fail(MOONS_ERR_ARG, "moons_pcg_set_unet: null argument");
  if (S->prec != MOONS_PREC_UNET) return fail(MOONS_ERR_STATE, "moons_pcg_set_unet: solver was not created with MOONS_PREC_UNET");
  if (S->unet) unet_prec_destroy(S->unet);
  std::string err;
  S->unet = unet_prec_create(net, &S->xbc, &err);
  if (!S->unet) return fail(MOONS_ERR_ARG, "moons_pcg_set_unet: " + err);
  return MOONS_OK;
}

int moons_pcg_destroy(moons_pcg_t S) {
  if (!S) return MOONS_OK;
  if (stream()) (void)hipStreamSynchronize(stream());
  if (S->mg) mg_destroy(S->mg);
  if (S->unet) unet_prec_destroy(S->unet);
  if (S->uf0) (void)hipFree(S->uf0);
  if (S->uz0) (void)hipFree(S->uz0);
  if (S->xbak) (void)hipFree(S->xbak);
  double *bufs[] = {S->r, S->p, S->q, S->partA, S->partB, S->partM, S->partC, S->kf[0], S->kf[1], S->kf[2]};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (S->ctrl) (void)hipFree(S->ctrl);
  if (S->host_ctrl) (void)hipHostFree(S->host_ctrl);
  if (S->host_init) (void)hipHostFree(S->host_init);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  delete S;
  return MOONS_OK;
}

int moons_pcg_sf_set_k(moons_pcg_t S, moons_field_t k[3]) {
  if (!S || !k) return fail(MOONS_ERR_ARG, "moons_pcg_sf_set_k: null argument");
  if (S->op != MOONS_OP_NRG_DIFFUSION) return fail(MOONS_ERR_STATE, "moons_pcg_sf_set_k: not an nrg_diffusion solver");
  for (int d = 0; d < 3; ++d)
    if (!k[d] || k[d]->m != S->m || k[d]->dl != (1 << d))
      return fail(MOONS_ERR_ARG, "moons_pcg_sf_set_k: k must be the x-, y- and z-face fields of the solver's mesh");
  for (int d = 0; d < 3; ++d) {
    if (!S->kf[d] && hipMalloc(&S->kf[d], k[d]->n * sizeof(double)) != hipSuccess)
      return fail(MOONS_ERR_HIP, "moons_pcg_sf_set_k: allocation failed");
    MOONS_HIP(hipMemcpyAsync(S->kf[d], k[d]->d, k[d]->n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  }
  return MOONS_OK;
}

int moons_pcg_sf_update_mfp(moons_pcg_t S, double coeff_implicit_time_split, int update_preconditioner) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_pcg_sf_update_mfp: null solver");
  if (S->op != MOONS_OP_NRG_DIFFUSION)
    return fail(MOONS_ERR_STATE, "moons_pcg_sf_update_mfp: Lap_uniform_SF has no implicit coefficient");
  // prec_Lap_SF does not depend on the coefficient: Minv is rebuilt in
  // registers from the 1D arrays either way
  (void)update_preconditioner;
  S->cn = coeff_implicit_time_split;
  return MOONS_OK;
}

int moons_pcg_set_isp(moons_pcg_t S, int iter_max, double tol_rel, double tol_abs, int n_skip_check_res) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_pcg_set_isp: null solver");
  if (iter_max < 1 || n_skip_check_res < 1) return fail(MOONS_ERR_ARG, "moons_pcg_set_isp: iter_max and n_skip must be >= 1");
  S->iter_max = iter_max;
  S->tol_rel = tol_rel;
  S->tol_abs = tol_abs;
  S->n_skip = n_skip_check_res;
  return MOONS_OK;
}



static int pcg_solve_impl(moons_pcg_s *S, moons_field_t x, moons_field_t b) {
  int st = pcg_upload_ctrl(S);
  if (st) return st;
  if (S->mg) mg_begin_solve(S->mg);
  const dim3 grid(S->G.nblocks), block(kNT);
  const int alln = S->xbc.all_neumann ? 1 : 0;
  const bool mg = flex(S);
  if (alln) hipLaunchKernelGGL(k_pcg_mean, grid, block, 0, stream(), b->d, S->G, S->partM);
  if (!mg) {
    if (S->op == MOONS_OP_NRG_DIFFUSION)
      hipLaunchKernelGGL((k_pcg_rhs<false, true>), grid, block, 0, stream(), b->d, x->d, S->r, S->p, S->G, S->ctrl,
                         S->partM, S->partB, alln, (float *)nullptr, flex_map(S), nrg_k(S));
    else
      hipLaunchKernelGGL((k_pcg_rhs<false, false>), grid, block, 0, stream(), b->d, x->d, S->r, S->p, S->G, S->ctrl,
                         S->partM, S->partB, alln, (float *)nullptr, flex_map(S), nrg_k(S));
    hipLaunchKernelGGL(k_pcg_init, dim3(1), block, 0, stream(), S->ctrl, S->partB, S->G.nblocks);
    MOONS_LAUNCH_CHECK();
  } else {
    // z0 = V(r0/vol), rho0 = (z0,r0), p = z0
    hipLaunchKernelGGL((k_pcg_rhs<true, false>), grid, block, 0, stream(), b->d, x->d, S->r, S->p, S->G, S->ctrl,
                       S->partM, S->partB, alln, flex_f0(S), flex_map(S), nrg_k(S));
    MOONS_LAUNCH_CHECK();
    if ((st = flex_apply(S))) return st;
    if ((st = pcg_launch(S, x->d, 0, 8))) return st;
    hipLaunchKernelGGL(k_mg_init, dim3(1), block, 0, stream(), S->ctrl, S->partB, S->partC, S->G.nblocks);
    MOONS_LAUNCH_CHECK();
    if ((st = pcg_launch(S, x->d, 0, 16))) return st;
  }
  if (sync_free()) {
    // exactly iter_max iterations, gated on the device; nothing read back
    for (int it = 1; it <= S->iter_max; ++it) {
      st = mg ? pcg_launch_iteration_mg(S, x->d, it) : pcg_launch_iteration(S, x->d, it);
      if (st) return st;
    }
    return launch_apply_bcs(x, false);
  }
  // Chunks of iterations; every kernel returns at once after the exit decision
  // (the multigrid's kernels only skip their stores, so a V-cycle after the
  // exit costs nearly a full one).  The multigrid iteration that ends a chunk
  // launches only its head (matvec, update, exit test); its tail (V-cycle,
  // (z,r),(z,q), p update) waits for the exit flag and is launched before the
  // next chunk only if the solve goes on.
  int launched = 0;
  bool tail_pending = false;
  for (;;) {
    MOONS_HIP(hipMemcpyAsync(S->host_ctrl, S->ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, stream()));
    MOONS_HIP(hipStreamSynchronize(stream()));
    if (S->host_ctrl->done || launched >= S->iter_max) break;
    if (tail_pending) {
      if ((st = pcg_launch_iteration_mg_tail(S, x->d, launched))) return st;
      tail_pending = false;
    }
    int chunk = launched == 0 ? (mg ? 3 : 16)
                              : next_chunk(*S->host_ctrl, S->tol_rel, mg ? 1 : 4, mg ? 32 : 512);
    chunk = std::min(chunk, S->iter_max - launched);
    for (int c = 0; c < chunk; ++c) {
      const int it = launched + c + 1;
      if (mg && c == chunk - 1) {
        st = pcg_launch_iteration_mg_head(S, x->d, it);
        tail_pending = true;
      } else {
        st = mg ? pcg_launch_iteration_mg(S, x->d, it) : pcg_launch_iteration(S, x->d, it);
      }
      if (st) return st;
    }
    launched += chunk;
  }
  // apply_BCs(x) after the last update (PCG_solver_algorithm.f90:77,98)
  st = launch_apply_bcs(x, false);
  if (st) return st;
  MOONS_HIP(hipMemcpyAsync(S->host_ctrl, S->ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, stream()));
  MOONS_HIP(hipStreamSynchronize(stream()));
  return MOONS_OK;
}

int moons_pcg_sf_solve(moons_pcg_t S, moons_field_t x, moons_field_t b, moons_pcg_info *info) {
  if (!S || !x || !b) return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve: null argument");
  if (x->m != S->m || b->m != S->m || x->dl != S->xbc.dl || b->dl != S->xbc.dl)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve: x and b must match the solver's mesh and location");
  if (S->prec == MOONS_PREC_UNET && !S->unet)
    return fail(MOONS_ERR_STATE, "moons_pcg_sf_solve: bind a network with moons_pcg_set_unet first");
  if (S->op == MOONS_OP_NRG_DIFFUSION && !S->kf[0])
    return fail(MOONS_ERR_STATE, "moons_pcg_sf_solve: set k with moons_pcg_sf_set_k first");
  const bool learned = S->prec == MOONS_PREC_UNET;
  if (learned) {
    if (!S->xbak && hipMalloc(&S->xbak, x->n * sizeof(double)) != hipSuccess)
      return fail(MOONS_ERR_HIP, "moons_pcg_sf_solve: allocation failed");
    MOONS_HIP(hipMemcpyAsync(S->xbak, x->d, x->n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  }
  if (sync_free()) {
    if (learned) return fail(MOONS_ERR_STATE, "moons_pcg_sf_solve: the learned solver cannot run inside a graph");
    S->jacobi_only = false;
    int st = pcg_solve_impl(S, x, b);
    if (st) return st;
    if (info) {
      info->iterations = -1;
      info->exit_loop = 0;
      info->res_norm0 = info->res_norm = info->ms = 0.0;
    }
    return MOONS_OK;
  }
  MOONS_HIP(hipEventRecord(S->ev0, stream()));
  S->jacobi_only = false;
  int st = pcg_solve_impl(S, x, b);
  if (st) return st;
  int fell_back = 0;
  if (learned && S->fallback) {
    // Jacobi-PCG is the fallback and the judge: a learned solve that breaks
    // down (NaN) or stops at iter_max without meeting a tolerance is redone
    // from the same x with the Jacobi preconditioner.
    const PcgCtrl &h = *S->host_ctrl;
    const bool bad = std::isnan(h.res) || std::isnan(h.res0) || std::isinf(h.res) || !(h.exit_loop & 3);
    if (bad) {
      MOONS_HIP(hipMemcpyAsync(x->d, S->xbak, x->n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
      S->jacobi_only = true;
      st = pcg_solve_impl(S, x, b);
      S->jacobi_only = false;
      if (st) return st;
      fell_back = 8;
    }
  }
  MOONS_HIP(hipEventRecord(S->ev1, stream()));
  MOONS_HIP(hipEventSynchronize(S->ev1));
  const PcgCtrl &h = *S->host_ctrl;
  if (std::isnan(h.res) || std::isnan(h.res0))
    return fail(MOONS_ERR_NAN, "Error: NaN in PCG residual (check_nans, PCG_solver_algorithm.f90:223-248)");
  if (info) {
    info->iterations = h.iter;
    info->exit_loop = h.exit_loop | fell_back;
    info->res_norm0 = h.res0;
    info->res_norm = h.res;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, S->ev0, S->ev1);
    info->ms = ms;
  }
  return MOONS_OK;
}

int moons_lap_uniform_sf(moons_field_t Ax, moons_field_t x) {
  if (!Ax || !x) return fail(MOONS_ERR_ARG, "moons_lap_uniform_sf: null argument");
  if (Ax->dl != x->dl || Ax->m != x->m) return fail(MOONS_ERR_ARG, "moons_lap_uniform_sf: location mismatch");
  int st = launch_apply_bcs(x, true);
  if (st) return st;
  bool ax_periodic = false;
  for (int f = 0; f < 6; ++f) ax_periodic |= Ax->has_bcs && Ax->bct[f] == MOONS_BC_PERIODIC;
  if (x->dl == MOONS_DL_CC && !ax_periodic) {
    CCGeom G = make_geom(x, pick_kz(x->s[2] - 2));
    MOONS_HIP(hipMemsetAsync(Ax->d, 0, Ax->n * sizeof(double), stream()));
    hipLaunchKernelGGL(k_lap_cc, dim3(G.nblocks), dim3(kNT), 0, stream(), x->d, Ax->d, G);
    MOONS_LAUNCH_CHECK();
    return MOONS_OK;
  }
  // General location: the reference composition grad then div through del.
  moons_field_s *g[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < 3 && !st; ++d) st = moons_field_create(x->m, x->dl ^ (1 << d), &g[d]);
  if (!st) st = moons_grad(g[0], g[1], g[2], x);
  if (!st) st = moons_div(Ax, g[0], g[1], g[2]);
  for (int d = 0; d < 3; ++d) moons_field_destroy(g[d]);
  return st;
}

int moons_clean_div(moons_pcg_t S, moons_field_t X[3], moons_field_t Xstar[3], moons_field_t phi, double scale,
                    moons_field_t temp_F[3], moons_field_t temp_CC, moons_pcg_info *info) {
  if (!S || !X || !Xstar || !phi || !temp_F || !temp_CC) return fail(MOONS_ERR_ARG, "moons_clean_div: null argument");
  for (int d = 0; d < 3; ++d)
    if (!X[d] || !Xstar[d] || !temp_F[d]) return fail(MOONS_ERR_ARG, "moons_clean_div: null component");
  int st;
  if ((st = moons_div(temp_CC, Xstar[0], Xstar[1], Xstar[2]))) return st;
  if ((st = launch_scale(temp_CC, scale))) return st;
  if ((st = moons_pcg_sf_solve(S, phi, temp_CC, info))) return st;
  if ((st = moons_grad(temp_F[0], temp_F[1], temp_F[2], phi))) return st;
  const double inv = 1.0 / scale;
  for (int d = 0; d < 3; ++d) {
    if ((st = launch_scale(temp_F[d], inv))) return st;
    if ((st = launch_sub(X[d], Xstar[d], temp_F[d]))) return st;
  }
  for (int d = 0; d < 3; ++d)
    if ((st = launch_apply_bcs(X[d], false))) return st;
  return MOONS_OK;
}

int moons_pcg_set_mg(moons_pcg_t S, int max_levels, int nu1, int nu2, int coarse_sweeps) {
  if (!S) return fail(MOONS_ERR_ARG, "moons_pcg_set_mg: null solver");
  if (!S->mg) return fail(MOONS_ERR_STATE, "moons_pcg_set_mg: solver was not created with MOONS_PREC_MG_FP32");
  if (max_levels < 1 || nu1 < 1 || nu2 < 0 || coarse_sweeps < 1)
    return fail(MOONS_ERR_ARG, "moons_pcg_set_mg: need max_levels >= 1, nu1 >= 1, nu2 >= 0, coarse_sweeps >= 1");
  mg_destroy(S->mg);
  S->mg_max_levels = max_levels;
  S->mg_nu1 = nu1;
  S->mg_nu2 = nu2;
  S->mg_coarse = coarse_sweeps;
  std::string err;
  S->mg = mg_create(&S->xbc, max_levels, nu1, nu2, coarse_sweeps, &err);
  if (!S->mg) return fail(MOONS_ERR_HIP, "moons_pcg_set_mg: " + err);
  if (S->mg_net && mg_set_learned(S->mg, S->mg_net, S->mg_net_level, &err)) {
    S->mg_net = nullptr;
    S->mg_net_level = -1;
    return fail(MOONS_ERR_ARG, "moons_pcg_set_mg: the new hierarchy cannot hold the learned level: " + err);
  }
  return MOONS_OK;
}

int moons_pcg_mg_info(moons_pcg_t S, int *levels, int n_coarsest[3]) {
  if (!S || !S->mg) return fail(MOONS_ERR_STATE, "moons_pcg_mg_info: not a multigrid solver");
  const int nl = mg_levels(S->mg);
  if (levels) *levels = nl;
  if (n_coarsest) mg_level_cells(S->mg, nl - 1, n_coarsest);
  return MOONS_OK;
}

int moons_pcg_apply_prec(moons_pcg_t S, moons_field_t r, moons_field_t z) {
  if (!S || !r || !z) return fail(MOONS_ERR_ARG, "moons_pcg_apply_prec: null argument");
  if (r->dl != S->xbc.dl || z->dl != S->xbc.dl || r->m != S->m || z->m != S->m)
    return fail(MOONS_ERR_ARG, "moons_pcg_apply_prec: fields must match the solver");
  const dim3 grid(S->G.nblocks), block(kNT);
  if (flex(S)) {
    if (S->prec == MOONS_PREC_UNET && !S->unet)
      return fail(MOONS_ERR_STATE, "moons_pcg_apply_prec: bind a network with moons_pcg_set_unet first");
    int st = pcg_upload_ctrl(S); // done = 0
    if (st) return st;
    hipLaunchKernelGGL(k_mg_r2f, grid, block, 0, stream(), r->d, flex_f0(S), S->G, flex_map(S));
    MOONS_LAUNCH_CHECK();
    if ((st = flex_apply(S))) return st;
    MOONS_HIP(hipMemsetAsync(z->d, 0, z->n * sizeof(double), stream()));
    hipLaunchKernelGGL(k_mg_x2z, grid, block, 0, stream(), flex_z0(S), z->d, S->G, flex_map(S));
    MOONS_LAUNCH_CHECK();
    return MOONS_OK;
  }
  return fail(MOONS_ERR_ARG, "moons_pcg_apply_prec: the Jacobi preconditioner is applied inside the PCG kernels only");
}

int moons_pcg_time_mg(moons_pcg_t S, int reps, double ms[10]) {
  if (!S || !ms || reps < 1) return fail(MOONS_ERR_ARG, "moons_pcg_time_mg: bad argument");
  if (!S->mg) return fail(MOONS_ERR_ARG, "moons_pcg_time_mg: the solver has no multigrid preconditioner");
  return mg_time_kernels(S->mg, reps, ms);
}

int moons_pcg_time_kernels(moons_pcg_t S, int reps, double ms[3]) {
  if (!S || !ms || reps < 1) return fail(MOONS_ERR_ARG, "moons_pcg_time_kernels: bad argument");
  if (S->op == MOONS_OP_NRG_DIFFUSION && !S->kf[0])
    return fail(MOONS_ERR_STATE, "moons_pcg_time_kernels: set k with moons_pcg_sf_set_k first");
  // One plain iteration sets up p, q, r and the partial sums; then each
  // kernel is launched reps times back to back between two events (the exit
  // test disabled, a scratch x updated), so the average is the kernel's
  // duration plus the dispatch gap between consecutive launches.
  double *xs = nullptr;
  const size_t bytes = (size_t)S->G.s0 * S->G.s1 * S->G.s2 * sizeof(double);
  MOONS_HIP(hipMalloc(&xs, bytes));
  MOONS_HIP(hipMemsetAsync(xs, 0, bytes, stream()));
  PcgCtrl &h = *S->host_ctrl;
  std::memset(&h, 0, sizeof(PcgCtrl));
  h.iter_max = 1 << 30;
  h.n_skip = 1 << 30;
  h.rho[0] = h.rho[1] = 1.0;
  h.res0 = 1.0;
  MOONS_HIP(hipMemcpyAsync(S->ctrl, &h, sizeof(PcgCtrl), hipMemcpyHostToDevice, stream()));
  hipEvent_t e[2];
  for (int i = 0; i < 2; ++i) MOONS_HIP(hipEventCreate(&e[i]));
  int st;
  if ((st = pcg_launch(S, xs, 1, 1 | 2 | 4))) return st;
  for (int kk = 0; kk < 3; ++kk) {
    // warm-up: as many launches as are timed (the first measurement after
    // other work otherwise catches the device's clock and cache transients)
    for (int rep = 0; rep < reps; ++rep)
      if ((st = pcg_launch(S, xs, 2, 1 << kk))) return st;
    MOONS_HIP(hipEventRecord(e[0], stream()));
    for (int rep = 0; rep < reps; ++rep)
      if ((st = pcg_launch(S, xs, 2, 1 << kk))) return st;
    MOONS_HIP(hipEventRecord(e[1], stream()));
    MOONS_HIP(hipEventSynchronize(e[1]));
    float tms = 0.f;
    MOONS_HIP(hipEventElapsedTime(&tms, e[0], e[1]));
    ms[kk] = tms / reps;
  }
  for (int i = 0; i < 2; ++i) (void)hipEventDestroy(e[i]);
  (void)hipFree(xs);
  return MOONS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// The allocation-free PPE solve: r, p, q, the block partial sums and the
// control block of one Jacobi-PCG solve live in this file-scope array (about
// 15.7 MB on every GPU the library runs on, held until the process ends).
// Host code reaches it only through the per-device address taken at
// moons_init (hipGetSymbolAddress); the array's name is not a device address
// in host code.
namespace {
constexpr size_t kNaCells = size_t(1) << 19;   // padded cells of x at most (e.g. 78^3 interior)
constexpr size_t kNaBlocks = size_t(1) << 16;  // tiles of the solve's launches at most
constexpr size_t kNaCtrl = 32;                 // doubles reserved for the control block
constexpr size_t kNaDoubles = kNaCtrl + 6 * kNaBlocks + 3 * kNaCells;
constexpr int kNaMaxDevices = 64;
static_assert(sizeof(PcgCtrl) <= kNaCtrl * sizeof(double), "control block does not fit its slot");

static __device__ __align__(256) double scratch[kNaDoubles];
double *g_scratch_dev[kNaMaxDevices] = {};  // address of scratch on each device

__global__ void __launch_bounds__(256) k_na_zero(double *a, size_t n) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) a[t] = 0.0;
}
} // namespace

namespace moons {
// Called by moons_init with the device current: the device address of scratch.
int pcg_noalloc_setup(int device) {
  if (device < 0 || device >= kNaMaxDevices) return fail(MOONS_ERR_ARG, "moons_init: device index out of range");
  if (g_scratch_dev[device]) return MOONS_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(MOONS_ERR_HIP, std::string("moons_init: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_dev[device] = static_cast<double *>(ptr);
  return MOONS_OK;
}
} // namespace moons

extern "C" {

int moons_pcg_sf_solve_noalloc(moons_field_t x, moons_field_t b, int iter_max, double tol_rel, double tol_abs,
                               int n_skip_check_res, moons_pcg_info *info) {
  if (!stream()) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!x || !b) return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve_noalloc: null argument");
  if (sync_free()) return fail(MOONS_ERR_STATE, "moons_pcg_sf_solve_noalloc: not inside a graph capture");
  if (x->dl != MOONS_DL_CC || b->dl != MOONS_DL_CC || x->m != b->m)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve_noalloc: x and b must be cell-centred fields of one mesh");
  if (iter_max < 1 || n_skip_check_res < 1)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve_noalloc: iter_max and n_skip must be >= 1");
  if (x->n > kNaCells)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve_noalloc: the field has " + std::to_string(x->n) +
                                   " padded cells, more than the scratch's " + std::to_string(kNaCells));
  int dev = -1;
  MOONS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices || !g_scratch_dev[dev])
    return fail(MOONS_ERR_STATE, "moons_pcg_sf_solve_noalloc: no scratch on this device (call moons_init)");
  moons_pcg_s S;
  S.noalloc = true;
  S.m = x->m;
  S.xbc = *x;
  S.xbc.d = nullptr;
  S.op = MOONS_OP_LAP_UNIFORM_SF;
  S.prec = MOONS_PREC_LAP_SF;
  S.iter_max = iter_max;
  S.tol_rel = tol_rel;
  S.tol_abs = tol_abs;
  S.n_skip = n_skip_check_res;
  S.G = make_geom(x, pick_kz(x->s[2] - 2));
  const size_t nb = (size_t)S.G.nblocks;
  if (nb > kNaBlocks)
    return fail(MOONS_ERR_ARG, "moons_pcg_sf_solve_noalloc: the field needs " + std::to_string(nb) +
                                   " tiles, more than the scratch's " + std::to_string(kNaBlocks));
  {
    const HostCoords *c = x->m->hc;
    double vs = 0.0;
    for (int k = 1; k < x->s[2] - 1; ++k)
      for (int j = 1; j < x->s[1] - 1; ++j)
        for (int i = 1; i < x->s[0] - 1; ++i) vs += c[0].dhn[i] * c[1].dhn[j] * c[2].dhn[k];
    S.volsum = vs;
  }
  double *base = g_scratch_dev[dev];
  S.ctrl = reinterpret_cast<PcgCtrl *>(base);
  S.partA = base + kNaCtrl;
  S.partB = S.partA + kNaBlocks;
  S.partM = S.partB + 2 * kNaBlocks;
  S.partC = S.partM + kNaBlocks;
  S.r = S.partC + 2 * kNaBlocks;
  S.p = S.r + kNaCells;
  S.q = S.p + kNaCells;
  PcgCtrl hc;
  std::memset(&hc, 0, sizeof hc);
  S.host_ctrl = &hc;
  // the previous call's vectors stay in the scratch: start from zero, as a new solver does
  hipLaunchKernelGGL(k_na_zero, dim3((unsigned)((3 * kNaCells + 255) / 256)), dim3(256), 0, stream(), S.r,
                     3 * kNaCells);
  MOONS_LAUNCH_CHECK();
  const auto t0 = std::chrono::steady_clock::now();
  const int st = pcg_solve_impl(&S, x, b);
  if (st) return st;
  MOONS_HIP(hipStreamSynchronize(stream()));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (std::isnan(hc.res) || std::isnan(hc.res0))
    return fail(MOONS_ERR_NAN, "Error: NaN in PCG residual (check_nans, PCG_solver_algorithm.f90:223-248)");
  if (info) {
    info->iterations = hc.iter;
    info->exit_loop = hc.exit_loop;
    info->res_norm0 = hc.res0;
    info->res_norm = hc.res;
    info->ms = ms;
  }
  return MOONS_OK;
}

} // extern "C"
