// cc_stencil.hpp -- cell-centred 7-point Laplacian pieces shared by the
// FP64 PCG and the multigrid kernels: per-direction coefficient pointers,
// XCD-aware z-marching tiles, the stencil in the reference's operation order,
// face-ghost rules and the on-the-fly Jacobi inverse.
#pragma once

#include "common.hpp"

namespace moons {

template <typename T>
struct Coef1T {
  const T *cD, *cU; // staggered CC->N (sc-1)
  const T *nD, *nU; // staggered N->CC (sn-1)
  const T *dh;      // dhn (cell widths, the CC dual spacing)
  const T *lA, *lB; // Laplacian diagonal terms (0 at ghosts)
};

template <typename T>
struct CCGeomT {
  int s0, s1, s2, n0, n1, n2;
  int ntx, nty, ntz, kz, nblocks;
  Coef1T<T> c[3];
  int bct[6];
  double bval[6];
  double dh_nhat[6];
  double robc[6]; // Robin ghost factor of a cell-centred field (direct and implicit)
};

using Coef1 = Coef1T<double>;
using CCGeom = CCGeomT<double>;

// 7-point cell-centred Laplacian in the reference's operation order:
// gx(node) = f(i-1)*D + f(i)*U, div = gx(i)*D' + gx(i+1)*U', then the y and z
// terms added left to right (code/handwritten/stencils/stencils.f90:47-59,108-120,
// composed as in lap_centered_SF_new, code/handwritten/ops/ops_discrete.f90:124-132).
// The six stencil coefficients of one direction at index i: CC->N at i - 1
// and i, N->CC at i (loaded once where a thread reuses them).
template <typename T>
struct St6 {
  T cDm, cUm, cD, cU, nD, nU;
};
template <typename T>
__device__ __forceinline__ St6<T> st6(const Coef1T<T> &X, int i) {
  return St6<T>{X.cD[i - 1], X.cU[i - 1], X.cD[i], X.cU[i], X.nD[i], X.nU[i]};
}

template <typename T>
__device__ __forceinline__ T lap7s(T c, T xm, T xp, T ym, T yp, T zm, T zp, const St6<T> &X, const St6<T> &Y,
                                   const St6<T> &Z) {
  const T gxm = xm * X.cDm + c * X.cUm;
  const T gxp = c * X.cD + xp * X.cU;
  T v = gxm * X.nD + gxp * X.nU;
  const T gym = ym * Y.cDm + c * Y.cUm;
  const T gyp = c * Y.cD + yp * Y.cU;
  v = v + gym * Y.nD + gyp * Y.nU;
  const T gzm = zm * Z.cDm + c * Z.cUm;
  const T gzp = c * Z.cD + zp * Z.cU;
  v = v + gzm * Z.nD + gzp * Z.nU;
  return v;
}

template <typename T>
__device__ __forceinline__ T lap7(T c, T xm, T xp, T ym, T yp, T zm, T zp, const Coef1T<T> &X, int i,
                                  const Coef1T<T> &Y, int j, const Coef1T<T> &Z, int k) {
  return lap7s(c, xm, xp, ym, yp, zm, zp, st6(X, i), st6(Y, j), st6(Z, k));
}

// Face-ghost value of a cell-centred field (never an edge or corner ghost).
// implicit_: homogeneous rules; else the BC value rule.  vI: first interior
// value behind the ghost; vIo: first interior value behind the opposite face
// (periodic); vmem: the value stored in the ghost (faces without BCs); robc:
// the face's Robin factor (the direct rule's value term has factor 0).
__device__ __forceinline__ double ghost_cc(int type, bool implicit_, double bval, double dh_nhat, double robc,
                                           double vI, double vIo, double vmem) {
  switch (type) {
    case MOONS_BC_DIRICHLET: return implicit_ ? -vI : 2.0 * bval - vI;
    case MOONS_BC_NEUMANN: return implicit_ ? vI : vI + bval * dh_nhat;
    case MOONS_BC_PERIODIC: return vIo;
    case MOONS_BC_SYMMETRIC: return vI;
    case MOONS_BC_ANTISYMMETRIC: return -vI;
    case MOONS_BC_ROBIN: return implicit_ ? vI * robc : vI * robc + bval * 0.0;
    default: return vmem;
  }
}

struct TileIdx {
  int i, j, k0, k1;
  bool active;
};

template <typename G_>
__device__ __forceinline__ TileIdx tile_of(const G_ &G) {
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = L % G.ntx;
  const int tj = (L / G.ntx) % G.nty;
  const int tk = L / (G.ntx * G.nty);
  TileIdx t;
  t.i = 1 + ti * kTX + (threadIdx.x % kTX);
  t.j = 1 + tj * kTY + (threadIdx.x / kTX);
  t.k0 = 1 + tk * G.kz;
  t.k1 = min(t.k0 + G.kz, G.n2 + 1);
  t.active = (t.i <= G.n0) && (t.j <= G.n1);
  return t;
}

// Minv = 1/(diag*vol) of prec_Lap_SF (preconditioners.f90:55-71), with
// diag summed in laplacian_diagonal's term order (GF_diagonals.f90:15-35).
template <typename T, typename G_>
__device__ __forceinline__ T minv_at(const G_ &G, T dxy, T axy, int k) {
  const T diag = axy + G.c[2].lA[k] + G.c[2].lB[k];
  const T vol = dxy * G.c[2].dh[k];
  return T(1) / (diag * vol);
}

// Writes the face ghosts of p next to interior cell (i,j,k) with the implicit
// (homogeneous) BC rules, which apply_BCs_implicit sets before each operator
// application (code/handwritten/solvers/matrix_free_operators.f90:60).
// Periodic: the thread at the far interior plane feeds the near ghost.
template <typename T, typename G_>
__device__ __forceinline__ void fill_ghosts_implicit(T *p, const G_ &G, int i, int j, int k, T v) {
  const size_t sx = G.s0, sxy = (size_t)G.s0 * G.s1;
  auto put = [&](int face, size_t at, size_t at_opp) {
    const int t = G.bct[face];
    if (t == MOONS_BC_NONE) return;
    if (t == MOONS_BC_PERIODIC) p[at_opp] = v;
    else if (t == MOONS_BC_DIRICHLET || t == MOONS_BC_ANTISYMMETRIC) p[at] = -v;
    else if (t == MOONS_BC_ROBIN) p[at] = v * (T)G.robc[face];
    else p[at] = v;
  };
  const size_t base = (size_t)j * sx + (size_t)k * sxy;
  if (i == 1) put(0, base, base + (size_t)(G.n0 + 1));
  if (i == G.n0) put(1, base + (size_t)(G.n0 + 1), base);
  const size_t bi = (size_t)i + (size_t)k * sxy;
  if (j == 1) put(2, bi, bi + (size_t)(G.n1 + 1) * sx);
  if (j == G.n1) put(3, bi + (size_t)(G.n1 + 1) * sx, bi);
  const size_t bj = (size_t)i + (size_t)j * sx;
  if (k == 1) put(4, bj, bj + (size_t)(G.n2 + 1) * sxy);
  if (k == G.n2) put(5, bj + (size_t)(G.n2 + 1) * sxy, bj);
}

// Block sum broadcast to every thread (fixed order, see block_sum).
template <int NT>
__device__ __forceinline__ double block_sum_all(double v, double *sh) {
  double r = block_sum<NT>(v, sh);
  if (threadIdx.x == 0) sh[0] = r;
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// Sum of nb partials written by the previous kernel, identical in every
// block (fixed order), broadcast to all threads.
template <int NT>
__device__ __forceinline__ double prologue_sum(const double *part, int nb, int stride, int off, double *sh) {
  double v = 0.0;
  for (int q = threadIdx.x; q < nb; q += NT) v += part[(size_t)q * stride + off];
  return block_sum_all<NT>(v, sh);
}

// Two prologue sums at once, r0 over pa[q stride + oa] and r1 over
// pb[q stride + ob]: each has prologue_sum's arithmetic and order (so the
// same result); their partial loads are in flight together and the block
// reduction's barriers are shared.  sh holds 2 NT / kWave + 2 doubles.
template <int NT>
__device__ __forceinline__ void prologue_sum2(const double *pa, int oa, const double *pb, int ob, int nb, int stride,
                                              double *sh, double &r0, double &r1) {
  constexpr int NW = NT / kWave;
  double v0 = 0.0, v1 = 0.0;
  for (int q = threadIdx.x; q < nb; q += NT) {
    v0 += pa[(size_t)q * stride + oa];
    v1 += pb[(size_t)q * stride + ob];
  }
  const int tid = threadIdx.x;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    v0 += __shfl_down(v0, off, kWave);
    v1 += __shfl_down(v1, off, kWave);
  }
  if ((tid & (kWave - 1)) == 0) {
    sh[tid / kWave] = v0;
    sh[NW + tid / kWave] = v1;
  }
  __syncthreads();
  if (tid == 0) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s0 += sh[w];
#pragma unroll
    for (int w = 0; w < NW; ++w) s1 += sh[NW + w];
    sh[2 * NW] = s0;
    sh[2 * NW + 1] = s1;
  }
  __syncthreads();
  r0 = sh[2 * NW];
  r1 = sh[2 * NW + 1];
  __syncthreads();
}

CCGeom make_geom(const moons_field_s *x, int kz);
int pick_kz(int n2);

} // namespace moons
