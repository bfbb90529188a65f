// mg.hip -- FP32 geometric multigrid V-cycle for the cell-centred Laplacian.
//
// The reference has no compiled multigrid (its V-cycle exists only in
// code_not_compiled/9.6.2017/MG/MG.f90:241-396: smooth, residual, restrict,
// recurse, prolongate, correct, smooth).  This V-cycle follows those steps
// with operators written from the finite-volume mathematics:
//   operator    L x = lap(x) with the homogeneous (implicit) BCs folded into
//               per-direction weights; L = A/vol for the PCG's A = vol*lap
//   smoother    red-black Gauss-Seidel, red then black before the coarse
//               correction and black then red after it (a symmetric cycle)
//   restriction volume-weighted sum of the 8 children's residuals
//   prolongation trilinear in the true cell-centre coordinates (stretched
//               grids), coarse ghosts from the BC rule
//   coarsest    one workgroup, a fixed number of red-black sweeps
// Everything is FP32 (4 B/cell per array): this is the preconditioner of an
// FP64 flexible PCG whose residual, dot products and BCs stay FP64.
#include "mg.hpp"
#include "unet.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cc_stencil.hpp"

namespace moons {

struct LevelDev {
  int n0, n1, n2, s0, s1;
  int per0, per1, per2;
  float *x, *f;
  const float *wm[3], *wp[3], *dg[3];
  const float *rf[3];          // restriction fractions into the next coarser level
  const int *pI[3], *pI2[3];   // prolongation sources in the next coarser level
  const float *pa[3], *pb[3];  // and their weights
};

// The coarse end of the hierarchy (levels whose x and f fit in one
// workgroup's LDS together) is cycled by a single workgroup in one launch:
// the same smoothing, restriction, coarse sweeps and prolongation as the
// per-level launches, in the same order, so the result is identical.
constexpr int kTailThreads = 1024;
constexpr int kMaxTail = 8;
constexpr size_t kTailLdsBytes = 60 * 1024;
constexpr long long kTailMaxCells = 8192;

struct TailDesc {
  int nt;                      // tail levels
  int per0, per1, per2;
  int n[kMaxTail][3];
  int xo[kMaxTail], fo[kMaxTail], co[kMaxTail], cs[kMaxTail]; // LDS float offsets; coefficient stride
  int nfloat;                  // LDS floats in use
};

struct MGHier {
  std::vector<LevelDev> L;
  std::vector<float *> arrays; // device allocations to free
  int nu1 = 2, nu2 = 2, coarse_sweeps = 40;
  int tail0 = 0;               // first level of the LDS-resident tail (== L.size(): none)
  size_t tail_lds = 0;         // LDS bytes of the tail's x and f arrays
  LevelDev *dlev = nullptr;    // level descriptors in device memory (for the tail kernel)
  TailDesc tail{};
  float *tailimg = nullptr;    // device copy of the tail's initial LDS image
  std::vector<std::vector<float>> hcoef; // host 1D coefficients: [(level*3 + d)*8 + c]
  std::vector<std::vector<double>> hn; // node arrays (ghosts included) of every level, 3 per level
  UnetPrec *learn = nullptr;   // learned initial guess of one grid level (or none)
  int learn_level = -1;
  float *gpart = nullptr;      // its Galerkin partial sums
};


__device__ __forceinline__ size_t lidx(const LevelDev &L, int i, int j, int k) {
  return (size_t)i + (size_t)L.s0 * ((size_t)j + (size_t)L.s1 * (size_t)k);
}

// Gauss-Seidel update of one cell: x = (f - sum_nb w x_nb) / diag.
__device__ __forceinline__ float gs_cell(const LevelDev &L, const float *__restrict__ x, int i, int j, int k,
                                         bool zero_guess) {
  const float d = L.dg[0][i] + L.dg[1][j] + L.dg[2][k];
  const float fv = L.f[lidx(L, i, j, k)];
  if (zero_guess) return fv / d;
  const int im = (i == 1) ? (L.per0 ? L.n0 : 1) : i - 1;
  const int ip = (i == L.n0) ? (L.per0 ? 1 : L.n0) : i + 1;
  const int jm = (j == 1) ? (L.per1 ? L.n1 : 1) : j - 1;
  const int jp = (j == L.n1) ? (L.per1 ? 1 : L.n1) : j + 1;
  const int km = (k == 1) ? (L.per2 ? L.n2 : 1) : k - 1;
  const int kp = (k == L.n2) ? (L.per2 ? 1 : L.n2) : k + 1;
  float s = L.wm[0][i] * x[lidx(L, im, j, k)] + L.wp[0][i] * x[lidx(L, ip, j, k)];
  s += L.wm[1][j] * x[lidx(L, i, jm, k)] + L.wp[1][j] * x[lidx(L, i, jp, k)];
  s += L.wm[2][k] * x[lidx(L, i, j, km)] + L.wp[2][k] * x[lidx(L, i, j, kp)];
  return (fv - s) / d;
}

// residual f - L x of one cell
__device__ __forceinline__ float res_cell(const LevelDev &L, int i, int j, int k) {
  const float *x = L.x;
  const int im = (i == 1) ? (L.per0 ? L.n0 : 1) : i - 1;
  const int ip = (i == L.n0) ? (L.per0 ? 1 : L.n0) : i + 1;
  const int jm = (j == 1) ? (L.per1 ? L.n1 : 1) : j - 1;
  const int jp = (j == L.n1) ? (L.per1 ? 1 : L.n1) : j + 1;
  const int km = (k == 1) ? (L.per2 ? L.n2 : 1) : k - 1;
  const int kp = (k == L.n2) ? (L.per2 ? 1 : L.n2) : k + 1;
  const float d = L.dg[0][i] + L.dg[1][j] + L.dg[2][k];
  float s = L.wm[0][i] * x[lidx(L, im, j, k)] + L.wp[0][i] * x[lidx(L, ip, j, k)];
  s += L.wm[1][j] * x[lidx(L, i, jm, k)] + L.wp[1][j] * x[lidx(L, i, jp, k)];
  s += L.wm[2][k] * x[lidx(L, i, j, km)] + L.wp[2][k] * x[lidx(L, i, j, kp)];
  s += d * x[lidx(L, i, j, k)];
  return L.f[lidx(L, i, j, k)] - s;
}

// Cells of one colour, (i+j+k) % 2 == color, one thread each.
__device__ __forceinline__ bool color_cell(const LevelDev &L, long long t, int color, int &i, int &j, int &k) {
  const int h0 = (L.n0 + 1) / 2;
  const long long per_k = (long long)h0 * L.n1;
  if (t >= per_k * L.n2) return false;
  k = 1 + (int)(t / per_k);
  const long long r = t % per_k;
  j = 1 + (int)(r / h0);
  const int ipair = (int)(r % h0);
  i = 2 * ipair + 1 + ((color + j + k + 1) & 1);
  return i <= L.n0;
}


__global__ void __launch_bounds__(256) k_mg_smooth(LevelDev L, int color, int zero_guess, const int *done) {
  if (done && *done) return;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const long long t = (long long)b * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!color_cell(L, t, color, i, j, k)) return;
  L.x[lidx(L, i, j, k)] = gs_cell(L, L.x, i, j, k, zero_guess != 0);
}

// Galerkin scaling of a learned initial guess n on one level: partial sums of
// (n, f) and (n, L n) over the interior (L the level operator).
__global__ void __launch_bounds__(256) k_mg_galerkin_part(LevelDev L, float *part, const int *done) {
  if (done && *done) return;
  __shared__ float sa[256], sb[256];
  const int nc = L.n0 * L.n1 * L.n2;
  const int t = blockIdx.x * 256 + threadIdx.x;
  float a = 0.f, bb = 0.f;
  if (t < nc) {
    const int i = 1 + t % L.n0, j = 1 + (t / L.n0) % L.n1, k = 1 + t / (L.n0 * L.n1);
    const size_t c = lidx(L, i, j, k);
    const float x = L.x[c], f = L.f[c];
    const float Lx = f - res_cell(L, i, j, k);
    a = x * f;
    bb = x * Lx;
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = bb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sa[0];
    part[2 * blockIdx.x + 1] = sb[0];
  }
}

// x *= c with c = (n, f) / (n, L n) (the energy-optimal multiple of n as an
// approximation of L^-1 f), or 0 when that is not a positive number; every
// block sums the partials in the same order.
__global__ void __launch_bounds__(256) k_mg_galerkin_scale(LevelDev L, const float *part, int nb, const int *done) {
  if (done && *done) return;
  __shared__ float sa[256], sb[256];
  float a = 0.f, bb = 0.f;
  for (int q = threadIdx.x; q < nb; q += 256) {
    a += part[2 * q];
    bb += part[2 * q + 1];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = bb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  const float q = sb[0] != 0.f ? sa[0] / sb[0] : 0.f;
  const float c = (q > 0.f && q < 3.0e38f) ? q : 0.f;
  const int nc = L.n0 * L.n1 * L.n2;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < nc) {
    const int i = 1 + t % L.n0, j = 1 + (t / L.n0) % L.n1, k = 1 + t / (L.n0 * L.n1);
    L.x[lidx(L, i, j, k)] *= c;
  }
}

// Coarse f = sum over the 8 children of (child vol / coarse vol) * residual.
__global__ void __launch_bounds__(256) k_mg_restrict(LevelDev F, LevelDev C, const int *done) {
  if (done && *done) return;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const long long t = (long long)b * blockDim.x + threadIdx.x;
  const long long nc = (long long)C.n0 * C.n1 * C.n2;
  if (t >= nc) return;
  const int I = 1 + (int)(t % C.n0);
  const int J = 1 + (int)((t / C.n0) % C.n1);
  const int K = 1 + (int)(t / ((long long)C.n0 * C.n1));
  float acc = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int i = 2 * I - 1 + (c & 1), j = 2 * J - 1 + ((c >> 1) & 1), k = 2 * K - 1 + (c >> 2);
    acc += F.rf[0][i] * F.rf[1][j] * F.rf[2][k] * res_cell(F, i, j, k);
  }
  C.f[lidx(C, I, J, K)] = acc;
}

// x_fine += trilinear interpolation of x_coarse.
__global__ void __launch_bounds__(256) k_mg_prolong(LevelDev F, LevelDev C, const int *done) {
  if (done && *done) return;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const long long t = (long long)b * blockDim.x + threadIdx.x;
  const long long nf = (long long)F.n0 * F.n1 * F.n2;
  if (t >= nf) return;
  const int i = 1 + (int)(t % F.n0);
  const int j = 1 + (int)((t / F.n0) % F.n1);
  const int k = 1 + (int)(t / ((long long)F.n0 * F.n1));
  const int I = F.pI[0][i], I2 = F.pI2[0][i];
  const int J = F.pI[1][j], J2 = F.pI2[1][j];
  const int K = F.pI[2][k], K2 = F.pI2[2][k];
  const float ax = F.pa[0][i], bx = F.pb[0][i];
  const float ay = F.pa[1][j], by = F.pb[1][j];
  const float az = F.pa[2][k], bz = F.pb[2][k];
  const float *c = C.x;
  const float e = ax * (ay * (az * c[lidx(C, I, J, K)] + bz * c[lidx(C, I, J, K2)]) +
                        by * (az * c[lidx(C, I, J2, K)] + bz * c[lidx(C, I, J2, K2)])) +
                  bx * (ay * (az * c[lidx(C, I2, J, K)] + bz * c[lidx(C, I2, J, K2)]) +
                        by * (az * c[lidx(C, I2, J2, K)] + bz * c[lidx(C, I2, J2, K2)]));
  F.x[lidx(F, i, j, k)] += e;
}

// Coarsest level: one workgroup, `sweeps` red-black sweeps from x = 0.
__global__ void __launch_bounds__(1024) k_mg_coarse(LevelDev L, int sweeps, const int *done) {
  if (done && *done) return;
  const int h0 = (L.n0 + 1) / 2;
  const long long ncol = (long long)h0 * L.n1 * L.n2;
  for (int s = 0; s < sweeps; ++s) {
    for (int color = 0; color < 2; ++color) {
      for (long long t = threadIdx.x; t < ncol; t += blockDim.x) {
        int i, j, k;
        if (color_cell(L, t, color, i, j, k)) L.x[lidx(L, i, j, k)] = gs_cell(L, L.x, i, j, k, s == 0 && color == 0);
      }
      __syncthreads();
    }
  }
}

// ---- LDS-resident tail.  Level l of the tail keeps x, f and its 1D
// coefficient arrays in LDS; everything is indexed off the one extern
// __shared__ array so the accesses are LDS instructions.

// coefficient c (0 wm, 1 wp, 2 dg, 3 rf, 4 pa, 5 pb, 6 pI, 7 pI2) of direction d at index i
#define TCOEF(l, c, d, i) sm[T.co[l] + ((c) * 3 + (d)) * T.cs[l] + (i)]
#define TCOEFI(l, c, d, i) __float_as_int(TCOEF(l, c, d, i))
#define TIDX(l, i, j, k) ((i) + (T.n[l][0] + 2) * ((j) + (T.n[l][1] + 2) * (k)))

__device__ __forceinline__ bool tail_color_cell(int n0, int n1, int n2, int t, int color, int &i, int &j, int &k) {
  const int h0 = (n0 + 1) / 2;
  const int per_k = h0 * n1;
  if (t >= per_k * n2) return false;
  k = 1 + t / per_k;
  const int r = t % per_k;
  j = 1 + r / h0;
  i = 2 * (r % h0) + 1 + ((color + j + k + 1) & 1);
  return i <= n0;
}

// neighbour indices with the periodic wrap / folded-BC clamp of gs_cell
__device__ __forceinline__ void tail_nb(const TailDesc &T, int l, int i, int j, int k, int nb[6]) {
  const int n0 = T.n[l][0], n1 = T.n[l][1], n2 = T.n[l][2];
  const int im = (i == 1) ? (T.per0 ? n0 : 1) : i - 1;
  const int ip = (i == n0) ? (T.per0 ? 1 : n0) : i + 1;
  const int jm = (j == 1) ? (T.per1 ? n1 : 1) : j - 1;
  const int jp = (j == n1) ? (T.per1 ? 1 : n1) : j + 1;
  const int km = (k == 1) ? (T.per2 ? n2 : 1) : k - 1;
  const int kp = (k == n2) ? (T.per2 ? 1 : n2) : k + 1;
  nb[0] = TIDX(l, im, j, k); nb[1] = TIDX(l, ip, j, k);
  nb[2] = TIDX(l, i, jm, k); nb[3] = TIDX(l, i, jp, k);
  nb[4] = TIDX(l, i, j, km); nb[5] = TIDX(l, i, j, kp);
}

// One colour of red-black Gauss-Seidel on tail level l (same arithmetic as gs_cell).
__device__ __forceinline__ void tail_smooth(const TailDesc &T, float *sm, int l, int color, bool zero_guess) {
  const int n0 = T.n[l][0], n1 = T.n[l][1], n2 = T.n[l][2];
  const int ncol = ((n0 + 1) / 2) * n1 * n2;
  float *x = sm + T.xo[l];
  const float *f = sm + T.fo[l];
  for (int t = threadIdx.x; t < ncol; t += kTailThreads) {
    int i, j, k;
    if (!tail_color_cell(n0, n1, n2, t, color, i, j, k)) continue;
    const float d = TCOEF(l, 2, 0, i) + TCOEF(l, 2, 1, j) + TCOEF(l, 2, 2, k);
    const int c = TIDX(l, i, j, k);
    const float fv = f[c];
    if (zero_guess) {
      x[c] = fv / d;
      continue;
    }
    int nb[6];
    tail_nb(T, l, i, j, k, nb);
    float s = TCOEF(l, 0, 0, i) * x[nb[0]] + TCOEF(l, 1, 0, i) * x[nb[1]];
    s += TCOEF(l, 0, 1, j) * x[nb[2]] + TCOEF(l, 1, 1, j) * x[nb[3]];
    s += TCOEF(l, 0, 2, k) * x[nb[4]] + TCOEF(l, 1, 2, k) * x[nb[5]];
    x[c] = (fv - s) / d;
  }
  __syncthreads();
}

// nsweeps red-black sweeps of tail level l (colour c_first first; the very
// first half-sweep from a zero guess when zero_first) with each thread's cells
// (at most CPT per colour) and their coefficients held in registers.
// WAVE: the level has at most 64 cells per colour; wave 0 alone runs the
// sweeps, ordered by wave-level fences instead of workgroup barriers.
template <int CPT, bool WAVE = false>
__device__ __forceinline__ void tail_sweeps(const TailDesc &T, float *sm, int l, int nsweeps, int c_first,
                                            bool zero_first) {
  if (WAVE && threadIdx.x >= 64) return;
  const int n0 = T.n[l][0], n1 = T.n[l][1], n2 = T.n[l][2];
  float *x = sm + T.xo[l];
  const float *f = sm + T.fo[l];
  bool v[2][CPT];
  int c[2][CPT], nb[2][CPT][6];
  float fv[2][CPT], dd[2][CPT], w[2][CPT][6];
#pragma unroll
  for (int col = 0; col < 2; ++col)
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      int i = 1, j = 1, k = 1;
      v[col][q] = tail_color_cell(n0, n1, n2, threadIdx.x + q * kTailThreads, col, i, j, k);
      if (!v[col][q]) i = j = k = 1;
      dd[col][q] = TCOEF(l, 2, 0, i) + TCOEF(l, 2, 1, j) + TCOEF(l, 2, 2, k);
      c[col][q] = TIDX(l, i, j, k);
      fv[col][q] = f[c[col][q]];
      tail_nb(T, l, i, j, k, nb[col][q]);
      w[col][q][0] = TCOEF(l, 0, 0, i); w[col][q][1] = TCOEF(l, 1, 0, i);
      w[col][q][2] = TCOEF(l, 0, 1, j); w[col][q][3] = TCOEF(l, 1, 1, j);
      w[col][q][4] = TCOEF(l, 0, 2, k); w[col][q][5] = TCOEF(l, 1, 2, k);
    }
  for (int s = 0; s < nsweeps; ++s) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = h == 0 ? c_first : 1 - c_first;
      const bool zg = zero_first && s == 0 && h == 0;
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        // col is uniform across the workgroup: select the register set
        const bool vv = col ? v[1][q] : v[0][q];
        if (!vv) continue;
        const int cc = col ? c[1][q] : c[0][q];
        const float ff = col ? fv[1][q] : fv[0][q];
        const float d = col ? dd[1][q] : dd[0][q];
        if (zg) {
          x[cc] = ff / d;
        } else {
          float a;
          if (col) {
            a = w[1][q][0] * x[nb[1][q][0]] + w[1][q][1] * x[nb[1][q][1]];
            a += w[1][q][2] * x[nb[1][q][2]] + w[1][q][3] * x[nb[1][q][3]];
            a += w[1][q][4] * x[nb[1][q][4]] + w[1][q][5] * x[nb[1][q][5]];
          } else {
            a = w[0][q][0] * x[nb[0][q][0]] + w[0][q][1] * x[nb[0][q][1]];
            a += w[0][q][2] * x[nb[0][q][2]] + w[0][q][3] * x[nb[0][q][3]];
            a += w[0][q][4] * x[nb[0][q][4]] + w[0][q][5] * x[nb[0][q][5]];
          }
          x[cc] = (ff - a) / d;
        }
      }
      if (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      } else {
        __syncthreads();
      }
    }
  }
}

// sweeps of tail level l by the register path when its colours fit, else the generic one
__device__ __forceinline__ void tail_level_sweeps(const TailDesc &T, float *sm, int l, int nsweeps, int c_first,
                                                  bool zero_first) {
  const int ncol = ((T.n[l][0] + 1) / 2) * T.n[l][1] * T.n[l][2];
  if (ncol <= 64) {
    tail_sweeps<1, true>(T, sm, l, nsweeps, c_first, zero_first);
    __syncthreads();
  } else if (ncol <= kTailThreads) {
    tail_sweeps<1>(T, sm, l, nsweeps, c_first, zero_first);
  } else if (ncol <= 2 * kTailThreads) {
    tail_sweeps<2>(T, sm, l, nsweeps, c_first, zero_first);
  } else {
    for (int s = 0; s < nsweeps; ++s) {
      tail_smooth(T, sm, l, c_first, zero_first && s == 0);
      tail_smooth(T, sm, l, 1 - c_first, false);
    }
  }
}

__global__ void __launch_bounds__(kTailThreads) k_mg_tail(const LevelDev *lev, const float *__restrict__ img, TailDesc T,
                                                          int t0, int nu1, int nu2, int coarse_sweeps, const int *done) {
  if (done && *done) return;
  extern __shared__ float sm[];
  const int nt = T.nt;
  // the tail's LDS image (coefficients in place, x and f zero), then f of the
  // first tail level: independent coalesced loads, all in flight together
  {
    const int nimg = T.nfloat;
    for (int t0i = 0; t0i < nimg; t0i += 4 * kTailThreads) {
      float r[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int t = t0i + q * kTailThreads + threadIdx.x;
        r[q] = t < nimg ? img[t] : 0.f;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int t = t0i + q * kTailThreads + threadIdx.x;
        if (t < nimg) sm[t] = r[q];
      }
    }
    __syncthreads();
    const float *f0 = lev[t0].f;
    const int n = (T.n[0][0] + 2) * (T.n[0][1] + 2) * (T.n[0][2] + 2);
    for (int t = threadIdx.x; t < n; t += kTailThreads) sm[T.fo[0] + t] = f0[t];
  }
  __syncthreads();
  for (int l = 0; l < nt - 1; ++l) {
    tail_level_sweeps(T, sm, l, nu1, 0, true);
    // restriction of the residual into level l+1 (res_cell / k_mg_restrict arithmetic)
    const int C0 = T.n[l + 1][0], C1 = T.n[l + 1][1], C2 = T.n[l + 1][2];
    const float *x = sm + T.xo[l];
    const float *f = sm + T.fo[l];
    const int nc = C0 * C1 * C2;
    for (int t = threadIdx.x; t < nc; t += kTailThreads) {
      const int I = 1 + t % C0, J = 1 + (t / C0) % C1, K = 1 + t / (C0 * C1);
      float acc = 0.f;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = 2 * I - 1 + (q & 1), j = 2 * J - 1 + ((q >> 1) & 1), k = 2 * K - 1 + (q >> 2);
        int nb[6];
        tail_nb(T, l, i, j, k, nb);
        const float d = TCOEF(l, 2, 0, i) + TCOEF(l, 2, 1, j) + TCOEF(l, 2, 2, k);
        float s = TCOEF(l, 0, 0, i) * x[nb[0]] + TCOEF(l, 1, 0, i) * x[nb[1]];
        s += TCOEF(l, 0, 1, j) * x[nb[2]] + TCOEF(l, 1, 1, j) * x[nb[3]];
        s += TCOEF(l, 0, 2, k) * x[nb[4]] + TCOEF(l, 1, 2, k) * x[nb[5]];
        s += d * x[TIDX(l, i, j, k)];
        const float r = f[TIDX(l, i, j, k)] - s;
        acc += TCOEF(l, 3, 0, i) * TCOEF(l, 3, 1, j) * TCOEF(l, 3, 2, k) * r;
      }
      sm[T.fo[l + 1] + TIDX(l + 1, I, J, K)] = acc;
    }
    __syncthreads();
  }
  tail_level_sweeps(T, sm, nt - 1, coarse_sweeps, 0, true);
  for (int l = nt - 2; l >= 0; --l) {
    const int n0 = T.n[l][0], n1 = T.n[l][1], n2 = T.n[l][2];
    float *x = sm + T.xo[l];
    const float *c = sm + T.xo[l + 1];
    const int nf = n0 * n1 * n2;
    for (int t = threadIdx.x; t < nf; t += kTailThreads) {
      const int i = 1 + t % n0, j = 1 + (t / n0) % n1, k = 1 + t / (n0 * n1);
      const int I = TCOEFI(l, 6, 0, i), I2 = TCOEFI(l, 7, 0, i);
      const int J = TCOEFI(l, 6, 1, j), J2 = TCOEFI(l, 7, 1, j);
      const int K = TCOEFI(l, 6, 2, k), K2 = TCOEFI(l, 7, 2, k);
      const float ax = TCOEF(l, 4, 0, i), bx = TCOEF(l, 5, 0, i);
      const float ay = TCOEF(l, 4, 1, j), by = TCOEF(l, 5, 1, j);
      const float az = TCOEF(l, 4, 2, k), bz = TCOEF(l, 5, 2, k);
      const int m = l + 1;
      const float e = ax * (ay * (az * c[TIDX(m, I, J, K)] + bz * c[TIDX(m, I, J, K2)]) +
                            by * (az * c[TIDX(m, I, J2, K)] + bz * c[TIDX(m, I, J2, K2)])) +
                      bx * (ay * (az * c[TIDX(m, I2, J, K)] + bz * c[TIDX(m, I2, J, K2)]) +
                            by * (az * c[TIDX(m, I2, J2, K)] + bz * c[TIDX(m, I2, J2, K2)]));
      x[TIDX(l, i, j, k)] += e;
    }
    __syncthreads();
    tail_level_sweeps(T, sm, l, nu2, 1, false);
  }
  {
    const LevelDev &G = lev[t0];
    const int n0 = T.n[0][0], n1 = T.n[0][1], n2 = T.n[0][2];
    const int n = n0 * n1 * n2;
    for (int t = threadIdx.x; t < n; t += kTailThreads) {
      const int i = 1 + t % n0, j = 1 + (t / n0) % n1, k = 1 + t / (n0 * n1);
      G.x[lidx(G, i, j, k)] = sm[T.xo[0] + TIDX(0, i, j, k)];
    }
  }
}

#undef TCOEF
#undef TCOEFI
#undef TIDX

// ------------------------------------------------------------------- host
namespace {

struct HostLevel {
  int n[3];
  std::vector<double> hn[3]; // nodes incl. ghosts
};

std::vector<double> coarsen_nodes(const std::vector<double> &hn) {
  // physical nodes hn[1..sn-2]; keep every other one, then mirror ghosts
  std::vector<double> phys(hn.begin() + 1, hn.end() - 1);
  std::vector<double> c;
  for (size_t m = 0; m < phys.size(); m += 2) c.push_back(phys[m]);
  std::vector<double> out(c.size() + 2);
  out[0] = c[0] - (c[1] - c[0]);
  for (size_t m = 0; m < c.size(); ++m) out[m + 1] = c[m];
  out[c.size() + 1] = c.back() + (c.back() - c[c.size() - 2]);
  return out;
}

bool dirichlet_like(int t) { return t == MOONS_BC_DIRICHLET || t == MOONS_BC_ANTISYMMETRIC; }

} // namespace

constexpr int kCoarseCells = 4;

MGHier *mg_create(const moons_field_s *x, int max_levels, int nu1, int nu2, int coarse_sweeps, std::string *err) {
  if (x->dl != MOONS_DL_CC) {
    if (err) *err = "multigrid needs a cell-centred field";
    return nullptr;
  }
  int bct[6];
  for (int f = 0; f < 6; ++f) bct[f] = x->has_bcs ? x->bct[f] : MOONS_BC_NEUMANN;
  std::vector<HostLevel> H;
  {
    HostLevel l0;
    for (int d = 0; d < 3; ++d) {
      l0.hn[d] = x->m->hc[d].hn;
      l0.n[d] = x->s[d] - 2;
    }
    H.push_back(l0);
  }
  // red-black ordering is a proper two-colouring across a periodic boundary
  // only for an even cell count (coarsening keeps periodic counts even)
  for (int d = 0; d < 3; ++d)
    if (bct[2 * d] == MOONS_BC_PERIODIC && H[0].n[d] % 2 != 0) {
      if (err) *err = "multigrid red-black smoothing needs an even cell count along periodic directions";
      return nullptr;
    }
  while ((int)H.size() < max_levels) {
    const HostLevel &f = H.back();
    bool ok = true;
    for (int d = 0; d < 3; ++d) ok = ok && (f.n[d] % 2 == 0) && (f.n[d] / 2 >= 2);
    for (int d = 0; d < 3; ++d)
      if (bct[2 * d] == MOONS_BC_PERIODIC) ok = ok && ((f.n[d] / 2) % 2 == 0);
    // a level of at most kCoarseCells per direction is solved by the
    // single-block sweeps; coarser levels cost launches and gain nothing
    if (std::max(f.n[0], std::max(f.n[1], f.n[2])) <= kCoarseCells) ok = false;
    if (!ok) break;
    HostLevel c;
    for (int d = 0; d < 3; ++d) {
      c.hn[d] = coarsen_nodes(f.hn[d]);
      c.n[d] = f.n[d] / 2;
    }
    H.push_back(c);
  }
  MGHier *h = new MGHier();
  for (const HostLevel &hl : H)
    for (int d = 0; d < 3; ++d) h->hn.push_back(hl.hn[d]);
  h->nu1 = nu1;
  h->nu2 = nu2;
  h->coarse_sweeps = coarse_sweeps;
  const int nl = (int)H.size();
  h->L.resize(nl);
  h->hcoef.resize((size_t)nl * 3 * 8);
  auto dalloc = [&](size_t n) -> float * {
    float *p = nullptr;
    if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) return nullptr;
    (void)hipMemsetAsync(p, 0, n * sizeof(float), stream());
    h->arrays.push_back(p);
    return p;
  };
  for (int l = 0; l < nl; ++l) {
    const HostLevel &hl = H[l];
    LevelDev &L = h->L[l];
    L.n0 = hl.n[0]; L.n1 = hl.n[1]; L.n2 = hl.n[2];
    L.s0 = L.n0 + 2; L.s1 = L.n1 + 2;
    L.per0 = bct[0] == MOONS_BC_PERIODIC;
    L.per1 = bct[2] == MOONS_BC_PERIODIC;
    L.per2 = bct[4] == MOONS_BC_PERIODIC;
    const size_t ncell = (size_t)(L.n0 + 2) * (L.n1 + 2) * (L.n2 + 2);
    L.x = dalloc(ncell);
    L.f = dalloc(ncell);
    if (!L.x || !L.f) {
      if (err) *err = "multigrid: hipMalloc failed";
      mg_destroy(h);
      return nullptr;
    }
    for (int d = 0; d < 3; ++d) {
      const int n = hl.n[d], s = n + 2;
      const std::vector<double> &hn = hl.hn[d];
      std::vector<double> dhn(s), hc(s), dhc(s - 1);
      for (int i = 0; i < s; ++i) {
        dhn[i] = hn[i + 1] - hn[i];
        hc[i] = 0.5 * (hn[i + 1] + hn[i]);
      }
      for (int i = 0; i < s - 1; ++i) dhc[i] = hc[i + 1] - hc[i];
      std::vector<float> wm(s, 0.f), wp(s, 0.f), dg(s, 0.f), rf(s, 0.f), pa(s, 0.f), pb(s, 0.f);
      std::vector<int> pI(s, 1), pI2(s, 1);
      const int bmin = bct[2 * d], bmax = bct[2 * d + 1];
      for (int i = 1; i <= n; ++i) {
        double a = 1.0 / (dhc[i - 1] * dhn[i]);
        double c = 1.0 / (dhc[i] * dhn[i]);
        double diag = -(a + c);
        if (i == 1 && bmin != MOONS_BC_PERIODIC) {
          diag += dirichlet_like(bmin) ? -a : a; // ghost = -x or +x folded in
          a = 0.0;
        }
        if (i == n && bmax != MOONS_BC_PERIODIC) {
          diag += dirichlet_like(bmax) ? -c : c;
          c = 0.0;
        }
        wm[i] = (float)a;
        wp[i] = (float)c;
        dg[i] = (float)diag;
      }
      if (l + 1 < nl) {
        const std::vector<double> &hnc = H[l + 1].hn[d];
        for (int i = 1; i <= n; ++i) rf[i] = (float)(dhn[i] / (hnc[(i + 1) / 2 + 1] - hnc[(i + 1) / 2]));
      }
      // prolongation weights for this level from its coarser level l+1
      if (l + 1 < nl) {
        const int nc = H[l + 1].n[d];
        const std::vector<double> &hnc = H[l + 1].hn[d];
        std::vector<double> hcc(nc + 2);
        for (int I = 0; I < nc + 2; ++I) hcc[I] = 0.5 * (hnc[I + 1] + hnc[I]);
        for (int i = 1; i <= n; ++i) {
          const int I = (i + 1) / 2;
          const bool left = (i % 2) == 1;
          int I2 = left ? I - 1 : I + 1;
          double a = (hc[i] - hcc[I2]) / (hcc[I] - hcc[I2]);
          double bw = 1.0 - a;
          if (I2 == 0 || I2 == nc + 1) {
            const int bt = left ? bmin : bmax;
            if (bt == MOONS_BC_PERIODIC) I2 = (I2 == 0) ? nc : 1;
            else {
              I2 = I;
              if (dirichlet_like(bt)) bw = -bw;
            }
          }
          pI[i] = I;
          pI2[i] = I2;
          pa[i] = (float)a;
          pb[i] = (float)bw;
        }
      }
      {
        std::vector<float> pIf(s), pI2f(s);
        for (int i = 0; i < s; ++i) {
          std::memcpy(&pIf[i], &pI[i], 4);
          std::memcpy(&pI2f[i], &pI2[i], 4);
        }
        const std::vector<float> *arr[8] = {&wm, &wp, &dg, &rf, &pa, &pb, &pIf, &pI2f};
        for (int c = 0; c < 8; ++c) h->hcoef[(l * 3 + d) * 8 + c] = *arr[c];
      }
      float *wmd = dalloc(s), *wpd = dalloc(s), *dgd = dalloc(s), *rfd = dalloc(s), *pad = dalloc(s), *pbd = dalloc(s);
      int *pId = nullptr, *pI2d = nullptr;
      float *tmp1 = dalloc(s), *tmp2 = dalloc(s);
      pId = reinterpret_cast<int *>(tmp1);
      pI2d = reinterpret_cast<int *>(tmp2);
      if (!wmd || !wpd || !dgd || !rfd || !pad || !pbd || !tmp1 || !tmp2) {
        if (err) *err = "multigrid: hipMalloc failed";
        mg_destroy(h);
        return nullptr;
      }
      (void)hipMemcpyAsync(wmd, wm.data(), s * sizeof(float), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(wpd, wp.data(), s * sizeof(float), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(dgd, dg.data(), s * sizeof(float), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(rfd, rf.data(), s * sizeof(float), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(pad, pa.data(), s * sizeof(float), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(pbd, pb.data(), s * sizeof(float), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(pId, pI.data(), s * sizeof(int), hipMemcpyHostToDevice, stream());
      (void)hipMemcpyAsync(pI2d, pI2.data(), s * sizeof(int), hipMemcpyHostToDevice, stream());
      (void)hipStreamSynchronize(stream());
      L.wm[d] = wmd; L.wp[d] = wpd; L.dg[d] = dgd; L.rf[d] = rfd;
      L.pa[d] = pad; L.pb[d] = pbd; L.pI[d] = pId; L.pI2[d] = pI2d;
    }
  }
  // LDS-resident tail: the longest coarse suffix whose x, f and 1D
  // coefficients fit in LDS (floats: 2 padded arrays + 8 arrays x 3 directions)
  auto level_floats = [&](int l) {
    const LevelDev &L = h->L[l];
    const int smax = std::max(L.n0, std::max(L.n1, L.n2)) + 2;
    return 2 * (size_t)(L.n0 + 2) * (L.n1 + 2) * (L.n2 + 2) + 24 * (size_t)smax;
  };
  h->tail0 = nl;
  size_t floats = 0;
  for (int l = nl - 1; l >= 0 && nl - l <= kMaxTail; --l) {
    const LevelDev &L = h->L[l];
    const size_t b = level_floats(l);
    if ((floats + b) * sizeof(float) > kTailLdsBytes || (long long)L.n0 * L.n1 * L.n2 > kTailMaxCells) break;
    floats += b;
    h->tail0 = l;
  }
  h->tail_lds = floats * sizeof(float);
  {
    TailDesc &T = h->tail;
    T.nt = nl - h->tail0;
    T.per0 = bct[0] == MOONS_BC_PERIODIC;
    T.per1 = bct[2] == MOONS_BC_PERIODIC;
    T.per2 = bct[4] == MOONS_BC_PERIODIC;
    int off = 0;
    for (int q = 0; q < T.nt; ++q) {
      const LevelDev &L = h->L[h->tail0 + q];
      T.n[q][0] = L.n0; T.n[q][1] = L.n1; T.n[q][2] = L.n2;
      const int pad = (L.n0 + 2) * (L.n1 + 2) * (L.n2 + 2);
      T.xo[q] = off; off += pad;
      T.fo[q] = off; off += pad;
      T.cs[q] = std::max(L.n0, std::max(L.n1, L.n2)) + 2;
      T.co[q] = off; off += 24 * T.cs[q];
    }
    T.nfloat = off;
    // initial LDS image: the coefficient blocks in place, zeros elsewhere
    std::vector<float> img((size_t)std::max(off, 1), 0.f);
    for (int q = 0; q < T.nt; ++q)
      for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 8; ++c) {
          const std::vector<float> &a = h->hcoef[((h->tail0 + q) * 3 + d) * 8 + c];
          std::copy(a.begin(), a.end(), img.begin() + T.co[q] + (c * 3 + d) * T.cs[q]);
        }
    h->tailimg = dalloc(img.size());
    if (!h->tailimg) {
      if (err) *err = "multigrid: hipMalloc failed";
      mg_destroy(h);
      return nullptr;
    }
    (void)hipMemcpyAsync(h->tailimg, img.data(), img.size() * sizeof(float), hipMemcpyHostToDevice, stream());
  }
  {
    float *p = dalloc((nl * sizeof(LevelDev) + sizeof(float) - 1) / sizeof(float));
    if (!p) {
      if (err) *err = "multigrid: hipMalloc failed";
      mg_destroy(h);
      return nullptr;
    }
    h->dlev = reinterpret_cast<LevelDev *>(p);
    (void)hipMemcpyAsync(h->dlev, h->L.data(), nl * sizeof(LevelDev), hipMemcpyHostToDevice, stream());
    (void)hipStreamSynchronize(stream());
  }
  return h;
}

void mg_destroy(MGHier *h) {
  if (!h) return;
  if (stream()) (void)hipStreamSynchronize(stream());
  if (h->learn) unet_prec_destroy(h->learn);
  if (h->gpart) (void)hipFree(h->gpart);
  for (float *p : h->arrays) (void)hipFree(p);
  delete h;
}

int mg_levels(const MGHier *h) { return (int)h->L.size(); }

int mg_level_cells(const MGHier *h, int l, int n[3]) {
  if (l < 0 || l >= (int)h->L.size()) return -1;
  n[0] = h->L[l].n0; n[1] = h->L[l].n1; n[2] = h->L[l].n2;
  return 0;
}

float *mg_f0(MGHier *h) { return h->L[0].f; }
float *mg_x0(MGHier *h) { return h->L[0].x; }

static unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

int mg_set_learned(MGHier *h, moons_unet_s *net, int level, std::string *err) {
  if (h->learn) unet_prec_destroy(h->learn);
  h->learn = nullptr;
  h->learn_level = -1;
  if (!net) return MOONS_OK;
  const int g = h->tail0 < (int)h->L.size() ? h->tail0 : (int)h->L.size() - 1;
  if (level < 0 || level >= g) {
    if (err) *err = "the learned level must be one of the grid-launch levels (above the LDS-resident coarse end)";
    return MOONS_ERR_ARG;
  }
  const std::vector<double> hn[3] = {h->hn[3 * level], h->hn[3 * level + 1], h->hn[3 * level + 2]};
  h->learn = unet_prec_create_nodes(net, hn, err);
  if (!h->learn) return MOONS_ERR_ARG;
  const LevelDev &L = h->L[level];
  const size_t nb = blocks_for((long long)L.n0 * L.n1 * L.n2);
  if (!h->gpart && hipMalloc(&h->gpart, 2 * 65536 * sizeof(float)) != hipSuccess) {
    if (err) *err = "hipMalloc failed";
    return MOONS_ERR_HIP;
  }
  if (nb > 65536) {
    if (err) *err = "learned level too large";
    return MOONS_ERR_ARG;
  }
  h->learn_level = level;
  return MOONS_OK;
}

void mg_set_params(MGHier *h, int nu1, int nu2, int coarse_sweeps) {
  h->nu1 = nu1;
  h->nu2 = nu2;
  h->coarse_sweeps = coarse_sweeps;
}



static void smooth(const LevelDev &L, int color, bool zero_guess, const int *done) {
  const long long nt = (long long)((L.n0 + 1) / 2) * L.n1 * L.n2;
  hipLaunchKernelGGL(k_mg_smooth, dim3(blocks_for(nt)), dim3(256), 0, stream(), L, color, zero_guess ? 1 : 0, done);
}

int mg_vcycle(MGHier *h, const int *done) {
  const int nl = (int)h->L.size();
  const bool tail = h->tail0 < nl;
  const int g = tail ? h->tail0 : nl - 1; // levels [0, g) run as grid launches
  // descend
  for (int l = 0; l < g; ++l) {
    LevelDev &L = h->L[l];
    const bool learned = h->learn && l == h->learn_level;
    if (learned) {
      // learned initial guess x = c N(f), then smoothing from it
      int st = unet_prec_apply(h->learn, L.f, L.x, done);
      if (st) return st;
      const int nc = L.n0 * L.n1 * L.n2;
      const unsigned nb = blocks_for(nc);
      hipLaunchKernelGGL(k_mg_galerkin_part, dim3(nb), dim3(256), 0, stream(), L, h->gpart, done);
      hipLaunchKernelGGL(k_mg_galerkin_scale, dim3(nb), dim3(256), 0, stream(), L, h->gpart, (int)nb, done);
    }
    for (int s = 0; s < h->nu1; ++s) {
      smooth(L, 0, s == 0 && !learned, done);
      smooth(L, 1, false, done);
    }
    const LevelDev &C = h->L[l + 1];
    const long long nc = (long long)C.n0 * C.n1 * C.n2;
    hipLaunchKernelGGL(k_mg_restrict, dim3(blocks_for(nc)), dim3(256), 0, stream(), L, C, done);
  }
  if (tail)
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), h->tail_lds, stream(), h->dlev, h->tailimg, h->tail,
                       h->tail0, h->nu1, h->nu2, h->coarse_sweeps, done);
  else
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(1024), 0, stream(), h->L[nl - 1], h->coarse_sweeps, done);
  // ascend
  for (int l = g - 1; l >= 0; --l) {
    LevelDev &L = h->L[l];
    const LevelDev &C = h->L[l + 1];
    const long long nf = (long long)L.n0 * L.n1 * L.n2;
    hipLaunchKernelGGL(k_mg_prolong, dim3(blocks_for(nf)), dim3(256), 0, stream(), L, C, done);
    for (int s = 0; s < h->nu2; ++s) {
      smooth(L, 1, false, done);
      smooth(L, 0, false, done);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MOONS_ERR_HIP, std::string("multigrid launch: ") + hipGetErrorString(e));
  return MOONS_OK;
}

} // namespace moons
