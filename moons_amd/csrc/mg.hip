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
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cc_stencil.hpp"

namespace moons {

struct LevelDev {
  int n0, n1, n2;
  int h0, half;                // cells per colour and row ((n0+1)/2), per colour (h0*n1*n2)
  int per0, per1, per2;
  int line1, line2;            // smoothed by zebra lines along y / z (stretched, non-periodic directions)
  float *x, *f;
  float *xalt;                 // the other buffer of x for the fused sweeps (swapped with x on the host)
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

constexpr int kTailLoads = 8; // per-thread loads of f and of the coefficients at entry

struct TailDesc {
  int nt;                      // tail levels
  int per0, per1, per2;
  int n[kMaxTail][3];
  int xo[kMaxTail], fo[kMaxTail], co[kMaxTail], cs[kMaxTail]; // LDS float offsets; coefficient stride
  int coff[kMaxTail];          // offset of each level's coefficient block in the packed device copy
  int ncoef;                   // floats of coefficients in all
  int nfloat;                  // LDS floats in use
  const float *f0;             // f of the first tail level (colour-split, in the level arrays)
  float *x0;                   // and its x, written at exit
};

struct MGHier {
  std::vector<LevelDev> L;
  std::vector<float *> arrays; // device allocations to free
  int nu1 = 2, nu2 = 2, coarse_sweeps = 40;
  int tail0 = 0;               // first level of the LDS-resident tail (== L.size(): none)
  size_t tail_lds = 0;         // LDS bytes of the tail's x and f arrays
  TailDesc tail{};
  float *tailimg = nullptr;    // the tail's coefficient blocks, packed
  std::vector<std::vector<float>> hcoef; // host 1D coefficients: [(level*3 + d)*8 + c]
  std::vector<std::vector<double>> hn; // node arrays (ghosts included) of every level, 3 per level
  UnetPrec *learn = nullptr;   // learned initial guess of one grid level (or none)
  int learn_level = -1;
  int learn_cycles = 0;        // the first learn_cycles V-cycles of a solve use it (0: every one)
  int gamma_level = -1;        // levels from this one down are cycled gamma_cycles times per visit
  int gamma_cycles = 1;
  int cycle = 0;               // V-cycles since the solve began
  float *lpf = nullptr, *lpx = nullptr; // its padded (n+2)^3 input and output images
  float *gpart = nullptr;      // its Galerkin partial sums
};


// Colour-split storage of a level's x and f: the cells of colour
// c = (i+j+k) % 2 fill the c-th half of the array, each half x-fastest over
// the pair index (i-1)/2, so that a red-black half-sweep reads and writes
// whole cache lines of one colour.  No ghost cells: the smoother and the
// residual fold the BCs into their weights.  mg_create keeps every level
// below 2^31 cells, so 32-bit arithmetic is exact.
__device__ __forceinline__ int lidx(const LevelDev &L, int i, int j, int k) {
  return ((i + j + k) & 1) * L.half + ((i - 1) >> 1) + L.h0 * ((j - 1) + L.n1 * (k - 1));
}

// Per-level grid kernels run on tiles of blockDim.x (x) by blockDim.y (y)
// threads, each thread marching kz planes along z.  The linear tile order (x
// fastest) is remapped so that each XCD gets a contiguous range of z-slabs,
// which keeps the z-neighbour planes of a tile in that XCD's L2.
__device__ __forceinline__ void tile_origin(int kz, int &i0, int &j, int &k0) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int b = xcd_remap(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
  const int r = b / gx;
  i0 = (b - r * gx) * blockDim.x + threadIdx.x;
  j = 1 + (r % gy) * blockDim.y + threadIdx.y;
  k0 = 1 + (r / gy) * kz;
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


// One half-sweep: the cells of one colour, (i+j+k) % 2 == color.  Thread x
// of the tile is the pair index p of row (j, k), i.e. the cell of this colour
// among (2p+1, 2p+2).  The arithmetic is gs_cell's; the neighbours are cells
// of the other colour (a periodic wrap keeps that, every count along a
// periodic direction being even), except a wall's clamp to the cell itself.
__global__ void __launch_bounds__(256) k_mg_smooth(LevelDev L, int color, int zero_guess, int kz, const int *done) {
  // *done is tested at the stores, so that its load overlaps the operand loads
  const bool skip = done && *done;
  int p, j, k0;
  tile_origin(kz, p, j, k0);
  if (p >= L.h0 || j > L.n1) return;
  float *__restrict__ xc = L.x + color * L.half;
  const float *__restrict__ xo = L.x + (1 - color) * L.half;
  const float *__restrict__ fc = L.f + color * L.half;
  const int h0 = L.h0, n1 = L.n1;
  const bool jlo = j == 1 && !L.per1, jhi = j == n1 && !L.per1;
  const int jm = (j == 1) ? n1 : j - 1, jp = (j == n1) ? 1 : j + 1;
  const float wmj = L.wm[1][j], wpj = L.wp[1][j], dgj = L.dg[1][j];
  const int kend = min(k0 + kz, L.n2 + 1);
  for (int k = k0; k < kend; ++k) {
    const int i = 2 * p + 1 + ((color + j + k + 1) & 1);
    if (i > L.n0) continue;
    const int row = h0 * ((j - 1) + n1 * (k - 1));
    const int c = row + p;
    const float d = L.dg[0][i] + dgj + L.dg[2][k];
    const float fv = fc[c];
    if (zero_guess) {
      if (!skip) xc[c] = fv / d;
      continue;
    }
    const float xim = (i == 1) ? (L.per0 ? xo[row + h0 - 1] : xc[c]) : xo[row + ((i - 2) >> 1)];
    const float xip = (i == L.n0) ? (L.per0 ? xo[row] : xc[c]) : xo[row + (i >> 1)];
    const float xjm = jlo ? xc[c] : xo[h0 * ((jm - 1) + n1 * (k - 1)) + p];
    const float xjp = jhi ? xc[c] : xo[h0 * ((jp - 1) + n1 * (k - 1)) + p];
    const int km = (k == 1) ? L.n2 : k - 1, kp = (k == L.n2) ? 1 : k + 1;
    const float xkm = (k == 1 && !L.per2) ? xc[c] : xo[h0 * ((j - 1) + n1 * (km - 1)) + p];
    const float xkp = (k == L.n2 && !L.per2) ? xc[c] : xo[h0 * ((j - 1) + n1 * (kp - 1)) + p];
    float s = L.wm[0][i] * xim + L.wp[0][i] * xip;
    s += wmj * xjm + wpj * xjp;
    s += L.wm[2][k] * xkm + L.wp[2][k] * xkp;
    if (!skip) xc[c] = (fv - s) / d;
  }
}

// Zebra line Gauss-Seidel along direction D (1: y, 2: z) on the lines of
// colour c -- lines (i, k) with (i + k) % 2 == c for y-lines, (i, j) with
// (i + j) % 2 == c for z-lines -- which couple only to lines of the other
// colour.  Each line's tridiagonal system (the weights of D along the line,
// diagonal dg_x + dg_y + dg_z, the off-line neighbours moved to the
// right-hand side; with zero_guess they are zero) is solved exactly.  A
// thread owns the C cells Cu .. Cu + C - 1 of a line (NP >= the line length,
// a power of two; NP / C threads per line) with their rows normalised to a
// unit diagonal.  Cyclic reduction inside the thread (log2 C strides) leaves
// one row per thread, coupling x[Cu] to x[C(u - 1)] and x[C(u + 1)]: rows
// Cu + t, t > 0, need only the thread's own rows, and row Cu the
// predecessor's rows C(u + 1) - h as eliminated at each stride h (one LDS
// exchange).  Those rows are solved by log2(NP / C) parallel cyclic reduction
// steps (rows (a, c, d) in two struct-of-arrays LDS buffers, one barrier per
// step); the eliminated rows then follow, widest stride first, with the
// successor's x[C(u + 1)].  Cells past the line's end hold the identity
// equation.  Threads of a block take consecutive lines fastest, contiguous in
// the colour-split layout, so the gather and the store are coalesced across
// lines.  A block of TB threads holds C TB / NP lines; blocks are numbered
// so that each XCD (the hardware deals consecutive block ids round-robin over
// the 8 XCDs) sweeps a contiguous slab of planes q, whose neighbouring blocks
// share its L2: the other colour's x read by the planes q - 1, q, q + 1 and
// both halves of a 128-byte row are then fetched from HBM once.
constexpr int kLineMax = 128; // longest line (at least 16 lines per block)
constexpr long long kLineMinCells = 32LL * 32 * 32; // levels this small smooth by points
// MOONS_MG_LINE_MIN_CELLS overrides kLineMinCells, for tuning runs.
static long long line_min_cells() {
  const char *e = getenv("MOONS_MG_LINE_MIN_CELLS");
  return e ? atoll(e) : kLineMinCells;
}

// Element e of a level array by a 32-bit byte offset from its base (the
// scalar-base + 32-bit vector-offset load form, without 64-bit address
// arithmetic per load).  Valid while 4 e < 2^32: line smoothing is used only
// on levels below 2^30 cells.
__device__ __forceinline__ float ld32(const float *__restrict__ base, unsigned e) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (e << 2));
}

// Row a x[t-1] + x[t] + c x[t+1] = d of cell t (0-based) of a line (normalised).
// A row past the line's end (t >= nl) or of a line past the level's edge
// (!ok) is the identity row (0, 0, 0); its loads go to an in-bounds cell and
// are discarded, so that no branch separates the rows' loads and all of a
// thread's loads are in flight together.
template <int D>
__device__ __forceinline__ float3 line_row(const LevelDev &L, int i, int m, int q, int t, int nl, bool ok,
                                           bool zero_guess, int &at) {
  const bool live = ok && t < nl;
  const int ic = ok ? i : 1, mc = ok ? m : 0, tc = t < nl ? t : nl - 1;
  const int j = D == 1 ? tc + 1 : q, k = D == 1 ? q : tc + 1;
  // plane offsets: h0 (j - 1) and h0 n1 (k - 1), one of them block-uniform;
  // every factor is below 2^24 (lines of at most 128 cells, levels below 2^30)
  const int h0 = L.h0, hn = L.h0 * L.n1;
  const int oj = (int)__umul24((unsigned)h0, (unsigned)(j - 1)), ok_ = (int)__umul24((unsigned)hn, (unsigned)(k - 1));
  const int col = (ic + j + k) & 1;
  const int row = oj + ok_;
  at = (col ? L.half : 0) + row + mc;
  const unsigned xo = (unsigned)(col ? 0 : L.half); // the other colour's half of x
  const int im = (ic == 1) ? (L.per0 ? L.n0 : 1) : ic - 1;
  const int ip = (ic == L.n0) ? (L.per0 ? 1 : L.n0) : ic + 1;
  float s = ld32(L.wm[0], ic) * ld32(L.x, xo + row + ((im - 1) >> 1)) +
            ld32(L.wp[0], ic) * ld32(L.x, xo + row + ((ip - 1) >> 1));
  if (D == 1) { // off-line neighbours along z (k is block-uniform)
    const int km = (k == 1) ? (L.per2 ? L.n2 : 1) : k - 1, kp = (k == L.n2) ? (L.per2 ? 1 : L.n2) : k + 1;
    s += ld32(L.wm[2], k) * ld32(L.x, xo + oj + hn * (km - 1) + mc) +
         ld32(L.wp[2], k) * ld32(L.x, xo + oj + hn * (kp - 1) + mc);
  } else {      // off-line neighbours along y (j is block-uniform)
    const int jm = (j == 1) ? (L.per1 ? L.n1 : 1) : j - 1, jp = (j == L.n1) ? (L.per1 ? 1 : L.n1) : j + 1;
    s += ld32(L.wm[1], j) * ld32(L.x, xo + h0 * (jm - 1) + ok_ + mc) +
         ld32(L.wp[1], j) * ld32(L.x, xo + h0 * (jp - 1) + ok_ + mc);
  }
  const float d = ld32(L.f, at) - (zero_guess ? 0.f : s);
  const float r = __builtin_amdgcn_rcpf(ld32(L.dg[0], ic) + ld32(L.dg[1], j) + ld32(L.dg[2], k));
  const float3 row3 = make_float3(ld32(L.wm[D], tc + 1) * r, ld32(L.wp[D], tc + 1) * r, d * r);
  return live ? row3 : make_float3(0.f, 0.f, 0.f);
}

template <int D, int NP, int TB, int C>
__global__ void __launch_bounds__(TB) k_mg_line(LevelDev L, int c, int zero_guess, const int *done, int gx) {
  constexpr int NQ = NP / C;             // threads per line
  constexpr int LPB = TB / NQ;           // lines per block
  constexpr int LV = C == 2 ? 1 : C == 4 ? 2 : C == 8 ? 3 : 4; // log2(C) local reduction levels
  static_assert(C == 2 || C == 4 || C == 8 || C == 16, "cells per thread");
  __shared__ float E[3 * LV][TB];        // rows C - h (h = 1, 2, .., C / 2) as eliminated, read by the successor
  __shared__ float S[2][3][TB];          // PCR rows (a, c, d), two buffers
  const int nl = D == 1 ? L.n1 : L.n2;   // line length, <= NP
  // XCD-contiguous numbering: the first 8 * (nb / 8) ids are dealt as slabs
  // of nb / 8 consecutive logical blocks per XCD, the rest keep their id
  const int nb = (int)gridDim.x, per = nb >> 3, b = (int)blockIdx.x;
  const int lb = b < (per << 3) ? (b & 7) * per + (b >> 3) : b;
  const int bx = lb % gx;
  const int q = lb / gx + 1;             // the other line index (k for y-lines, j for z-lines)
  const int p = (c + 1 + q) & 1;         // i-parity of the lines of colour c at this q
  const int l = threadIdx.x % LPB, u = threadIdx.x / LPB; // line in the block, cell group
  const int m = bx * LPB + l;
  const int i = 2 * m + 1 + p;
  const bool ok = i <= L.n0;
  const int me = threadIdx.x;
  int at[C];
  float a[C], cc[C], d[C];
#pragma unroll
  for (int s4 = 0; s4 < C; ++s4) {
    const float3 r = line_row<D>(L, i, m, q, C * u + s4, nl, ok, zero_guess, at[s4]);
    a[s4] = r.x;
    cc[s4] = r.y;
    d[s4] = r.z;
  }
  // Cyclic reduction inside the thread: at stride h every row t (a multiple
  // of 2h, t > 0) drops x[t - h] and x[t + h] with those rows (as left by
  // stride h / 2), coupling to x[t - 2h] and x[t + 2h].  Rows are updated in
  // place; a row keeps the values it had when it was eliminated.
#pragma unroll
  for (int h = 1; h < C; h <<= 1)
#pragma unroll
    for (int t = 2 * h; t < C; t += 2 * h) {
      const float rr = __builtin_amdgcn_rcpf(1.f - a[t] * cc[t - h] - cc[t] * a[t + h]);
      const float an = -a[t] * a[t - h] * rr, cn = -cc[t] * cc[t + h] * rr;
      d[t] = (d[t] - a[t] * d[t - h] - cc[t] * d[t + h]) * rr;
      a[t] = an;
      cc[t] = cn;
    }
  // row 0 needs the predecessor's row C - h at each stride h (one exchange)
#pragma unroll
  for (int v = 0; v < LV; ++v) {
    E[3 * v][me] = a[C - (1 << v)];
    E[3 * v + 1][me] = cc[C - (1 << v)];
    E[3 * v + 2][me] = d[C - (1 << v)];
  }
  __syncthreads();
  float a0 = a[0], c0 = cc[0], d0 = d[0];
#pragma unroll
  for (int v = 0; v < LV; ++v) {
    const int h = 1 << v;
    float am = 0.f, cm = 0.f, dm = 0.f; // the predecessor's row C - h
    if (u > 0) {
      const int o = me - LPB;
      am = E[3 * v][o];
      cm = E[3 * v + 1][o];
      dm = E[3 * v + 2][o];
    }
    const float rr = __builtin_amdgcn_rcpf(1.f - a0 * cm - c0 * a[h]);
    const float an = -a0 * am * rr, cn = -c0 * cc[h] * rr;
    d0 = (d0 - a0 * dm - c0 * d[h]) * rr;
    a0 = an;
    c0 = cn;
  }
  // rows C u couple to x[C (u - 1)] and x[C (u + 1)]: parallel cyclic reduction
  float aa = a0, ac = c0, dd = d0;
  int buf = 0;
#pragma unroll
  for (int st = 1; st < NQ; st <<= 1) {
    // ping-pong buffers: a buffer is rewritten two steps later, after the
    // barrier that follows every read of it
    S[buf][0][me] = aa;
    S[buf][1][me] = ac;
    S[buf][2][me] = dd;
    __syncthreads();
    float am = 0.f, cm = 0.f, dm = 0.f, ap = 0.f, cp = 0.f, dp = 0.f;
    if (u - st >= 0) {
      const int o = me - st * LPB;
      am = S[buf][0][o];
      cm = S[buf][1][o];
      dm = S[buf][2][o];
    }
    if (u + st < NQ) {
      const int o = me + st * LPB;
      ap = S[buf][0][o];
      cp = S[buf][1][o];
      dp = S[buf][2][o];
    }
    const float rr = __builtin_amdgcn_rcpf(1.f - aa * cm - ac * ap);
    const float an = -aa * am * rr, cn = -ac * cp * rr;
    dd = (dd - aa * dm - ac * dp) * rr;
    aa = an;
    ac = cn;
    buf ^= 1;
  }
  // dd is x[C u]; x[C (u + 1)] from the successor, then the eliminated rows
  // from the widest stride down
  S[buf][0][me] = dd;
  __syncthreads();
  float x[C + 1];
  x[0] = dd;
  x[C] = u + 1 < NQ ? S[buf][0][me + LPB] : 0.f;
#pragma unroll
  for (int h = C / 2; h >= 1; h >>= 1)
#pragma unroll
    for (int t = h; t < C; t += 2 * h) x[t] = d[t] - a[t] * x[t - h] - cc[t] * x[t + h];
  const bool skip = done && *done; // stores skipped once the solver has exited
  if (ok && !skip) {
#pragma unroll
    for (int s4 = 0; s4 < C; ++s4)
      if (C * u + s4 < nl) L.x[at[s4]] = x[s4];
  }
}

// One whole red-black sweep (colour c_first, then the other) in one launch,
// from x = L.x into xn, every value read from L.x: a thread owns the pair
// (2p+1, 2p+2) of row (j, k), computes the new value of its c_first cell,
// then that of its other cell from the new values of that cell's neighbours,
// each recomputed from L.x (its pair partner's is the value just computed).
// The arithmetic is that of the two half-sweeps (gs_cell, the same operand
// order), so the result is identical; the separate output buffer removes the
// hazard between tiles, and the recomputation costs loads the L1 and L2 serve.
// A neighbour that a wall clamps to the cell itself is read from L.x, as the
// second half-sweep reads the cell's own old value.
__global__ void __launch_bounds__(256) k_mg_sweep2(LevelDev L, float *__restrict__ xn, int c_first, int zero_first,
                                                   int kz, const int *done) {
  const bool skip = done && *done; // tested at the stores (overlaps the loads)
  int p, j, k0;
  tile_origin(kz, p, j, k0);
  if (p >= L.h0 || j > L.n1) return;
  const float *xo = L.x;
  const bool zg = zero_first != 0;
  const int kend = min(k0 + kz, L.n2 + 1);
  for (int k = k0; k < kend; ++k) {
    const int ia = 2 * p + 1 + ((c_first + j + k + 1) & 1); // the pair's cell of colour c_first
    const int ib = 2 * p + 1 + ((c_first + j + k) & 1);     // and its cell of the other colour
    const bool has_a = ia <= L.n0, has_b = ib <= L.n0;
    const float va = has_a ? gs_cell(L, xo, ia, j, k, zg) : 0.f;
    float vb = 0.f;
    if (has_b) {
      const int im = (ib == 1) ? (L.per0 ? L.n0 : 1) : ib - 1;
      const int ip = (ib == L.n0) ? (L.per0 ? 1 : L.n0) : ib + 1;
      const int jm = (j == 1) ? (L.per1 ? L.n1 : 1) : j - 1;
      const int jp = (j == L.n1) ? (L.per1 ? 1 : L.n1) : j + 1;
      const int km = (k == 1) ? (L.per2 ? L.n2 : 1) : k - 1;
      const int kp = (k == L.n2) ? (L.per2 ? 1 : L.n2) : k + 1;
      // the value the second half-sweep reads at neighbour (ii, jj, kk) of ib
      auto nv = [&](int ii, int jj, int kk) -> float {
        if (ii == ib && jj == j && kk == k) return xo[lidx(L, ib, j, k)];
        if (ii == ia && jj == j && kk == k) return va;
        return gs_cell(L, xo, ii, jj, kk, zg);
      };
      const float d = L.dg[0][ib] + L.dg[1][j] + L.dg[2][k];
      const float fv = L.f[lidx(L, ib, j, k)];
      float s = L.wm[0][ib] * nv(im, j, k) + L.wp[0][ib] * nv(ip, j, k);
      s += L.wm[1][j] * nv(ib, jm, k) + L.wp[1][j] * nv(ib, jp, k);
      s += L.wm[2][k] * nv(ib, j, km) + L.wp[2][k] * nv(ib, j, kp);
      vb = (fv - s) / d;
    }
    if (!skip) {
      if (has_a) xn[lidx(L, ia, j, k)] = va;
      if (has_b) xn[lidx(L, ib, j, k)] = vb;
    }
  }
}

// Level arrays <-> a padded (n+2)^3 FP32 image (the learned initial guess
// reads and writes that layout); ghosts of the image are left as they are.
__global__ void __launch_bounds__(256) k_mg_to_padded(LevelDev L, float *__restrict__ pf, const int *done) {
  if (done && *done) return;
  const int nc = L.n0 * L.n1 * L.n2;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  const int i = 1 + t % L.n0, j = 1 + (t / L.n0) % L.n1, k = 1 + t / (L.n0 * L.n1);
  pf[i + (L.n0 + 2) * (j + (L.n1 + 2) * k)] = L.f[lidx(L, i, j, k)];
}
__global__ void __launch_bounds__(256) k_mg_from_padded(LevelDev L, const float *__restrict__ px, const int *done) {
  if (done && *done) return;
  const int nc = L.n0 * L.n1 * L.n2;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  const int i = 1 + t % L.n0, j = 1 + (t / L.n0) % L.n1, k = 1 + t / (L.n0 * L.n1);
  L.x[lidx(L, i, j, k)] = px[i + (L.n0 + 2) * (j + (L.n1 + 2) * k)];
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
__global__ void __launch_bounds__(256) k_mg_restrict(LevelDev F, LevelDev C, int kz, const int *done) {
  const bool skip = done && *done; // tested at the store (overlaps the loads)
  int I, J, K0;
  tile_origin(kz, I, J, K0);
  ++I;
  if (I > C.n0 || J > C.n1) return;
  const int Kend = min(K0 + kz, C.n2 + 1);
  for (int K = K0; K < Kend; ++K) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int i = 2 * I - 1 + (c & 1), j = 2 * J - 1 + ((c >> 1) & 1), k = 2 * K - 1 + (c >> 2);
      acc += F.rf[0][i] * F.rf[1][j] * F.rf[2][k] * res_cell(F, i, j, k);
    }
    if (!skip) C.f[lidx(C, I, J, K)] = acc;
  }
}

// x_fine += trilinear interpolation of x_coarse, one thread per coarse cell
// (I,J,K) and its 8 children (2I-1..2I, 2J-1..2J, 2K-1..2K).  The two
// x-children have the same pair index I-1 in their colour halves, so a wave's
// read-modify-writes are coalesced, and the 27 coarse values the children
// interpolate from are loaded once.  Each child's arithmetic is the
// per-fine-cell formula (source I and pI2, weights pa and pb per direction).
// MODE selects the children corrected: 0 all; otherwise only the cells that
// the fine level's first post-smoothing half-sweep reads, since that
// half-sweep overwrites its own colour without reading it (a wall's clamped
// neighbour is read with weight zero): 1 the z-line colour 0, (i + j) even
// (cx + cy even); 2 the y-line colour 0, (i + k) even (cx + cz even); 3 the
// red-black colour 0, (i + j + k) even (cx + cy + cz odd).
template <int MODE>
__device__ __forceinline__ constexpr bool prolong_child(int cx, int cy, int cz) {
  return MODE == 0 || (MODE == 1 && ((cx + cy) & 1) == 0) || (MODE == 2 && ((cx + cz) & 1) == 0) ||
         (MODE == 3 && ((cx + cy + cz) & 1) == 1);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_mg_prolong(LevelDev F, LevelDev C, int kz, const int *done) {
  const bool skip = done && *done; // tested at the stores (overlaps the loads)
  int I, J, K0;
  tile_origin(kz, I, J, K0);
  ++I;
  if (I > C.n0 || J > C.n1) return;
  const float *c = C.x;
  int xs[3], ys[3];
  float ax[2], bx[2], ay[2], by[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int i = 2 * I - 1 + a, j = 2 * J - 1 + a;
    xs[2 * a] = F.pI2[0][i];
    ys[2 * a] = F.pI2[1][j];
    ax[a] = F.pa[0][i]; bx[a] = F.pb[0][i];
    ay[a] = F.pa[1][j]; by[a] = F.pb[1][j];
  }
  xs[1] = I;
  ys[1] = J;
  const int Kend = min(K0 + kz, C.n2 + 1);
  for (int K = K0; K < Kend; ++K) {
    int zs[3];
    float az[2], bz[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int k = 2 * K - 1 + a;
      zs[2 * a] = F.pI2[2][k];
      az[a] = F.pa[2][k]; bz[a] = F.pb[2][k];
    }
    zs[1] = K;
    float v[3][3][3], xf[2][2][2];
    // the coarse values some corrected child interpolates from
    auto used = [](int x, int y, int z) {
      for (int cz = 0; cz < 2; ++cz)
        for (int cy = 0; cy < 2; ++cy)
          for (int cx = 0; cx < 2; ++cx)
            if (prolong_child<MODE>(cx, cy, cz) && (x == 1 || x == 2 * cx) && (y == 1 || y == 2 * cy) &&
                (z == 1 || z == 2 * cz))
              return true;
      return false;
    };
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
      for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) v[z][y][x] = used(x, y, z) ? c[lidx(C, xs[x], ys[y], zs[z])] : 0.f;
#pragma unroll
    for (int cz = 0; cz < 2; ++cz)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cx = 0; cx < 2; ++cx)
          xf[cz][cy][cx] =
              prolong_child<MODE>(cx, cy, cz) ? F.x[lidx(F, 2 * I - 1 + cx, 2 * J - 1 + cy, 2 * K - 1 + cz)] : 0.f;
#pragma unroll
    for (int cz = 0; cz < 2; ++cz)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cx = 0; cx < 2; ++cx) {
          if (!prolong_child<MODE>(cx, cy, cz)) continue;
          const int z2 = 2 * cz, y2 = 2 * cy, x2 = 2 * cx; // the pI2 source of this child
          const float e = ax[cx] * (ay[cy] * (az[cz] * v[1][1][1] + bz[cz] * v[z2][1][1]) +
                                    by[cy] * (az[cz] * v[1][y2][1] + bz[cz] * v[z2][y2][1])) +
                          bx[cx] * (ay[cy] * (az[cz] * v[1][1][x2] + bz[cz] * v[z2][1][x2]) +
                                    by[cy] * (az[cz] * v[1][y2][x2] + bz[cz] * v[z2][y2][x2]));
          if (!skip) F.x[lidx(F, 2 * I - 1 + cx, 2 * J - 1 + cy, 2 * K - 1 + cz)] = xf[cz][cy][cx] + e;
        }
  }
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

__global__ void __launch_bounds__(kTailThreads) k_mg_tail(const float *__restrict__ cimg, TailDesc T, int nu1,
                                                          int nu2, int coarse_sweeps, const int *done) {
  const bool skip = done && *done; // tested once the entry loads are in flight
  extern __shared__ float sm[];
  const int nt = T.nt;
  // f of the first tail level (from the colour-split level array) and the
  // coefficient blocks: independent loads, all in flight while LDS is cleared
  const int n0 = T.n[0][0], n1 = T.n[0][1], nf0 = n0 * n1 * T.n[0][2];
  const int h0 = (n0 + 1) / 2, half = h0 * n1 * T.n[0][2];
  float fr[kTailLoads], cr[kTailLoads];
#pragma unroll
  for (int q = 0; q < kTailLoads; ++q) {
    const int t = q * kTailThreads + threadIdx.x;
    fr[q] = 0.f;
    cr[q] = 0.f;
    if (t < nf0) {
      const int i = 1 + t % n0, j = 1 + (t / n0) % n1, k = 1 + t / (n0 * n1);
      fr[q] = T.f0[((i + j + k) & 1) * half + ((i - 1) >> 1) + h0 * ((j - 1) + n1 * (k - 1))];
    }
    if (t < T.ncoef) cr[q] = cimg[t];
  }
  for (int t = threadIdx.x; t < T.nfloat; t += kTailThreads) sm[t] = 0.f;
  __syncthreads();
  if (skip) return;
#pragma unroll
  for (int q = 0; q < kTailLoads; ++q) {
    const int t = q * kTailThreads + threadIdx.x;
    if (t < nf0) {
      const int i = 1 + t % n0, j = 1 + (t / n0) % n1, k = 1 + t / (n0 * n1);
      sm[T.fo[0] + TIDX(0, i, j, k)] = fr[q];
    }
    if (t < T.ncoef) {
      int l = 0;
      while (l + 1 < nt && t >= T.coff[l + 1]) ++l;
      sm[T.co[l] + (t - T.coff[l])] = cr[q];
    }
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
    // with post-sweeps to follow, only the red-black colour 0 is corrected:
    // the first post half-sweep rewrites colour 1 without reading it (a
    // wall's clamped neighbour has weight zero)
    const bool colour0 = nu2 > 0;
    const int nf = colour0 ? ((n0 + 1) / 2) * n1 * n2 : n0 * n1 * n2;
    for (int t = threadIdx.x; t < nf; t += kTailThreads) {
      int i = 1 + t % n0, j = 1 + (t / n0) % n1, k = 1 + t / (n0 * n1);
      if (colour0 && !tail_color_cell(n0, n1, n2, t, 0, i, j, k)) continue;
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
  for (int t = threadIdx.x; t < nf0; t += kTailThreads) {
    const int i = 1 + t % n0, j = 1 + (t / n0) % n1, k = 1 + t / (n0 * n1);
    T.x0[((i + j + k) & 1) * half + ((i - 1) >> 1) + h0 * ((j - 1) + n1 * (k - 1))] = sm[T.xo[0] + TIDX(0, i, j, k)];
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
  for (int f = 0; f < 6; ++f)
    if (bct[f] == MOONS_BC_ROBIN) {
      // the Robin ghost factor depends on the boundary cell width of each level
      if (err) *err = "multigrid: Robin BCs are not supported (use the FP64 Jacobi-PCG)";
      return nullptr;
    }
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
    L.h0 = (L.n0 + 1) / 2;
    L.half = (int)((size_t)L.h0 * L.n1 * L.n2);
    L.per0 = bct[0] == MOONS_BC_PERIODIC;
    L.per1 = bct[2] == MOONS_BC_PERIODIC;
    L.per2 = bct[4] == MOONS_BC_PERIODIC;
    L.line1 = L.line2 = 0;
    const size_t ncell = (size_t)2 * ((L.n0 + 1) / 2) * L.n1 * L.n2;
    if (ncell >= ((size_t)1 << 31)) {
      if (err) *err = "multigrid: a level of 2^31 or more cells is not supported";
      mg_destroy(h);
      return nullptr;
    }
    L.x = dalloc(ncell);
    L.xalt = dalloc(ncell);
    L.f = dalloc(ncell);
    if (!L.x || !L.xalt || !L.f) {
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
      if (d > 0) {
        // zebra lines along a stretched (cell widths varying by more than 1.5x),
        // non-periodic y or z direction of a level of more than 32^3 cells
        // (below that the point sweeps' single fused launches cost less and
        // the iteration count does not change) whose lines fit one block
        double dmin = dhn[1], dmax = dhn[1];
        for (int i = 1; i <= n; ++i) {
          dmin = std::min(dmin, dhn[i]);
          dmax = std::max(dmax, dhn[i]);
        }
        const long long cells = (long long)hl.n[0] * hl.n[1] * hl.n[2];
        const bool per = bct[2 * d] == MOONS_BC_PERIODIC;
        // (the line kernels address the level by 32-bit byte offsets: below 2^30 cells)
        const int ln = !per && n >= 4 && n <= kLineMax && cells > line_min_cells() && cells < (1LL << 30) &&
                       dmax > 1.5 * dmin;
        (d == 1 ? L.line1 : L.line2) = ln;
      }
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
  for (int l = h->tail0; l < nl; ++l) h->L[l].line1 = h->L[l].line2 = 0; // the tail smooths by points
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
    // the coefficient blocks of the tail levels, packed (the kernel places
    // each at its LDS offset co)
    int nco = 0;
    for (int q = 0; q < T.nt; ++q) {
      T.coff[q] = nco;
      nco += 24 * T.cs[q];
    }
    T.ncoef = nco;
    if (T.n[0][0] * T.n[0][1] * T.n[0][2] > kTailLoads * kTailThreads || nco > kTailLoads * kTailThreads) {
      if (err) *err = "multigrid: internal error (tail larger than its entry loads)";
      mg_destroy(h);
      return nullptr;
    }
    std::vector<float> img((size_t)std::max(nco, 1), 0.f);
    for (int q = 0; q < T.nt; ++q)
      for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 8; ++c) {
          const std::vector<float> &a = h->hcoef[((h->tail0 + q) * 3 + d) * 8 + c];
          std::copy(a.begin(), a.end(), img.begin() + T.coff[q] + (c * 3 + d) * T.cs[q]);
        }
    if (T.nt > 0) {
      T.f0 = h->L[h->tail0].f;
      T.x0 = h->L[h->tail0].x;
    }
    h->tailimg = dalloc(img.size());
    if (!h->tailimg) {
      if (err) *err = "multigrid: hipMalloc failed";
      mg_destroy(h);
      return nullptr;
    }
    (void)hipMemcpyAsync(h->tailimg, img.data(), img.size() * sizeof(float), hipMemcpyHostToDevice, stream());
    (void)hipStreamSynchronize(stream()); // img and the coefficient vectors go out of scope
  }
  return h;
}

void mg_destroy(MGHier *h) {
  if (!h) return;
  if (stream()) (void)hipStreamSynchronize(stream());
  if (h->learn) unet_prec_destroy(h->learn);
  if (h->gpart) (void)hipFree(h->gpart);
  if (h->lpf) (void)hipFree(h->lpf);
  if (h->lpx) (void)hipFree(h->lpx);
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
void mg_f32_map(const MGHier *h, int map[3]) {
  map[0] = h->L[0].h0;
  map[1] = h->L[0].n1;
  map[2] = h->L[0].half;
}
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
  const size_t npad = (size_t)(L.n0 + 2) * (L.n1 + 2) * (L.n2 + 2);
  if (h->lpf) (void)hipFree(h->lpf);
  if (h->lpx) (void)hipFree(h->lpx);
  h->lpf = h->lpx = nullptr;
  if (hipMalloc(&h->lpf, npad * sizeof(float)) != hipSuccess || hipMalloc(&h->lpx, npad * sizeof(float)) != hipSuccess) {
    if (err) *err = "hipMalloc failed";
    return MOONS_ERR_HIP;
  }
  (void)hipMemsetAsync(h->lpf, 0, npad * sizeof(float), stream());
  (void)hipMemsetAsync(h->lpx, 0, npad * sizeof(float), stream());
  h->learn_level = level;
  return MOONS_OK;
}

void mg_set_params(MGHier *h, int nu1, int nu2, int coarse_sweeps) {
  h->nu1 = nu1;
  h->nu2 = nu2;
  h->coarse_sweeps = coarse_sweeps;
}



// Launch shape of a per-level grid kernel over nx by ny by nz points: a
// 256-thread tile no wider than the row needs (so narrow levels leave no lanes
// idle), and the largest z-march (1, 2 or 4 planes) that still gives at least
// 2048 tiles (8 per CU).
struct TileLaunch {
  dim3 grid, block;
  int kz;
};

static TileLaunch tiles(int nx, int ny, int nz, int bx_max = 64) {
  int bx = bx_max;
  while (bx > 8 && bx / 2 >= nx) bx /= 2;
  const int by = 256 / bx;
  TileLaunch t;
  t.block = dim3(bx, by);
  auto grid = [&](int kz) { return dim3((nx + bx - 1) / bx, (ny + by - 1) / by, (nz + kz - 1) / kz); };
  t.kz = 1;
  while (t.kz < 4) {
    const dim3 g = grid(2 * t.kz);
    if ((long long)g.x * g.y * g.z < 2048) break;
    t.kz *= 2;
  }
  t.grid = grid(t.kz);
  return t;
}

static void smooth(const LevelDev &L, int color, bool zero_guess, const int *done) {
  // 32 x 8 tiles: a wave covers two rows, and the y-neighbour rows of a
  // tile are shared by more of its threads (measured faster than 64 x 4)
  const TileLaunch t = tiles((L.n0 + 1) / 2, L.n1, L.n2, 32);
  hipLaunchKernelGGL(k_mg_smooth, t.grid, t.block, 0, stream(), L, color, zero_guess ? 1 : 0, t.kz, done);
}

// Levels of at most this many cells are smoothed by fused whole sweeps (one
// launch each, x swapped between its two buffers); larger levels by two
// in-place half-sweep launches, whose loads the recomputation would multiply.
constexpr long long kFusedSweepMaxCells = 64LL * 64 * 64;

static void sweep2(LevelDev &L, int c_first, bool zero_first, const int *done) {
  const TileLaunch t = tiles((L.n0 + 1) / 2, L.n1, L.n2);
  hipLaunchKernelGGL(k_mg_sweep2, t.grid, t.block, 0, stream(), L, L.xalt, c_first, zero_first ? 1 : 0, t.kz, done);
  std::swap(L.x, L.xalt);
}

// Zebra line half-sweep of colour c along direction D of level L: C cells
// a thread, NP / C threads a line, TB threads a block.
template <int D, int NP, int TB, int C>
static void line_launch(const LevelDev &L, int c, int zg, const int *done, int o) {
  const int lpb = TB / (NP / C);
  const int gx = (L.h0 + lpb - 1) / lpb;
  hipLaunchKernelGGL((k_mg_line<D, NP, TB, C>), dim3((unsigned)(gx * o)), dim3(TB), 0, stream(), L, c, zg, done,
                     gx);
}

template <int D>
static void line_half_d(const LevelDev &L, int c, bool zero_guess, const int *done) {
  const int nl = D == 1 ? L.n1 : L.n2, o = D == 1 ? L.n2 : L.n1;
  const int zg = zero_guess ? 1 : 0;
  if (nl <= 16) line_launch<D, 16, 64, 4>(L, c, zg, done, o);
  else if (nl <= 32) line_launch<D, 32, 128, 4>(L, c, zg, done, o);
  else if (nl <= 64) line_launch<D, 64, 128, 4>(L, c, zg, done, o);
  else line_launch<D, kLineMax, 512, 8>(L, c, zg, done, o);
}

static void line_half(const LevelDev &L, int D, int c, bool zero_guess, const int *done) {
  if (D == 1) line_half_d<1>(L, c, zero_guess, done);
  else line_half_d<2>(L, c, zero_guess, done);
}

// One line sweep of level L: pre-smoothing (forward) y-lines colour 0, 1,
// then z-lines colour 0, 1; post-smoothing (backward) the reverse order, so
// the V-cycle stays a symmetric preconditioner.
// timer: the kernel-timer class of these launches (0: none).
static void line_sweep(LevelDev &L, bool forward, bool zero_first, const int *done, int timer) {
  int ds[2], n = 0;
  if (L.line1) ds[n++] = 1;
  if (L.line2) ds[n++] = 2;
  // one timer pair around the whole sweep: the event overhead spreads over its launches
  ktimer_mark(timer);
  if (forward) {
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < 2; ++c) line_half(L, ds[a], c, zero_first && a == 0 && c == 0, done);
  } else {
    for (int a = n - 1; a >= 0; --a)
      for (int c = 1; c >= 0; --c) line_half(L, ds[a], c, false, done);
  }
  ktimer_mark(timer, 2 * n);
}

// One red-black sweep of level L, colour c_first first (lines, where the
// level has line directions: c_first 0 is the forward order).
static void sweep(LevelDev &L, int c_first, bool zero_first, const int *done, int timer = 0) {
  if (L.line1 || L.line2) {
    line_sweep(L, c_first == 0, zero_first, done, timer);
    return;
  }
  if ((long long)L.n0 * L.n1 * L.n2 <= kFusedSweepMaxCells) {
    sweep2(L, c_first, zero_first, done);
  } else {
    smooth(L, c_first, zero_first, done);
    smooth(L, 1 - c_first, false, done);
  }
}

static void restrict_to(const LevelDev &F, const LevelDev &C, const int *done) {
  const TileLaunch t = tiles(C.n0, C.n1, C.n2);
  hipLaunchKernelGGL(k_mg_restrict, t.grid, t.block, 0, stream(), F, C, t.kz, done);
}

// post: the fine level's post-smoothing sweeps that follow (with none, every
// cell is corrected)
static void prolong_from(const LevelDev &F, const LevelDev &C, const int *done, int post) {
  const TileLaunch t = tiles(C.n0, C.n1, C.n2);
  // the first post-smoothing half-sweep: z-lines (or y-lines) of colour 1,
  // or the red-black colour 1
  const int mode = post < 1 ? 0 : F.line2 ? 1 : F.line1 ? 2 : 3;
  switch (mode) {
    case 1: hipLaunchKernelGGL(k_mg_prolong<1>, t.grid, t.block, 0, stream(), F, C, t.kz, done); break;
    case 2: hipLaunchKernelGGL(k_mg_prolong<2>, t.grid, t.block, 0, stream(), F, C, t.kz, done); break;
    case 3: hipLaunchKernelGGL(k_mg_prolong<3>, t.grid, t.block, 0, stream(), F, C, t.kz, done); break;
    default: hipLaunchKernelGGL(k_mg_prolong<0>, t.grid, t.block, 0, stream(), F, C, t.kz, done); break;
  }
}

// MOONS_MG_COARSE_NU="a,b[,l]": a pre- and b post-smoothing sweeps on level
// l (1 when not given) and every level below it, the finest level's counts
// elsewhere (tuning runs); -1: not set.
struct CoarseNu {
  int nu1 = -1, nu2 = -1, from = 1;
};

static CoarseNu coarse_nu() {
  CoarseNu c;
  if (const char *e = getenv("MOONS_MG_COARSE_NU")) {
    c.nu1 = atoi(e);
    if (const char *p = strchr(e, ',')) {
      c.nu2 = atoi(p + 1);
      if (const char *q = strchr(p + 1, ',')) c.from = std::max(1, atoi(q + 1));
    }
  }
  return c;
}

int mg_vcycle(MGHier *h, const int *done) {
  const int nl = (int)h->L.size();
  const bool tail = h->tail0 < nl;
  const int g = tail ? h->tail0 : nl - 1; // levels [0, g) run as grid launches
  // levels [c, g) and the coarse end are cycled K times per visit (a
  // gamma-cycle from level c; K = 1, c = g: the plain V-cycle)
  const int c = h->gamma_level >= 0 && h->gamma_level < g ? h->gamma_level : g;
  const int K = c < g ? std::max(1, h->gamma_cycles) : 1;
  // sweeps per level: nu1/nu2, or MOONS_MG_COARSE_NU's on its levels
  static const CoarseNu cnu = coarse_nu();
  auto nu1_at = [&](int l) { return l >= cnu.from && cnu.nu1 >= 0 ? cnu.nu1 : h->nu1; };
  auto nu2_at = [&](int l) { return l >= cnu.from && cnu.nu2 >= 0 ? cnu.nu2 : h->nu2; };
  auto descend = [&](int l, bool zero_first) -> int {
    LevelDev &L = h->L[l];
    const bool learned = zero_first && h->learn && l == h->learn_level &&
                         (h->learn_cycles == 0 || h->cycle < h->learn_cycles);
    if (learned) {
      // learned initial guess x = c N(f), then smoothing from it
      const int nc = L.n0 * L.n1 * L.n2;
      const unsigned nb = blocks_for(nc);
      hipLaunchKernelGGL(k_mg_to_padded, dim3(nb), dim3(256), 0, stream(), L, h->lpf, done);
      int st = unet_prec_apply(h->learn, h->lpf, h->lpx, done);
      if (st) return st;
      hipLaunchKernelGGL(k_mg_from_padded, dim3(nb), dim3(256), 0, stream(), L, h->lpx, done);
      hipLaunchKernelGGL(k_mg_galerkin_part, dim3(nb), dim3(256), 0, stream(), L, h->gpart, done);
      hipLaunchKernelGGL(k_mg_galerkin_scale, dim3(nb), dim3(256), 0, stream(), L, h->gpart, (int)nb, done);
    }
    for (int s = 0; s < nu1_at(l); ++s)
      sweep(L, 0, s == 0 && zero_first && !learned, done, l == 0 ? MOONS_KTIMER_MG_LINE_FINE : 0);
    restrict_to(L, h->L[l + 1], done);
    return MOONS_OK;
  };
  auto ascend = [&](int l) {
    LevelDev &L = h->L[l];
    prolong_from(L, h->L[l + 1], done, nu2_at(l));
    for (int s = 0; s < nu2_at(l); ++s) sweep(L, 1, false, done, l == 0 ? MOONS_KTIMER_MG_LINE_FINE : 0);
  };
  int st;
  for (int l = 0; l < c; ++l)
    if ((st = descend(l, true))) return st;
  for (int k = 0; k < K; ++k) {
    for (int l = c; l < g; ++l)
      if ((st = descend(l, l > c || k == 0))) return st;
    if (tail)
      hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), h->tail_lds, stream(), h->tailimg, h->tail,
                         nu1_at(g), nu2_at(g), h->coarse_sweeps, done);
    else
      hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(1024), 0, stream(), h->L[nl - 1], h->coarse_sweeps, done);
    for (int l = g - 1; l >= c; --l) ascend(l);
  }
  for (int l = c - 1; l >= 0; --l) ascend(l);
  ++h->cycle;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MOONS_ERR_HIP, std::string("multigrid launch: ") + hipGetErrorString(e));
  return MOONS_OK;
}

void mg_begin_solve(MGHier *h) { h->cycle = 0; }
void mg_set_learned_cycles(MGHier *h, int cycles) { h->learn_cycles = cycles; }
int mg_set_gamma(MGHier *h, int level, int cycles) {
  const int nl = (int)h->L.size();
  const int g = h->tail0 < nl ? h->tail0 : nl - 1; // grid-launched levels
  if (level >= g) return MOONS_ERR_ARG;
  h->gamma_level = level;
  h->gamma_cycles = cycles;
  return MOONS_OK;
}

int mg_time_kernels(MGHier *h, int reps, double ms[10]) {
  const int nl = (int)h->L.size();
  hipEvent_t e[2];
  for (int i = 0; i < 2; ++i)
    if (hipEventCreate(&e[i]) != hipSuccess) return fail(MOONS_ERR_HIP, "hipEventCreate failed");
  int st = MOONS_OK;
  for (int w = 0; w < 10 && st == MOONS_OK; ++w) {
    ms[w] = 0.0;
    if ((w == 1 || w == 2) && nl < 2) continue;
    // 5, 7, 9: a fused whole sweep of level 0, 1, 2; 6, 8: a half-sweep of level 1, 2
    const int lw = w == 5 ? 0 : (w == 6 || w == 7) ? 1 : 2;
    const int g = h->tail0 < nl ? h->tail0 : nl - 1; // grid-launched levels
    if (w >= 5 && lw >= g) continue;
    // levels smoothed by lines: their half-sweep entry times a line half-sweep
    // (the first line direction), and they have no fused sweeps
    auto lines = [&](int l) { return l < g && (h->L[l].line1 || h->L[l].line2); };
    auto ldir = [&](int l) { return h->L[l].line1 ? 1 : 2; };
    if ((w == 5 || w == 7 || w == 9) && lines(lw)) continue;
    auto launch = [&]() -> int {
      switch (w) {
        case 0:
          if (lines(0)) line_half(h->L[0], ldir(0), 0, false, nullptr);
          else smooth(h->L[0], 0, false, nullptr);
          return MOONS_OK;
        case 1: restrict_to(h->L[0], h->L[1], nullptr); return MOONS_OK;
        case 2: prolong_from(h->L[0], h->L[1], nullptr, h->nu2); return MOONS_OK;
        case 5: case 7: case 9: sweep2(h->L[lw], 0, false, nullptr); return MOONS_OK;
        case 6: case 8:
          if (lines(lw)) line_half(h->L[lw], ldir(lw), 0, false, nullptr);
          else smooth(h->L[lw], 0, false, nullptr);
          return MOONS_OK;
        case 3:
          if (h->tail0 < nl)
            hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), h->tail_lds, stream(), h->tailimg, h->tail,
                               h->nu1, h->nu2, h->coarse_sweeps, nullptr);
          else
            hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(1024), 0, stream(), h->L[nl - 1], h->coarse_sweeps, nullptr);
          return MOONS_OK;
        case 4: return mg_vcycle(h, nullptr);
        default: return MOONS_OK;
      }
    };
    for (int r = 0; r < 2 && st == MOONS_OK; ++r) st = launch(); // warm-up
    if (st) break;
    (void)hipEventRecord(e[0], stream());
    for (int r = 0; r < reps && st == MOONS_OK; ++r) st = launch();
    (void)hipEventRecord(e[1], stream());
    if (st) break;
    hipError_t he = hipEventSynchronize(e[1]);
    float t = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e[0], e[1]);
    if (he != hipSuccess) {
      st = fail(MOONS_ERR_HIP, std::string("multigrid kernel timing: ") + hipGetErrorString(he));
      break;
    }
    ms[w] = (double)t / reps;
    // the fused sweeps swapped x of level lw (reps + 2) times: restore its buffer
    if ((w == 5 || w == 7 || w == 9) && (reps % 2) != 0) std::swap(h->L[lw].x, h->L[lw].xalt);
  }
  for (int i = 0; i < 2; ++i) (void)hipEventDestroy(e[i]);
  return st;
}

} // namespace moons
