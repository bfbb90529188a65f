// unet.hip -- learned preconditioner for the pressure-Poisson PCG: a 3-level
// 3D U-net (16/32/64 channels, 3x3x3 convolutions, ReLU, 2x2x2 average
// pooling, nearest upsampling with skip concatenation), BF16 weights and
// activations, FP32 accumulation on the MFMA units, trained on the GPU by its
// own forward and backward kernels (Adam on FP32 master weights).
//
// No counterpart in the reference (its preconditioners are the Jacobi ones of
// code/handwritten/solvers/preconditioners/preconditioners.f90:55-129); it
// plugs into the flexible PCG like the FP32 multigrid.
//
// Layout: activations are channels-last BF16 with a zero halo of one cell,
// index ((k*(n1+2) + j)*(n0+2) + i)*C + c (i,j,k padded).  A convolution is an
// implicit GEMM: M = voxels (16 per MFMA tile), N = output channels, K = 27
// taps x input channels; v_mfma_f32_16x16x16_bf16 with A = 16 voxels x 16
// input channels (one 8-byte load per lane), B = weights pre-packed in
// fragment order.  Layers with 4 input or 1 output channel are plain FP32
// kernels (0.4% of the flops).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "unet.hpp"

namespace moons {

typedef unsigned short bf16;
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u); // round to nearest even (finite values)
  return (bf16)(u >> 16);
}
static inline float h_bf2f(bf16 h) {
  unsigned u = ((unsigned)h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static inline bf16 h_f2bf(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

constexpr int kUT = 256;
static inline unsigned ublk(size_t n) { return (unsigned)((n + kUT - 1) / kUT); }

__device__ __forceinline__ bool gated(const int *done) { return done && *done; }

// padded linear voxel index of interior voxel v (0-based over n0*n1*n2)
__device__ __forceinline__ int pvox(int v, int n0, int n1) {
  const int i = v % n0, j = (v / n0) % n1, k = v / (n0 * n1);
  return ((k + 1) * (n1 + 2) + (j + 1)) * (n0 + 2) + (i + 1);
}

// ------------------------------------------------------------ MFMA conv
// Y = act(conv(X) + bias), Cout = 16*NB.  Each wave: 32 voxels (2 M-tiles) x
// all output channels.  WB: fragments ((tap*KB + kb)*NB + nb)*64 + lane.
template <int NB>
__global__ void __launch_bounds__(kUT) k_conv_mfma(const bf16 *__restrict__ X, int Cin, const s16x4 *__restrict__ WB,
                                                  const float *__restrict__ bias, bf16 *__restrict__ Y, int n0, int n1,
                                                  int n2, int relu, const int *done) {
  if (gated(done)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nvox = n0 * n1 * n2;
  const int base = (blockIdx.x * (kUT / 64) + wave) * 32;
  if (base >= nvox) return;
  const int s0 = n0 + 2, s1 = n1 + 2;
  int vb[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const int v = base + mt * 16 + (lane & 15);
    vb[mt] = pvox(v < nvox ? v : 0, n0, n1);
  }
  const int kq = (lane >> 4) * 4;
  f32x4 acc[2][NB];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mt][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int KB = Cin >> 4;
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int off = (dz * s1 + dy) * s0 + dx;
    const bf16 *x0 = X + (size_t)(vb[0] + off) * Cin + kq;
    const bf16 *x1 = X + (size_t)(vb[1] + off) * Cin + kq;
    const s16x4 *wt = WB + (size_t)tap * KB * NB * 64 + lane;
    for (int kb = 0; kb < KB; ++kb) {
      const s16x4 a0 = *reinterpret_cast<const s16x4 *>(x0 + kb * 16);
      const s16x4 a1 = *reinterpret_cast<const s16x4 *>(x1 + kb * 16);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const s16x4 b = wt[(kb * NB + nb) * 64];
        acc[0][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b, acc[0][nb], 0, 0, 0);
        acc[1][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, acc[1][nb], 0, 0, 0);
      }
    }
  }
  const int Cout = NB * 16;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = base + mt * 16 + (lane >> 4) * 4 + j;
      if (v >= nvox) continue;
      const int pv = pvox(v, n0, n1);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int co = nb * 16 + (lane & 15);
        float o = acc[mt][nb][j] + (bias ? bias[co] : 0.f);
        if (relu) o = fmaxf(o, 0.f);
        Y[(size_t)pv * Cout + co] = f2bf(o);
      }
    }
}

// ------------------------------------------------ LDS-tiled MFMA conv
// The finest level's convolutions (n0 a multiple of 16) on gfx950's
// v_mfma_f32_16x16x32_bf16.  A workgroup (4 waves) owns a 16 x 2 x 2 voxel
// tile: it stages the tile's input halo (18 x 4 x 4 voxels, all Cin = 16 CG
// channels) in LDS once with 16-byte coalesced loads, then each wave computes
// one 16-voxel x-row for all Cout = 16 NB channels.  K runs over the
// 27 CG 16-channel groups (tap-major), two groups per MFMA: lane l holds
// A[voxel l % 16][channels 8 (l / 16 % 2) ..+8 of group 2c + l / 32] read
// from LDS as one 16-byte word, and B pre-packed in the same k order
// (WB32[(c NB + nb) 64 + lane]); an odd group count pads the last MFMA with
// zeros.  Every input voxel is read from HBM once per tile (1.5x for the
// halo) instead of 27 times through L2; the 16 -> 16 layer's weight
// fragments are staged in LDS beside the halo, so its MFMA chain reads only
// LDS, and a workgroup loops over several tiles along z on large grids so
// that they are staged once for all of them.
constexpr int kCTX = 16, kCTY = 2, kCTZ = 2; // output tile (one row of 16 voxels a wave)
constexpr int kHX = kCTX + 2, kHY = kCTY + 2, kHZ = kCTZ + 2;

template <int CG, int NB>
__global__ void __launch_bounds__(256) k_conv_tile(const bf16 *__restrict__ X, const u16x8 *__restrict__ WB,
                                                  const float *__restrict__ bias, bf16 *__restrict__ Y, int n0,
                                                  int n1, int n2, int relu, int zt, const int *done) {
  if (gated(done)) return;
  constexpr int CIN = 16 * CG, COUT = 16 * NB;
  constexpr int V8 = CIN / 8;                  // 16-byte words per voxel
  constexpr int NH = kHX * kHY * kHZ;          // halo voxels
  constexpr int G = 27 * CG, NC = (G + 1) / 2; // channel groups, MFMA steps
  constexpr int NW = NC * NB * 64;             // weight fragments
  // the weights join the halo in LDS when both fit in 32 KB (the 16 -> 16
  // layer; the larger ones keep their occupancy and read the weights via L2)
  constexpr bool WLDS = (NH * V8 + NW) * 16 <= 32 * 1024;
  __shared__ u16x8 sx[NH * V8 + (WLDS ? NW : 0)];
  const int x0 = blockIdx.x * kCTX, y0 = blockIdx.y * kCTY;
  const int s0 = n0 + 2, s1 = n1 + 2;
  if constexpr (WLDS)
    for (int q = threadIdx.x; q < NW; q += 256) sx[NH * V8 + q] = WB[q];
  const u16x8 *wb = WLDS ? sx + NH * V8 : WB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ry = wave & 1, rz = wave >> 1, m = lane & 15, kh = (lane >> 4) & 1, gs = lane >> 5;
  // the workgroup's zt tiles along z, one after another (the weights staged once)
  for (int t = 0; t < zt; ++t) {
  const int z0 = (blockIdx.z * zt + t) * kCTZ;
  if (t > 0) __syncthreads(); // the previous tile's halo reads are done
  // stage the halo: padded voxel (x0 + hx, y0 + hy, z0 + hz) is interior
  // voxel (x0 + hx - 1, ...), so the halo's first cell is the padded origin
  for (int q = threadIdx.x; q < NH * V8; q += 256) {
    const int v = q / V8, w = q - v * V8;
    const int hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
    const size_t pv = ((size_t)(z0 + hz) * s1 + (y0 + hy)) * s0 + (x0 + hx);
    sx[q] = *reinterpret_cast<const u16x8 *>(X + pv * CIN + w * 8);
  }
  __syncthreads();
  f32x4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int g = 2 * c + gs;
    u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g < G) {
      const int tap = g / CG, cg = g - tap * CG;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      a = sx[(((rz + dz) * kHY + (ry + dy)) * kHX + (m + dx)) * V8 + cg * 2 + kh];
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const u16x8 b = wb[(c * NB + nb) * 64 + lane];
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                        acc[nb], 0, 0, 0);
    }
  }
  const size_t prow = ((size_t)(z0 + rz + 1) * s1 + (y0 + ry + 1)) * s0 + (x0 + 1);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const size_t pv = prow + (lane >> 4) * 4 + j;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = nb * 16 + m;
      float o = acc[nb][j] + (bias ? bias[co] : 0.f);
      if (relu) o = fmaxf(o, 0.f);
      Y[pv * COUT + co] = f2bf(o);
    }
  }
  }
}

// The same convolution with its 27 taps split into groups, for layers whose
// few voxels give too few waves to hide the load latency of the tap loop:
// blockIdx.y = group g takes taps [g T, min(27, g T + T)) and writes its FP32
// partial sums P[(g nvox + v) Cout + co]; k_conv_reduce adds the groups in
// order, then the bias and the activation.
template <int NB>
__global__ void __launch_bounds__(kUT) k_conv_mfma_split(const bf16 *__restrict__ X, int Cin,
                                                        const s16x4 *__restrict__ WB, float *__restrict__ P, int n0,
                                                        int n1, int n2, int T, const int *done) {
  if (gated(done)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nvox = n0 * n1 * n2;
  const int base = (blockIdx.x * (kUT / 64) + wave) * 32;
  if (base >= nvox) return;
  const int g = blockIdx.y;
  const int t0 = g * T, t1 = min(27, t0 + T);
  const int s0 = n0 + 2, s1 = n1 + 2;
  int vb[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const int v = base + mt * 16 + (lane & 15);
    vb[mt] = pvox(v < nvox ? v : 0, n0, n1);
  }
  const int kq = (lane >> 4) * 4;
  f32x4 acc[2][NB];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mt][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int KB = Cin >> 4;
  for (int tap = t0; tap < t1; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int off = (dz * s1 + dy) * s0 + dx;
    const bf16 *x0 = X + (size_t)(vb[0] + off) * Cin + kq;
    const bf16 *x1 = X + (size_t)(vb[1] + off) * Cin + kq;
    const s16x4 *wt = WB + (size_t)tap * KB * NB * 64 + lane;
    for (int kb = 0; kb < KB; ++kb) {
      const s16x4 a0 = *reinterpret_cast<const s16x4 *>(x0 + kb * 16);
      const s16x4 a1 = *reinterpret_cast<const s16x4 *>(x1 + kb * 16);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const s16x4 b = wt[(kb * NB + nb) * 64];
        acc[0][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b, acc[0][nb], 0, 0, 0);
        acc[1][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, acc[1][nb], 0, 0, 0);
      }
    }
  }
  const int Cout = NB * 16;
  float *Pg = P + (size_t)g * nvox * Cout;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = base + mt * 16 + (lane >> 4) * 4 + j;
      if (v >= nvox) continue;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) Pg[(size_t)v * Cout + nb * 16 + (lane & 15)] = acc[mt][nb][j];
    }
}

__global__ void __launch_bounds__(kUT) k_conv_reduce(const float *__restrict__ P, int S, int nvox, int Cout,
                                                    const float *__restrict__ bias, int relu, bf16 *__restrict__ Y,
                                                    int n0, int n1, const int *done) {
  if (gated(done)) return;
  const size_t n = (size_t)nvox * Cout;
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= n) return;
  float o = 0.f;
#pragma unroll 9
  for (int g = 0; g < S; ++g) o += P[(size_t)g * n + t];
  const int co = (int)(t % Cout), v = (int)(t / Cout);
  if (bias) o += bias[co];
  if (relu) o = fmaxf(o, 0.f);
  Y[(size_t)pvox(v, n0, n1) * Cout + co] = f2bf(o);
}

// dW partials: part[chunk][tap][ci][co] = sum over the chunk's voxels of
// X[v + tap][ci] * dY[v][co] (MFMA with M = ci, N = co, K = voxels).
template <int NB>
__global__ void __launch_bounds__(64) k_conv_dw(const bf16 *__restrict__ X, int Cin, const bf16 *__restrict__ dY,
                                               int n0, int n1, int n2, int chunk, float *__restrict__ part) {
  const int lane = threadIdx.x;
  const int c = blockIdx.x;
  const int KB = Cin >> 4;
  const int tap = blockIdx.y / KB, kb = blockIdx.y % KB;
  const int nvox = n0 * n1 * n2;
  const int s0 = n0 + 2, s1 = n1 + 2;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int off = (dz * s1 + dy) * s0 + dx;
  const int Cout = NB * 16;
  f32x4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int v0 = c * chunk, v1 = min(nvox, v0 + chunk);
  const int ci = kb * 16 + (lane & 15);
  for (int vs = v0; vs < v1; vs += 16) {
    s16x4 a;
    int pv[4];
    bool ok[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = vs + (lane >> 4) * 4 + j;
      ok[j] = v < v1;
      pv[j] = pvox(ok[j] ? v : 0, n0, n1);
      a[j] = ok[j] ? (short)X[(size_t)(pv[j] + off) * Cin + ci] : (short)0;
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      s16x4 b;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = ok[j] ? (short)dY[(size_t)pv[j] * Cout + nb * 16 + (lane & 15)] : (short)0;
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nb], 0, 0, 0);
    }
  }
  float *out = part + ((size_t)c * 27 + tap) * Cin * Cout;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[(size_t)(kb * 16 + (lane >> 4) * 4 + j) * Cout + nb * 16 + (lane & 15)] = acc[nb][j];
}

// dW (+)= sum over chunks of the partials, in chunk order
__global__ void __launch_bounds__(kUT) k_reduce_chunks(const float *part, int nchunk, size_t n, float *dW, int accumulate) {
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= n) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * n + t];
  dW[t] = accumulate ? dW[t] + s : s;
}

// db (+)= sum over voxels of dY[v][co]; one block per output channel
__global__ void __launch_bounds__(kUT) k_dbias(const bf16 *dY, int C, int n0, int n1, int n2, float *db, int accumulate) {
  __shared__ float sh[kUT];
  const int co = blockIdx.x;
  const int nvox = n0 * n1 * n2;
  float s = 0.f;
  for (int v = threadIdx.x; v < nvox; v += kUT) s += bf2f(dY[(size_t)pvox(v, n0, n1) * C + co]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kUT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = accumulate ? db[co] + sh[0] : sh[0];
}

// ------------------------------------------------------ small-channel layers
// first layer: Y[v][0..15] = relu(b + sum_tap,ci W[tap][ci][co] X[v+tap][ci]), Cin = 4
__global__ void __launch_bounds__(kUT) k_conv_in(const bf16 *__restrict__ X, const float *__restrict__ W,
                                                const float *__restrict__ bias, bf16 *__restrict__ Y, int n0, int n1,
                                                int n2, const int *done) {
  // a thread per voxel and output channel (the channel fastest: the 16
  // threads of a voxel share its input loads and store 32 contiguous bytes)
  if (gated(done)) return;
  const int t = blockIdx.x * kUT + threadIdx.x;
  const int v = t >> 4, co = t & 15;
  if (v >= n0 * n1 * n2) return;
  const int s0 = n0 + 2, s1 = n1 + 2;
  const int pv = pvox(v, n0, n1);
  float acc = bias[co];
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    // the voxel's 4 channels in one 8-byte load
    const uint2 q = *reinterpret_cast<const uint2 *>(X + (size_t)(pv + (dz * s1 + dy) * s0 + dx) * 4);
    const unsigned qw[2] = {q.x, q.y};
#pragma unroll
    for (int ci = 0; ci < 4; ++ci) {
      const float xv = __uint_as_float((ci & 1) ? (qw[ci >> 1] & 0xFFFF0000u) : (qw[ci >> 1] << 16));
      acc += xv * W[(tap * 4 + ci) * 16 + co];
    }
  }
  Y[(size_t)pv * 16 + co] = f2bf(fmaxf(acc, 0.f));
}

// last layer: out[pv] = scale * (b + sum_ci sum_tap W[tap][ci] X[v+tap][ci]),
// Cin = 16, FP32 padded output.  16 threads a voxel, one per input channel,
// each summing its channel over the taps; the 16 sums are added in a fixed
// butterfly (deterministic) and lane ci = 0 writes.
__global__ void __launch_bounds__(kUT) k_conv_out(const bf16 *__restrict__ X, const float *__restrict__ W,
                                                 const float *__restrict__ bias, float *__restrict__ out, int n0, int n1,
                                                 int n2, const float *scale, const int *done) {
  if (gated(done)) return;
  const int t = blockIdx.x * kUT + threadIdx.x;
  const int v = t >> 4, ci = t & 15;
  const bool live = v < n0 * n1 * n2;
  const int s0 = n0 + 2, s1 = n1 + 2;
  const int pv = pvox(live ? v : 0, n0, n1);
  float acc = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    acc += bf2f(X[(size_t)(pv + (dz * s1 + dy) * s0 + dx) * 16 + ci]) * W[tap * 16 + ci];
  }
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 16);
  if (live && ci == 0) {
    const float r = bias[0] + acc;
    out[pv] = scale ? r * (*scale) : r;
  }
}

// last layer backward: dX[v][ci] = sum_tap W[tap][ci] * dout[v - tap] (dout zero outside)
__global__ void __launch_bounds__(kUT) k_conv_out_dx(const float *__restrict__ dout, const float *__restrict__ W,
                                                    bf16 *__restrict__ dX, int n0, int n1, int n2) {
  const int v = blockIdx.x * kUT + threadIdx.x;
  if (v >= n0 * n1 * n2) return;
  const int s0 = n0 + 2, s1 = n1 + 2;
  const int pv = pvox(v, n0, n1);
  float acc[16];
#pragma unroll
  for (int ci = 0; ci < 16; ++ci) acc[ci] = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const float g = dout[pv - ((dz * s1 + dy) * s0 + dx)]; // halo of dout is zero
    const float *w = W + tap * 16;
#pragma unroll
    for (int ci = 0; ci < 16; ++ci) acc[ci] += w[ci] * g;
  }
#pragma unroll
  for (int ci = 0; ci < 16; ++ci) dX[(size_t)pv * 16 + ci] = f2bf(acc[ci]);
}

// small-channel dW partials (first and last layers): one block per (chunk, tap),
// each thread a voxel range, CI x CO sums reduced across the block.
// X: CI channels BF16; G: CO channels, BF16 (gb) or FP32 padded scalar (gf, CO = 1).
template <int CI, int CO>
__global__ void __launch_bounds__(kUT) k_small_dw(const bf16 *__restrict__ X, const bf16 *__restrict__ gb,
                                                 const float *__restrict__ gf, int n0, int n1, int n2, int chunk,
                                                 float *__restrict__ part) {
  __shared__ float sh[kUT];
  const int c = blockIdx.x, tap = blockIdx.y;
  const int nvox = n0 * n1 * n2;
  const int s0 = n0 + 2, s1 = n1 + 2;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int off = (dz * s1 + dy) * s0 + dx;
  float acc[CI * CO];
#pragma unroll
  for (int q = 0; q < CI * CO; ++q) acc[q] = 0.f;
  const int v0 = c * chunk, v1 = min(nvox, v0 + chunk);
  for (int v = v0 + threadIdx.x; v < v1; v += kUT) {
    const int pv = pvox(v, n0, n1);
    float g[CO];
#pragma unroll
    for (int co = 0; co < CO; ++co) g[co] = gf ? gf[pv] : bf2f(gb[(size_t)pv * CO + co]);
#pragma unroll
    for (int ci = 0; ci < CI; ++ci) {
      const float xv = bf2f(X[(size_t)(pv + off) * CI + ci]);
#pragma unroll
      for (int co = 0; co < CO; ++co) acc[ci * CO + co] += xv * g[co];
    }
  }
  float *out = part + ((size_t)c * 27 + tap) * CI * CO;
#pragma unroll
  for (int q = 0; q < CI * CO; ++q) {
    sh[threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = kUT / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
    __syncthreads();
  }
}

// --------------------------------------------------------- data movement
// 2x2x2 average pooling (n -> n/2)
__global__ void __launch_bounds__(kUT) k_pool(const bf16 *X, bf16 *Y, int C, int N0, int N1, int N2, const int *done) {
  if (gated(done)) return;
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  const size_t n = (size_t)N0 * N1 * N2 * C;
  if (t >= n) return;
  const int c = (int)(t % C);
  const int V = (int)(t / C);
  const int I = V % N0, J = (V / N0) % N1, K = V / (N0 * N1);
  const int n0 = 2 * N0, n1 = 2 * N1;
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
    s += bf2f(X[(size_t)(((k + 1) * (n1 + 2) + (j + 1)) * (n0 + 2) + (i + 1)) * C + c]);
  }
  Y[(size_t)pvox(V, N0, N1) * C + c] = f2bf(0.125f * s);
}

// pooling backward: dX[child] (+)= dY[parent] / 8
__global__ void __launch_bounds__(kUT) k_pool_bwd(const bf16 *dY, bf16 *dX, int C, int n0, int n1, int n2, int accumulate) {
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  const size_t n = (size_t)n0 * n1 * n2 * C;
  if (t >= n) return;
  const int c = (int)(t % C);
  const int v = (int)(t / C);
  const int i = v % n0, j = (v / n0) % n1, k = v / (n0 * n1);
  const int N0 = n0 / 2, N1 = n1 / 2;
  const float g = 0.125f * bf2f(dY[(size_t)(((k / 2 + 1) * (N1 + 2) + (j / 2 + 1)) * (N0 + 2) + (i / 2 + 1)) * C + c]);
  const size_t o = (size_t)pvox(v, n0, n1) * C + c;
  dX[o] = f2bf(accumulate ? bf2f(dX[o]) + g : g);
}

// Y[v] = [ up(U)[v] (CU channels) , S[v] (CS channels) ] at the fine level
__global__ void __launch_bounds__(kUT) k_upcat(const bf16 *U, int CU, const bf16 *S, int CS, bf16 *Y, int n0, int n1,
                                              int n2, const int *done) {
  if (gated(done)) return;
  const int C = CU + CS;
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= (size_t)n0 * n1 * n2 * C) return;
  const int c = (int)(t % C);
  const int v = (int)(t / C);
  const int i = v % n0, j = (v / n0) % n1, k = v / (n0 * n1);
  const int pv = pvox(v, n0, n1);
  bf16 val;
  if (c < CU) {
    const int N0 = n0 / 2, N1 = n1 / 2;
    val = U[(size_t)(((k / 2 + 1) * (N1 + 2) + (j / 2 + 1)) * (N0 + 2) + (i / 2 + 1)) * CU + c];
  } else {
    val = S[(size_t)pv * CS + (c - CU)];
  }
  Y[(size_t)pv * C + c] = val;
}

// upcat backward: dU[V][c] = sum of the 8 children's dY[child][c]; dS[v][c] = dY[v][CU + c]
__global__ void __launch_bounds__(kUT) k_upcat_bwd_up(const bf16 *dY, int C, bf16 *dU, int CU, int N0, int N1, int N2) {
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= (size_t)N0 * N1 * N2 * CU) return;
  const int c = (int)(t % CU);
  const int V = (int)(t / CU);
  const int I = V % N0, J = (V / N0) % N1, K = V / (N0 * N1);
  const int n0 = 2 * N0, n1 = 2 * N1;
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
    s += bf2f(dY[(size_t)(((k + 1) * (n1 + 2) + (j + 1)) * (n0 + 2) + (i + 1)) * C + c]);
  }
  dU[(size_t)pvox(V, N0, N1) * CU + c] = f2bf(s);
}

__global__ void __launch_bounds__(kUT) k_upcat_bwd_skip(const bf16 *dY, int C, int CU, bf16 *dS, int n0, int n1, int n2) {
  const int CS = C - CU;
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= (size_t)n0 * n1 * n2 * CS) return;
  const int c = (int)(t % CS);
  const int pv = pvox((int)(t / CS), n0, n1);
  dS[(size_t)pv * CS + c] = dY[(size_t)pv * C + CU + c];
}

// g = g * (act > 0), interior only (halos stay zero)
__global__ void __launch_bounds__(kUT) k_relu_bwd(bf16 *g, const bf16 *act, int C, int n0, int n1, int n2) {
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= (size_t)n0 * n1 * n2 * C) return;
  const int c = (int)(t % C);
  const size_t o = (size_t)pvox((int)(t / C), n0, n1) * C + c;
  if (!(bf2f(act[o]) > 0.f)) g[o] = 0;
}

// ----------------------------------------------------------- weights
// pack FP32 W[tap][ci][co] into MFMA B fragments (transposed/flipped for the
// data gradient when flip: W'[tap][co][ci] = W[26 - tap][ci][co])
__global__ void __launch_bounds__(kUT) k_pack(const float *W, int Cin, int Cout, int flip, s16x4 *WB) {
  const int Ci = flip ? Cout : Cin, Co = flip ? Cin : Cout;
  const int KB = Ci / 16, NB = Co / 16;
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  const size_t n = (size_t)27 * KB * NB * 64;
  if (t >= n) return;
  const int lane = (int)(t % 64);
  const int nb = (int)((t / 64) % NB);
  const int kb = (int)((t / 64 / NB) % KB);
  const int tap = (int)(t / 64 / NB / KB);
  s16x4 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int ci = kb * 16 + (lane >> 4) * 4 + j, co = nb * 16 + (lane & 15);
    // flip: ci ranges over the layer's Cout, co over its Cin
    const float w = flip ? W[((size_t)(26 - tap) * Cin + co) * Cout + ci] : W[((size_t)tap * Cin + ci) * Cout + co];
    r[j] = (short)f2bf(w);
  }
  WB[t] = r;
}

// pack FP32 W into the K = 32 B fragments of k_conv_tile: MFMA step c holds
// the 16-channel groups 2c and 2c + 1 (group g = tap Ci/16 + channel block),
// lane l the k values 8 (l / 16) ..+8; flip as in k_pack.
__global__ void __launch_bounds__(kUT) k_pack32(const float *W, int Cin, int Cout, int flip, u16x8 *WB) {
  const int Ci = flip ? Cout : Cin, Co = flip ? Cin : Cout;
  const int CG = Ci / 16, NB = Co / 16, G = 27 * CG, NC = (G + 1) / 2;
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= (size_t)NC * NB * 64) return;
  const int lane = (int)(t % 64), nb = (int)((t / 64) % NB), c = (int)(t / 64 / NB);
  u16x8 r;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = (lane >> 4) * 8 + j;
    const int g = 2 * c + (k >> 4);
    float w = 0.f;
    if (g < G) {
      const int tap = g / CG, ci = (g - tap * CG) * 16 + (k & 15), co = nb * 16 + (lane & 15);
      w = flip ? W[((size_t)(26 - tap) * Cin + co) * Cout + ci] : W[((size_t)tap * Cin + ci) * Cout + co];
    }
    r[j] = f2bf(w);
  }
  WB[t] = r;
}

// BF16-rounded FP32 copy of the small layers' weights (the values the network uses)
__global__ void __launch_bounds__(kUT) k_round_bf16(const float *W, float *Wr, size_t n) {
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t < n) Wr[t] = bf2f(f2bf(W[t]));
}

__global__ void __launch_bounds__(kUT) k_adam(float *w, float *m, float *v, const float *g, size_t n, float lr, float b1,
                                             float b2, float eps, float c1, float c2, float gscale) {
  const size_t t = (size_t)blockIdx.x * kUT + threadIdx.x;
  if (t >= n) return;
  const float gg = g[t] * gscale;
  const float mm = b1 * m[t] + (1.f - b1) * gg;
  const float vv = b2 * v[t] + (1.f - b2) * gg * gg;
  m[t] = mm;
  v[t] = vv;
  w[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// -------------------------------------------------------- loss, I/O
// loss partials of sum (out - target)^2 over interior, dout = 2 (out - target) / nvox
__global__ void __launch_bounds__(kUT) k_mse(const float *out, const float *target, float *dout, int n0, int n1, int n2,
                                            float *part) {
  __shared__ float sh[kUT];
  const int nvox = n0 * n1 * n2;
  const int v = blockIdx.x * kUT + threadIdx.x;
  float e = 0.f;
  if (v < nvox) {
    const int pv = pvox(v, n0, n1);
    const float d = out[pv] - target[pv];
    dout[pv] = 2.f * d / (float)nvox;
    e = d * d;
  }
  sh[threadIdx.x] = e;
  __syncthreads();
  for (int w = kUT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// network input from FP32 padded channels (4 arrays of (n+2)^3) -> BF16 channels-last
__global__ void __launch_bounds__(kUT) k_input(const float *in4, bf16 *X, int n0, int n1, int n2) {
  const int v = blockIdx.x * kUT + threadIdx.x;
  if (v >= n0 * n1 * n2) return;
  const int pv = pvox(v, n0, n1);
  const size_t np = (size_t)(n0 + 2) * (n1 + 2) * (n2 + 2);
#pragma unroll
  for (int c = 0; c < 4; ++c) X[(size_t)pv * 4 + c] = f2bf(in4[c * np + pv]);
}

// preconditioner input: channel 0 = (f * vol^(2/3)) / s with s the RMS of
// f * vol^(2/3) (from part), geometry channels 1..3 precomputed
__global__ void __launch_bounds__(kUT) k_prec_norm(const float *f0, const float *vol23, int n0, int n1, int n2,
                                                  float *part, const int *done) {
  if (gated(done)) return;
  __shared__ float sh[kUT];
  const int v = blockIdx.x * kUT + threadIdx.x;
  float e = 0.f;
  if (v < n0 * n1 * n2) {
    const int pv = pvox(v, n0, n1);
    const float g = f0[pv] * vol23[pv];
    e = g * g;
  }
  sh[threadIdx.x] = e;
  __syncthreads();
  for (int w = kUT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kUT) k_prec_scale(const float *part, int nb, int nvox, float *scale, const int *done) {
  if (gated(done)) return;
  __shared__ float sh[kUT];
  float s = 0.f;
  for (int q = threadIdx.x; q < nb; q += kUT) s += part[q];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kUT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float rms = sqrtf(sh[0] / (float)nvox);
    scale[0] = rms > 0.f ? rms : 1.f;       // output scale
    scale[1] = rms > 0.f ? 1.f / rms : 1.f; // input scale
  }
}

__global__ void __launch_bounds__(kUT) k_prec_input(const float *f0, const float *vol23, const float *geom,
                                                   const float *scale, bf16 *X, int n0, int n1, int n2, const int *done) {
  if (gated(done)) return;
  const int v = blockIdx.x * kUT + threadIdx.x;
  if (v >= n0 * n1 * n2) return;
  const int pv = pvox(v, n0, n1);
  const size_t np = (size_t)(n0 + 2) * (n1 + 2) * (n2 + 2);
  X[(size_t)pv * 4] = f2bf(f0[pv] * vol23[pv] * scale[1]);
#pragma unroll
  for (int c = 0; c < 3; ++c) X[(size_t)pv * 4 + 1 + c] = f2bf(geom[c * np + pv]);
}

// bias gradient of the out layer: sum of dout over the interior
__global__ void __launch_bounds__(kUT) k_sum_interior(const float *a, int n0, int n1, int n2, float *out, int accumulate) {
  __shared__ float sh[kUT];
  const int nvox = n0 * n1 * n2;
  float s = 0.f;
  for (int v = threadIdx.x; v < nvox; v += kUT) s += a[pvox(v, n0, n1)];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kUT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = accumulate ? out[0] + sh[0] : sh[0];
}

// ---- lane-layout self check of the MFMA fragments the kernels assume
__global__ void __launch_bounds__(64) k_mfma_check(const float *A, const float *B, float *D) {
  const int l = threadIdx.x;
  s16x4 a, b;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a[j] = (short)f2bf(A[(l & 15) * 16 + (l >> 4) * 4 + j]); // A[i][k]
    b[j] = (short)f2bf(B[((l >> 4) * 4 + j) * 16 + (l & 15)]); // B[k][n]
  }
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) D[((l >> 4) * 4 + j) * 16 + (l & 15)] = c[j];
}

// the same for v_mfma_f32_16x16x32_bf16: A 16 x 32, B 32 x 16, lane l holding
// A[l % 16][8 (l / 16) ..+8] and B[8 (l / 16) ..+8][l % 16]
__global__ void __launch_bounds__(64) k_mfma_check32(const float *A, const float *B, float *D) {
  const int l = threadIdx.x;
  u16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    a[j] = f2bf(A[(l & 15) * 32 + (l >> 4) * 8 + j]);
    b[j] = f2bf(B[((l >> 4) * 8 + j) * 16 + (l & 15)]);
  }
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) D[((l >> 4) * 4 + j) * 16 + (l & 15)] = c[j];
}

// ------------------------------------------------------------------ host
struct Layer {
  int Cin, Cout, lev; // level 0 = finest
  size_t w_off, b_off;
};

// conv layers in order: in, e1b, e2a, e2b, ba, bb, d2a, d2b, d1a, d1b, out
static const int kLayers[11][3] = {{4, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2}, {64, 64, 2},
                                   {96, 32, 1}, {32, 32, 1}, {48, 16, 0}, {16, 16, 0}, {16, 1, 0}};

struct Buf {
  bf16 *d = nullptr;
  int C = 0, lev = 0;
};

} // namespace moons

using namespace moons;

struct moons_unet_s {
  Layer L[11];
  size_t nparam = 0;
  float *W = nullptr, *G = nullptr, *Am = nullptr, *Av = nullptr, *Wr = nullptr; // master, grad, Adam m/v, rounded
  s16x4 *WB[11] = {}, *WT[11] = {};
  u16x8 *WB32[11] = {}, *WT32[11] = {}; // k_conv_tile fragments (finest-level layers)
  int adam_t = 0;
  // activations for one grid size
  int n[3] = {0, 0, 0};
  Buf in, a0, e1, p1, a2, e2, p2, a4, b, c2, a6, d2, c1, a8, d1;
  Buf g_d1, g_a8, g_c1, g_d2, g_a6, g_c2, g_b, g_a4, g_p2, g_e2, g_a2, g_p1, g_e1, g_a0;
  float *out = nullptr, *dout = nullptr, *part = nullptr;
  size_t part_n = 0;
  std::vector<void *> act_allocs;
};

namespace {

int lev_n(const moons_unet_s *U, int lev, int d) { return U->n[d] >> lev; }

size_t padded(const moons_unet_s *U, int lev) {
  return (size_t)(lev_n(U, lev, 0) + 2) * (lev_n(U, lev, 1) + 2) * (lev_n(U, lev, 2) + 2);
}

void free_acts(moons_unet_s *U) {
  for (void *p : U->act_allocs) (void)hipFree(p);
  U->act_allocs.clear();
  U->n[0] = U->n[1] = U->n[2] = 0;
}

int alloc_buf(moons_unet_s *U, Buf &b, int C, int lev) {
  b.C = C;
  b.lev = lev;
  const size_t bytes = padded(U, lev) * C * sizeof(bf16);
  if (hipMalloc(&b.d, bytes) != hipSuccess) return fail(MOONS_ERR_HIP, "unet: hipMalloc failed");
  U->act_allocs.push_back(b.d);
  if (hipMemsetAsync(b.d, 0, bytes, stream()) != hipSuccess) return fail(MOONS_ERR_HIP, "unet: memset failed");
  return MOONS_OK;
}

// activations and gradients for an n0 x n1 x n2 grid (each a multiple of 4)
int ensure_acts(moons_unet_s *U, int n0, int n1, int n2) {
  if (U->n[0] == n0 && U->n[1] == n1 && U->n[2] == n2) return MOONS_OK;
  if (n0 % 4 || n1 % 4 || n2 % 4 || n0 < 4 || n1 < 4 || n2 < 4)
    return fail(MOONS_ERR_ARG, "unet: grid dimensions must be multiples of 4");
  free_acts(U);
  U->n[0] = n0; U->n[1] = n1; U->n[2] = n2;
  struct { Buf *b; int C, lev; } spec[] = {
      {&U->in, 4, 0},    {&U->a0, 16, 0},   {&U->e1, 16, 0},   {&U->p1, 16, 1},   {&U->a2, 32, 1},
      {&U->e2, 32, 1},   {&U->p2, 32, 2},   {&U->a4, 64, 2},   {&U->b, 64, 2},    {&U->c2, 96, 1},
      {&U->a6, 32, 1},   {&U->d2, 32, 1},   {&U->c1, 48, 0},   {&U->a8, 16, 0},   {&U->d1, 16, 0},
      {&U->g_d1, 16, 0}, {&U->g_a8, 16, 0}, {&U->g_c1, 48, 0}, {&U->g_d2, 32, 1}, {&U->g_a6, 32, 1},
      {&U->g_c2, 96, 1}, {&U->g_b, 64, 2},  {&U->g_a4, 64, 2}, {&U->g_p2, 32, 2}, {&U->g_e2, 32, 1},
      {&U->g_a2, 32, 1}, {&U->g_p1, 16, 1}, {&U->g_e1, 16, 0}, {&U->g_a0, 16, 0}};
  for (auto &s : spec) {
    int st = alloc_buf(U, *s.b, s.C, s.lev);
    if (st) return st;
  }
  const size_t np = padded(U, 0);
  if (hipMalloc(&U->out, np * sizeof(float)) != hipSuccess || hipMalloc(&U->dout, np * sizeof(float)) != hipSuccess)
    return fail(MOONS_ERR_HIP, "unet: hipMalloc failed");
  U->act_allocs.push_back(U->out);
  U->act_allocs.push_back(U->dout);
  (void)hipMemsetAsync(U->out, 0, np * sizeof(float), stream());
  (void)hipMemsetAsync(U->dout, 0, np * sizeof(float), stream());
  // partial sums: dW chunks (at most 64 chunks x 27 x 96 x 64) and loss blocks
  U->part_n = std::max<size_t>((size_t)64 * 27 * 96 * 64, ublk((size_t)n0 * n1 * n2) + 16);
  if (hipMalloc(&U->part, U->part_n * sizeof(float)) != hipSuccess) return fail(MOONS_ERR_HIP, "unet: hipMalloc failed");
  U->act_allocs.push_back(U->part);
  return MOONS_OK;
}

// fragments of k_pack32 for a convolution of Ci input and Co output channels
size_t pack32_n(int Ci, int Co) { return (size_t)((27 * (Ci / 16) + 1) / 2) * (Co / 16) * 64; }

// the (Cin, Cout) pairs k_conv_tile is instantiated for
bool tile_conv(int Ci, int Co) { return (Ci == 16 && Co == 16) || (Ci == 48 && Co == 16) || (Ci == 16 && Co == 48); }

int pack_weights(moons_unet_s *U) {
  for (int l = 1; l <= 9; ++l) {
    const Layer &L = U->L[l];
    const size_t nf = (size_t)27 * (L.Cin / 16) * (L.Cout / 16) * 64;
    hipLaunchKernelGGL(k_pack, dim3(ublk(nf)), dim3(kUT), 0, stream(), U->W + L.w_off, L.Cin, L.Cout, 0, U->WB[l]);
    hipLaunchKernelGGL(k_pack, dim3(ublk(nf)), dim3(kUT), 0, stream(), U->W + L.w_off, L.Cin, L.Cout, 1, U->WT[l]);
    if (U->WB32[l]) {
      hipLaunchKernelGGL(k_pack32, dim3(ublk(pack32_n(L.Cin, L.Cout))), dim3(kUT), 0, stream(), U->W + L.w_off, L.Cin,
                         L.Cout, 0, U->WB32[l]);
      hipLaunchKernelGGL(k_pack32, dim3(ublk(pack32_n(L.Cout, L.Cin))), dim3(kUT), 0, stream(), U->W + L.w_off, L.Cin,
                         L.Cout, 1, U->WT32[l]);
    }
  }
  hipLaunchKernelGGL(k_round_bf16, dim3(ublk(U->nparam)), dim3(kUT), 0, stream(), U->W, U->Wr, U->nparam);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// Waves below which a convolution splits its taps (about two per SIMD; the
// 32^3 layers, 1024 waves, measured no faster split).
constexpr int kSplitWaves = 512;

template <int NB>
static void conv_split(const moons_unet_s *U, const s16x4 *WB, const float *bias, const Buf &X, Buf &Y, int n0, int n1,
                       int n2, int relu, const int *done, int S, int T) {
  const int nvox = n0 * n1 * n2;
  const dim3 grid((unsigned)((nvox + 127) / 128), (unsigned)S);
  hipLaunchKernelGGL(k_conv_mfma_split<NB>, grid, dim3(kUT), 0, stream(), X.d, X.C, WB, U->part, n0, n1, n2, T, done);
  hipLaunchKernelGGL(k_conv_reduce, dim3(ublk((size_t)nvox * Y.C)), dim3(kUT), 0, stream(), U->part, S, nvox, Y.C, bias,
                     relu, Y.d, n0, n1, done);
}

void conv_mfma(const moons_unet_s *U, const s16x4 *WB, const float *bias, const Buf &X, Buf &Y, int n0, int n1, int n2,
               int relu, const int *done, const u16x8 *WB32 = nullptr) {
  const int nvox = n0 * n1 * n2;
  if (WB32 && n0 % kCTX == 0 && n1 % kCTY == 0 && n2 % kCTZ == 0 && tile_conv(X.C, Y.C)) {
    // tiles per workgroup along z: keep at least 2048 workgroups (8 a CU)
    const int tz = n2 / kCTZ, wxy = (n0 / kCTX) * (n1 / kCTY);
    int zt = 1;
    while (tz % (2 * zt) == 0 && (size_t)wxy * (tz / (2 * zt)) >= 2048) zt *= 2;
    const dim3 grid((unsigned)(n0 / kCTX), (unsigned)(n1 / kCTY), (unsigned)(tz / zt));
    if (X.C == 16 && Y.C == 16) {
      const int timer = bias ? MOONS_KTIMER_UNET_CONV16 : 0; // forward layers carry a bias
      ktimer_mark(timer);
      hipLaunchKernelGGL((k_conv_tile<1, 1>), grid, dim3(256), 0, stream(), X.d, WB32, bias, Y.d, n0, n1, n2, relu, zt, done);
      ktimer_mark(timer);
    }
    else if (X.C == 48)
      hipLaunchKernelGGL((k_conv_tile<3, 1>), grid, dim3(256), 0, stream(), X.d, WB32, bias, Y.d, n0, n1, n2, relu, zt, done);
    else
      hipLaunchKernelGGL((k_conv_tile<1, 3>), grid, dim3(256), 0, stream(), X.d, WB32, bias, Y.d, n0, n1, n2, relu, zt, done);
    return;
  }
  const unsigned grid = (unsigned)((nvox + 127) / 128);
  const int waves = (nvox + 31) / 32;
  // a wide input (48 channels or more: 81+ MFMA steps a wave) splits up to
  // four times as many waves
  const int target = X.C >= 48 ? 16 * kSplitWaves : 4 * kSplitWaves;
  if (waves < kSplitWaves || (X.C >= 48 && waves < 4 * kSplitWaves)) {
    // groups of T taps so that the launch has about target waves, within the
    // partial-sum buffer
    int S = std::min(27, (target + waves - 1) / waves);
    while (S > 1 && (size_t)S * nvox * Y.C > U->part_n) --S;
    const int T = (27 + S - 1) / S;
    S = (27 + T - 1) / T;
    if (S > 1) {
      switch (Y.C / 16) {
        case 1: conv_split<1>(U, WB, bias, X, Y, n0, n1, n2, relu, done, S, T); break;
        case 2: conv_split<2>(U, WB, bias, X, Y, n0, n1, n2, relu, done, S, T); break;
        case 3: conv_split<3>(U, WB, bias, X, Y, n0, n1, n2, relu, done, S, T); break;
        case 4: conv_split<4>(U, WB, bias, X, Y, n0, n1, n2, relu, done, S, T); break;
        default: conv_split<6>(U, WB, bias, X, Y, n0, n1, n2, relu, done, S, T); break;
      }
      return;
    }
  }
  switch (Y.C / 16) {
    case 1: hipLaunchKernelGGL(k_conv_mfma<1>, dim3(grid), dim3(kUT), 0, stream(), X.d, X.C, WB, bias, Y.d, n0, n1, n2, relu, done); break;
    case 2: hipLaunchKernelGGL(k_conv_mfma<2>, dim3(grid), dim3(kUT), 0, stream(), X.d, X.C, WB, bias, Y.d, n0, n1, n2, relu, done); break;
    case 3: hipLaunchKernelGGL(k_conv_mfma<3>, dim3(grid), dim3(kUT), 0, stream(), X.d, X.C, WB, bias, Y.d, n0, n1, n2, relu, done); break;
    case 4: hipLaunchKernelGGL(k_conv_mfma<4>, dim3(grid), dim3(kUT), 0, stream(), X.d, X.C, WB, bias, Y.d, n0, n1, n2, relu, done); break;
    default: hipLaunchKernelGGL(k_conv_mfma<6>, dim3(grid), dim3(kUT), 0, stream(), X.d, X.C, WB, bias, Y.d, n0, n1, n2, relu, done); break;
  }
}

// forward pass from U->in (BF16, 4 channels) to out (FP32 padded, times *scale if given)
int forward(moons_unet_s *U, float *out, const float *scale, const int *done) {
  const int n0 = U->n[0], n1 = U->n[1], n2 = U->n[2];
  const float *W = U->Wr;
  auto nv = [&](int lev) { return (size_t)(n0 >> lev) * (n1 >> lev) * (n2 >> lev); };
  hipLaunchKernelGGL(k_conv_in, dim3(ublk(16 * nv(0))), dim3(kUT), 0, stream(), U->in.d, W + U->L[0].w_off,
                     U->W + U->L[0].b_off, U->a0.d, n0, n1, n2, done);
  conv_mfma(U, U->WB[1], U->W + U->L[1].b_off, U->a0, U->e1, n0, n1, n2, 1, done, U->WB32[1]);
  hipLaunchKernelGGL(k_pool, dim3(ublk(nv(1) * 16)), dim3(kUT), 0, stream(), U->e1.d, U->p1.d, 16, n0 / 2, n1 / 2, n2 / 2, done);
  conv_mfma(U, U->WB[2], U->W + U->L[2].b_off, U->p1, U->a2, n0 / 2, n1 / 2, n2 / 2, 1, done);
  conv_mfma(U, U->WB[3], U->W + U->L[3].b_off, U->a2, U->e2, n0 / 2, n1 / 2, n2 / 2, 1, done);
  hipLaunchKernelGGL(k_pool, dim3(ublk(nv(2) * 32)), dim3(kUT), 0, stream(), U->e2.d, U->p2.d, 32, n0 / 4, n1 / 4, n2 / 4, done);
  conv_mfma(U, U->WB[4], U->W + U->L[4].b_off, U->p2, U->a4, n0 / 4, n1 / 4, n2 / 4, 1, done);
  conv_mfma(U, U->WB[5], U->W + U->L[5].b_off, U->a4, U->b, n0 / 4, n1 / 4, n2 / 4, 1, done);
  hipLaunchKernelGGL(k_upcat, dim3(ublk(nv(1) * 96)), dim3(kUT), 0, stream(), U->b.d, 64, U->e2.d, 32, U->c2.d, n0 / 2, n1 / 2, n2 / 2, done);
  conv_mfma(U, U->WB[6], U->W + U->L[6].b_off, U->c2, U->a6, n0 / 2, n1 / 2, n2 / 2, 1, done);
  conv_mfma(U, U->WB[7], U->W + U->L[7].b_off, U->a6, U->d2, n0 / 2, n1 / 2, n2 / 2, 1, done);
  hipLaunchKernelGGL(k_upcat, dim3(ublk(nv(0) * 48)), dim3(kUT), 0, stream(), U->d2.d, 32, U->e1.d, 16, U->c1.d, n0, n1, n2, done);
  conv_mfma(U, U->WB[8], U->W + U->L[8].b_off, U->c1, U->a8, n0, n1, n2, 1, done, U->WB32[8]);
  conv_mfma(U, U->WB[9], U->W + U->L[9].b_off, U->a8, U->d1, n0, n1, n2, 1, done, U->WB32[9]);
  hipLaunchKernelGGL(k_conv_out, dim3(ublk(16 * nv(0))), dim3(kUT), 0, stream(), U->d1.d, W + U->L[10].w_off,
                     U->W + U->L[10].b_off, out, n0, n1, n2, scale, done);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// dW and db of an MFMA layer (accumulated into G when acc)
void grad_mfma(moons_unet_s *U, int l, const Buf &X, const Buf &dY, int n0, int n1, int n2, int acc) {
  const Layer &L = U->L[l];
  const int nvox = n0 * n1 * n2;
  int nchunk = std::max(1, std::min(64, nvox / 2048));
  const int chunk = ((nvox + nchunk - 1) / nchunk + 15) / 16 * 16;
  nchunk = (nvox + chunk - 1) / chunk;
  const dim3 grid(nchunk, 27 * (L.Cin / 16));
  switch (L.Cout / 16) {
    case 1: hipLaunchKernelGGL(k_conv_dw<1>, grid, dim3(64), 0, stream(), X.d, L.Cin, dY.d, n0, n1, n2, chunk, U->part); break;
    case 2: hipLaunchKernelGGL(k_conv_dw<2>, grid, dim3(64), 0, stream(), X.d, L.Cin, dY.d, n0, n1, n2, chunk, U->part); break;
    default: hipLaunchKernelGGL(k_conv_dw<4>, grid, dim3(64), 0, stream(), X.d, L.Cin, dY.d, n0, n1, n2, chunk, U->part); break;
  }
  const size_t nw = (size_t)27 * L.Cin * L.Cout;
  hipLaunchKernelGGL(k_reduce_chunks, dim3(ublk(nw)), dim3(kUT), 0, stream(), U->part, nchunk, nw, U->G + L.w_off, acc);
  hipLaunchKernelGGL(k_dbias, dim3(L.Cout), dim3(kUT), 0, stream(), dY.d, L.Cout, n0, n1, n2, U->G + L.b_off, acc);
}

// full backward pass from dout; parameter gradients into G (accumulated when acc)
int backward(moons_unet_s *U, int acc) {
  const int n0 = U->n[0], n1 = U->n[1], n2 = U->n[2];
  const int h0 = n0 / 2, h1 = n1 / 2, h2 = n2 / 2, q0 = n0 / 4, q1 = n1 / 4, q2 = n2 / 4;
  auto nv = [&](int lev) { return (size_t)(n0 >> lev) * (n1 >> lev) * (n2 >> lev); };
  const float *Wr = U->Wr;
  // out layer (16 -> 1)
  {
    const Layer &L = U->L[10];
    const int nvox = (int)nv(0);
    int nchunk = std::max(1, std::min(64, nvox / 4096));
    const int chunk = (nvox + nchunk - 1) / nchunk;
    nchunk = (nvox + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_small_dw<16, 1>), dim3(nchunk, 27), dim3(kUT), 0, stream(), U->d1.d, nullptr, U->dout, n0, n1,
                       n2, chunk, U->part);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(ublk(27 * 16)), dim3(kUT), 0, stream(), U->part, nchunk, (size_t)27 * 16,
                       U->G + L.w_off, acc);
    hipLaunchKernelGGL(k_conv_out_dx, dim3(ublk(nv(0))), dim3(kUT), 0, stream(), U->dout, Wr + L.w_off, U->g_d1.d, n0, n1, n2);
  }
  auto relu_bwd = [&](Buf &g, const Buf &act, int lev) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(ublk(nv(lev) * g.C)), dim3(kUT), 0, stream(), g.d, act.d, g.C, n0 >> lev,
                       n1 >> lev, n2 >> lev);
  };
  auto dgrad = [&](int l, const Buf &dY, Buf &dX, int a0, int a1, int a2) {
    // dX = conv(dY; W flipped/transposed): Cin' = Cout, Cout' = Cin
    conv_mfma(U, U->WT[l], nullptr, dY, dX, a0, a1, a2, 0, nullptr, U->WT32[l]);
  };
  relu_bwd(U->g_d1, U->d1, 0);
  grad_mfma(U, 9, U->a8, U->g_d1, n0, n1, n2, acc);
  dgrad(9, U->g_d1, U->g_a8, n0, n1, n2);
  relu_bwd(U->g_a8, U->a8, 0);
  grad_mfma(U, 8, U->c1, U->g_a8, n0, n1, n2, acc);
  dgrad(8, U->g_a8, U->g_c1, n0, n1, n2);
  hipLaunchKernelGGL(k_upcat_bwd_up, dim3(ublk(nv(1) * 32)), dim3(kUT), 0, stream(), U->g_c1.d, 48, U->g_d2.d, 32, h0, h1, h2);
  hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(ublk(nv(0) * 16)), dim3(kUT), 0, stream(), U->g_c1.d, 48, 32, U->g_e1.d, n0, n1, n2);
  relu_bwd(U->g_d2, U->d2, 1);
  grad_mfma(U, 7, U->a6, U->g_d2, h0, h1, h2, acc);
  dgrad(7, U->g_d2, U->g_a6, h0, h1, h2);
  relu_bwd(U->g_a6, U->a6, 1);
  grad_mfma(U, 6, U->c2, U->g_a6, h0, h1, h2, acc);
  dgrad(6, U->g_a6, U->g_c2, h0, h1, h2);
  hipLaunchKernelGGL(k_upcat_bwd_up, dim3(ublk(nv(2) * 64)), dim3(kUT), 0, stream(), U->g_c2.d, 96, U->g_b.d, 64, q0, q1, q2);
  hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(ublk(nv(1) * 32)), dim3(kUT), 0, stream(), U->g_c2.d, 96, 64, U->g_e2.d, h0, h1, h2);
  relu_bwd(U->g_b, U->b, 2);
  grad_mfma(U, 5, U->a4, U->g_b, q0, q1, q2, acc);
  dgrad(5, U->g_b, U->g_a4, q0, q1, q2);
  relu_bwd(U->g_a4, U->a4, 2);
  grad_mfma(U, 4, U->p2, U->g_a4, q0, q1, q2, acc);
  dgrad(4, U->g_a4, U->g_p2, q0, q1, q2);
  hipLaunchKernelGGL(k_pool_bwd, dim3(ublk(nv(1) * 32)), dim3(kUT), 0, stream(), U->g_p2.d, U->g_e2.d, 32, h0, h1, h2, 1);
  relu_bwd(U->g_e2, U->e2, 1);
  grad_mfma(U, 3, U->a2, U->g_e2, h0, h1, h2, acc);
  dgrad(3, U->g_e2, U->g_a2, h0, h1, h2);
  relu_bwd(U->g_a2, U->a2, 1);
  grad_mfma(U, 2, U->p1, U->g_a2, h0, h1, h2, acc);
  dgrad(2, U->g_a2, U->g_p1, h0, h1, h2);
  hipLaunchKernelGGL(k_pool_bwd, dim3(ublk(nv(0) * 16)), dim3(kUT), 0, stream(), U->g_p1.d, U->g_e1.d, 16, n0, n1, n2, 1);
  relu_bwd(U->g_e1, U->e1, 0);
  grad_mfma(U, 1, U->a0, U->g_e1, n0, n1, n2, acc);
  dgrad(1, U->g_e1, U->g_a0, n0, n1, n2);
  relu_bwd(U->g_a0, U->a0, 0);
  {
    const Layer &L = U->L[0];
    const int nvox = (int)nv(0);
    int nchunk = std::max(1, std::min(64, nvox / 4096));
    const int chunk = (nvox + nchunk - 1) / nchunk;
    nchunk = (nvox + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_small_dw<4, 16>), dim3(nchunk, 27), dim3(kUT), 0, stream(), U->in.d, U->g_a0.d, nullptr, n0,
                       n1, n2, chunk, U->part);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(ublk(27 * 64)), dim3(kUT), 0, stream(), U->part, nchunk, (size_t)27 * 64,
                       U->G + L.w_off, acc);
    hipLaunchKernelGGL(k_dbias, dim3(16), dim3(kUT), 0, stream(), U->g_a0.d, 16, n0, n1, n2, U->G + L.b_off, acc);
  }
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int backward_all(moons_unet_s *U, int acc) {
  int st = backward(U, acc);
  if (st) return st;
  hipLaunchKernelGGL(k_sum_interior, dim3(1), dim3(kUT), 0, stream(), U->dout, U->n[0], U->n[1], U->n[2],
                     U->G + U->L[10].b_off, acc);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// loss of the current out against target (both FP32 padded), dout set
int loss(moons_unet_s *U, const float *target, double *l) {
  const int n0 = U->n[0], n1 = U->n[1], n2 = U->n[2];
  const int nvox = n0 * n1 * n2;
  const unsigned nb = ublk(nvox);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(kUT), 0, stream(), U->out, target, U->dout, n0, n1, n2, U->part);
  MOONS_LAUNCH_CHECK();
  if (l) {
    std::vector<float> h(nb);
    MOONS_HIP(hipMemcpyAsync(h.data(), U->part, nb * sizeof(float), hipMemcpyDeviceToHost, stream()));
    MOONS_HIP(hipStreamSynchronize(stream()));
    double s = 0.0;
    for (float v : h) s += v;
    *l = s / nvox;
  }
  return MOONS_OK;
}

} // namespace

// ------------------------------------------------ preconditioner interface
namespace moons {

struct UnetPrec {
  moons_unet_s *net = nullptr;
  float *vol23 = nullptr, *geom = nullptr, *scale = nullptr, *part = nullptr;
  int n[3];
};

// z0 (FP32 padded) = s * N(f0 * vol^(2/3) / s), every kernel gated on *done
int unet_prec_apply(UnetPrec *P, const float *f0, float *z0, const int *done) {
  moons_unet_s *U = P->net;
  int st = ensure_acts(U, P->n[0], P->n[1], P->n[2]);
  if (st) return st;
  const int n0 = P->n[0], n1 = P->n[1], n2 = P->n[2];
  const int nvox = n0 * n1 * n2;
  const unsigned nb = ublk(nvox);
  hipLaunchKernelGGL(k_prec_norm, dim3(nb), dim3(kUT), 0, stream(), f0, P->vol23, n0, n1, n2, P->part, done);
  hipLaunchKernelGGL(k_prec_scale, dim3(1), dim3(kUT), 0, stream(), P->part, (int)nb, nvox, P->scale, done);
  hipLaunchKernelGGL(k_prec_input, dim3(nb), dim3(kUT), 0, stream(), f0, P->vol23, P->geom, P->scale, U->in.d, n0, n1,
                     n2, done);
  return forward(U, z0, P->scale, done);
}

// Geometry of a cell-centred field's mesh for the network input:
// vol^(2/3) and log(dh_d / vol^(1/3)), d = x, y, z, per padded cell.
UnetPrec *unet_prec_create_nodes(moons_unet_s *net, const std::vector<double> hn[3], std::string *err) {
  UnetPrec *P = new UnetPrec();
  P->net = net;
  int s[3];
  std::vector<double> dh[3];
  for (int d = 0; d < 3; ++d) {
    s[d] = (int)hn[d].size() - 1; // padded cells
    P->n[d] = s[d] - 2;
    dh[d].resize(s[d]);
    for (int i = 0; i < s[d]; ++i) dh[d][i] = hn[d][i + 1] - hn[d][i];
  }
  if (P->n[0] % 4 || P->n[1] % 4 || P->n[2] % 4 || P->n[0] < 4 || P->n[1] < 4 || P->n[2] < 4) {
    if (err) *err = "the learned preconditioner needs cell counts that are multiples of 4";
    delete P;
    return nullptr;
  }
  const size_t np = (size_t)s[0] * s[1] * s[2];
  std::vector<float> v23(np), geom(3 * np);
  for (int k = 0; k < s[2]; ++k)
    for (int j = 0; j < s[1]; ++j)
      for (int i = 0; i < s[0]; ++i) {
        const size_t o = (size_t)i + (size_t)s[0] * ((size_t)j + (size_t)s[1] * k);
        const double w[3] = {dh[0][i], dh[1][j], dh[2][k]};
        const double vol = w[0] * w[1] * w[2];
        const double h = std::cbrt(vol);
        v23[o] = (float)(h * h);
        for (int d = 0; d < 3; ++d) geom[d * np + o] = (float)std::log(w[d] / h);
      }
  bool ok = hipMalloc(&P->vol23, np * 4) == hipSuccess && hipMalloc(&P->geom, 3 * np * 4) == hipSuccess &&
            hipMalloc(&P->scale, 2 * 4) == hipSuccess &&
            hipMalloc(&P->part, (ublk((size_t)P->n[0] * P->n[1] * P->n[2]) + 1) * 4) == hipSuccess;
  if (!ok) {
    if (err) *err = "hipMalloc failed";
    unet_prec_destroy(P);
    return nullptr;
  }
  (void)hipMemcpyAsync(P->vol23, v23.data(), np * 4, hipMemcpyHostToDevice, stream());
  (void)hipMemcpyAsync(P->geom, geom.data(), 3 * np * 4, hipMemcpyHostToDevice, stream());
  (void)hipStreamSynchronize(stream());
  return P;
}

UnetPrec *unet_prec_create(moons_unet_s *net, const moons_field_s *x, std::string *err) {
  const std::vector<double> hn[3] = {x->m->hc[0].hn, x->m->hc[1].hn, x->m->hc[2].hn};
  return unet_prec_create_nodes(net, hn, err);
}

void unet_prec_destroy(UnetPrec *P) {
  if (!P) return;
  for (float *q : {P->vol23, P->geom, P->scale, P->part})
    if (q) (void)hipFree(q);
  delete P;
}

} // namespace moons

extern "C" {

int moons_unet_create(unsigned long long seed, moons_unet_t *out) {
  if (!stream()) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!out) return fail(MOONS_ERR_ARG, "moons_unet_create: null argument");
  moons_unet_s *U = new moons_unet_s();
  size_t off = 0;
  for (int l = 0; l < 11; ++l) {
    U->L[l].Cin = kLayers[l][0];
    U->L[l].Cout = kLayers[l][1];
    U->L[l].lev = kLayers[l][2];
    U->L[l].w_off = off;
    off += (size_t)27 * U->L[l].Cin * U->L[l].Cout;
    U->L[l].b_off = off;
    off += U->L[l].Cout;
  }
  U->nparam = off;
  // He initialisation (normal, std sqrt(2 / fan_in)), biases zero
  std::vector<float> w(off, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < 11; ++l) {
    const float sd = std::sqrt(2.f / (27.f * U->L[l].Cin)) * (l == 10 ? 0.5f : 1.f);
    std::normal_distribution<float> nd(0.f, sd);
    for (size_t q = 0; q < (size_t)27 * U->L[l].Cin * U->L[l].Cout; ++q) w[U->L[l].w_off + q] = nd(rng);
  }
  bool ok = hipMalloc(&U->W, off * 4) == hipSuccess && hipMalloc(&U->G, off * 4) == hipSuccess &&
            hipMalloc(&U->Am, off * 4) == hipSuccess && hipMalloc(&U->Av, off * 4) == hipSuccess &&
            hipMalloc(&U->Wr, off * 4) == hipSuccess;
  for (int l = 1; l <= 9 && ok; ++l) {
    const size_t nf = (size_t)27 * (U->L[l].Cin / 16) * (U->L[l].Cout / 16) * 64;
    ok = hipMalloc(&U->WB[l], nf * sizeof(s16x4)) == hipSuccess && hipMalloc(&U->WT[l], nf * sizeof(s16x4)) == hipSuccess;
    if (ok && U->L[l].lev == 0 && tile_conv(U->L[l].Cin, U->L[l].Cout))
      ok = hipMalloc(&U->WB32[l], pack32_n(U->L[l].Cin, U->L[l].Cout) * sizeof(u16x8)) == hipSuccess &&
           hipMalloc(&U->WT32[l], pack32_n(U->L[l].Cout, U->L[l].Cin) * sizeof(u16x8)) == hipSuccess;
  }
  if (!ok) {
    moons_unet_destroy(U);
    return fail(MOONS_ERR_HIP, "moons_unet_create: allocation failed");
  }
  MOONS_HIP(hipMemcpyAsync(U->W, w.data(), off * 4, hipMemcpyHostToDevice, stream()));
  MOONS_HIP(hipMemsetAsync(U->G, 0, off * 4, stream()));
  MOONS_HIP(hipMemsetAsync(U->Am, 0, off * 4, stream()));
  MOONS_HIP(hipMemsetAsync(U->Av, 0, off * 4, stream()));
  int st = pack_weights(U);
  if (st) {
    moons_unet_destroy(U);
    return st;
  }
  MOONS_HIP(hipStreamSynchronize(stream()));
  *out = U;
  return MOONS_OK;
}

int moons_unet_destroy(moons_unet_t U) {
  if (!U) return MOONS_OK;
  if (stream()) (void)hipStreamSynchronize(stream());
  free_acts(U);
  for (float *p : {U->W, U->G, U->Am, U->Av, U->Wr})
    if (p) (void)hipFree(p);
  for (int l = 0; l < 11; ++l) {
    if (U->WB[l]) (void)hipFree(U->WB[l]);
    if (U->WT[l]) (void)hipFree(U->WT[l]);
    if (U->WB32[l]) (void)hipFree(U->WB32[l]);
    if (U->WT32[l]) (void)hipFree(U->WT32[l]);
  }
  delete U;
  return MOONS_OK;
}

size_t moons_unet_num_params(moons_unet_t U) { return U ? U->nparam : 0; }

int moons_unet_get_weights(moons_unet_t U, float *host, size_t n) {
  if (!U || !host || n != U->nparam) return fail(MOONS_ERR_ARG, "moons_unet_get_weights: bad argument");
  MOONS_HIP(hipMemcpyAsync(host, U->W, n * 4, hipMemcpyDeviceToHost, stream()));
  MOONS_HIP(hipStreamSynchronize(stream()));
  return MOONS_OK;
}

int moons_unet_set_weights(moons_unet_t U, const float *host, size_t n) {
  if (!U || !host || n != U->nparam) return fail(MOONS_ERR_ARG, "moons_unet_set_weights: bad argument");
  MOONS_HIP(hipMemcpyAsync(U->W, host, n * 4, hipMemcpyHostToDevice, stream()));
  int st = pack_weights(U);
  if (st) return st;
  MOONS_HIP(hipStreamSynchronize(stream()));
  return MOONS_OK;
}

// in4: 4 FP32 arrays of (n0+2)(n1+2)(n2+2) (halo ignored); out: (n0+2)(n1+2)(n2+2) FP32
int moons_unet_forward_host(moons_unet_t U, int n0, int n1, int n2, const float *in4, float *out) {
  if (!U || !in4 || !out) return fail(MOONS_ERR_ARG, "moons_unet_forward_host: null argument");
  int st = ensure_acts(U, n0, n1, n2);
  if (st) return st;
  const size_t np = padded(U, 0);
  float *d = nullptr;
  MOONS_HIP(hipMalloc(&d, 4 * np * sizeof(float)));
  MOONS_HIP(hipMemcpyAsync(d, in4, 4 * np * sizeof(float), hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(k_input, dim3(ublk((size_t)n0 * n1 * n2)), dim3(kUT), 0, stream(), d, U->in.d, n0, n1, n2);
  st = forward(U, U->out, nullptr, nullptr);
  if (!st) {
    MOONS_HIP(hipMemcpyAsync(out, U->out, np * sizeof(float), hipMemcpyDeviceToHost, stream()));
    MOONS_HIP(hipStreamSynchronize(stream()));
  }
  (void)hipFree(d);
  return st;
}

// loss = mean (N(in) - target)^2 over the interior and its gradient w.r.t. every parameter
int moons_unet_grad_host(moons_unet_t U, int n0, int n1, int n2, const float *in4, const float *target, double *l,
                         float *grad) {
  if (!U || !in4 || !target || !grad) return fail(MOONS_ERR_ARG, "moons_unet_grad_host: null argument");
  int st = ensure_acts(U, n0, n1, n2);
  if (st) return st;
  const size_t np = padded(U, 0);
  float *d = nullptr, *t = nullptr;
  MOONS_HIP(hipMalloc(&d, 4 * np * sizeof(float)));
  MOONS_HIP(hipMalloc(&t, np * sizeof(float)));
  MOONS_HIP(hipMemcpyAsync(d, in4, 4 * np * sizeof(float), hipMemcpyHostToDevice, stream()));
  MOONS_HIP(hipMemcpyAsync(t, target, np * sizeof(float), hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(k_input, dim3(ublk((size_t)n0 * n1 * n2)), dim3(kUT), 0, stream(), d, U->in.d, n0, n1, n2);
  if (!(st = forward(U, U->out, nullptr, nullptr)) && !(st = loss(U, t, l)) && !(st = backward_all(U, 0))) {
    MOONS_HIP(hipMemcpyAsync(grad, U->G, U->nparam * 4, hipMemcpyDeviceToHost, stream()));
    MOONS_HIP(hipStreamSynchronize(stream()));
  }
  (void)hipFree(d);
  (void)hipFree(t);
  return st;
}

// Adam training on a data set of npair samples of one grid size:
// in4: npair x 4 x padded FP32 (input channels), target: npair x padded FP32.
// Each epoch visits the samples in a seeded random order, batch samples per
// Adam step; loss_hist[e] = mean training loss of epoch e.
int moons_unet_train(moons_unet_t U, int n0, int n1, int n2, int npair, const float *in4, const float *target,
                     int epochs, int batch, double lr, unsigned long long seed, double *loss_hist) {
  if (!U || !in4 || !target || npair < 1 || epochs < 1 || batch < 1)
    return fail(MOONS_ERR_ARG, "moons_unet_train: bad argument");
  int st = ensure_acts(U, n0, n1, n2);
  if (st) return st;
  const size_t np = padded(U, 0);
  float *din = nullptr, *dt = nullptr;
  if (hipMalloc(&din, (size_t)npair * 4 * np * sizeof(float)) != hipSuccess ||
      hipMalloc(&dt, (size_t)npair * np * sizeof(float)) != hipSuccess) {
    if (din) (void)hipFree(din);
    return fail(MOONS_ERR_HIP, "moons_unet_train: data set allocation failed");
  }
  MOONS_HIP(hipMemcpyAsync(din, in4, (size_t)npair * 4 * np * sizeof(float), hipMemcpyHostToDevice, stream()));
  MOONS_HIP(hipMemcpyAsync(dt, target, (size_t)npair * np * sizeof(float), hipMemcpyHostToDevice, stream()));
  std::mt19937_64 rng(seed);
  std::vector<int> order(npair);
  for (int q = 0; q < npair; ++q) order[q] = q;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs && !st; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    double esum = 0.0;
    for (int q0 = 0; q0 < npair && !st; q0 += batch) {
      const int qn = std::min(batch, npair - q0);
      for (int q = 0; q < qn && !st; ++q) {
        const int s = order[q0 + q];
        hipLaunchKernelGGL(k_input, dim3(ublk((size_t)n0 * n1 * n2)), dim3(kUT), 0, stream(), din + (size_t)s * 4 * np,
                           U->in.d, n0, n1, n2);
        double l = 0.0;
        if (!(st = forward(U, U->out, nullptr, nullptr)) && !(st = loss(U, dt + (size_t)s * np, &l)))
          st = backward_all(U, q > 0 ? 1 : 0);
        esum += l;
      }
      if (st) break;
      ++U->adam_t;
      const float c1 = 1.f - std::pow(b1, (float)U->adam_t), c2 = 1.f - std::pow(b2, (float)U->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(ublk(U->nparam)), dim3(kUT), 0, stream(), U->W, U->Am, U->Av, U->G, U->nparam,
                         (float)lr, b1, b2, eps, c1, c2, 1.f / (float)qn);
      MOONS_LAUNCH_CHECK();
      st = pack_weights(U);
    }
    if (loss_hist) loss_hist[e] = esum / npair;
  }
  (void)hipStreamSynchronize(stream());
  (void)hipFree(din);
  (void)hipFree(dt);
  return st;
}

int moons_mfma_check(const float *A, const float *B, float *D) {
  if (!A || !B || !D) return fail(MOONS_ERR_ARG, "moons_mfma_check: null argument");
  float *d = nullptr;
  MOONS_HIP(hipMalloc(&d, 3 * 256 * sizeof(float)));
  MOONS_HIP(hipMemcpyAsync(d, A, 256 * 4, hipMemcpyHostToDevice, stream()));
  MOONS_HIP(hipMemcpyAsync(d + 256, B, 256 * 4, hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(k_mfma_check, dim3(1), dim3(64), 0, stream(), d, d + 256, d + 512);
  MOONS_HIP(hipMemcpyAsync(D, d + 512, 256 * 4, hipMemcpyDeviceToHost, stream()));
  MOONS_HIP(hipStreamSynchronize(stream()));
  (void)hipFree(d);
  return MOONS_OK;
}

int moons_mfma_check32(const float *A, const float *B, float *D) {
  if (!A || !B || !D) return fail(MOONS_ERR_ARG, "moons_mfma_check32: null argument");
  float *d = nullptr;
  MOONS_HIP(hipMalloc(&d, (512 + 512 + 256) * sizeof(float)));
  MOONS_HIP(hipMemcpyAsync(d, A, 512 * 4, hipMemcpyHostToDevice, stream()));
  MOONS_HIP(hipMemcpyAsync(d + 512, B, 512 * 4, hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(k_mfma_check32, dim3(1), dim3(64), 0, stream(), d, d + 512, d + 1024);
  MOONS_HIP(hipMemcpyAsync(D, d + 1024, 256 * 4, hipMemcpyDeviceToHost, stream()));
  MOONS_HIP(hipStreamSynchronize(stream()));
  (void)hipFree(d);
  return MOONS_OK;
}

} // extern "C"
