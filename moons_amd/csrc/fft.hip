// fft.hip -- direct solver of the cell-centred pressure Poisson problem by
// cosine transforms: the reference's FFT solver (init_FFT,
// code/handwritten/FFT_solver/FFT_solver_extend.f90:43-126; solve_FFT_SF_x/y/z,
// code/handwritten/FFT_solver/FFT_algorithm.f90:27-210).
//
// Each plane normal to `dir` is an independent 2D problem in the two
// transverse directions a < b.  The interior of f is transformed by the
// staggered-grid DCT-II along a and b (code/handwritten/ops/ops_dct.f90:36-70:
// Y(k) = sum_j X(j) cos(pi (k-1) (j-1/2) / N)), every mode (p, q) is divided
// by the eigenvalue of the 5-point Neumann Laplacian on the uniform transverse
// grid, lambda = (2 cos(pi p/Na) - 2)/ha^2 + (2 cos(pi q/Nb) - 2)/hb^2, the
// (0, 0) mode is set to zero ("pin down pressure"), the inverse transform
// (ops_idct.f90: x(j) = Y(1)/N + 2/N sum_{k>=2} Y(k) cos(...)) gives u's
// interior, and u's BCs set its ghosts.
//
// On the GPU each transform along one direction is one FP64 GEMM with the
// N x N cosine matrix (rocBLAS, on the matrix cores): along x the matrix
// multiplies the n0 x (n1 n2) view from the left, along y a strided batch of
// n0 x n1 planes from the right, along z the (n0 n1) x n2 view from the right.
// The reference computes the same transforms with a radix-2 FFT, so the two
// agree to rounding.
//
// Mode factors.  By default they are the reference's (init_FFT,
// FFT_solver_extend.f90:62-112): 0.5 ha hb / (cos(pi p/Na) + cos(pi q/Nb) - 2)
// with ha, hb the first spacings of the transverse directions, which is
// 1/lambda for equal spacings and 1/lambda of the 5-point Laplacian with both
// spacings sqrt(ha hb) otherwise; for dir = 1 the reference guards the zero
// mode with the indices (i, j) instead of (j, k), so the modes (0, q > 0) of
// the first x-plane get the factor 1, which is reproduced.  Where the
// reference's guard leaves 0.5 ha hb / 0 (the (0, 0) mode of every other
// x-plane) this solver stores 0; the mode is set to 0 right after ("pin down
// pressure"), so the result is the same.  MOONS_FFT_COEF_EXACT selects
// 1/lambda for any spacings and no dir = 1 special case.
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "common.hpp"

struct moons_fft_s {
  moons_mesh_s *m = nullptr;
  int dir = 0;                   // 0, 1, 2
  int n[3] = {0, 0, 0};          // interior cells
  int a = 1, b = 2;              // transverse directions
  double *W = nullptr, *V = nullptr; // compact n0 n1 n2 work arrays
  double *C[3] = {nullptr, nullptr, nullptr};  // DCT-II matrices of the transverse directions
  double *Ci[3] = {nullptr, nullptr, nullptr}; // and their inverses
  double *coef = nullptr;        // factor per mode (p, q), p fastest; 0 at (0, 0)
  int kind = MOONS_FFT_COEF_REFERENCE;
  rocblas_handle blas = nullptr;
};

namespace moons {
namespace {

constexpr int kFT = 256;

// compact interior <- padded field, and back
__global__ void __launch_bounds__(kFT) k_fft_pack(const double *__restrict__ f, double *__restrict__ w, int n0, int n1,
                                                  int n2) {
  const long long t = (long long)blockIdx.x * kFT + threadIdx.x;
  const long long nc = (long long)n0 * n1 * n2;
  if (t >= nc) return;
  const int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((long long)n0 * n1));
  w[t] = f[(size_t)(i + 1) + (size_t)(n0 + 2) * ((size_t)(j + 1) + (size_t)(n1 + 2) * (size_t)(k + 1))];
}

__global__ void __launch_bounds__(kFT) k_fft_unpack(const double *__restrict__ w, double *__restrict__ u, int n0,
                                                    int n1, int n2) {
  const long long t = (long long)blockIdx.x * kFT + threadIdx.x;
  const long long nc = (long long)n0 * n1 * n2;
  if (t >= nc) return;
  const int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((long long)n0 * n1));
  u[(size_t)(i + 1) + (size_t)(n0 + 2) * ((size_t)(j + 1) + (size_t)(n1 + 2) * (size_t)(k + 1))] = w[t];
}

// y *= the factor of the cell's transverse mode (p along a, q along b);
// x_quirk: the reference's dir = 1 guard (modes (0, q > 0) of the first
// x-plane keep their value)
__global__ void __launch_bounds__(kFT) k_fft_scale(double *__restrict__ y, const double *__restrict__ coef, int n0,
                                                   int n1, int n2, int a, int b, int x_quirk) {
  const long long t = (long long)blockIdx.x * kFT + threadIdx.x;
  const long long nc = (long long)n0 * n1 * n2;
  if (t >= nc) return;
  const int idx[3] = {(int)(t % n0), (int)((t / n0) % n1), (int)(t / ((long long)n0 * n1))};
  const int na = a == 0 ? n0 : (a == 1 ? n1 : n2);
  const double c = (x_quirk && idx[0] == 0 && idx[a] == 0 && idx[b] > 0) ? 1.0 : coef[idx[a] + na * idx[b]];
  y[t] = y[t] * c;
}

unsigned fblocks(long long n) { return (unsigned)((n + kFT - 1) / kFT); }

const char *blas_msg(rocblas_status s) { return rocblas_status_to_string(s); }

// Y = transform of X along direction t with the n_t x n_t matrix M
// (row = output index, column = input index, column-major).
int transform(moons_fft_s *F, int t, const double *M, const double *X, double *Y) {
  const int n0 = F->n[0], n1 = F->n[1], n2 = F->n[2];
  const double one = 1.0, zero = 0.0;
  rocblas_status s;
  if (t == 0) {
    s = rocblas_dgemm(F->blas, rocblas_operation_none, rocblas_operation_none, n0, n1 * n2, n0, &one, M, n0, X, n0,
                      &zero, Y, n0);
  } else if (t == 1) {
    s = rocblas_dgemm_strided_batched(F->blas, rocblas_operation_none, rocblas_operation_transpose, n0, n1, n1, &one,
                                      X, n0, (rocblas_stride)n0 * n1, M, n1, 0, &zero, Y, n0,
                                      (rocblas_stride)n0 * n1, n2);
  } else {
    s = rocblas_dgemm(F->blas, rocblas_operation_none, rocblas_operation_transpose, n0 * n1, n2, n2, &one, X, n0 * n1,
                      M, n2, &zero, Y, n0 * n1);
  }
  if (s != rocblas_status_success) return fail(MOONS_ERR_HIP, std::string("DCT transform (rocBLAS dgemm): ") + blas_msg(s));
  return MOONS_OK;
}

bool uniform(const std::vector<double> &dhn, int n) {
  // interior cell widths dhn[1..n] (one ghost cell on each side)
  const double h = dhn[1];
  for (int i = 1; i <= n; ++i)
    if (std::fabs(dhn[i] - h) > 1e-12 * std::fabs(h)) return false;
  return true;
}

// The mode factors of kind (MOONS_FFT_COEF_*) uploaded to F->coef.
int fft_set_coefficients(moons_fft_s *F, int kind) {
  const double pi = 3.14159265358979323846;
  const int Na = F->n[F->a], Nb = F->n[F->b];
  const HostCoords &ca = F->m->hc[F->a], &cb = F->m->hc[F->b];
  std::vector<double> coef((size_t)Na * Nb);
  for (int q = 0; q < Nb; ++q)
    for (int p = 0; p < Na; ++p) {
      double c;
      if (kind == MOONS_FFT_COEF_EXACT) {
        const double ha = ca.dhn[1], hb = cb.dhn[1];
        const double lam = (2.0 * std::cos(pi * p / Na) - 2.0) / (ha * ha) + (2.0 * std::cos(pi * q / Nb) - 2.0) / (hb * hb);
        c = 1.0 / lam;
      } else {
        // init_FFT (FFT_solver_extend.f90:62-112), the first spacing dhn(1) of each transverse direction
        const double ha = ca.dhn[0], hb = cb.dhn[0];
        const double cosa = std::cos(pi * (double)p / (double)Na), cosb = std::cos(pi * (double)q / (double)Nb);
        c = 0.5 * ha * hb / (cosa + cosb - 2.0);
      }
      coef[(size_t)p + (size_t)Na * q] = (p == 0 && q == 0) ? 0.0 : c;
    }
  MOONS_HIP(hipMemcpy(F->coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
  F->kind = kind;
  return MOONS_OK;
}

} // namespace
} // namespace moons

using namespace moons;

extern "C" {

int moons_fft_create(moons_mesh_t m, int dir, moons_fft_t *out) {
  if (!m || !out) return fail(MOONS_ERR_ARG, "moons_fft_create: null argument");
  *out = nullptr;
  if (dir < 1 || dir > 3) return fail(MOONS_ERR_ARG, "Error: dir must = 1,2,3 in solve_FFT_SF in FFT_algorithm.f90");
  moons_fft_s *F = new moons_fft_s();
  F->m = m;
  F->dir = dir - 1;
  for (int d = 0; d < 3; ++d) F->n[d] = m->hc[d].sc - 2;
  F->a = F->dir == 0 ? 1 : 0;
  F->b = F->dir == 2 ? 1 : 2;
  auto bail = [&](int code, const std::string &msg) {
    moons_fft_destroy(F);
    return fail(code, msg);
  };
  for (int t : {F->a, F->b})
    if (!uniform(m->hc[t].dhn, F->n[t]))
      return bail(MOONS_ERR_ARG, "moons_fft_create: the cosine-transform solver needs uniform spacing in the two "
                                 "directions transverse to dir");
  const size_t nc = (size_t)F->n[0] * F->n[1] * F->n[2];
  if (hipMalloc(&F->W, nc * sizeof(double)) != hipSuccess || hipMalloc(&F->V, nc * sizeof(double)) != hipSuccess)
    return bail(MOONS_ERR_HIP, "moons_fft_create: hipMalloc failed");
  const double pi = 3.14159265358979323846;
  for (int t : {F->a, F->b}) {
    const int N = F->n[t];
    std::vector<double> C((size_t)N * N), Ci((size_t)N * N);
    for (int p = 0; p < N; ++p)
      for (int j = 0; j < N; ++j) {
        const double c = std::cos(pi * p * (j + 0.5) / N);
        C[(size_t)p + (size_t)N * j] = c;                              // row p (mode), column j (cell)
        Ci[(size_t)j + (size_t)N * p] = (p == 0 ? 1.0 : 2.0) / N * c; // row j (cell), column p (mode)
      }
    if (hipMalloc(&F->C[t], C.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&F->Ci[t], Ci.size() * sizeof(double)) != hipSuccess)
      return bail(MOONS_ERR_HIP, "moons_fft_create: hipMalloc failed");
    if (hipMemcpy(F->C[t], C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(F->Ci[t], Ci.data(), Ci.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return bail(MOONS_ERR_HIP, "moons_fft_create: hipMemcpy failed");
  }
  if (hipMalloc(&F->coef, (size_t)F->n[F->a] * F->n[F->b] * sizeof(double)) != hipSuccess)
    return bail(MOONS_ERR_HIP, "moons_fft_create: allocation failed");
  if (int st = fft_set_coefficients(F, MOONS_FFT_COEF_REFERENCE)) {
    moons_fft_destroy(F);
    return st;
  }
  if (rocblas_create_handle(&F->blas) != rocblas_status_success) {
    F->blas = nullptr;
    return bail(MOONS_ERR_HIP, "moons_fft_create: rocblas_create_handle failed");
  }
  *out = F;
  return MOONS_OK;
}

int moons_fft_set_coefficients(moons_fft_t F, int kind) {
  if (!F) return fail(MOONS_ERR_ARG, "moons_fft_set_coefficients: null solver");
  if (kind != MOONS_FFT_COEF_REFERENCE && kind != MOONS_FFT_COEF_EXACT)
    return fail(MOONS_ERR_ARG, "moons_fft_set_coefficients: kind must be MOONS_FFT_COEF_REFERENCE or _EXACT");
  if (stream()) MOONS_HIP(hipStreamSynchronize(stream()));
  return fft_set_coefficients(F, kind);
}

int moons_fft_destroy(moons_fft_t F) {
  if (!F) return MOONS_OK;
  if (stream()) (void)hipStreamSynchronize(stream());
  if (F->blas) (void)rocblas_destroy_handle(F->blas);
  for (double *p : {F->W, F->V, F->coef, F->C[0], F->C[1], F->C[2], F->Ci[0], F->Ci[1], F->Ci[2]})
    if (p) (void)hipFree(p);
  delete F;
  return MOONS_OK;
}

int moons_fft_solve(moons_fft_t F, moons_field_t u, moons_field_t f) {
  if (!F || !u || !f) return fail(MOONS_ERR_ARG, "moons_fft_solve: null argument");
  if (u->m != F->m || f->m != F->m || u->dl != MOONS_DL_CC || f->dl != MOONS_DL_CC)
    return fail(MOONS_ERR_ARG, "moons_fft_solve: u and f must be cell-centred fields of the solver's mesh");
  if (rocblas_set_stream(F->blas, stream()) != rocblas_status_success)
    return fail(MOONS_ERR_HIP, "moons_fft_solve: rocblas_set_stream failed");
  const int n0 = F->n[0], n1 = F->n[1], n2 = F->n[2];
  const long long nc = (long long)n0 * n1 * n2;
  hipLaunchKernelGGL(k_fft_pack, dim3(fblocks(nc)), dim3(kFT), 0, stream(), f->d, F->W, n0, n1, n2);
  MOONS_LAUNCH_CHECK();
  int st;
  // forward: along a (W -> V), along b (V -> W)
  if ((st = transform(F, F->a, F->C[F->a], F->W, F->V))) return st;
  if ((st = transform(F, F->b, F->C[F->b], F->V, F->W))) return st;
  const int x_quirk = F->kind == MOONS_FFT_COEF_REFERENCE && F->dir == 0;
  hipLaunchKernelGGL(k_fft_scale, dim3(fblocks(nc)), dim3(kFT), 0, stream(), F->W, F->coef, n0, n1, n2, F->a, F->b,
                     x_quirk);
  MOONS_LAUNCH_CHECK();
  // inverse: along a (W -> V), along b (V -> W)
  if ((st = transform(F, F->a, F->Ci[F->a], F->W, F->V))) return st;
  if ((st = transform(F, F->b, F->Ci[F->b], F->V, F->W))) return st;
  hipLaunchKernelGGL(k_fft_unpack, dim3(fblocks(nc)), dim3(kFT), 0, stream(), F->W, u->d, n0, n1, n2);
  MOONS_LAUNCH_CHECK();
  return launch_apply_bcs(u, false);
}

int moons_clean_div_fft(moons_fft_t F, moons_field_t X[3], moons_field_t Xstar[3], moons_field_t phi, double scale,
                        moons_field_t temp_F[3], moons_field_t temp_CC) {
  if (!F || !X || !Xstar || !phi || !temp_F || !temp_CC) return fail(MOONS_ERR_ARG, "moons_clean_div_fft: null argument");
  for (int d = 0; d < 3; ++d)
    if (!X[d] || !Xstar[d] || !temp_F[d]) return fail(MOONS_ERR_ARG, "moons_clean_div_fft: null component");
  int st;
  if ((st = moons_div(temp_CC, Xstar[0], Xstar[1], Xstar[2]))) return st;
  if ((st = launch_scale(temp_CC, scale))) return st;
  if ((st = moons_fft_solve(F, phi, temp_CC))) return st;
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

} // extern "C"
