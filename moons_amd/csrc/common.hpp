// common.hpp -- shared definitions of libmoons_amd (HIP, gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/moons_amd.h"

namespace moons {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
hipStream_t stream();
bool sync_free();
void set_sync_free(bool on);
// Records the next event of the armed kernel timer when `which` is the kernel
// class it times (moons_ktimer_start): call before and after a group of
// launches of that class, the closing call with the group's launch count.
void ktimer_mark(int which, int launches = 1);

#define MOONS_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::moons::fail(MOONS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define MOONS_LAUNCH_CHECK()                                                             \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return ::moons::fail(MOONS_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------- mesh
// Device copies of one direction's 1D coordinate and stencil arrays
// (reference coordinates type, code/generated_code/grid/coordinates.f90:41-70).
struct DirCoefDev {
  const double *hn, *hc, *dhn, *dhc;
  const double *cc2n_D, *cc2n_U; // sc-1
  const double *n2cc_D, *n2cc_U; // sn-1
  const double *theta_D, *theta_U;
  const double *colCC_L[2], *colCC_D[2], *colCC_U[2]; // sc-2
  const double *colN_L[2], *colN_D[2], *colN_U[2];    // sn-2
  // Laplacian diagonal terms on cell centres, laplacian_diagonal
  // (code/handwritten/grid_field/GF_diagonals.f90:15-35) split per direction:
  //   lapA[i] = n2cc_D[i]*cc2n_U[i-1], lapB[i] = n2cc_U[i]*cc2n_D[i]  (0 at ghosts)
  const double *lapA, *lapB; // sc
};

struct MeshDev {
  int sn[3], sc[3];
  DirCoefDev c[3];
};

struct HostCoords {
  int sn, sc;
  std::vector<double> hn, hc, dhn, dhc, cc2n_D, cc2n_U, n2cc_D, n2cc_U, theta_D, theta_U;
  std::vector<double> colCC_L[2], colCC_D[2], colCC_U[2], colN_L[2], colN_D[2], colN_U[2];
  std::vector<double> lapA, lapB;
};

void coords_init(HostCoords &c, const double *hn, int sn);

// ---------------------------------------------------------------- BCs
struct BCDev {
  int type[6];
  double val[6];
  double dh_nhat[6]; // dh at the boundary times the outward normal sign
  // Robin faces (u + c_w du/dh = 0), by the field's location along the face
  // normal: direct rule ug = ui*rob1 + bval*rob2 (cell-centred) or
  // ug = ui + ub*rob1 + bval*rob2 (node-centred); the implicit rule has robi
  // in place of rob1 and no value term.
  double rob1[6], rob2[6], robi[6];
};

// ------------------------------------------------------ reductions
constexpr int kWave = 64;

// Deterministic block sum: wave shuffle tree, then waves summed in fixed
// order by thread 0.  Result valid in thread 0 only.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  if ((tid & (kWave - 1)) == 0) sh[tid / kWave] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0) {
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) r += sh[w];
  }
  __syncthreads();
  return r;
}

// ------------------------------------------------------ XCD-aware tiling
// Workgroups are dispatched round-robin over the 8 XCDs (b -> XCD b%8);
// remap so that each XCD processes one contiguous range of logical tiles
// (contiguous z-slabs), keeping halo planes in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int nx = 8;
  int q = nb / nx, r = nb % nx;
  int x = b % nx, m = b / nx;
  int start = x * q + (x < r ? x : r);
  return start + m;
}

// Interior tile of a cell-centred field: 64 (x) by 4 (y) threads, each
// thread marching kz cells along z.
constexpr int kTX = 64;
constexpr int kTY = 4;
constexpr int kNT = kTX * kTY;

struct CCTiles {
  int n0, n1, n2;   // interior cells per direction
  int s0, s1, s2;   // padded shape
  int ntx, nty, ntz, kz;
  int nblocks;
};

inline CCTiles make_cc_tiles(int s0, int s1, int s2, int kz) {
  CCTiles t;
  t.s0 = s0; t.s1 = s1; t.s2 = s2;
  t.n0 = s0 - 2; t.n1 = s1 - 2; t.n2 = s2 - 2;
  t.kz = kz;
  t.ntx = (t.n0 + kTX - 1) / kTX;
  t.nty = (t.n1 + kTY - 1) / kTY;
  t.ntz = (t.n2 + kz - 1) / kz;
  t.nblocks = t.ntx * t.nty * t.ntz;
  return t;
}

} // namespace moons

// ---------------------------------------------------------------- handles
struct moons_mesh_s {
  moons::HostCoords hc[3];
  moons::MeshDev dev;
  double *dbuf = nullptr; // one allocation holding every 1D array
};

struct moons_field_s {
  moons_mesh_s *m = nullptr;
  int dl = 0;
  int s[3] = {0, 0, 0};
  size_t n = 0;
  double *d = nullptr;
  bool has_bcs = false;
  bool all_neumann = false;
  int bct[6] = {0, 0, 0, 0, 0, 0};
  double bval[6] = {0, 0, 0, 0, 0, 0};
  double c_w[6] = {0, 0, 0, 0, 0, 0};         // wall conductance ratio per face (Robin faces)
  double robin_coeff[6] = {0, 0, 0, 0, 0, 0}; // kept with the BC properties
};

// Copies the BCs of src (types, values, all_Neumann and the BC properties
// c_w and Robin coefficients) to dst, a field of the same location.
inline void copy_bcs_of(moons_field_s *dst, const moons_field_s *src) {
  dst->has_bcs = src->has_bcs;
  dst->all_neumann = src->all_neumann;
  for (int f = 0; f < 6; ++f) {
    dst->bct[f] = src->bct[f];
    dst->bval[f] = src->bval[f];
    dst->c_w[f] = src->c_w[f];
    dst->robin_coeff[f] = src->robin_coeff[f];
  }
}

// Where an interior mesh lies in a total mesh: per direction the ranges of
// coinciding cell centres (C) and nodes (N), 0-based inclusive:
// interior lo, interior hi, total lo, total hi.
struct moons_domain_s {
  moons_mesh_s *mi = nullptr, *mt = nullptr;
  int C[3][4], N[3][4];
};

namespace moons {
// subarray copies over a domain's ranges (the field's location picks C or N per direction)
int launch_embed(moons_field_s *total, const moons_field_s *interior, const moons_domain_s *d);
int launch_extract(moons_field_s *interior, const moons_field_s *total, const moons_domain_s *d);
BCDev make_bcdev(const moons_field_s *f);
int launch_apply_bcs(moons_field_s *u, bool implicit_);
int launch_apply_bcs3(moons_field_s *const u[3], bool implicit_); // three fields (e.g. a VF)
int launch_del(moons_field_s *out, const moons_field_s *f, int n, int dir, int pad, int gen);
int launch_assign_ghost_xperiodic(moons_field_s *u, double val);
int launch_scale(moons_field_s *u, double s);
int launch_interp(moons_field_s *out, const moons_field_s *in, int dir);
// Gate for the general kernels (apply BCs, del, plane assignment): while set,
// those kernels return at once when the device flag is non-zero.
void set_gate(const int *gate);
const int *current_gate();
int launch_plane_set(moons_field_s *u, int d, int idx, double val);
// The device address of the allocation-free solve's scratch (pcg.hip), taken
// once per device with that device current.
int pcg_noalloc_setup(int device);
int launch_plane_scale(moons_field_s *u, int d, int idx, double val);
int launch_cross(moons_field_s *out, const moons_field_s *a, const moons_field_s *b, const moons_field_s *c,
                 const moons_field_s *d, int sign);
int launch_sub(moons_field_s *out, const moons_field_s *a, const moons_field_s *b);
}
