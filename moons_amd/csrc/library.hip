// library.hip -- context, mesh, fields, BCs and the general staggered
// operators of libmoons_amd.  gfx950 only.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace moons {

static thread_local std::string g_err;
static hipStream_t g_stream = nullptr;
static int g_device = -1;
// Device flag that gates the general kernels launched by an iterative
// solver: when *g_gate != 0 they return at once (the loop has exited).
static const int *g_gate = nullptr;

void set_gate(const int *gate) { g_gate = gate; }
const int *current_gate() { return g_gate; }

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
hipStream_t stream() { return g_stream; }

// Set while a caller captures a HIP graph: solvers then launch exactly
// iter_max iterations (each returns at once after convergence, on the
// device) and read nothing back to the host.
static bool g_sync_free = false;
bool sync_free() { return g_sync_free; }
void set_sync_free(bool on) { g_sync_free = on; }

// In-situ kernel timer: event pairs recorded on the library stream around
// every launch of one kernel class while it is armed.
static int g_kt_which = 0;
static std::vector<hipEvent_t> g_kt_ev;
static size_t g_kt_used = 0;
static long long g_kt_launches = 0; // launches inside the recorded pairs
void ktimer_mark(int which, int launches) {
  if (which == 0 || which != g_kt_which || g_kt_used >= g_kt_ev.size() || g_sync_free) return;
  if (g_kt_used % 2 == 1) g_kt_launches += launches; // closing a pair
  (void)hipEventRecord(g_kt_ev[g_kt_used++], g_stream);
}

// ------------------------------------------------------------------ 1D grids
static void add_ghost_nodes(const std::vector<double> &h, double *out) {
  size_t np = h.size();
  double dh0 = h[1] - h[0];
  double dhe = h[np - 1] - h[np - 2];
  out[0] = h[0] - dh0;
  for (size_t i = 0; i < np; ++i) out[i + 1] = h[i];
  out[np + 1] = h[np - 1] + dhe;
}

// ------------------------------------------------------------ coordinates
// init_c + init_stencils_c: code/handwritten/grid/coordinates_extend.f90:78-173,506-527.
// Stencil formulas: code/handwritten/grid/derivative_stencils.f90:31-208,
// code/handwritten/grid/interpolation_stencils.f90:12-35.
void coords_init(HostCoords &c, const double *hn, int sn) {
  c.sn = sn;
  c.sc = sn - 1;
  const int sc = c.sc;
  c.hn.assign(hn, hn + sn);
  c.dhn.resize(sn - 1);
  for (int i = 0; i < sn - 1; ++i) c.dhn[i] = hn[i + 1] - hn[i];
  c.hc.resize(sc);
  for (int i = 0; i < sc; ++i) c.hc[i] = (hn[i + 1] + hn[i]) / 2.0;
  c.dhc.resize(sc - 1);
  for (int i = 0; i < sc - 1; ++i) c.dhc[i] = c.hc[i + 1] - c.hc[i];
  c.theta_D.resize(sc - 1);
  c.theta_U.resize(sc - 1);
  for (int i = 0; i < sc - 1; ++i) {
    c.theta_D[i] = (c.hn[i + 1] - c.hc[i]) / (c.hc[i + 1] - c.hc[i]);
    c.theta_U[i] = 1.0 - c.theta_D[i];
  }
  c.cc2n_D.resize(sc - 1);
  c.cc2n_U.resize(sc - 1);
  for (int i = 0; i < sc - 1; ++i) {
    c.cc2n_U[i] = 1.0 / c.dhc[i];
    c.cc2n_D[i] = -c.cc2n_U[i];
  }
  c.n2cc_D.resize(sn - 1);
  c.n2cc_U.resize(sn - 1);
  for (int i = 0; i < sn - 1; ++i) {
    c.n2cc_U[i] = 1.0 / c.dhn[i];
    c.n2cc_D[i] = -c.n2cc_U[i];
  }
  // Collocated 3-point stencils, one-sided/half-cell rows at the walls.
  // Written with 1-based helpers to follow the reference's index algebra.
  for (int n = 0; n < 2; ++n) {
    c.colCC_L[n].assign(sc - 2, 0.0); c.colCC_D[n].assign(sc - 2, 0.0); c.colCC_U[n].assign(sc - 2, 0.0);
    c.colN_L[n].assign(sn - 2, 0.0); c.colN_D[n].assign(sn - 2, 0.0); c.colN_U[n].assign(sn - 2, 0.0);
  }
  {
    // nodes (dh = dhn, s = sn)
    const double *dh = c.dhn.data();
    auto DH = [&](int i) { return dh[i - 1]; };
    int s = sn;
    double *L1 = c.colN_L[0].data(), *D1 = c.colN_D[0].data(), *U1 = c.colN_U[0].data();
    double *L2 = c.colN_L[1].data(), *D2 = c.colN_D[1].data(), *U2 = c.colN_U[1].data();
    if (s > 3) {
      int i = 2;
      L1[0] = -(DH(i + 1) + 2.0 * DH(i)) / (DH(i) * (DH(i + 1) + DH(i)));
      D1[0] = (DH(i + 1) + DH(i)) / (DH(i + 1) * DH(i));
      U1[0] = -DH(i) / (DH(i + 1) * (DH(i + 1) + DH(i)));
      L2[0] = 2.0 / (DH(i) * (DH(i + 1) + DH(i)));
      D2[0] = -2.0 / (DH(i + 1) * DH(i));
      U2[0] = 2.0 / (DH(i + 1) * (DH(i + 1) + DH(i)));
      for (i = 3; i <= s - 2; ++i) {
        L1[i - 2] = -(DH(i) / (DH(i - 1) * (DH(i - 1) + DH(i))));
        D1[i - 2] = ((-DH(i - 1) + DH(i)) / (DH(i - 1) * DH(i)));
        U1[i - 2] = (DH(i - 1) / (DH(i) * (DH(i - 1) + DH(i))));
        L2[i - 2] = 2.0 / (DH(i - 1) * (DH(i - 1) + DH(i)));
        D2[i - 2] = -2.0 / (DH(i - 1) * DH(i));
        U2[i - 2] = 2.0 / (DH(i) * (DH(i - 1) + DH(i)));
      }
      i = s - 1;
      L1[s - 3] = DH(i - 1) / (DH(i - 2) * (DH(i - 1) + DH(i - 2)));
      D1[s - 3] = -(DH(i - 1) + DH(i - 2)) / (DH(i - 1) * DH(i - 2));
      U1[s - 3] = (2.0 * DH(i - 1) + DH(i - 2)) / (DH(i - 1) * (DH(i - 1) + DH(i - 2)));
      L2[s - 3] = 2.0 / (DH(i - 2) * (DH(i - 1) + DH(i - 2)));
      D2[s - 3] = -2.0 / (DH(i - 1) * DH(i - 2));
      U2[s - 3] = 2.0 / (DH(i - 1) * (DH(i - 1) + DH(i - 2)));
    }
  }
  {
    // cell centres (dh = dhc, s = sc); wall rows use half spacings
    const double *dh = c.dhc.data();
    auto DH = [&](int i) { return dh[i - 1]; };
    int s = sc;
    double *L1 = c.colCC_L[0].data(), *D1 = c.colCC_D[0].data(), *U1 = c.colCC_U[0].data();
    double *L2 = c.colCC_L[1].data(), *D2 = c.colCC_D[1].data(), *U2 = c.colCC_U[1].data();
    if (s > 2) {
      int i = 2;
      double h = 0.5 * DH(i - 1);
      L1[0] = (-(DH(i)) / (h * (h + DH(i))));
      D1[0] = ((-h + DH(i)) / (h * DH(i)));
      U1[0] = (h / (DH(i) * (h + DH(i))));
      L2[0] = 2.0 / (h * (h + DH(i)));
      D2[0] = -2.0 / (h * DH(i));
      U2[0] = 2.0 / (DH(i) * (h + DH(i)));
      for (i = 3; i <= s - 2; ++i) {
        L1[i - 2] = (-DH(i) / (DH(i - 1) * (DH(i - 1) + DH(i))));
        D1[i - 2] = ((-DH(i - 1) + DH(i)) / (DH(i - 1) * DH(i)));
        U1[i - 2] = (DH(i - 1) / (DH(i) * (DH(i - 1) + DH(i))));
        L2[i - 2] = 2.0 / (DH(i - 1) * (DH(i - 1) + DH(i)));
        D2[i - 2] = -(2.0 / (DH(i - 1) * DH(i)));
        U2[i - 2] = 2.0 / (DH(i) * (DH(i - 1) + DH(i)));
      }
      i = s - 1;
      double g = 0.5 * DH(i);
      L1[s - 3] = (-0.5 * DH(i) / (DH(i - 1) * (DH(i - 1) + g)));
      D1[s - 3] = ((-DH(i - 1) + g) / (DH(i - 1) * g));
      U1[s - 3] = (DH(i - 1) / (g * (DH(i - 1) + g)));
      L2[s - 3] = 2.0 / (DH(i - 1) * (DH(i - 1) + g));
      D2[s - 3] = -2.0 / (DH(i - 1) * g);
      U2[s - 3] = 2.0 / (g * (DH(i - 1) + g));
    }
  }
  c.lapA.assign(sc, 0.0);
  c.lapB.assign(sc, 0.0);
  for (int i = 1; i < sc - 1; ++i) {
    c.lapA[i] = c.n2cc_D[i] * c.cc2n_U[i - 1];
    c.lapB[i] = c.n2cc_U[i] * c.cc2n_D[i];
  }
}

// dh of a face's sub-domain (FSD%dh, code/handwritten/sub_domain/face_SD_extend.f90:63):
// the spacing between the ghost and boundary node at a min face, the last
// interior cell at a max face (get_dh_boundary over the face grid get_face_GI
// keeps, code/handwritten/sub_domain/sub_domain_extend.f90:125-134,
// code/handwritten/grid/grid_extend.f90:260-268).
static double face_dh(const HostCoords &c, int is_max) { return is_max ? c.dhn[c.sn - 3] : c.dhn[0]; }

BCDev make_bcdev(const moons_field_s *f) {
  BCDev b;
  for (int face = 0; face < 6; ++face) {
    b.type[face] = f->has_bcs ? f->bct[face] : MOONS_BC_NONE;
    b.val[face] = f->bval[face];
    int d = face / 2;
    const HostCoords &c = f->m->hc[d];
    const double dh = face_dh(c, face % 2);
    const double nhat = (face % 2) ? 1.0 : -1.0;
    b.dh_nhat[face] = dh * nhat;
    // Robin coefficients in the reference's operation order: apply_Robin_C/N
    // (code/handwritten/BCs/direct/apply_BCs_faces_raw.f90:236-282, whose bvals
    // coefficient of the C rule is overwritten by 0) and the implicit twins
    // (code/handwritten/BCs/implicit/apply_BCs_faces_raw_implicit.f90:213-257).
    const double cw = f->c_w[face];
    if ((f->dl >> d) & 1) {
      b.rob1[face] = -(2.0 * dh / cw * nhat);
      b.rob2[face] = 2.0 * dh / cw * nhat;
      b.robi[face] = -(2.0 * dh * nhat / cw);
    } else {
      const double c1 = (2.0 * cw * nhat / dh - 1.0) / (2.0 * cw * nhat / dh + 1.0);
      b.rob1[face] = c1;
      b.rob2[face] = 0.0;
      b.robi[face] = c1;
    }
  }
  return b;
}

// ---------------------------------------------------------------- kernels

__device__ __forceinline__ size_t idx3(const int s[3], int i, int j, int k) {
  return (size_t)i + (size_t)s[0] * ((size_t)j + (size_t)s[1] * (size_t)k);
}

struct Shape3 {
  int s[3];
};

// One face of apply_BCs / apply_BCs_implicit over the full plane (tangential
// ghosts included), raw rules of code/handwritten/BCs/direct/apply_BCs_faces_raw.f90
// Up to three fields' faces of one direction (k_apply_bc_multi).
struct BcMulti {
  int on[3];
  double *u[3];
  Shape3 S[3];
  int nalong[3];
  int type[3][2];
  double val[3][2], dh[3][2], r1[3][2], r2[3][2];
};

// and code/handwritten/BCs/implicit/apply_BCs_faces_raw_implicit.f90.
// r1, r2: the face's Robin coefficients for the rule applied (implicit: robi, 0).
__device__ __forceinline__ void apply_bc_point(double *u, const Shape3 &S, int face, int type, int nalong, double bval,
                                               double dh_nhat, double r1, double r2, int implicit_, int ja, int jb) {
  const int d = face / 2, is_max = face % 2;
  const int a = (d + 1) % 3, b = (d + 2) % 3;
  const int s = S.s[d];
  int G, B, I, Io;
  if (!nalong) {
    G = is_max ? s - 1 : 0; B = -1; I = is_max ? s - 2 : 1; Io = is_max ? 1 : s - 2;
  } else {
    G = is_max ? s - 1 : 0; B = is_max ? s - 2 : 1; I = is_max ? s - 3 : 2; Io = is_max ? 2 : s - 3;
  }
  int ii[3];
  ii[a] = ja; ii[b] = jb;
  ii[d] = G; const size_t ig = idx3(S.s, ii[0], ii[1], ii[2]);
  ii[d] = I; const size_t iin = idx3(S.s, ii[0], ii[1], ii[2]);
  switch (type) {
    case MOONS_BC_DIRICHLET:
      if (!nalong) {
        u[ig] = implicit_ ? -u[iin] : 2.0 * bval - u[iin];
      } else {
        ii[d] = B; const size_t ib = idx3(S.s, ii[0], ii[1], ii[2]);
        double ub = implicit_ ? 0.0 : bval;
        u[ib] = ub;
        u[ig] = implicit_ ? -u[iin] : 2.0 * ub - u[iin];
      }
      break;
    case MOONS_BC_NEUMANN:
      if (implicit_) u[ig] = u[iin];
      else if (!nalong) u[ig] = u[iin] + bval * dh_nhat;
      else u[ig] = u[iin] + 2.0 * bval * dh_nhat;
      break;
    case MOONS_BC_PERIODIC: {
      int Gp = G, Iop = Io;
      if (nalong && is_max && s > 4) { Gp = B; Iop = 1; }
      ii[d] = Gp; const size_t igp = idx3(S.s, ii[0], ii[1], ii[2]);
      ii[d] = Iop; const size_t io = idx3(S.s, ii[0], ii[1], ii[2]);
      u[igp] = u[io];
    } break;
    case MOONS_BC_SYMMETRIC: u[ig] = u[iin]; break;
    case MOONS_BC_ANTISYMMETRIC: u[ig] = -u[iin]; break;
    case MOONS_BC_ROBIN:
      if (!nalong) {
        u[ig] = implicit_ ? u[iin] * r1 : u[iin] * r1 + bval * r2;
      } else {
        ii[d] = B; const size_t ib = idx3(S.s, ii[0], ii[1], ii[2]);
        u[ig] = implicit_ ? u[iin] + u[ib] * r1 : u[iin] + u[ib] * r1 + bval * r2;
      }
      break;
    default: break;
  }
}

// Robin coefficients of a face for the direct or implicit rule.
static inline void robin_pair(const BCDev &b, int face, bool implicit_, double &r1, double &r2) {
  r1 = implicit_ ? b.robi[face] : b.rob1[face];
  r2 = implicit_ ? 0.0 : b.rob2[face];
}

__global__ void __launch_bounds__(256) k_apply_bc_face(double *u, Shape3 S, int face, int type, int nalong,
                                                       double bval, double dh_nhat, double r1, double r2,
                                                       int implicit_, const int *gate) {
  if (gate && *gate) return;
  const int d = face / 2;
  const int a = (d + 1) % 3, b = (d + 2) % 3;
  const int na = S.s[a], nb = S.s[b];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na * nb) return;
  apply_bc_point(u, S, face, type, nalong, bval, dh_nhat, r1, r2, implicit_, t % na, t / na);
}

// Both faces of direction d in one launch: the first na*nb threads the min
// face, the others the max face.  Used when the direction has at least 6
// points, where the planes one face writes are none of those the other reads,
// so the result equals the min-then-max sequence.
__global__ void __launch_bounds__(256) k_apply_bc_pair(double *u, Shape3 S, int d, int type0, int type1, int nalong,
                                                       double bval0, double bval1, double dh0, double dh1,
                                                       double r10, double r11, double r20, double r21,
                                                       int implicit_, const int *gate) {
  if (gate && *gate) return;
  const int a = (d + 1) % 3, b = (d + 2) % 3;
  const int na = S.s[a], nb = S.s[b];
  const int np = na * nb;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * np) return;
  const int side = t >= np ? 1 : 0;
  const int q = t - side * np;
  apply_bc_point(u, S, 2 * d + side, side ? type1 : type0, nalong, side ? bval1 : bval0, side ? dh1 : dh0,
                 side ? r11 : r10, side ? r21 : r20, implicit_, q % na, q / na);
}

// Assigns val to one whole plane idx along direction d.
__global__ void __launch_bounds__(256) k_assign_plane(double *u, Shape3 S, int d, int idx, double val,
                                                      const int *gate) {
  if (gate && *gate) return;
  const int a = (d + 1) % 3, b = (d + 2) % 3;
  const int na = S.s[a], nb = S.s[b];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na * nb) return;
  int ii[3];
  ii[a] = t % na; ii[b] = t / na; ii[d] = idx;
  u[idx3(S.s, ii[0], ii[1], ii[2])] = val;
}

// Assigns val to the ghost planes of the faces in mask (bit f: face f = 2d +
// side, plane 0 or s_d - 1) in one launch: thread t walks the faces' planes
// one after another.  Edges shared by two faces get the same value twice.
__global__ void __launch_bounds__(256) k_assign_faces(double *u, Shape3 S, int mask, double val, const int *gate) {
  if (gate && *gate) return;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int f = 0; f < 6; ++f) {
    if (!((mask >> f) & 1)) continue;
    const int d = f / 2, a = (d + 1) % 3, b = (d + 2) % 3;
    const int np = S.s[a] * S.s[b];
    if (t < np) {
      int ii[3];
      ii[a] = t % S.s[a]; ii[b] = t / S.s[a]; ii[d] = (f & 1) ? S.s[d] - 1 : 0;
      u[idx3(S.s, ii[0], ii[1], ii[2])] = val;
      return;
    }
    t -= np;
  }
}

static int launch_assign_faces(moons_field_s *u, int mask, double val) {
  if (!mask) return MOONS_OK;
  Shape3 S{{u->s[0], u->s[1], u->s[2]}};
  size_t n = 0;
  for (int f = 0; f < 6; ++f)
    if ((mask >> f) & 1) n += (size_t)u->s[((f / 2) + 1) % 3] * u->s[((f / 2) + 2) % 3];
  hipLaunchKernelGGL(k_assign_faces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g_stream, u->d, S, mask, val,
                     g_gate);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// Multiplies one whole plane idx along direction d by val.
__global__ void __launch_bounds__(256) k_scale_plane(double *u, Shape3 S, int d, int idx, double val,
                                                     const int *gate) {
  if (gate && *gate) return;
  const int a = (d + 1) % 3, b = (d + 2) % 3;
  const int na = S.s[a], nb = S.s[b];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na * nb) return;
  int ii[3];
  ii[a] = t % na; ii[b] = t / na; ii[d] = idx;
  const size_t q = idx3(S.s, ii[0], ii[1], ii[2]);
  u[q] = u[q] * val;
}

// del along dir for every output element of the padded range, one thread per
// output element.  kind: 0 stag CC->N (gt=1), 1 stag N->CC (gt=0), 2 col CC, 3 col N.
// Follows code/handwritten/stencils/stencils.f90 element by element, including
// the end entries the staggered kernels zero before accumulating.
template <int DIR, int KIND>
__global__ void __launch_bounds__(256) k_del(double *__restrict__ out, const double *__restrict__ f, Shape3 So,
                                             Shape3 Sf, const double *__restrict__ L, const double *__restrict__ D,
                                             const double *__restrict__ U, int gen, int pad, const int *gate) {
  if (gate && *gate) return;
  constexpr int dir = DIR, kind = KIND;
  const int a = (dir + 1) % 3, b = (dir + 2) % 3;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const int nd = So.s[dir];
  const int nlo = Sf.s[lo] - 2 * pad, nhi = Sf.s[hi] - 2 * pad;
  // 32-bit: the launcher rejects fields of 2^31 or more elements
  const int total = nd * nlo * nhi;
  // XCD-contiguous block numbering: each XCD takes one contiguous range of
  // planes, so the neighbouring input planes a derivative reads share its L2
  const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x) * (int)blockDim.x + (int)threadIdx.x;
  if (t >= total || nlo <= 0 || nhi <= 0) return;
  // threads take the output elements in memory order (x fastest), so that a
  // wave's loads and stores are contiguous whatever the direction
  int ext[3], io[3];
  ext[dir] = nd; ext[lo] = nlo; ext[hi] = nhi;
  io[0] = t % ext[0];
  io[1] = (t / ext[0]) % ext[1];
  io[2] = t / (ext[0] * ext[1]);
  io[lo] += pad;
  io[hi] += pad;
  const int o = io[dir], plo = io[lo], phi = io[hi];
  const size_t oo = idx3(So.s, io[0], io[1], io[2]);
  const size_t fs = (dir == 0) ? 1 : (dir == 1 ? (size_t)Sf.s[0] : (size_t)Sf.s[0] * Sf.s[1]);
  int iff[3];
  iff[dir] = 0; iff[lo] = plo; iff[hi] = phi;
  const double *fp = f + idx3(Sf.s, iff[0], iff[1], iff[2]);
  const int s = Sf.s[dir];
  double cur = 0.0; // the output's old value, read only where the result uses it
  double res;
  if (kind == 0 || kind == 1) {
    const int gt = (kind == 0) ? 1 : 0;
    const bool zero_end = (o == 0 || o == nd - 1);
    const int i = o - gt;
    const bool written = (i >= 0 && i <= s - 2);
    if (!zero_end && (gen != 1 || !written)) cur = out[oo];
    if (!written) {
      res = cur;
    } else {
      const double x0 = fp[(size_t)i * fs] * D[i];
      const double x1 = fp[(size_t)(i + 1) * fs] * U[i];
      if (gen == 1) res = x0 + x1;
      else if (gen == 2) res = cur + x0 + x1;
      else res = cur - (x0 + x1);
    }
  } else {
    // collocated, pad1 = pad2 = 0
    if (gen != 1) cur = out[oo];
#define FF(q) fp[(size_t)(q) * fs]
    if (o == 0 || o == s - 1) {
      res = 0.0;
    } else if (o == 1) {
      const double t1 = 0.5 * (FF(0) + FF(1)) * L[0] * 1.0;
      const double t2 = FF(0) * L[0] * 0.0;
      if (kind == 2) {
        if (gen == 1) res = t1 + t2 + FF(1) * D[0] + FF(2) * U[0];
        else if (gen == 2) res = cur + t1 + t2 + FF(1) * D[0] + FF(2) * U[0];
        else res = cur - (t1 + t2 + FF(1) * D[0] + FF(2) * U[0]);
      } else {
        if (gen == 1) res = FF(1) * L[0] + FF(2) * D[0] + FF(3) * U[0];
        else if (gen == 2) res = cur + FF(1) * L[0] + FF(2) * D[0] + FF(3) * U[0];
        else res = cur - (FF(1) * L[0] + FF(2) * D[0] + FF(3) * U[0]);
      }
    } else if (o == s - 2) {
      if (kind == 2) {
        const double t3 = 0.5 * (FF(s - 2) + FF(s - 1)) * U[s - 3] * 1.0;
        const double t4 = FF(s - 1) * U[s - 3] * 0.0;
        if (gen == 1) res = FF(s - 3) * L[s - 3] + FF(s - 2) * D[s - 3] + t3 + t4;
        else if (gen == 2) res = cur + FF(s - 3) * L[s - 3] + FF(s - 2) * D[s - 3] + t3 + t4;
        else res = cur - (FF(s - 3) * L[s - 3] + FF(s - 2) * D[s - 3] + t3 + t4);
      } else {
        if (gen == 1) res = FF(s - 4) * L[s - 3] + FF(s - 3) * D[s - 3] + FF(s - 2) * U[s - 3];
        else if (gen == 2) res = cur + FF(s - 4) * L[s - 3] + FF(s - 3) * D[s - 3] + FF(s - 2) * U[s - 3];
        else res = cur - (FF(s - 4) * L[s - 3] + FF(s - 3) * D[s - 3] + FF(s - 2) * U[s - 3]);
      }
    } else {
      const int i = o;
      const double x0 = FF(i - 1) * L[i - 1];
      if (gen == 1) res = x0 + FF(i) * D[i - 1] + FF(i + 1) * U[i - 1];
      else if (gen == 2) res = cur + x0 + FF(i) * D[i - 1] + FF(i + 1) * U[i - 1];
      else res = cur - (x0 + FF(i) * D[i - 1] + FF(i + 1) * U[i - 1]);
    }
#undef FF
  }
  out[oo] = res;
}

// interpO2_SF along dir, one thread per output element: C->N with theta
// weights and extrapolated end nodes (GF_interp.f90:19-47, GF_extrap.f90:17-46)
// or N->C half sums (GF_interp.f90:49-74).
template <int DIR, int C2N>
__global__ void __launch_bounds__(256) k_interp(double *__restrict__ out, const double *__restrict__ in, Shape3 So,
                                                Shape3 Si, const double *__restrict__ tD,
                                                const double *__restrict__ tU) {
  constexpr int dir = DIR, c2n = C2N;
  const int a = (dir + 1) % 3, b = (dir + 2) % 3;
  const int nd = So.s[dir];
  // 32-bit: the launcher rejects fields of 2^31 or more elements
  const int total = nd * So.s[a] * So.s[b];
  const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x) * (int)blockDim.x + (int)threadIdx.x; // as in k_del
  if (t >= total) return;
  // threads take the output elements in memory order (x fastest)
  int io[3], ii[3];
  io[0] = t % So.s[0];
  io[1] = (t / So.s[0]) % So.s[1];
  io[2] = t / (So.s[0] * So.s[1]);
  const int o = io[dir], pa = io[a], pb = io[b];
  ii[dir] = 0; ii[a] = pa; ii[b] = pb;
  const double *C = in + idx3(Si.s, ii[0], ii[1], ii[2]);
  const size_t fs = (dir == 0) ? 1 : (dir == 1 ? (size_t)Si.s[0] : (size_t)Si.s[0] * Si.s[1]);
  double v;
  if (c2n) {
    const int sC = Si.s[dir], sN = nd;
    if (o >= 1 && o <= sC - 1) {
      v = C[(size_t)o * fs] * tD[o - 1] + C[(size_t)(o - 1) * fs] * tU[o - 1];
    } else if (o == 0) {
      const double n1 = C[fs] * tD[0] + C[0] * tU[0];
      v = 2.0 * C[0] - n1;
    } else { // o == sN - 1
      const double nl = C[(size_t)(sC - 1) * fs] * tD[sC - 2] + C[(size_t)(sC - 2) * fs] * tU[sC - 2];
      v = 2.0 * C[(size_t)(sC - 1) * fs] - nl;
      (void)sN;
    }
  } else {
    v = 0.5 * (C[(size_t)o * fs] + C[(size_t)(o + 1) * fs]);
  }
  out[idx3(So.s, io[0], io[1], io[2])] = v;
}

// out = a*b - c*d (sign > 0) or -(a*b - c*d) (sign < 0): the cross-product
// components of GF_cross_product.f90 (x and z: Ay Bz - Az By; y: -(Ax Bz - Az Bx)).
__global__ void __launch_bounds__(256) k_cross(double *out, const double *a, const double *b, const double *c,
                                               const double *d, size_t n, int sign) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double v = a[t] * b[t] - c[t] * d[t];
  out[t] = sign > 0 ? v : -v;
}

__global__ void __launch_bounds__(256) k_scale(double *u, size_t n, double s) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) u[t] = u[t] * s;
}
__global__ void __launch_bounds__(256) k_sub(double *out, const double *a, const double *b, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = a[t] - b[t];
}
__global__ void __launch_bounds__(256) k_fill(double *u, size_t n, double v) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) u[t] = v;
}

// volume_DL_GF: code/handwritten/grid_field/GF_distributions.f90:61-88
__global__ void __launch_bounds__(256) k_volume(double *vol, Shape3 S, int e0, int e1, int e2,
                                                const double *dh0, const double *dh1, const double *dh2) {
  size_t n = (size_t)S.s[0] * S.s[1] * S.s[2];
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int i = (int)(t % S.s[0]);
  int j = (int)((t / S.s[0]) % S.s[1]);
  int k = (int)(t / ((size_t)S.s[0] * S.s[1]));
  double v = 0.0;
  if (i >= 1 && i < S.s[0] - 1 && j >= 1 && j < S.s[1] - 1 && k >= 1 && k < S.s[2] - 1)
    v = dh0[i - e0] * dh1[j - e1] * dh2[k - e2];
  vol[t] = v;
}

// --------------------------------------------------------------- launchers
static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

int launch_apply_bcs(moons_field_s *u, bool implicit_) {
  if (!u->has_bcs) return MOONS_OK;
  Shape3 S{{u->s[0], u->s[1], u->s[2]}};
  BCDev b = make_bcdev(u);
  for (int d = 0; d < 3; ++d) {
    const int a = (d + 1) % 3, bb = (d + 2) % 3;
    const size_t np = (size_t)u->s[a] * u->s[bb];
    const int nalong = (u->dl >> d) & 1;
    const int f0 = 2 * d, f1 = 2 * d + 1;
    double r1[2], r2[2];
    robin_pair(b, f0, implicit_, r1[0], r2[0]);
    robin_pair(b, f1, implicit_, r1[1], r2[1]);
    if (b.type[f0] != MOONS_BC_NONE && b.type[f1] != MOONS_BC_NONE && u->s[d] >= 6) {
      hipLaunchKernelGGL(k_apply_bc_pair, dim3(nblk(2 * np)), dim3(256), 0, g_stream, u->d, S, d, b.type[f0],
                         b.type[f1], nalong, b.val[f0], b.val[f1], b.dh_nhat[f0], b.dh_nhat[f1], r1[0], r1[1], r2[0],
                         r2[1], implicit_ ? 1 : 0, g_gate);
      MOONS_LAUNCH_CHECK();
      continue;
    }
    for (int face : {f0, f1}) {
      if (b.type[face] == MOONS_BC_NONE) continue;
      hipLaunchKernelGGL(k_apply_bc_face, dim3(nblk(np)), dim3(256), 0, g_stream, u->d, S, face, b.type[face],
                         nalong, b.val[face], b.dh_nhat[face], r1[face - f0], r2[face - f0], implicit_ ? 1 : 0,
                         g_gate);
      MOONS_LAUNCH_CHECK();
    }
  }
  return MOONS_OK;
}

// apply_BCs of three fields: per direction, the faces of every field whose
// direction has at least 6 points in one launch (fields are independent, so
// only each field's own direction order matters), the others face by face.
__global__ void __launch_bounds__(256) k_apply_bc_multi(BcMulti M, int d, int implicit_, const int *gate) {
  if (gate && *gate) return;
  const int f = blockIdx.y;
  if (!M.on[f]) return;
  const Shape3 &S = M.S[f];
  const int a = (d + 1) % 3, b = (d + 2) % 3;
  const int na = S.s[a], nb = S.s[b];
  const int np = na * nb;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * np) return;
  const int side = t >= np ? 1 : 0;
  const int q = t - side * np;
  apply_bc_point(M.u[f], S, 2 * d + side, M.type[f][side], M.nalong[f], M.val[f][side], M.dh[f][side],
                 M.r1[f][side], M.r2[f][side], implicit_, q % na, q / na);
}

int launch_apply_bcs3(moons_field_s *const u[3], bool implicit_) {
  for (int d = 0; d < 3; ++d) {
    BcMulti M{};
    size_t npmax = 0;
    for (int f = 0; f < 3; ++f) {
      const moons_field_s *x = u[f];
      if (!x->has_bcs) continue;
      BCDev b = make_bcdev(x);
      const int f0 = 2 * d, f1 = 2 * d + 1;
      const int a = (d + 1) % 3, bb = (d + 2) % 3;
      const size_t np = (size_t)x->s[a] * x->s[bb];
      if (b.type[f0] != MOONS_BC_NONE && b.type[f1] != MOONS_BC_NONE && x->s[d] >= 6) {
        M.on[f] = 1;
        M.u[f] = x->d;
        M.S[f] = Shape3{{x->s[0], x->s[1], x->s[2]}};
        M.nalong[f] = (x->dl >> d) & 1;
        for (int side = 0; side < 2; ++side) {
          M.type[f][side] = b.type[2 * d + side];
          M.val[f][side] = b.val[2 * d + side];
          M.dh[f][side] = b.dh_nhat[2 * d + side];
          robin_pair(b, 2 * d + side, implicit_, M.r1[f][side], M.r2[f][side]);
        }
        npmax = std::max(npmax, np);
      } else {
        Shape3 S{{x->s[0], x->s[1], x->s[2]}};
        for (int face : {f0, f1}) {
          if (b.type[face] == MOONS_BC_NONE) continue;
          double r1, r2;
          robin_pair(b, face, implicit_, r1, r2);
          hipLaunchKernelGGL(k_apply_bc_face, dim3(nblk(np)), dim3(256), 0, g_stream, x->d, S, face, b.type[face],
                             (x->dl >> d) & 1, b.val[face], b.dh_nhat[face], r1, r2, implicit_ ? 1 : 0, g_gate);
          MOONS_LAUNCH_CHECK();
        }
      }
    }
    if (npmax > 0) {
      hipLaunchKernelGGL(k_apply_bc_multi, dim3(nblk(2 * npmax), 3), dim3(256), 0, g_stream, M, d, implicit_ ? 1 : 0,
                         g_gate);
      MOONS_LAUNCH_CHECK();
    }
  }
  return MOONS_OK;
}

int launch_assign_ghost_xperiodic(moons_field_s *u, double val) {
  int mask = 0;
  for (int face = 0; face < 6; ++face)
    if (!(u->has_bcs && u->bct[face] == MOONS_BC_PERIODIC)) mask |= 1 << face;
  return launch_assign_faces(u, mask, val);
}

static int launch_plane(moons_field_s *u, int d, int idx, double val, bool scale) {
  if (d < 0 || d > 2 || idx < 0 || idx >= u->s[d]) return fail(MOONS_ERR_ARG, "plane index out of range");
  Shape3 S{{u->s[0], u->s[1], u->s[2]}};
  const int a = (d + 1) % 3, bb = (d + 2) % 3;
  const size_t np = (size_t)u->s[a] * u->s[bb];
  if (scale)
    hipLaunchKernelGGL(k_scale_plane, dim3(nblk(np)), dim3(256), 0, g_stream, u->d, S, d, idx, val, g_gate);
  else
    hipLaunchKernelGGL(k_assign_plane, dim3(nblk(np)), dim3(256), 0, g_stream, u->d, S, d, idx, val, g_gate);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int launch_plane_set(moons_field_s *u, int d, int idx, double val) { return launch_plane(u, d, idx, val, false); }
int launch_plane_scale(moons_field_s *u, int d, int idx, double val) { return launch_plane(u, d, idx, val, true); }

int launch_del(moons_field_s *out, const moons_field_s *f, int n, int dir, int pad, int gen) {
  if (dir < 0 || dir > 2 || (n != 1 && n != 2) || gen < 1 || gen > 3 || pad < 0 || pad > 1)
    return fail(MOONS_ERR_ARG, "del: bad n/dir/pad/gen");
  if (out->m != f->m) return fail(MOONS_ERR_ARG, "del: fields on different meshes");
  int nin = (f->dl >> dir) & 1, nout = (out->dl >> dir) & 1;
  if ((f->dl & ~(1 << dir)) != (out->dl & ~(1 << dir)))
    return fail(MOONS_ERR_ARG, "Error: diffType undetermined in ops_del (locations differ off dir)");
  const DirCoefDev &c = f->m->dev.c[dir];
  int kind;
  const double *L = nullptr, *D, *U;
  if (!nin && nout) { kind = 0; D = c.cc2n_D; U = c.cc2n_U; }
  else if (nin && !nout) { kind = 1; D = c.n2cc_D; U = c.n2cc_U; }
  else if (!nin) { kind = 2; L = c.colCC_L[n - 1]; D = c.colCC_D[n - 1]; U = c.colCC_U[n - 1]; }
  else { kind = 3; L = c.colN_L[n - 1]; D = c.colN_D[n - 1]; U = c.colN_U[n - 1]; }
  Shape3 So{{out->s[0], out->s[1], out->s[2]}}, Sf{{f->s[0], f->s[1], f->s[2]}};
  int a = (dir + 1) % 3, b = (dir + 2) % 3;
  size_t total = (size_t)out->s[dir] * (size_t)(f->s[a] - 2 * pad) * (size_t)(f->s[b] - 2 * pad);
  if (total >= ((size_t)1 << 31) || out->n >= ((size_t)1 << 31) || f->n >= ((size_t)1 << 31))
    return fail(MOONS_ERR_ARG, "del: fields of 2^31 or more elements are not supported");
  // direction and kind as template arguments: no run-time branch on them per element
  using KDel = void (*)(double *, const double *, Shape3, Shape3, const double *, const double *, const double *, int,
                        int, const int *);
  static const KDel kdel[3][4] = {{k_del<0, 0>, k_del<0, 1>, k_del<0, 2>, k_del<0, 3>},
                                  {k_del<1, 0>, k_del<1, 1>, k_del<1, 2>, k_del<1, 3>},
                                  {k_del<2, 0>, k_del<2, 1>, k_del<2, 2>, k_del<2, 3>}};
  hipLaunchKernelGGL(kdel[dir][kind], dim3(nblk(total)), dim3(256), 0, g_stream, out->d, f->d, So, Sf, L, D, U, gen,
                     pad, g_gate);
  MOONS_LAUNCH_CHECK();
  if (gen == 1 && pad > 0) {
    // diff(): assign with padding zeroes the ghost planes of the other directions
    int st = launch_assign_faces(out, (3 << (2 * a)) | (3 << (2 * b)), 0.0);
    if (st) return st;
  }
  return MOONS_OK;
}

int launch_interp(moons_field_s *out, const moons_field_s *in, int dir) {
  if (dir < 0 || dir > 2 || out->m != in->m) return fail(MOONS_ERR_ARG, "interp: bad dir or mesh");
  const int nout = (out->dl >> dir) & 1, nin = (in->dl >> dir) & 1;
  if (nout == nin || (out->dl & ~(1 << dir)) != (in->dl & ~(1 << dir)))
    return fail(MOONS_ERR_ARG, "Error: bad DLs in interpO2_SF (ops_interp.f90)");
  const DirCoefDev &c = in->m->dev.c[dir];
  Shape3 So{{out->s[0], out->s[1], out->s[2]}}, Si{{in->s[0], in->s[1], in->s[2]}};
  if (out->n >= ((size_t)1 << 31) || in->n >= ((size_t)1 << 31))
    return fail(MOONS_ERR_ARG, "interp: fields of 2^31 or more elements are not supported");
  using KInterp = void (*)(double *, const double *, Shape3, Shape3, const double *, const double *);
  static const KInterp kint[3][2] = {{k_interp<0, 0>, k_interp<0, 1>}, {k_interp<1, 0>, k_interp<1, 1>},
                                     {k_interp<2, 0>, k_interp<2, 1>}};
  hipLaunchKernelGGL(kint[dir][nout], dim3(nblk(out->n)), dim3(256), 0, g_stream, out->d, in->d, So, Si, c.theta_D,
                     c.theta_U);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int launch_cross(moons_field_s *out, const moons_field_s *a, const moons_field_s *b, const moons_field_s *c,
                 const moons_field_s *d, int sign) {
  if (a->n != out->n || b->n != out->n || c->n != out->n || d->n != out->n)
    return fail(MOONS_ERR_ARG, "cross_product: fields must be collocated");
  hipLaunchKernelGGL(k_cross, dim3(nblk(out->n)), dim3(256), 0, g_stream, out->d, a->d, b->d, c->d, d->d, out->n, sign);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int launch_scale(moons_field_s *u, double s) {
  hipLaunchKernelGGL(k_scale, dim3(nblk(u->n)), dim3(256), 0, g_stream, u->d, u->n, s);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int launch_sub(moons_field_s *out, const moons_field_s *a, const moons_field_s *b) {
  hipLaunchKernelGGL(k_sub, dim3(nblk(out->n)), dim3(256), 0, g_stream, out->d, a->d, b->d, out->n);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

} // namespace moons

using namespace moons;

// =================================================================== C ABI
extern "C" {

int moons_init(int device) {
  int n = 0;
  MOONS_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(MOONS_ERR_ARG, "moons_init: no such device");
  MOONS_HIP(hipSetDevice(device));
  if (int st = pcg_noalloc_setup(device)) return st;
  if (g_stream && g_device == device) return MOONS_OK;
  if (g_stream) MOONS_HIP(hipStreamDestroy(g_stream));
  MOONS_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  return MOONS_OK;
}

int moons_ktimer_start(int which, int max_launches) {
  if (which < 1 || max_launches < 1) return fail(MOONS_ERR_ARG, "moons_ktimer_start: bad argument");
  if (g_kt_which) return fail(MOONS_ERR_STATE, "moons_ktimer_start: a timer is already armed");
  g_kt_ev.assign((size_t)2 * max_launches, nullptr);
  for (auto &e : g_kt_ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (auto &f : g_kt_ev)
        if (f) (void)hipEventDestroy(f);
      g_kt_ev.clear();
      return fail(MOONS_ERR_HIP, "moons_ktimer_start: hipEventCreate failed");
    }
  g_kt_used = 0;
  g_kt_launches = 0;
  g_kt_which = which;
  return MOONS_OK;
}

int moons_ktimer_stop(int *launches, double *avg_ms) {
  if (!g_kt_which) return fail(MOONS_ERR_STATE, "moons_ktimer_stop: no timer armed");
  double sum = 0.0;
  const size_t pairs = g_kt_used / 2;
  int st = MOONS_OK;
  if (pairs > 0 && hipEventSynchronize(g_kt_ev[2 * pairs - 1]) != hipSuccess)
    st = fail(MOONS_ERR_HIP, "moons_ktimer_stop: hipEventSynchronize failed");
  for (size_t q = 0; q < pairs && !st; ++q) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_kt_ev[2 * q], g_kt_ev[2 * q + 1]) != hipSuccess) {
      st = fail(MOONS_ERR_HIP, "moons_ktimer_stop: hipEventElapsedTime failed");
      break;
    }
    sum += ms;
  }
  for (auto &e : g_kt_ev) (void)hipEventDestroy(e);
  g_kt_ev.clear();
  g_kt_which = 0;
  g_kt_used = 0;
  if (st) return st;
  if (launches) *launches = (int)g_kt_launches;
  if (avg_ms) *avg_ms = g_kt_launches ? sum / (double)g_kt_launches : 0.0;
  return MOONS_OK;
}

int moons_stream_create(void **out) {
  if (!out) return fail(MOONS_ERR_ARG, "moons_stream_create: null argument");
  hipStream_t st = nullptr;
  MOONS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *out = st;
  return MOONS_OK;
}

int moons_stream_destroy(void *st) {
  if (!st) return MOONS_OK;
  MOONS_HIP(hipStreamSynchronize(static_cast<hipStream_t>(st)));
  MOONS_HIP(hipStreamDestroy(static_cast<hipStream_t>(st)));
  return MOONS_OK;
}

int moons_stream_synchronize(void *st) {
  MOONS_HIP(hipStreamSynchronize(st ? static_cast<hipStream_t>(st) : g_stream));
  return MOONS_OK;
}

int moons_finalize(void) {
  if (g_stream) {
    MOONS_HIP(hipStreamSynchronize(g_stream));
    MOONS_HIP(hipStreamDestroy(g_stream));
  }
  g_stream = nullptr;
  g_device = -1;
  return MOONS_OK;
}

const char *moons_last_error(void) { return g_err.c_str(); }

int moons_synchronize(void) {
  if (!g_stream) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  MOONS_HIP(hipStreamSynchronize(g_stream));
  return MOONS_OK;
}

void *moons_stream(void) { return (void *)g_stream; }

int moons_nodes_uniform(double hmin, double hmax, int N, double *hn) {
  if (N < 2 || !hn || !(hmax > hmin)) return fail(MOONS_ERR_ARG, "moons_nodes_uniform: need N >= 2, hmax > hmin");
  std::vector<double> h(N + 1);
  double dh = (hmax - hmin) / (double)N;
  for (int i = 0; i <= N; ++i) h[i] = hmin + (double)i * dh;
  add_ghost_nodes(h, hn);
  return MOONS_OK;
}

int moons_nodes_roberts_both(double hmin, double hmax, int N, double beta, double *hn) {
  if (N < 2 || !hn || !(hmax > hmin) || !(beta > 1.0))
    return fail(MOONS_ERR_ARG, "moons_nodes_roberts_both: need N >= 2, hmax > hmin, beta > 1");
  const double one = 1.0, two = 2.0, a = 0.5, b = beta;
  const double g = (b + one) / (b - one);
  const double dh = (hmax - hmin) / (double)N;
  std::vector<double> h(N + 1);
  for (int i = 0; i <= N; ++i) {
    double hb = hmin + (double)i * dh;
    hb = (hb - hmin) / (hmax - hmin);
    double e = std::pow(g, (hb - a) / (one - a));
    h[i] = ((b + two * a) * e - b + two * a) / ((two * a + one) * (one + e));
  }
  for (int i = 0; i <= N; ++i) h[i] = hmin + (hmax - hmin) * h[i];
  add_ghost_nodes(h, hn);
  return MOONS_OK;
}

int moons_mesh_create(const double *hx, int sx, const double *hy, int sy, const double *hz, int sz,
                      moons_mesh_t *out) {
  if (!g_stream) return fail(MOONS_ERR_STATE, "moons_init has not been called");
  if (!out || !hx || !hy || !hz) return fail(MOONS_ERR_ARG, "moons_mesh_create: null argument");
  if (sx < 5 || sy < 5 || sz < 5) return fail(MOONS_ERR_ARG, "moons_mesh_create: need sn >= 5 (N >= 2) per direction");
  const double *h[3] = {hx, hy, hz};
  const int sn[3] = {sx, sy, sz};
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i + 1 < sn[d]; ++i)
      if (!(h[d][i + 1] > h[d][i])) return fail(MOONS_ERR_ARG, "moons_mesh_create: node coordinates must increase");
  moons_mesh_s *m = new moons_mesh_s();
  size_t total = 0;
  for (int d = 0; d < 3; ++d) {
    coords_init(m->hc[d], h[d], sn[d]);
    const HostCoords &c = m->hc[d];
    total += c.hn.size() + c.hc.size() + c.dhn.size() + c.dhc.size() + c.cc2n_D.size() * 2 + c.n2cc_D.size() * 2 +
             c.theta_D.size() * 2 + 3 * (c.colCC_L[0].size() * 2 + c.colN_L[0].size() * 2) + 2 * c.lapA.size();
  }
  if (hipMalloc(&m->dbuf, total * sizeof(double)) != hipSuccess) {
    delete m;
    return fail(MOONS_ERR_HIP, "moons_mesh_create: hipMalloc failed");
  }
  std::vector<double> host(total);
  size_t off = 0;
  auto put = [&](const std::vector<double> &v) -> const double * {
    const double *p = m->dbuf + off;
    std::memcpy(host.data() + off, v.data(), v.size() * sizeof(double));
    off += v.size();
    return p;
  };
  for (int d = 0; d < 3; ++d) {
    HostCoords &c = m->hc[d];
    DirCoefDev &dc = m->dev.c[d];
    m->dev.sn[d] = c.sn;
    m->dev.sc[d] = c.sc;
    dc.hn = put(c.hn); dc.hc = put(c.hc); dc.dhn = put(c.dhn); dc.dhc = put(c.dhc);
    dc.cc2n_D = put(c.cc2n_D); dc.cc2n_U = put(c.cc2n_U);
    dc.n2cc_D = put(c.n2cc_D); dc.n2cc_U = put(c.n2cc_U);
    dc.theta_D = put(c.theta_D); dc.theta_U = put(c.theta_U);
    for (int n = 0; n < 2; ++n) {
      dc.colCC_L[n] = put(c.colCC_L[n]); dc.colCC_D[n] = put(c.colCC_D[n]); dc.colCC_U[n] = put(c.colCC_U[n]);
      dc.colN_L[n] = put(c.colN_L[n]); dc.colN_D[n] = put(c.colN_D[n]); dc.colN_U[n] = put(c.colN_U[n]);
    }
    dc.lapA = put(c.lapA); dc.lapB = put(c.lapB);
  }
  if (hipMemcpy(m->dbuf, host.data(), total * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(m->dbuf);
    delete m;
    return fail(MOONS_ERR_HIP, "moons_mesh_create: hipMemcpy failed");
  }
  *out = m;
  return MOONS_OK;
}

int moons_mesh_destroy(moons_mesh_t m) {
  if (!m) return MOONS_OK;
  if (g_stream) (void)hipStreamSynchronize(g_stream);
  if (m->dbuf) MOONS_HIP(hipFree(m->dbuf));
  delete m;
  return MOONS_OK;
}

int moons_mesh_coords(moons_mesh_t m, int dir, int which, double *out, int *n) {
  if (!m || dir < 0 || dir > 2 || which < 0 || which > 3) return fail(MOONS_ERR_ARG, "moons_mesh_coords: bad argument");
  const HostCoords &c = m->hc[dir];
  const std::vector<double> *v = which == 0 ? &c.hn : which == 1 ? &c.hc : which == 2 ? &c.dhn : &c.dhc;
  if (n) *n = (int)v->size();
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(double));
  return MOONS_OK;
}

int moons_field_create(moons_mesh_t m, int dl, moons_field_t *out) {
  if (!m || !out || dl < 0 || dl > 7) return fail(MOONS_ERR_ARG, "moons_field_create: bad argument");
  moons_field_s *f = new moons_field_s();
  f->m = m;
  f->dl = dl;
  for (int d = 0; d < 3; ++d) f->s[d] = ((dl >> d) & 1) ? m->hc[d].sn : m->hc[d].sc;
  f->n = (size_t)f->s[0] * f->s[1] * f->s[2];
  if (hipMalloc(&f->d, f->n * sizeof(double)) != hipSuccess) {
    delete f;
    return fail(MOONS_ERR_HIP, "moons_field_create: hipMalloc failed");
  }
  if (hipMemsetAsync(f->d, 0, f->n * sizeof(double), g_stream) != hipSuccess) {
    (void)hipFree(f->d);
    delete f;
    return fail(MOONS_ERR_HIP, "moons_field_create: hipMemsetAsync failed");
  }
  *out = f;
  return MOONS_OK;
}

int moons_field_destroy(moons_field_t f) {
  if (!f) return MOONS_OK;
  if (g_stream) (void)hipStreamSynchronize(g_stream);
  if (f->d) MOONS_HIP(hipFree(f->d));
  delete f;
  return MOONS_OK;
}

int moons_field_shape(moons_field_t f, int s[3]) {
  if (!f || !s) return fail(MOONS_ERR_ARG, "moons_field_shape: null argument");
  for (int d = 0; d < 3; ++d) s[d] = f->s[d];
  return MOONS_OK;
}

int moons_field_set_bcs(moons_field_t f, const int type[6], const double val[6], int all_neumann) {
  if (!f || !type) return fail(MOONS_ERR_ARG, "moons_field_set_bcs: null argument");
  for (int face = 0; face < 6; ++face) {
    if (type[face] < MOONS_BC_NONE || type[face] > MOONS_BC_ROBIN)
      return fail(MOONS_ERR_ARG, "moons_field_set_bcs: unknown BC type");
  }
  for (int d = 0; d < 3; ++d) {
    bool p0 = type[2 * d] == MOONS_BC_PERIODIC, p1 = type[2 * d + 1] == MOONS_BC_PERIODIC;
    if (p0 != p1) return fail(MOONS_ERR_ARG, "moons_field_set_bcs: periodic BCs must pair opposite faces");
  }
  f->has_bcs = true;
  f->all_neumann = all_neumann != 0;
  for (int face = 0; face < 6; ++face) {
    f->bct[face] = type[face];
    f->bval[face] = val ? val[face] : 0.0;
  }
  return MOONS_OK;
}

int moons_field_set_bc_props(moons_field_t f, const double c_w[6], const double robin_coeff[6]) {
  if (!f || !c_w) return fail(MOONS_ERR_ARG, "moons_field_set_bc_props: null argument");
  for (int face = 0; face < 6; ++face) {
    f->c_w[face] = c_w[face];
    f->robin_coeff[face] = robin_coeff ? robin_coeff[face] : 0.0;
  }
  return MOONS_OK;
}

int moons_field_copy_bcs(moons_field_t dst, moons_field_t src) {
  if (!dst || !src) return fail(MOONS_ERR_ARG, "moons_field_copy_bcs: null argument");
  if (dst->dl != src->dl) return fail(MOONS_ERR_ARG, "moons_field_copy_bcs: fields of different data locations");
  copy_bcs_of(dst, src);
  return MOONS_OK;
}

int moons_field_upload(moons_field_t f, const double *host, size_t n) {
  if (!f || !host || n != f->n) return fail(MOONS_ERR_ARG, "moons_field_upload: size mismatch");
  MOONS_HIP(hipMemcpyAsync(f->d, host, n * sizeof(double), hipMemcpyHostToDevice, g_stream));
  MOONS_HIP(hipStreamSynchronize(g_stream));
  return MOONS_OK;
}

int moons_field_download(moons_field_t f, double *host, size_t n) {
  if (!f || !host || n != f->n) return fail(MOONS_ERR_ARG, "moons_field_download: size mismatch");
  MOONS_HIP(hipMemcpyAsync(host, f->d, n * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  MOONS_HIP(hipStreamSynchronize(g_stream));
  return MOONS_OK;
}

int moons_field_fill(moons_field_t f, double val) {
  if (!f) return fail(MOONS_ERR_ARG, "moons_field_fill: null argument");
  hipLaunchKernelGGL(k_fill, dim3(nblk(f->n)), dim3(256), 0, g_stream, f->d, f->n, val);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int moons_field_copy(moons_field_t dst, moons_field_t src) {
  if (!dst || !src || dst->n != src->n || dst->dl != src->dl) return fail(MOONS_ERR_ARG, "moons_field_copy: shape mismatch");
  MOONS_HIP(hipMemcpyAsync(dst->d, src->d, dst->n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
  return MOONS_OK;
}

int moons_field_device_ptr(moons_field_t f, void **dptr) {
  if (!f || !dptr) return fail(MOONS_ERR_ARG, "moons_field_device_ptr: null argument");
  *dptr = f->d;
  return MOONS_OK;
}

int moons_apply_bcs(moons_field_t u) {
  if (!u) return fail(MOONS_ERR_ARG, "moons_apply_bcs: null argument");
  return launch_apply_bcs(u, false);
}

int moons_apply_bcs_implicit(moons_field_t u) {
  if (!u) return fail(MOONS_ERR_ARG, "moons_apply_bcs_implicit: null argument");
  return launch_apply_bcs(u, true);
}

int moons_del(moons_field_t dfdh, moons_field_t f, int n, int dir, int pad, int gen) {
  if (!dfdh || !f) return fail(MOONS_ERR_ARG, "moons_del: null argument");
  return launch_del(dfdh, f, n, dir, pad, gen);
}

int moons_grad(moons_field_t gx, moons_field_t gy, moons_field_t gz, moons_field_t u) {
  moons_field_t g[3] = {gx, gy, gz};
  if (!u) return fail(MOONS_ERR_ARG, "moons_grad: null argument");
  for (int d = 0; d < 3; ++d) {
    if (!g[d]) continue;
    if (g[d]->dl != (u->dl ^ (1 << d))) return fail(MOONS_ERR_ARG, "moons_grad: output location mismatch");
    int st = launch_del(g[d], u, 1, d, 0, 1);
    if (st) return st;
  }
  return MOONS_OK;
}

int moons_div(moons_field_t divU, moons_field_t u, moons_field_t v, moons_field_t w) {
  if (!divU || !u || !v || !w) return fail(MOONS_ERR_ARG, "moons_div: null argument");
  if (u->dl != (divU->dl ^ 1) || v->dl != (divU->dl ^ 2) || w->dl != (divU->dl ^ 4))
    return fail(MOONS_ERR_ARG, "moons_div: component locations do not match the output");
  int st;
  if ((st = launch_del(divU, u, 1, 0, 0, 1))) return st;
  if ((st = launch_del(divU, v, 1, 1, 0, 2))) return st;
  if ((st = launch_del(divU, w, 1, 2, 0, 2))) return st;
  return launch_assign_ghost_xperiodic(divU, 0.0);
}

int moons_curl(moons_field_t curlU, moons_field_t u, moons_field_t v, moons_field_t w, int dir) {
  if (!curlU || !u || !v || !w || dir < 0 || dir > 2) return fail(MOONS_ERR_ARG, "moons_curl: bad argument");
  int st;
  switch (dir) {
    case 0:
      if ((st = launch_del(curlU, w, 1, 1, 0, 1))) return st;
      return launch_del(curlU, v, 1, 2, 0, 3);
    case 1:
      if ((st = launch_del(curlU, u, 1, 2, 0, 1))) return st;
      return launch_del(curlU, w, 1, 0, 0, 3);
    default:
      if ((st = launch_del(curlU, v, 1, 0, 0, 1))) return st;
      return launch_del(curlU, u, 1, 1, 0, 3);
  }
}

int moons_interp(moons_field_t out, moons_field_t in, int dir) {
  if (!out || !in) return fail(MOONS_ERR_ARG, "moons_interp: null argument");
  return launch_interp(out, in, dir);
}

// face2Edge_TF_no_diag (ops_interp.f90:531-544) + cross_product_TF
// (TF_extend.f90:385-400).  tmp: the six face-to-edge interpolants
// (B_x on y,z-edges, B_y on x,z-edges, B_z on x,y-edges).
int moons_edge_cross_face_E(moons_field_t E[3], moons_field_t U_E[6], moons_field_t B[3], moons_field_t tmp[6]) {
  if (!E || !U_E || !B || !tmp) return fail(MOONS_ERR_ARG, "moons_edge_cross_face_E: null argument");
  for (int q = 0; q < 6; ++q)
    if (!U_E[q] || !tmp[q] || (q < 3 && (!E[q] || !B[q]))) return fail(MOONS_ERR_ARG, "moons_edge_cross_face_E: null component");
  const int edge_dl[3] = {MOONS_DL_EDGE_X, MOONS_DL_EDGE_Y, MOONS_DL_EDGE_Z};
  // tmp layout: 0 Bx@y, 1 Bx@z, 2 By@x, 3 By@z, 4 Bz@x, 5 Bz@y ; U_E: Uy@x, Uz@x, Ux@y, Uz@y, Ux@z, Uy@z
  const int tmp_face[6] = {0, 0, 1, 1, 2, 2}, tmp_edge[6] = {1, 2, 0, 2, 0, 1};
  const int ue_edge[6] = {0, 0, 1, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    if (tmp[q]->dl != edge_dl[tmp_edge[q]] || U_E[q]->dl != edge_dl[ue_edge[q]])
      return fail(MOONS_ERR_ARG, "moons_edge_cross_face_E: component locations do not match");
    if (B[tmp_face[q]]->dl != (1 << tmp_face[q])) return fail(MOONS_ERR_ARG, "moons_edge_cross_face_E: B must be on faces");
  }
  for (int d = 0; d < 3; ++d)
    if (E[d]->dl != edge_dl[d]) return fail(MOONS_ERR_ARG, "moons_edge_cross_face_E: E must be on edges");
  int st;
  for (int q = 0; q < 6; ++q) {
    const int orth = 3 - tmp_face[q] - tmp_edge[q];
    if ((st = launch_interp(tmp[q], B[tmp_face[q]], orth))) return st;
  }
  // x: Uy@x * Bz@x - Uz@x * By@x ; y: -(Ux@y * Bz@y - Uz@y * Bx@y) ; z: Ux@z * By@z - Uy@z * Bx@z
  if ((st = launch_cross(E[0], U_E[0], tmp[4], U_E[1], tmp[2], 1))) return st;
  if ((st = launch_cross(E[1], U_E[2], tmp[5], U_E[3], tmp[0], -1))) return st;
  return launch_cross(E[2], U_E[4], tmp[3], U_E[5], tmp[1], 1);
}

// advect_B (ops_advect.f90:186-200): adv = curl(U_E x face2Edge(B)).
int moons_advect_B(moons_field_t adv[3], moons_field_t U_E[6], moons_field_t B[3], moons_field_t temp_E_TF[6],
                   moons_field_t temp_E[3]) {
  int st = moons_edge_cross_face_E(temp_E, U_E, B, temp_E_TF);
  if (st) return st;
  for (int d = 0; d < 3; ++d)
    if ((st = moons_curl(adv[d], temp_E[0], temp_E[1], temp_E[2], d))) return st;
  return MOONS_OK;
}

int moons_volume(moons_field_t vol) {
  if (!vol) return fail(MOONS_ERR_ARG, "moons_volume: null argument");
  Shape3 S{{vol->s[0], vol->s[1], vol->s[2]}};
  const MeshDev &M = vol->m->dev;
  int e[3];
  const double *dh[3];
  for (int d = 0; d < 3; ++d) {
    e[d] = (vol->dl >> d) & 1;
    dh[d] = e[d] ? M.c[d].dhc : M.c[d].dhn;
  }
  hipLaunchKernelGGL(k_volume, dim3(nblk(vol->n)), dim3(256), 0, g_stream, vol->d, S, e[0], e[1], e[2], dh[0], dh[1],
                     dh[2]);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

} // extern "C"
