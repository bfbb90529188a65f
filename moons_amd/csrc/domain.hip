// domain.hip -- multi-domain meshes of the induction solve: the uniform mesh
// extension, the overlap of an interior mesh inside a total mesh, embedding
// and extraction of fields between the two, and the conductivity of the
// total mesh.
//
//   ext_uniform_IO   code/handwritten/grid/grid_continue.f90:41-47,94-107,197-222
//   ext_Roberts_*, ext_prep/app_*
//                    code/handwritten/grid/grid_continue.f90:49-77,107-193,224-306 with the
//                    stretch-parameter matching of coordinate_stretch_param_match.f90:47-95,176-233
//                    and the Roberts transformations of coordinate_distribution_funcs.f90:145-248
//   overlap          get_C_overlap_OL / get_N_overlap_OL with the physical
//                    padding p = 1 and tol 1e-12,
//                    code/handwritten/sub_domain/overlap_extend.f90:204-278,364-391,
//                    code/handwritten/physical_domain/physical_sub_domain_extend.f90:36-45
//   embed / extract  embed_GF / extract_GF over EE_shape's ranges,
//                    code/handwritten/grid_field/GF_embed_extract/GF_embed_extract.f90:14-60,
//                    code/handwritten/ops/ops_embedExtract.f90:280-320
//   sigma_inv        set_sigma_inv_VF, code/handwritten/solvers/induction/induction_aux.f90:178-194
#include "common.hpp"

#include <cmath>
#include <vector>

namespace moons {

constexpr int kDT = 256;

// dst box (dlo + t) <- src box (slo + t) for t in [0, ext)
__global__ void __launch_bounds__(kDT) k_subcopy(double *__restrict__ dst, int dsx, int dsy,
                                                const double *__restrict__ src, int ssx, int ssy, int3 dlo, int3 slo,
                                                int3 ext) {
  const size_t n = (size_t)ext.x * ext.y * ext.z;
  const size_t t = (size_t)blockIdx.x * kDT + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % ext.x), j = (int)((t / ext.x) % ext.y), k = (int)(t / ((size_t)ext.x * ext.y));
  dst[(size_t)(dlo.x + i) + (size_t)dsx * ((size_t)(dlo.y + j) + (size_t)dsy * (size_t)(dlo.z + k))] =
      src[(size_t)(slo.x + i) + (size_t)ssx * ((size_t)(slo.y + j) + (size_t)ssy * (size_t)(slo.z + k))];
}

__global__ void __launch_bounds__(kDT) k_subfill(double *__restrict__ dst, int dsx, int dsy, int3 lo, int3 ext,
                                                double v) {
  const size_t n = (size_t)ext.x * ext.y * ext.z;
  const size_t t = (size_t)blockIdx.x * kDT + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % ext.x), j = (int)((t / ext.x) % ext.y), k = (int)(t / ((size_t)ext.x * ext.y));
  dst[(size_t)(lo.x + i) + (size_t)dsx * ((size_t)(lo.y + j) + (size_t)dsy * (size_t)(lo.z + k))] = v;
}

// Ranges of a location dl: the N overlap along node-centred directions, C otherwise.
static void box_of(const moons_domain_s *d, int dl, int3 &lo_i, int3 &lo_t, int3 &ext) {
  int li[3], lt[3], e[3];
  for (int q = 0; q < 3; ++q) {
    const int *r = ((dl >> q) & 1) ? d->N[q] : d->C[q];
    li[q] = r[0];
    lt[q] = r[2];
    e[q] = r[1] - r[0] + 1;
  }
  lo_i = make_int3(li[0], li[1], li[2]);
  lo_t = make_int3(lt[0], lt[1], lt[2]);
  ext = make_int3(e[0], e[1], e[2]);
}

static int check_pair(const moons_field_s *tot, const moons_field_s *in, const moons_domain_s *d, const char *who) {
  if (!tot || !in || !d) return fail(MOONS_ERR_ARG, std::string(who) + ": null argument");
  if (tot->dl != in->dl) return fail(MOONS_ERR_ARG, std::string(who) + ": fields of different locations");
  if (tot->m != d->mt || in->m != d->mi)
    return fail(MOONS_ERR_ARG, std::string(who) + ": fields are not on the domain's total and interior meshes");
  return MOONS_OK;
}

static inline unsigned dblk(const int3 &e) { return (unsigned)(((size_t)e.x * e.y * e.z + kDT - 1) / kDT); }

int launch_embed(moons_field_s *total, const moons_field_s *interior, const moons_domain_s *d) {
  if (int st = check_pair(total, interior, d, "embed")) return st;
  int3 li, lt, e;
  box_of(d, total->dl, li, lt, e);
  hipLaunchKernelGGL(k_subcopy, dim3(dblk(e)), dim3(kDT), 0, stream(), total->d, total->s[0], total->s[1],
                     interior->d, interior->s[0], interior->s[1], lt, li, e);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

int launch_extract(moons_field_s *interior, const moons_field_s *total, const moons_domain_s *d) {
  if (int st = check_pair(total, interior, d, "extract")) return st;
  int3 li, lt, e;
  box_of(d, total->dl, li, lt, e);
  hipLaunchKernelGGL(k_subcopy, dim3(dblk(e)), dim3(kDT), 0, stream(), interior->d, interior->s[0], interior->s[1],
                     total->d, total->s[0], total->s[1], li, lt, e);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

// get_C_overlap_OL / get_N_overlap_OL: points of a (interior) and b (total)
// equal within tol, p points left out at each end of both; false when none.
static bool overlap(const std::vector<double> &a, const std::vector<double> &b, int r[4]) {
  const double tol = 1e-12;
  const int na = (int)a.size(), nb = (int)b.size();
  const int pa = na - 1 > 0 ? 1 : 0, pb = nb - 1 > 0 ? 1 : 0;
  r[0] = r[2] = 1 << 30;
  r[1] = r[3] = -1;
  for (int i = pa; i < na - pa; ++i)
    for (int j = pb; j < nb - pb; ++j)
      if (std::fabs(b[j] - a[i]) < tol) {
        r[0] = std::min(r[0], i);
        r[1] = std::max(r[1], i);
        r[2] = std::min(r[2], j);
        r[3] = std::max(r[3], j);
      }
  return r[1] >= 0;
}

// ------------------------------------------------------------------
// Roberts mesh extensions (host code).  Scalar libm arithmetic in the
// reference's evaluation order; x**2 is x*x.
// ------------------------------------------------------------------
namespace ext {

// T2_h_i: Roberts transformation 2 at node i of N, stretching alpha, beta
double t2_h(double b, double a, int i, int N) {
  const double g = (b + 1.0) / (b - 1.0);
  const double t = ((double)i / (double)N - a) / (1.0 - a);
  const double gt = std::pow(g, t);
  return ((b + 2.0 * a) * gt - b + 2.0 * a) / ((2.0 * a + 1.0) * (1.0 + gt));
}

// d T2_h_i / d beta
double t2_prime(double b, double a, int i, int N) {
  const double g = (b + 1.0) / (b - 1.0);
  const double t = ((double)i / (double)N - a) / (1.0 - a);
  const double gt = std::pow(g, t), g2t = std::pow(g, 2.0 * t);
  const double num = -(b * b) * g2t + b * b + 4.0 * b * gt * t + g2t - 1.0;
  return -num / ((2.0 * a + 1.0) * (b - 1.0) * (b + 1.0) * ((gt + 1.0) * (gt + 1.0)));
}

// Newton for the beta whose end cell (near hmax, or near hmin) has width dh:
// newtonT2 (tol 10 eps, giving up above 1000 or below 1 + 1e-10)
bool match_beta(bool near_hmax, double a, double hmin, double hmax, int N, double dh, double *beta_out) {
  const double tol = 10.0 * 2.220446049250313e-16;
  const double beta_low = 1.0 + std::pow(10.0, -10.0);
  double beta = 1.0 + std::pow(10.0, -10.0), fb = 0.0;
  const int i1 = near_hmax ? N : 1, i0 = near_hmax ? N - 1 : 0;
  for (int k = 0; k < 1000000; ++k) {
    fb = t2_h(beta, a, i1, N) - t2_h(beta, a, i0, N) - dh / (hmax - hmin);
    const double fp = t2_prime(beta, a, i1, N) - t2_prime(beta, a, i0, N);
    if (std::fabs(fb) < tol) break;
    if (beta > 1000.0) break;
    if (beta < beta_low) {
      beta = 1.1;
      break;
    }
    beta = beta - fb / fp;
  }
  if (std::isnan(beta) || std::fabs(fb) > tol) return false;
  *beta_out = beta;
  return true;
}

std::vector<double> unit_nodes(double hmin, double hmax, int N) {
  std::vector<double> x(N + 1);
  const double dh = (hmax - hmin) / (double)N;
  for (int i = 0; i <= N; ++i) x[i] = (hmin + (double)i * dh - hmin) / (hmax - hmin);
  return x;
}

// every spacing ratio >= 1 (increasing) or <= 1 (decreasing)
bool ratios_ok(const std::vector<double> &h, bool increasing) {
  for (size_t i = 0; i + 2 < h.size(); ++i) {
    const double R = (h[i + 2] - h[i + 1]) / (h[i + 1] - h[i]);
    if ((increasing && R < 1.0) || (!increasing && R > 1.0)) return false;
  }
  return true;
}

// robertsLeft (transformation 1), robertsRight / robertsBoth (transformation 2, alpha 0 / 1/2)
bool roberts(char kind, double hmin, double hmax, int N, double beta, std::vector<double> &out) {
  const std::vector<double> x = unit_nodes(hmin, hmax, N);
  std::vector<double> h(N + 1);
  const double b = beta, g = (b + 1.0) / (b - 1.0);
  if (kind == 'L') {
    const double a = 1.0;
    for (int i = 0; i <= N; ++i) {
      const double e = std::pow(g, a - x[i]);
      h[i] = ((b + a) - (b - a) * e) / (e + a);
    }
    if (!ratios_ok(h, true)) return false;
  } else {
    const double a = kind == 'R' ? 0.0 : 0.5;
    for (int i = 0; i <= N; ++i) {
      const double e = std::pow(g, (x[i] - a) / (1.0 - a));
      h[i] = ((b + 2.0 * a) * e - b + 2.0 * a) / ((2.0 * a + 1.0) * (1.0 + e));
    }
    if (std::fabs(a) < std::pow(10.0, -10.0) && !ratios_ok(h, false)) return false;
  }
  out.resize(N + 1);
  for (int i = 0; i <= N; ++i) out[i] = hmin + (hmax - hmin) * h[i];
  return true;
}

// needs_more_points: some neighbouring spacing ratio above max_ratio
bool needs_more_points(const std::vector<double> &b, double max_ratio) {
  if (b.size() <= 2) return false;
  double R = 0.0;
  for (size_t i = 0; i + 2 < b.size(); ++i) {
    const double d0 = b[i + 1] - b[i], d1 = b[i + 2] - b[i + 1];
    R = std::max(R, std::max(d0 / d1, d1 / d0));
  }
  return R > max_ratio;
}

// One Roberts segment on [lo, hi]: the N_iter loop of ext_prep/app_Roberts_*
// (N_final = the points of the last segment, one more when it fails the
// quality check).  match: 's' dh near hmax alpha 0, 'g' near hmin alpha 0,
// 'b' near hmax alpha 1/2.
int segment(char match, char dist, double lo, double hi, int N, double dh, int n_iter, double max_ratio,
            std::vector<double> &b) {
  int N_final = N;
  for (int it = 0; it < n_iter; ++it) {
    double beta;
    if (!match_beta(match != 'g', match == 'b' ? 0.5 : 0.0, lo, hi, N_final, dh, &beta))
      return fail(MOONS_ERR_ARG, "moons_nodes_ext: the stretch parameter matching did not converge");
    if (!roberts(dist, lo, hi, N_final, beta, b))
      return fail(MOONS_ERR_ARG, "moons_nodes_ext: the Roberts segment fails its spacing-ratio check");
    N_final = (int)b.size();
    if (needs_more_points(b, max_ratio)) N_final = N_final + 1;
    else break;
  }
  return MOONS_OK;
}

// prepend (kind 'U' uniform, 'L', 'R', 'B', 'C' C2F) of N cells, length L
int prep(std::vector<double> &h, char kind, double L, int N, int n_iter, double max_ratio) {
  const double hmin = h[1], dh = h[1] - h[0];
  std::vector<double> b;
  if (kind == 'U') {
    for (int i = N; i >= 0; --i) b.push_back(hmin + (-1.0 * (double)i) * dh);
  } else {
    const char match = kind == 'B' ? 'b' : (kind == 'C' ? 'g' : 's');
    const char dist = kind == 'C' ? 'L' : kind;
    if (int st = segment(match, dist, hmin - L, hmin, N, dh, n_iter, max_ratio, b)) return st;
  }
  std::vector<double> a(b);
  a.insert(a.end(), h.begin() + 2, h.end());
  a.insert(a.begin(), a[0] - (a[1] - a[0])); // prepend_ghost_c
  h.swap(a);
  return MOONS_OK;
}

// append (kinds as prep) of N cells, length L
int app(std::vector<double> &h, char kind, double L, int N, int n_iter, double max_ratio) {
  const size_t n = h.size();
  const double hmax = h[n - 2], dh = h[n - 1] - h[n - 2];
  std::vector<double> b;
  if (kind == 'U') {
    for (int i = 0; i <= N; ++i) b.push_back(hmax + (1.0 * (double)i) * dh);
  } else {
    const char match = kind == 'B' ? 'b' : (kind == 'C' ? 'g' : 's');
    const char dist = kind == 'C' ? 'R' : kind;
    if (int st = segment(match, dist, hmax, hmax + L, N, dh, n_iter, max_ratio, b)) return st;
  }
  h.resize(n - 2);
  h.insert(h.end(), b.begin(), b.end());
  h.push_back(h.back() + (h.back() - h[h.size() - 2])); // append_ghost_c
  return MOONS_OK;
}

} // namespace ext

} // namespace moons

using namespace moons;

extern "C" {

int moons_nodes_ext(const double *hn, int sn, int kind, double L, int N, int n_iter, double max_stretch_ratio,
                    double *out, int cap, int *n_out) {
  if (!hn || !out || !n_out || sn < 4 || N < 1 || n_iter < 1)
    return fail(MOONS_ERR_ARG, "moons_nodes_ext: need sn >= 4, N >= 1 and n_iter >= 1");
  if (kind < MOONS_EXT_UNIFORM || kind > MOONS_EXT_APP_ROBERTS_C2F)
    return fail(MOONS_ERR_ARG, "moons_nodes_ext: unknown kind");
  if (kind != MOONS_EXT_UNIFORM && kind != MOONS_EXT_PREP_UNIFORM && kind != MOONS_EXT_APP_UNIFORM && !(L > 0.0))
    return fail(MOONS_ERR_ARG, "moons_nodes_ext: a Roberts extension needs L > 0");
  std::vector<double> h(hn, hn + sn);
  // prepend kind, append kind ('-': none)
  static const char P[14] = {'U', 'R', 'L', 'B', 'U', 'L', 'R', 'B', 'C', '-', '-', '-', '-', '-'};
  static const char A[14] = {'U', 'L', 'R', 'B', '-', '-', '-', '-', '-', 'U', 'L', 'R', 'B', 'C'};
  int st;
  if (P[kind] != '-' && (st = ext::prep(h, P[kind], L, N, n_iter, max_stretch_ratio))) return st;
  if (A[kind] != '-' && (st = ext::app(h, A[kind], L, N, n_iter, max_stretch_ratio))) return st;
  *n_out = (int)h.size();
  if ((int)h.size() > cap) return fail(MOONS_ERR_ARG, "moons_nodes_ext: out has room for fewer nodes than the result");
  for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
  return MOONS_OK;
}


int moons_nodes_ext_uniform(const double *hn, int sn, int N, double *out) {
  if (!hn || !out || sn < 4 || N < 1) return fail(MOONS_ERR_ARG, "moons_nodes_ext_uniform: need sn >= 4 and N >= 1");
  std::vector<double> a;
  // prepend (ext_prep_uniform_IO): uniformLeft(hmin, dhn(1), N) replaces the ghost and hmin
  const double hmin = hn[1], dh = hn[1] - hn[0];
  for (int i = N; i >= 0; --i) a.push_back(hmin + (-1.0 * (double)i) * dh);
  for (int i = 2; i < sn; ++i) a.push_back(hn[i]);
  a.insert(a.begin(), a[0] - (a[1] - a[0])); // prepend_ghost_c
  // append (ext_app_uniform_IO): uniformRight(hmax, dhn_e, N) replaces hmax and the ghost
  const size_t n = a.size();
  const double hmax = a[n - 2], dhe = a[n - 1] - a[n - 2];
  a.resize(n - 2);
  for (int i = 0; i <= N; ++i) a.push_back(hmax + (1.0 * (double)i) * dhe);
  a.push_back(a.back() + (a.back() - a[a.size() - 2])); // append_ghost_c
  for (size_t i = 0; i < a.size(); ++i) out[i] = a[i];
  return MOONS_OK;
}

int moons_domain_create(moons_mesh_t interior, moons_mesh_t total, moons_domain_t *out) {
  if (!interior || !total || !out) return fail(MOONS_ERR_ARG, "moons_domain_create: null argument");
  moons_domain_s *D = new moons_domain_s();
  D->mi = interior;
  D->mt = total;
  for (int q = 0; q < 3; ++q) {
    const HostCoords &a = interior->hc[q], &b = total->hc[q];
    const bool okC = overlap(a.hc, b.hc, D->C[q]);
    const bool okN = overlap(a.hn, b.hn, D->N[q]);
    const bool same = okC && okN && D->C[q][1] - D->C[q][0] == D->C[q][3] - D->C[q][2] &&
                      D->N[q][1] - D->N[q][0] == D->N[q][3] - D->N[q][2];
    if (!same) {
      delete D;
      return fail(MOONS_ERR_ARG, "moons_domain_create: the meshes do not overlap in a common range of cells "
                                 "and nodes along direction " + std::to_string(q + 1));
    }
  }
  *out = D;
  return MOONS_OK;
}

int moons_domain_destroy(moons_domain_t d) {
  delete d;
  return MOONS_OK;
}

int moons_domain_overlap(moons_domain_t d, int dir, int node, int r[4]) {
  if (!d || !r || dir < 0 || dir > 2) return fail(MOONS_ERR_ARG, "moons_domain_overlap: bad argument");
  for (int q = 0; q < 4; ++q) r[q] = node ? d->N[dir][q] : d->C[dir][q];
  return MOONS_OK;
}

int moons_embed(moons_field_t total, moons_field_t interior, moons_domain_t d) {
  return launch_embed(total, interior, d);
}

int moons_extract(moons_field_t interior, moons_field_t total, moons_domain_t d) {
  return launch_extract(interior, total, d);
}

int moons_set_sigma_inv(moons_field_t sig, moons_domain_t d, double sig_local_over_sig_f) {
  if (!sig || !d) return fail(MOONS_ERR_ARG, "moons_set_sigma_inv: null argument");
  if (sig->m != d->mt) return fail(MOONS_ERR_ARG, "moons_set_sigma_inv: field is not on the domain's total mesh");
  if (!(sig_local_over_sig_f != 0.0)) return fail(MOONS_ERR_ARG, "moons_set_sigma_inv: ratio must be non-zero");
  if (int st = moons_field_fill(sig, 1.0 / sig_local_over_sig_f)) return st;
  int3 li, lt, e;
  box_of(d, sig->dl, li, lt, e);
  hipLaunchKernelGGL(k_subfill, dim3(dblk(e)), dim3(kDT), 0, stream(), sig->d, sig->s[0], sig->s[1], lt, e, 1.0);
  MOONS_LAUNCH_CHECK();
  return MOONS_OK;
}

} // extern "C"
