/*
 * moons_oracle.c -- CPU FP64 restatement of the MOONS per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 *
 * Every function below cites the reference file:line it restates (paths are
 * relative to the reference root, code/handwritten abbreviated as hw/).  The
 * arithmetic follows the reference's expression order term by term, so that
 * element-wise results of a GPU build compiled without FP contraction agree
 * bit for bit; only global reductions are summed in a different order.
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   grids (uniform, Roberts)       closed forms + compiled reference
 *                                   coordinate_distribution_funcs.f90 (oracle/_ref)
 *   stencil coefficients           closed forms (+-1/dh, theta = 1/2 on uniform grids,
 *                                   exactness on quadratics) + compiled reference
 *                                   derivative_stencils.f90 / interpolation_stencils.f90
 *   pencil kernels (del)           compiled reference stencils.f90 (oracle/_ref)
 *   BC ghost rules                 compiled reference apply_BCs_faces_raw*.f90 (oracle/_ref)
 *   grad/div/curl/lap_centered     manufactured polynomials, div(curl)=0, curl(grad)=0,
 *                                   symmetry of vol*A, dense assembly by unit vectors
 *   Lap diagonal / Jacobi Minv     dense-matrix diagonal away from wall rows
 *   PCG (PPE)                      dense numpy solve of the assembled system on
 *                                   4^3-8^3; residual-drop to 1e-12; textbook PCG
 *                                   recurrence on the assembled matrix
 *   interp, face2Edge, E = U x B,  exactness on linear fields (theta weights on
 *   advect_B                        stretched grids); sign and pairing of the cross
 *                                   product on axis-aligned fields; div(curl E) = 0
 *   VF operators, VF PCG           symmetry and positivity of vol*A with the wall
 *                                   factors; Jacobi diagonal = dense diagonal off the
 *                                   walls; PCG = dense solve; lid BC through the RHS
 */
#include "moons_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the OpenMP loops (the CPU timing baseline).  With 1 (the
   default) every loop runs sequentially.  With more, the stencil pencils and
   element-wise loops split over threads with unchanged results, and the
   padded sums add per-plane partial sums in plane order, which changes their
   last bits (used for timing only). */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 1 ? n : 1; }

#define IDX(s, i, j, k) ((size_t)(i) + (size_t)(s)[0] * ((size_t)(j) + (size_t)(s)[1] * (size_t)(k)))

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ------------------------------------------------------------------------- */
/* 1D grids                                                                  */
/* ------------------------------------------------------------------------- */

/* Ghost nodes mirror the first/last spacing: hw/grid/coordinates_extend.f90:529-538 */
static void add_ghost_nodes(const double *hn_phys, int np, double *out) {
  double dh0 = hn_phys[1] - hn_phys[0];
  double dhe = hn_phys[np - 1] - hn_phys[np - 2];
  out[0] = hn_phys[0] - dh0;
  memcpy(out + 1, hn_phys, sizeof(double) * (size_t)np);
  out[np + 1] = hn_phys[np - 1] + dhe;
}

/* hw/grid/coordinate_distribution_funcs.f90:34-61 (uniform_func) then ghost nodes */
int orc_uniform_nodes(double hmin, double hmax, int N, double *hn_out) {
  if (N < 1) return -1;
  double *h = dalloc((size_t)N + 1);
  double dh = (hmax - hmin) / (double)N;
  for (int i = 0; i <= N; ++i) h[i] = hmin + (double)i * dh;
  add_ghost_nodes(h, N + 1, hn_out);
  free(h);
  return 0;
}

/* Tannehill transformation 2: hw/grid/coordinate_distribution_funcs.f90:192-248 */
static void transformation2(double hmin, double hmax, int N, double alpha, double beta, double *hn) {
  const double one = 1.0, two = 2.0;
  double a = alpha, b = beta;
  double g = (b + one) / (b - one);
  double dh = (hmax - hmin) / (double)N;
  for (int i = 0; i <= N; ++i) {
    double hb = hmin + (double)i * dh;
    hb = (hb - hmin) / (hmax - hmin);
    double e = pow(g, (hb - a) / (one - a));
    double v = ((b + two * a) * e - b + two * a) / ((two * a + one) * (one + e));
    hn[i] = v;
  }
  for (int i = 0; i <= N; ++i) hn[i] = hmin + (hmax - hmin) * hn[i];
}

/* Tannehill transformation 1: hw/grid/coordinate_distribution_funcs.f90:143-177 */
static void transformation1(double hmin, double hmax, int N, double beta, double *hn) {
  double a = 1.0, b = beta;
  double g = (b + 1.0) / (b - 1.0);
  double dh = (hmax - hmin) / (double)N;
  for (int i = 0; i <= N; ++i) {
    double hb = hmin + (double)i * dh;
    hb = (hb - hmin) / (hmax - hmin);
    double e = pow(g, a - hb);
    hn[i] = ((b + a) - (b - a) * e) / (e + a);
  }
  for (int i = 0; i <= N; ++i) hn[i] = hmin + (hmax - hmin) * hn[i];
}

/* robertsBoth_func (alpha = 1/2): hw/grid/coordinate_distribution_funcs.f90:325-331;
   grid_Roberts_B adds ghost nodes: hw/grid/grid_init.f90:60-68 */
int orc_roberts_both_nodes(double hmin, double hmax, int N, double beta, double *hn_out) {
  if (N < 1 || !(beta > 1.0)) return -1;
  double *h = dalloc((size_t)N + 1);
  transformation2(hmin, hmax, N, 0.5, beta, h);
  add_ghost_nodes(h, N + 1, hn_out);
  free(h);
  return 0;
}
int orc_roberts_right_nodes(double hmin, double hmax, int N, double beta, double *hn_out) {
  if (N < 1 || !(beta > 1.0)) return -1;
  double *h = dalloc((size_t)N + 1);
  transformation2(hmin, hmax, N, 0.0, beta, h);
  add_ghost_nodes(h, N + 1, hn_out);
  free(h);
  return 0;
}
int orc_roberts_left_nodes(double hmin, double hmax, int N, double beta, double *hn_out) {
  if (N < 1 || !(beta > 1.0)) return -1;
  double *h = dalloc((size_t)N + 1);
  transformation1(hmin, hmax, N, beta, h);
  add_ghost_nodes(h, N + 1, hn_out);
  free(h);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* coordinates: hw/grid/coordinates_extend.f90:78-173 (init_c), :506-527      */
/* ------------------------------------------------------------------------- */

/* collocated first derivative on nodes: hw/grid/derivative_stencils.f90:79-107 */
static void col_node_1(const double *dh, int s, double *L, double *D, double *U) {
#define DH(i) dh[(i)-1]
#define SL(k) L[(k)-1]
#define SD(k) D[(k)-1]
#define SU(k) U[(k)-1]
  int i;
  if (s > 3) {
    i = 2;
    SL(1) = -(DH(i + 1) + 2.0 * DH(i)) / (DH(i) * (DH(i + 1) + DH(i)));
    SD(1) = (DH(i + 1) + DH(i)) / (DH(i + 1) * DH(i));
    SU(1) = -DH(i) / (DH(i + 1) * (DH(i + 1) + DH(i)));
    for (i = 3; i <= s - 2; ++i) {
      SL(i - 1) = -(DH(i) / (DH(i - 1) * (DH(i - 1) + DH(i))));
      SD(i - 1) = ((-DH(i - 1) + DH(i)) / (DH(i - 1) * DH(i)));
      SU(i - 1) = (DH(i - 1) / (DH(i) * (DH(i - 1) + DH(i))));
    }
    i = s - 1;
    SL(s - 2) = DH(i - 1) / (DH(i - 2) * (DH(i - 1) + DH(i - 2)));
    SD(s - 2) = -(DH(i - 1) + DH(i - 2)) / (DH(i - 1) * DH(i - 2));
    SU(s - 2) = (2.0 * DH(i - 1) + DH(i - 2)) / (DH(i - 1) * (DH(i - 1) + DH(i - 2)));
  }
}

/* collocated second derivative on nodes: hw/grid/derivative_stencils.f90:109-137 */
static void col_node_2(const double *dh, int s, double *L, double *D, double *U) {
  int i;
  if (s > 3) {
    i = 2;
    SL(1) = 2.0 / (DH(i) * (DH(i + 1) + DH(i)));
    SD(1) = -2.0 / (DH(i + 1) * DH(i));
    SU(1) = 2.0 / (DH(i + 1) * (DH(i + 1) + DH(i)));
    for (i = 3; i <= s - 2; ++i) {
      SL(i - 1) = 2.0 / (DH(i - 1) * (DH(i - 1) + DH(i)));
      SD(i - 1) = -2.0 / (DH(i - 1) * DH(i));
      SU(i - 1) = 2.0 / (DH(i) * (DH(i - 1) + DH(i)));
    }
    i = s - 1;
    SL(s - 2) = 2.0 / (DH(i - 2) * (DH(i - 1) + DH(i - 2)));
    SD(s - 2) = -2.0 / (DH(i - 1) * DH(i - 2));
    SU(s - 2) = 2.0 / (DH(i - 1) * (DH(i - 1) + DH(i - 2)));
  }
}

/* collocated first derivative on cell centres: hw/grid/derivative_stencils.f90:139-170 */
static void col_cc_1(const double *dh, int s, double *L, double *D, double *U) {
  int i;
  if (s > 2) {
    i = 2;
    SL(1) = (-(DH(i)) / ((0.5 * DH(i - 1)) * ((0.5 * DH(i - 1)) + DH(i))));
    SD(1) = ((-(0.5 * DH(i - 1)) + DH(i)) / ((0.5 * DH(i - 1)) * DH(i)));
    SU(1) = ((0.5 * DH(i - 1)) / (DH(i) * ((0.5 * DH(i - 1)) + DH(i))));
    for (i = 3; i <= s - 2; ++i) {
      SL(i - 1) = (-DH(i) / (DH(i - 1) * (DH(i - 1) + DH(i))));
      SD(i - 1) = ((-DH(i - 1) + DH(i)) / (DH(i - 1) * DH(i)));
      SU(i - 1) = (DH(i - 1) / (DH(i) * (DH(i - 1) + DH(i))));
    }
    i = s - 1;
    SL(s - 2) = (-0.5 * DH(i) / (DH(i - 1) * (DH(i - 1) + (0.5 * DH(i)))));
    SD(s - 2) = ((-DH(i - 1) + (0.5 * DH(i))) / (DH(i - 1) * (0.5 * DH(i))));
    SU(s - 2) = (DH(i - 1) / ((0.5 * DH(i)) * (DH(i - 1) + (0.5 * DH(i)))));
  }
}

/* collocated second derivative on cell centres: hw/grid/derivative_stencils.f90:172-208 */
static void col_cc_2(const double *dh, int s, double *L, double *D, double *U) {
  int i;
  if (s > 2) {
    i = 2;
    SL(1) = 2.0 / ((0.5 * DH(i - 1)) * ((0.5 * DH(i - 1)) + DH(i)));
    SD(1) = -2.0 / ((0.5 * DH(i - 1)) * DH(i));
    SU(1) = 2.0 / (DH(i) * ((0.5 * DH(i - 1)) + DH(i)));
    for (i = 3; i <= s - 2; ++i) {
      SL(i - 1) = 2.0 / (DH(i - 1) * (DH(i - 1) + DH(i)));
      SD(i - 1) = -(2.0 / (DH(i - 1) * DH(i)));
      SU(i - 1) = 2.0 / (DH(i) * (DH(i - 1) + DH(i)));
    }
    i = s - 1;
    SL(s - 2) = 2.0 / (DH(i - 1) * (DH(i - 1) + (0.5 * DH(i))));
    SD(s - 2) = -2.0 / (DH(i - 1) * (0.5 * DH(i)));
    SU(s - 2) = 2.0 / ((0.5 * DH(i)) * (DH(i - 1) + (0.5 * DH(i))));
  }
#undef DH
#undef SL
#undef SD
#undef SU
}

static void coords_init(orc_coords *c, const double *hn, int sn) {
  memset(c, 0, sizeof(*c));
  c->sn = sn;
  c->sc = sn - 1;
  int sc = c->sc;
  c->hn = dalloc((size_t)sn);
  c->hc = dalloc((size_t)sc);
  c->dhn = dalloc((size_t)sn - 1);
  c->dhc = dalloc((size_t)sc - 1);
  memcpy(c->hn, hn, sizeof(double) * (size_t)sn);
  for (int i = 0; i < sn - 1; ++i) c->dhn[i] = hn[i + 1] - hn[i];
  for (int i = 0; i < sc; ++i) c->hc[i] = (hn[i + 1] + hn[i]) / 2.0;
  for (int i = 0; i < sc - 1; ++i) c->dhc[i] = c->hc[i + 1] - c->hc[i];
  /* interpolation_stencil: hw/grid/interpolation_stencils.f90:12-35 */
  c->theta_D = dalloc((size_t)sc - 1);
  c->theta_U = dalloc((size_t)sc - 1);
  for (int i = 0; i < sc - 1; ++i) {
    c->theta_D[i] = (c->hn[i + 1] - c->hc[i]) / (c->hc[i + 1] - c->hc[i]);
    c->theta_U[i] = 1.0 - c->theta_D[i];
  }
  /* staggered_CC2N / staggered_N2CC: hw/grid/derivative_stencils.f90:31-73 */
  c->cc2n_D = dalloc((size_t)sc - 1);
  c->cc2n_U = dalloc((size_t)sc - 1);
  for (int i = 0; i < sc - 1; ++i) {
    c->cc2n_D[i] = -(1.0 / c->dhc[i]);
    c->cc2n_U[i] = 1.0 / c->dhc[i];
  }
  c->n2cc_D = dalloc((size_t)sn - 1);
  c->n2cc_U = dalloc((size_t)sn - 1);
  for (int i = 0; i < sn - 1; ++i) {
    c->n2cc_D[i] = -(1.0 / c->dhn[i]);
    c->n2cc_U[i] = 1.0 / c->dhn[i];
  }
  for (int n = 0; n < 2; ++n) {
    c->colCC_L[n] = dalloc((size_t)sc - 2);
    c->colCC_D[n] = dalloc((size_t)sc - 2);
    c->colCC_U[n] = dalloc((size_t)sc - 2);
    c->colN_L[n] = dalloc((size_t)sn - 2);
    c->colN_D[n] = dalloc((size_t)sn - 2);
    c->colN_U[n] = dalloc((size_t)sn - 2);
  }
  col_cc_1(c->dhc, sc, c->colCC_L[0], c->colCC_D[0], c->colCC_U[0]);
  col_cc_2(c->dhc, sc, c->colCC_L[1], c->colCC_D[1], c->colCC_U[1]);
  col_node_1(c->dhn, sn, c->colN_L[0], c->colN_D[0], c->colN_U[0]);
  col_node_2(c->dhn, sn, c->colN_L[1], c->colN_D[1], c->colN_U[1]);
}

static void coords_free(orc_coords *c) {
  double **ps[] = {&c->hn, &c->hc, &c->dhn, &c->dhc, &c->cc2n_D, &c->cc2n_U, &c->n2cc_D,
                   &c->n2cc_U, &c->theta_D, &c->theta_U};
  for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); ++i) free(*ps[i]);
  for (int n = 0; n < 2; ++n) {
    free(c->colCC_L[n]); free(c->colCC_D[n]); free(c->colCC_U[n]);
    free(c->colN_L[n]); free(c->colN_D[n]); free(c->colN_U[n]);
  }
}

orc_grid *orc_grid_create(const double *hx, int sx, const double *hy, int sy, const double *hz, int sz) {
  if (sx < 4 || sy < 4 || sz < 4) return NULL;
  orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
  coords_init(&g->c[0], hx, sx);
  coords_init(&g->c[1], hy, sy);
  coords_init(&g->c[2], hz, sz);
  return g;
}

void orc_grid_destroy(orc_grid *g) {
  if (!g) return;
  for (int d = 0; d < 3; ++d) coords_free(&g->c[d]);
  free(g);
}

void orc_field_shape(const orc_grid *g, int dl, int s[3]) {
  for (int d = 0; d < 3; ++d) s[d] = ((dl >> d) & 1) ? g->c[d].sn : g->c[d].sc;
}

size_t orc_field_size(const orc_grid *g, int dl) {
  int s[3];
  orc_field_shape(g, dl, s);
  return (size_t)s[0] * (size_t)s[1] * (size_t)s[2];
}

int orc_coord_array(const orc_grid *g, int dir, int which, double *out) {
  const orc_coords *c = &g->c[dir];
  const double *src = NULL;
  int n = 0;
  switch (which) {
    case 0: src = c->hn; n = c->sn; break;
    case 1: src = c->hc; n = c->sc; break;
    case 2: src = c->dhn; n = c->sn - 1; break;
    case 3: src = c->dhc; n = c->sc - 1; break;
    case 4: src = c->cc2n_D; n = c->sc - 1; break;
    case 5: src = c->cc2n_U; n = c->sc - 1; break;
    case 6: src = c->n2cc_D; n = c->sn - 1; break;
    case 7: src = c->n2cc_U; n = c->sn - 1; break;
    case 8: src = c->theta_D; n = c->sc - 1; break;
    case 9: src = c->theta_U; n = c->sc - 1; break;
    default:
      if (which >= 10 && which < 16) {
        int nn = (which - 10) / 3, t = (which - 10) % 3;
        src = t == 0 ? c->colCC_L[nn] : t == 1 ? c->colCC_D[nn] : c->colCC_U[nn];
        n = c->sc - 2;
      } else if (which >= 16 && which < 22) {
        int nn = (which - 16) / 3, t = (which - 16) % 3;
        src = t == 0 ? c->colN_L[nn] : t == 1 ? c->colN_D[nn] : c->colN_U[nn];
        n = c->sn - 2;
      } else {
        return -1;
      }
  }
  if (out) memcpy(out, src, sizeof(double) * (size_t)n);
  return n;
}

/* ------------------------------------------------------------------------- */
/* pencil kernels: hw/stencils/stencils.f90                                  */
/* ------------------------------------------------------------------------- */

/* stag_assign/add/subtract (stencils.f90:47-59,108-120,169-181).  gen 1/2/3. */
static void pencil_stag(double *out, size_t os, int sd, const double *f, size_t fs, int s,
                        const double *D, const double *U, int gt, int gen) {
  out[0] = 0.0;
  out[(size_t)(sd - 1) * os] = 0.0;
  for (int i = 0; i < s - 1; ++i) {
    double a = f[(size_t)i * fs], b = f[(size_t)(i + 1) * fs];
    double *o = &out[(size_t)(i + gt) * os];
    if (gen == 1) *o = a * D[i] + b * U[i];
    else if (gen == 2) *o = *o + a * D[i] + b * U[i];
    else *o = *o - (a * D[i] + b * U[i]);
  }
}

/* col_CC_assign/add/subtract with pad1 = pad2 = 0 (stencils.f90:60-84,121-145,182-206) */
static void pencil_col_cc(double *out, size_t os, const double *f, size_t fs, int s,
                          const double *L, const double *D, const double *U, int gen) {
#define F(i) f[(size_t)(i) * fs]
#define O(i) out[(size_t)(i) * os]
  for (int i = 2; i <= s - 3; ++i) {
    double t = F(i - 1) * L[i - 1];
    if (gen == 1) O(i) = t + F(i) * D[i - 1] + F(i + 1) * U[i - 1];
    else if (gen == 2) O(i) = O(i) + t + F(i) * D[i - 1] + F(i + 1) * U[i - 1];
    else O(i) = O(i) - (t + F(i) * D[i - 1] + F(i + 1) * U[i - 1]);
  }
  {
    /* near-wall rows: 0.5*(f(1)+f(2))*L(1)*(1-pad1) + f(1)*L(1)*pad1 + f(2)*D(1) + f(3)*U(1) */
    double t1 = 0.5 * (F(0) + F(1)) * L[0] * 1.0;
    double t2 = F(0) * L[0] * 0.0;
    double v = t1 + t2 + F(1) * D[0] + F(2) * U[0];
    double t3 = 0.5 * (F(s - 2) + F(s - 1)) * U[s - 3] * 1.0;
    double t4 = F(s - 1) * U[s - 3] * 0.0;
    double w = F(s - 3) * L[s - 3] + F(s - 2) * D[s - 3] + t3 + t4;
    if (gen == 1) { O(1) = v; O(s - 2) = w; }
    else if (gen == 2) {
      O(1) = O(1) + t1 + t2 + F(1) * D[0] + F(2) * U[0];
      O(s - 2) = O(s - 2) + F(s - 3) * L[s - 3] + F(s - 2) * D[s - 3] + t3 + t4;
    } else { O(1) = O(1) - v; O(s - 2) = O(s - 2) - w; }
  }
  O(0) = 0.0;
  O(s - 1) = 0.0;
}

/* col_N_assign/add/subtract with pad1 = pad2 = 0 (stencils.f90:85-106,146-167,207-228) */
static void pencil_col_n(double *out, size_t os, const double *f, size_t fs, int s,
                         const double *L, const double *D, const double *U, int gen) {
  for (int i = 2; i <= s - 3; ++i) {
    double t = F(i - 1) * L[i - 1];
    if (gen == 1) O(i) = t + F(i) * D[i - 1] + F(i + 1) * U[i - 1];
    else if (gen == 2) O(i) = O(i) + t + F(i) * D[i - 1] + F(i + 1) * U[i - 1];
    else O(i) = O(i) - (t + F(i) * D[i - 1] + F(i + 1) * U[i - 1]);
  }
  double v = F(1) * L[0] + F(2) * D[0] + F(3) * U[0];
  double w = F(s - 4) * L[s - 3] + F(s - 3) * D[s - 3] + F(s - 2) * U[s - 3];
  if (gen == 1) { O(1) = v; O(s - 2) = w; }
  else if (gen == 2) {
    O(1) = O(1) + F(1) * L[0] + F(2) * D[0] + F(3) * U[0];
    O(s - 2) = O(s - 2) + F(s - 4) * L[s - 3] + F(s - 3) * D[s - 3] + F(s - 2) * U[s - 3];
  } else { O(1) = O(1) - v; O(s - 2) = O(s - 2) - w; }
  O(0) = 0.0;
  O(s - 1) = 0.0;
#undef F
#undef O
}

static void zero_plane(double *u, const int s[3], int d, int idx, double val) {
  int a = (d + 1) % 3, b = (d + 2) % 3;
  int ii[3];
  for (int jb = 0; jb < s[b]; ++jb)
    for (int ja = 0; ja < s[a]; ++ja) {
      ii[d] = idx; ii[a] = ja; ii[b] = jb;
      u[IDX(s, ii[0], ii[1], ii[2])] = val;
    }
}

/* diff / diff_tree_search / diff_stag / diff_col: hw/ops/ops_del.f90:59-230 */
void orc_del(double *dfdh, int dl_out, const double *f, int dl_in, const orc_grid *g, int n,
             int dir, int pad, int gen) {
  int sf[3], so[3];
  orc_field_shape(g, dl_in, sf);
  orc_field_shape(g, dl_out, so);
  const orc_coords *c = &g->c[dir];
  int nin = (dl_in >> dir) & 1, nout = (dl_out >> dir) & 1;
  size_t strf[3] = {1, (size_t)sf[0], (size_t)sf[0] * sf[1]};
  size_t stro[3] = {1, (size_t)so[0], (size_t)so[0] * so[1]};
  int a = (dir + 1) % 3, b = (dir + 2) % 3;
  int lo = (a < b) ? a : b, hi = (a < b) ? b : a; /* reference loops the outer dims in k,j / k,i / j,i */
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
  for (int q = pad; q < sf[hi] - pad; ++q)
    for (int p = pad; p < sf[lo] - pad; ++p) {
      size_t offf = (size_t)p * strf[lo] + (size_t)q * strf[hi];
      size_t offo = (size_t)p * stro[lo] + (size_t)q * stro[hi];
      if (!nin && nout)
        pencil_stag(dfdh + offo, stro[dir], so[dir], f + offf, strf[dir], sf[dir], c->cc2n_D, c->cc2n_U, 1, gen);
      else if (nin && !nout)
        pencil_stag(dfdh + offo, stro[dir], so[dir], f + offf, strf[dir], sf[dir], c->n2cc_D, c->n2cc_U, 0, gen);
      else if (!nin && !nout)
        pencil_col_cc(dfdh + offo, stro[dir], f + offf, strf[dir], sf[dir], c->colCC_L[n - 1], c->colCC_D[n - 1], c->colCC_U[n - 1], gen);
      else
        pencil_col_n(dfdh + offo, stro[dir], f + offf, strf[dir], sf[dir], c->colN_L[n - 1], c->colN_D[n - 1], c->colN_U[n - 1], gen);
    }
  if (gen == 1 && pad > 0) {
    zero_plane(dfdh, so, a, 0, 0.0); zero_plane(dfdh, so, a, so[a] - 1, 0.0);
    zero_plane(dfdh, so, b, 0, 0.0); zero_plane(dfdh, so, b, so[b] - 1, 0.0);
  }
}

/* ------------------------------------------------------------------------- */
/* interpolation                                                             */
/* ------------------------------------------------------------------------- */

/* interpO2_SF: hw/ops/ops_interp.f90:190-221 over interp_C2N_GF / interp_N2C_GF
   (hw/grid_field/GF_interp/GF_interp.f90:19-76) and extrap_staggered_GF
   (GF_extrap.f90:17-46).  Loops span the whole other two dimensions. */
void orc_interp(double *out, int dl_out, const double *in, int dl_in, const orc_grid *g, int dir) {
  int so[3], si[3];
  orc_field_shape(g, dl_out, so);
  orc_field_shape(g, dl_in, si);
  int nout = (dl_out >> dir) & 1, nin = (dl_in >> dir) & 1;
  const orc_coords *c = &g->c[dir];
  int a = (dir + 1) % 3, b = (dir + 2) % 3;
  size_t stro[3] = {1, (size_t)so[0], (size_t)so[0] * so[1]};
  size_t stri[3] = {1, (size_t)si[0], (size_t)si[0] * si[1]};
  for (int pb = 0; pb < si[b]; ++pb)
    for (int pa = 0; pa < si[a]; ++pa) {
      const double *C = in + (size_t)pa * stri[a] + (size_t)pb * stri[b];
      double *N = out + (size_t)pa * stro[a] + (size_t)pb * stro[b];
      size_t fs = stri[dir], os = stro[dir];
      if (nout && !nin) { /* C -> N */
        int sC = si[dir], sN = so[dir];
        for (int i = 1; i <= sC - 1; ++i)
          N[(size_t)i * os] = C[(size_t)i * fs] * c->theta_D[i - 1] + C[(size_t)(i - 1) * fs] * c->theta_U[i - 1];
        N[0] = 2.0 * C[0] - N[os];
        N[(size_t)(sN - 1) * os] = 2.0 * C[(size_t)(sC - 1) * fs] - N[(size_t)(sN - 2) * os];
      } else { /* N -> C */
        int sN = si[dir];
        for (int i = 0; i < sN - 1; ++i) N[(size_t)i * os] = 0.5 * (C[(size_t)i * fs] + C[(size_t)(i + 1) * fs]);
      }
    }
}

/* face2Edge_SF_1 with edgeDir != faceDir (ops_interp.f90:242-270): one
   interpolation along the remaining direction. */
static void face2edge(double *edge, int edge_dir, const double *face, int face_dir, const orc_grid *g) {
  int orth = 3 - edge_dir - face_dir;
  int dl_face = 1 << face_dir;
  int dl_edge = 7 ^ (1 << edge_dir);
  orc_interp(edge, dl_edge, face, dl_face, g, orth);
}

/* edgeCrossFace_E (ops_advect.f90:202-218) = face2Edge_TF_no_diag
   (ops_interp.f90:531-544) then cross_product_TF (hw/fields/TF_extend.f90:385-400,
   GF_cross_product.f90: x = Ay Bz - Az By, y = -(Ax Bz - Az Bx), z = Ax By - Ay Bx). */
void orc_edge_cross_face_E(double *Ex, double *Ey, double *Ez, double *const U_E[6], double *const B[3],
                           const orc_grid *g) {
  size_t nx = orc_field_size(g, ORC_DL_EX), ny = orc_field_size(g, ORC_DL_EY), nz = orc_field_size(g, ORC_DL_EZ);
  /* B components on edges: Bz_x, By_x (x-edges), Bx_y, Bz_y (y-edges), Bx_z, By_z (z-edges) */
  double *Bzx = dalloc(nx), *Byx = dalloc(nx), *Bxy = dalloc(ny), *Bzy = dalloc(ny), *Bxz = dalloc(nz), *Byz = dalloc(nz);
  face2edge(Bxy, 1, B[0], 0, g);
  face2edge(Bxz, 2, B[0], 0, g);
  face2edge(Byx, 0, B[1], 1, g);
  face2edge(Byz, 2, B[1], 1, g);
  face2edge(Bzx, 0, B[2], 2, g);
  face2edge(Bzy, 1, B[2], 2, g);
  const double *Uyx = U_E[0], *Uzx = U_E[1], *Uxy = U_E[2], *Uzy = U_E[3], *Uxz = U_E[4], *Uyz = U_E[5];
  for (size_t q = 0; q < nx; ++q) Ex[q] = Uyx[q] * Bzx[q] - Uzx[q] * Byx[q];
  for (size_t q = 0; q < ny; ++q) Ey[q] = -(Uxy[q] * Bzy[q] - Uzy[q] * Bxy[q]);
  for (size_t q = 0; q < nz; ++q) Ez[q] = Uxz[q] * Byz[q] - Uyz[q] * Bxz[q];
  free(Bzx); free(Byx); free(Bxy); free(Bzy); free(Bxz); free(Byz);
}

/* advect_B (ops_advect.f90:186-200): adv = curl(edgeCrossFace_E(U_E, B)),
   curl_VF of an edge field (ops_discrete.f90:357-365). */
void orc_advect_B(double *adv[3], double *const U_E[6], double *const B[3], const orc_grid *g) {
  double *E[3];
  E[0] = dalloc(orc_field_size(g, ORC_DL_EX));
  E[1] = dalloc(orc_field_size(g, ORC_DL_EY));
  E[2] = dalloc(orc_field_size(g, ORC_DL_EZ));
  orc_edge_cross_face_E(E[0], E[1], E[2], U_E, B, g);
  for (int d = 0; d < 3; ++d) orc_curl_component(adv[d], 1 << d, E[0], E[1], E[2], 0, g, d);
  for (int d = 0; d < 3; ++d) free(E[d]);
}

/* ------------------------------------------------------------------------- */
/* BCs                                                                       */
/* ------------------------------------------------------------------------- */

/* Ghost/boundary/first-interior plane indices per face, hw/sub_domain/face_SD_extend.f90:37-90 */
static void face_planes(int nalong, int s, int is_max, int *G, int *B, int *I, int *Iopp) {
  if (!nalong) {
    *G = is_max ? s - 1 : 0; *B = -1; *I = is_max ? s - 2 : 1; *Iopp = is_max ? 1 : s - 2;
  } else {
    *G = is_max ? s - 1 : 0; *B = is_max ? s - 2 : 1; *I = is_max ? s - 3 : 2; *Iopp = is_max ? 2 : s - 3;
  }
}

/* dh of a face's sub-domain (FSD%dh, hw/sub_domain/face_SD_extend.f90:63):
   get_dh_boundary (hw/sub_domain/sub_domain_extend.f90:125-134) returns dhn(1)
   of the three-node face grid get_face_GI keeps (hw/grid/grid_extend.f90:260-268,
   hw/grid/coordinates_extend.f90:472-485): ghost and boundary node at the min
   face, the last interior cell at the max face. */
double orc_face_dh(const orc_coords *c, int is_max) {
  return is_max ? c->dhn[c->sn - 3] : c->dhn[0];
}

double orc_grid_face_dh(const orc_grid *g, int d, int is_max) { return orc_face_dh(&g->c[d], is_max); }

/* apply_BCs_faces (hw/BCs/apply/apply_BCs.f90:42-64) over raw kernels
   hw/BCs/direct/apply_BCs_faces_raw.f90:29-283 and
   hw/BCs/implicit/apply_BCs_faces_raw_implicit.f90:27-257.  Faces are applied
   in order 1..6 (the reference runs them in an OpenMP loop; only edge/corner
   ghosts, shared by two faces, depend on that order). */
void orc_apply_bcs(double *u, int dl, const orc_grid *g, const orc_bcs *bc, int implicit_) {
  int s[3];
  orc_field_shape(g, dl, s);
  for (int face = 0; face < 6; ++face) {
    int t = bc->type[face];
    if (t == ORC_BC_NONE) continue;
    int d = face / 2, is_max = face % 2;
    int nal = (dl >> d) & 1;
    int G, B, I, Io;
    face_planes(nal, s[d], is_max, &G, &B, &I, &Io);
    double nhat = is_max ? 1.0 : -1.0;
    const orc_coords *c = &g->c[d];
    double dh = orc_face_dh(c, is_max);
    double bv = implicit_ ? 0.0 : bc->val[face];
    int Gp = G, Iop = Io;
    if (nal && t == ORC_BC_PERIODIC && is_max && s[d] > 4) { Gp = B; Iop = 1; }
    int a = (d + 1) % 3, bb = (d + 2) % 3;
    int ii[3];
    for (int jb = 0; jb < s[bb]; ++jb)
      for (int ja = 0; ja < s[a]; ++ja) {
        ii[a] = ja; ii[bb] = jb;
        ii[d] = G; size_t ig = IDX(s, ii[0], ii[1], ii[2]);
        ii[d] = I; size_t iin = IDX(s, ii[0], ii[1], ii[2]);
        switch (t) {
          case ORC_BC_DIRICHLET:
            if (!nal) {
              u[ig] = implicit_ ? -u[iin] : 2.0 * bv - u[iin];
            } else {
              ii[d] = B; size_t ib = IDX(s, ii[0], ii[1], ii[2]);
              u[ib] = bv;
              u[ig] = implicit_ ? -u[iin] : 2.0 * u[ib] - u[iin];
            }
            break;
          case ORC_BC_NEUMANN:
            if (implicit_) u[ig] = u[iin];
            else if (!nal) u[ig] = u[iin] + bv * (dh * nhat);
            else u[ig] = u[iin] + 2.0 * bv * (dh * nhat);
            break;
          case ORC_BC_PERIODIC: {
            ii[d] = Gp; size_t igp = IDX(s, ii[0], ii[1], ii[2]);
            ii[d] = Iop; size_t io = IDX(s, ii[0], ii[1], ii[2]);
            u[igp] = u[io];
          } break;
          case ORC_BC_SYMMETRIC: u[ig] = u[iin]; break;
          case ORC_BC_ANTISYMMETRIC: u[ig] = -u[iin]; break;
          case ORC_BC_ROBIN: {
            /* u + c_w du/dh = 0: apply_Robin_C/N (hw/BCs/direct/apply_BCs_faces_raw.f90:236-282)
               and apply_Robin_C/N_implicit (hw/BCs/implicit/apply_BCs_faces_raw_implicit.f90:213-257),
               with the face sub-domain's dh and nhat. */
            double dr = dh, cw = bc->cw[face];
            if (!nal) {
              double c1 = (2.0 * cw * nhat / dr - 1.0) / (2.0 * cw * nhat / dr + 1.0);
              /* the direct rule's bvals coefficient is assigned 0 before use (:244-245) */
              u[ig] = implicit_ ? u[iin] * c1 : u[iin] * c1 + bv * 0.0;
            } else {
              ii[d] = B; size_t ib = IDX(s, ii[0], ii[1], ii[2]);
              if (implicit_) {
                double c1 = -(2.0 * dr * nhat / cw);
                u[ig] = u[iin] + u[ib] * c1;
              } else {
                double c1 = -(2.0 * dr / cw * nhat), c2 = 2.0 * dr / cw * nhat;
                u[ig] = u[iin] + u[ib] * c1 + bv * c2;
              }
            }
          } break;
          default: break;
        }
      }
  }
}

/* assign_ghost_XPeriodic: hw/block_field/block_field_extend.f90:233-271.
   bc == NULL means the field carries no BCs: all six ghost planes are assigned. */
void orc_assign_ghost_xperiodic(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val) {
  int s[3];
  orc_field_shape(g, dl, s);
  for (int face = 0; face < 6; ++face) {
    if (bc && bc->type[face] == ORC_BC_PERIODIC) continue;
    int d = face / 2;
    zero_plane(u, s, d, face % 2 ? s[d] - 1 : 0, val);
  }
}

static void assign_ghost_n_xperiodic(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val) {
  int s[3];
  orc_field_shape(g, dl, s);
  if (!bc) return;
  for (int face = 0; face < 6; ++face) {
    if (bc->type[face] == ORC_BC_PERIODIC) continue;
    int d = face / 2;
    if (!((dl >> d) & 1)) continue;
    zero_plane(u, s, d, face % 2 ? s[d] - 1 : 0, val);
  }
}

/* set_assign_wall_Dirichlet: hw/block_field/block_field_extend.f90:273-308 */
void orc_assign_wall_dirichlet(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val) {
  int s[3];
  orc_field_shape(g, dl, s);
  for (int face = 0; face < 6; ++face) {
    int d = face / 2;
    if (!((dl >> d) & 1)) continue;
    if (bc && bc->type[face] != ORC_BC_DIRICHLET) continue;
    zero_plane(u, s, d, face % 2 ? s[d] - 2 : 1, val);
  }
}

/* set_multiply_wall_Neumann: hw/block_field/block_field_extend.f90:348-377 */
void orc_multiply_wall_neumann(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val) {
  int s[3];
  orc_field_shape(g, dl, s);
  for (int face = 0; face < 6; ++face) {
    int d = face / 2;
    if (!((dl >> d) & 1)) continue;
    if (bc && bc->type[face] != ORC_BC_NEUMANN) continue;
    int idx = face % 2 ? s[d] - 2 : 1;
    int a = (d + 1) % 3, b = (d + 2) % 3, ii[3];
    for (int jb = 0; jb < s[b]; ++jb)
      for (int ja = 0; ja < s[a]; ++ja) {
        ii[d] = idx; ii[a] = ja; ii[b] = jb;
        size_t q = IDX(s, ii[0], ii[1], ii[2]);
        u[q] = u[q] * val;
      }
  }
}

/* set_assign_wall_Periodic_single: hw/block_field/block_field_extend.f90:310-346 */
static void assign_wall_periodic_single(double *u, int dl, const orc_grid *g, const orc_bcs *bc, double val) {
  int s[3];
  orc_field_shape(g, dl, s);
  if (!bc) return;
  for (int d = 0; d < 3; ++d) {
    if (!((dl >> d) & 1)) continue;
    if (bc->type[2 * d] != ORC_BC_PERIODIC || bc->type[2 * d + 1] != ORC_BC_PERIODIC) continue;
    if (!(s[d] > 4)) continue;
    zero_plane(u, s, d, s[d] - 2, val);
  }
}

/* ------------------------------------------------------------------------- */
/* discrete operators: hw/ops/ops_discrete.f90                               */
/* ------------------------------------------------------------------------- */

/* grad_SF: ops_discrete.f90:173-182 (pad 0) */
void orc_grad(double *gx, double *gy, double *gz, const double *u, int dl_u, const orc_grid *g) {
  double *out[3] = {gx, gy, gz};
  for (int d = 0; d < 3; ++d)
    if (out[d]) orc_del(out[d], dl_u ^ (1 << d), u, dl_u, g, 1, d, 0, 1);
}

/* div_SF: ops_discrete.f90:158-171.  dl_u is the location of u (x comp.);
   components are u (N along x), v, w.  bc_out: BCs of divU (NULL = none). */
void orc_div(double *divU, int dl_out, const double *u, const double *v, const double *w, int dl_u,
             const orc_grid *g, const orc_bcs *bc_out) {
  (void)dl_u;
  orc_del(divU, dl_out, u, dl_out ^ 1, g, 1, 0, 0, 1);
  orc_del(divU, dl_out, v, dl_out ^ 2, g, 1, 1, 0, 2);
  orc_del(divU, dl_out, w, dl_out ^ 4, g, 1, 2, 0, 2);
  orc_assign_ghost_xperiodic(divU, dl_out, g, bc_out, 0.0);
}

/* curl_SF: ops_discrete.f90:184-202.  Component dir of curl(u,v,w); the
   inputs' locations follow from the output location by toggling. */
void orc_curl_component(double *curlU, int dl_out, const double *u, const double *v, const double *w,
                        int dl_in, const orc_grid *g, int dir) {
  (void)dl_in;
  switch (dir) {
    case 0:
      orc_del(curlU, dl_out, w, dl_out ^ 2, g, 1, 1, 0, 1);
      orc_del(curlU, dl_out, v, dl_out ^ 4, g, 1, 2, 0, 3);
      break;
    case 1:
      orc_del(curlU, dl_out, u, dl_out ^ 4, g, 1, 2, 0, 1);
      orc_del(curlU, dl_out, w, dl_out ^ 1, g, 1, 0, 0, 3);
      break;
    default:
      orc_del(curlU, dl_out, v, dl_out ^ 1, g, 1, 0, 0, 1);
      orc_del(curlU, dl_out, u, dl_out ^ 2, g, 1, 1, 0, 3);
      break;
  }
}

/* lap_centered_SF_new: ops_discrete.f90:124-132 (grad then div) */
void orc_lap_centered_sf(double *lapU, const double *U, int dl, const orc_grid *g, const orc_bcs *bc_out) {
  double *t[3];
  for (int d = 0; d < 3; ++d) t[d] = dalloc(orc_field_size(g, dl ^ (1 << d)));
  orc_grad(t[0], t[1], t[2], U, dl, g);
  orc_div(lapU, dl, t[0], t[1], t[2], dl ^ 1, g, bc_out);
  for (int d = 0; d < 3; ++d) free(t[d]);
}

/* ------------------------------------------------------------------------- */
/* auxiliaries                                                               */
/* ------------------------------------------------------------------------- */

/* volume_DL_GF: hw/grid_field/GF_distributions.f90:61-88 */
void orc_volume(double *vol, int dl, const orc_grid *g) {
  int s[3];
  orc_field_shape(g, dl, s);
  memset(vol, 0, sizeof(double) * orc_field_size(g, dl));
  const double *dh[3];
  int e[3];
  for (int d = 0; d < 3; ++d) {
    e[d] = (dl >> d) & 1;
    dh[d] = e[d] ? g->c[d].dhc : g->c[d].dhn;
  }
  for (int k = 1; k < s[2] - 1; ++k)
    for (int j = 1; j < s[1] - 1; ++j)
      for (int i = 1; i < s[0] - 1; ++i)
        vol[IDX(s, i, j, k)] = dh[0][i - e[0]] * dh[1][j - e[1]] * dh[2][k - e[2]];
}

/* laplacian_diagonal: hw/grid_field/GF_diagonals.f90:15-35 */
void orc_lap_diagonal(double *diag, int dl, const orc_grid *g) {
  int s[3];
  orc_field_shape(g, dl, s);
  memset(diag, 0, sizeof(double) * orc_field_size(g, dl));
  int p[3];
  for (int d = 0; d < 3; ++d) p[d] = (dl >> d) & 1;
  const orc_coords *cx = &g->c[0], *cy = &g->c[1], *cz = &g->c[2];
  for (int k = 1; k < s[2] - 1; ++k)
    for (int j = 1; j < s[1] - 1; ++j)
      for (int i = 1; i < s[0] - 1; ++i)
        diag[IDX(s, i, j, k)] = cx->n2cc_D[i] * cx->cc2n_U[i - 1] + cx->n2cc_U[i - p[0]] * cx->cc2n_D[i - p[0]] +
                                cy->n2cc_D[j] * cy->cc2n_U[j - 1] + cy->n2cc_U[j - p[1]] * cy->cc2n_D[j - p[1]] +
                                cz->n2cc_D[k] * cz->cc2n_U[k - 1] + cz->n2cc_U[k - p[2]] * cz->cc2n_D[k - p[2]];
}

/* sum_GF_pad: hw/grid_field/GF_aux.f90:140-158 (sequential order k, j, i) */
double orc_sum_pad(const double *u, const int s[3], int pad) {
  double m = 0.0;
  if (g_threads > 1) {
    const int nk = s[2] - 2 * pad;
    double *part = (double *)calloc(nk > 0 ? (size_t)nk : 1, sizeof(double));
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = pad; k < s[2] - pad; ++k) {
      double a = 0.0;
      for (int j = pad; j < s[1] - pad; ++j)
        for (int i = pad; i < s[0] - pad; ++i) a += u[IDX(s, i, j, k)];
      part[k - pad] = a;
    }
    for (int k = 0; k < nk; ++k) m += part[k];
    free(part);
    return m;
  }
  for (int k = pad; k < s[2] - pad; ++k)
    for (int j = pad; j < s[1] - pad; ++j)
      for (int i = pad; i < s[0] - pad; ++i) m += u[IDX(s, i, j, k)];
  return m;
}

double orc_sum_all(const double *u, size_t n) {
  double m = 0.0;
  for (size_t q = 0; q < n; ++q) m += u[q];
  return m;
}

/* dot_product_SF: hw/ops/ops_aux.f90:209-217 */
static double dot_sf(const double *A, const double *B, int dl, const orc_grid *g, const orc_bcs *bcx, double *tmp) {
  int s[3];
  orc_field_shape(g, dl, s);
  size_t n = orc_field_size(g, dl);
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
  for (size_t q = 0; q < n; ++q) tmp[q] = A[q] * B[q];
  orc_assign_wall_dirichlet(tmp, dl, g, bcx, 0.0);
  return orc_sum_pad(tmp, s, 1);
}

/* Lap_uniform_SF_explicit (matrix_free_operators.f90:35-50): lap_centered;
   the output carries no BCs so all its ghost planes are zeroed by div_SF. */
static void lap_explicit(double *Ax, const double *x, int dl, const orc_grid *g) {
  orc_lap_centered_sf(Ax, x, dl, g, NULL);
}

/* nrg_diffusion_explicit (matrix_free_operators.f90:148-161):
   grad(tempk%x, x); multiply(tempk%x, k%x); div(Ax, tempk%x); multiply(Ax, c);
   add(Ax, x).  k holds three face arrays (x-, y-, z-faces). */
static void nrg_explicit(double *Ax, const double *x, double *const k[3], double c, const orc_grid *g) {
  double *t[3];
  size_t nf[3];
  for (int d = 0; d < 3; ++d) {
    nf[d] = orc_field_size(g, 1 << d);
    t[d] = dalloc(nf[d]);
  }
  orc_grad(t[0], t[1], t[2], x, ORC_DL_CC, g);
  for (int d = 0; d < 3; ++d)
    for (size_t q = 0; q < nf[d]; ++q) t[d][q] = t[d][q] * k[d][q];
  orc_div(Ax, ORC_DL_CC, t[0], t[1], t[2], ORC_DL_FX, g, NULL);
  size_t n = orc_field_size(g, ORC_DL_CC);
  for (size_t q = 0; q < n; ++q) Ax[q] = Ax[q] * c;
  for (size_t q = 0; q < n; ++q) Ax[q] = Ax[q] + x[q];
  for (int d = 0; d < 3; ++d) free(t[d]);
}

/* nrg_diffusion (matrix_free_operators.f90:163-174): implicit BCs on x first */
void orc_nrg_op(double *Ax, double *x, double *const k[3], double c, const orc_grid *g, const orc_bcs *bc) {
  orc_apply_bcs(x, ORC_DL_CC, g, bc, 1);
  nrg_explicit(Ax, x, k, c, g);
}

/* the operator of an SF PCG: Lap_uniform_SF_explicit (kind 0) or
   nrg_diffusion_explicit with k and c (kind 1) */
typedef struct {
  int kind;
  double *const *k;
  double c;
} sf_op;

static void op_explicit(const sf_op *op, double *Ax, const double *x, int dl, const orc_grid *g) {
  if (op->kind == 1) nrg_explicit(Ax, x, op->k, op->c, g);
  else lap_explicit(Ax, x, dl, g);
}

/* ------------------------------------------------------------------------- */
/* PCG: hw/solvers/PCG/PCG_solver_algorithm.f90:37-107, PCG_aux.f90:30-123    */
/* ------------------------------------------------------------------------- */
static int pcg_sf(const sf_op *op, const orc_grid *g, int dl, const orc_bcs *bc, int all_neumann, double *x,
                  const double *b, int iter_max, double tol_rel, double tol_abs, int n_skip, double *out,
                  double *res_hist);

int orc_pcg_lap_sf(const orc_grid *g, int dl, const orc_bcs *bc, int all_neumann, double *x,
                   const double *b, int iter_max, double tol_rel, double tol_abs, int n_skip, double *out,
                   double *res_hist) {
  sf_op op = {0, NULL, 0.0};
  return pcg_sf(&op, g, dl, bc, all_neumann, x, b, iter_max, tol_rel, tol_abs, n_skip, out, res_hist);
}

/* The energy equation's solver (energy_extend.f90:129-130): nrg_diffusion with
   prec_lap_SF, c = coeff_implicit_time_split set by update_MFP
   (PCG_solver_extend.f90:269-279). */
int orc_pcg_nrg_sf(const orc_grid *g, const orc_bcs *bc, int all_neumann, double *const k[3], double c, double *x,
                   const double *b, int iter_max, double tol_rel, double tol_abs, int n_skip, double *out,
                   double *res_hist) {
  sf_op op = {1, k, c};
  return pcg_sf(&op, g, ORC_DL_CC, bc, all_neumann, x, b, iter_max, tol_rel, tol_abs, n_skip, out, res_hist);
}

static int pcg_sf(const sf_op *op, const orc_grid *g, int dl, const orc_bcs *bc, int all_neumann, double *x,
                  const double *b, int iter_max, double tol_rel, double tol_abs, int n_skip, double *out,
                  double *res_hist) {
  if (n_skip < 1) return -1;
  int s[3];
  orc_field_shape(g, dl, s);
  size_t n = orc_field_size(g, dl);
  double *r = dalloc(n), *z = dalloc(n), *p = dalloc(n), *Ax = dalloc(n), *tmp = dalloc(n);
  double *vol = dalloc(n), *Minv = dalloc(n);
  orc_volume(vol, dl, g);
  double volsum = orc_sum_all(vol, n); /* volume_SF: SF_extend.f90:571-581 */
  /* prec_Lap_SF: preconditioners.f90:55-71 */
  orc_lap_diagonal(Minv, dl, g);
  for (size_t q = 0; q < n; ++q) Minv[q] = Minv[q] * vol[q];
  for (size_t q = 0; q < n; ++q) Minv[q] = 1.0 / Minv[q];
  orc_assign_ghost_xperiodic(Minv, dl, g, NULL, 0.0);

  /* modify_RHS_SF: PCG_aux.f90:30-54 */
  memcpy(r, b, sizeof(double) * n);
  if (all_neumann) { /* subtract_phys_mean_vol_SF: ops_aux.f90:146-167 */
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) tmp[q] = r[q] * vol[q];
    double mean = orc_sum_pad(tmp, s, 1) / volsum;
#pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) r[q] = r[q] - mean;
    orc_assign_ghost_xperiodic(r, dl, g, NULL, 0.0);
  }
  /* compute_Ax_BC / compute_x_BC: PCG_aux.f90:81-123 */
  memset(tmp, 0, sizeof(double) * n);
  orc_apply_bcs(tmp, dl, g, bc, 0);
  assign_ghost_n_xperiodic(tmp, dl, g, bc, 0.0);
  op_explicit(op, Ax, tmp, dl, g);
  orc_assign_ghost_xperiodic(Ax, dl, g, bc, 0.0);
  for (size_t q = 0; q < n; ++q) r[q] = r[q] - Ax[q];
  /* operator(Ax,x): the operator applies the implicit BCs to x first */
  orc_apply_bcs(x, dl, g, bc, 1);
  op_explicit(op, Ax, x, dl, g);
  for (size_t q = 0; q < n; ++q) r[q] = r[q] - Ax[q];
  for (size_t q = 0; q < n; ++q) r[q] = r[q] * vol[q];
  if (dl != ORC_DL_CC) {
    assign_wall_periodic_single(r, dl, g, bc, 0.0);
    orc_assign_wall_dirichlet(r, dl, g, bc, 0.0);
    orc_multiply_wall_neumann(r, dl, g, bc, 0.5);
  }

  /* norms compute: hw/PCG/norms_extend.f90:131-139, GF_norms.f90:24-41 (all cells) */
  double res0 = 0.0;
  for (size_t q = 0; q < n; ++q) res0 += r[q] * r[q];
  res0 = sqrt(res0);
  for (size_t q = 0; q < n; ++q) z[q] = Minv[q] * r[q];
  memcpy(p, z, sizeof(double) * n);
  double rhok = dot_sf(r, z, dl, g, bc, tmp);
  double res = sqrt(fabs(rhok));
  int nh = 0;
  if (res_hist) res_hist[nh++] = res;
  int iter = 0;
  /* update_exit_loop_ISP: hw/var_set/iter_solver_params_extend.f90:86-93 */
  int e1 = res / res0 < tol_rel, e2 = res < tol_abs, e3 = iter >= iter_max;
  (void)e1; (void)e3;
  if (!e2) {
    for (int it = 1; it <= iter_max; ++it) {
      orc_apply_bcs(p, dl, g, bc, 1);
      op_explicit(op, Ax, p, dl, g);
      orc_multiply_wall_neumann(Ax, dl, g, bc, 0.5);
      orc_assign_wall_dirichlet(Ax, dl, g, bc, 0.0);
      assign_wall_periodic_single(Ax, dl, g, bc, 0.0);
  #pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) Ax[q] = Ax[q] * vol[q];
      double alpha = rhok / dot_sf(p, Ax, dl, g, bc, tmp);
  #pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) x[q] = x[q] + p[q] * alpha;
      orc_apply_bcs(x, dl, g, bc, 0);
      iter++;
  #pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) r[q] = r[q] + Ax[q] * (-alpha);
      if (iter % n_skip == 0) {
        res = sqrt(fabs(dot_sf(r, r, dl, g, bc, tmp)));
        if (res_hist) res_hist[nh++] = res;
        e1 = res / res0 < tol_rel; e2 = res < tol_abs; e3 = iter >= iter_max;
        if (e1 || e2 || e3) break;
      }
  #pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) z[q] = Minv[q] * r[q];
      double rhokp1 = dot_sf(z, r, dl, g, bc, tmp);
      double beta = rhokp1 / rhok;
  #pragma omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (size_t q = 0; q < n; ++q) p[q] = p[q] * beta + z[q];
      rhok = rhokp1;
    }
  } else {
    orc_apply_bcs(x, dl, g, bc, 0);
    iter++;
  }
  if (out) { out[0] = (double)iter; out[1] = res0; out[2] = res; out[3] = (double)nh; }
  free(r); free(z); free(p); free(Ax); free(tmp); free(vol); free(Minv);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* vector-field PCG                                                          */
/* ------------------------------------------------------------------------- */

/* lap_centered_SF_given_both (ops_discrete.f90:112-122): staggered derivative
   out and back along each direction, accumulated into lapU. */
static void lap_given_both(double *lapU, const double *U, int dl, const orc_grid *g) {
  for (int d = 0; d < 3; ++d) {
    int dlt = dl ^ (1 << d);
    double *t = dalloc(orc_field_size(g, dlt));
    orc_del(t, dlt, U, dl, g, 1, d, 0, 1);
    orc_del(lapU, dl, t, dlt, g, 1, d, 0, d == 0 ? 1 : 2);
    free(t);
  }
}

/* mom_diffusion(_explicit) / ind_diffusion(_explicit):
   matrix_free_operators.f90:120-146,176-202 */
void orc_op_vf(double *Ax[3], double *x[3], int kind, double c, double *const k[3], const orc_grid *g,
               const orc_bcs bc[3], int implicit_) {
  if (implicit_)
    for (int d = 0; d < 3; ++d) orc_apply_bcs(x[d], 1 << d, g, &bc[d], 1);
  if (kind == ORC_OP_MOM_DIFFUSION) {
    for (int d = 0; d < 3; ++d) lap_given_both(Ax[d], x[d], 1 << d, g);
  } else {
    const int edl[3] = {ORC_DL_EX, ORC_DL_EY, ORC_DL_EZ};
    double *E[3];
    for (int d = 0; d < 3; ++d) E[d] = dalloc(orc_field_size(g, edl[d]));
    for (int d = 0; d < 3; ++d) orc_curl_component(E[d], edl[d], x[0], x[1], x[2], 0, g, d);
    for (int d = 0; d < 3; ++d) {
      size_t n = orc_field_size(g, edl[d]);
      for (size_t q = 0; q < n; ++q) E[d][q] = E[d][q] * k[d][q];
    }
    for (int d = 0; d < 3; ++d) orc_curl_component(Ax[d], 1 << d, E[0], E[1], E[2], 0, g, d);
    for (int d = 0; d < 3; ++d) free(E[d]);
  }
  for (int d = 0; d < 3; ++d) {
    size_t n = orc_field_size(g, 1 << d);
    for (size_t q = 0; q < n; ++q) Ax[d][q] = Ax[d][q] * c;
    for (size_t q = 0; q < n; ++q) Ax[d][q] = Ax[d][q] + x[d][q];
  }
}

/* curl_curl_diagonal: hw/grid_field/GF_diagonals.f90:37-71 */
static void curl_curl_diag(double *dg[3], double *const sig[3], const orc_grid *g) {
  const orc_coords *C[3] = {&g->c[0], &g->c[1], &g->c[2]};
  int sd[3][3], ss[3][3];
  for (int d = 0; d < 3; ++d) orc_field_shape(g, 1 << d, sd[d]);
  orc_field_shape(g, ORC_DL_EX, ss[0]);
  orc_field_shape(g, ORC_DL_EY, ss[1]);
  orc_field_shape(g, ORC_DL_EZ, ss[2]);
  for (int d = 0; d < 3; ++d) memset(dg[d], 0, sizeof(double) * orc_field_size(g, 1 << d));
  /* diag_x: y and z terms with sig_z and sig_y */
  for (int kk = 1; kk < sd[0][2] - 1; ++kk)
    for (int j = 1; j < sd[0][1] - 1; ++j)
      for (int i = 1; i < sd[0][0] - 1; ++i)
        dg[0][IDX(sd[0], i, j, kk)] =
            -C[1]->n2cc_D[j] * sig[2][IDX(ss[2], i, j, kk)] * C[1]->cc2n_U[j - 1] -
            C[1]->n2cc_U[j] * sig[2][IDX(ss[2], i, j + 1, kk)] * C[1]->cc2n_D[j] -
            C[2]->n2cc_D[kk] * sig[1][IDX(ss[1], i, j, kk)] * C[2]->cc2n_U[kk - 1] -
            C[2]->n2cc_U[kk] * sig[1][IDX(ss[1], i, j, kk + 1)] * C[2]->cc2n_D[kk];
  for (int kk = 1; kk < sd[1][2] - 1; ++kk)
    for (int j = 1; j < sd[1][1] - 1; ++j)
      for (int i = 1; i < sd[1][0] - 1; ++i)
        dg[1][IDX(sd[1], i, j, kk)] =
            -C[0]->n2cc_D[i] * sig[2][IDX(ss[2], i, j, kk)] * C[0]->cc2n_U[i - 1] -
            C[0]->n2cc_U[i] * sig[2][IDX(ss[2], i + 1, j, kk)] * C[0]->cc2n_D[i] -
            C[2]->n2cc_D[kk] * sig[0][IDX(ss[0], i, j, kk)] * C[2]->cc2n_U[kk - 1] -
            C[2]->n2cc_U[kk] * sig[0][IDX(ss[0], i, j, kk + 1)] * C[2]->cc2n_D[kk];
  for (int kk = 1; kk < sd[2][2] - 1; ++kk)
    for (int j = 1; j < sd[2][1] - 1; ++j)
      for (int i = 1; i < sd[2][0] - 1; ++i)
        dg[2][IDX(sd[2], i, j, kk)] =
            -C[0]->n2cc_D[i] * sig[1][IDX(ss[1], i, j, kk)] * C[0]->cc2n_U[i - 1] -
            C[0]->n2cc_U[i] * sig[1][IDX(ss[1], i + 1, j, kk)] * C[0]->cc2n_D[i] -
            C[1]->n2cc_D[j] * sig[0][IDX(ss[0], i, j, kk)] * C[1]->cc2n_U[j - 1] -
            C[1]->n2cc_U[j] * sig[0][IDX(ss[0], i, j + 1, kk)] * C[1]->cc2n_D[j];
}

void orc_curl_curl_diagonal(double *dg[3], double *const sig[3], const orc_grid *g) { curl_curl_diag(dg, sig, g); }

/* prec_mom_SF / prec_ind_VF: preconditioners.f90:86-129 */
void orc_minv_vf(double *Minv[3], int kind, double c, double *const k[3], const orc_grid *g) {
  if (kind == ORC_OP_IND_DIFFUSION) curl_curl_diag(Minv, k, g);
  for (int d = 0; d < 3; ++d) {
    int dl = 1 << d;
    size_t n = orc_field_size(g, dl);
    if (kind == ORC_OP_MOM_DIFFUSION) orc_lap_diagonal(Minv[d], dl, g);
    double *vol = dalloc(n);
    orc_volume(vol, dl, g);
    for (size_t q = 0; q < n; ++q) Minv[d][q] = Minv[d][q] * c;
    for (size_t q = 0; q < n; ++q) Minv[d][q] = Minv[d][q] + 1.0;
    for (size_t q = 0; q < n; ++q) Minv[d][q] = Minv[d][q] * vol[q];
    for (size_t q = 0; q < n; ++q) Minv[d][q] = 1.0 / Minv[d][q];
    orc_assign_ghost_xperiodic(Minv[d], dl, g, NULL, 0.0);
    free(vol);
  }
}

/* dot_product_VF: hw/ops/ops_aux.f90:199-207 */
static double dot_vf(double *const A[3], double *const B[3], const orc_grid *g, const orc_bcs bc[3]) {
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    int dl = 1 << d, sh[3];
    orc_field_shape(g, dl, sh);
    size_t n = orc_field_size(g, dl);
    double *t = dalloc(n);
    for (size_t q = 0; q < n; ++q) t[q] = A[d][q] * B[d][q];
    orc_assign_wall_dirichlet(t, dl, g, &bc[d], 0.0);
    s = s + orc_sum_pad(t, sh, 1);
    free(t);
  }
  return s;
}

int orc_pcg_vf(const orc_grid *g, int kind, double c, double *const k[3], const orc_bcs bc[3], double *x[3],
               double *const b[3], int iter_max, double tol_rel, double tol_abs, int n_skip, double *out,
               double *res_hist) {
  if (n_skip < 1) return -1;
  double *r[3], *z[3], *p[3], *Ax[3], *t[3], *vol[3], *Minv[3];
  size_t n[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = orc_field_size(g, 1 << d);
    r[d] = dalloc(n[d]); z[d] = dalloc(n[d]); p[d] = dalloc(n[d]); Ax[d] = dalloc(n[d]);
    t[d] = dalloc(n[d]); vol[d] = dalloc(n[d]); Minv[d] = dalloc(n[d]);
    orc_volume(vol[d], 1 << d, g);
  }
  orc_minv_vf(Minv, kind, c, k, g);
  /* modify_RHS_VF */
  for (int d = 0; d < 3; ++d) memcpy(r[d], b[d], sizeof(double) * n[d]);
  for (int d = 0; d < 3; ++d) {
    memset(t[d], 0, sizeof(double) * n[d]);
    orc_apply_bcs(t[d], 1 << d, g, &bc[d], 0);
    assign_ghost_n_xperiodic(t[d], 1 << d, g, &bc[d], 0.0);
  }
  orc_op_vf(Ax, t, kind, c, k, g, bc, 0);
  for (int d = 0; d < 3; ++d) {
    orc_assign_ghost_xperiodic(Ax[d], 1 << d, g, &bc[d], 0.0);
    for (size_t q = 0; q < n[d]; ++q) r[d][q] = r[d][q] - Ax[d][q];
  }
  orc_op_vf(Ax, x, kind, c, k, g, bc, 1);
  for (int d = 0; d < 3; ++d) {
    for (size_t q = 0; q < n[d]; ++q) r[d][q] = r[d][q] - Ax[d][q];
    for (size_t q = 0; q < n[d]; ++q) r[d][q] = r[d][q] * vol[d][q];
    assign_wall_periodic_single(r[d], 1 << d, g, &bc[d], 0.0);
    orc_assign_wall_dirichlet(r[d], 1 << d, g, &bc[d], 0.0);
    orc_multiply_wall_neumann(r[d], 1 << d, g, &bc[d], 0.5);
  }
  double res0 = 0.0;
  for (int d = 0; d < 3; ++d) {
    double e = 0.0;
    for (size_t q = 0; q < n[d]; ++q) e += r[d][q] * r[d][q];
    res0 = res0 + e;
  }
  res0 = sqrt(res0);
  for (int d = 0; d < 3; ++d) {
    for (size_t q = 0; q < n[d]; ++q) z[d][q] = Minv[d][q] * r[d][q];
    memcpy(p[d], z[d], sizeof(double) * n[d]);
  }
  double rhok = dot_vf(r, z, g, bc);
  double res = sqrt(fabs(rhok));
  int nh = 0;
  if (res_hist) res_hist[nh++] = res;
  int iter = 0;
  int e2 = res < tol_abs;
  if (!e2) {
    for (int it = 1; it <= iter_max; ++it) {
      orc_op_vf(Ax, p, kind, c, k, g, bc, 1);
      for (int d = 0; d < 3; ++d) {
        orc_multiply_wall_neumann(Ax[d], 1 << d, g, &bc[d], 0.5);
        orc_assign_wall_dirichlet(Ax[d], 1 << d, g, &bc[d], 0.0);
        assign_wall_periodic_single(Ax[d], 1 << d, g, &bc[d], 0.0);
        for (size_t q = 0; q < n[d]; ++q) Ax[d][q] = Ax[d][q] * vol[d][q];
      }
      double alpha = rhok / dot_vf(p, Ax, g, bc);
      for (int d = 0; d < 3; ++d) {
        for (size_t q = 0; q < n[d]; ++q) x[d][q] = x[d][q] + p[d][q] * alpha;
        orc_apply_bcs(x[d], 1 << d, g, &bc[d], 0);
      }
      iter++;
      for (int d = 0; d < 3; ++d)
        for (size_t q = 0; q < n[d]; ++q) r[d][q] = r[d][q] + Ax[d][q] * (-alpha);
      if (iter % n_skip == 0) {
        res = sqrt(fabs(dot_vf(r, r, g, bc)));
        if (res_hist) res_hist[nh++] = res;
        if (res / res0 < tol_rel || res < tol_abs || iter >= iter_max) break;
      }
      for (int d = 0; d < 3; ++d)
        for (size_t q = 0; q < n[d]; ++q) z[d][q] = Minv[d][q] * r[d][q];
      double rhokp1 = dot_vf(z, r, g, bc);
      double beta = rhokp1 / rhok;
      for (int d = 0; d < 3; ++d)
        for (size_t q = 0; q < n[d]; ++q) p[d][q] = p[d][q] * beta + z[d][q];
      rhok = rhokp1;
    }
  } else {
    for (int d = 0; d < 3; ++d) orc_apply_bcs(x[d], 1 << d, g, &bc[d], 0);
    iter++;
  }
  if (out) { out[0] = (double)iter; out[1] = res0; out[2] = res; out[3] = (double)nh; }
  for (int d = 0; d < 3; ++d) {
    free(r[d]); free(z[d]); free(p[d]); free(Ax[d]); free(t[d]); free(vol[d]); free(Minv[d]);
  }
  return 0;
}

/* SplitMix64 (Steele, Lea & Flood, OOPSLA 2014), counter form: value q uses
   state seed + (q+1)*golden, identical to the sequential generator. */
void orc_splitmix_fill(double *out, size_t n, unsigned long long seed, double amplitude) {
  for (size_t q = 0; q < n; ++q) {
    unsigned long long zz = seed + (unsigned long long)(q + 1) * 0x9E3779B97F4A7C15ULL;
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ULL;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBULL;
    zz = zz ^ (zz >> 31);
    double u = (double)(zz >> 11) * (1.0 / 9007199254740992.0); /* [0,1) */
    out[q] = amplitude * (2.0 * u - 1.0);
  }
}
