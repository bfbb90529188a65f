// mg.hpp -- FP32 geometric multigrid V-cycle for the cell-centred Laplacian,
// used as the preconditioner of the mixed-precision flexible PCG.
#pragma once

#include <string>

#include "common.hpp"

namespace moons {

struct MGHier;

// Builds the level hierarchy for a cell-centred field's mesh and BCs:
// coarse nodes are every other node of the finer level (the reference's
// coordinate restriction, code/handwritten/grid/coordinates_extend.f90:336-355),
// coarsening stops at an odd or < 4 cell count or after max_levels levels.
MGHier *mg_create(const moons_field_s *x, int max_levels, int nu1, int nu2, int coarse_sweeps, std::string *err);
void mg_destroy(MGHier *h);
int mg_levels(const MGHier *h);
int mg_level_cells(const MGHier *h, int l, int n[3]);
// Fine-level right-hand side f0 = r/vol and output x0, FP32, in the
// colour-split layout: cell (i,j,k) (1-based interior indices) at
// ((i+j+k) & 1) * map[2] + (i-1)/2 + map[0] * ((j-1) + map[1] * (k-1)).
float *mg_f0(MGHier *h);
float *mg_x0(MGHier *h);
void mg_f32_map(const MGHier *h, int map[3]);
// One V-cycle x0 = V f0 on the library stream; every kernel returns at once
// when *done (device) is non-zero.
int mg_vcycle(MGHier *h, const int *done);
// Start of a solve (the learned level's cycle count restarts).
void mg_begin_solve(MGHier *h);
// The learned initial guess in the first `cycles` V-cycles of each solve (0: every cycle).
void mg_set_learned_cycles(MGHier *h, int cycles);
// Levels from `level` down (and the coarse end) are cycled `cycles` times per
// visit; level < 0 or cycles = 1: the plain V-cycle.  MOONS_ERR_ARG when
// `level` is not a grid-launched level (the single-workgroup coarse end
// starts from zero at each launch).
int mg_set_gamma(MGHier *h, int level, int cycles);
void mg_set_params(MGHier *h, int nu1, int nu2, int coarse_sweeps);
// Average duration (ms) of one launch, reps launches back to back on the
// library stream: [0] a half-sweep of level 0, [1] restriction 0 -> 1,
// [2] prolongation 1 -> 0 (0 when there is one level), [3] the coarse end,
// [4] a whole V-cycle, [5] a fused whole sweep of level 0, [6]/[7] a
// half-sweep / fused sweep of level 1, [8]/[9] of level 2 (0 for a level
// that does not exist).  Overwrites the level arrays.
int mg_time_kernels(MGHier *h, int reps, double ms[10]);
// Learned initial guess on grid level `level` (net == nullptr removes it):
// before that level's pre-smoothing, x = c N(f) with the Galerkin factor c.
int mg_set_learned(MGHier *h, moons_unet_s *net, int level, std::string *err);

} // namespace moons
