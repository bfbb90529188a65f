// unet.hpp -- the learned preconditioner as seen by the flexible PCG.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace moons {

struct UnetPrec;

// Binds a network to a cell-centred field's mesh (geometry input channels).
UnetPrec *unet_prec_create(moons_unet_s *net, const moons_field_s *x, std::string *err);
// ... or to a grid given by its three node arrays (ghost nodes included).
UnetPrec *unet_prec_create_nodes(moons_unet_s *net, const std::vector<double> hn[3], std::string *err);
void unet_prec_destroy(UnetPrec *P);
// z0 = s N(f0 vol^(2/3) / s) on FP32 padded arrays; every kernel returns at
// once when *done is non-zero.
int unet_prec_apply(UnetPrec *P, const float *f0, float *z0, const int *done);

} // namespace moons
