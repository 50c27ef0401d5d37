// sia_host.hpp -- host-side helpers shared by the C ABI's source files.
#pragma once

#include "../../include/pism_sia.h"
#include "sia_device.hpp"

namespace pism_sia {

// DevParams of a grid and parameter set: every constant the kernels use,
// precomputed on the host (the bed smoother's half-widths are left at -1)
void make_dev_params(const sia_grid &g, const sia_params &p, DevParams &d);
// sia_create's argument checks; SIA_OK, SIA_ERR_ARGUMENT or SIA_ERR_CONFIG
int check_grid_params(const sia_grid *grid, const sia_params *params);
// BedSmoother's half-widths into d.Nx, d.Ny (-1: no smoothing) and its moment
// scales s2, s3, s4 into s; SIA_ERR_CONFIG when the box exceeds the grid
int bed_smoother_setup(const sia_grid &g, const sia_params &p, DevParams &d, double s[3]);
// FlowLawUsesGrainSize (flowlaws.cc:33-42): only the Goldsby-Kohlstedt law
// changes with grain size among the laws of flowlaw_factory.cc:77-83
inline bool law_uses_grain_size(int law) { return law == SIA_FLOW_GK; }

}  // namespace pism_sia
