// sia_noalloc.hip -- sia_flux_update_noalloc: the flux-only SIA update of a
// handle-free caller, with every intermediate field in a fixed device array
// of this library instead of memory allocated per call or per handle.
//
// The kernels are the ones sia_update(fast = 1) runs (bed smoother, gradient,
// SIAFD_Regional's override, theta/thksmooth, the FP64 column kernel and the
// D_max reduction), so the flux and D_max are bitwise those of sia_update +
// sia_diffusive_flux + sia_max_diffusivity_device on the same inputs.
#include "../../include/pism_sia.h"
#include "sia_host.hpp"
#include "sia_kernels.hpp"

#include <mutex>

using namespace pism_sia;

namespace {

// Layout of the scratch array, in doubles, for a grid of at most
// SIA_NOALLOC_MAX_COLUMNS map columns and MAX_MZ levels:
//   z [MAX_MZ] | topgsmooth, maxtl, C2, C3, C4, theta, thksmooth [n2 each] |
//   h_x, h_y, D_stag [2 n2 each] | D_max partials [n2 / 32] | 3 words
// (CFL bound, accumulated flags, flags when the caller passes none).
constexpr long MAXC = SIA_NOALLOC_MAX_COLUMNS;
constexpr long N_PART = MAXC / 32;  // at least the column blocks (2 n2 staggered columns, 128 or more per block)
constexpr long SCRATCH_DOUBLES = MAX_MZ + 7 * MAXC + 3 * 2 * MAXC + N_PART + 8;

}  // namespace

// The library's scratch memory: one copy per GPU, held until the process ends.
// Host code never uses this name as an address: it takes the device address
// from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[MAX_DEVICES] = {nullptr};

// device address of this device's scratch array (looked up once per device)
int scratch_for_device(int dev, double **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return SIA_ERR_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SIA_ERR_HIP;
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return SIA_OK;
}

}  // namespace

extern "C" int sia_flux_update_noalloc(const sia_grid *grid, const sia_params *params, const sia_inputs *in,
                                       double *flux, double *D_max, int32_t *err_flags, void *hip_stream) {
  if (!grid || !params || !in || !flux || !D_max) return SIA_ERR_ARGUMENT;
  int rc = check_grid_params(grid, params);
  if (rc) return rc;
  const long n2 = (long)grid->Mx * grid->My;
  if (n2 > MAXC) return SIA_ERR_ARGUMENT;  // larger than the scratch array holds
  if (!in->land_ice_thickness || !in->bedrock_altitude || !in->surface_altitude || !in->mask || !in->enthalpy) {
    return SIA_ERR_ARGUMENT;
  }
  if (params->compute_grain_size_using_age && !params->do_age) return SIA_ERR_CONFIG;  // SIAFD.cc:570-573
  const bool use_age = law_uses_grain_size(params->sia_flow_law) && params->compute_grain_size_using_age &&
                       params->do_age;
  if (use_age && !in->age) return SIA_ERR_ARGUMENT;
  DevParams P;
  make_dev_params(*grid, *params, P);
  P.use_age = use_age ? 1 : 0;
  double sc[3] = {0.0, 0.0, 0.0};
  if ((rc = bed_smoother_setup(*grid, *params, P, sc))) return rc;

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SIA_ERR_HIP;
  double *base = nullptr;
  if ((rc = scratch_for_device(dev, &base))) return rc;
  double *d_z = base;
  double *topgs = d_z + MAX_MZ, *maxtl = topgs + n2, *C2 = maxtl + n2, *C3 = C2 + n2, *C4 = C3 + n2;
  double *theta = C4 + n2, *thks = theta + n2;
  double *hx = thks + n2, *hy = hx + 2 * n2, *Dstag = hy + 2 * n2;
  double *dmax_part = Dstag + 2 * n2;
  unsigned long long *words = (unsigned long long *)(dmax_part + N_PART);
  unsigned long long *cfl_bits = words;
  int *err_acc = (int *)(words + 1);
  int *err = err_flags ? (int *)err_flags : (int *)(words + 2);

  hipStream_t st = (hipStream_t)hip_stream;
  if (hipMemcpyAsync(d_z, grid->z, sizeof(double) * (size_t)grid->Mz, hipMemcpyHostToDevice, st) != hipSuccess) {
    return SIA_ERR_HIP;
  }
  // D_max is reduced straight into the caller's scalar (its bits are the double)
  unsigned long long *dmax_bits = (unsigned long long *)D_max;
  launch_reset(dmax_bits, err, err_acc, cfl_bits, P.max_dt, st);
  // BedSmoother::preprocess_bed on every call: the entry point keeps no state
  if (P.Nx < 0) {
    if (hipMemcpyAsync(topgs, in->bedrock_altitude, sizeof(double) * (size_t)n2, hipMemcpyDeviceToDevice, st) !=
        hipSuccess) {
      return SIA_ERR_HIP;
    }
    launch_fill(maxtl, n2, 0.0, st);
    launch_fill(C2, n2, 0.0, st);
    launch_fill(C3, n2, 0.0, st);
    launch_fill(C4, n2, 0.0, st);
  } else {
    launch_bed_smoother(P, sc[0], sc[1], sc[2], in->bedrock_altitude, topgs, maxtl, C2, C3, C4, st);
  }
  const ThetaArgs T{topgs, maxtl, C2, C3, C4, theta, thks, err};
  launch_gradient(P, in->land_ice_thickness, in->bedrock_altitude, in->surface_altitude, in->mask, hx, hy, T, st);
  if (in->no_model_mask && in->usurfstore) {  // SIAFD_Regional (regional.cc:45-109)
    launch_regional_gradient(P, in->no_model_mask, in->usurfstore, hx, hy, st);
  }
  ColumnArgs A{};
  A.z = d_z;
  A.E = in->enthalpy;
  A.age = use_age ? in->age : nullptr;
  A.thks = thks;
  A.theta = theta;
  A.hx = hx;
  A.hy = hy;
  A.flux = flux;
  A.Dstag = Dstag;
  A.dmax_bits = dmax_bits;
  A.dmax_part = dmax_part;
  A.err = err;
  launch_column(P, A, false, st, 0, false);  // staggered order: its partials fit N_PART
  return hipGetLastError() == hipSuccess ? SIA_OK : SIA_ERR_HIP;
}
