// workload.hip -- device generator of the synthetic polythermal benchmark sheet
// (include/sia_workload.h). Mirrors pismv0.7.3_amd/workloads.py:greenland_like
// operation for operation so that host and device windows agree to rounding.
#include "../../include/sia_workload.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace {

constexpr uint64_t M1 = 0xBF58476D1CE4E5B9ull, M2 = 0x94D049BB133111EBull;
constexpr uint64_t G1 = 0x9E3779B97F4A7C15ull, G2 = 0xD1B54A32D192ED03ull;
enum { S_BUMP_A, S_BUMP_W, S_BUMP_CX, S_BUMP_CY, S_DOME, S_TEMPERATE, S_OMEGA, S_TSURF };

__host__ __device__ inline double uniform01(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t z = seed * G1 + stream * G2 + idx;
  z = (z ^ (z >> 30)) * M1;
  z = (z ^ (z >> 27)) * M2;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

struct GeomArgs {
  int Mx, My, Mx_full, My_full, i0, j0, n_bumps, is_dry;
  uint64_t seed;
  double dx, Lx, Ly, Hmax, x0f, y0f, ph1, ph2;
  double alpha_float, sea_level, icefree, floating;
  const double *bumps;  // [n_bumps][4]: a, w, cx, cy
};

__global__ void k_geometry(GeomArgs A, double *H, double *b, double *h, double *mask) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)A.Mx * A.My) return;
  const int il = (int)(q / A.My), jl = (int)(q % A.My);
  const double X = -A.Lx + (double)(A.i0 + il) * A.dx;
  const double Y = -A.Ly + (double)(A.j0 + jl) * A.dx;
  double bb = 0.0;
  for (int m = 0; m < A.n_bumps; m++) {
    const double a = A.bumps[4 * m], w = A.bumps[4 * m + 1], cx = A.bumps[4 * m + 2], cy = A.bumps[4 * m + 3];
    const double ddx = X - cx, ddy = Y - cy;
    const double d2 = ddx * ddx + ddy * ddy;
    bb += a * exp(-d2 / (2.0 * w * w));
  }
  const double Rx = 0.78 * A.Lx, Ry = 0.78 * A.Ly;
  const double xn = (X - A.x0f * A.Lx) / Rx, yn = (Y - A.y0f * A.Ly) / Ry;
  const double rho = sqrt(xn * xn + yn * yn);
  const double phi = atan2(yn, xn);
  const double sc = 1.0 + 0.06 * sin(3.0 * phi + A.ph1) + 0.04 * sin(5.0 * phi + A.ph2);
  const double r = rho / sc;
  double HH = 0.0;
  if (r < 1.0) {
    double t = 1.0 - pow(r, 4.0 / 3.0);
    if (t < 0.0) t = 0.0;
    HH = A.Hmax * pow(t, 0.375);
  }
  // GeometryCalculator (Mask.hh:82-118)
  const double hgrounded = bb + HH, hfloating = A.sea_level + A.alpha_float * HH;
  const bool floating = (hfloating > hgrounded + A.floating) && !A.is_dry;
  const bool ice_free = HH < A.icefree;
  H[q] = HH;
  b[q] = bb;
  h[q] = floating ? hfloating : hgrounded;
  mask[q] = floating ? (ice_free ? 4.0 : 3.0) : (ice_free ? 0.0 : 2.0);
}

struct EnthArgs {
  int Mx, My, Mz, My_full, i0, j0;
  uint64_t seed;
  double ci, T0, L, rho_g, beta, Tmelt, pair;
  const double *z;
};

__global__ void k_enthalpy(EnthArgs A, const double *__restrict__ H, double *__restrict__ E) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)A.Mx * A.My * A.Mz;
  if (e >= n) return;
  const long q = e / A.Mz;
  const int k = (int)(e % A.Mz);
  const int il = (int)(q / A.My), jl = (int)(q % A.My);
  const uint64_t col = (uint64_t)(A.i0 + il) * (uint64_t)A.My_full + (uint64_t)(A.j0 + jl);
  const double HH = H[q];
  const double f_t = 0.1 * uniform01(A.seed, S_TEMPERATE, col);
  const double omega = 0.01 * uniform01(A.seed, S_OMEGA, col);
  const double T_s = 238.15 + 10.0 * uniform01(A.seed, S_TSURF, col);
  const double z_t = f_t * HH;
  const double p_t = (HH - z_t > 0.0) ? A.pair + A.rho_g * (HH - z_t) : A.pair;
  const double T_base = A.Tmelt - A.beta * p_t - 1.0;
  const double E_air = A.ci * (T_s - A.T0);
  const double denom = (HH - z_t > 0.0) ? HH - z_t : 1.0;
  const double zk = A.z[k];
  const double depth = HH - zk;
  const double p = (depth > 0.0) ? A.pair + A.rho_g * depth : A.pair;
  const double E_cts = A.ci * ((A.Tmelt - A.beta * p) - A.T0);
  const double E_temp = E_cts + omega * A.L;
  double frac = depth / denom;
  frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
  const double T_cold = T_s + (T_base - T_s) * frac * frac;
  const double E_cold = A.ci * (T_cold - A.T0);
  double Ek = (zk <= z_t) ? E_temp : E_cold;
  if (zk > HH) Ek = E_air;
  E[e] = (HH > 0.0) ? Ek : E_air;
}

}  // namespace

extern "C" int sia_workload_generate(const sia_workload_spec *spec, const sia_grid *grid,
                                     const sia_params *params, double *H, double *b, double *h,
                                     double *mask, double *E, void *hip_stream) {
  if (!spec || !grid || !params || !H || !b || !h || !mask || !grid->z) return SIA_ERR_ARGUMENT;
  if (spec->i0 < 0 || spec->j0 < 0 || spec->i0 + grid->Mx > spec->Mx_full || spec->j0 + grid->My > spec->My_full)
    return SIA_ERR_ARGUMENT;
  hipStream_t st = (hipStream_t)hip_stream;
  const double Lx = 0.5 * (spec->Mx_full - 1) * spec->dx, Ly = 0.5 * (spec->My_full - 1) * spec->dx;
  std::vector<double> bumps(4 * (size_t)spec->n_bumps);
  for (int m = 0; m < spec->n_bumps; m++) {
    bumps[4 * m] = -500.0 + 2000.0 * uniform01(spec->seed, S_BUMP_A, m);
    bumps[4 * m + 1] = 5e3 + 75e3 * uniform01(spec->seed, S_BUMP_W, m);
    bumps[4 * m + 2] = -Lx + 2 * Lx * uniform01(spec->seed, S_BUMP_CX, m);
    bumps[4 * m + 3] = -Ly + 2 * Ly * uniform01(spec->seed, S_BUMP_CY, m);
  }
  double *d_bumps = nullptr, *d_z = nullptr;
  if (hipMalloc((void **)&d_bumps, sizeof(double) * bumps.size() + 8) != hipSuccess) return SIA_ERR_OUT_OF_MEMORY;
  if (hipMalloc((void **)&d_z, sizeof(double) * grid->Mz) != hipSuccess) {
    (void)hipFree(d_bumps);
    return SIA_ERR_OUT_OF_MEMORY;
  }
  int rc = SIA_OK;
  if (hipMemcpyAsync(d_bumps, bumps.data(), sizeof(double) * bumps.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_z, grid->z, sizeof(double) * grid->Mz, hipMemcpyHostToDevice, st) != hipSuccess) {
    rc = SIA_ERR_HIP;
  }
  if (rc == SIA_OK) {
    GeomArgs G{};
    G.Mx = grid->Mx;
    G.My = grid->My;
    G.Mx_full = spec->Mx_full;
    G.My_full = spec->My_full;
    G.i0 = spec->i0;
    G.j0 = spec->j0;
    G.n_bumps = spec->n_bumps;
    G.is_dry = params->is_dry_simulation;
    G.seed = spec->seed;
    G.dx = spec->dx;
    G.Lx = Lx;
    G.Ly = Ly;
    G.Hmax = spec->Hmax;
    G.x0f = -0.05 + 0.1 * uniform01(spec->seed, S_DOME, 0);
    G.y0f = -0.05 + 0.1 * uniform01(spec->seed, S_DOME, 1);
    G.ph1 = 2 * M_PI * uniform01(spec->seed, S_DOME, 2);
    G.ph2 = 2 * M_PI * uniform01(spec->seed, S_DOME, 3);
    G.alpha_float = 1 - params->ice_density / params->sea_water_density;
    G.sea_level = params->sea_level;
    G.icefree = params->mask_icefree_thickness_standard;
    G.floating = params->mask_is_floating_thickness_standard;
    G.bumps = d_bumps;
    const long n2 = (long)grid->Mx * grid->My;
    hipLaunchKernelGGL(k_geometry, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, G, H, b, h, mask);
    if (E) {
      EnthArgs A{};
      A.Mx = grid->Mx;
      A.My = grid->My;
      A.Mz = grid->Mz;
      A.My_full = spec->My_full;
      A.i0 = spec->i0;
      A.j0 = spec->j0;
      A.seed = spec->seed;
      A.ci = params->ice_specific_heat_capacity;
      A.T0 = params->enthalpy_converter_reference_temperature;
      A.L = params->water_latent_heat_fusion;
      A.rho_g = params->ice_density * params->standard_gravity;
      A.beta = params->beta_CC;
      A.Tmelt = params->water_melting_point_temperature;
      A.pair = params->surface_pressure;
      A.z = d_z;
      const long n3 = n2 * grid->Mz;
      hipLaunchKernelGGL(k_enthalpy, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, A, H, E);
    }
    if (hipGetLastError() != hipSuccess) rc = SIA_ERR_HIP;
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = SIA_ERR_HIP;
  (void)hipFree(d_bumps);
  (void)hipFree(d_z);
  return rc;
}
