// siafd_hip_test.cc -- stand-alone driver of the C++ host class SIAFD_HIP,
// in the spirit of src/base/stressbalance/sia/siafd_test.cc: it builds the
// 61x61x31 Vialov dome (isothermal Glen, A = 1e-16 Pa-3 a-1, flat bed,
// exactTestsABCDE.c:28-53), runs one full update on the GPU and reports the
// flux error against the steady-state closed form |Q| = M0 r / 2, D_max and
// the error path for a negative thickness. Exit status 0 means all checks passed.
#include "SIAFD_HIP.hh"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

using pism::stressbalance::RuntimeError;
using pism::stressbalance::SIAFD_HIP;

static void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "HIP error in %s: %s\n", what, hipGetErrorString(e));
    std::exit(2);
  }
}

int main() {
  const int Mx = 61, My = 61, Mz = 31;
  const double L = 750e3, dx = 2 * L / (Mx - 1), Lz = 4000.0, secpera = 3.15569259747e7;
  std::vector<double> z(Mz);
  for (int k = 0; k < Mz - 1; k++) z[k] = (Lz / (Mz - 1.0)) * k;
  z[Mz - 1] = Lz;
  sia_grid grid{Mx, My, Mz, 0, dx, dx, z.data()};
  sia_params P;
  sia_params_default(&P);
  P.sia_flow_law = SIA_FLOW_ISOTHERMAL_GLEN;
  P.ice_softness = 1e-16 / secpera;
  P.is_dry_simulation = 1;

  const size_t n2 = (size_t)Mx * My, n3 = n2 * Mz;
  std::vector<double> H(n2), b(n2, 0.0), h(n2), mask(n2), E(n3, 2009.0 * 30.0);
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      const double x = -L + i * dx, y = -L + j * dx, r = std::sqrt(x * x + y * y);
      const double t = 1.0 - std::pow(r / L, 4.0 / 3.0);
      const double HH = (r < L && t > 0) ? 3600.0 * std::pow(t, 3.0 / 8.0) : 0.0;
      H[i * My + j] = HH;
      h[i * My + j] = HH;
      mask[i * My + j] = HH < 0.01 ? 0.0 : 2.0;
    }
  }
  double *dH, *db, *dh, *dm, *dE;
  hip_check(hipMalloc((void **)&dH, n2 * 8), "malloc");
  hip_check(hipMalloc((void **)&db, n2 * 8), "malloc");
  hip_check(hipMalloc((void **)&dh, n2 * 8), "malloc");
  hip_check(hipMalloc((void **)&dm, n2 * 8), "malloc");
  hip_check(hipMalloc((void **)&dE, n3 * 8), "malloc");
  hip_check(hipMemcpy(dH, H.data(), n2 * 8, hipMemcpyHostToDevice), "copy");
  hip_check(hipMemcpy(db, b.data(), n2 * 8, hipMemcpyHostToDevice), "copy");
  hip_check(hipMemcpy(dh, h.data(), n2 * 8, hipMemcpyHostToDevice), "copy");
  hip_check(hipMemcpy(dm, mask.data(), n2 * 8, hipMemcpyHostToDevice), "copy");
  hip_check(hipMemcpy(dE, E.data(), n3 * 8, hipMemcpyHostToDevice), "copy");

  int failures = 0;
  {
    SIAFD_HIP sia(grid, P);
    sia.init();
    pism::stressbalance::SIAInputs in;
    in.land_ice_thickness = dH;
    in.bedrock_altitude = db;
    in.surface_altitude = dh;
    in.mask = dm;
    in.enthalpy = dE;
    sia.set_inputs(in);
    sia.update(pism::stressbalance::VelocityInput(), false);
    const double D_max = sia.max_diffusivity();
    std::vector<double> flux(2 * n2);
    hip_check(hipMemcpy(flux.data(), sia.diffusive_flux(), 2 * n2 * 8, hipMemcpyDeviceToHost), "copy");
    const double n = 3.0, rho = 910.0, g = 9.81;
    const double Gamma = 2 * std::pow(rho * g, n) * P.ice_softness / (n + 2);
    const double M0 = Gamma * std::pow(3600.0 / std::sqrt(L), 2 * n + 2) / std::pow(2.0, n - 1);
    double max_rel = 0.0;
    for (int i = 34; i < 56; i++) {
      const double xs = -L + i * dx + 0.5 * dx;
      const double Q = flux[2 * ((size_t)i * My + 30)];
      max_rel = std::fmax(max_rel, std::fabs(Q - M0 * xs / 2) / (M0 * xs / 2));
    }
    std::printf("SIAFD_HIP dome: D_max = %.6f m2/s, max rel. flux error vs M0 r/2 = %.4f\n", D_max, max_rel);
    if (!(D_max > 1.40 && D_max < 1.50)) failures++;
    if (!(max_rel < 1e-2)) failures++;
    const double *u = sia.velocity_u();
    if (!u) failures++;
    // error path: negative thickness (PISMBedSmoother.cc:312-314)
    H[30 * My + 30] = -1.0;
    hip_check(hipMemcpy(dH, H.data(), n2 * 8, hipMemcpyHostToDevice), "copy");
    bool thrown = false;
    try {
      sia.update(pism::stressbalance::VelocityInput(), true);
    } catch (RuntimeError &e) {
      thrown = (e.status() == SIA_ERR_NEGATIVE_THICKNESS);
      std::printf("expected error: %s\n", e.what());
    }
    if (!thrown) failures++;
  }
  hip_check(hipFree(dH), "free");
  hip_check(hipFree(db), "free");
  hip_check(hipFree(dh), "free");
  hip_check(hipFree(dm), "free");
  hip_check(hipFree(dE), "free");
  std::printf("%s\n", failures ? "FAILED" : "PASSED");
  return failures ? 1 : 0;
}
