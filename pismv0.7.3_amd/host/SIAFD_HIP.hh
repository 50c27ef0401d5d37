// SIAFD_HIP.hh -- C++ host class over the C ABI (include/pism_sia.h) that
// re-creates the SSB_Modifier / SIAFD surface of PISM v0.7.3
// (src/base/stressbalance/SSB_Modifier.hh:37-67, sia/SIAFD.hh:51-116), so a
// caller written against SIAFD can swap it in. Status codes become exceptions
// with context, as ParallelSection::check / RuntimeError do in the reference
// (error_handling.cc:153-192, PISMStressBalance.cc:112-115).
#ifndef PISM_SIAFD_HIP_HH
#define PISM_SIAFD_HIP_HH

#include <pism_sia.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace pism {
namespace stressbalance {

class RuntimeError : public std::runtime_error {
public:
  RuntimeError(int status, const std::string &msg) : std::runtime_error(msg), m_status(status) {}
  int status() const { return m_status; }

private:
  int m_status;
};

//! The registry fields SIAFD::update reads (PISMVars names, SIAFD.cc:122, 292, 552-556, 605).
struct SIAInputs {
  const double *land_ice_thickness = nullptr;
  const double *bedrock_altitude = nullptr;
  const double *surface_altitude = nullptr;
  const double *mask = nullptr;
  const double *enthalpy = nullptr;
  const double *age = nullptr;
  const double *basal_melt_rate = nullptr;
  long long bed_state_counter = 0;
};

//! 2-D sliding velocity (IceModelVec2V vel_input); both NULL = ZeroSliding.
struct VelocityInput {
  const double *u = nullptr, *v = nullptr;
};

class SIAFD_HIP {
public:
  //! SIAFD::SIAFD(grid, EC) (SIAFD.cc:35-75); allocates on the current HIP device.
  SIAFD_HIP(const sia_grid &grid, const sia_params &params) {
    check(sia_create(&grid, &params, &m_h), "creating the SIA stress balance modifier");
  }
  ~SIAFD_HIP() { sia_destroy(m_h); }
  SIAFD_HIP(const SIAFD_HIP &) = delete;
  SIAFD_HIP &operator=(const SIAFD_HIP &) = delete;

  //! SIAFD::init (SIAFD.cc:86-98): the first update preprocesses the bed.
  void init() {}

  void set_stream(void *hip_stream) { check(sia_set_stream(m_h, hip_stream), "setting the stream"); }

  //! Registry of inputs (m_grid->variables() in the reference).
  void set_inputs(const SIAInputs &in) { m_in = in; }

  //! SSB_Modifier::update(vel_input, fast) (SSB_Modifier.hh:44, SIAFD.cc:119-148).
  void update(const VelocityInput &vel_input, bool fast) {
    sia_inputs in{};
    in.land_ice_thickness = m_in.land_ice_thickness;
    in.bedrock_altitude = m_in.bedrock_altitude;
    in.surface_altitude = m_in.surface_altitude;
    in.mask = m_in.mask;
    in.enthalpy = m_in.enthalpy;
    in.age = m_in.age;
    in.vel_input_u = vel_input.u;
    in.vel_input_v = vel_input.v;
    in.basal_melt_rate = m_in.basal_melt_rate;
    in.bed_state_counter = m_in.bed_state_counter;
    check(sia_update(m_h, &in, fast ? 1 : 0), "updating the stress balance");
  }

  //! SSB_Modifier::diffusive_flux() (SSB_Modifier.hh:47): device [Mx][My][2].
  const double *diffusive_flux() {
    const double *p = nullptr;
    check(sia_diffusive_flux(m_h, &p), "diffusive_flux");
    return p;
  }
  //! SSB_Modifier::max_diffusivity() (SSB_Modifier.hh:50), m2 s-1.
  double max_diffusivity() {
    double d = 0.0;
    check(sia_max_diffusivity(m_h, &d), "max_diffusivity");
    return d;
  }
  //! SSB_Modifier::velocity_u()/velocity_v() (SSB_Modifier.hh:52-54): device [Mx][My][Mz].
  const double *velocity_u() { return field3(sia_velocity_u, "velocity_u"); }
  const double *velocity_v() { return field3(sia_velocity_v, "velocity_v"); }
  //! StressBalance::velocity_w() (PISMStressBalance.cc:138-140).
  const double *velocity_w() { return field3(sia_velocity_w, "velocity_w"); }
  // StressBalance::volumetric_strain_heating() (PISMStressBalance.cc:146-148)
  const double *volumetric_strain_heating() {
    return field3(sia_volumetric_strain_heating, "volumetric_strain_heating");
  }

  //! SSB_Modifier::stdout_report() (SSB_Modifier.hh:58).
  std::string stdout_report() { return ""; }

  sia_handle *handle() { return m_h; }

private:
  typedef int (*Getter)(sia_handle *, const double **);
  const double *field3(Getter g, const char *what) {
    const double *p = nullptr;
    check(g(m_h, &p), what);
    return p;
  }
  void check(int status, const char *context) {
    if (status != SIA_OK) {
      std::string msg = std::string(m_h ? sia_last_error(m_h) : "") + "\n  while " + context;
      throw RuntimeError(status, msg);
    }
  }

  sia_handle *m_h = nullptr;
  SIAInputs m_in;
};

}  // namespace stressbalance
}  // namespace pism

#endif
