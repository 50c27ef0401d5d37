/*
 * sia_oracle.c -- CPU oracle for the SIA hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Plain C99 restatement of PISM v0.7.3 (paths relative to the reference root).
 * Each function names the lines it follows and keeps their order of
 * floating-point operations. Built with -ffp-contract=off so that no FMA
 * contraction changes the rounding of the reference's expressions.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 */
#include "sia_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

static int fail(int code, const char *msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

const char *or_last_error(void) { return g_err; }

/* periodic index helpers: PISM's DMDA ghosts wrap (IceGrid.cc:807-819) */
static inline int wrap(int i, int M) {
  i %= M;
  return i < 0 ? i + M : i;
}
static inline size_t I2(const sia_grid *g, int i, int j) {
  return (size_t)wrap(i, g->Mx) * (size_t)g->My + (size_t)wrap(j, g->My);
}
static inline size_t IS(const sia_grid *g, int i, int j, int o) { return I2(g, i, j) * 2 + (size_t)o; }
static inline size_t I3(const sia_grid *g, int i, int j) { return I2(g, i, j) * (size_t)g->Mz; }

/* IceModelVec2Int::as_int (IceModelVec_inline.hh:94-100) */
static inline int as_int(double a) { return (int)floor(a + 0.5); }

/* mask predicates (Mask.hh:28-65) */
enum { M_FREE_BEDROCK = 0, M_GROUNDED = 2, M_FLOATING = 3, M_FREE_OCEAN = 4 };
static inline int m_ocean(int M) { return M >= M_FLOATING; }
static inline int m_grounded(int M) { return !m_ocean(M); }
static inline int m_icy(int M) { return M == M_GROUNDED || M == M_FLOATING; }
static inline int m_grounded_ice(int M) { return m_icy(M) && m_grounded(M); }
static inline int m_floating_ice(int M) { return m_icy(M) && m_ocean(M); }
static inline int m_ice_free(int M) { return !m_icy(M); }
static inline int m_ice_free_ocean(int M) { return m_ocean(M) && m_ice_free(M); }
static inline int m_ice_free_land(int M) { return m_grounded(M) && m_ice_free(M); }

/* ------------------------------------------------------------------------ */
/* Enthalpy converter: enthalpyConverter.cc:56-357 (default, cold, Kirchhoff) */

/* EnthalpyConverter::pressure, enthalpyConverter.cc:92-98 */
double or_ec_pressure(const sia_params *P, double depth) {
  if (depth > 0.0) {
    return P->surface_pressure + P->ice_density * P->standard_gravity * depth;
  }
  return P->surface_pressure;
}

/* melting_temperature_impl: enthalpyConverter.cc:128-130; cold: 347-349 */
double or_ec_melting_temperature(const sia_params *P, double p) {
  if (P->enthalpy_converter == SIA_EC_COLD) {
    return P->water_melting_point_temperature;
  }
  return P->water_melting_point_temperature - P->beta_CC * p;
}

/* L_impl: enthalpyConverter.cc:115-118; Kirchhoff: 433-435 */
static double ec_L(const sia_params *P, double T_pm) {
  if (P->enthalpy_converter == SIA_EC_KIRCHHOFF) {
    return P->water_latent_heat_fusion +
           (P->water_specific_heat_capacity - P->ice_specific_heat_capacity) * (T_pm - 273.15);
  }
  return P->water_latent_heat_fusion;
}

/* enthalpy_cts_impl: enthalpyConverter.cc:141-143 */
double or_ec_enthalpy_cts(const sia_params *P, double p) {
  return P->ice_specific_heat_capacity *
         (or_ec_melting_temperature(P, p) - P->enthalpy_converter_reference_temperature);
}

/* temperature_impl: enthalpyConverter.cc:175-189; cold: 355-357 */
double or_ec_temperature(const sia_params *P, double E, double p) {
  if (P->enthalpy_converter == SIA_EC_COLD) {
    return (E / P->ice_specific_heat_capacity) + P->enthalpy_converter_reference_temperature;
  }
  if (E < or_ec_enthalpy_cts(P, p)) {
    return (E / P->ice_specific_heat_capacity) + P->enthalpy_converter_reference_temperature;
  }
  return or_ec_melting_temperature(P, p);
}

/* enthalpyConverter.cc:197-199 */
double or_ec_pressure_adjusted_temperature(const sia_params *P, double E, double p) {
  return or_ec_temperature(P, E, p) - or_ec_melting_temperature(P, p) +
         P->water_melting_point_temperature;
}

/* water_fraction_impl: enthalpyConverter.cc:219-234; cold: 342-345 */
double or_ec_water_fraction(const sia_params *P, double E, double p) {
  if (P->enthalpy_converter == SIA_EC_COLD) {
    return 0.0;
  }
  double E_s = or_ec_enthalpy_cts(P, p);
  if (E <= E_s) {
    return 0.0;
  }
  return (E - E_s) / ec_L(P, or_ec_melting_temperature(P, p));
}

/* enthalpy_impl: enthalpyConverter.cc:256-281; cold: 337-340 */
double or_ec_enthalpy(const sia_params *P, double T, double omega, double p) {
  if (P->enthalpy_converter == SIA_EC_COLD) {
    return P->ice_specific_heat_capacity * (T - P->enthalpy_converter_reference_temperature);
  }
  const double T_melting = or_ec_melting_temperature(P, p);
  if (T < T_melting) {
    return P->ice_specific_heat_capacity * (T - P->enthalpy_converter_reference_temperature);
  }
  return or_ec_enthalpy_cts(P, p) + omega * ec_L(P, T_melting);
}

/* EnthalpyConverter::enthalpy_permissive (enthalpyConverter.cc:307-320; cold: 331-335) */
double or_ec_enthalpy_permissive(const sia_params *P, double T, double omega, double p) {
  if (P->enthalpy_converter == SIA_EC_COLD) {
    return P->ice_specific_heat_capacity * (T - P->enthalpy_converter_reference_temperature);
  }
  const double T_m = or_ec_melting_temperature(P, p);
  if (T < T_m) return or_ec_enthalpy(P, T, 0.0, p);
  return or_ec_enthalpy(P, T_m, fmax(0.0, fmin(omega, 1.0)), p);
}

/* is_temperate_impl with do_cold_ice_methods = no: enthalpyConverter.cc:156-162; cold: 351-353 */
int or_ec_is_temperate(const sia_params *P, double E, double p) {
  if (P->enthalpy_converter == SIA_EC_COLD) {
    return 0;
  }
  return E >= or_ec_enthalpy_cts(P, p);
}

/* ------------------------------------------------------------------------ */
/* Flow laws */

/* FlowLaw::softness_parameter_paterson_budd, flowlaws.cc:90-95 */
static double softness_pb(const sia_params *P, double T_pa) {
  if (T_pa < P->Paterson_Budd_critical_temperature) {
    return P->Paterson_Budd_A_cold * exp(-P->Paterson_Budd_Q_cold / (P->ideal_gas_constant * T_pa));
  }
  return P->Paterson_Budd_A_warm * exp(-P->Paterson_Budd_Q_warm / (P->ideal_gas_constant * T_pa));
}

/* m_beta_CC_grad / (m_rho * m_standard_gravity), flowlaws.cc:71 and 242 */
static double pb_pressure_coeff(const sia_params *P) {
  const double beta_CC_grad = P->beta_CC * P->ice_density * P->standard_gravity;
  return beta_CC_grad / (P->ice_density * P->standard_gravity);
}

/* GPBLD::softness_parameter, flowlaws.cc:209-221 */
static double gpbld_softness(const sia_params *P, double E, double p) {
  const double E_s = or_ec_enthalpy_cts(P, p);
  if (E < E_s) {
    double T_pa = or_ec_pressure_adjusted_temperature(P, E, p);
    return softness_pb(P, T_pa);
  }
  double omega = or_ec_water_fraction(P, E, p);
  /* std::min(omega, limit) */
  if (P->gpbld_water_frac_observed_limit < omega) {
    omega = P->gpbld_water_frac_observed_limit;
  }
  return softness_pb(P, P->water_melting_point_temperature) *
         (1.0 + P->gpbld_water_frac_coeff * omega);
}

/* Hooke::softness_parameter_from_temp, flowlaws.cc:267-270 */
static double hooke_softness(const sia_params *P, double T_pa) {
  return P->Hooke_A * exp(-P->Hooke_Q / (P->ideal_gas_constant * T_pa) +
                          3.0 * P->Hooke_C * pow(P->Hooke_Tr - T_pa, -P->Hooke_k));
}

/* GoldsbyKohlstedt::flow_from_temp, flowlaws.cc:274-312 (constants), 366-402 */
static double gk_flow_from_temp(const sia_params *P, double stress, double temp,
                                double pressure, double gs) {
  const double V_act_vol = -13.e-6;
  const double disl_crit_temp = 258.0, disl_A_cold = 4.0e-19, disl_A_warm = 6.0e4,
               disl_n = 4.0, disl_Q_cold = 60.e3, disl_Q_warm = 180.e3;
  const double gbs_crit_temp = 255.0, gbs_A_cold = 6.1811e-14, gbs_A_warm = 4.7547e15,
               gbs_n = 1.8, gbs_Q_cold = 49.e3, gbs_Q_warm = 192.e3, p_grain_sz_exp = 1.4;
  const double basal_A = 2.1896e-7, basal_n = 2.4, basal_Q = 60.e3;
  const double diff_crit_temp = 258.0, diff_V_m = 1.97e-5, diff_D_0v = 9.10e-4,
               diff_Q_v = 59.4e3, diff_D_0b = 5.8e-4, diff_Q_b = 49.e3, diff_delta = 9.04e-10;
  const double R = P->ideal_gas_constant;
  double eps_diff, eps_disl, eps_basal, eps_gbs, diff_D_b;

  if (fabs(stress) < 1e-10) {
    return 0;
  }
  const double beta_CC_grad = P->beta_CC * P->ice_density * P->standard_gravity;
  const double T = temp + (beta_CC_grad / (P->ice_density * P->standard_gravity)) * pressure;
  const double pV = pressure * V_act_vol;
  const double RT = R * T;
  const double diff_D_v = diff_D_0v * exp(-diff_Q_v / RT);
  diff_D_b = diff_D_0b * exp(-diff_Q_b / RT);
  if (T > diff_crit_temp) {
    diff_D_b *= 1000;
  }
  eps_diff = 14 * diff_V_m * (diff_D_v + M_PI * diff_delta * diff_D_b / gs) / (RT * (gs * gs));
  if (T > disl_crit_temp) {
    eps_disl = disl_A_warm * pow(stress, disl_n - 1) * exp(-(disl_Q_warm + pV) / RT);
  } else {
    eps_disl = disl_A_cold * pow(stress, disl_n - 1) * exp(-(disl_Q_cold + pV) / RT);
  }
  eps_basal = basal_A * pow(stress, basal_n - 1) * exp(-(basal_Q + pV) / RT);
  if (T > gbs_crit_temp) {
    eps_gbs = gbs_A_warm * (pow(stress, gbs_n - 1) / pow(gs, p_grain_sz_exp)) *
              exp(-(gbs_Q_warm + pV) / RT);
  } else {
    eps_gbs = gbs_A_cold * (pow(stress, gbs_n - 1) / pow(gs, p_grain_sz_exp)) *
              exp(-(gbs_Q_cold + pV) / RT);
  }
  return eps_diff + eps_disl + (eps_basal * eps_gbs) / (eps_basal + eps_gbs);
}

/* FlowLaw::flow and overrides: flowlaws.cc:98-101 (base), 232-244 (PB),
   flowlaws.hh:227-229 (isothermal Glen), flowlaws.hh:296-305 (arr, arrwarm),
   flowlaws.cc:314-318 (GK). */
double or_flow(const sia_params *P, double stress, double E, double p, double gs) {
  const double n = P->sia_Glen_exponent;
  switch (P->sia_flow_law) {
  case SIA_FLOW_ISOTHERMAL_GLEN:
    return P->ice_softness * pow(stress, n - 1);
  case SIA_FLOW_GPBLD:
    return gpbld_softness(P, E, p) * pow(stress, n - 1);
  case SIA_FLOW_PB: {
    const double temp = or_ec_temperature(P, E, p);
    const double T_pa = temp + pb_pressure_coeff(P) * p;
    return softness_pb(P, T_pa) * pow(stress, n - 1);
  }
  case SIA_FLOW_HOOKE: {
    const double temp = or_ec_temperature(P, E, p);
    const double T_pa = temp + pb_pressure_coeff(P) * p;
    return hooke_softness(P, T_pa) * pow(stress, n - 1);
  }
  case SIA_FLOW_ARR: {
    const double temp = or_ec_temperature(P, E, p);
    return P->Paterson_Budd_A_cold * exp(-P->Paterson_Budd_Q_cold / (P->ideal_gas_constant * temp)) *
           pow(stress, n - 1);
  }
  case SIA_FLOW_ARRWARM: {
    const double temp = or_ec_temperature(P, E, p);
    return P->Paterson_Budd_A_warm * exp(-P->Paterson_Budd_Q_warm / (P->ideal_gas_constant * temp)) *
           pow(stress, n - 1);
  }
  case SIA_FLOW_GK: {
    const double temp = or_ec_temperature(P, E, p);
    return gk_flow_from_temp(P, stress, temp, p, gs);
  }
  default:
    return NAN;
  }
}

/* FlowLaw::hardness_parameter = softness^(-1/n) (flowlaws.cc:76, 103-105) with
   each law's softness_parameter: PatersonBudd and its cold/warm variants use
   the converter's pressure-adjusted temperature (flowlaws.cc:226-229; hooke
   267-270; arr/arrwarm via flowlaws.hh:281-306), GPBLD its own
   (flowlaws.cc:209-221), isothermal Glen the constant B (flowlaws.hh:235-237,
   flowlaws.cc:249-253), Goldsby-Kohlstedt the Paterson-Budd relation
   (flowlaws.cc:336-351). */
double or_hardness(const sia_params *P, double E, double p) {
  const double hp = -1.0 / P->sia_Glen_exponent;
  const double R = P->ideal_gas_constant;
  switch (P->sia_flow_law) {
  case SIA_FLOW_ISOTHERMAL_GLEN:
    return pow(P->ice_softness, hp);
  case SIA_FLOW_GPBLD:
    return pow(gpbld_softness(P, E, p), hp);
  case SIA_FLOW_PB:
  case SIA_FLOW_GK:
    return pow(softness_pb(P, or_ec_pressure_adjusted_temperature(P, E, p)), hp);
  case SIA_FLOW_HOOKE:
    return pow(hooke_softness(P, or_ec_pressure_adjusted_temperature(P, E, p)), hp);
  case SIA_FLOW_ARR: {
    const double T_pa = or_ec_pressure_adjusted_temperature(P, E, p);
    return pow(P->Paterson_Budd_A_cold * exp(-P->Paterson_Budd_Q_cold / (R * T_pa)), hp);
  }
  case SIA_FLOW_ARRWARM: {
    const double T_pa = or_ec_pressure_adjusted_temperature(P, E, p);
    return pow(P->Paterson_Budd_A_warm * exp(-P->Paterson_Budd_Q_warm / (R * T_pa)), hp);
  }
  default:
    return NAN;
  }
}

/* FlowLawUsesGrainSize, flowlaws.cc:33-42 */
int or_flow_uses_grain_size(const sia_params *P) {
  static const double gs[] = {1e-4, 1e-3, 1e-2, 1}, s = 1e4, E = 400000, p = 1e6;
  double ref = or_flow(P, s, E, p, gs[0]);
  for (int i = 1; i < 4; i++) {
    if (or_flow(P, s, E, p, gs[i]) != ref) {
      return 1;
    }
  }
  return 0;
}

/* SIAFD::grainSizeVostok, SIAFD.cc:980-1018 (m_second_to_kiloyear, SIAFD.cc:69) */
double or_grain_size_vostok(const sia_params *P, double age_seconds) {
  enum { numPoints = 22 };
  static const double ageAt[numPoints] = {
      0.0000e+00, 5.0000e+01, 1.0000e+02, 1.2500e+02, 1.5000e+02, 1.5800e+02, 1.6500e+02, 1.7000e+02,
      1.8000e+02, 1.8800e+02, 2.0000e+02, 2.2500e+02, 2.4500e+02, 2.6000e+02, 3.0000e+02, 3.2000e+02,
      3.5000e+02, 4.0000e+02, 5.0000e+02, 6.0000e+02, 8.0000e+02, 1.0000e+04};
  static const double gsAt[numPoints] = {
      1.8000e-03, 2.2000e-03, 3.0000e-03, 4.0000e-03, 4.3000e-03, 3.0000e-03, 3.0000e-03, 4.6000e-03,
      3.4000e-03, 3.3000e-03, 5.9000e-03, 6.2000e-03, 5.4000e-03, 6.8000e-03, 3.5000e-03, 6.0000e-03,
      8.0000e-03, 8.3000e-03, 3.6000e-03, 3.8000e-03, 9.5000e-03, 1.0000e-02};
  const double second_to_kiloyear = 1.0 / (1000.0 * P->seconds_per_year);
  const double a = age_seconds * second_to_kiloyear;
  int l = 0, r = numPoints - 1;
  if (a < ageAt[l]) {
    return gsAt[l];
  } else if (a > ageAt[r]) {
    return gsAt[r];
  }
  while (r > l + 1) {
    const int j = (r + l) / 2;
    if (a < ageAt[j]) {
      r = j;
    } else {
      l = j;
    }
  }
  return gsAt[l] + (a - ageAt[l]) * (gsAt[r] - gsAt[l]) / (ageAt[r] - ageAt[l]);
}

/* ------------------------------------------------------------------------ */
/* IceGrid::kBelowHeight, IceGrid.cc:379-391. gsl_interp_accel_find returns the
   i with z[i] <= x < z[i+1], clamped to [0, Mz-2] (the GSL binary search
   never returns its upper bound). */
int or_kbelow(const double *z, int Mz, double height) {
  const double Lz = z[Mz - 1];
  if (height < 0.0 - 1.0e-6) {
    return -1;
  }
  if (height > Lz + 1.0e-6) {
    return -2;
  }
  int ilo = 0, ihi = Mz - 1;
  while (ihi > ilo + 1) {
    int i = (ihi + ilo) / 2;
    if (z[i] > height) {
      ihi = i;
    } else {
      ilo = i;
    }
  }
  return ilo;
}

/* ------------------------------------------------------------------------ */
/* BedSmoother */

/* preprocess_bed (PISMBedSmoother.cc:96-149), smooth_the_bed_on_proc0
   (162-193), compute_coefficients_on_proc0 (196-269). The box average does
   not wrap (in-grid points only). */
int or_bed_smoother(const sia_grid *g, const sia_params *P, const double *topg,
                    double *topgsmooth, double *maxtl, double *C2, double *C3,
                    double *C4, int *Nx_out, int *Ny_out) {
  const int Mx = g->Mx, My = g->My;
  const size_t N2 = (size_t)Mx * My;
  if (P->bed_smoother_range <= 0.0) {
    memcpy(topgsmooth, topg, N2 * sizeof(double));
    /* maxtl, C2..C4 keep their (zero) initial values; theta = 1 */
    memset(maxtl, 0, N2 * sizeof(double));
    memset(C2, 0, N2 * sizeof(double));
    memset(C3, 0, N2 * sizeof(double));
    memset(C4, 0, N2 * sizeof(double));
    *Nx_out = -1;
    *Ny_out = -1;
    return SIA_OK;
  }
  int Nx = (int)ceil(P->bed_smoother_range / g->dx);
  int Ny = (int)ceil(P->bed_smoother_range / g->dy);
  if (Nx < 1) Nx = 1;
  if (Ny < 1) Ny = 1;
  if (Nx >= Mx || Ny >= My) {
    return fail(SIA_ERR_CONFIG, "input Nx, Ny in bed smoother is too large because domain of "
                                "smoothing exceeds IceGrid domain");
  }
  *Nx_out = Nx;
  *Ny_out = Ny;

  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      double sum = 0.0, count = 0.0;
      for (int r = -Nx; r <= Nx; r++) {
        for (int s = -Ny; s <= Ny; s++) {
          if ((i + r >= 0) && (i + r < Mx) && (j + s >= 0) && (j + s < My)) {
            sum += topg[(size_t)(i + r) * My + (j + s)];
            count += 1.0;
          }
        }
      }
      topgsmooth[(size_t)i * My + j] = sum / count;
    }
  }

  const double n = P->sia_Glen_exponent, k = (n + 2) / n, s2 = k * (2 * n + 2) / (2 * n),
               s3 = s2 * (3 * n + 2) / (3 * n), s4 = s3 * (4 * n + 2) / (4 * n);

  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      const double topgs = topgsmooth[(size_t)i * My + j];
      double maxtltemp = 0.0, sum2 = 0.0, sum3 = 0.0, sum4 = 0.0, count = 0.0;
      for (int r = -Nx; r <= Nx; r++) {
        for (int s = -Ny; s <= Ny; s++) {
          if ((i + r >= 0) && (i + r < Mx) && (j + s >= 0) && (j + s < My)) {
            const double tl = topg[(size_t)(i + r) * My + (j + s)] - topgs;
            maxtltemp = (maxtltemp < tl) ? tl : maxtltemp; /* std::max */
            const double tl2 = tl * tl;
            sum2 += tl2;
            sum3 += tl2 * tl;
            sum4 += tl2 * tl2;
            count += 1.0;
          }
        }
      }
      const size_t q = (size_t)i * My + j;
      maxtl[q] = maxtltemp;
      /* c = sum/count, then VecScale(C, s) */
      C2[q] = (sum2 / count) * s2;
      C3[q] = (sum3 / count) * s3;
      C4[q] = (sum4 / count) * s4;
    }
  }
  return SIA_OK;
}

/* BedSmoother::get_theta, PISMBedSmoother.cc:360-420 */
int or_theta(const sia_grid *g, const sia_params *P, int Nx, int Ny, const double *usurf,
             const double *topgsmooth, const double *maxtl, const double *C2,
             const double *C3, const double *C4, double *theta) {
  const size_t N2 = (size_t)g->Mx * g->My;
  if (Nx < 0 || Ny < 0) {
    for (size_t q = 0; q < N2; q++) theta[q] = 1.0;
    return SIA_OK;
  }
  for (size_t q = 0; q < N2; q++) {
    const double H = usurf[q] - topgsmooth[q];
    if (H > maxtl[q]) {
      const double Hinv = 1.0 / ((H < 1.0) ? 1.0 : H); /* 1/std::max(H, 1.0) */
      double omega = 1.0 + Hinv * Hinv * (C2[q] + Hinv * (C3[q] + Hinv * C4[q]));
      if (omega <= 0) {
        return fail(SIA_ERR_OMEGA_NEGATIVE, "omega is negative in BedSmoother.get_theta()");
      }
      if (omega < 0.001) {
        omega = 0.001;
      }
      double t = pow(omega, -P->sia_Glen_exponent);
      if (t > 1.0) t = 1.0;
      if (t < 0.0) t = 0.0;
      theta[q] = t;
    } else {
      theta[q] = 0.00;
    }
  }
  return SIA_OK;
}

/* BedSmoother::get_smoothed_thk, PISMBedSmoother.cc:286-336 */
int or_thksmooth(const sia_grid *g, const double *usurf, const double *thk,
                 const double *mask, const double *maxtl, const double *topgsmooth,
                 double *result) {
  const size_t N2 = (size_t)g->Mx * g->My;
  for (size_t q = 0; q < N2; q++) {
    if (thk[q] < 0.0) {
      return fail(SIA_ERR_NEGATIVE_THICKNESS, "BedSmoother detects negative original thickness");
    } else if (thk[q] == 0.0) {
      result[q] = 0.0;
    } else if (maxtl[q] >= thk[q]) {
      result[q] = thk[q];
    } else {
      if (m_grounded(as_int(mask[q]))) {
        const double thks_try = usurf[q] - topgsmooth[q];
        result[q] = (thks_try > 0.0) ? thks_try : 0.0;
      } else {
        result[q] = thk[q];
      }
    }
  }
  return SIA_OK;
}

/* ------------------------------------------------------------------------ */
/* Surface gradient: SIAFD.cc:187-503 */

/* surface_gradient_eta, SIAFD.cc:210-284 (b.diff_*_stag*: iceModelVec2.cc:473-495) */
static int gradient_eta(const sia_grid *g, const sia_params *P, const double *H,
                        const double *b, double *h_x, double *h_y) {
  const int Mx = g->Mx, My = g->My;
  const double n = P->sia_Glen_exponent, etapow = (2.0 * n + 2.0) / n, invpow = 1.0 / etapow,
               dinvpow = (-n - 2.0) / (2.0 * n + 2.0);
  const double dx = g->dx, dy = g->dy;
  const size_t N2 = (size_t)Mx * My;
  double *eta = (double *)malloc(N2 * sizeof(double));
  if (!eta) return fail(SIA_ERR_OUT_OF_MEMORY, "out of memory");
  for (size_t q = 0; q < N2; q++) eta[q] = pow(H[q], etapow);
#define ETA(a, c) eta[I2(g, (a), (c))]
#define B(a, c) b[I2(g, (a), (c))]
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      /* o = 0 */
      {
        const double mean_eta = 0.5 * (ETA(i + 1, j) + ETA(i, j));
        double hx, hy;
        if (mean_eta > 0.0) {
          const double factor = invpow * pow(mean_eta, dinvpow);
          hx = factor * (ETA(i + 1, j) - ETA(i, j)) / dx;
          hy = factor * (+ETA(i + 1, j + 1) + ETA(i, j + 1) - ETA(i + 1, j - 1) - ETA(i, j - 1)) /
               (4.0 * dy);
        } else {
          hx = 0.0;
          hy = 0.0;
        }
        hx += (B(i + 1, j) - B(i, j)) / dx;
        hy += (B(i + 1, j + 1) + B(i, j + 1) - B(i + 1, j - 1) - B(i, j - 1)) / (4 * dy);
        h_x[IS(g, i, j, 0)] = hx;
        h_y[IS(g, i, j, 0)] = hy;
      }
      /* o = 1 */
      {
        const double mean_eta = 0.5 * (ETA(i, j + 1) + ETA(i, j));
        double hx, hy;
        if (mean_eta > 0.0) {
          const double factor = invpow * pow(mean_eta, dinvpow);
          hy = factor * (ETA(i, j + 1) - ETA(i, j)) / dy;
          hx = factor * (+ETA(i + 1, j + 1) + ETA(i + 1, j) - ETA(i - 1, j + 1) - ETA(i - 1, j)) /
               (4.0 * dx);
        } else {
          hy = 0.0;
          hx = 0.0;
        }
        hy += (B(i, j + 1) - B(i, j)) / dy;
        hx += (B(i + 1, j + 1) + B(i + 1, j) - B(i - 1, j + 1) - B(i - 1, j)) / (4 * dx);
        h_x[IS(g, i, j, 1)] = hx;
        h_y[IS(g, i, j, 1)] = hy;
      }
    }
  }
#undef ETA
#undef B
  free(eta);
  return SIA_OK;
}

/* surface_gradient_mahaffy, SIAFD.cc:289-317 */
static int gradient_mahaffy(const sia_grid *g, const double *h, double *h_x, double *h_y) {
  const double dx = g->dx, dy = g->dy;
#define HH(a, c) h[I2(g, (a), (c))]
  for (int i = 0; i < g->Mx; i++) {
    for (int j = 0; j < g->My; j++) {
      h_x[IS(g, i, j, 0)] = (HH(i + 1, j) - HH(i, j)) / dx;
      h_y[IS(g, i, j, 0)] = (+HH(i + 1, j + 1) + HH(i, j + 1) - HH(i + 1, j - 1) - HH(i, j - 1)) / (4.0 * dy);
      h_y[IS(g, i, j, 1)] = (HH(i, j + 1) - HH(i, j)) / dy;
      h_x[IS(g, i, j, 1)] = (+HH(i + 1, j + 1) + HH(i + 1, j) - HH(i - 1, j + 1) - HH(i - 1, j)) / (4.0 * dx);
    }
  }
#undef HH
  return SIA_OK;
}

/* surface_gradient_haseloff, SIAFD.cc:366-503 */
static int gradient_haseloff(const sia_grid *g, const double *h, const double *b,
                             const double *mask, double *h_x, double *h_y) {
  const int Mx = g->Mx, My = g->My;
  const double dx = g->dx, dy = g->dy;
  const size_t N2 = (size_t)Mx * My;
  double *w_i = (double *)malloc(N2 * sizeof(double));
  double *w_j = (double *)malloc(N2 * sizeof(double));
  if (!w_i || !w_j) {
    free(w_i);
    free(w_j);
    return fail(SIA_ERR_OUT_OF_MEMORY, "out of memory");
  }
#define HH(a, c) h[I2(g, (a), (c))]
#define BB(a, c) b[I2(g, (a), (c))]
#define MM(a, c) as_int(mask[I2(g, (a), (c))])
#define WI(a, c) w_i[I2(g, (a), (c))]
#define WJ(a, c) w_j[I2(g, (a), (c))]
#define HX(a, c, o) h_x[IS(g, (a), (c), (o))]
#define HY(a, c, o) h_y[IS(g, (a), (c), (o))]
  /* first loop (SIAFD.cc:399-439); ghost values are periodic copies */
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      {
        const int m0 = MM(i, j), m1 = MM(i + 1, j);
        if ((m_floating_ice(m0) && m_ice_free_ocean(m1)) || (m_ice_free_ocean(m0) && m_floating_ice(m1))) {
          HX(i, j, 0) = 0;
          WI(i, j) = 0;
        } else if ((m_icy(m0) && m_ice_free(m1) && BB(i + 1, j) > HH(i, j)) ||
                   (m_ice_free(m0) && m_icy(m1) && BB(i, j) > HH(i + 1, j))) {
          HX(i, j, 0) = 0.0;
          WI(i, j) = 0;
        } else {
          HX(i, j, 0) = (HH(i + 1, j) - HH(i, j)) / dx;
          WI(i, j) = 1;
        }
      }
      {
        const int m0 = MM(i, j), m1 = MM(i, j + 1);
        if ((m_floating_ice(m0) && m_ice_free_ocean(m1)) || (m_ice_free_ocean(m0) && m_floating_ice(m1))) {
          HY(i, j, 1) = 0.0;
          WJ(i, j) = 0.0;
        } else if ((m_icy(m0) && m_ice_free(m1) && BB(i, j + 1) > HH(i, j)) ||
                   (m_ice_free(m0) && m_icy(m1) && BB(i, j) > HH(i, j + 1))) {
          HY(i, j, 1) = 0.0;
          WJ(i, j) = 0.0;
        } else {
          HY(i, j, 1) = (HH(i, j + 1) - HH(i, j)) / dy;
          WJ(i, j) = 1.0;
        }
      }
    }
  }
  /* second loop (SIAFD.cc:441-499) */
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      {
        if (WJ(i, j) > 0) {
          double W = WI(i, j) + WI(i - 1, j) + WI(i - 1, j + 1) + WI(i, j + 1);
          if (W > 0) {
            HX(i, j, 1) = 1.0 / W * (HX(i, j, 0) + HX(i - 1, j, 0) + HX(i - 1, j + 1, 0) + HX(i, j + 1, 0));
          } else {
            HX(i, j, 1) = 0.0;
          }
        } else {
          if (m_icy(MM(i, j))) {
            double W = WI(i, j) + WI(i - 1, j);
            if (W > 0) {
              HX(i, j, 1) = 1.0 / W * (HX(i, j, 0) + HX(i - 1, j, 0));
            } else {
              HX(i, j, 1) = 0.0;
            }
          } else {
            double W = WI(i, j + 1) + WI(i - 1, j + 1);
            if (W > 0) {
              HX(i, j, 1) = 1.0 / W * (HX(i - 1, j + 1, 0) + HX(i, j + 1, 0));
            } else {
              HX(i, j, 1) = 0.0;
            }
          }
        }
      }
      {
        if (WI(i, j) > 0) {
          double W = WJ(i, j) + WJ(i, j - 1) + WJ(i + 1, j - 1) + WJ(i + 1, j);
          if (W > 0) {
            HY(i, j, 0) = 1.0 / W * (HY(i, j, 1) + HY(i, j - 1, 1) + HY(i + 1, j - 1, 1) + HY(i + 1, j, 1));
          } else {
            HY(i, j, 0) = 0.0;
          }
        } else {
          if (m_icy(MM(i, j))) {
            double W = WJ(i, j) + WJ(i, j - 1);
            if (W > 0) {
              HY(i, j, 0) = 1.0 / W * (HY(i, j, 1) + HY(i, j - 1, 1));
            } else {
              HY(i, j, 0) = 0.0;
            }
          } else {
            double W = WJ(i + 1, j - 1) + WJ(i + 1, j);
            if (W > 0) {
              HY(i, j, 0) = 1.0 / W * (HY(i + 1, j - 1, 1) + HY(i + 1, j, 1));
            } else {
              HY(i, j, 0) = 0.0;
            }
          }
        }
      }
    }
  }
#undef HH
#undef BB
#undef MM
#undef WI
#undef WJ
#undef HX
#undef HY
  free(w_i);
  free(w_j);
  return SIA_OK;
}

/* SIAFD::compute_surface_gradient, SIAFD.cc:187-207 */
int or_surface_gradient(const sia_grid *g, const sia_params *P, const double *H,
                        const double *b, const double *h, const double *mask,
                        double *h_x, double *h_y) {
  switch (P->surface_gradient_method) {
  case SIA_GRADIENT_ETA:
    return gradient_eta(g, P, H, b, h_x, h_y);
  case SIA_GRADIENT_HASELOFF:
    return gradient_haseloff(g, h, b, mask, h_x, h_y);
  case SIA_GRADIENT_MAHAFFY:
    return gradient_mahaffy(g, h, h_x, h_y);
  default:
    return fail(SIA_ERR_CONFIG, "value of surface_gradient_method is not valid");
  }
}

/* ------------------------------------------------------------------------ */
/* SIAFD::compute_diffusive_flux, SIAFD.cc:546-722. D_stag (optional) is D
   before the domain-edge override, which is what compute_diffusivity_staggered
   (SIAFD.cc:750-805) recomputes from the stored delta. */
int or_diffusive_flux(const sia_grid *g, const sia_params *P, const double *h_x,
                      const double *h_y, const double *thksmooth, const double *theta,
                      const double *E, const double *age, int full_update,
                      double *flux, double *D_stag, double *delta0, double *delta1,
                      double *D_max_out) {
  const int Mx = g->Mx, My = g->My, Mz = g->Mz;
  const double *z = g->z;
  const double enhancement_factor = P->sia_enhancement_factor;
  double ice_grain_size = P->ice_grain_size;

  if (P->sia_flow_law < 0 || P->sia_flow_law > SIA_FLOW_GK) {
    return fail(SIA_ERR_CONFIG, "Selected ice flow law is not available");
  }
  if (P->compute_grain_size_using_age && !P->do_age) {
    return fail(SIA_ERR_CONFIG, "SIAFD::compute_diffusive_flux(): do_age not set but age is "
                                "needed for grain-size-based flow law");
  }
  const int use_age = or_flow_uses_grain_size(P) && P->compute_grain_size_using_age && P->do_age;
  if (use_age && age == NULL) {
    return fail(SIA_ERR_ARGUMENT, "age field is required");
  }

  double *delta_ij = (double *)malloc((size_t)Mz * sizeof(double));
  if (!delta_ij) return fail(SIA_ERR_OUT_OF_MEMORY, "out of memory");

  double my_D_max = 0.0;
  for (int o = 0; o < 2; o++) {
    double *delta = (o == 0) ? delta0 : delta1;
    for (int i = 0; i < Mx; i++) {
      for (int j = 0; j < My; j++) {
        const int oi = 1 - o, oj = o;
        const double thk = 0.5 * (thksmooth[I2(g, i, j)] + thksmooth[I2(g, i + oi, j + oj)]);
        const size_t qs = IS(g, i, j, o);

        if (thk == 0.0) {
          flux[qs] = 0.0;
          if (D_stag) D_stag[qs] = 0.0;
          if (full_update) {
            for (int k = 0; k < Mz; k++) delta[I3(g, i, j) + k] = 0.0;
          }
          continue;
        }

        const double *E_ij = E + I3(g, i, j), *E_off = E + I3(g, i + oi, j + oj);
        const double *age_ij = use_age ? age + I3(g, i, j) : NULL;
        const double *age_off = use_age ? age + I3(g, i + oi, j + oj) : NULL;

        const double slope = (o == 0) ? h_x[qs] : h_y[qs];
        const int ks = or_kbelow(z, Mz, thk);
        if (ks < 0) {
          free(delta_ij);
          return fail(SIA_ERR_THICKNESS_ABOVE_LZ, "height is outside the computational grid");
        }
        const double alpha = sqrt(h_x[qs] * h_x[qs] + h_y[qs] * h_y[qs]);
        const double theta_local = 0.5 * (theta[I2(g, i, j)] + theta[I2(g, i + oi, j + oj)]);

        double Dfoffset = 0.0;
        for (int k = 0; k <= ks; ++k) {
          double depth = thk - z[k];
          const double pressure = or_ec_pressure(P, depth);
          if (use_age) {
            ice_grain_size = or_grain_size_vostok(P, 0.5 * (age_ij[k] + age_off[k]));
          }
          double Ek = 0.5 * (E_ij[k] + E_off[k]);
          double fl = or_flow(P, alpha * pressure, Ek, pressure, ice_grain_size);
          delta_ij[k] = enhancement_factor * theta_local * 2.0 * pressure * fl;
          if (k > 0) {
            const double dz = z[k] - z[k - 1];
            Dfoffset += 0.5 * dz * ((depth + dz) * delta_ij[k - 1] + depth * delta_ij[k]);
          }
        }
        {
          const double dz = thk - z[ks];
          Dfoffset += 0.5 * dz * dz * delta_ij[ks];
        }
        if (D_stag) D_stag[qs] = Dfoffset;

        if (i < 0 || i >= Mx - 1 || j < 0 || j >= My - 1) {
          Dfoffset = 0.0;
        }
        my_D_max = (my_D_max < Dfoffset) ? Dfoffset : my_D_max; /* std::max */
        flux[qs] = -Dfoffset * slope;

        if (full_update) {
          for (int k = ks + 1; k < Mz; ++k) delta_ij[k] = 0.0;
          memcpy(delta + I3(g, i, j), delta_ij, (size_t)Mz * sizeof(double));
        }
      }
    }
  }
  free(delta_ij);
  *D_max_out = my_D_max;
  return SIA_OK;
}

/* SIAFD::compute_I, SIAFD.cc:819-879 */
int or_compute_I(const sia_grid *g, const double *thksmooth, const double *delta0,
                 const double *delta1, double *I0, double *I1) {
  const int Mx = g->Mx, My = g->My, Mz = g->Mz;
  const double *z = g->z;
  for (int o = 0; o < 2; ++o) {
    const double *delta = (o == 0) ? delta0 : delta1;
    double *I = (o == 0) ? I0 : I1;
    for (int i = 0; i < Mx; i++) {
      for (int j = 0; j < My; j++) {
        const int oi = 1 - o, oj = o;
        const double thk = 0.5 * (thksmooth[I2(g, i, j)] + thksmooth[I2(g, i + oi, j + oj)]);
        const double *delta_ij = delta + I3(g, i, j);
        double *I_ij = I + I3(g, i, j);
        const int ks = or_kbelow(z, Mz, thk);
        if (ks < 0) {
          return fail(SIA_ERR_THICKNESS_ABOVE_LZ, "height is outside the computational grid");
        }
        I_ij[0] = 0.0;
        double I_current = 0.0;
        for (int k = 1; k <= ks; ++k) {
          const double dz = z[k] - z[k - 1];
          I_current += 0.5 * dz * (delta_ij[k - 1] + delta_ij[k]);
          I_ij[k] = I_current;
        }
        for (int k = ks + 1; k < Mz; ++k) {
          I_ij[k] = I_current;
        }
      }
    }
  }
  return SIA_OK;
}

/* SIAFD::compute_3d_horizontal_velocity, SIAFD.cc:899-963 */
int or_horizontal_velocity(const sia_grid *g, const double *h_x, const double *h_y,
                           const double *I0, const double *I1, const double *ub,
                           const double *vb, double *u, double *v) {
  const int Mx = g->Mx, My = g->My, Mz = g->Mz;
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      const double *I_e = I0 + I3(g, i, j), *I_w = I0 + I3(g, i - 1, j), *I_n = I1 + I3(g, i, j),
                   *I_s = I1 + I3(g, i, j - 1);
      double *u_ij = u + I3(g, i, j), *v_ij = v + I3(g, i, j);
      const double h_x_w = h_x[IS(g, i - 1, j, 0)], h_x_e = h_x[IS(g, i, j, 0)],
                   h_x_n = h_x[IS(g, i, j, 1)], h_x_s = h_x[IS(g, i, j - 1, 1)];
      const double h_y_w = h_y[IS(g, i - 1, j, 0)], h_y_e = h_y[IS(g, i, j, 0)],
                   h_y_n = h_y[IS(g, i, j, 1)], h_y_s = h_y[IS(g, i, j - 1, 1)];
      const double vel_input_u = ub ? ub[I2(g, i, j)] : 0.0;
      const double vel_input_v = vb ? vb[I2(g, i, j)] : 0.0;
      for (int k = 0; k < Mz; ++k) {
        u_ij[k] = -0.25 * (I_e[k] * h_x_e + I_w[k] * h_x_w + I_n[k] * h_x_n + I_s[k] * h_x_s);
        v_ij[k] = -0.25 * (I_e[k] * h_y_e + I_w[k] * h_y_w + I_n[k] * h_y_n + I_s[k] * h_y_s);
        u_ij[k] += vel_input_u;
        v_ij[k] += vel_input_v;
      }
    }
  }
  return SIA_OK;
}

/* StressBalance::compute_vertical_velocity, PISMStressBalance.cc:192-287 */
int or_vertical_velocity(const sia_grid *g, const double *mask, const double *u,
                         const double *v, const double *bmr, double *w) {
  const int Mx = g->Mx, My = g->My, Mz = g->Mz;
  const double *z = g->z;
#define MM(a, c) as_int(mask[I2(g, (a), (c))])
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      double *w_ij = w + I3(g, i, j);
      const double *u_w = u + I3(g, i - 1, j), *u_ij = u + I3(g, i, j), *u_e = u + I3(g, i + 1, j);
      const double *v_s = v + I3(g, i, j - 1), *v_ij = v + I3(g, i, j), *v_n = v + I3(g, i, j + 1);
      double west = 1, east = 1, south = 1, north = 1, D_x = 0, D_y = 0;
      const int m = MM(i, j);
      if ((m_icy(m) && m_ice_free(MM(i + 1, j))) || (m_ice_free(m) && m_icy(MM(i + 1, j)))) east = 0;
      if ((m_icy(m) && m_ice_free(MM(i - 1, j))) || (m_ice_free(m) && m_icy(MM(i - 1, j)))) west = 0;
      if (east + west > 0) {
        D_x = 1.0 / (g->dx * (east + west));
      } else {
        D_x = 0.0;
      }
      if ((m_icy(m) && m_ice_free(MM(i, j + 1))) || (m_ice_free(m) && m_icy(MM(i, j + 1)))) north = 0;
      if ((m_icy(m) && m_ice_free(MM(i, j - 1))) || (m_ice_free(m) && m_icy(MM(i, j - 1)))) south = 0;
      if (north + south > 0) {
        D_y = 1.0 / (g->dy * (north + south));
      } else {
        D_y = 0.0;
      }
      double u_x = D_x * (west * (u_ij[0] - u_w[0]) + east * (u_e[0] - u_ij[0])),
             v_y = D_y * (south * (v_ij[0] - v_s[0]) + north * (v_n[0] - v_ij[0]));
      if (bmr != NULL) {
        w_ij[0] = -bmr[I2(g, i, j)];
      } else {
        w_ij[0] = 0.0;
      }
      double old_integrand = u_x + v_y;
      for (int k = 1; k < Mz; ++k) {
        u_x = D_x * (west * (u_ij[k] - u_w[k]) + east * (u_e[k] - u_ij[k]));
        v_y = D_y * (south * (v_ij[k] - v_s[k]) + north * (v_n[k] - v_ij[k]));
        const double new_integrand = u_x + v_y;
        const double dz = z[k] - z[k - 1];
        w_ij[k] = w_ij[k - 1] - 0.5 * (new_integrand + old_integrand) * dz;
        old_integrand = new_integrand;
      }
    }
  }
#undef MM
  return SIA_OK;
}

/* SIA_Sliding::surface_gradient_haseloff (SIA_Sliding.cc:374-498): the
   thickness-threshold form (H <= 0 is ice-free) with pseudo-surface
   elevations; its mahaffy and eta variants equal SIAFD's (SIA_Sliding.cc:
   292-372, 501-531). */
static int sliding_gradient_haseloff(const sia_grid *g, const double *H, const double *b, const double *h,
                                     double *h_x, double *h_y) {
  const double Hicefree = 0.0, dx = g->dx, dy = g->dy;
#define HH(a, c) H[I2(g, (a), (c))]
#define BB(a, c) b[I2(g, (a), (c))]
#define SS(a, c) h[I2(g, (a), (c))]
  for (int i = 0; i < g->Mx; i++) {
    for (int j = 0; j < g->My; j++) {
      { /* i-offset */
        const int icefreeP = HH(i, j) <= Hicefree, icefreeE = HH(i + 1, j) <= Hicefree,
                  icefreeN = HH(i, j + 1) <= Hicefree, icefreeS = HH(i, j - 1) <= Hicefree,
                  icefreeNE = HH(i + 1, j + 1) <= Hicefree, icefreeSE = HH(i + 1, j - 1) <= Hicefree;
        double hhE = SS(i + 1, j);
        if (icefreeE && (BB(i + 1, j) > SS(i, j))) hhE = SS(i, j);
        if (icefreeP && (BB(i, j) > SS(i + 1, j))) hhE = SS(i, j);
        h_x[IS(g, i, j, 0)] = (hhE - SS(i, j)) / dx;
        double hhN = SS(i, j + 1);
        if (icefreeN && (BB(i, j + 1) > SS(i, j))) hhN = SS(i, j);
        if (icefreeP && (BB(i, j) > SS(i, j + 1))) hhN = SS(i, j);
        double hhS = SS(i, j - 1);
        if (icefreeS && (BB(i, j - 1) > SS(i, j))) hhS = SS(i, j);
        if (icefreeP && (BB(i, j) > SS(i, j - 1))) hhS = SS(i, j);
        double hhNE = SS(i + 1, j + 1);
        if (icefreeNE && (BB(i + 1, j + 1) > SS(i + 1, j))) hhNE = SS(i + 1, j);
        if (icefreeE && (BB(i + 1, j) > SS(i + 1, j + 1))) hhNE = SS(i + 1, j);
        double hhSE = SS(i + 1, j - 1);
        if (icefreeSE && (BB(i + 1, j - 1) > SS(i + 1, j))) hhSE = SS(i + 1, j);
        if (icefreeE && (BB(i + 1, j) > SS(i + 1, j - 1))) hhSE = SS(i + 1, j);
        h_y[IS(g, i, j, 0)] = (hhNE + hhN - hhSE - hhS) / (4.0 * dy);
      }
      { /* j-offset */
        const int icefreeP = HH(i, j) <= Hicefree, icefreeN = HH(i, j + 1) <= Hicefree,
                  icefreeE = HH(i + 1, j) <= Hicefree, icefreeW = HH(i - 1, j) <= Hicefree,
                  icefreeNE = HH(i + 1, j + 1) <= Hicefree, icefreeNW = HH(i - 1, j + 1) <= Hicefree;
        double hhN = SS(i, j + 1);
        if (icefreeN && (BB(i, j + 1) > SS(i, j))) hhN = SS(i, j);
        if (icefreeP && (BB(i, j) > SS(i, j + 1))) hhN = SS(i, j);
        h_y[IS(g, i, j, 1)] = (hhN - SS(i, j)) / dy;
        double hhE = SS(i + 1, j);
        if (icefreeE && (BB(i + 1, j) > SS(i, j))) hhE = SS(i, j);
        if (icefreeP && (BB(i, j) > SS(i + 1, j))) hhE = SS(i, j);
        double hhW = SS(i - 1, j);
        if (icefreeW && (BB(i - 1, j) > SS(i, j))) hhW = SS(i, j);
        if (icefreeP && (BB(i, j) > SS(i - 1, j))) hhW = SS(i, j);
        double hhNE = SS(i + 1, j + 1);
        if (icefreeNE && (BB(i + 1, j + 1) > SS(i, j + 1))) hhNE = SS(i, j + 1);
        if (icefreeN && (BB(i, j + 1) > SS(i + 1, j + 1))) hhNE = SS(i, j + 1);
        double hhNW = SS(i - 1, j + 1);
        if (icefreeNW && (BB(i - 1, j + 1) > SS(i, j + 1))) hhNW = SS(i, j + 1);
        if (icefreeN && (BB(i, j + 1) > SS(i - 1, j + 1))) hhNW = SS(i, j + 1);
        h_x[IS(g, i, j, 1)] = (hhNE + hhE - hhNW - hhW) / (4.0 * dx);
      }
    }
  }
#undef HH
#undef BB
#undef SS
  return SIA_OK;
}

/* SIA_Sliding::basalVelocitySIA (SIA_Sliding.cc:209-267): the coefficient C of
   U_b = -C grad h */
static double basal_velocity_sia(const sia_params *P, const or_sliding *sp, double x_in, double y_in, double H,
                                 double T) {
  const double ice_rho = P->ice_density, g = P->standard_gravity,
               beta_CC_grad = P->beta_CC * ice_rho * P->standard_gravity, secpera = P->seconds_per_year;
  if (sp->mode == OR_SLIDING_TEST_E) {
    const double r1 = 200e3, r2 = 700e3, theta1 = 10 * (M_PI / 180), theta2 = 40 * (M_PI / 180);
    const double x = fabs(x_in), y = fabs(y_in);
    const double r = sqrt(x * x + y * y);
    double theta;
    if (x < 1.0) {
      theta = M_PI / 2.0;
    } else {
      theta = atan(y / x);
    }
    if ((r > r1) && (r < r2) && (theta > theta1) && (theta < theta2)) {
      const double rbot = (r2 - r1) * (r2 - r1), thetabot = (theta2 - theta1) * (theta2 - theta1);
      const double mu_max = 2.5e-11;
      const double muE = mu_max * (4.0 * (r - r1) * (r2 - r) / rbot) *
                         (4.0 * (theta - theta1) * (theta2 - theta) / thetabot);
      return muE * ice_rho * g * H;
    }
    return 0.0;
  }
  if (sp->mode == OR_SLIDING_EISMINT2_G || sp->mode == OR_SLIDING_EISMINT2_H) {
    const double Bfactor = 1e-3 / secpera;
    const double pressure = or_ec_pressure(P, H);
    const double E = or_ec_enthalpy_permissive(P, T, 0.0, pressure);
    if (sp->mode == OR_SLIDING_EISMINT2_G) return Bfactor * ice_rho * g * H;
    return or_ec_is_temperate(P, E, pressure) ? Bfactor * ice_rho * g * H : 0.0;
  }
  if (T + beta_CC_grad * H > sp->minimum_temperature_for_sliding) {
    const double p_over = ice_rho * g * H;
    return sp->mu_sliding * p_over;
  }
  return 0;
}

/* SIA_Sliding::update (SIA_Sliding.cc:106-206) */
int or_sia_sliding(const sia_grid *g, const sia_params *P, const or_sliding *sp, const double *x, const double *y,
                   const double *H, const double *h, const double *b, const double *mask, const double *E,
                   double *h_x, double *h_y, double *ub, double *vb, double *heat) {
  int rc;
  switch (P->surface_gradient_method) {
  case SIA_GRADIENT_ETA:
    rc = gradient_eta(g, P, H, b, h_x, h_y);
    break;
  case SIA_GRADIENT_HASELOFF:
    rc = sliding_gradient_haseloff(g, H, b, h, h_x, h_y);
    break;
  case SIA_GRADIENT_MAHAFFY:
    rc = gradient_mahaffy(g, h, h_x, h_y);
    break;
  default:
    return fail(SIA_ERR_CONFIG, "value of surface_gradient_method is not valid");
  }
  if (rc) return rc;
  const double ice_rho = P->ice_density;
  for (int i = 0; i < g->Mx; i++) {
    for (int j = 0; j < g->My; j++) {
      const size_t q = I2(g, i, j);
      if (m_ocean(as_int(mask[q]))) {
        ub[q] = 0.0;
        vb[q] = 0.0;
        heat[q] = 0.0;
        continue;
      }
      const double myhx = 0.25 * (h_x[IS(g, i, j, 0)] + h_x[IS(g, i - 1, j, 0)] + h_x[IS(g, i, j, 1)] +
                                  h_x[IS(g, i, j - 1, 1)]),
                   myhy = 0.25 * (h_y[IS(g, i, j, 0)] + h_y[IS(g, i - 1, j, 0)] + h_y[IS(g, i, j, 1)] +
                                  h_y[IS(g, i, j - 1, 1)]);
      const double Hq = h[q] - b[q];
      const double T = or_ec_temperature(P, E[q * (size_t)g->Mz], or_ec_pressure(P, Hq));
      const double basalC = basal_velocity_sia(P, sp, x[i], y[j], Hq, T);
      ub[q] = -basalC * myhx;
      vb[q] = -basalC * myhy;
      const double Pov = ice_rho * P->standard_gravity * Hq;
      heat[q] = -(Pov * myhx) * ub[q] - (Pov * myhy) * vb[q];
    }
  }
  return SIA_OK;
}

/* SIAFD_Regional::compute_surface_gradient (regional.cc:45-109): where the
   no-model mask is set on any point of a staggered gradient's stencil, the
   gradient comes from the stored surface usurfstore instead (zero where the
   stencil leaves the domain). Mask and usurfstore neighbours wrap
   periodically, as PISM's ghosts do. */
int or_regional_gradient(const sia_grid *g, const double *nmm, const double *hst, double *h_x,
                         double *h_y) {
  const int Mx = g->Mx, My = g->My;
  const double dx = g->dx, dy = g->dy;
#define NM(a, c) nmm[I2(g, (a), (c))]
#define HS(a, c) hst[I2(g, (a), (c))]
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      if (NM(i, j) > 0.5 || NM(i + 1, j) > 0.5) {
        if (i < 0 || i + 1 > Mx - 1) {
          h_x[IS(g, i, j, 0)] = 0.0;
        } else {
          h_x[IS(g, i, j, 0)] = (HS(i + 1, j) - HS(i, j)) / dx;
        }
      }
      if (NM(i - 1, j + 1) > 0.5 || NM(i + 1, j + 1) > 0.5 || NM(i - 1, j) > 0.5 || NM(i + 1, j) > 0.5) {
        if (i - 1 < 0 || j + 1 > My - 1 || i + 1 > Mx - 1) {
          h_x[IS(g, i, j, 1)] = 0.0;
        } else {
          h_x[IS(g, i, j, 1)] = (+HS(i + 1, j + 1) + HS(i + 1, j) - HS(i - 1, j + 1) - HS(i - 1, j)) / (4.0 * dx);
        }
      }
      if (NM(i, j + 1) > 0.5 || NM(i + 1, j + 1) > 0.5 || NM(i, j - 1) > 0.5 || NM(i + 1, j - 1) > 0.5) {
        if (i < 0 || j + 1 > My - 1 || i + 1 > Mx - 1 || j - 1 < 0) {
          h_y[IS(g, i, j, 0)] = 0.0;
        } else {
          h_y[IS(g, i, j, 0)] = (+HS(i + 1, j + 1) + HS(i, j + 1) - HS(i + 1, j - 1) - HS(i, j - 1)) / (4.0 * dy);
        }
      }
      if (NM(i, j) > 0.5 || NM(i, j + 1) > 0.5) {
        if (j < 0 || j + 1 > My - 1) {
          h_y[IS(g, i, j, 1)] = 0.0;
        } else {
          h_y[IS(g, i, j, 1)] = (HS(i, j + 1) - HS(i, j)) / dy;
        }
      }
    }
  }
#undef NM
#undef HS
  return SIA_OK;
}

/* D^2 of the shallow strain-rate approximation, PISMStressBalance.cc:309-311 */
static double D2(double u_x, double u_y, double u_z, double v_x, double v_y, double v_z) {
  return 0.5 * ((u_x + v_y) * (u_x + v_y) + u_x * u_x + v_y * v_y +
                0.5 * ((u_y + v_x) * (u_y + v_x) + u_z * u_z + v_z * v_z));
}

/* StressBalance::compute_volumetric_strain_heating, PISMStressBalance.cc:370-500 */
int or_strain_heating(const sia_grid *g, const sia_params *P, const double *H, const double *mask,
                      const double *E, const double *u, const double *v, double *Sigma) {
  const int Mx = g->Mx, My = g->My, Mz = g->Mz;
  const double *z = g->z;
  const double n = P->sia_Glen_exponent, exponent = 0.5 * (1.0 / n + 1.0),
               e_to_a_power = pow(P->sia_enhancement_factor, -1.0 / n);
#define MM(a, c) as_int(mask[I2(g, (a), (c))])
  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      const double Hij = H[I2(g, i, j)];
      const int ks = or_kbelow(z, Mz, Hij);
      if (ks < 0) return fail(SIA_ERR_THICKNESS_ABOVE_LZ, "strain heating: thickness outside the grid");
      double west = 1, east = 1, south = 1, north = 1, D_x = 0, D_y = 0;
      const int m = MM(i, j);
      if ((m_icy(m) && m_ice_free(MM(i + 1, j))) || (m_ice_free(m) && m_icy(MM(i + 1, j)))) east = 0;
      if ((m_icy(m) && m_ice_free(MM(i - 1, j))) || (m_ice_free(m) && m_icy(MM(i - 1, j)))) west = 0;
      D_x = (east + west > 0) ? 1.0 / (g->dx * (east + west)) : 0.0;
      if ((m_icy(m) && m_ice_free(MM(i, j + 1))) || (m_ice_free(m) && m_icy(MM(i, j + 1)))) north = 0;
      if ((m_icy(m) && m_ice_free(MM(i, j - 1))) || (m_ice_free(m) && m_icy(MM(i, j - 1)))) south = 0;
      D_y = (north + south > 0) ? 1.0 / (g->dy * (north + south)) : 0.0;
      const double *u_ij = u + I3(g, i, j), *u_w = u + I3(g, i - 1, j), *u_e = u + I3(g, i + 1, j),
                   *u_s = u + I3(g, i, j - 1), *u_n = u + I3(g, i, j + 1);
      const double *v_ij = v + I3(g, i, j), *v_w = v + I3(g, i - 1, j), *v_e = v + I3(g, i + 1, j),
                   *v_s = v + I3(g, i, j - 1), *v_n = v + I3(g, i, j + 1);
      const double *E_ij = E + I3(g, i, j);
      double *S = Sigma + I3(g, i, j);
      for (int k = 0; k <= ks; ++k) {
        const double pressure = or_ec_pressure(P, Hij - z[k]);
        const double B = or_hardness(P, E_ij[k], pressure);
        double dz, u_z, v_z;
        const double u_x = D_x * (west * (u_ij[k] - u_w[k]) + east * (u_e[k] - u_ij[k])),
                     u_y = D_y * (south * (u_ij[k] - u_s[k]) + north * (u_n[k] - u_ij[k])),
                     v_x = D_x * (west * (v_ij[k] - v_w[k]) + east * (v_e[k] - v_ij[k])),
                     v_y = D_y * (south * (v_ij[k] - v_s[k]) + north * (v_n[k] - v_ij[k]));
        if (k > 0) {
          dz = z[k + 1] - z[k - 1];
          u_z = (u_ij[k + 1] - u_ij[k - 1]) / dz;
          v_z = (v_ij[k + 1] - v_ij[k - 1]) / dz;
        } else {
          dz = z[1] - z[0];
          u_z = (u_ij[1] - u_ij[0]) / dz;
          v_z = (v_ij[1] - v_ij[0]) / dz;
        }
        S[k] = 2.0 * e_to_a_power * B * pow(D2(u_x, u_y, u_z, v_x, v_y, v_z), exponent);
      }
      for (int k = ks + 1; k < Mz; ++k) S[k] = 0.0;
    }
  }
#undef MM
  return SIA_OK;
}

/* SIAFD::update (SIAFD.cc:119-148) + StressBalance::update's strain heating and
   w (PISMStressBalance.cc:98-110) */
int or_sia_update(const sia_grid *g, const sia_params *P, or_fields *f, int full) {
  int rc;
  rc = or_bed_smoother(g, P, f->b, f->topgsmooth, f->maxtl, f->C2, f->C3, f->C4, &f->Nx, &f->Ny);
  if (rc) return rc;
  rc = or_surface_gradient(g, P, f->H, f->b, f->h, f->mask, f->h_x, f->h_y);
  if (rc) return rc;
  if (f->no_model_mask && f->usurfstore) {  /* SIAFD_Regional */
    rc = or_regional_gradient(g, f->no_model_mask, f->usurfstore, f->h_x, f->h_y);
    if (rc) return rc;
  }
  rc = or_theta(g, P, f->Nx, f->Ny, f->h, f->topgsmooth, f->maxtl, f->C2, f->C3, f->C4, f->theta);
  if (rc) return rc;
  rc = or_thksmooth(g, f->h, f->H, f->mask, f->maxtl, f->topgsmooth, f->thksmooth);
  if (rc) return rc;
  rc = or_diffusive_flux(g, P, f->h_x, f->h_y, f->thksmooth, f->theta, f->E, f->age, full,
                         f->flux, f->D_stag, f->delta0, f->delta1, &f->D_max);
  if (rc) return rc;
  if (full) {
    rc = or_compute_I(g, f->thksmooth, f->delta0, f->delta1, f->I0, f->I1);
    if (rc) return rc;
    rc = or_horizontal_velocity(g, f->h_x, f->h_y, f->I0, f->I1, f->ub, f->vb, f->u, f->v);
    if (rc) return rc;
    if (f->Sigma) {
      rc = or_strain_heating(g, P, f->H, f->mask, f->E, f->u, f->v, f->Sigma);
      if (rc) return rc;
    }
    rc = or_vertical_velocity(g, f->mask, f->u, f->v, f->bmr, f->w);
    if (rc) return rc;
  }
  return SIA_OK;
}

/* ------------------------------------------------------------------------ */
/* GeometryCalculator::compute (Mask.hh:82-118) applied by update_mask and
   update_surface_elevation (iMgeometry.cc:80-149) */
int or_update_mask_surface(const sia_grid *g, const sia_params *P, const double *bed,
                           const double *H, double *mask, double *h) {
  const double alpha = 1 - P->ice_density / P->sea_water_density;
  const size_t N2 = (size_t)g->Mx * g->My;
  for (size_t q = 0; q < N2; q++) {
    const double thickness = H[q];
    if (thickness < 0) {
      return fail(SIA_ERR_NEGATIVE_THICKNESS, "Thickness negative");
    }
    const double hgrounded = bed[q] + thickness;
    const double hfloating = P->sea_level + alpha * thickness;
    const int is_floating = (hfloating > hgrounded + P->mask_is_floating_thickness_standard),
              ice_free = (thickness < P->mask_icefree_thickness_standard);
    int mask_result;
    double surface_result;
    if (is_floating && !P->is_dry_simulation) {
      surface_result = hfloating;
      mask_result = ice_free ? M_FREE_OCEAN : M_FLOATING;
    } else {
      surface_result = hgrounded;
      mask_result = ice_free ? M_FREE_BEDROCK : M_GROUNDED;
    }
    mask[q] = (double)mask_result;
    h[q] = surface_result;
  }
  return SIA_OK;
}

/* IceModel::adjust_flow, iMgeometry.cc:172-281. dir: 0=N,1=E,2=S,3=W */
static void adjust_flow(int mask_current, const int mask_nb[4], double SSA[4], double SIA[4]) {
  for (int n = 0; n < 4; ++n) {
    const int c = mask_current, nb = mask_nb[n];
    if (m_grounded_ice(c) && m_grounded_ice(nb)) continue;
    if ((m_grounded_ice(c) && m_floating_ice(nb)) || (m_floating_ice(c) && m_grounded_ice(nb))) continue;
    if ((m_grounded_ice(c) && m_ice_free_land(nb)) || (m_ice_free_land(c) && m_grounded_ice(nb))) continue;
    if ((m_grounded_ice(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_grounded_ice(nb))) continue;
    if (m_floating_ice(c) && m_floating_ice(nb)) {
      SIA[n] = 0.0;
      continue;
    }
    if ((m_floating_ice(c) && m_ice_free_land(nb)) || (m_ice_free_land(c) && m_floating_ice(nb))) {
      SIA[n] = 0.0;
      SSA[n] = 0.0;
      continue;
    }
    if ((m_floating_ice(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_floating_ice(nb))) {
      SIA[n] = 0.0;
      continue;
    }
    if (m_ice_free_land(c) && m_ice_free_land(nb)) {
      SIA[n] = 0.0;
      SSA[n] = 0.0;
      continue;
    }
    if ((m_ice_free_land(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_ice_free_land(nb))) {
      SIA[n] = 0.0;
      SSA[n] = 0.0;
      continue;
    }
    if (m_ice_free_ocean(c) && m_ice_free_ocean(nb)) {
      SIA[n] = 0.0;
      SSA[n] = 0.0;
      continue;
    }
  }
}

/* IceModel::massContExplicitStep, iMgeometry.cc:497-822, with
   cell_interface_fluxes (309-415) and no part_grid / Dirichlet BC. */
int or_mass_continuity(const sia_grid *g, const sia_params *P, double dt,
                       const double *H, const double *flux, const double *ub,
                       const double *vb, const double *mask, const double *smb,
                       const double *bmr, double *H_new, double *acc) {
  const int Mx = g->Mx, My = g->My;
  const double dx = g->dx, dy = g->dy;
  const double ice_density = P->ice_density;
  const double cell_area = dx * dy; /* IceModel::compute_cell_areas, iMcell_areas.cc:53 */
  double proc_grounded_basal_ice_flux = 0, proc_nonneg_rule_flux = 0, proc_sub_shelf_ice_flux = 0,
         proc_surface_ice_flux = 0, proc_sum_divQ_SIA = 0, proc_sum_divQ_SSA = 0,
         proc_Href_to_H_flux = 0, proc_H_to_Href_flux = 0;

  for (int i = 0; i < Mx; i++) {
    for (int j = 0; j < My; j++) {
      const size_t q = I2(g, i, j);
      const double meter_to_kg = cell_area * ice_density, meter_per_s_to_kg = meter_to_kg * dt;
      double divQ_SIA = 0.0, divQ_SSA = 0.0;
      double surface_mass_balance = (smb ? smb[q] : 0.0) / ice_density, my_basal_melt_rate = 0.0,
             H_to_Href_flux = 0.0, Href_to_H_flux = 0.0, nonneg_rule_flux = 0.0;
      if (P->include_bmr_in_continuity) {
        my_basal_melt_rate = bmr ? bmr[q] : 0.0;
      }
      /* directions N, E, S, W */
      const int ni[4] = {i, i + 1, i, i - 1}, nj[4] = {j + 1, j, j - 1, j};
      const int mc = as_int(mask[q]);
      int mn[4];
      for (int d = 0; d < 4; d++) mn[d] = as_int(mask[I2(g, ni[d], nj[d])]);
      /* SIA flux at the interfaces (IceModelVec2Stag::star, IceModelVec_inline.hh:75-92) */
      double Q[4] = {flux[IS(g, i, j, 1)], flux[IS(g, i, j, 0)], flux[IS(g, i, j - 1, 1)],
                     flux[IS(g, i - 1, j, 0)]};
      /* SSA velocity at the interfaces (cell_interface_fluxes, iMgeometry.cc:329-371) */
      double V[4];
      for (int d = 0; d < 4; d++) {
        const int ew = (d == 1 || d == 3);
        const double *vel = ew ? ub : vb;
        const double vij = vel ? vel[q] : 0.0;
        const double vnb = vel ? vel[I2(g, ni[d], nj[d])] : 0.0;
        if (m_icy(mc) && m_icy(mn[d])) {
          V[d] = 0.5 * (vij + vnb);
        } else if (m_icy(mc) && m_ice_free(mn[d])) {
          V[d] = vij;
        } else if (m_ice_free(mc) && m_icy(mn[d])) {
          V[d] = vnb;
        } else {
          V[d] = 0.0;
        }
      }
      adjust_flow(mc, mn, V, Q);
      const double Qn = Q[0], Qe = Q[1], Qs = Q[2], Qw = Q[3];
      const double vn = V[0], ve = V[1], vs = V[2], vw = V[3];
      divQ_SIA = (Qe - Qw) / dx + (Qn - Qs) / dy;
      divQ_SSA += (ve * (ve > 0 ? H[q] : H[I2(g, i + 1, j)]) - vw * (vw > 0 ? H[I2(g, i - 1, j)] : H[q])) / dx;
      divQ_SSA += (vn * (vn > 0 ? H[q] : H[I2(g, i, j + 1)]) - vs * (vs > 0 ? H[I2(g, i, j - 1)] : H[q])) / dy;

      if (m_ice_free_ocean(mc)) {
        surface_mass_balance = 0.0;
        my_basal_melt_rate = 0.0;
      }
      double Hn = H[q];
      Hn += (dt * (surface_mass_balance - my_basal_melt_rate - (divQ_SIA + divQ_SSA)) + Href_to_H_flux);
      if (Hn < 0.0) {
        nonneg_rule_flux += -Hn;
        Hn = 0.0;
      }
      H_new[q] = Hn;
      if (m_grounded(mc)) {
        proc_grounded_basal_ice_flux += -my_basal_melt_rate * meter_per_s_to_kg;
      } else {
        proc_sub_shelf_ice_flux += -my_basal_melt_rate * meter_per_s_to_kg;
      }
      proc_surface_ice_flux += surface_mass_balance * meter_per_s_to_kg;
      proc_sum_divQ_SIA += -divQ_SIA * meter_per_s_to_kg;
      proc_sum_divQ_SSA += -divQ_SSA * meter_per_s_to_kg;
      proc_nonneg_rule_flux += nonneg_rule_flux * meter_to_kg;
      proc_H_to_Href_flux += -H_to_Href_flux * meter_to_kg;
      proc_Href_to_H_flux += Href_to_H_flux * meter_to_kg;
    }
  }
  if (acc) {
    acc[0] = proc_grounded_basal_ice_flux;
    acc[1] = proc_nonneg_rule_flux;
    acc[2] = proc_sub_shelf_ice_flux;
    acc[3] = proc_surface_ice_flux;
    acc[4] = proc_sum_divQ_SIA;
    acc[5] = proc_sum_divQ_SSA;
    acc[6] = proc_Href_to_H_flux;
    acc[7] = proc_H_to_Href_flux;
  }
  return SIA_OK;
}

/* IceModel::max_timestep for the SIA-only path: "max", "end of the run",
   "2D CFL" (max_timestep_cfl_2d, iMadaptive.cc:118-144) and "diffusivity"
   (max_timestep_diffusivity, iMadaptive.cc:157-171); the smallest wins
   (iMadaptive.cc:311-327). */
double or_max_timestep(const sia_grid *g, const sia_params *P, double D_max,
                       double time_to_end, const double *ub, const double *vb,
                       const double *mask) {
  double dt = P->maximum_time_step;
  if (time_to_end > 0.0 && time_to_end < dt) dt = time_to_end;
  /* 2D CFL */
  double cfl = P->maximum_time_step;
  if (ub && vb) {
    const size_t N2 = (size_t)g->Mx * g->My;
    for (size_t q = 0; q < N2; q++) {
      if (m_icy(as_int(mask[q]))) {
        const double denom = fabs(ub[q]) / g->dx + fabs(vb[q]) / g->dy;
        if (denom > 0.0) {
          const double c = 1.0 / denom;
          cfl = (c < cfl) ? c : cfl;
        }
      }
    }
  }
  if (cfl < dt) dt = cfl;
  double dt_D = P->maximum_time_step;
  if (D_max > 0.0) {
    const double grid_factor = 1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy);
    dt_D = P->adaptive_timestepping_ratio * 2.0 / (D_max * grid_factor);
  }
  if (dt_D < dt) dt = dt_D;
  return dt;
}
