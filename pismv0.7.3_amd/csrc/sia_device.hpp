// sia_device.hpp -- device-side physics of the SIA hot path for gfx950.
//
// Flow laws and enthalpy converters of PISM v0.7.3 written for the GPU. The
// host precomputes every constant into DevParams (passed by value as a kernel
// argument, so it lives in scalar registers/SGPR-loaded constant memory).
// The order of floating-point operations follows the reference expressions so
// that results stay within a few ulp of PISM's; FMA contraction is disabled
// for the whole library (-ffp-contract=off) and fused multiply-adds appear
// only where written explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pism_sia {

enum Law { LAW_ISOGLEN = 0, LAW_PB = 1, LAW_GPBLD = 2, LAW_HOOKE = 3, LAW_ARR = 4, LAW_ARRWARM = 5, LAW_GK = 6 };
enum EcKind { EC_DEFAULT = 0, EC_COLD = 1, EC_KIRCHHOFF = 2 };
enum ErrBits { ERR_NEG_THK = 1, ERR_THK_LZ = 2, ERR_OMEGA = 4 };

// largest Mz the kernels stage in LDS (z levels)
constexpr int MAX_MZ = 4096;

struct DevParams {
  int32_t Mx, My, Mz;
  int32_t law, ec, use_age;
  int32_t nm1_int;          // n-1 when it is 1, 2 or 3 (integer power), else -1
  int32_t Nx, Ny;           // bed smoother half-widths, -1 = inactive
  int32_t grad;             // surface_gradient_method
  int32_t is_dry, include_bmr;
  double dx, dy, Lz;
  // SIA
  double e, n, nm1, gs;
  // enthalpy converter (enthalpyConverter.cc:56-68)
  double p_air, rho_i, g, beta, ci, T_melt, T0, L, cw;
  // Paterson-Budd (flowlaws.cc:60-86)
  double A_cold, A_warm, Q_cold, Q_warm, R, T_crit;
  double pb_coeff;          // m_beta_CC_grad / (m_rho * m_standard_gravity)
  double A_T0;              // softness_parameter_paterson_budd(water_melting_point_temperature)
  double wf_coeff, wf_limit;
  double A_glen;            // ice_softness
  double hooke_A, hooke_Q, hooke_C, hooke_k, hooke_Tr;
  double sec_to_ka;
  double inv_ci, inv_L;      // 1 / c_i, 1 / L (the surrogate's per-level pass, flow_gpbld_fast)
  double QR_cold, QR_warm;   // Q / R of the Paterson-Budd cold and warm branches
  double exp_c[11];          // 1/12!, ..., 1/2! for exp_fast (ExpCoef)
  double log_c[10];          // 2/19, 2/17, ..., 2/3, 2 for log_fast (LogCoef)
  // strain heating (PISMStressBalance.cc:384-388; flowlaws.cc:76, 249-253)
  double hard_pow;           // -1 / n
  double e_pow;              // e^(-1/n)
  double heat_exp;           // (1/n + 1) / 2
  double B_glen;             // ice_softness^(-1/n)
  double ln_A_cold, ln_A_warm, ln_A_T0, ln_A_glen, ln_hooke_A;  // logs of the softness constants
  // geometry / time stepping
  double alpha_float, sea_level, icefree_thk, floating_thk, max_dt, adapt_ratio;
};

__device__ __forceinline__ int wrapi(int i, int M) {
  return i < 0 ? i + M : (i >= M ? i - M : i);
}

// (i, j) of map point q = i My + j; grids hold fewer than 2^31 points, so the
// division is 32-bit
__device__ __forceinline__ void map_ij(long q, int My, int &i, int &j) {
  const unsigned u = (unsigned)q;
  i = (int)(u / (unsigned)My);
  j = (int)(u - (unsigned)i * (unsigned)My);
}

// IceModelVec2Int::as_int (IceModelVec_inline.hh:94-100)
__device__ __forceinline__ int as_int(double a) { return (int)floor(a + 0.5); }

// mask predicates (Mask.hh:28-65)
__device__ __forceinline__ bool m_ocean(int M) { return M >= 3; }
__device__ __forceinline__ bool m_grounded(int M) { return !m_ocean(M); }
__device__ __forceinline__ bool m_icy(int M) { return M == 2 || M == 3; }
__device__ __forceinline__ bool m_grounded_ice(int M) { return m_icy(M) && m_grounded(M); }
__device__ __forceinline__ bool m_floating_ice(int M) { return m_icy(M) && m_ocean(M); }
__device__ __forceinline__ bool m_ice_free(int M) { return !m_icy(M); }
__device__ __forceinline__ bool m_ice_free_ocean(int M) { return m_ocean(M) && m_ice_free(M); }
__device__ __forceinline__ bool m_ice_free_land(int M) { return m_grounded(M) && m_ice_free(M); }

// ---------------------------------------------------------------- EC
// EnthalpyConverter::pressure (enthalpyConverter.cc:92-98)
__device__ __forceinline__ double ec_pressure(const DevParams &P, double depth) {
  return depth > 0.0 ? P.p_air + P.rho_i * P.g * depth : P.p_air;
}
// melting_temperature_impl (enthalpyConverter.cc:128-130; cold: 347-349)
__device__ __forceinline__ double ec_melting_temperature(const DevParams &P, double p) {
  return P.ec == EC_COLD ? P.T_melt : P.T_melt - P.beta * p;
}
// L_impl (enthalpyConverter.cc:115-118; Kirchhoff: 433-435)
__device__ __forceinline__ double ec_L(const DevParams &P, double T_pm) {
  return P.ec == EC_KIRCHHOFF ? P.L + (P.cw - P.ci) * (T_pm - 273.15) : P.L;
}
// enthalpy_cts_impl (enthalpyConverter.cc:141-143)
__device__ __forceinline__ double ec_enthalpy_cts(const DevParams &P, double p) {
  return P.ci * (ec_melting_temperature(P, p) - P.T0);
}
// temperature_impl (enthalpyConverter.cc:175-189; cold: 355-357)
__device__ __forceinline__ double ec_temperature(const DevParams &P, double E, double p) {
  if (P.ec == EC_COLD || E < ec_enthalpy_cts(P, p)) {
    return (E / P.ci) + P.T0;
  }
  return ec_melting_temperature(P, p);
}
// enthalpyConverter.cc:197-199
__device__ __forceinline__ double ec_pressure_adjusted_temperature(const DevParams &P, double E, double p) {
  return ec_temperature(P, E, p) - ec_melting_temperature(P, p) + P.T_melt;
}
// water_fraction_impl (enthalpyConverter.cc:219-234; cold: 342-345)
__device__ __forceinline__ double ec_water_fraction(const DevParams &P, double E, double p) {
  if (P.ec == EC_COLD) return 0.0;
  const double E_s = ec_enthalpy_cts(P, p);
  if (E <= E_s) return 0.0;
  return (E - E_s) / ec_L(P, ec_melting_temperature(P, p));
}

// enthalpy_impl (enthalpyConverter.cc:256-281; cold: 337-340)
__device__ __forceinline__ double ec_enthalpy(const DevParams &P, double T, double omega, double p) {
  if (P.ec == EC_COLD) return P.ci * (T - P.T0);
  const double T_m = ec_melting_temperature(P, p);
  if (T < T_m) return P.ci * (T - P.T0);
  return ec_enthalpy_cts(P, p) + omega * ec_L(P, T_m);
}
// enthalpy_permissive (enthalpyConverter.cc:307-320; cold: 331-335)
__device__ __forceinline__ double ec_enthalpy_permissive(const DevParams &P, double T, double omega, double p) {
  if (P.ec == EC_COLD) return P.ci * (T - P.T0);
  const double T_m = ec_melting_temperature(P, p);
  if (T < T_m) return ec_enthalpy(P, T, 0.0, p);
  return ec_enthalpy(P, T_m, fmax(0.0, fmin(omega, 1.0)), p);
}
// is_temperate_impl (enthalpyConverter.cc:156-162; cold: 351-353)
__device__ __forceinline__ bool ec_is_temperate(const DevParams &P, double E, double p) {
  return P.ec != EC_COLD && E >= ec_enthalpy_cts(P, p);
}

// ---------------------------------------------------------------- flow laws
// FlowLaw::softness_parameter_paterson_budd (flowlaws.cc:90-95)
__device__ __forceinline__ double softness_pb(const DevParams &P, double T_pa) {
  const bool cold = T_pa < P.T_crit;
  const double A = cold ? P.A_cold : P.A_warm;
  const double Q = cold ? P.Q_cold : P.Q_warm;
  return A * exp(-Q / (P.R * T_pa));
}

__device__ __forceinline__ double stress_pow(const DevParams &P, double s) {
  switch (P.nm1_int) {
  case 1: return s;
  case 2: return s * s;               // == correctly rounded pow(s, 2)
  case 3: return s * s * s;
  default: return pow(s, P.nm1);
  }
}

// GoldsbyKohlstedt::flow_from_temp (flowlaws.cc:274-312 constants, 366-402)
__device__ inline double gk_flow_from_temp(const DevParams &P, double stress, double temp,
                                           double pressure, double gs) {
  const double V_act_vol = -13.e-6;
  const double disl_crit_temp = 258.0, disl_A_cold = 4.0e-19, disl_A_warm = 6.0e4,
               disl_n = 4.0, disl_Q_cold = 60.e3, disl_Q_warm = 180.e3;
  const double gbs_crit_temp = 255.0, gbs_A_cold = 6.1811e-14, gbs_A_warm = 4.7547e15,
               gbs_n = 1.8, gbs_Q_cold = 49.e3, gbs_Q_warm = 192.e3, p_grain_sz_exp = 1.4;
  const double basal_A = 2.1896e-7, basal_n = 2.4, basal_Q = 60.e3;
  const double diff_crit_temp = 258.0, diff_V_m = 1.97e-5, diff_D_0v = 9.10e-4,
               diff_Q_v = 59.4e3, diff_D_0b = 5.8e-4, diff_Q_b = 49.e3, diff_delta = 9.04e-10;
  if (fabs(stress) < 1e-10) return 0;
  const double T = temp + P.pb_coeff * pressure;
  const double pV = pressure * V_act_vol;
  const double RT = P.R * T;
  const double diff_D_v = diff_D_0v * exp(-diff_Q_v / RT);
  double diff_D_b = diff_D_0b * exp(-diff_Q_b / RT);
  if (T > diff_crit_temp) diff_D_b *= 1000;
  const double eps_diff = 14 * diff_V_m * (diff_D_v + M_PI * diff_delta * diff_D_b / gs) / (RT * (gs * gs));
  double eps_disl;
  if (T > disl_crit_temp) {
    eps_disl = disl_A_warm * pow(stress, disl_n - 1) * exp(-(disl_Q_warm + pV) / RT);
  } else {
    eps_disl = disl_A_cold * pow(stress, disl_n - 1) * exp(-(disl_Q_cold + pV) / RT);
  }
  const double eps_basal = basal_A * pow(stress, basal_n - 1) * exp(-(basal_Q + pV) / RT);
  double eps_gbs;
  if (T > gbs_crit_temp) {
    eps_gbs = gbs_A_warm * (pow(stress, gbs_n - 1) / pow(gs, p_grain_sz_exp)) * exp(-(gbs_Q_warm + pV) / RT);
  } else {
    eps_gbs = gbs_A_cold * (pow(stress, gbs_n - 1) / pow(gs, p_grain_sz_exp)) * exp(-(gbs_Q_cold + pV) / RT);
  }
  return eps_diff + eps_disl + (eps_basal * eps_gbs) / (eps_basal + eps_gbs);
}

// FlowLaw::flow and the overrides selected by sia_flow_law
// (flowlaws.cc:98-101, 209-244, 267-270, 314-318; flowlaws.hh:227-229, 296-326)
template <int LAW>
__device__ __forceinline__ double flow(const DevParams &P, double stress, double E, double p, double gs) {
  if (LAW == LAW_ISOGLEN) {
    return P.A_glen * stress_pow(P, stress);
  } else if (LAW == LAW_GPBLD) {
    const double E_s = ec_enthalpy_cts(P, p);
    double A;
    if (E < E_s) {
      A = softness_pb(P, ec_pressure_adjusted_temperature(P, E, p));
    } else {
      double omega = ec_water_fraction(P, E, p);
      if (P.wf_limit < omega) omega = P.wf_limit;
      A = P.A_T0 * (1.0 + P.wf_coeff * omega);
    }
    return A * stress_pow(P, stress);
  } else if (LAW == LAW_PB) {
    const double T_pa = ec_temperature(P, E, p) + P.pb_coeff * p;
    return softness_pb(P, T_pa) * stress_pow(P, stress);
  } else if (LAW == LAW_HOOKE) {
    const double T_pa = ec_temperature(P, E, p) + P.pb_coeff * p;
    const double A = P.hooke_A * exp(-P.hooke_Q / (P.R * T_pa) + 3.0 * P.hooke_C * pow(P.hooke_Tr - T_pa, -P.hooke_k));
    return A * stress_pow(P, stress);
  } else if (LAW == LAW_ARR) {
    const double temp = ec_temperature(P, E, p);
    return P.A_cold * exp(-P.Q_cold / (P.R * temp)) * stress_pow(P, stress);
  } else if (LAW == LAW_ARRWARM) {
    const double temp = ec_temperature(P, E, p);
    return P.A_warm * exp(-P.Q_warm / (P.R * temp)) * stress_pow(P, stress);
  } else {  // LAW_GK
    return gk_flow_from_temp(P, stress, ec_temperature(P, E, p), p, gs);
  }
}

// FlowLaw::hardness_parameter = softness^(-1/n) with each law's
// softness_parameter (flowlaws.cc:103-105, 209-229, 249-253, 267-270, 336-351;
// flowlaws.hh:235-237, 281-306): the Paterson-Budd family uses the
// converter's pressure-adjusted temperature here, not temp + beta p as in flow
template <int LAW>
__device__ __forceinline__ double hardness(const DevParams &P, double E, double p) {
  if (LAW == LAW_ISOGLEN) {
    return P.B_glen;
  } else if (LAW == LAW_GPBLD) {
    const double E_s = ec_enthalpy_cts(P, p);
    double A;
    if (E < E_s) {
      A = softness_pb(P, ec_pressure_adjusted_temperature(P, E, p));
    } else {
      double omega = ec_water_fraction(P, E, p);
      if (P.wf_limit < omega) omega = P.wf_limit;
      A = P.A_T0 * (1.0 + P.wf_coeff * omega);
    }
    return pow(A, P.hard_pow);
  } else if (LAW == LAW_PB || LAW == LAW_GK) {
    return pow(softness_pb(P, ec_pressure_adjusted_temperature(P, E, p)), P.hard_pow);
  } else if (LAW == LAW_HOOKE) {
    const double T_pa = ec_pressure_adjusted_temperature(P, E, p);
    return pow(P.hooke_A * exp(-P.hooke_Q / (P.R * T_pa) + 3.0 * P.hooke_C * pow(P.hooke_Tr - T_pa, -P.hooke_k)),
               P.hard_pow);
  } else if (LAW == LAW_ARR) {
    const double T_pa = ec_pressure_adjusted_temperature(P, E, p);
    return pow(P.A_cold * exp(-P.Q_cold / (P.R * T_pa)), P.hard_pow);
  } else {  // LAW_ARRWARM
    const double T_pa = ec_pressure_adjusted_temperature(P, E, p);
    return pow(P.A_warm * exp(-P.Q_warm / (P.R * T_pa)), P.hard_pow);
  }
}

// ln of each law's softness_parameter (the A whose -1/n power is the hardness,
// flowlaws.cc:103-105): for the Paterson-Budd family ln A_{c,w} - Q / (R T_pa)
// in closed form, so that B = exp(-ln A / n) takes one exp instead of an exp
// and a pow (equal to B to a few ulp)
template <int LAW>
__device__ __forceinline__ double log_softness(const DevParams &P, double E, double p) {
  if (LAW == LAW_ISOGLEN) {
    return P.ln_A_glen;
  } else if (LAW == LAW_GPBLD) {
    const double E_s = ec_enthalpy_cts(P, p);
    if (E < E_s) {
      const double T_pa = ec_pressure_adjusted_temperature(P, E, p);
      const bool cold = T_pa < P.T_crit;
      return (cold ? P.ln_A_cold : P.ln_A_warm) - (cold ? P.Q_cold : P.Q_warm) / (P.R * T_pa);
    }
    double omega = ec_water_fraction(P, E, p);
    if (P.wf_limit < omega) omega = P.wf_limit;
    return P.ln_A_T0 + log(1.0 + P.wf_coeff * omega);
  } else if (LAW == LAW_PB || LAW == LAW_GK) {
    const double T_pa = ec_pressure_adjusted_temperature(P, E, p);
    const bool cold = T_pa < P.T_crit;
    return (cold ? P.ln_A_cold : P.ln_A_warm) - (cold ? P.Q_cold : P.Q_warm) / (P.R * T_pa);
  } else if (LAW == LAW_HOOKE) {
    const double T_pa = ec_pressure_adjusted_temperature(P, E, p);
    return P.ln_hooke_A - P.hooke_Q / (P.R * T_pa) + 3.0 * P.hooke_C * pow(P.hooke_Tr - T_pa, -P.hooke_k);
  } else if (LAW == LAW_ARR) {
    return P.ln_A_cold - P.Q_cold / (P.R * ec_pressure_adjusted_temperature(P, E, p));
  } else {  // LAW_ARRWARM
    return P.ln_A_warm - P.Q_warm / (P.R * ec_pressure_adjusted_temperature(P, E, p));
  }
}

// ---------------------------------------------------------------- fast FP64 math
// 1 / t for t in the normal range: the hardware reciprocal estimate refined by
// two Newton steps (each doubles the correct bits); within 2 ulp of 1 / t
__device__ __forceinline__ double rcp_nr(double t) {
  double y = __builtin_amdgcn_rcp(t);
  double e = fma(-t, y, 1.0);
  y = fma(y, e, y);
  e = fma(-t, y, 1.0);
  return fma(y, e, y);
}

// e^x for |x| <= 700 (the Arrhenius factors exp(-Q / (R T)), the strain heating):
// x = n ln 2 + r with n = rint(x / ln 2) and |r| <= ln 2 / 2, the product
// n ln 2 taken in two parts (a high part with trailing zero bits, exact for
// |n| < 2^11, and the remainder); e^r by its Taylor series to degree 12,
// whose truncation is below 2.5e-16 relative for |r| <= 0.347; scaled by 2^n.
// Within a few ulp of e^x.
// The coefficients 1/k! are passed in wave-uniform (scalar) registers so that
// each Horner step is one v_fma_f64 with a scalar operand (as literals they
// would be copied into the accumulator before every step).
struct ExpCoef {
  double c[11];  // 1/12!, 1/11!, ..., 1/2!
};

__device__ __forceinline__ double exp_fast(double x, const ExpCoef &C) {
  const double n = rint(x * 1.4426950408889634074);
  double r = fma(-n, 6.93147180369123816490e-01, x);
  r = fma(-n, 1.90821492927058770002e-10, r);
  double q = C.c[0];
#pragma unroll
  for (int i = 1; i < 11; i++) q = fma(q, r, C.c[i]);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, (int)n);
}

__host__ __device__ inline ExpCoef exp_coef() {
  ExpCoef C;
  double f = 1.0;
  for (int k = 2; k <= 12; k++) f *= (double)k;  // 12!
  for (int i = 0; i < 11; i++) {
    C.c[i] = 1.0 / f;   // 1/(12-i)!
    f /= (double)(12 - i);
  }
  return C;
}

// ln x for finite x > 0 (the strain heating's ln D^2, the temperate softness's
// ln(1 + c omega)): x = 2^e m with m in [sqrt(1/2), sqrt(2)), so that
// ln x = e ln 2 + 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716; the
// series 2 atanh(s) = sum_k 2 s^(2k+1) / (2k+1) to k = 9 leaves a truncation
// below 3e-17 relative; e ln 2 in two parts as in exp_fast. Within a few ulp of
// ln x. The coefficients 2 / (2k + 1) come in scalar registers, as exp_fast's.
struct LogCoef {
  double c[10];  // 2/19, 2/17, ..., 2/3, 2
};

__device__ __forceinline__ double log_fast(double x, const LogCoef &C) {
  int e;
  double m = frexp(x, &e);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m += m;
    e -= 1;
  }
  const double f = m - 1.0;  // exact
  const double s = f * rcp_nr(2.0 + f);
  const double s2 = s * s;
  double q = C.c[0];
#pragma unroll
  for (int i = 1; i < 10; i++) q = fma(q, s2, C.c[i]);
  const double de = (double)e;
  return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, s * q));
}

__host__ __device__ inline LogCoef log_coef() {
  LogCoef C;
  for (int i = 0; i < 10; i++) C.c[i] = 2.0 / (double)(2 * (9 - i) + 1);
  return C;
}

// GPBLD's flow (flowlaws.cc:209-221 through FlowLaw::flow, 98-101) with the
// default enthalpy converter (enthalpyConverter.cc:92-234) and n - 1 = NM1 in
// {1, 2, 3}, branch-free: the cold and the temperate softness are both
// evaluated and the one E selects is kept. Against flow<LAW_GPBLD> the
// divisions by c_i and L become products with their reciprocals, and the cold
// softness A exp(-Q / (R T_pa)) becomes exp(ln A - (Q / R) (1 / T_pa)) by
// exp_fast: the softness agrees to a relative 1e-14 at most (the rounding of
// an exponent of size 60).
template <int NM1>
__device__ __forceinline__ double flow_gpbld_fast(const DevParams &P, double stress, double E, double p) {
  const double Tm = P.T_melt - P.beta * p;                 // melting_temperature_impl
  const double E_s = P.ci * (Tm - P.T0);                   // enthalpy_cts_impl
  const double T_pa = ((E * P.inv_ci) + P.T0) - Tm + P.T_melt;  // temperature_impl, pressure_adjusted_temperature
  const double y = rcp_nr(T_pa);
  // the warm Paterson-Budd branch and the temperate softness are evaluated only
  // when some lane of the wavefront takes them (wave-uniform branches); each
  // lane's value is the one it selects either way
  double x = P.ln_A_cold - P.QR_cold * y;
  if (__ballot(!(T_pa < P.T_crit))) {
    const double x_warm = P.ln_A_warm - P.QR_warm * y;
    x = (T_pa < P.T_crit) ? x : x_warm;
  }
  double A = exp_fast(x, *(const ExpCoef *)P.exp_c);  // softness_parameter_paterson_budd
  if (__ballot(!(E < E_s))) {
    const double omega = fmin((E - E_s) * P.inv_L, P.wf_limit);  // water_fraction_impl, capped
    const double A_t = P.A_T0 * (1.0 + P.wf_coeff * omega);
    A = (E < E_s) ? A : A_t;
  }
  const double sp = (NM1 == 1) ? stress : ((NM1 == 2) ? stress * stress : stress * stress * stress);
  return A * sp;
}

// log_softness<LAW_GPBLD> with the default enthalpy converter, the divisions
// replaced as in flow_gpbld_fast (ln A agrees to 1e-14 absolute)
__device__ __forceinline__ double log_softness_gpbld_fast(const DevParams &P, double E, double p) {
  const double Tm = P.T_melt - P.beta * p;
  const double E_s = P.ci * (Tm - P.T0);
  if (E < E_s) {
    const double T_pa = ((E * P.inv_ci) + P.T0) - Tm + P.T_melt;
    const bool cold = T_pa < P.T_crit;
    return (cold ? P.ln_A_cold : P.ln_A_warm) - (cold ? P.QR_cold : P.QR_warm) * rcp_nr(T_pa);
  }
  double omega = (E - E_s) * P.inv_L;
  if (P.wf_limit < omega) omega = P.wf_limit;
  return P.ln_A_T0 + log_fast(1.0 + P.wf_coeff * omega, *(const LogCoef *)P.log_c);
}

// SIAFD::grainSizeVostok (SIAFD.cc:980-1018)
__device__ inline double grain_size_vostok(const DevParams &P, double age_seconds) {
  const int numPoints = 22;
  const double ageAt[22] = {0.0000e+00, 5.0000e+01, 1.0000e+02, 1.2500e+02, 1.5000e+02, 1.5800e+02,
                            1.6500e+02, 1.7000e+02, 1.8000e+02, 1.8800e+02, 2.0000e+02, 2.2500e+02,
                            2.4500e+02, 2.6000e+02, 3.0000e+02, 3.2000e+02, 3.5000e+02, 4.0000e+02,
                            5.0000e+02, 6.0000e+02, 8.0000e+02, 1.0000e+04};
  const double gsAt[22] = {1.8000e-03, 2.2000e-03, 3.0000e-03, 4.0000e-03, 4.3000e-03, 3.0000e-03,
                           3.0000e-03, 4.6000e-03, 3.4000e-03, 3.3000e-03, 5.9000e-03, 6.2000e-03,
                           5.4000e-03, 6.8000e-03, 3.5000e-03, 6.0000e-03, 8.0000e-03, 8.3000e-03,
                           3.6000e-03, 3.8000e-03, 9.5000e-03, 1.0000e-02};
  const double a = age_seconds * P.sec_to_ka;
  int l = 0, r = numPoints - 1;
  if (a < ageAt[l]) return gsAt[l];
  if (a > ageAt[r]) return gsAt[r];
  while (r > l + 1) {
    const int j = (r + l) / 2;
    if (a < ageAt[j]) r = j; else l = j;
  }
  return gsAt[l] + (a - ageAt[l]) * (gsAt[r] - gsAt[l]) / (ageAt[r] - ageAt[l]);
}

// IceGrid::kBelowHeight (IceGrid.cc:379-391): GSL binary search, clamped to Mz-2.
// Returns -1 when height is outside [-1e-6, Lz + 1e-6].
__device__ __forceinline__ int kbelow(const double *__restrict__ z, int Mz, double Lz, double height) {
  if (height < 0.0 - 1.0e-6 || height > Lz + 1.0e-6) return -1;
  int ilo = 0, ihi = Mz - 1;
  while (ihi > ilo + 1) {
    const int i = (ihi + ilo) >> 1;
    if (z[i] > height) ihi = i; else ilo = i;
  }
  return ilo;
}

}  // namespace pism_sia
