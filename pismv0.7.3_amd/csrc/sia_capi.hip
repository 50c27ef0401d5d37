// sia_capi.hip -- the C ABI of include/pism_sia.h: handle, device buffers,
// orchestration of SIAFD::update, vertical velocity, time step and mass
// continuity on one HIP stream.
#include "../../include/pism_sia.h"
#include "../../include/sia_surrogate.h"
#include "../../include/sia_ensemble.h"
#include "sia_host.hpp"
#include "sia_kernels.hpp"
#include "sia_surrogate.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace pism_sia;

struct sia_handle {
  sia_grid grid{};
  std::vector<double> z;
  sia_params params{};
  DevParams dp{};
  hipStream_t stream = nullptr;
  int device = 0;
  bool sync_check = true;
  std::string last_error;

  double *d_z = nullptr;
  double *d_idz = nullptr;  // 1 / (z_{k+1} - z_{k-1}) per level (VelArgs::idz)
  // bed smoother state (BedSmoother members, PISMBedSmoother.hh:79-124)
  double *topgs = nullptr, *maxtl = nullptr, *C2 = nullptr, *C3 = nullptr, *C4 = nullptr;
  int64_t bed_counter = -1;  // SIAFD::init sets it to -1 (SIAFD.cc:95-97)
  bool bed_ready = false;
  // SIAFD work and output fields
  double *hx = nullptr, *hy = nullptr, *theta = nullptr, *thks = nullptr;
  double *flux = nullptr, *Dstag = nullptr;
  double *I0 = nullptr, *I1 = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *Sigma = nullptr;
  int *ks = nullptr;
  // scalars / scratch
  unsigned long long *dmax_bits = nullptr, *cfl_bits = nullptr;
  int *err = nullptr;      // flags of the current update
  int *err_acc = nullptr;  // flags of the earlier updates since the last check
  double *dt_dev = nullptr, *acc_dev = nullptr, *partials = nullptr, *H_tmp = nullptr;
  double *dmax_part = nullptr;
  double *sl_hx = nullptr, *sl_hy = nullptr, *sl_x = nullptr, *sl_y = nullptr;  // SIA_Sliding work
  const double *last_ub = nullptr, *last_vb = nullptr, *last_mask = nullptr;
  bool have_flux = false, have_3d = false;
  // learned surrogate for fast updates (borrowed) and its per-update buffers
  sia_surrogate *sg = nullptr;
  unsigned char *sg_status = nullptr;
  int *sg_block_count = nullptr, *sg_list = nullptr, *sg_nlist = nullptr, *sg_counts = nullptr;
  int *sg_fb_list = nullptr, *sg_nfb = nullptr;  // FP64 fallback columns and their count
  float *sg_F = nullptr, *sg_CQb = nullptr;
  double *sg_CSt = nullptr;
  float *sg_score = nullptr;     // per-slot fallback scores, written while calibrating
  bool sg_calibrating = false;
  // profiling events (Profiling::begin/end of the reference)
  struct Rec {
    int ev;
    hipEvent_t a, b;
  };
  bool prof_on = false;
  std::vector<Rec> recs;
  std::vector<hipEvent_t> event_pool;
};

// learned column surrogate: FP32 master weights, Adam state, packed inference
// weights, normalization and the training samples (all on one device)
struct sia_surrogate {
  int device = 0;
  hipStream_t stream = nullptr;
  float *master = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  unsigned short *frag = nullptr;
  SgNorm norm{};
  bool have_norm = false;
  double glen_n = 0.0;  // Glen exponent of the training samples (0: none yet)
  int version = 0;      // bumped when the normalization or tau changes (ensembles copy them)
  float tau = INFINITY; // fallback score threshold (calibrated; +inf: none)
  int bound = 0;        // handles bound to this model (sia_use_surrogate); destroy refuses while > 0
  SgSample *samples = nullptr;
  long long *d_n = nullptr;
  int64_t cap = 0, n = 0;
  long long adam_t = 0;
  // training buffers for batch size tb.B
  SgTrainBuffers tb{};
  std::vector<void *> tb_mem;
  double *d_loss = nullptr;
  int64_t d_loss_len = 0;
};

namespace pism_sia {

static double softness_pb_host(const sia_params &p, double T_pa) {
  if (T_pa < p.Paterson_Budd_critical_temperature) {
    return p.Paterson_Budd_A_cold * std::exp(-p.Paterson_Budd_Q_cold / (p.ideal_gas_constant * T_pa));
  }
  return p.Paterson_Budd_A_warm * std::exp(-p.Paterson_Budd_Q_warm / (p.ideal_gas_constant * T_pa));
}

void make_dev_params(const sia_grid &g, const sia_params &p, DevParams &d) {
  d = DevParams{};
  d.Mx = g.Mx;
  d.My = g.My;
  d.Mz = g.Mz;
  d.law = p.sia_flow_law;
  d.ec = p.enthalpy_converter;
  d.grad = p.surface_gradient_method;
  d.is_dry = p.is_dry_simulation;
  d.include_bmr = p.include_bmr_in_continuity;
  d.dx = g.dx;
  d.dy = g.dy;
  d.Lz = g.z[g.Mz - 1];
  d.e = p.sia_enhancement_factor;
  d.n = p.sia_Glen_exponent;
  d.nm1 = p.sia_Glen_exponent - 1;
  d.nm1_int = (d.nm1 == 1.0) ? 1 : (d.nm1 == 2.0) ? 2 : (d.nm1 == 3.0) ? 3 : -1;
  d.gs = p.ice_grain_size;
  d.p_air = p.surface_pressure;
  d.rho_i = p.ice_density;
  d.g = p.standard_gravity;
  d.beta = p.beta_CC;
  d.ci = p.ice_specific_heat_capacity;
  d.T_melt = p.water_melting_point_temperature;
  d.T0 = p.enthalpy_converter_reference_temperature;
  d.L = p.water_latent_heat_fusion;
  d.cw = p.water_specific_heat_capacity;
  d.A_cold = p.Paterson_Budd_A_cold;
  d.A_warm = p.Paterson_Budd_A_warm;
  d.Q_cold = p.Paterson_Budd_Q_cold;
  d.Q_warm = p.Paterson_Budd_Q_warm;
  d.R = p.ideal_gas_constant;
  d.T_crit = p.Paterson_Budd_critical_temperature;
  {
    const double beta_CC_grad = p.beta_CC * p.ice_density * p.standard_gravity;  // flowlaws.cc:71
    d.pb_coeff = beta_CC_grad / (p.ice_density * p.standard_gravity);            // flowlaws.cc:242
  }
  d.A_T0 = softness_pb_host(p, p.water_melting_point_temperature);  // flowlaws.cc:219
  d.ln_A_T0 = std::log(d.A_T0);
  d.wf_coeff = p.gpbld_water_frac_coeff;
  d.wf_limit = p.gpbld_water_frac_observed_limit;
  d.A_glen = p.ice_softness;
  d.hooke_A = p.Hooke_A;
  d.hooke_Q = p.Hooke_Q;
  d.hooke_C = p.Hooke_C;
  d.hooke_k = p.Hooke_k;
  d.hooke_Tr = p.Hooke_Tr;
  d.sec_to_ka = 1.0 / (1000.0 * p.seconds_per_year);
  d.hard_pow = -1.0 / p.sia_Glen_exponent;                         // flowlaws.cc:76
  d.e_pow = std::pow(p.sia_enhancement_factor, -1.0 / p.sia_Glen_exponent);  // PISMStressBalance.cc:388
  d.heat_exp = 0.5 * (1.0 / p.sia_Glen_exponent + 1.0);             // PISMStressBalance.cc:387
  d.B_glen = std::pow(p.ice_softness, d.hard_pow);                  // flowlaws.cc:252
  d.ln_A_cold = std::log(p.Paterson_Budd_A_cold);
  d.ln_A_warm = std::log(p.Paterson_Budd_A_warm);
  d.ln_A_glen = std::log(p.ice_softness);
  d.ln_hooke_A = std::log(p.Hooke_A);
  d.inv_ci = 1.0 / p.ice_specific_heat_capacity;
  d.inv_L = 1.0 / p.water_latent_heat_fusion;
  d.QR_cold = p.Paterson_Budd_Q_cold / p.ideal_gas_constant;
  d.QR_warm = p.Paterson_Budd_Q_warm / p.ideal_gas_constant;
  {
    const ExpCoef ec = exp_coef();
    for (int i = 0; i < 11; i++) d.exp_c[i] = ec.c[i];
    const LogCoef lc = log_coef();
    for (int i = 0; i < 10; i++) d.log_c[i] = lc.c[i];
  }
  d.alpha_float = 1 - p.ice_density / p.sea_water_density;
  d.sea_level = p.sea_level;
  d.icefree_thk = p.mask_icefree_thickness_standard;
  d.floating_thk = p.mask_is_floating_thickness_standard;
  d.max_dt = p.maximum_time_step;
  d.adapt_ratio = p.adaptive_timestepping_ratio;
  d.Nx = -1;
  d.Ny = -1;
  d.use_age = 0;
}

int check_grid_params(const sia_grid *grid, const sia_params *params) {
  if (!grid || !params || !grid->z) return SIA_ERR_ARGUMENT;
  if (grid->Mx < 3 || grid->My < 3 || grid->Mz < 2 || !(grid->dx > 0) || !(grid->dy > 0)) return SIA_ERR_ARGUMENT;
  // indices of staggered columns are 32-bit; z levels are staged in LDS
  if (2.0 * grid->Mx * grid->My >= 2147483647.0 || grid->Mz > MAX_MZ) return SIA_ERR_ARGUMENT;
  if (params->sia_flow_law < 0 || params->sia_flow_law > SIA_FLOW_GK) return SIA_ERR_CONFIG;
  if (params->surface_gradient_method < 0 || params->surface_gradient_method > SIA_GRADIENT_ETA) return SIA_ERR_CONFIG;
  if (params->enthalpy_converter < 0 || params->enthalpy_converter > SIA_EC_KIRCHHOFF) return SIA_ERR_CONFIG;
  for (int k = 1; k < grid->Mz; k++) {
    if (!(grid->z[k] > grid->z[k - 1])) return SIA_ERR_ARGUMENT;
  }
  return SIA_OK;
}

// BedSmoother's half-widths and moment scales (PISMBedSmoother.cc:96-149,
// 247-263): Nx = Ny = -1 when the range is not positive (no smoothing)
int bed_smoother_setup(const sia_grid &g, const sia_params &p, DevParams &d, double s[3]) {
  d.Nx = -1;
  d.Ny = -1;
  if (p.bed_smoother_range <= 0.0) return SIA_OK;
  int Nx = (int)std::ceil(p.bed_smoother_range / g.dx);
  int Ny = (int)std::ceil(p.bed_smoother_range / g.dy);
  if (Nx < 1) Nx = 1;
  if (Ny < 1) Ny = 1;
  if (Nx >= g.Mx || Ny >= g.My) return SIA_ERR_CONFIG;
  d.Nx = Nx;
  d.Ny = Ny;
  const double n = p.sia_Glen_exponent, k = (n + 2) / n;
  s[0] = k * (2 * n + 2) / (2 * n);
  s[1] = s[0] * (3 * n + 2) / (3 * n);
  s[2] = s[1] * (4 * n + 2) / (4 * n);
  return SIA_OK;
}

}  // namespace pism_sia

namespace {

int set_err(sia_handle *h, int code, const std::string &msg) {
  if (h) h->last_error = msg;
  return code;
}

#define HIP_TRY(h, call)                                                                        \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      return set_err((h), SIA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    }                                                                                           \
  } while (0)

hipEvent_t take_event(sia_handle *h) {
  if (!h->event_pool.empty()) {
    hipEvent_t e = h->event_pool.back();
    h->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one profiling event with hipEvents on the handle's stream.
struct Stage {
  sia_handle *h;
  size_t idx;
  bool on;
  Stage(sia_handle *hh, int ev) : h(hh), idx(0), on(hh->prof_on) {
    if (!on) return;
    sia_handle::Rec r{ev, take_event(h), take_event(h)};
    if (!r.a || !r.b) {
      on = false;
      return;
    }
    (void)hipEventRecord(r.a, h->stream);
    idx = h->recs.size();
    h->recs.push_back(r);
  }
  ~Stage() {
    if (on) (void)hipEventRecord(h->recs[idx].b, h->stream);
  }
};

template <typename T>
int dalloc(sia_handle *h, T **p, size_t count) {
  if (*p) return SIA_OK;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(h, e == hipErrorOutOfMemory ? SIA_ERR_OUT_OF_MEMORY : SIA_ERR_HIP,
                   std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return SIA_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}


// Reports the flags raised by every update since the last check (the
// reference throws at once, SIAFD.cc:654; here a step that does not
// synchronise leaves them to the next check) and clears them.
int check_flags(sia_handle *h) {
  int f[2] = {0, 0};
  HIP_TRY(h, hipMemcpyAsync(&f[0], h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipMemcpyAsync(&f[1], h->err_acc, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  launch_clear_flags(h->err, h->err_acc, h->stream);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int flags = f[0] | f[1];
  if (flags & ERR_NEG_THK) {
    return set_err(h, SIA_ERR_NEGATIVE_THICKNESS, "negative ice thickness (BedSmoother::get_smoothed_thk)");
  }
  if (flags & ERR_THK_LZ) {
    return set_err(h, SIA_ERR_THICKNESS_ABOVE_LZ, "ice thickness is above the top of the computational grid");
  }
  if (flags & ERR_OMEGA) {
    return set_err(h, SIA_ERR_OMEGA_NEGATIVE, "omega is negative in BedSmoother::get_theta()");
  }
  return SIA_OK;
}

// BedSmoother::preprocess_bed (PISMBedSmoother.cc:96-149)
int preprocess_bed(sia_handle *h, const double *bed) {
  const size_t n2 = (size_t)h->grid.Mx * h->grid.My;
  const sia_params &p = h->params;
  if (p.bed_smoother_range <= 0.0) {
    HIP_TRY(h, hipMemcpyAsync(h->topgs, bed, n2 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    // fill kernels, not memset nodes, so that a captured step replays correctly
    launch_fill(h->maxtl, (long)n2, 0.0, h->stream);
    launch_fill(h->C2, (long)n2, 0.0, h->stream);
    launch_fill(h->C3, (long)n2, 0.0, h->stream);
    launch_fill(h->C4, (long)n2, 0.0, h->stream);
    HIP_TRY(h, hipGetLastError());
    h->dp.Nx = -1;
    h->dp.Ny = -1;
    return SIA_OK;
  }
  double sc[3];
  if (bed_smoother_setup(h->grid, p, h->dp, sc)) {
    return set_err(h, SIA_ERR_CONFIG,
                   "input Nx, Ny in bed smoother is too large because domain of smoothing exceeds IceGrid domain");
  }
  const double s2 = sc[0], s3 = sc[1], s4 = sc[2];
  launch_bed_smoother(h->dp, s2, s3, s4, bed, h->topgs, h->maxtl, h->C2, h->C3, h->C4, h->stream);
  HIP_TRY(h, hipGetLastError());
  return SIA_OK;
}

int surrogate_flux(sia_handle *h, const sia_inputs *in, ColumnArgs A);

int update_impl(sia_handle *h, const sia_inputs *in, int fast, bool allow_sync) {
  if (!h || !in) return set_err(h, SIA_ERR_ARGUMENT, "NULL argument");
  if (!in->land_ice_thickness || !in->bedrock_altitude || !in->surface_altitude || !in->mask || !in->enthalpy) {
    return set_err(h, SIA_ERR_ARGUMENT, "sia_update: a required input field is NULL");
  }
  const sia_params &p = h->params;
  if (p.compute_grain_size_using_age && !p.do_age) {
    return set_err(h, SIA_ERR_CONFIG,
                   "SIAFD::compute_diffusive_flux(): do_age not set but age is needed for grain-size-based flow law");
  }
  const bool use_age = law_uses_grain_size(p.sia_flow_law) && p.compute_grain_size_using_age && p.do_age;
  if (use_age && !in->age) return set_err(h, SIA_ERR_ARGUMENT, "sia_update: age field is required");
  h->dp.use_age = use_age ? 1 : 0;
  const bool full = (fast == 0);
  const size_t n2 = (size_t)h->grid.Mx * h->grid.My, n3 = n2 * (size_t)h->grid.Mz;
  int rc;
  if (full) {
    const size_t nI = n2 * (size_t)i_stride(h->grid.Mz);
    if ((rc = dalloc(h, &h->I0, nI)) || (rc = dalloc(h, &h->I1, nI)) || (rc = dalloc(h, &h->u, n3)) ||
        (rc = dalloc(h, &h->v, n3)) || (rc = dalloc(h, &h->w, n3)) || (rc = dalloc(h, &h->Sigma, n3)) ||
        (rc = dalloc(h, &h->ks, 2 * n2))) {
      return rc;
    }
  }
  hipStream_t st = h->stream;
  h->have_flux = false;
  launch_reset(h->dmax_bits, h->err, h->err_acc, h->cfl_bits, h->dp.max_dt, st);
  // SIAFD.cc:128-133: re-smooth the bed when its state counter increased
  if (in->bed_state_counter > h->bed_counter || !h->bed_ready) {
    Stage stage(h, SIA_EV_BED_SMOOTHER);
    if ((rc = preprocess_bed(h, in->bedrock_altitude))) return rc;
    h->bed_counter = in->bed_state_counter;
    h->bed_ready = true;
  }
  const DevParams &P = h->dp;
  {
    Stage stage(h, SIA_EV_GRADIENT);
    const ThetaArgs T{h->topgs, h->maxtl, h->C2, h->C3, h->C4, h->theta, h->thks, h->err};
    launch_gradient(P, in->land_ice_thickness, in->bedrock_altitude, in->surface_altitude, in->mask, h->hx,
                    h->hy, T, st);
    if (in->no_model_mask && in->usurfstore) {  // SIAFD_Regional (regional.cc:45-109)
      launch_regional_gradient(P, in->no_model_mask, in->usurfstore, h->hx, h->hy, st);
    }
  }
  ColumnArgs A{};
  A.z = h->d_z;
  A.E = in->enthalpy;
  A.age = use_age ? in->age : nullptr;
  A.thks = h->thks;
  A.theta = h->theta;
  A.hx = h->hx;
  A.hy = h->hy;
  A.flux = h->flux;
  A.Dstag = h->Dstag;
  A.I0 = h->I0;
  A.I1 = h->I1;
  A.ks = h->ks;
  A.dmax_bits = h->dmax_bits;
  A.dmax_part = h->dmax_part;
  A.err = h->err;
  {
    Stage stage(h, SIA_EV_FLUX);
    if (!full && h->sg) {
      if ((rc = surrogate_flux(h, in, A))) return rc;
    } else {
      launch_column(P, A, full, st);
    }
  }
  if (full) {
    // "SIA 3D hor. vel.", "SB vert. vel." and the strain heating run as one fused kernel
    Stage stage(h, SIA_EV_3D_HOR_VEL);
    VelArgs V{};
    V.z = h->d_z;
    V.idz = h->d_idz;
    V.mask = in->mask;
    V.I0 = h->I0;
    V.I1 = h->I1;
    V.ks = h->ks;
    V.hx = h->hx;
    V.hy = h->hy;
    V.ub = in->vel_input_u;
    V.vb = in->vel_input_v;
    V.bmr = in->basal_melt_rate;
    V.u = h->u;
    V.v = h->v;
    V.w = h->w;
    // StressBalance::update computes the strain heating on every full update
    // (PISMStressBalance.cc:102-105); it runs in the same pass
    V.H = in->land_ice_thickness;
    V.E = in->enthalpy;
    V.Sigma = h->Sigma;
    V.err = h->err;
    launch_velocity_uvw(P, V, st);
  }
  if (in->vel_input_u && in->vel_input_v) {
    launch_cfl2d(P, in->mask, in->vel_input_u, in->vel_input_v, h->cfl_bits, st);
  }
  HIP_TRY(h, hipGetLastError());
  h->last_ub = in->vel_input_u;
  h->last_vb = in->vel_input_v;
  h->last_mask = in->mask;
  h->have_flux = true;
  h->have_3d = full;
  if (allow_sync && h->sync_check) return check_flags(h);
  return SIA_OK;
}

int ensure_sg_buffers(sia_handle *h) {
  const size_t n2 = (size_t)h->grid.Mx * h->grid.My, ns = 2 * n2;
  int rc;
  if ((rc = dalloc(h, &h->sg_status, ns)) || (rc = dalloc(h, &h->sg_block_count, 3 * (size_t)sg_classify_blocks(h->dp))) ||
      (rc = dalloc(h, &h->sg_list, ns)) || (rc = dalloc(h, &h->sg_nlist, 1)) || (rc = dalloc(h, &h->sg_counts, 3)) ||
      (rc = dalloc(h, &h->sg_F, ns * SG_FSTRIDE)) || (rc = dalloc(h, &h->sg_CQb, ns * SG_NCELL)) ||
      (rc = dalloc(h, &h->sg_CSt, ns)) || (rc = dalloc(h, &h->sg_fb_list, ns)) || (rc = dalloc(h, &h->sg_nfb, 1))) {
    return rc;
  }
  return SIA_OK;
}

SgColumnArgs sg_column_args(sia_handle *h, const sia_inputs *in) {
  return SgColumnArgs{h->d_z, in->enthalpy, h->thks, h->theta, h->hx, h->hy, h->err};
}

SgListArgs sg_list_args(sia_handle *h) {
  return SgListArgs{h->sg_status, h->sg_block_count, h->sg_list, h->sg_nlist, h->sg_counts, h->flux, h->Dstag,
                    h->sg_fb_list, h->sg_nfb, 1};
}

// flux, D_stag and D_max of a fast update through the surrogate: classify and
// list the staggered columns, one feature pass over their levels, the fused
// MLP, then the FP64 column kernel on the fallback columns only
int surrogate_flux(sia_handle *h, const sia_inputs *in, ColumnArgs A) {
  int rc;
  if ((rc = ensure_sg_buffers(h))) return rc;
  const DevParams &P = h->dp;
  hipStream_t st = h->stream;
  const SgColumnArgs C = sg_column_args(h, in);
  const SgListArgs L = sg_list_args(h);
  launch_sg_list(P, C, L, st);
  SgFeatureArgs F{};
  F.list = h->sg_list;
  F.n_list = h->sg_nlist;
  F.F = h->sg_F;
  F.CQb = h->sg_CQb;
  F.CSt = h->sg_CSt;
  F.status = h->sg_status;
  {
    Stage stage(h, SIA_EV_SG_FEATURES);
    launch_sg_features(P, C, F, false, st);
  }
  SgMlpArgs M{};
  M.list = h->sg_list;
  M.n_list = h->sg_nlist;
  M.F = h->sg_F;
  M.CQb = h->sg_CQb;
  M.CSt = h->sg_CSt;
  M.frag = h->sg->frag;
  M.bias = h->sg->master;
  M.norm = h->sg->norm;
  M.hx = h->hx;
  M.hy = h->hy;
  M.status = h->sg_status;
  M.flux = h->flux;
  M.Dstag = h->Dstag;
  M.dmax_part = h->dmax_part + column_list_blocks(P);
  M.counts = h->sg_counts;
  M.tau = h->sg_calibrating ? INFINITY : h->sg->tau;
  M.score = h->sg_calibrating ? h->sg_score : nullptr;
  M.fb_list = h->sg_fb_list;
  M.n_fb = h->sg_nfb;
  {
    Stage stage(h, SIA_EV_SG_MLP);
    launch_sg_mlp(P, M, st);
  }
  // the FP64 kernel on the listed fallback columns only
  launch_column_list(P, A, h->sg_fb_list, h->sg_nfb, st, sg_mlp_blocks());
  HIP_TRY(h, hipGetLastError());
  return SIA_OK;
}

// Xavier-uniform initial weights, zero biases (host, seeded)
void sg_init_params(std::vector<float> &w, uint64_t seed) {
  w.assign(SG_NPARAM, 0.0f);
  uint64_t x = seed;
  auto next = [&x]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const int off[5] = {SG_W1, SG_W2, SG_W3, SG_W4, SG_W5};
  const int fin[5] = {SG_NIN, SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_WIDTH};
  const int fout[5] = {SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_NOUT};
  for (int l = 0; l < 5; l++) {
    const double a = std::sqrt(6.0 / (fin[l] + fout[l]));
    for (int e = 0; e < fin[l] * fout[l]; e++) w[off[l] + e] = (float)((2.0 * next() - 1.0) * a);
  }
}

// (re)allocates the training buffers for batch size B
int sg_train_buffers(sia_surrogate *sg, int B) {
  if (sg->tb.B == B) return SIA_OK;
  for (void *p : sg->tb_mem) dfree(p);
  sg->tb_mem.clear();
  sg->tb = SgTrainBuffers{};
  auto alloc = [sg](size_t bytes, void **p) {
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return false;
    }
    sg->tb_mem.push_back(*p);
    return true;
  };
  SgTrainBuffers &T = sg->tb;
  const size_t b = (size_t)B;
  bool ok = alloc(b * SG_KIN * 2, (void **)&T.X0);
  const int N[SG_NLAYER] = {SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_NOUT};
  for (int l = 0; l < SG_NLAYER && ok; l++) ok = alloc(b * N[l] * 4, (void **)&T.Z[l]);
  T.Abf[0] = T.X0;
  for (int l = 1; l < SG_NLAYER && ok; l++) ok = alloc(b * SG_WIDTH * 2, (void **)&T.Abf[l]);
  ok = ok && alloc(b * SG_NCELL * 4, (void **)&T.T) && alloc(b * SG_NCELL * 4, (void **)&T.S) &&
       alloc(b * SG_NCELL * 4, (void **)&T.Hm) && alloc(b * SG_WIDTH * 4, (void **)&T.dZ) &&
       alloc(b * SG_WIDTH * 2, (void **)&T.dZbf) && alloc(b * SG_WIDTH * 4, (void **)&T.dA) &&
       alloc((size_t)SG_WIDTH * SG_KIN * 4, (void **)&T.dWpad) &&
       alloc((size_t)(SG_WIDTH * SG_KIN + 3 * SG_WIDTH * SG_WIDTH + SG_NOUT * SG_WIDTH) * 2, (void **)&T.Wbf);
  if (!ok) {
    for (void *p : sg->tb_mem) dfree(p);
    sg->tb_mem.clear();
    sg->tb = SgTrainBuffers{};
    return SIA_ERR_OUT_OF_MEMORY;
  }
  T.grad = sg->grad;
  T.B = B;
  return SIA_OK;
}

int sg_loss_buffer(sia_surrogate *sg, int64_t n) {
  if (sg->d_loss_len >= n) return SIA_OK;
  dfree(sg->d_loss);
  sg->d_loss = nullptr;
  sg->d_loss_len = 0;
  if (hipMalloc((void **)&sg->d_loss, sizeof(double) * (size_t)n) != hipSuccess) {
    sg->d_loss = nullptr;
    return SIA_ERR_OUT_OF_MEMORY;
  }
  sg->d_loss_len = n;
  return SIA_OK;
}

constexpr char SG_MAGIC[8] = {'P', 'S', 'I', 'A', 'S', 'G', '0', '1'};
struct SgBlobHeader {
  char magic[8];
  int32_t nparam, nin;
  double glen_n;
  int32_t have_norm, has_tau;  // has_tau: the fallback threshold follows the parameters
};

}  // namespace

extern "C" {

void sia_params_default(sia_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  const double secpera = 3.15569259747e7;  // UDUNITS "year"
  p->sia_flow_law = SIA_FLOW_GPBLD;                 // cdl:1094
  p->surface_gradient_method = SIA_GRADIENT_HASELOFF;  // cdl:1272
  p->enthalpy_converter = SIA_EC_DEFAULT;
  p->compute_grain_size_using_age = 0;               // cdl:309
  p->do_age = 0;                                     // cdl:34
  p->is_dry_simulation = 0;                          // cdl:136
  p->include_bmr_in_continuity = 1;                  // cdl:148
  p->sia_enhancement_factor = 1.0;                   // cdl:281
  p->sia_Glen_exponent = 3.0;                        // cdl:1009
  p->bed_smoother_range = 5.0e3;                     // cdl:560
  p->ice_grain_size = 1.0;                           // cdl:304
  p->ice_softness = 3.1689e-24;                      // cdl:1020
  p->Paterson_Budd_A_cold = 3.61e-13;                // cdl:1060
  p->Paterson_Budd_A_warm = 1.73e3;                  // cdl:1065
  p->Paterson_Budd_Q_cold = 6.0e4;                   // cdl:1070
  p->Paterson_Budd_Q_warm = 13.9e4;                  // cdl:1075
  p->Paterson_Budd_critical_temperature = 263.15;    // cdl:1080
  p->gpbld_water_frac_coeff = 181.25;                // cdl:1115
  p->gpbld_water_frac_observed_limit = 0.01;         // cdl:1120
  p->Hooke_A = 4.42165e-9;                           // cdl:1025
  p->Hooke_Q = 7.88e4;                               // cdl:1030
  p->Hooke_C = 0.16612;                              // cdl:1035
  p->Hooke_k = 1.17;                                 // cdl:1040
  p->Hooke_Tr = 273.39;                              // cdl:1045
  p->ideal_gas_constant = 8.31441;                   // cdl:1193
  p->beta_CC = 7.9e-8;                               // cdl:963
  p->ice_specific_heat_capacity = 2009.0;            // cdl:1003
  p->water_specific_heat_capacity = 4170.0;          // cdl:988
  p->water_latent_heat_fusion = 3.34e5;              // cdl:983
  p->surface_pressure = 0.0;                         // cdl:968
  p->ice_density = 910.0;                            // cdl:993
  p->standard_gravity = 9.81;                        // cdl:1188
  p->water_melting_point_temperature = 273.15;       // cdl:973
  p->enthalpy_converter_reference_temperature = 223.15;  // cdl:978
  p->sea_water_density = 1028.0;                     // cdl:1140
  p->sea_level = 0.0;
  p->mask_icefree_thickness_standard = 0.01;         // cdl:1225
  p->mask_is_floating_thickness_standard = 1.0;      // cdl:1230
  p->adaptive_timestepping_ratio = 0.12;             // cdl:330
  p->maximum_time_step = 60.0 * secpera;             // cdl:341
  p->seconds_per_year = secpera;
}

int sia_create(const sia_grid *grid, const sia_params *params, sia_handle **out) {
  if (!grid || !params || !out || !grid->z) return SIA_ERR_ARGUMENT;
  *out = nullptr;
  const int rc0 = check_grid_params(grid, params);
  if (rc0) return rc0;
  sia_handle *h = new (std::nothrow) sia_handle();
  if (!h) return SIA_ERR_OUT_OF_MEMORY;
  h->grid = *grid;
  h->z.assign(grid->z, grid->z + grid->Mz);
  h->grid.z = h->z.data();
  h->params = *params;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return SIA_ERR_HIP;
  }
  make_dev_params(h->grid, h->params, h->dp);
  const size_t n2 = (size_t)grid->Mx * grid->My;
  int rc;
  if ((rc = dalloc(h, &h->d_z, (size_t)grid->Mz)) || (rc = dalloc(h, &h->d_idz, (size_t)grid->Mz)) ||
      (rc = dalloc(h, &h->topgs, n2)) ||
      (rc = dalloc(h, &h->maxtl, n2)) || (rc = dalloc(h, &h->C2, n2)) || (rc = dalloc(h, &h->C3, n2)) ||
      (rc = dalloc(h, &h->C4, n2)) || (rc = dalloc(h, &h->hx, 2 * n2)) || (rc = dalloc(h, &h->hy, 2 * n2)) ||
      (rc = dalloc(h, &h->theta, n2)) || (rc = dalloc(h, &h->thks, n2)) || (rc = dalloc(h, &h->flux, 2 * n2)) ||
      (rc = dalloc(h, &h->Dstag, 2 * n2)) || (rc = dalloc(h, &h->dmax_bits, 1)) ||
      (rc = dalloc(h, &h->cfl_bits, 1)) || (rc = dalloc(h, &h->err, 1)) || (rc = dalloc(h, &h->err_acc, 1)) ||
      (rc = dalloc(h, &h->dt_dev, 1)) ||
      (rc = dalloc(h, &h->acc_dev, 8)) || (rc = dalloc(h, &h->H_tmp, n2)) ||
      (rc = dalloc(h, &h->partials, (size_t)mass_partial_blocks(h->dp) * 8)) ||
      (rc = dalloc(h, &h->dmax_part,
                   (size_t)(std::max(column_blocks(h->dp), column_strip_blocks(h->dp)) + sg_mlp_blocks())))) {
    sia_destroy(h);
    return rc;
  }
  // the strain heating's centred vertical differences (PISMStressBalance.cc:464-476)
  std::vector<double> idz(grid->Mz);
  for (int k = 0; k < grid->Mz; k++) {
    const double *z = h->z.data();
    idz[k] = (k == 0) ? 1.0 / (z[1] - z[0]) : ((k + 1 < grid->Mz) ? 1.0 / (z[k + 1] - z[k - 1]) : 0.0);
  }
  if (hipMemcpy(h->d_z, h->z.data(), sizeof(double) * grid->Mz, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_idz, idz.data(), sizeof(double) * grid->Mz, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->err, 0, sizeof(int)) != hipSuccess || hipMemset(h->err_acc, 0, sizeof(int)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    sia_destroy(h);
    return SIA_ERR_HIP;
  }
  *out = h;
  return SIA_OK;
}

int sia_destroy(sia_handle *h) {
  if (!h) return SIA_OK;
  if (h->sg) h->sg->bound--;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  else (void)hipDeviceSynchronize();
  void *bufs[] = {h->d_z, h->d_idz, h->topgs, h->maxtl, h->C2, h->C3, h->C4, h->hx, h->hy, h->theta, h->thks,
                  h->flux, h->Dstag, h->I0, h->I1, h->u, h->v, h->w, h->Sigma, h->ks, h->dmax_bits, h->cfl_bits, h->err, h->err_acc,
                  h->dt_dev, h->acc_dev, h->partials, h->H_tmp, h->dmax_part, h->sl_hx, h->sl_hy,
                  h->sl_x, h->sl_y, h->sg_status,
                  h->sg_block_count, h->sg_list, h->sg_nlist, h->sg_counts, h->sg_F, h->sg_CQb, h->sg_CSt,
                  h->sg_score, h->sg_fb_list, h->sg_nfb};
  for (void *b : bufs) dfree(b);
  for (auto &r : h->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
  delete h;
  return SIA_OK;
}

int sia_set_stream(sia_handle *h, void *hip_stream) {
  if (!h) return SIA_ERR_ARGUMENT;
  h->stream = (hipStream_t)hip_stream;
  return SIA_OK;
}

int sia_set_sync_check(sia_handle *h, int check) {
  if (!h) return SIA_ERR_ARGUMENT;
  h->sync_check = check != 0;
  return SIA_OK;
}

int sia_check_errors(sia_handle *h) {
  if (!h) return SIA_ERR_ARGUMENT;
  return check_flags(h);
}

int sia_update(sia_handle *h, const sia_inputs *in, int fast) { return update_impl(h, in, fast, true); }

int sia_diffusive_flux(sia_handle *h, const double **flux) {
  if (!h || !flux) return SIA_ERR_ARGUMENT;
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  *flux = h->flux;
  return SIA_OK;
}

int sia_max_diffusivity(sia_handle *h, double *D_max) {
  if (!h || !D_max) return SIA_ERR_ARGUMENT;
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  unsigned long long bits = 0;
  HIP_TRY(h, hipMemcpyAsync(&bits, h->dmax_bits, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
  int rc = check_flags(h);  // synchronises
  double d;
  std::memcpy(&d, &bits, sizeof(d));
  *D_max = d;
  return rc;
}

int sia_max_diffusivity_device(sia_handle *h, const double **D_max) {
  if (!h || !D_max) return SIA_ERR_ARGUMENT;
  *D_max = (const double *)h->dmax_bits;
  return SIA_OK;
}

static int get3d(sia_handle *h, const double **out, double *buf) {
  if (!h || !out) return SIA_ERR_ARGUMENT;
  if (!h->have_3d) return set_err(h, SIA_ERR_STATE, "3-D velocities need a full (fast = 0) update");
  *out = buf;
  return SIA_OK;
}

int sia_velocity_u(sia_handle *h, const double **u) { return get3d(h, u, h ? h->u : nullptr); }
int sia_velocity_v(sia_handle *h, const double **v) { return get3d(h, v, h ? h->v : nullptr); }
int sia_velocity_w(sia_handle *h, const double **w) { return get3d(h, w, h ? h->w : nullptr); }
int sia_volumetric_strain_heating(sia_handle *h, const double **Sigma) {
  return get3d(h, Sigma, h ? h->Sigma : nullptr);
}

int sia_surface_gradient(sia_handle *h, const double **h_x, const double **h_y) {
  if (!h || !h_x || !h_y) return SIA_ERR_ARGUMENT;
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  *h_x = h->hx;
  *h_y = h->hy;
  return SIA_OK;
}

int sia_diffusivity_staggered(sia_handle *h, const double **D) {
  if (!h || !D) return SIA_ERR_ARGUMENT;
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  *D = h->Dstag;
  return SIA_OK;
}

int sia_theta_thksmooth(sia_handle *h, const double **theta, const double **thksmooth) {
  if (!h || !theta || !thksmooth) return SIA_ERR_ARGUMENT;
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  *theta = h->theta;
  *thksmooth = h->thks;
  return SIA_OK;
}

int sia_bed_smoother_fields(sia_handle *h, const double **topgsmooth, const double **maxtl, const double **C2,
                            const double **C3, const double **C4, int32_t *Nx, int32_t *Ny) {
  if (!h) return SIA_ERR_ARGUMENT;
  if (!h->bed_ready) return set_err(h, SIA_ERR_STATE, "the bed has not been preprocessed yet");
  if (topgsmooth) *topgsmooth = h->topgs;
  if (maxtl) *maxtl = h->maxtl;
  if (C2) *C2 = h->C2;
  if (C3) *C3 = h->C3;
  if (C4) *C4 = h->C4;
  if (Nx) *Nx = h->dp.Nx;
  if (Ny) *Ny = h->dp.Ny;
  return SIA_OK;
}

int sia_max_timestep(sia_handle *h, double time_to_end, double *dt) {
  if (!h || !dt) return SIA_ERR_ARGUMENT;
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  launch_dt(h->dp, h->dmax_bits, h->cfl_bits, 0.0, nullptr, h->dt_dev, h->stream);
  HIP_TRY(h, hipGetLastError());
  double d = 0.0;
  HIP_TRY(h, hipMemcpyAsync(&d, h->dt_dev, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (time_to_end > 0.0 && time_to_end < d) d = time_to_end;
  *dt = d;
  return SIA_OK;
}

static int mass_step_impl(sia_handle *h, const sia_geometry *g, const double *dt_dev, sia_mass_accounting *acc) {
  if (!h || !g || !g->land_ice_thickness || !g->surface_altitude || !g->mask || !g->bedrock_altitude) {
    return set_err(h, SIA_ERR_ARGUMENT, "sia_mass_continuity_step: NULL geometry field");
  }
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  MassArgs M{};
  M.dt = dt_dev;
  M.H = g->land_ice_thickness;
  M.flux = h->flux;
  M.mask = g->mask;
  M.ub = h->last_ub;
  M.vb = h->last_vb;
  M.smb = g->climatic_mass_balance;
  M.bmr = g->basal_melt_rate;
  M.H_new = h->H_tmp;
  M.partials = h->partials;
  {
    Stage stage(h, SIA_EV_MASS_TRANSPORT);
    launch_mass_continuity(h->dp, M, h->acc_dev, h->stream);
    launch_mask_surface(h->dp, h->H_tmp, g->bedrock_altitude, g->land_ice_thickness, g->mask,
                        g->surface_altitude, h->err, h->stream);
  }
  HIP_TRY(h, hipGetLastError());
  if (acc) {
    double a[8];
    HIP_TRY(h, hipMemcpyAsync(a, h->acc_dev, sizeof(a), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    acc->grounded_basal_ice_flux = a[0];
    acc->nonneg_rule_flux = a[1];
    acc->sub_shelf_ice_flux = a[2];
    acc->surface_ice_flux = a[3];
    acc->sum_divQ_SIA = a[4];
    acc->sum_divQ_SSA = a[5];
    acc->Href_to_H_flux = a[6];
    acc->H_to_Href_flux = a[7];
  }
  return SIA_OK;
}

int sia_mass_continuity_step(sia_handle *h, const sia_geometry *g, double dt, sia_mass_accounting *acc) {
  if (!h) return SIA_ERR_ARGUMENT;
  if (!(dt >= 0.0)) return set_err(h, SIA_ERR_ARGUMENT, "dt must be non-negative");
  HIP_TRY(h, hipMemcpyAsync(h->dt_dev, &dt, sizeof(double), hipMemcpyHostToDevice, h->stream));
  int rc = mass_step_impl(h, g, h->dt_dev, acc);
  if (rc) return rc;
  // the host dt must stay alive until the copy ran
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->sync_check) return check_flags(h);
  return SIA_OK;
}

int sia_step(sia_handle *h, const sia_geometry *g, const double *enthalpy, int fast, double t_end,
             double *t_device, double *dt_device) {
  if (!h || !g || !enthalpy) return set_err(h, SIA_ERR_ARGUMENT, "sia_step: NULL argument");
  sia_inputs in{};
  in.land_ice_thickness = g->land_ice_thickness;
  in.bedrock_altitude = g->bedrock_altitude;
  in.surface_altitude = g->surface_altitude;
  in.mask = g->mask;
  in.enthalpy = enthalpy;
  in.basal_melt_rate = g->basal_melt_rate;
  in.bed_state_counter = h->bed_ready ? h->bed_counter : 0;
  int rc = update_impl(h, &in, fast, false);
  if (rc) return rc;
  double *dt_out = dt_device ? dt_device : h->dt_dev;
  launch_dt(h->dp, h->dmax_bits, h->cfl_bits, t_end, t_device, dt_out, h->stream);
  HIP_TRY(h, hipGetLastError());
  return mass_step_impl(h, g, dt_out, nullptr);
}

int sia_surface_values(sia_handle *h, const double *field3d, const double *thickness, double *out2d) {
  if (!h || !field3d || !thickness || !out2d) return set_err(h, SIA_ERR_ARGUMENT, "NULL argument");
  launch_surface_values(h->dp, h->d_z, field3d, thickness, out2d, h->stream);
  HIP_TRY(h, hipGetLastError());
  return SIA_OK;
}

int sia_diffusivity(sia_handle *h, double *out2d) {
  if (!h || !out2d) return set_err(h, SIA_ERR_ARGUMENT, "NULL argument");
  if (!h->have_flux) return set_err(h, SIA_ERR_STATE, "no update has run");
  launch_diffusivity_regular(h->dp, h->Dstag, out2d, h->stream);
  HIP_TRY(h, hipGetLastError());
  return SIA_OK;
}

int sia_sliding_update(sia_handle *h, const sia_inputs *in, const sia_sliding_params *sp, double *u_b,
                       double *v_b, double *basal_frictional_heating) {
  if (!h || !in || !sp || !u_b || !v_b || !basal_frictional_heating) return set_err(h, SIA_ERR_ARGUMENT, "NULL argument");
  if (!in->land_ice_thickness || !in->bedrock_altitude || !in->surface_altitude || !in->mask || !in->enthalpy) {
    return set_err(h, SIA_ERR_ARGUMENT, "sia_sliding_update: a required input field is NULL");
  }
  if (sp->mode < SIA_SLIDING_PRESSURE_MELTING || sp->mode > SIA_SLIDING_EISMINT2_H) {
    return set_err(h, SIA_ERR_CONFIG, "unknown SIA sliding mode");
  }
  if (sp->mode == SIA_SLIDING_TEST_E && (!sp->x || !sp->y)) {
    return set_err(h, SIA_ERR_ARGUMENT, "test E sliding needs the grid coordinates x, y");
  }
  const size_t n2 = (size_t)h->grid.Mx * h->grid.My;
  int rc;
  if ((rc = dalloc(h, &h->sl_hx, 2 * n2)) || (rc = dalloc(h, &h->sl_hy, 2 * n2))) return rc;
  if (sp->mode == SIA_SLIDING_TEST_E) {
    if ((rc = dalloc(h, &h->sl_x, (size_t)h->grid.Mx)) || (rc = dalloc(h, &h->sl_y, (size_t)h->grid.My))) return rc;
    HIP_TRY(h, hipMemcpyAsync(h->sl_x, sp->x, sizeof(double) * h->grid.Mx, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(h->sl_y, sp->y, sizeof(double) * h->grid.My, hipMemcpyHostToDevice, h->stream));
  }
  SlideArgs A{};
  A.F = Fields2D{in->land_ice_thickness, in->bedrock_altitude, in->surface_altitude, in->mask};
  A.E = in->enthalpy;
  A.x = sp->mode == SIA_SLIDING_TEST_E ? h->sl_x : nullptr;
  A.y = sp->mode == SIA_SLIDING_TEST_E ? h->sl_y : nullptr;
  A.hx = h->sl_hx;
  A.hy = h->sl_hy;
  A.ub = u_b;
  A.vb = v_b;
  A.heat = basal_frictional_heating;
  A.mode = sp->mode;
  A.mu = sp->mu_sliding;
  A.min_T = sp->minimum_temperature_for_sliding;
  A.secpera = h->params.seconds_per_year;
  launch_sia_sliding(h->dp, A, h->stream);
  HIP_TRY(h, hipGetLastError());
  // the host coordinates must stay valid until the copies ran
  if (sp->mode == SIA_SLIDING_TEST_E) HIP_TRY(h, hipStreamSynchronize(h->stream));
  return SIA_OK;
}

int sia_profiling(sia_handle *h, int on) {
  if (!h) return SIA_ERR_ARGUMENT;
  h->prof_on = on != 0;
  return SIA_OK;
}

int sia_profiling_read(sia_handle *h, double *ms, int64_t *count) {
  if (!h || !ms || !count) return SIA_ERR_ARGUMENT;
  for (int e = 0; e < SIA_EV_COUNT; e++) {
    ms[e] = 0.0;
    count[e] = 0;
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (auto &r : h->recs) {
    float t = 0.0f;
    HIP_TRY(h, hipEventElapsedTime(&t, r.a, r.b));
    ms[r.ev] += (double)t;
    count[r.ev] += 1;
    h->event_pool.push_back(r.a);
    h->event_pool.push_back(r.b);
  }
  h->recs.clear();
  return SIA_OK;
}

const char *sia_event_name(int event) {
  static const char *names[SIA_EV_COUNT] = {"SIA bed smoother", "SIA gradient",       "SIA flux",
                                            "SIA 3D hor. vel.", "SB vert. vel.",      "mass transport",
                                            "surrogate features", "surrogate MLP",    "SB strain heat"};
  return (event >= 0 && event < SIA_EV_COUNT) ? names[event] : "unknown";
}

const char *sia_last_error(const sia_handle *h) { return h ? h->last_error.c_str() : "NULL handle"; }


// ---------------------------------------------------------------- learned surrogate
int sia_surrogate_create(uint64_t seed, int64_t capacity, sia_surrogate **out) {
  if (!out || capacity < 0) return SIA_ERR_ARGUMENT;
  *out = nullptr;
  sia_surrogate *sg = new (std::nothrow) sia_surrogate();
  if (!sg) return SIA_ERR_OUT_OF_MEMORY;
  if (hipGetDevice(&sg->device) != hipSuccess) {
    delete sg;
    return SIA_ERR_HIP;
  }
  std::vector<float> w;
  sg_init_params(w, seed);
  bool ok = hipMalloc((void **)&sg->master, sizeof(float) * SG_NPARAM) == hipSuccess &&
            hipMalloc((void **)&sg->m, sizeof(float) * SG_NPARAM) == hipSuccess &&
            hipMalloc((void **)&sg->v, sizeof(float) * SG_NPARAM) == hipSuccess &&
            hipMalloc((void **)&sg->grad, sizeof(float) * SG_NPARAM) == hipSuccess &&
            hipMalloc((void **)&sg->frag, sizeof(unsigned short) * SG_NFRAG * 512) == hipSuccess &&
            hipMalloc((void **)&sg->d_n, sizeof(long long)) == hipSuccess &&
            (capacity == 0 || hipMalloc((void **)&sg->samples, sizeof(SgSample) * (size_t)capacity) == hipSuccess);
  if (ok) {
    const long long zero = 0;
    ok = hipMemcpy(sg->master, w.data(), sizeof(float) * SG_NPARAM, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(sg->m, 0, sizeof(float) * SG_NPARAM) == hipSuccess &&
         hipMemset(sg->v, 0, sizeof(float) * SG_NPARAM) == hipSuccess &&
         hipMemcpy(sg->d_n, &zero, sizeof(zero), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    sia_surrogate_destroy(sg);
    return SIA_ERR_OUT_OF_MEMORY;
  }
  launch_sg_pack(sg->master, sg->frag, nullptr);
  if (hipDeviceSynchronize() != hipSuccess) {
    sia_surrogate_destroy(sg);
    return SIA_ERR_HIP;
  }
  sg->cap = capacity;
  *out = sg;
  return SIA_OK;
}

int sia_surrogate_destroy(sia_surrogate *sg) {
  if (!sg) return SIA_OK;
  // a bound handle would read the freed model on its next fast update
  if (sg->bound > 0) return SIA_ERR_STATE;
  (void)hipDeviceSynchronize();
  void *bufs[] = {sg->master, sg->m, sg->v, sg->grad, sg->frag, sg->samples, sg->d_n, sg->d_loss};
  for (void *b : bufs) dfree(b);
  for (void *p : sg->tb_mem) dfree(p);
  delete sg;
  return SIA_OK;
}

int sia_surrogate_set_stream(sia_surrogate *sg, void *hip_stream) {
  if (!sg) return SIA_ERR_ARGUMENT;
  sg->stream = (hipStream_t)hip_stream;
  return SIA_OK;
}

int sia_surrogate_collect(sia_surrogate *sg, sia_handle *h, const sia_inputs *in, int stride, int64_t *n_total) {
  if (!sg || !h || !in || stride < 1) return set_err(h, SIA_ERR_ARGUMENT, "sia_surrogate_collect: bad argument");
  const sia_params &p = h->params;
  if (p.sia_flow_law != SIA_FLOW_GPBLD || p.enthalpy_converter != SIA_EC_DEFAULT) {
    return set_err(h, SIA_ERR_CONFIG, "the surrogate models GPBLD with the default enthalpy converter");
  }
  if (h->dp.nm1_int < 1) return set_err(h, SIA_ERR_CONFIG, "the surrogate needs n in {2, 3, 4}");
  if (sg->glen_n != 0.0 && sg->glen_n != p.sia_Glen_exponent) {
    return set_err(h, SIA_ERR_CONFIG, "samples with another Glen exponent are already held");
  }
  if (!sg->samples) return set_err(h, SIA_ERR_CONFIG, "the surrogate was created without sample capacity");
  // the FP64 path provides thksmooth, theta and the surface gradient
  sia_surrogate *saved = h->sg;
  h->sg = nullptr;
  int rc = update_impl(h, in, 1, true);
  h->sg = saved;
  if (rc) return rc;
  if ((rc = ensure_sg_buffers(h))) return rc;
  const SgColumnArgs C = sg_column_args(h, in);
  SgListArgs L = sg_list_args(h);
  L.strips = 0;  // samples: every stride-th column in staggered index order
  launch_sg_list(h->dp, C, L, h->stream);
  SgFeatureArgs F{};
  F.list = h->sg_list;
  F.n_list = h->sg_nlist;
  F.status = h->sg_status;
  F.samples = sg->samples;
  F.n_samples = sg->d_n;
  F.cap = sg->cap;
  F.stride = stride;
  launch_sg_features(h->dp, C, F, true, h->stream);
  HIP_TRY(h, hipGetLastError());
  int nlist = 0;
  HIP_TRY(h, hipMemcpyAsync(&nlist, h->sg_nlist, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int64_t nsel = (nlist + stride - 1) / stride;
  sg->n = std::min<int64_t>(sg->cap, sg->n + nsel);
  const long long nn = sg->n;
  HIP_TRY(h, hipMemcpy(sg->d_n, &nn, sizeof(nn), hipMemcpyHostToDevice));
  sg->glen_n = p.sia_Glen_exponent;
  if (n_total) *n_total = sg->n;
  return SIA_OK;
}

int sia_surrogate_fit_normalization(sia_surrogate *sg) {
  if (!sg) return SIA_ERR_ARGUMENT;
  if (sg->n < 2) return SIA_ERR_STATE;
  const size_t np = (size_t)SG_STAT_WAVES * SG_NIN * 4 + (size_t)SG_STAT_WAVES * 3;
  double *d = nullptr;
  if (hipMalloc((void **)&d, sizeof(double) * np) != hipSuccess) return SIA_ERR_OUT_OF_MEMORY;
  launch_sg_norm_stats(sg->samples, sg->n, d, sg->stream);
  std::vector<double> hbuf(np);
  const bool ok = hipMemcpyAsync(hbuf.data(), d, sizeof(double) * np, hipMemcpyDeviceToHost, sg->stream) == hipSuccess &&
                  hipStreamSynchronize(sg->stream) == hipSuccess;
  dfree(d);
  if (!ok) return SIA_ERR_HIP;
  const int nw = std::min<int64_t>(SG_STAT_WAVES, sg->n);  // waves past n hold no sample
  for (int f = 0; f < SG_NIN; f++) {
    double s = 0, s2 = 0, lo = 1e300, hi = -1e300;
    for (int w = 0; w < nw; w++) {
      const double *o = &hbuf[((size_t)w * SG_NIN + f) * 4];
      s += o[0];
      s2 += o[1];
      lo = std::min(lo, o[2]);
      hi = std::max(hi, o[3]);
    }
    const double mean = s / (double)sg->n;
    const double var = std::max(s2 / (double)sg->n - mean * mean, 0.0);
    const double sd = std::max(std::sqrt(var), 1e-6);
    const double margin = 0.02 * (hi - lo) + 1e-6;
    sg->norm.mean[f] = (float)mean;
    sg->norm.inv_std[f] = (float)(1.0 / sd);
    sg->norm.lo[f] = (float)(lo - margin);
    sg->norm.hi[f] = (float)(hi + margin);
  }
  // output unit: 4 x the RMS target log-correction (robust to rare outliers)
  double rmax = 0.0, r2 = 0.0, rn = 0.0;
  for (int w = 0; w < nw; w++) {
    const double *o = &hbuf[(size_t)SG_STAT_WAVES * SG_NIN * 4 + (size_t)w * 3];
    rmax = std::max(rmax, o[0]);
    r2 += o[1];
    rn += o[2];
  }
  const double rms = rn > 0.0 ? std::sqrt(r2 / rn) : 0.0;
  sg->norm.rscale = (float)(rms > 0.0 ? 4.0 * rms : (rmax > 0.0 ? rmax : 1.0));
  sg->have_norm = true;
  sg->version++;
  return SIA_OK;
}

int sia_surrogate_train(sia_surrogate *sg, int64_t steps, int batch, double lr_start, double lr_end, uint64_t seed,
                        double *loss_out) {
  if (!sg || steps < 1 || batch < 32 || batch % 32 != 0) return SIA_ERR_ARGUMENT;
  if (!sg->have_norm || sg->n < 1) return SIA_ERR_STATE;
  int rc;
  if ((rc = sg_train_buffers(sg, batch)) || (rc = sg_loss_buffer(sg, steps))) return rc;
  if (hipMemsetAsync(sg->d_loss, 0, sizeof(double) * (size_t)steps, sg->stream) != hipSuccess) return SIA_ERR_HIP;
  const double pi = 3.14159265358979323846;
  for (int64_t k = 0; k < steps; k++) {
    const double lr = lr_end + 0.5 * (lr_start - lr_end) * (1.0 + std::cos(pi * (double)k / (double)steps));
    launch_sg_gather(sg->samples, sg->n, sg->norm, seed, k, sg->tb, sg->stream);
    SgTrainBuffers T = sg->tb;
    T.loss = sg->d_loss + k;
    launch_sg_forward_backward(sg->master, T, sg->norm.rscale, 1e-3f, sg->stream);
    launch_sg_adam(sg->master, sg->m, sg->v, sg->grad, (float)lr, 0.9f, 0.999f, 1e-8f, ++sg->adam_t, sg->stream);
  }
  launch_sg_pack(sg->master, sg->frag, sg->stream);
  if (hipGetLastError() != hipSuccess) return SIA_ERR_HIP;
  if (loss_out) {
    if (hipMemcpyAsync(loss_out, sg->d_loss, sizeof(double) * (size_t)steps, hipMemcpyDeviceToHost, sg->stream) !=
        hipSuccess)
      return SIA_ERR_HIP;
  }
  if (hipStreamSynchronize(sg->stream) != hipSuccess) return SIA_ERR_HIP;
  return SIA_OK;
}

int sia_surrogate_get_params(const sia_surrogate *sg, float *host, size_t n) {
  if (!sg || !host || n != (size_t)SG_NPARAM) return SIA_ERR_ARGUMENT;
  if (hipMemcpy(host, sg->master, sizeof(float) * SG_NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return SIA_ERR_HIP;
  return SIA_OK;
}

int sia_surrogate_set_params(sia_surrogate *sg, const float *host, size_t n) {
  if (!sg || !host || n != (size_t)SG_NPARAM) return SIA_ERR_ARGUMENT;
  if (hipMemcpy(sg->master, host, sizeof(float) * SG_NPARAM, hipMemcpyHostToDevice) != hipSuccess) return SIA_ERR_HIP;
  launch_sg_pack(sg->master, sg->frag, nullptr);
  if (hipDeviceSynchronize() != hipSuccess) return SIA_ERR_HIP;
  return SIA_OK;
}

int sia_surrogate_grad(sia_surrogate *sg, int batch, const float *x, const float *r, const float *share,
                       const float *has, double *loss_host, float *grad_host) {
  if (!sg || !x || !r || !share || !has || batch < 32 || batch % 32 != 0) return SIA_ERR_ARGUMENT;
  int rc;
  if ((rc = sg_train_buffers(sg, batch)) || (rc = sg_loss_buffer(sg, 1))) return rc;
  const float rscale = sg->have_norm ? sg->norm.rscale : 1.0f;
  if (hipMemsetAsync(sg->d_loss, 0, sizeof(double), sg->stream) != hipSuccess) return SIA_ERR_HIP;
  launch_sg_set_inputs(x, r, share, has, batch, sg->tb, sg->stream);
  SgTrainBuffers T = sg->tb;
  T.loss = sg->d_loss;
  launch_sg_forward_backward(sg->master, T, rscale, 1e-3f, sg->stream);
  if (hipGetLastError() != hipSuccess) return SIA_ERR_HIP;
  if (loss_host && hipMemcpyAsync(loss_host, sg->d_loss, sizeof(double), hipMemcpyDeviceToHost, sg->stream) != hipSuccess)
    return SIA_ERR_HIP;
  if (grad_host &&
      hipMemcpyAsync(grad_host, sg->grad, sizeof(float) * SG_NPARAM, hipMemcpyDeviceToHost, sg->stream) != hipSuccess)
    return SIA_ERR_HIP;
  if (hipStreamSynchronize(sg->stream) != hipSuccess) return SIA_ERR_HIP;
  return SIA_OK;
}

int sia_surrogate_save(const sia_surrogate *sg, void *blob, size_t *size) {
  if (!sg || !size) return SIA_ERR_ARGUMENT;
  const size_t need = sizeof(SgBlobHeader) + sizeof(SgNorm) + sizeof(float) * (SG_NPARAM + 1);
  if (!blob) {
    *size = need;
    return SIA_OK;
  }
  if (*size < need) return SIA_ERR_ARGUMENT;
  SgBlobHeader hdr{};
  std::memcpy(hdr.magic, SG_MAGIC, 8);
  hdr.nparam = SG_NPARAM;
  hdr.nin = SG_NIN;
  hdr.glen_n = sg->glen_n;
  hdr.have_norm = sg->have_norm ? 1 : 0;
  hdr.has_tau = 1;
  char *b = (char *)blob;
  std::memcpy(b, &hdr, sizeof(hdr));
  std::memcpy(b + sizeof(hdr), &sg->norm, sizeof(SgNorm));
  if (hipMemcpy(b + sizeof(hdr) + sizeof(SgNorm), sg->master, sizeof(float) * SG_NPARAM, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return SIA_ERR_HIP;
  std::memcpy(b + sizeof(hdr) + sizeof(SgNorm) + sizeof(float) * SG_NPARAM, &sg->tau, sizeof(float));
  *size = need;
  return SIA_OK;
}

int sia_surrogate_load(sia_surrogate *sg, const void *blob, size_t size) {
  if (!sg || !blob) return SIA_ERR_ARGUMENT;
  const size_t need = sizeof(SgBlobHeader) + sizeof(SgNorm) + sizeof(float) * SG_NPARAM;
  if (size < need) return SIA_ERR_ARGUMENT;
  SgBlobHeader hdr;
  std::memcpy(&hdr, blob, sizeof(hdr));
  if (std::memcmp(hdr.magic, SG_MAGIC, 8) != 0 || hdr.nparam != SG_NPARAM || hdr.nin != SG_NIN) {
    return SIA_ERR_CONFIG;
  }
  const char *b = (const char *)blob;
  std::memcpy(&sg->norm, b + sizeof(hdr), sizeof(SgNorm));
  sg->have_norm = hdr.have_norm != 0;
  sg->tau = INFINITY;
  if (hdr.has_tau && size >= need + sizeof(float)) {
    std::memcpy(&sg->tau, b + sizeof(hdr) + sizeof(SgNorm) + sizeof(float) * SG_NPARAM, sizeof(float));
  }
  sg->version++;
  sg->glen_n = hdr.glen_n;
  return sia_surrogate_set_params(sg, (const float *)(b + sizeof(hdr) + sizeof(SgNorm)), SG_NPARAM);
}

int sia_surrogate_samples(const sia_surrogate *sg, const void **samples, int64_t *n) {
  if (!sg || !samples || !n) return SIA_ERR_ARGUMENT;
  *samples = sg->samples;
  *n = sg->n;
  return SIA_OK;
}

// binds sg to h (nullptr: unbinds), keeping the models' bound-handle counts
static void bind_surrogate(sia_handle *h, sia_surrogate *sg) {
  if (h->sg) h->sg->bound--;
  h->sg = sg;
  if (sg) sg->bound++;
}

int sia_use_surrogate(sia_handle *h, sia_surrogate *sg) {
  if (!h) return SIA_ERR_ARGUMENT;
  if (!sg) {
    bind_surrogate(h, nullptr);
    return SIA_OK;
  }
  const sia_params &p = h->params;
  if (!sg->have_norm) return set_err(h, SIA_ERR_CONFIG, "the surrogate has no normalization (not trained)");
  if (p.sia_flow_law != SIA_FLOW_GPBLD || p.enthalpy_converter != SIA_EC_DEFAULT) {
    return set_err(h, SIA_ERR_CONFIG, "the surrogate models GPBLD with the default enthalpy converter");
  }
  if (p.sia_Glen_exponent != sg->glen_n || h->dp.nm1_int < 1) {
    return set_err(h, SIA_ERR_CONFIG, "the surrogate was trained for another Glen exponent");
  }
  if (sg->device != h->device) return set_err(h, SIA_ERR_CONFIG, "the surrogate lives on another device");
  int rc;
  if ((rc = ensure_sg_buffers(h))) return rc;
  bind_surrogate(h, sg);
  return SIA_OK;
}

int sia_surrogate_set_fallback_score(sia_surrogate *sg, double tau) {
  if (!sg || std::isnan(tau) || tau < 0.0) return SIA_ERR_ARGUMENT;
  sg->tau = (float)tau;
  sg->version++;
  return SIA_OK;
}

// a surrogate update of `in` on h with every in-range column kept and its
// fallback score recorded; tau = the (1 - fraction) quantile of those scores
int sia_surrogate_calibrate_fallback(sia_surrogate *sg, sia_handle *h, const sia_inputs *in, double fraction,
                                     double *tau) {
  if (!h) return SIA_ERR_ARGUMENT;
  if (!sg || !in || !(fraction > 0.0 && fraction < 1.0)) {
    return set_err(h, SIA_ERR_ARGUMENT, "sia_surrogate_calibrate_fallback: bad argument");
  }
  sia_surrogate *saved = h->sg;
  int rc = sia_use_surrogate(h, sg);
  if (rc == SIA_OK) rc = dalloc(h, &h->sg_score, 2 * (size_t)h->grid.Mx * h->grid.My);
  if (rc == SIA_OK) {
    h->sg_calibrating = true;
    rc = update_impl(h, in, 1, true);
    h->sg_calibrating = false;
  }
  bind_surrogate(h, saved);
  if (rc) return rc;
  int nlist = 0;
  HIP_TRY(h, hipMemcpyAsync(&nlist, h->sg_nlist, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  std::vector<float> sc((size_t)nlist);
  if (nlist > 0) HIP_TRY(h, hipMemcpy(sc.data(), h->sg_score, sizeof(float) * (size_t)nlist, hipMemcpyDeviceToHost));
  sc.erase(std::remove_if(sc.begin(), sc.end(), [](float v) { return !(v >= 0.0f); }), sc.end());
  if (sc.empty()) return set_err(h, SIA_ERR_STATE, "calibration: the surrogate handles no column of these inputs");
  const size_t k = std::min(sc.size() - 1, (size_t)std::floor((1.0 - fraction) * (double)sc.size()));
  std::nth_element(sc.begin(), sc.begin() + (long)k, sc.end());
  sg->tau = sc[k];
  sg->version++;
  if (tau) *tau = sg->tau;
  return SIA_OK;
}

int sia_surrogate_counts(sia_handle *h, int64_t counts[3]) {
  if (!h || !counts) return SIA_ERR_ARGUMENT;
  if (!h->sg_counts) return set_err(h, SIA_ERR_STATE, "no surrogate update has run");
  int c[3];
  HIP_TRY(h, hipMemcpyAsync(c, h->sg_counts, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int k = 0; k < 3; k++) counts[k] = c[k];
  return SIA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- ensembles
struct sia_ensemble {
  int device = 0, n = 0, law = 0, max_mz = 0, cpw = 16;
  std::vector<sia_handle *> members;
  std::vector<sia_geometry> geom;
  std::vector<EnsMember> table;        // host copy; A.dmax_part / n_part follow cpw
  EnsMember *d_table = nullptr;
  std::vector<int> pt_start, col_start[3];
  int *d_pt_start = nullptr, *d_col_start[3] = {nullptr, nullptr, nullptr};
  double *d_t = nullptr, *d_dt = nullptr, *d_part = nullptr;
  int *d_err_acc = nullptr;
  std::vector<size_t> part_off;        // member offsets into d_part (sized for 8 columns per wave + MLP)
  // members bound to a surrogate (sia_use_surrogate on the member handle before
  // sia_ensemble_create): one group per surrogate, whose stages run between
  // the gradient and the (fallback-only) column stage
  struct SgGroup {
    sia_surrogate *sg = nullptr;
    std::vector<int> idx;               // ensemble member indices
    std::vector<EnsSgMember> table;
    EnsSgMember *d_table = nullptr;
    std::vector<EnsListMember> ltable;  // the FP64 pass over each member's fallback list
    EnsListMember *d_ltable = nullptr;
    std::vector<int> start[4];          // block prefixes: classify, features, MLP, fallback list
    int *d_start[4] = {nullptr, nullptr, nullptr, nullptr};
    int feat_waves = 4, nm1 = 2, max_mz = 0;
    int version = 0;                    // sg->version the table's normalization copies
  };
  std::vector<SgGroup> groups;
  std::vector<sia_surrogate *> bound;   // each member's surrogate at creation (nullptr: FP64)
  std::vector<int> mlp_part;            // MLP D_max partials of each member
  std::vector<int> lst_part;            // fallback-list D_max partials of each member
  hipStream_t own = nullptr, user = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0, graph_cpw = 0;
  std::string last_error;
};

namespace {

const int ENS_CPW[3] = {8, 16, 32};

int cpw_slot(int cpw) { return cpw == 8 ? 0 : (cpw == 32 ? 2 : 1); }

int ens_err(sia_ensemble *e, int code, const std::string &msg) {
  if (e) e->last_error = msg;
  return code;
}

#define ENS_TRY(e, call)                                                                       \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return ens_err((e), SIA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// member table for the current tile width, uploaded on the ensemble's stream
int ens_upload(sia_ensemble *e) {
  const int c = cpw_slot(e->cpw);
  // member partials: column stage (none for surrogate-bound members), then MLP, then fallback list
  for (int i = 0; i < e->n; i++) {
    const int ncol = e->col_start[c][i + 1] - e->col_start[c][i];
    e->table[i].A.dmax_part = e->d_part + e->part_off[i];
    e->table[i].n_part = ncol + e->mlp_part[i] + e->lst_part[i];
  }
  for (auto &g : e->groups) {
    for (size_t k = 0; k < g.idx.size(); k++) {
      const int i = g.idx[k];
      const int ncol = e->col_start[c][i + 1] - e->col_start[c][i];
      g.table[k].M.dmax_part = e->table[i].A.dmax_part + ncol;
      g.ltable[k].A.dmax_part = e->table[i].A.dmax_part + ncol + e->mlp_part[i];
    }
    ENS_TRY(e, hipMemcpyAsync(g.d_table, g.table.data(), sizeof(EnsSgMember) * g.table.size(),
                              hipMemcpyHostToDevice, e->own));
    ENS_TRY(e, hipMemcpyAsync(g.d_ltable, g.ltable.data(), sizeof(EnsListMember) * g.ltable.size(),
                              hipMemcpyHostToDevice, e->own));
  }
  ENS_TRY(e, hipMemcpyAsync(e->d_table, e->table.data(), sizeof(EnsMember) * e->n, hipMemcpyHostToDevice, e->own));
  ENS_TRY(e, hipStreamSynchronize(e->own));
  return SIA_OK;
}

EnsLaunch ens_launch(const sia_ensemble *e) {
  const int c = cpw_slot(e->cpw);
  EnsLaunch L{};
  L.table = e->d_table;
  L.pt_start = e->d_pt_start;
  L.col_start = e->d_col_start[c];
  L.nm = e->n;
  L.pt_blocks = e->pt_start[e->n];
  L.col_blocks = e->col_start[c][e->n];
  L.cpw = e->cpw;
  L.law = e->law;
  L.max_mz = e->max_mz;
  return L;
}

// one ensemble step: gradient, the surrogate stages of each group, then the
// column stage (fallback columns only for surrogate-bound members), dt, mass
// continuity and the surface/mask update
void ens_issue_step(const sia_ensemble *e, const EnsLaunch &L, hipStream_t st) {
  launch_ens_front(L, st);
  for (const auto &g : e->groups) {
    EnsSgLaunch G{};
    G.table = g.d_table;
    G.cls_start = g.d_start[0];
    G.feat_start = g.d_start[1];
    G.mlp_start = g.d_start[2];
    G.nm = (int)g.idx.size();
    G.cls_blocks = g.start[0].back();
    G.feat_blocks = g.start[1].back();
    G.mlp_blocks = g.start[2].back();
    G.feat_waves = g.feat_waves;
    G.nm1 = g.nm1;
    G.max_mz = g.max_mz;
    launch_ens_sg(G, st);
    launch_ens_column_list(g.d_ltable, g.d_start[3], (int)g.idx.size(), g.start[3].back(), g.max_mz, e->law, st);
  }
  launch_ens_back(L, st);
}

void ens_drop_graph(sia_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->graph_steps = 0;
}

void ens_mark_members(sia_ensemble *e) {
  for (int i = 0; i < e->n; i++) {
    sia_handle *h = e->members[i];
    h->have_flux = true;
    h->have_3d = false;
    h->last_ub = nullptr;
    h->last_vb = nullptr;
    h->last_mask = e->geom[i].mask;
  }
}

}  // namespace

extern "C" {

int sia_ensemble_create(sia_handle *const *members, const sia_geometry *geometry, const double *const *enthalpy,
                        const double *t_end, int n, sia_ensemble **out) {
  if (!members || !geometry || !enthalpy || !out || n < 1) return SIA_ERR_ARGUMENT;
  *out = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SIA_ERR_HIP;
  sia_ensemble *e = new (std::nothrow) sia_ensemble();
  if (!e) return SIA_ERR_OUT_OF_MEMORY;
  e->device = dev;
  e->n = n;
  e->members.assign(members, members + n);
  e->geom.assign(geometry, geometry + n);
  e->table.resize(n);
  e->pt_start.assign(n + 1, 0);
  for (int c = 0; c < 3; c++) e->col_start[c].assign(n + 1, 0);
  e->part_off.assign(n, 0);
  e->bound.assign(n, nullptr);
  e->mlp_part.assign(n, 0);
  e->lst_part.assign(n, 0);
  size_t part_total = 0;
  long long pt_total = 0, col_total[3] = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    sia_handle *h = members[i];
    const sia_geometry &g = geometry[i];
    if (!h || !enthalpy[i] || !g.land_ice_thickness || !g.surface_altitude || !g.mask || !g.bedrock_altitude) {
      delete e;
      return SIA_ERR_ARGUMENT;
    }
    if (h->device != dev) {
      e->last_error = "member " + std::to_string(i) + " lives on another device";
      sia_ensemble_destroy(e);
      return SIA_ERR_ARGUMENT;
    }
    if (i == 0) e->law = h->params.sia_flow_law;
    if (h->params.sia_flow_law != e->law ||
        (law_uses_grain_size(h->params.sia_flow_law) && h->params.compute_grain_size_using_age)) {
      delete e;
      return SIA_ERR_CONFIG;
    }
    if (!h->bed_ready) {
      delete e;
      return SIA_ERR_STATE;
    }
    e->max_mz = std::max(e->max_mz, h->grid.Mz);
    e->bound[i] = h->sg;
    if (h->sg) {
      size_t gi = 0;
      while (gi < e->groups.size() && e->groups[gi].sg != h->sg) gi++;
      if (gi == e->groups.size()) {
        e->groups.emplace_back();
        e->groups.back().sg = h->sg;
        e->groups.back().nm1 = h->dp.nm1_int;
        e->groups.back().version = h->sg->version;
      }
      e->groups[gi].idx.push_back(i);
      e->groups[gi].max_mz = std::max(e->groups[gi].max_mz, h->grid.Mz);
    }
    e->pt_start[i] = (int)pt_total;
    pt_total += ens_point_blocks(h->dp);
    for (int c = 0; c < 3; c++) {
      // a surrogate-bound member's FP64 columns run in its group's fallback-list pass
      e->col_start[c][i] = (int)col_total[c];
      col_total[c] += h->sg ? 0 : ens_column_blocks(h->dp, ENS_CPW[c]);
    }
  }
  // surrogate groups: per-member block shares of the feature pass and the MLP
  // in proportion to the member's staggered columns (at least one block each)
  for (auto &g : e->groups) {
    long long tot = 0;
    for (int i : g.idx) tot += 2LL * members[i]->grid.Mx * members[i]->grid.My;
    g.feat_waves = sg_features_waves(g.max_mz);
    const int feat_total = 8192 / g.feat_waves;
    for (int k = 0; k < 4; k++) g.start[k].assign(1, 0);
    for (int i : g.idx) {
      const double share = (double)(2LL * members[i]->grid.Mx * members[i]->grid.My) / (double)tot;
      const int nb[4] = {sg_classify_blocks(members[i]->dp), std::max(1, (int)std::lround(feat_total * share)),
                         std::max(1, (int)std::lround(sg_mlp_blocks() * share)),
                         std::max(1, (int)std::lround(1024 * share))};
      for (int k = 0; k < 4; k++) g.start[k].push_back(g.start[k].back() + nb[k]);
      e->mlp_part[i] = nb[2];
      e->lst_part[i] = nb[3];
    }
  }
  for (int i = 0; i < n; i++) {
    e->part_off[i] = part_total;
    int ncol_max = 0;
    for (int c = 0; c < 3; c++) ncol_max = std::max(ncol_max, ens_column_blocks(members[i]->dp, ENS_CPW[c]));
    part_total += (size_t)(members[i]->sg ? 0 : ncol_max) + (size_t)e->mlp_part[i] +
                  (size_t)e->lst_part[i];
  }
  if (pt_total >= 2147483647LL || col_total[0] >= 2147483647LL) {
    delete e;
    return SIA_ERR_ARGUMENT;
  }
  e->pt_start[n] = (int)pt_total;
  for (int c = 0; c < 3; c++) e->col_start[c][n] = (int)col_total[c];
  auto fail = [&](hipError_t err, const char *what) {
    e->last_error = std::string(what) + ": " + hipGetErrorString(err);
    sia_ensemble_destroy(e);
    return SIA_ERR_HIP;
  };
  hipError_t err;
  if ((err = hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking)) != hipSuccess) return fail(err, "stream");
  if ((err = hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming)) != hipSuccess ||
      (err = hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming)) != hipSuccess) {
    return fail(err, "events");
  }
  if ((err = hipMalloc((void **)&e->d_table, sizeof(EnsMember) * n)) != hipSuccess ||
      (err = hipMalloc((void **)&e->d_pt_start, sizeof(int) * (n + 1))) != hipSuccess ||
      (err = hipMalloc((void **)&e->d_t, sizeof(double) * n)) != hipSuccess ||
      (err = hipMalloc((void **)&e->d_dt, sizeof(double) * n)) != hipSuccess ||
      (err = hipMalloc((void **)&e->d_err_acc, sizeof(int) * n)) != hipSuccess ||
      (err = hipMalloc((void **)&e->d_part, sizeof(double) * part_total)) != hipSuccess) {
    return fail(err, "hipMalloc");
  }
  for (int c = 0; c < 3; c++) {
    if ((err = hipMalloc((void **)&e->d_col_start[c], sizeof(int) * (n + 1))) != hipSuccess) return fail(err, "hipMalloc");
    if ((err = hipMemcpy(e->d_col_start[c], e->col_start[c].data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice)) !=
        hipSuccess) {
      return fail(err, "hipMemcpy");
    }
  }
  if ((err = hipMemcpy(e->d_pt_start, e->pt_start.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (err = hipMemset(e->d_t, 0, sizeof(double) * n)) != hipSuccess ||
      (err = hipMemset(e->d_dt, 0, sizeof(double) * n)) != hipSuccess ||
      (err = hipMemset(e->d_err_acc, 0, sizeof(int) * n)) != hipSuccess) {
    return fail(err, "hipMemcpy");
  }
  for (int i = 0; i < n; i++) {
    sia_handle *h = members[i];
    const sia_geometry &g = geometry[i];
    EnsMember &M = e->table[i];
    M = EnsMember{};
    M.P = h->dp;
    M.P.use_age = 0;
    M.F = Fields2D{g.land_ice_thickness, g.bedrock_altitude, g.surface_altitude, g.mask};
    M.topgs = h->topgs;
    M.maxtl = h->maxtl;
    M.C2 = h->C2;
    M.C3 = h->C3;
    M.C4 = h->C4;
    M.hx = h->hx;
    M.hy = h->hy;
    M.theta = h->theta;
    M.thks = h->thks;
    M.A.z = h->d_z;
    M.A.E = enthalpy[i];
    M.A.thks = h->thks;
    M.A.theta = h->theta;
    M.A.hx = h->hx;
    M.A.hy = h->hy;
    M.A.flux = h->flux;
    M.A.Dstag = h->Dstag;
    M.A.dmax_bits = h->dmax_bits;
    M.A.err = h->err;
    M.M.dt = e->d_dt + i;
    M.M.H = g.land_ice_thickness;
    M.M.flux = h->flux;
    M.M.mask = g.mask;
    M.M.smb = g.climatic_mass_balance;
    M.M.bmr = g.basal_melt_rate;
    M.M.H_new = h->H_tmp;
    M.H = g.land_ice_thickness;
    M.mask = g.mask;
    M.h = g.surface_altitude;
    M.cfl_bits = h->cfl_bits;
    M.t = e->d_t + i;
    M.dt = e->d_dt + i;
    M.err_acc = e->d_err_acc + i;
    M.t_end = t_end ? t_end[i] : 0.0;
  }
  for (auto &g : e->groups) {
    for (int i : g.idx) {
      sia_handle *h = members[i];
      if (ensure_sg_buffers(h) != SIA_OK) {
        e->last_error = h->last_error;
        sia_ensemble_destroy(e);
        return SIA_ERR_OUT_OF_MEMORY;
      }
      EnsSgMember S{};
      S.P = h->dp;
      S.C = SgColumnArgs{h->d_z, enthalpy[i], h->thks, h->theta, h->hx, h->hy, h->err};
      S.L = sg_list_args(h);
      S.F.list = h->sg_list;
      S.F.n_list = h->sg_nlist;
      S.F.F = h->sg_F;
      S.F.CQb = h->sg_CQb;
      S.F.CSt = h->sg_CSt;
      S.F.status = h->sg_status;
      S.F.stride = 1;
      S.M.list = h->sg_list;
      S.M.n_list = h->sg_nlist;
      S.M.F = h->sg_F;
      S.M.CQb = h->sg_CQb;
      S.M.CSt = h->sg_CSt;
      S.M.frag = g.sg->frag;
      S.M.bias = g.sg->master;
      S.M.norm = g.sg->norm;
      S.M.hx = h->hx;
      S.M.hy = h->hy;
      S.M.status = h->sg_status;
      S.M.flux = h->flux;
      S.M.Dstag = h->Dstag;
      S.M.counts = h->sg_counts;
      S.M.tau = g.sg->tau;
      S.M.score = nullptr;
      S.M.fb_list = h->sg_fb_list;
      S.M.n_fb = h->sg_nfb;
      g.table.push_back(S);
      EnsListMember Lm{};
      Lm.P = e->table[i].P;
      Lm.A = e->table[i].A;
      Lm.list = h->sg_fb_list;
      Lm.n_list = h->sg_nfb;
      g.ltable.push_back(Lm);
    }
    if ((err = hipMalloc((void **)&g.d_table, sizeof(EnsSgMember) * g.table.size())) != hipSuccess ||
        (err = hipMalloc((void **)&g.d_ltable, sizeof(EnsListMember) * g.ltable.size())) != hipSuccess) {
      return fail(err, "hipMalloc");
    }
    for (int k = 0; k < 4; k++) {
      const size_t b = sizeof(int) * g.start[k].size();
      if ((err = hipMalloc((void **)&g.d_start[k], b)) != hipSuccess ||
          (err = hipMemcpy(g.d_start[k], g.start[k].data(), b, hipMemcpyHostToDevice)) != hipSuccess) {
        return fail(err, "hipMalloc");
      }
    }
  }
  int rc = ens_upload(e);
  if (rc) {
    sia_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return SIA_OK;
}

int sia_ensemble_destroy(sia_ensemble *e) {
  if (!e) return SIA_OK;
  if (e->own) (void)hipStreamSynchronize(e->own);
  ens_drop_graph(e);
  void *bufs[] = {e->d_table, e->d_pt_start, e->d_col_start[0], e->d_col_start[1], e->d_col_start[2],
                  e->d_t, e->d_dt, e->d_err_acc, e->d_part};
  for (void *b : bufs) dfree(b);
  for (auto &g : e->groups) {
    dfree(g.d_table);
    dfree(g.d_ltable);
    for (int k = 0; k < 4; k++) dfree(g.d_start[k]);
  }
  if (e->ev_in) (void)hipEventDestroy(e->ev_in);
  if (e->ev_out) (void)hipEventDestroy(e->ev_out);
  if (e->own) (void)hipStreamDestroy(e->own);
  delete e;
  return SIA_OK;
}

int sia_ensemble_set_stream(sia_ensemble *e, void *hip_stream) {
  if (!e) return SIA_ERR_ARGUMENT;
  e->user = (hipStream_t)hip_stream;
  return SIA_OK;
}

int sia_ensemble_set_tile(sia_ensemble *e, int cols_per_wave) {
  if (!e) return SIA_ERR_ARGUMENT;
  if (cols_per_wave != 8 && cols_per_wave != 16 && cols_per_wave != 32) {
    return ens_err(e, SIA_ERR_ARGUMENT, "columns per wavefront must be 8, 16 or 32");
  }
  if (cols_per_wave == e->cpw) return SIA_OK;
  // the previous steps may still read the table
  ENS_TRY(e, hipStreamSynchronize(e->own));
  e->cpw = cols_per_wave;
  return ens_upload(e);
}

int sia_ensemble_step(sia_ensemble *e, int n_steps, int use_graph) {
  if (!e || n_steps < 0) return SIA_ERR_ARGUMENT;
  if (n_steps == 0) return SIA_OK;
  for (int i = 0; i < e->n; i++) {
    if (e->members[i]->sg != e->bound[i]) {
      return ens_err(e, SIA_ERR_STATE, "member " + std::to_string(i) +
                                           ": its surrogate binding changed after sia_ensemble_create");
    }
  }
  bool renorm = false;
  for (auto &g : e->groups) {
    if (g.version != g.sg->version) {
      for (auto &S : g.table) {
        S.M.norm = g.sg->norm;
        S.M.tau = g.sg->tau;
      }
      g.version = g.sg->version;
      renorm = true;
    }
  }
  if (renorm) {
    ENS_TRY(e, hipStreamSynchronize(e->own));  // earlier steps may still read the tables
    const int rc = ens_upload(e);
    if (rc) return rc;
  }
  ENS_TRY(e, hipEventRecord(e->ev_in, e->user));
  ENS_TRY(e, hipStreamWaitEvent(e->own, e->ev_in, 0));
  const EnsLaunch L = ens_launch(e);
  if (use_graph) {
    if (!e->exec || e->graph_steps != n_steps || e->graph_cpw != e->cpw) {
      ens_drop_graph(e);
      ENS_TRY(e, hipStreamBeginCapture(e->own, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < n_steps; s++) ens_issue_step(e, L, e->own);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(e->own, &g);
      if (ce != hipSuccess) return ens_err(e, SIA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      e->graph = g;
      ENS_TRY(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
      e->graph_steps = n_steps;
      e->graph_cpw = e->cpw;
    }
    ENS_TRY(e, hipGraphLaunch(e->exec, e->own));
  } else {
    for (int s = 0; s < n_steps; s++) ens_issue_step(e, L, e->own);
    ENS_TRY(e, hipGetLastError());
  }
  ENS_TRY(e, hipEventRecord(e->ev_out, e->own));
  ENS_TRY(e, hipStreamWaitEvent(e->user, e->ev_out, 0));
  ens_mark_members(e);
  return SIA_OK;
}

int sia_ensemble_autotune(sia_ensemble *e, int steps, int *chosen, double *ms) {
  if (!e || steps < 1) return SIA_ERR_ARGUMENT;
  // save every member's geometry, time and dt
  std::vector<double *> saved(e->n, nullptr);
  double *t_save = nullptr;
  auto release = [&]() {
    for (double *p : saved) dfree(p);
    dfree(t_save);
  };
  ENS_TRY(e, hipEventRecord(e->ev_in, e->user));
  ENS_TRY(e, hipStreamWaitEvent(e->own, e->ev_in, 0));
  hipError_t err = hipMalloc((void **)&t_save, sizeof(double) * 2 * e->n);
  for (int i = 0; err == hipSuccess && i < e->n; i++) {
    const size_t n2 = (size_t)e->table[i].P.Mx * e->table[i].P.My;
    err = hipMalloc((void **)&saved[i], sizeof(double) * 3 * n2);
  }
  if (err != hipSuccess) {
    release();
    return ens_err(e, SIA_ERR_OUT_OF_MEMORY, "autotune: no room to save the ensemble state");
  }
  auto copy_state = [&](bool save) -> hipError_t {
    hipError_t r = hipSuccess;
    for (int i = 0; r == hipSuccess && i < e->n; i++) {
      const size_t n2 = (size_t)e->table[i].P.Mx * e->table[i].P.My, b = n2 * sizeof(double);
      double *f[3] = {e->geom[i].land_ice_thickness, e->geom[i].surface_altitude, e->geom[i].mask};
      for (int k = 0; r == hipSuccess && k < 3; k++) {
        r = save ? hipMemcpyAsync(saved[i] + k * n2, f[k], b, hipMemcpyDeviceToDevice, e->own)
                 : hipMemcpyAsync(f[k], saved[i] + k * n2, b, hipMemcpyDeviceToDevice, e->own);
      }
    }
    if (r == hipSuccess) {
      r = save ? hipMemcpyAsync(t_save, e->d_t, sizeof(double) * e->n, hipMemcpyDeviceToDevice, e->own)
               : hipMemcpyAsync(e->d_t, t_save, sizeof(double) * e->n, hipMemcpyDeviceToDevice, e->own);
    }
    if (r == hipSuccess) {
      r = save ? hipMemcpyAsync(t_save + e->n, e->d_dt, sizeof(double) * e->n, hipMemcpyDeviceToDevice, e->own)
               : hipMemcpyAsync(e->d_dt, t_save + e->n, sizeof(double) * e->n, hipMemcpyDeviceToDevice, e->own);
    }
    return r;
  };
  hipEvent_t a = nullptr, b = nullptr;
  double best = 1e300, res[3] = {0, 0, 0};
  int best_cpw = e->cpw;
  err = copy_state(true);
  if (err == hipSuccess) err = hipEventCreate(&a);
  if (err == hipSuccess) err = hipEventCreate(&b);
  for (int c = 0; err == hipSuccess && c < 3; c++) {
    if (sia_ensemble_set_tile(e, ENS_CPW[c]) != SIA_OK) {
      err = hipErrorUnknown;
      break;
    }
    const EnsLaunch L = ens_launch(e);
    ens_issue_step(e, L, e->own);  // warm-up
    err = hipEventRecord(a, e->own);
    for (int s = 0; err == hipSuccess && s < steps; s++) ens_issue_step(e, L, e->own);
    if (err == hipSuccess) err = hipEventRecord(b, e->own);
    if (err == hipSuccess) err = hipEventSynchronize(b);
    float t = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, a, b);
    res[c] = (double)t / steps;
    if (res[c] < best) {
      best = res[c];
      best_cpw = ENS_CPW[c];
    }
    if (err == hipSuccess) err = copy_state(false);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (err == hipSuccess) err = hipStreamSynchronize(e->own);
  release();
  if (err != hipSuccess) return ens_err(e, SIA_ERR_HIP, std::string("autotune: ") + hipGetErrorString(err));
  int rc = sia_ensemble_set_tile(e, best_cpw);
  if (rc) return rc;
  if (chosen) *chosen = best_cpw;
  if (ms) std::copy(res, res + 3, ms);
  ENS_TRY(e, hipEventRecord(e->ev_out, e->own));
  ENS_TRY(e, hipStreamWaitEvent(e->user, e->ev_out, 0));
  return SIA_OK;
}

int sia_ensemble_time(sia_ensemble *e, int member, const double **t_device, const double **dt_device) {
  if (!e || member < 0 || member >= e->n) return SIA_ERR_ARGUMENT;
  if (t_device) *t_device = e->d_t + member;
  if (dt_device) *dt_device = e->d_dt + member;
  return SIA_OK;
}

int sia_ensemble_check_errors(sia_ensemble *e) {
  if (!e) return SIA_ERR_ARGUMENT;
  ENS_TRY(e, hipStreamSynchronize(e->own));
  std::vector<int> acc(e->n, 0);
  ENS_TRY(e, hipMemcpy(acc.data(), e->d_err_acc, sizeof(int) * e->n, hipMemcpyDeviceToHost));
  for (int i = 0; i < e->n; i++) {
    int now = 0;
    ENS_TRY(e, hipMemcpy(&now, e->members[i]->err, sizeof(int), hipMemcpyDeviceToHost));
    const int f = acc[i] | now;
    const std::string who = "member " + std::to_string(i) + ": ";
    if (f & ERR_NEG_THK) return ens_err(e, SIA_ERR_NEGATIVE_THICKNESS, who + "negative ice thickness");
    if (f & ERR_THK_LZ) return ens_err(e, SIA_ERR_THICKNESS_ABOVE_LZ, who + "ice thickness is above the top of the computational grid");
    if (f & ERR_OMEGA) return ens_err(e, SIA_ERR_OMEGA_NEGATIVE, who + "omega is negative in BedSmoother::get_theta()");
  }
  return SIA_OK;
}

const char *sia_ensemble_last_error(const sia_ensemble *e) { return e ? e->last_error.c_str() : "NULL ensemble"; }

}  // extern "C"
