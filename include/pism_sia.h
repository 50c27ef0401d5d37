/*
 * pism_sia.h -- C ABI of the MI355X (gfx950) shallow-ice (SIA) stress balance.
 *
 * This is the drop-in boundary for PISM's SIA hot path. A maintainer replaces
 * the C++ objects below with the calls on the right; every call is plain C,
 * takes plain pointers and sizes, and returns an int status (no exception
 * crosses the boundary).
 *
 *   reference interface (PISM v0.7.3)                         replacement
 *   ---------------------------------------------------------------------------
 *   SIAFD::SIAFD(grid, EC) + SIAFD::init()                    sia_create
 *     src/base/stressbalance/sia/SIAFD.cc:35-98
 *   SIAFD::~SIAFD()  (SIAFD.cc:77-83)                          sia_destroy
 *   SSB_Modifier::update(vel_input, fast) = SIAFD::update      sia_update
 *     src/base/stressbalance/SSB_Modifier.hh:44; SIAFD.cc:119-148
 *     plus StressBalance::compute_vertical_velocity (w) when !fast
 *     src/base/stressbalance/PISMStressBalance.cc:98-110, 192-287
 *   SSB_Modifier::diffusive_flux()  (SSB_Modifier.hh:47)       sia_diffusive_flux
 *   SSB_Modifier::max_diffusivity() (SSB_Modifier.hh:50)       sia_max_diffusivity
 *   SSB_Modifier::velocity_u()/velocity_v() (SSB_Modifier.hh:52-54)
 *                                                              sia_velocity_u/_v
 *   StressBalance::velocity_w() (PISMStressBalance.cc:138-140) sia_velocity_w
 *   StressBalance::volumetric_strain_heating() (PISMStressBalance.cc:146-148,
 *     computed on !fast updates, 102-104, 370-500)            sia_volumetric_strain_heating
 *   SIAFD_Regional::compute_surface_gradient (regional.cc:45-109)
 *                                                              sia_inputs.no_model_mask + usurfstore
 *   SIA_Sliding::update (SIA_Sliding.cc:106-206)              sia_sliding_update
 *   BedSmoother::preprocess_bed (PISMBedSmoother.cc:96-149),
 *     triggered by the bed state counter (SIAFD.cc:128-133)   sia_update (bed_state_counter)
 *   SIAFD diagnostics h_x, h_y, diffusivity_staggered, schoofs_theta,
 *     thksmooth, topgsmooth (SIAFD_diagnostics.cc:26-35)      sia_surface_gradient,
 *                                                              sia_diffusivity_staggered,
 *                                                              sia_bed_smoother_fields,
 *                                                              sia_theta_thksmooth
 *   IceModel::max_timestep_diffusivity (iMadaptive.cc:157-171)
 *     and the dt choice of IceModel::max_timestep (214-388)    sia_max_timestep
 *   IceModel::massContExplicitStep (iMgeometry.cc:497-822)
 *     + updateSurfaceElevationAndMask (iMgeometry.cc:56-149)  sia_mass_continuity_step
 *   one SIA-only IceModel::step (iceModel.cc:601-794)          sia_step
 *
 * Data layout (the reference's PETSc DMDA layout, IceGrid.cc:183-195, 814):
 *   2-D fields        a[i*My + j]                 (j contiguous)
 *   staggered fields  a[(i*My + j)*2 + o]         o=0: (i+1/2,j), o=1: (i,j+1/2)
 *   3-D fields        a[(i*My + j)*Mz + k]        (column k contiguous)
 * Ghost cells are not stored: every stencil access wraps periodically, which
 * is what PISM's periodic DMDA ghosts give (IceGrid.cc:807-819).
 * All field pointers passed to or returned by the functions below are DEVICE
 * pointers (HBM) unless the comment says "host".
 *
 * Streams: one handle is bound to one HIP stream (sia_set_stream); all work
 * is stream-ordered. A handle is not thread-safe; separate handles are.
 */
#ifndef PISM_SIA_H
#define PISM_SIA_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the RuntimeError cases of the reference) ------------- */
enum sia_status {
  SIA_OK = 0,
  SIA_ERR_ARGUMENT = 1,          /* bad argument / size / NULL pointer */
  SIA_ERR_HIP = 2,               /* a HIP runtime call failed */
  SIA_ERR_NEGATIVE_THICKNESS = 3,/* PISMBedSmoother.cc:312-314, iMgeometry.cc:140-142 */
  SIA_ERR_THICKNESS_ABOVE_LZ = 4,/* IceGrid.cc:381-389 (kBelowHeight) */
  SIA_ERR_OMEGA_NEGATIVE = 5,    /* PISMBedSmoother.cc:395-398 */
  SIA_ERR_CONFIG = 6,            /* unknown flow law / gradient method (SIAFD.cc:203-206,
                                    flowlaw_factory.cc:100-104), grain size without age
                                    (SIAFD.cc:570-573), smoothing domain too large
                                    (PISMBedSmoother.cc:132-135) */
  SIA_ERR_OUT_OF_MEMORY = 7,
  SIA_ERR_STATE = 8              /* output requested before the update that makes it */
};

/* ---- configuration selections (pism_config.cdl keywords) ---------------- */
enum sia_flow_law_kind {         /* sia_flow_law, cdl:1091-1095; flowlaw_factory.cc:77-83 */
  SIA_FLOW_ISOTHERMAL_GLEN = 0,  /* "isothermal_glen" */
  SIA_FLOW_PB = 1,               /* "pb" */
  SIA_FLOW_GPBLD = 2,            /* "gpbld" (default) */
  SIA_FLOW_HOOKE = 3,            /* "hooke" */
  SIA_FLOW_ARR = 4,              /* "arr"  (Paterson-Budd cold) */
  SIA_FLOW_ARRWARM = 5,          /* "arrwarm" */
  SIA_FLOW_GK = 6                /* "gk"   (Goldsby-Kohlstedt) */
};

enum sia_gradient_method {       /* surface_gradient_method, cdl:1272; SIAFD.cc:187-207 */
  SIA_GRADIENT_HASELOFF = 0,     /* default */
  SIA_GRADIENT_MAHAFFY = 1,
  SIA_GRADIENT_ETA = 2
};

enum sia_enthalpy_converter_kind {
  SIA_EC_DEFAULT = 0,            /* EnthalpyConverter (enthalpyConverter.cc:56-320) */
  SIA_EC_COLD = 1,               /* ColdEnthalpyConverter, used by pismv (enthalpyConverter.cc:322-357) */
  SIA_EC_KIRCHHOFF = 2           /* KirchhoffEnthalpyConverter (enthalpyConverter.cc:422-435) */
};

/* ---- grid: IceGrid's Mx, My, Mz, dx, dy, z[] (IceGrid.hh:201-295) -------
   Mx, My >= 3, 2 Mx My < 2^31, 2 <= Mz <= 4096; sia_create returns
   SIA_ERR_ARGUMENT otherwise, before any device call. */
typedef struct sia_grid {
  int32_t Mx, My, Mz;
  int32_t reserved;
  double dx, dy;                 /* m */
  const double *z;               /* host pointer, Mz levels, z[0] = 0, strictly increasing,
                                    z[Mz-1] = Lz (IceGrid.cc:333-376) */
} sia_grid;

/* ---- parameters, named as in pism_config.cdl, already in SI units ------- */
typedef struct sia_params {
  int32_t sia_flow_law;                    /* enum sia_flow_law_kind */
  int32_t surface_gradient_method;         /* enum sia_gradient_method */
  int32_t enthalpy_converter;              /* enum sia_enthalpy_converter_kind */
  int32_t compute_grain_size_using_age;    /* cdl:309 */
  int32_t do_age;                          /* cdl (needed with the line above) */
  int32_t is_dry_simulation;               /* cdl:136 */
  int32_t include_bmr_in_continuity;       /* cdl:148 */
  int32_t reserved;
  double sia_enhancement_factor;           /* cdl:281 */
  double sia_Glen_exponent;                /* cdl:1009 */
  double bed_smoother_range;               /* cdl:560, m */
  double ice_grain_size;                   /* cdl:304, m */
  double ice_softness;                     /* cdl:1020, Pa-3 s-1 */
  double Paterson_Budd_A_cold;             /* cdl:1060 */
  double Paterson_Budd_A_warm;             /* cdl:1065 */
  double Paterson_Budd_Q_cold;             /* cdl:1070 */
  double Paterson_Budd_Q_warm;             /* cdl:1075 */
  double Paterson_Budd_critical_temperature;/* cdl:1080 */
  double gpbld_water_frac_coeff;           /* cdl:1115 */
  double gpbld_water_frac_observed_limit;  /* cdl:1120 */
  double Hooke_A, Hooke_Q, Hooke_C, Hooke_k, Hooke_Tr; /* cdl:1025-1045 */
  double ideal_gas_constant;               /* cdl:1193 */
  double beta_CC;                          /* cdl:963 */
  double ice_specific_heat_capacity;       /* cdl:1003 */
  double water_specific_heat_capacity;     /* cdl:988 */
  double water_latent_heat_fusion;         /* cdl:983 */
  double surface_pressure;                 /* cdl:968 */
  double ice_density;                      /* cdl:993 */
  double standard_gravity;                 /* cdl:1188 */
  double water_melting_point_temperature;  /* cdl:973 */
  double enthalpy_converter_reference_temperature; /* cdl:978 */
  double sea_water_density;                /* cdl:1140 */
  double sea_level;                        /* m; OceanModel::sea_level_elevation() */
  double mask_icefree_thickness_standard;  /* cdl:1225 */
  double mask_is_floating_thickness_standard; /* cdl:1230 */
  double adaptive_timestepping_ratio;      /* cdl:330 */
  double maximum_time_step;                /* cdl:341, converted to seconds */
  double seconds_per_year;                 /* UDUNITS "year" = 3.15569259747e7 s */
} sia_params;

/* Fills *p with the pism_config.cdl defaults (SI units). Host-only. */
void sia_params_default(sia_params *p);

/* ---- per-update inputs: the fields SIAFD reads from the Vars registry ---- */
typedef struct sia_inputs {      /* PISMVars names: SIAFD.cc:122, 220-222, 292, 370-377, 552-556, 605 */
  const double *land_ice_thickness;   /* H    [Mx][My] */
  const double *bedrock_altitude;     /* b    [Mx][My] */
  const double *surface_altitude;     /* h    [Mx][My] */
  const double *mask;                 /* mask [Mx][My], integer codes stored as double
                                         (Mask.hh:28-34; read as floor(x+0.5)) */
  const double *enthalpy;             /* E    [Mx][My][Mz], J kg-1 */
  const double *age;                  /* [Mx][My][Mz] s, or NULL (only with GK +
                                         compute_grain_size_using_age) */
  const double *vel_input_u;          /* sliding velocity u_b [Mx][My] or NULL (ZeroSliding) */
  const double *vel_input_v;          /* v_b, or NULL */
  const double *basal_melt_rate;      /* [Mx][My] m s-1 for w at the base, or NULL (w_0 = 0) */
  int64_t bed_state_counter;          /* IceModelVec::get_state_counter() of the bed; the
                                         smoother re-runs when it exceeds the stored one */
  const double *no_model_mask;        /* [Mx][My] or NULL; with usurfstore: SIAFD_Regional's */
  const double *usurfstore;           /* gradient override (regional.cc:45-109) */
} sia_inputs;

typedef struct sia_handle sia_handle;

/* SIAFD ctor + init: allocates the device work and output fields on the
   current HIP device. */
int sia_create(const sia_grid *grid, const sia_params *params, sia_handle **out);
int sia_destroy(sia_handle *h);

/* Bind the handle to a HIP stream (hipStream_t passed as void*; NULL = default). */
int sia_set_stream(sia_handle *h, void *hip_stream);

/* When check != 0 (the default), sia_update synchronises the stream and turns
   device-side error flags into a status, as ParallelSection::check does
   (error_handling.cc:153-192). When 0, sia_update never synchronises and the
   flags are reported by the next sia_check_errors / sia_max_diffusivity. */
int sia_set_sync_check(sia_handle *h, int check);
int sia_check_errors(sia_handle *h);

/* SIAFD::update(vel_input, fast) (SIAFD.cc:119-148). fast != 0 computes only
   the diffusive flux and D_max; fast == 0 also computes u, v (SIAFD.cc:899-963)
   and w (PISMStressBalance.cc:192-287). */
int sia_update(sia_handle *h, const sia_inputs *in, int fast);

/* Outputs owned by the handle; valid until the next sia_update. */
int sia_diffusive_flux(sia_handle *h, const double **flux);        /* [Mx][My][2] */
int sia_max_diffusivity(sia_handle *h, double *D_max);             /* host value, m2 s-1 */
int sia_max_diffusivity_device(sia_handle *h, const double **D_max);/* device scalar */
int sia_velocity_u(sia_handle *h, const double **u);               /* [Mx][My][Mz] m s-1 */
int sia_velocity_v(sia_handle *h, const double **v);
int sia_velocity_w(sia_handle *h, const double **w);
/* StressBalance::volumetric_strain_heating() (PISMStressBalance.cc:146-148,
   computed by compute_volumetric_strain_heating, 370-500, on every full
   update): [Mx][My][Mz] W m-3, zero above kBelowHeight(H). */
int sia_volumetric_strain_heating(sia_handle *h, const double **Sigma);

/* Diagnostics (SIAFD_diagnostics.cc:26-35). */
int sia_surface_gradient(sia_handle *h, const double **h_x, const double **h_y); /* [Mx][My][2] each */
int sia_diffusivity_staggered(sia_handle *h, const double **D);    /* [Mx][My][2] */
int sia_theta_thksmooth(sia_handle *h, const double **theta, const double **thksmooth);
int sia_bed_smoother_fields(sia_handle *h, const double **topgsmooth, const double **maxtl,
                            const double **C2, const double **C3, const double **C4,
                            int32_t *Nx, int32_t *Ny);

/* SIAFD diagnostic "diffusivity" (SIAFD_diagnostics.cc:108-127): the staggered
   D of the last update averaged to the regular grid (SIAFD.cc:736-744,
   iceModelVec2.cc:679-689), written to the caller's device array [Mx][My]. */
int sia_diffusivity(sia_handle *h, double *out2d);

/* IceModelVec3::getSurfaceValues (iceModelVec3.cc:131-155, 206-222): a 3-D
   device field [Mx][My][Mz] on the handle's levels interpolated linearly to
   z = thickness[i][j] (end levels beyond the column), as sia.py's
   computeSIASurfaceVelocities uses it (site-packages/PISM/sia.py:24-55). */
int sia_surface_values(sia_handle *h, const double *field3d, const double *thickness, double *out2d);

/* Maximum time step (s) after sia_update: min(maximum_time_step, time_to_end,
   diffusivity limit) as in IceModel::max_timestep with ZeroSliding
   (iMadaptive.cc:157-171, 214-388). time_to_end <= 0 means "no end". */
int sia_max_timestep(sia_handle *h, double time_to_end, double *dt);

/* Geometry state for mass continuity (device pointers, [Mx][My]). */
typedef struct sia_geometry {
  double *land_ice_thickness;   /* H, updated in place */
  double *surface_altitude;     /* h, recomputed */
  double *mask;                 /* recomputed */
  const double *bedrock_altitude;
  const double *climatic_mass_balance; /* kg m-2 s-1, or NULL (zero) */
  const double *basal_melt_rate;       /* m s-1, or NULL (zero) */
} sia_geometry;

/* Totals of the flux accounting (iMgeometry.cc:781-790), kg over one step. */
typedef struct sia_mass_accounting {
  double grounded_basal_ice_flux, nonneg_rule_flux, sub_shelf_ice_flux, surface_ice_flux,
         sum_divQ_SIA, sum_divQ_SSA, Href_to_H_flux, H_to_Href_flux;
} sia_mass_accounting;

/* IceModel::massContExplicitStep with the handle's current diffusive flux and
   the sliding velocity of the last update (iMgeometry.cc:497-822), followed by
   updateSurfaceElevationAndMask (iMgeometry.cc:56-149). dt in seconds.
   acc may be NULL (host pointer; forces a synchronisation when given). */
int sia_mass_continuity_step(sia_handle *h, const sia_geometry *g, double dt,
                             sia_mass_accounting *acc);

/* One SIA-only model step entirely on the device: full (fast == 0) or
   flux-only update, dt = min(maximum_time_step, time_to_end, diffusivity
   limit) computed on the device, mass continuity and surface/mask update.
   The dt taken is accumulated into *t_device (device scalar, seconds) so
   that steps can be captured in a HIP graph without host synchronisation.
   t_end <= 0 means no end. */
int sia_step(sia_handle *h, const sia_geometry *g, const double *enthalpy,
             int fast, double t_end, double *t_device, double *dt_device);

/* ---- SIA-type basal sliding: SIA_Sliding (SIA_Sliding.cc) --------------- */
enum sia_sliding_mode {
  SIA_SLIDING_PRESSURE_MELTING = 0, /* mu_sliding p_over where the base is within beta_CC rho g H of
                                       minimum_temperature_for_sliding (SIA_Sliding.cc:260-266) */
  SIA_SLIDING_TEST_E = 1,           /* sia_sliding_verification_mode (SIA_Sliding.cc:217-238) */
  SIA_SLIDING_EISMINT2_G = 2,       /* EISMINT_II_experiment "G" (SIA_Sliding.cc:240-258) */
  SIA_SLIDING_EISMINT2_H = 3        /* EISMINT_II_experiment "H" */
};

typedef struct sia_sliding_params {
  int32_t mode;                          /* enum sia_sliding_mode */
  int32_t reserved;
  double mu_sliding;                     /* cdl:549 */
  double minimum_temperature_for_sliding;/* cdl:477, K */
  const double *x, *y;                   /* host: IceGrid::x(), y() (Mx and My values, m);
                                            read for SIA_SLIDING_TEST_E only */
} sia_sliding_params;

/* SIA_Sliding::update (SIA_Sliding.cc:106-206), the ShallowStressBalance of
   pismv's test E and of EISMINT II G/H: its own surface gradient (the
   handle's surface_gradient_method), the sliding velocity U_b = -C grad h and
   the basal frictional heating, written to the caller's device arrays
   [Mx][My]. Pass u_b, v_b as vel_input_u/_v of the next sia_update, as
   StressBalance::update does (PISMStressBalance.cc:89-96). Reads H, b, h, mask
   and the basal enthalpy from in. */
int sia_sliding_update(sia_handle *h, const sia_inputs *in, const sia_sliding_params *sp, double *u_b,
                       double *v_b, double *basal_frictional_heating);

/* ---- profiling: the reference's Profiling events (Profiling.hh:29-39) --------
   Event names as in SIAFD.cc:129-146, PISMStressBalance.cc:103-109 and
   iceModel.cc:732-735. When on, each stage is bracketed by hipEvents on the
   handle's stream; sia_profiling_read synchronises, returns the accumulated
   milliseconds and launch counts per event, and clears them. Not for use
   while a stream is being captured into a graph. */
enum sia_event {
  SIA_EV_BED_SMOOTHER = 0,   /* "SIA bed smoother" */
  SIA_EV_GRADIENT = 1,       /* "SIA gradient" (+ theta, thksmooth) */
  SIA_EV_FLUX = 2,           /* "SIA flux" */
  SIA_EV_3D_HOR_VEL = 3,     /* "SIA 3D hor. vel." */
  SIA_EV_VERT_VEL = 4,       /* "SB vert. vel.": computed in the "SIA 3D hor. vel." kernel, not recorded apart */
  SIA_EV_MASS_TRANSPORT = 5, /* "mass transport" (+ time step, surface and mask) */
  SIA_EV_SG_FEATURES = 6,    /* "surrogate features": the learned surrogate's per-level pass (inside "SIA flux") */
  SIA_EV_SG_MLP = 7,         /* "surrogate MLP": its fused network (inside "SIA flux") */
  SIA_EV_STRAIN_HEAT = 8,    /* "SB strain heat" (PISMStressBalance.cc:102-104): computed in the
                                "SIA 3D hor. vel." kernel, not recorded apart */
  SIA_EV_COUNT = 9
};
int sia_profiling(sia_handle *h, int on);
int sia_profiling_read(sia_handle *h, double *ms /* [SIA_EV_COUNT] */, int64_t *count /* [SIA_EV_COUNT] */);
const char *sia_event_name(int event);

/* Human-readable message for the last failing call on this handle. */
const char *sia_last_error(const sia_handle *h);

/* ---- flux-only update that allocates nothing ---------------------------------
   No counterpart in the reference. For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): the fast (flux-only) SIAFD::update of sia_update(h, in, 1)
   without a handle. The bed smoother, gradient, theta/thksmooth and the column
   pass run on every call (no state is kept) with their intermediate fields in
   one scratch array that the library holds on each GPU it runs on, for the
   whole process (about 55 MB per GPU); the call makes no hipMalloc/hipFree.
   Results are bitwise those of sia_update(fast = 1) on a handle with the same
   grid and parameters (sia_diffusive_flux, sia_max_diffusivity_device).
     grid: at most SIA_NOALLOC_MAX_COLUMNS map columns (Mx * My) and 4096
           levels; larger grids return SIA_ERR_ARGUMENT. grid->z (host) is
           copied on the stream.
     in:   as for sia_update (device fields; bed_state_counter is ignored).
     flux: device [Mx][My][2], written.  D_max: device scalar, written.
     err_flags: device int or NULL: the update's error flags (bit 1: negative
           thickness, 2: thickness above Lz, 4: negative omega; the
           SIA_ERR_NEGATIVE_THICKNESS, _THICKNESS_ABOVE_LZ, _OMEGA_NEGATIVE
           cases), read by the caller after the stream completes.
     hip_stream: hipStream_t or NULL. The call does not synchronise.
   Every call on a GPU uses that GPU's one scratch array: calls that overlap
   on one GPU (other streams or threads) race and must be serialised by the
   caller. */
#define SIA_NOALLOC_MAX_COLUMNS 524288
int sia_flux_update_noalloc(const sia_grid *grid, const sia_params *params, const sia_inputs *in, double *flux,
                            double *D_max, int32_t *err_flags, void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* PISM_SIA_H */
