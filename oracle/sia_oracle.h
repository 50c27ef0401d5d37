/*
 * sia_oracle.h -- CPU oracle for the SIA hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C, loop-for-loop restatement of PISM v0.7.3's SIA stress balance,
 * vertical velocity and explicit mass-continuity step, used only as the
 * checker of the HIP implementation (tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py). The product never links or calls it.
 *
 * Ghosts: PISM's DMDAs are periodic (IceGrid.cc:807-819) and every ghost
 * value SIAFD reads is either a periodic copy or recomputed from periodic
 * copies with the same formula, so the oracle stores owned points only and
 * wraps every stencil index periodically.
 *
 * Parity pinning per function (see tests/test_oracle_*.py):
 *   or_flow (all 7 laws) + EC: golden tables test_20/21/24/25/26 and the
 *       invariants of test/enthalpy_converter.py; isothermal closed form.
 *   or_bed_smoother / or_theta: golden output of test_22 (bedrough_test).
 *   or_kbelow: GSL bsearch semantics, brute force on every level.
 *   or_surface_gradient: planar surfaces (exact), haseloff == mahaffy away
 *       from margins, eta exact for linear eta, brute-force margin cases.
 *   or_diffusive_flux / or_compute_I: isothermal closed form D and I to
 *       O(dz^2), exact discrete trapezoid in long double on tiny inputs.
 *   or_horizontal_velocity / or_vertical_velocity: closed-form SIA velocity
 *       field of a planar slab, incompressibility, test-F-style invariants.
 *   or_mass_continuity: mass conservation to 14 digits (test_12 invariant),
 *       Vialov steady state, exact flux divergence on tiny grids.
 */
#ifndef SIA_ORACLE_H
#define SIA_ORACLE_H

#include "../include/pism_sia.h"

#ifdef __cplusplus
extern "C" {
#endif

/* enthalpy converter (enthalpyConverter.cc:56-357) */
double or_ec_pressure(const sia_params *P, double depth);
double or_ec_melting_temperature(const sia_params *P, double p);
double or_ec_enthalpy_cts(const sia_params *P, double p);
double or_ec_temperature(const sia_params *P, double E, double p);
double or_ec_pressure_adjusted_temperature(const sia_params *P, double E, double p);
double or_ec_water_fraction(const sia_params *P, double E, double p);
double or_ec_enthalpy(const sia_params *P, double T, double omega, double p);
int    or_ec_is_temperate(const sia_params *P, double E, double p);

/* flow laws (flowlaws.cc:33-495, flowlaws.hh:165-404) */
double or_flow(const sia_params *P, double stress, double E, double p, double gs);
int    or_flow_uses_grain_size(const sia_params *P);
double or_grain_size_vostok(const sia_params *P, double age_seconds);

/* IceGrid::kBelowHeight (IceGrid.cc:379-391); returns -1 on error */
int or_kbelow(const double *z, int Mz, double height);

/* BedSmoother (PISMBedSmoother.cc:96-420) */
int or_bed_smoother(const sia_grid *g, const sia_params *P, const double *topg,
                    double *topgsmooth, double *maxtl, double *C2, double *C3,
                    double *C4, int *Nx, int *Ny);
int or_theta(const sia_grid *g, const sia_params *P, int Nx, int Ny, const double *usurf,
             const double *topgsmooth, const double *maxtl, const double *C2,
             const double *C3, const double *C4, double *theta);
int or_thksmooth(const sia_grid *g, const double *usurf, const double *thk,
                 const double *mask, const double *maxtl, const double *topgsmooth,
                 double *result);

/* SIAFD (SIAFD.cc:187-963) */
int or_surface_gradient(const sia_grid *g, const sia_params *P, const double *H,
                        const double *b, const double *h, const double *mask,
                        double *h_x, double *h_y);
int or_diffusive_flux(const sia_grid *g, const sia_params *P, const double *h_x,
                      const double *h_y, const double *thksmooth, const double *theta,
                      const double *E, const double *age, int full_update,
                      double *flux, double *D_stag, double *delta0, double *delta1,
                      double *D_max);
int or_compute_I(const sia_grid *g, const double *thksmooth, const double *delta0,
                 const double *delta1, double *I0, double *I1);
int or_horizontal_velocity(const sia_grid *g, const double *h_x, const double *h_y,
                           const double *I0, const double *I1, const double *ub,
                           const double *vb, double *u, double *v);

/* StressBalance::compute_vertical_velocity (PISMStressBalance.cc:192-287) */
int or_vertical_velocity(const sia_grid *g, const double *mask, const double *u,
                         const double *v, const double *bmr, double *w);

/* All fields of one SIA update; NULL members are skipped where optional. */
typedef struct or_fields {
  /* inputs */
  const double *H, *b, *h, *mask, *E, *age, *ub, *vb, *bmr;
  /* outputs / work ([Mx*My] for 2-D, [Mx*My*2] staggered, [Mx*My*Mz] 3-D) */
  double *topgsmooth, *maxtl, *C2, *C3, *C4;
  double *theta, *thksmooth, *h_x, *h_y;
  double *flux, *D_stag;
  double *delta0, *delta1, *I0, *I1;   /* required when full */
  double *u, *v, *w;                   /* required when full */
  double D_max;
  int Nx, Ny;
  double *Sigma;                       /* [Mx*My*Mz] strain heating when full, or NULL */
  const double *no_model_mask, *usurfstore; /* SIAFD_Regional inputs, or NULL */
} or_fields;

/* SIAFD::update + (if full) StressBalance::compute_vertical_velocity. The bed
   smoother always runs (the oracle keeps no state). */
int or_sia_update(const sia_grid *g, const sia_params *P, or_fields *f, int full);

/* SIA_Sliding (SIA_Sliding.cc): the sliding law's mode and parameters */
enum { OR_SLIDING_PRESSURE_MELTING = 0, OR_SLIDING_TEST_E = 1, OR_SLIDING_EISMINT2_G = 2,
       OR_SLIDING_EISMINT2_H = 3 };
typedef struct or_sliding {
  int mode;
  double mu_sliding;                       /* cdl:549 */
  double minimum_temperature_for_sliding;  /* cdl:477 */
} or_sliding;

/* SIA_Sliding::update (SIA_Sliding.cc:106-206): its own surface gradient
   (h_x, h_y staggered work arrays), the sliding velocity and the basal
   frictional heating; x[Mx], y[My] are the grid coordinates */
int or_sia_sliding(const sia_grid *g, const sia_params *P, const or_sliding *sp, const double *x, const double *y,
                   const double *H, const double *h, const double *b, const double *mask, const double *E,
                   double *h_x, double *h_y, double *ub, double *vb, double *heat);

double or_ec_enthalpy_permissive(const sia_params *P, double T, double omega, double p);

/* SIAFD_Regional::compute_surface_gradient's override (regional.cc:45-109) */
int or_regional_gradient(const sia_grid *g, const double *nmm, const double *hst, double *h_x,
                         double *h_y);

/* FlowLaw::hardness_parameter of the selected law (flowlaws.cc:103-105, 336-351) */
double or_hardness(const sia_params *P, double E, double p);

/* StressBalance::compute_volumetric_strain_heating (PISMStressBalance.cc:370-500) */
int or_strain_heating(const sia_grid *g, const sia_params *P, const double *H, const double *mask,
                      const double *E, const double *u, const double *v, double *Sigma);

/* GeometryCalculator::compute over the grid (Mask.hh:67-118, iMgeometry.cc:80-149) */
int or_update_mask_surface(const sia_grid *g, const sia_params *P, const double *bed,
                           const double *H, double *mask, double *h);

/* IceModel::massContExplicitStep (iMgeometry.cc:497-822) without part_grid and
   Dirichlet BCs; acc[8] in the order of iMgeometry.cc:781-788. */
int or_mass_continuity(const sia_grid *g, const sia_params *P, double dt,
                       const double *H, const double *flux, const double *ub,
                       const double *vb, const double *mask, const double *smb,
                       const double *bmr, double *H_new, double *acc);

/* IceModel::max_timestep restricted to the SIA-only path (iMadaptive.cc:118-171, 214-388) */
double or_max_timestep(const sia_grid *g, const sia_params *P, double D_max,
                       double time_to_end, const double *ub, const double *vb,
                       const double *mask);

const char *or_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
