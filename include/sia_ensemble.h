/* sia_ensemble.h -- batched SIA ensembles on one MI355X (C ABI).
 *
 * No counterpart in the reference: PISM v0.7.3 runs one model per process
 * (IceModel::step, iceModel.cc:651, drives one StressBalance). An ensemble
 * groups existing sia_handle members (any grid sizes, per-member enhancement
 * factor and Glen exponent, one flow law) and advances all of them with the
 * step of sia_step(fast = 1) -- SIA flux update, device-side dt
 * (iMadaptive.cc:157-171, 214-388), mass continuity and the surface/mask
 * update (iMgeometry.cc:56-149, 497-822) -- issuing ONE kernel launch per stage
 * for all members. Every member's results are bitwise those of calling
 * sia_step on it alone. Multi-step runs can be captured once into a HIP graph
 * and replayed; all state stays on the device.
 *
 * Threading and streams: an ensemble runs on a stream of its own and is
 * ordered after/before the stream given by sia_ensemble_set_stream (NULL: the
 * default stream) with events, so callers see ordinary stream semantics. The
 * member handles must not be used by other calls while an ensemble step is in
 * flight.
 */
#ifndef SIA_ENSEMBLE_H
#define SIA_ENSEMBLE_H

#include "pism_sia.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sia_ensemble sia_ensemble;

/* members[i]: a handle on the current device whose bed smoother is ready (one
   sia_update or sia_step ran on it); all members share the flow law, and the
   grain-size/age coupling is off. geometry[i], enthalpy[i]: member i's device
   fields ([Mx][My] and [Mx][My][Mz]; H, h and mask are updated in place by the
   steps). t_end: host array of n end times in seconds (<= 0: no end), or NULL.
   Member times start at 0.
   A member bound to a learned surrogate (sia_use_surrogate, include/
   sia_surrogate.h) takes its fast updates through it, as sia_step would: the
   surrogate stages run with one launch per stage for all members bound to the
   same surrogate, and the FP64 column stage takes only their fallback columns.
   The binding is read here; sia_ensemble_step returns SIA_ERR_STATE once a
   member's binding differs from it. A surrogate retrained while bound is
   picked up by the next sia_ensemble_step. */
int sia_ensemble_create(sia_handle *const *members, const sia_geometry *geometry, const double *const *enthalpy,
                        const double *t_end, int n, sia_ensemble **out);
int sia_ensemble_destroy(sia_ensemble *e);
int sia_ensemble_set_stream(sia_ensemble *e, void *hip_stream);

/* Columns per wavefront of the column stage: 8, 16 (default) or 32. */
int sia_ensemble_set_tile(sia_ensemble *e, int cols_per_wave);

/* Autotuner: times `steps` ensemble steps with each tile width (HIP events on
   the ensemble's stream), keeps the fastest and restores every member's H, h,
   mask, time and dt to what they were before the call. ms: host [3] (tile
   widths 8, 16, 32) or NULL; chosen: host or NULL. */
int sia_ensemble_autotune(sia_ensemble *e, int steps, int *chosen, double *ms);

/* n_steps steps of every member. use_graph != 0: the n_steps steps are
   captured once into a HIP graph (again when n_steps or the tile width
   changes) and the graph is launched. Never synchronises the host. */
int sia_ensemble_step(sia_ensemble *e, int n_steps, int use_graph);

/* Device scalars of member i: model time (s) and the last dt (s). */
int sia_ensemble_time(sia_ensemble *e, int member, const double **t_device, const double **dt_device);

/* Synchronises and returns the status of the first member whose device error
   flags were raised by any step since the ensemble was created. */
int sia_ensemble_check_errors(sia_ensemble *e);

/* Human-readable message for the last failing call on this ensemble. */
const char *sia_ensemble_last_error(const sia_ensemble *e);

#ifdef __cplusplus
}
#endif

#endif /* SIA_ENSEMBLE_H */
