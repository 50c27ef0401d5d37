/*
 * sia_surrogate.h -- C ABI of the learned column surrogate of the SIA
 * diffusivity (MI355X, gfx950). No counterpart in the reference: it stands in
 * for the per-level flow-law integral of SIAFD::compute_diffusive_flux
 * (src/base/stressbalance/sia/SIAFD.cc:620-721) on fast (flux-only) updates,
 * for the GPBLD flow law (src/base/rheology/flowlaws.cc:209-221) with the
 * default enthalpy converter (src/base/enthalpyConverter.cc:56-320).
 *
 * Model. D of a staggered column is the reference's discrete sum
 * D = C sum_k G_k A(E_k, p_k) (C = 2 e theta |grad h|^(n-1), G_k = trapezoid
 * weight times p_k^n). Temperate levels are summed exactly (A is linear in
 * the water fraction there). Cold levels are reduced to 32 cells in
 * zeta = z / thk; a second-order physics baseline gives each cell's sum, and
 * an MLP (34 -> 4 x 128 GELU -> 32; BF16 weights and activations on MFMA,
 * FP32 accumulate) predicts a log-correction per cell. Inputs: the 32 cell
 * mean temperatures (the resampled enthalpy profile), thickness, |grad h|.
 *
 * Columns whose inputs fall outside the training range, and columns thinner
 * than the second level, fall back to the FP64 kernel (per column). Training
 * samples come from the FP64 path; training runs on the GPU with this
 * library's own forward, backward and Adam kernels.
 *
 * All pointers are DEVICE pointers unless the comment says "host".
 */
#ifndef SIA_SURROGATE_H
#define SIA_SURROGATE_H

#include <stdint.h>
#include <stddef.h>

#include "pism_sia.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SIA_SURROGATE_NPARAM 58144 /* FP32 parameters: W1[128][34] b1 W2..W4[128][128] b2..b4 W5[32][128] b5 */
#define SIA_SURROGATE_NIN 34
#define SIA_SURROGATE_KIN 64       /* padded input width of the gradient hook */
#define SIA_SURROGATE_NOUT 32

typedef struct sia_surrogate sia_surrogate;

/* A model with Xavier-uniform weights (seeded) and room for `capacity`
   training samples on the current device. */
int sia_surrogate_create(uint64_t seed, int64_t capacity, sia_surrogate **out);
/* SIA_ERR_STATE (and nothing freed) while a handle is still bound to the model:
   unbind with sia_use_surrogate(h, NULL) or sia_destroy(h) first. */
int sia_surrogate_destroy(sia_surrogate *sg);
int sia_surrogate_set_stream(sia_surrogate *sg, void *hip_stream);

/* Runs the FP64 path of handle h on `in` (a fast update) and appends every
   stride-th eligible staggered column as a training sample: inputs, per-cell
   target log-corrections and the cell shares of D. *n_total (host) receives
   the number of samples held. */
int sia_surrogate_collect(sia_surrogate *sg, sia_handle *h, const sia_inputs *in, int stride, int64_t *n_total);
/* Input normalization, training range and output scale from the samples. */
int sia_surrogate_fit_normalization(sia_surrogate *sg);
/* `steps` Adam steps on random minibatches of `batch` samples (a multiple of
   32), learning rate cosine-annealed from lr_start to lr_end. loss_out (host,
   [steps] or NULL) receives the loss of every step. */
int sia_surrogate_train(sia_surrogate *sg, int64_t steps, int batch, double lr_start, double lr_end, uint64_t seed,
                        double *loss_out);

/* FP32 parameters (host arrays of SIA_SURROGATE_NPARAM). */
int sia_surrogate_get_params(const sia_surrogate *sg, float *host, size_t n);
int sia_surrogate_set_params(sia_surrogate *sg, const float *host, size_t n);
/* Loss and gradient of one batch (test hook): x [B][64] normalized inputs,
   r, share, has [B][32] (targets in output units); grad_host [NPARAM]. */
int sia_surrogate_grad(sia_surrogate *sg, int batch, const float *x, const float *r, const float *share,
                       const float *has, double *loss_host, float *grad_host);
/* Whole model (normalization + weights) as a host blob; size query with blob == NULL. */
int sia_surrogate_save(const sia_surrogate *sg, void *blob, size_t *size);
int sia_surrogate_load(sia_surrogate *sg, const void *blob, size_t size);
/* Training samples (device array of 416-byte records) and their count. */
int sia_surrogate_samples(const sia_surrogate *sg, const void **samples, int64_t *n);

/* Fast updates of h use the surrogate from now on (NULL: FP64 again). The
   model must outlive the binding: sia_surrogate_destroy refuses a bound model.
   SIA_ERR_CONFIG if the flow law, enthalpy converter or Glen exponent of h
   differ from the model's, or the model has no normalization yet. */
int sia_use_surrogate(sia_handle *h, sia_surrogate *sg);
/* Staggered columns of the last surrogate update (host): [0] D = 0,
   [1] surrogate, [2] FP64 fallback. */
int sia_surrogate_counts(sia_handle *h, int64_t counts[3]);

/* Score fallback. Besides columns outside the training range, a column whose
   score sum_m t_m |r_m| / D exceeds tau is computed by the FP64 kernel
   (t_m = C Qb_m exp(r_m): cell m's predicted share of D; r_m: the network's
   log-correction). The score is the predicted-share-weighted size of the
   correction; the columns it singles out carry the largest surrogate errors.
   tau is +inf (no score fallback) until set, and is saved with the model.
   calibrate_fallback runs one surrogate fast update of `in` on h (fields the
   model was trained on, not those it will be evaluated on) with every in-range
   column kept, and sets tau to the (1 - fraction) quantile of their scores:
   about `fraction` of such columns then fall back. tau (host) or NULL. */
int sia_surrogate_calibrate_fallback(sia_surrogate *sg, sia_handle *h, const sia_inputs *in, double fraction,
                                     double *tau);
/* tau >= 0, or INFINITY for no score fallback. */
int sia_surrogate_set_fallback_score(sia_surrogate *sg, double tau);

#ifdef __cplusplus
}
#endif

#endif /* SIA_SURROGATE_H */
