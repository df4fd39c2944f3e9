/* Learned starting points for the trial-period fits.
 *
 * None of these entry points has a counterpart in the reference, which starts
 * every trial period from the same ten poles and the ellipsoid shape
 * (period_search/period_search_BOINC.cpp:819-844).  They add:
 *   - ps_lc_features: phase-folded light-curve features of the engine's
 *     problem at given trial frequencies;
 *   - ps_fit_batch_from: LM fits from caller-given starting parameters (the
 *     same loop as ps_fit_batch, which takes the reference's starts);
 *   - ps_mlp_*: a small fully connected FP32 network (n_in -> H -> H -> n_out,
 *     ReLU) with GPU forward, backward and Adam kernels, which maps the
 *     features to a starting pole and low-order shape.
 * Return codes are the PS_* codes of periodsearch.h. */
#ifndef PS_PROPOSAL_H
#define PS_PROPOSAL_H

#include "periodsearch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Feature vector at trial frequency freq[i] (1/h), written to
 * out[i * (nbins + 2) ...]: the nbins-bin phase-folded light curve (phase
 * frac(24 * freq * t), each point's brightness over its light-curve mean
 * minus one, averaged per bin; empty bins 0), then the mean solar phase
 * angle (rad) and the mean |ecliptic latitude| of the Earth direction (rad).
 * 1 <= nbins <= 256. */
int ps_lc_features(ps_engine *e, const double *freq, int n, int nbins, float *out);

/* n fits from starting parameters cg[i * ma .. i * ma + ma - 1] (ma from
 * ps_dims; order as a[] of mrqmin: shape coefficients, 90 deg - beta [rad],
 * lambda [rad], omega [rad/day], phase parameters, log cl, cls), each run
 * through the LM loop of ps_fit_batch under the problem's stop rule.  If
 * cg_out is not null it receives the final parameters, same layout. */
int ps_fit_batch_from(ps_engine *e, const double *cg, int n, double conw_r, ps_fit_out *out, double *cg_out);

typedef struct ps_mlp ps_mlp;

int ps_mlp_create(int device, int n_in, int n_hidden, int n_out, int max_batch, ps_mlp **out);
int ps_mlp_destroy(ps_mlp *m);
const char *ps_mlp_last_error(const ps_mlp *m);
/* Number of FP32 parameters: W1 [n_in][H], b1 [H], W2 [H][H], b2 [H],
 * W3 [H][n_out], b3 [n_out], flat in that order. */
long long ps_mlp_param_count(const ps_mlp *m);
/* Set or read all parameters (setting also clears the Adam state). */
int ps_mlp_set_params(ps_mlp *m, const float *params, long long n);
int ps_mlp_get_params(ps_mlp *m, float *params, long long n);
/* y[batch][n_out] = network(x[batch][n_in]). */
int ps_mlp_forward(ps_mlp *m, const float *x, int batch, float *y);
/* Gradient of the loss (1/batch) sum_b sum_k w[k] (y[b][k] - t[b][k])^2
 * with respect to every parameter (same layout), and the loss. */
int ps_mlp_gradients(ps_mlp *m, const float *x, const float *t, const float *w_out, int batch, float *grad,
                     float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss; *loss is
 * the loss before the step. */
int ps_mlp_train_step(ps_mlp *m, const float *x, const float *t, const float *w_out, int batch, float lr,
                      float *loss);

#ifdef __cplusplus
}
#endif

#endif
