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
 *     features to a starting pole and low-order shape;
 *   - ps_scan_learned: the scan that screens every trial period from the
 *     network's start and has the plain ten-pole fits judge the best.
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

/* Amplitude-aspect features of one asteroid, the same at every trial
 * period: from the dense light curves (at least 10 points within one day),
 * amplitude A = (max - min) / mean brightness and the longitude lv of the
 * mean Earth direction; out = [c, s, hypot(c, s), mean A, count / 10] with
 * (c, s) = sum (A - mean A) (cos 2 lv, sin 2 lv) / (count * mean A) -- the
 * second harmonic of amplitude against viewing longitude, which points at
 * the pole longitude modulo 180 deg.  Zeros with fewer than two such light
 * curves. */
int ps_aspect_features(const ps_problem *p, float out[5]);

/* A trained proposal network: n_in = nbins + 2 + 5 inputs (ps_lc_features
 * then ps_aspect_features), scaled as (x - x_mean) / x_std; n_out = 12
 * outputs, unscaled as y * t_std + t_mean: the doubled-longitude pole
 * (cos b cos 2l, cos b sin 2l, sin b) and the 9 coefficients with l <= 2
 * (c00, c10, c20, c11 cos/sin, c21 cos/sin, c22 cos/sin; c00 is not used). */
typedef struct ps_proposal_net {
  ps_mlp *mlp;
  int n_in, n_out, nbins, max_batch;  /* max_batch: the value given to ps_mlp_create */
  const float *x_mean, *x_std, *t_mean, *t_std;
} ps_proposal_net;

/* Load a trained network from an uncompressed .npz (numpy.savez) holding
 * float32 1-D arrays params, x_mean, x_std, t_mean and t_std: n_in =
 * len(x_mean), n_out = len(t_mean), nbins = n_in - 7, the hidden width from
 * len(params) = (n_in + 1) H + (H + 1) H + (H + 1) n_out.  Creates the
 * network on `device` for batches up to max_batch and fills *net, which owns
 * the network and the scalings until ps_proposal_close.  PS_ERR_IO: the file
 * is missing, compressed or not in that form. */
int ps_proposal_open(const char *path, int device, int max_batch, ps_proposal_net *net);
void ps_proposal_close(ps_proposal_net *net);

/* How the learned scan screens and judges. */
typedef struct ps_learned_opts {
  int block;         /* trial periods per screening batch */
  int n_candidates;  /* lowest-rms periods per block sent to the judge (no two within halo) */
  int halo;          /* neighbours on either side of each candidate also judged */
  int mirror;        /* 1: fit the mirror pole (lambda + 180 deg) as well */
  int refine;        /* rounds that judge the unjudged neighbours of the best judged row, until it is a local
                      * minimum of the judged rows (0: none) */
  ps_engine *judge;  /* engine for the plain ten-pole judge fits (its precision, e.g. FP64, and the same
                      * problem); NULL: the screening engine.  Another engine runs a block's judge fits on
                      * its own thread while the next block is screened. */
} ps_learned_opts;

/* Rows n_from..n_to of a learned scan (no counterpart in the reference, whose
 * every trial period is fitted from ten poles, period_search_BOINC.cpp:807-924).
 * Each trial period is fitted by the full LM loop (the problem's stop rule)
 * from the network's proposed start (the reference's start with the proposed
 * pole and l <= 2 shape); per block the candidates and their neighbours are
 * refitted from the reference's ten poles by the judge, and those rows
 * replace the screened ones: judged[i] = 1 marks them, the rows in the
 * reference's sense; then the best judged row's unjudged neighbours are
 * judged until it is a local minimum of the judged rows (opt->refine rounds
 * at most).  The other rows are the screening fits' (period of the best
 * screening fit, its rms and chi2).  p is the problem given to
 * ps_set_problem of e (and of the judge).  screened (optional, [n]) receives
 * every period's screening row; *n_fits the LM fits run (screening + judge). */
int ps_scan_learned(ps_engine *e, const ps_problem *p, const ps_proposal_net *net, const ps_learned_opts *opt,
                    double freq_start, double freq_step, int n_from, int n_to, double conw_r, ps_row *rows,
                    int *judged, ps_row *screened, long long *n_fits);

#ifdef __cplusplus
}
#endif

#endif
