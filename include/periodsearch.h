/*
 * periodsearch.h -- C ABI of the MI355X period-search engine (libpsamd.so).
 *
 * The engine replaces the trial-period x initial-pole Levenberg-Marquardt fit
 * loop of the Asteroids@home period_search app.  Its only backend seam in the
 * reference is the C++ API of the CUDA app,
 *   period_search_cuda/period_search/start_CUDA.h:2-14
 *   (CUDAPrepare / CUDAPrecalc / CUDAStart / CUDAUnprepare),
 * which reads most of its inputs from extern globals
 * (period_search_cuda/period_search/globals.h:6-20).  This header exposes the
 * same roles as plain C functions over caller-owned buffers, with no globals:
 *
 *   ps_create + ps_set_problem   <- CUDAPrepare (start_CUDA.cu:119) and the
 *                                   static uploads of CUDAStart (start_CUDA.cu:558-647)
 *   ps_precalc                   <- CUDAPrecalc (start_CUDA.cu:248), one conw round
 *   ps_scan                      <- CUDAStart (start_CUDA.cu:523) without the text
 *                                   sink: it returns FP64 rows, the host writer
 *                                   (ps_format_row, ps_write_checkpoint) owns the text
 *   ps_destroy                   <- CUDAUnprepare (start_CUDA.cu:240)
 *   ps_fit_batch, ps_eval_mrqcof -- test seams with no reference counterpart
 *                                   (per-fit results, one normal-equation build)
 *
 * Host-side helpers replace the driver code that surrounds those calls in
 * period_search/period_search_BOINC.cpp (input parsing :277-485, shape
 * setup :737-782, row text :926-931, checkpoint :85-101).
 *
 * Conventions: arrays are 0-indexed and dense; the engine copies what it
 * needs, the caller keeps ownership.  Return value 0 = OK, negative = error
 * (ps_last_error gives the text).  One engine handle per device; handles are
 * independent, calls on one handle must not overlap.
 */
#ifndef PERIODSEARCH_H
#define PERIODSEARCH_H

#ifdef __cplusplus
extern "C" {
#endif

#define PS_OK 0
#define PS_ERR_ARG -1
#define PS_ERR_HIP -2
#define PS_ERR_LIMIT -3
#define PS_ERR_IO -4
#define PS_ERR_STATE -5

/* Arithmetic of the model/Jacobian stage.  PS_PREC_FP64 is the plain form
 * (everything FP64).  The mixed forms keep the per-point geometry (phase
 * reduction, rotation, phase function), the model values, residuals, chi2,
 * beta, the normal-equation solve and the LM acceptance in FP64; the facet
 * stage runs in FP32, the brightness/derivative contraction and the alpha
 * accumulation on the matrix cores with FP32 accumulation, with FP32
 * operands (PS_PREC_FP32), bf16 operands for the contraction
 * (PS_PREC_BF16; alpha stays FP32) or the contraction operands split into
 * bf16 high and low parts, three bf16 products (PS_PREC_BF16X3), or into
 * three bf16 parts, the six products down to third order (PS_PREC_BF16X6,
 * about 24 significant bits like FP32).  PS_PREC_BF16A is PS_PREC_BF16 with
 * alpha's products also from bf16-rounded rows (beta and chi2 as in
 * PS_PREC_BF16): a screening form, e.g. for ps_scan_learned, whose rows are
 * not results; alpha differs from PS_PREC_BF16's by at most 2^-8 of its
 * largest entry. */
#define PS_PREC_FP64 0
#define PS_PREC_FP32 1
#define PS_PREC_BF16 2
#define PS_PREC_BF16X3 3
#define PS_PREC_BF16X6 4
#define PS_PREC_BF16A 5

typedef struct ps_engine ps_engine;

/* One asteroid: points, shape basis and fit constants. */
typedef struct ps_problem {
  /* observations (the engine appends the 3 convexity rows itself) */
  int n_points;              /* real points, = ndata before the conv rows */
  int n_lc;                  /* real light curves */
  const int *lc_points;      /* [n_lc] points per light curve */
  const int *lc_inrel;       /* [n_lc] 1 = relative light curve (Inrel = 1 - flag) */
  const double *tim;         /* [n_points] JD - jd0 [days] */
  const double *brightness;  /* [n_points] */
  const double *sig;         /* [n_points] per-light-curve mean brightness */
  const double *weight;      /* [n_points] per-point weight */
  const double *ee;          /* [n_points*3] unit vector asteroid -> Earth */
  const double *ee0;         /* [n_points*3] unit vector asteroid -> Sun */
  /* shape basis on the Gaussian sphere (period_search/globals.h:5-19) */
  int lmax, mmax;
  int n_fac;                 /* F = 8*nrows^2 */
  const double *nor;         /* [F*3] facet normals */
  const double *darea;       /* [F] facet areas on the unit sphere */
  const double *pleg;        /* [F*(lmax+1)*(lmax+1)] P_l^m, index (f*(lmax+1)+l)*(lmax+1)+m */
  const double *fc;          /* [F*(mmax+1)] cos(m*phi) */
  const double *fs;          /* [F*(mmax+1)] sin(m*phi) */
  const double *dsph;        /* [F*ncoef] */
  const double *cg_first;    /* [ncoef] initial ellipsoid shape */
  /* fit constants (period_search_in header) */
  double phi0;               /* initial rotation angle [rad] */
  double par[3];             /* phase-function a, d, k */
  int ia_par[3];
  double cl;                 /* Lambert weight (fitted as log cl) */
  int ia_cl;
  int ia_prd;                /* period free */
  int onlyrel;               /* 1 if every light curve is relative (fixes shape coef 1) */
  double alamda_start, alamda_incr;
  double stop_condition;     /* >1: that many iterations; <1: dev threshold */
  int n_iter_min;
  int n_poles;               /* initial poles, normally 10 */
  const double *beta_pole;   /* [n_poles] deg */
  const double *lambda_pole; /* [n_poles] deg */
} ps_problem;

/* Result of one fit (one trial period from one initial pole). */
typedef struct ps_fit_out {
  double dev;      /* rms: sqrt(rchisq/(ndata-3)) */
  double chisq;    /* final Chisq */
  double rchisq;   /* Chisq minus the convexity penalty */
  double prd;      /* fitted period [days] */
  double la;       /* lambda [deg] (not wrapped) */
  double be;       /* beta [deg] */
  double dark;     /* |sum Area*Nor| */
  double totarea;  /* sum of Area of the last shape evaluation */
  double alamda;   /* final Marquardt lambda */
  int niter;       /* LM iterations run */
  int err;         /* last gauss_errc code (0 = none) */
  double latency_us; /* device wall clock from the fit taking its slot to its summary [us] */
} ps_fit_out;

/* One output row: the six numbers of period_search_out. */
typedef struct ps_row {
  double period_h;  /* 24 * per_best */
  double rms;       /* dev_best */
  double chisq;     /* dev_best^2 * (ndata-3) */
  double dark;      /* dark-facet % (row n=1 prints conw_r*escl^2 instead) */
  double lambda;    /* deg, wrapped to [0,360) */
  double beta;      /* deg */
} ps_row;

/* ---- engine ---------------------------------------------------------- */
int ps_create(int device, ps_engine **out);
int ps_destroy(ps_engine *e);
const char *ps_last_error(const ps_engine *e);
/* Select the arithmetic of the model stage (PS_PREC_*). Default FP64. */
int ps_set_precision(ps_engine *e, int precision);

int ps_set_problem(ps_engine *e, const ps_problem *p);

/* Several asteroids at once (no counterpart in the reference, which scans
 * one asteroid per process): the problems may differ in points, light
 * curves, lmax/mmax, facets and stop rule; their fits share the engine's
 * slots and work queue (ps_fit_batch_multi).  With more than one problem set,
 * the one-problem calls (ps_precalc, ps_scan, ps_fit_batch, ps_eval_mrqcof,
 * ps_lc_features, ps_fit_batch_from) return PS_ERR_STATE. */
int ps_set_problems(ps_engine *e, const ps_problem *p, int n);

/* n fits over the problems of ps_set_problems: item i fits problem prob[i]
 * at trial frequency freq[i] (1/h) from that problem's pole pole[i];
 * conw_r[k] is problem k's convexity weight.  Items run in any order through
 * the slots; out[i] is item i's result. */
int ps_fit_batch_multi(ps_engine *e, const int *prob, const double *freq, const int *pole, int n,
                       const double *conw_r, ps_fit_out *out);

/* Upper bound on concurrently resident fits (slots).  Default 2048.  Each
 * slot holds one fit's LM state and Jacobian in HBM. */
int ps_set_slots(ps_engine *e, int max_slots);

/* Slot groups, 1..4 (default 1): the slots are split into groups, each
 * driven on its own HIP stream, so the per-slot kernels of one group overlap
 * the model kernels of the others.  Results do not depend on it. */
int ps_set_groups(ps_engine *e, int groups);

/* One convexity-weight round (CUDAPrecalc): fits min(10, nPeriods) periods
 * x n_poles at conw_r, averages the best dark-facet %, sets *new_conw = 1 if
 * the average is < 1, else doubles *conw_r.  Frequencies are 1/P with P
 * in hours, as in the reference driver (period_search_BOINC.cpp:784-786). */
int ps_precalc(ps_engine *e, double freq_start, double freq_end, double freq_step, double *conw_r,
               int *new_conw);
/* The same round resumed after its first done_periods periods, whose best
 * dark-facet % sum to sum_dark_done: the state the optimised CPU app
 * checkpoints inside a round
 * (period_search_optimization_cpu/period_search/period_search_BOINC.cpp:627-640, 772-775).
 * ps_precalc is ps_precalc_resume with 0 and 0.0. */
int ps_precalc_resume(ps_engine *e, double freq_start, double freq_end, double freq_step, int done_periods,
                      double sum_dark_done, double *conw_r, int *new_conw);

/* Rows n_from..n_to (1-based, inclusive) of the scan at conw_r (CUDAStart).
 * rows: [n_to-n_from+1].  ndata_minus_3 is taken from the problem. */
int ps_scan(ps_engine *e, double freq_start, double freq_step, int n_from, int n_to, double conw_r,
            ps_row *rows);

/* Test seam: n independent fits; freq [n] (1/h), pole [n] (0-based index
 * into the problem's pole table). */
int ps_fit_batch(ps_engine *e, const double *freq, const int *pole, int n, double conw_r, ps_fit_out *out);

/* ps_fit_batch for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  freq, pole and out are DEVICE pointers
 * (caller-owned, n entries each, same meaning as in ps_fit_batch).  The fits'
 * slot memory, slot states and counters live in one fixed 64 MiB array per
 * device that the library holds until the process ends; the call allocates
 * nothing (no hipMalloc/hipFree, no graph, no timing events) and clears what
 * must start at zero with kernels.  Fewer slots fit there than the engine
 * normally uses, so large batches run in more refill rounds; results are
 * bitwise those of ps_fit_batch.  PS_ERR_LIMIT if one fit's slot memory
 * exceeds the array, PS_ERR_ARG on a pole index out of range (checked on the
 * device).  Calls that overlap on one device -- from any engine on it --
 * share the array and race: serialise them. */
int ps_fit_batch_noalloc(ps_engine *e, const double *freq, const int *pole, int n, double conw_r, ps_fit_out *out);
/* Test seam: one normal-equation build (mrqcof) at each of n parameter
 * vectors cg [n*ma] (ma = ncoef+8, reference order).  alpha [n*mfit*mfit]
 * (full symmetric), beta [n*mfit], chisq [n]; chck [n*3] = sum Area*Nor. */
int ps_eval_mrqcof(ps_engine *e, const double *cg, int n, double conw_r, double *alpha, double *beta,
                   double *chisq, double *chck);

/* Problem dimensions after ps_set_problem: ncoef, ma, mfit. */
int ps_dims(const ps_engine *e, int *ncoef, int *ma, int *mfit);

/* Time of the last fit-kernel launch [ms] measured with HIP events on the
 * engine's stream, and the number of fits it ran. */
int ps_last_kernel_ms(const ps_engine *e, double *ms, long long *fits);
/* The last batch of fits: GPU time (ms), fits, and LM iterations run over the
 * slots (every resident fit advances one iteration per LM iteration of the
 * batch).  No counterpart in the reference. */
int ps_last_batch(const ps_engine *e, double *ms, long long *items, long long *iterations);

/* Progress of a running call (ps_scan, ps_precalc, the fit batches): fn is
 * called on the calling thread between GPU polls with the fits finished so
 * far and the fits of the batch; the reference reports the same through
 * boinc_fraction_done (period_search/period_search_BOINC.cpp:809-810,
 * period_search_cuda/period_search/start_CUDA.cu:674-678).  fn = NULL turns
 * it off. */
typedef void (*ps_progress_fn)(void *user, long long fits_done, long long fits_total);
int ps_set_progress(ps_engine *e, ps_progress_fn fn, void *user);
/* Per-launch timing of the model kernel (the dominant kernel) with HIP
 * events on the engine's stream, and a count of model evaluations.  While
 * profiling, the LM iterations run as plain launches with events between
 * them instead of replaying the poll graph (the kernels are the same).
 * ps_set_profiling resets the counters. */
int ps_set_profiling(ps_engine *e, int on);
int ps_kernel_stats(const ps_engine *e, double *model_ms, long long *model_launches, long long *evals);

/* ---- host-side driver helpers ---------------------------------------- */
/* Parsed and normalised period_search_in (period_search_BOINC.cpp:277-485). */
typedef struct ps_input ps_input;
int ps_input_read(const char *path, ps_input **out);
void ps_input_free(ps_input *in);
/* Fill a ps_problem that points into the input's storage (valid until
 * ps_input_free).  Builds the shape basis (trifac, areanorm, sphfunc, ellfit). */
int ps_input_problem(ps_input *in, ps_problem *p);
/* Scalars of the input: per_start, per_end, per_step_coef, jd_min, jd_max,
 * escl, conw, freq_start, freq_end, freq_step; ndata (real points). */
int ps_input_scalars(const ps_input *in, double *out10, int *ndata);
/* period_search_out row text, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n";
 * round_pole != 0 prints round(lambda), round(beta) as the CUDA/OC apps do. */
int ps_format_row(const ps_row *r, int round_pole, char *buf, int buflen);

/* Shape basis alone: nrows, lmax, mmax, ellipsoid axes -> tables sized as in
 * ps_problem (nor [F*3], darea [F], pleg [F*(lmax+1)^2], fc/fs [F*(mmax+1)],
 * dsph [F*ncoef], cg_first [ncoef]).  Returns F, or <0. */
int ps_shape_basis(int nrows, int lmax, int mmax, double a, double b, double c, double *nor, double *darea,
                   double *pleg, double *fc, double *fs, double *dsph, double *cg_first);

/* One asteroid's request for ps_shape_basis_many: triangulation rows,
 * Laplace-series degree and order, ellipsoid axes, and host buffers for the
 * tables, sized as for ps_shape_basis. */
typedef struct ps_shape_job {
  int nrows, lmax, mmax;
  double a, b, c;
  double *nor, *darea, *pleg, *fc, *fs, *dsph, *cg_first;
} ps_shape_job;

/* Shape setup of n asteroids on the engine's GPU: the work of n
 * ps_shape_basis calls (trifac, areanorm, sphfunc, ellfit as called at
 * period_search_BOINC.cpp:507-552) in six launches.  The facet tables and
 * the ellipsoid normal matrix and its LU are built once per distinct
 * (nrows, lmax, mmax); per asteroid remain the ellipsoid's log-curvature, one
 * right-hand side and its substitution.  The arithmetic follows
 * ps_shape_basis's order; results differ only where the device's
 * sin/cos/acos/atan2/log round differently from the host's (<= 1 ulp).  Does
 * not touch the engine's problem.  Returns PS_OK or <0. */
int ps_shape_basis_many(ps_engine *e, int n, const ps_shape_job *jobs);

/* Post-scan selection over work-unit outputs (the project server's step
 * after the app; not part of the reference repository): rows of one or more
 * scans in one array.  *index = the row of least chi^2, the first of equal
 * minima; rows whose chi^2 is NaN are skipped.  PS_ERR_ARG if no row has a
 * number. */
int ps_best_row(const ps_row *rows, long long n, long long *index);
/* Read a period_search_out file (rows as ps_format_row writes them,
 * period_search_BOINC.cpp:926-931) into rows[0..cap); *n = rows read.
 * PS_ERR_IO on a malformed line, PS_ERR_LIMIT if the file has more than cap
 * rows. */
int ps_read_rows(const char *path, ps_row *rows, long long cap, long long *n);

/* ps_input_read without the shape basis: the tables are sized and zero until
 * ps_shape_basis_many fills them through ps_input_shape_job's pointers. */
int ps_input_read_deferred(const char *path, ps_input **out);
/* The input's shape request, pointing at its own table storage. */
int ps_input_shape_job(ps_input *in, ps_shape_job *job);

/* The best pole of each trial period (period_search_BOINC.cpp:900-924) over
 * fits from ps_fit_batch-like calls: per consecutive fits per period, the
 * first fit whose dev is strictly below the best so far (starting at 1e40)
 * wins, so a NaN dev never does and a period with none keeps dev 1e40 and
 * zeros; dark = |chck| / totarea * 100; lambda wrapped to [0, 360); chi2 =
 * dev^2 * n_points (n_points = ndata - 3, the real points). */
int ps_rows_from_fits(const ps_fit_out *fits, int n_periods, int per, int n_points, ps_row *rows);

/* ps_input_read with options:
 *   PS_READ_DEFER_SHAPE  as ps_input_read_deferred;
 *   PS_READ_CLAMP_MIN    each time and brightness below DBL_MIN is raised to
 *                        DBL_MIN as it is read, as the optimised CPU app does
 *                        (period_search_optimization_cpu/period_search/
 *                        period_search_BOINC.cpp:371-380); the base app
 *                        (default) keeps the values as read. */
#define PS_READ_DEFER_SHAPE 1
#define PS_READ_CLAMP_MIN 2
int ps_input_read_ex(const char *path, int flags, ps_input **out);

/* ---- checkpoint (period_search_state) ------------------------------- */
/* The three reference formats, told apart by their field count:
 *   PS_STATE_BASE  "nlines"                       period_search/period_search_BOINC.cpp:85-101
 *   PS_STATE_CUDA  "nlines newConw conw_r"        period_search_cuda/period_search/period_search_BOINC.cpp:96-112
 *   PS_STATE_OC    "nlines newConw conw_r sum_dark test_periods"
 *                                                 period_search_optimization_cpu/period_search/period_search_BOINC.cpp:86-101
 * nlines = rows written; newConw = 1 once the convexity weight is final;
 * sum_dark / test_periods = the OC app's progress inside a precalc round
 * (sum of the best dark-facet % over its first test_periods periods). */
#define PS_STATE_NONE 0
#define PS_STATE_BASE 1
#define PS_STATE_CUDA 2
#define PS_STATE_OC 3
typedef struct ps_state {
  int format;       /* PS_STATE_*; NONE = no usable checkpoint (start afresh, truncate the output) */
  int nlines;
  int new_conw;
  double conw_r;
  double sum_dark;
  int test_periods;
} ps_state;
/* Reads path.  A missing, empty, unparsable or inconsistent file (negative
 * nlines, non-positive conw_r) gives format PS_STATE_NONE and PS_OK, as the
 * reference apps fall back to a fresh start (period_search_BOINC.cpp:238-251). */
int ps_state_read(const char *path, ps_state *st);
/* Writes st in st->format to path + ".temp" and renames it over path, so a
 * reader never sees a partial file. */
int ps_state_write(const char *path, const ps_state *st);

#ifdef __cplusplus
}
#endif
#endif
