/*
 * ps_oracle.h -- FP64 CPU oracle for the trial-period fit loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference CPU app
 * (the period_search/ C sources and the scan loop of period_search/period_search_BOINC.cpp)
 * operation for operation, in the same order and precision, so that its
 * results are bit-identical to the reference built with gcc on x86-64.
 * It is the checker for the HIP engine: only tests/, the smoke() entry and
 * the cpu_baseline leg of bench.py may load it.  The product library never
 * links or calls it.
 *
 * State that the reference keeps in globals (period_search/globals.h:5-19)
 * lives in one orc_ctx, so several problems can coexist in one process.
 * Arrays stay 1-indexed, as in the reference, so each loop below reads like
 * the loop it restates.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Limits (period_search/constants.h:1-11).  ORC_POINTS_MAX follows the
 * CUDA/AVX apps (2000) rather than the base app (1000); it only sizes
 * buffers and does not change any arithmetic. */
#define ORC_POINTS_MAX 2000
#define ORC_MAX_N_OBS 20000
#define ORC_MAX_LC 200
#define ORC_MAX_N_FAC 1000
#define ORC_MAX_N_ITER 100
#define ORC_MAX_N_PAR 200
#define ORC_MAX_LM 10
#define ORC_N_PHOT_PAR 5
#define ORC_TINY 1e-8
#define ORC_N_POLES 10

typedef struct orc_ctx orc_ctx;

/* Result of one LM fit from one initial pole (period_search_BOINC.cpp:817-920). */
typedef struct orc_fit_result {
  double dev;      /* dev_new: sqrt(rchisq/(ndata-3)) after the last iteration */
  double chisq;    /* Chisq after the last iteration */
  double rchisq;   /* Chisq minus the convexity term */
  double prd;      /* 2*pi/omega [days] */
  double la;       /* lambda [deg] */
  double be;       /* beta [deg] */
  double dark;     /* |chck| */
  double totarea;  /* sum of Area of the last curv() call */
  double alamda;   /* final Marquardt lambda */
  int niter;       /* iterations run */
  int pad;
} orc_fit_result;

orc_ctx *orc_create(void);
void orc_destroy(orc_ctx *c);

/* Parse a period_search_in file and do everything the reference driver does
 * before the conw precalc (period_search_BOINC.cpp:277-485).  0 = OK. */
int orc_read_input(orc_ctx *c, const char *path);

/* Rebuild the shape basis and LM masks for the current conw_r
 * (period_search_BOINC.cpp:489-566 / 715-802).  Appends the 3 conv rows. */
void orc_prepare_round(orc_ctx *c, double conw_r);
/* Remove the 3 conv rows again (period_search_BOINC.cpp:713). */
void orc_end_round(orc_ctx *c);

/* One LM fit (period_search_BOINC.cpp:819-911). */
void orc_fit(orc_ctx *c, double freq, int pole, orc_fit_result *out);

/* One trial period over the 10 poles.  row[0..5] = per_best [days], dev_best,
 * dark_best, la_best, be_best, chisq column; returns via row.
 * (period_search_BOINC.cpp:812-924) */
void orc_period(orc_ctx *c, int n, double row[6], int precalc);

/* Convexity-weight precalc: the whole while loop (period_search_BOINC.cpp:487-715).
 * Returns the final conw_r. */
double orc_precalc(orc_ctx *c);

/* Scan rows n_from..n_to (1-based, inclusive) with the final conw_r; writes
 * the period_search_out text to out_path (mode "w" or "a").  Returns rows written. */
int orc_scan_to_file(orc_ctx *c, int n_from, int n_to, const char *out_path, const char *mode);

/* Number of trial periods: (int)((freq_start-freq_end)/freq_step)+1. */
int orc_n_periods(orc_ctx *c);

/* Accessors for tests (ctypes). */
double *orc_dptr(orc_ctx *c, const char *name);
int *orc_iptr(orc_ctx *c, const char *name);
double orc_get(orc_ctx *c, const char *name);
void orc_set(orc_ctx *c, const char *name, double v);

/* Function-level entry points that act on the context like the reference
 * functions act on their globals. */
void orc_trifac(orc_ctx *c, int nrows);
void orc_areanorm(orc_ctx *c, int ndir, int nfac);
void orc_sphfunc(orc_ctx *c, int ndir);
void orc_ellfit(orc_ctx *c, double *cg, double a, double b, double cc, int ndir, int ncoef);
void orc_ludcmp(double *a_rowmajor, int n, int *indx, double *d);
void orc_lubksb(const double *a_rowmajor, int n, const int *indx, double *b);
void orc_curv(orc_ctx *c, const double *cg);
void orc_blmatrix(orc_ctx *c, double bet, double lam);
void orc_matrix(orc_ctx *c, double omg, double t, double *tmat16, double *dtm64);
void orc_phasec(orc_ctx *c, double *dcdp, double alpha, const double *p);
double orc_bright(orc_ctx *c, const double *ee, const double *ee0, double t,
                  const double *cg, double *dyda, int ncoef);
double orc_conv(orc_ctx *c, int nc, double *dres, int ma);
double orc_mrqcof(orc_ctx *c, const double *cg, const int *ia, int ma,
                  double *alpha_rowmajor, double *beta);
int orc_gauss_errc(double *a_rowmajor, int n, double *b);

#ifdef __cplusplus
}
#endif
#endif
