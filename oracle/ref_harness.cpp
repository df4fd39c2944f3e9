// ref_harness.cpp -- test harness around the reference's own numeric sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the unmodified reference sources period_search/{trifac,areanorm,sphfunc,
// ellfit,ludcmp,lubksb,mrqmin,mrqcof,bright,curv,conv,blmatrix,phasec,matrix,
// gauss_errc,dot_product,memory,covsrt}.c (g++ -O3, as period_search/Makefile:15-18)
// into oracle/_ref/libpsref.so.  The reference's driver cannot be built here
// (it needs the BOINC client headers and libraries), so this file does what
// the reference's unit tests do (period_search_pure_cpu_win/UnitTest_pure_cpu_win/
// InnerFunctionsTests.cpp:31-63): it defines the globals the numeric files
// read, and exposes extern "C" entry points so tests can drive those files
// and compare them with the oracle restatement bit for bit.
//
// The LM loop in ref_fit() restates the scan loop of
// period_search/period_search_BOINC.cpp:819-911 around the reference mrqmin().
#include <math.h>
#include <string.h>

#include "constants.h"
#include "declarations.h"
#include "globals.h"

// Globals, as defined in period_search/period_search_BOINC.cpp:129-143.
int Lmax, Mmax, Niter, Lastcall, Ncoef, Numfac, Lcurves, Nphpar, Lpoints[MAX_LC + 1], Inrel[MAX_LC + 1],
    Deallocate, n_iter;
double Ochisq, Chisq, Alamda, Alamda_incr, Alamda_start, Phi_0, Scale, Area[MAX_N_FAC + 1],
    Darea[MAX_N_FAC + 1], Sclnw[MAX_LC + 1], Yout[MAX_N_OBS + 1], Fc[MAX_N_FAC + 1][MAX_LM + 1],
    Fs[MAX_N_FAC + 1][MAX_LM + 1], Tc[MAX_N_FAC + 1][MAX_LM + 1], Ts[MAX_N_FAC + 1][MAX_LM + 1],
    Dsph[MAX_N_FAC + 1][MAX_N_PAR + 1], Dg[MAX_N_FAC + 1][MAX_N_PAR + 1], Nor[MAX_N_FAC + 1][4], Blmat[4][4],
    Pleg[MAX_N_FAC + 1][MAX_LM + 1][MAX_LM + 1], Dblm[3][4][4], Weight[MAX_N_OBS + 1];

static double **g_ee, **g_ee0, *g_tim, *g_brightness, *g_sig, *g_cg, *g_cg_first;
static double **g_covar, **g_aalpha, *g_t, *g_f, *g_at, *g_af;
static int *g_ia, **g_ifp;
static int g_ndata, g_ndir;

static void ensure_alloc() {
  if (g_ee) return;
  g_ee = matrix_double(MAX_N_OBS, 3);
  g_ee0 = matrix_double(MAX_N_OBS, 3);
  g_covar = matrix_double(MAX_N_PAR, MAX_N_PAR);
  g_aalpha = matrix_double(MAX_N_PAR, MAX_N_PAR);
  g_ifp = matrix_int(MAX_N_FAC, 4);
  g_tim = vector_double(MAX_N_OBS);
  g_brightness = vector_double(MAX_N_OBS);
  g_sig = vector_double(MAX_N_OBS);
  g_cg = vector_double(MAX_N_PAR);
  g_cg_first = vector_double(MAX_N_PAR);
  g_t = vector_double(MAX_N_FAC);
  g_f = vector_double(MAX_N_FAC);
  g_at = vector_double(MAX_N_FAC);
  g_af = vector_double(MAX_N_FAC);
  g_ia = vector_int(MAX_N_PAR);
}

extern "C" {

struct ref_fit_result {
  double dev, chisq, rchisq, prd, la, be, dark, totarea, alamda;
  int niter, pad;
};

// Load one problem.  Arrays are 1-indexed with slot 0 unused, laid out like
// the reference's: ee/ee0 are (ndata+1) x 4 (row stride 4), ia is (ma+1).
// ndata counts the 3 conv rows; Lcurves counts the conv light curve.
int ref_set_problem(int lmax, int mmax, int nrows, int lcurves, const int *lpoints, const int *inrel, int ndata,
                    const double *tim, const double *brightness, const double *sig, const double *ee4,
                    const double *ee04, const double *weight, double phi0, double alamda_incr,
                    double alamda_start, const int *ia, int ma) {
  ensure_alloc();
  if (ndata > MAX_N_OBS || lcurves > MAX_LC || 8 * nrows * nrows > MAX_N_FAC) return -1;
  for (int i = 1; i <= lcurves; i++)
    if (lpoints[i] > POINTS_MAX) return -2;
  Lmax = lmax;
  Mmax = mmax;
  Lcurves = lcurves;
  for (int i = 1; i <= lcurves; i++) {
    Lpoints[i] = lpoints[i];
    Inrel[i] = inrel[i];
  }
  g_ndata = ndata;
  for (int i = 1; i <= ndata; i++) {
    g_tim[i] = tim[i];
    g_brightness[i] = brightness[i];
    g_sig[i] = sig[i];
    Weight[i] = weight[i];
    for (int k = 1; k <= 3; k++) {
      g_ee[i][k] = ee4[4 * i + k];
      g_ee0[i][k] = ee04[4 * i + k];
    }
  }
  Phi_0 = phi0;
  Alamda_incr = alamda_incr;
  Alamda_start = alamda_start;
  Nphpar = 3;
  Ncoef = 0;
  for (int m = 0; m <= Mmax; m++)
    for (int l = m; l <= Lmax; l++) {
      Ncoef++;
      if (m != 0) Ncoef++;
    }
  for (int i = 1; i <= ma; i++) g_ia[i] = ia[i];
  // Gaussian-sphere vertices (period_search_BOINC.cpp:507-542), then the
  // reference trifac/areanorm/sphfunc.
  double dth = M_PI / (2 * nrows), dph;
  int k = 1;
  g_t[1] = 0;
  g_f[1] = 0;
  for (int i = 1; i <= nrows; i++) {
    dph = M_PI / (2 * i);
    for (int j = 0; j <= 4 * i - 1; j++) {
      k++;
      g_t[k] = i * dth;
      g_f[k] = j * dph;
    }
  }
  for (int i = nrows - 1; i >= 1; i--) {
    dph = M_PI / (2 * i);
    for (int j = 0; j <= 4 * i - 1; j++) {
      k++;
      g_t[k] = M_PI - i * dth;
      g_f[k] = j * dph;
    }
  }
  g_ndir = k + 1;
  g_t[g_ndir] = M_PI;
  g_f[g_ndir] = 0;
  Numfac = 8 * nrows * nrows;
  trifac(nrows, g_ifp);
  areanorm(g_t, g_f, g_ndir, Numfac, g_ifp, g_at, g_af);
  sphfunc(Numfac, g_at, g_af);
  return 0;
}

void ref_ellfit(double *cg_first_out, double a, double b, double c) {
  ellfit(g_cg_first, a, b, c, Numfac, Ncoef, g_at, g_af);
  for (int i = 1; i <= Ncoef; i++) cg_first_out[i] = g_cg_first[i];
}

void ref_set_cg_first(const double *cg_first) {
  for (int i = 1; i <= Ncoef; i++) g_cg_first[i] = cg_first[i];
}

// One LM fit (period_search_BOINC.cpp:819-911).  pole_beta/pole_lambda in deg.
void ref_fit(double freq, double pole_beta, double pole_lambda, const double *par, const int *ia_par, double cl,
             double stop_condition, int n_iter_min, double conw_r, ref_fit_result *out) {
  int i, j, n_iter_max;
  double prd, dev_old, dev_new, iter_diff, iter_diff_max, rchisq, totarea, sum, dark, chck[4];
  double *cg = g_cg;
  prd = 1 / freq;
  for (i = 1; i <= Ncoef; i++) cg[i] = g_cg_first[i];
  cg[Ncoef + 1] = pole_beta;
  cg[Ncoef + 2] = pole_lambda;
  cg[Ncoef + 1] = 90 - cg[Ncoef + 1];
  cg[Ncoef + 1] = DEG2RAD * cg[Ncoef + 1];
  cg[Ncoef + 2] = DEG2RAD * cg[Ncoef + 2];
  cg[Ncoef + 3] = 24 * 2 * PI / prd;
  for (i = 1; i <= Nphpar; i++) {
    cg[Ncoef + 3 + i] = par[i];
    g_ia[Ncoef + 3 + i] = ia_par[i];
  }
  cg[Ncoef + 3 + Nphpar + 2] = 1;
  cg[Ncoef + 3 + Nphpar + 1] = log(cl);
  n_iter_max = 0;
  iter_diff_max = -1;
  rchisq = -1;
  if (stop_condition > 1) {
    n_iter_max = (int)stop_condition;
    iter_diff_max = 0;
    n_iter_min = 0;
  }
  if (stop_condition < 1) {
    n_iter_max = MAX_N_ITER;
    iter_diff_max = stop_condition;
  }
  Alamda = -1;
  Niter = 0;
  iter_diff = 1e40;
  dev_old = 1e30;
  dev_new = 0;
  Lastcall = 0;
  chck[1] = chck[2] = chck[3] = 0;
  while (((Niter < n_iter_max) && (iter_diff > iter_diff_max)) || (Niter < n_iter_min)) {
    mrqmin(g_ee, g_ee0, g_tim, g_brightness, g_sig, cg, g_ia, Ncoef + 5 + Nphpar, g_covar, g_aalpha);
    Niter++;
    if ((Niter == 1) || (Chisq < Ochisq)) {
      Ochisq = Chisq;
      curv(cg);
      for (i = 1; i <= 3; i++) {
        chck[i] = 0;
        for (j = 1; j <= Numfac; j++) chck[i] = chck[i] + Area[j] * Nor[j][i];
      }
      rchisq = Chisq - (pow(chck[1], 2) + pow(chck[2], 2) + pow(chck[3], 2)) * pow(conw_r, 2);
    }
    dev_new = sqrt(rchisq / (g_ndata - 3));
    if (dev_old - dev_new > 1e-10) {
      iter_diff = dev_old - dev_new;
      dev_old = dev_new;
    }
  }
  Deallocate = 1;
  mrqmin(g_ee, g_ee0, g_tim, g_brightness, g_sig, cg, g_ia, Ncoef + 5 + Nphpar, g_covar, g_aalpha);
  Deallocate = 0;
  totarea = 0;
  for (i = 1; i <= Numfac; i++) totarea = totarea + Area[i];
  sum = pow(chck[1], 2) + pow(chck[2], 2) + pow(chck[3], 2);
  dark = sqrt(sum);
  out->dev = dev_new;
  out->chisq = Chisq;
  out->rchisq = rchisq;
  out->prd = 2 * PI / cg[Ncoef + 3];
  out->la = RAD2DEG * cg[Ncoef + 2];
  out->be = 90 - RAD2DEG * cg[Ncoef + 1];
  out->dark = dark;
  out->totarea = totarea;
  out->alamda = Alamda;
  out->niter = Niter;
}

// Function-level access.
void ref_curv(const double *cg) { curv(const_cast<double *>(cg)); }
double ref_bright(const double *ee, const double *ee0, double t, const double *cg, double *dyda, int ncoef) {
  double e[4] = {0, ee[1], ee[2], ee[3]}, e0[4] = {0, ee0[1], ee0[2], ee0[3]};
  return bright(e, e0, t, const_cast<double *>(cg), dyda, ncoef);
}
void ref_blmatrix(double bet, double lam) { blmatrix(bet, lam); }
// alpha_out: (MAX_N_PAR+1)^2 row-major, 1-indexed.
double ref_mrqcof(const double *cg, int ma, double *alpha_out, double *beta_out) {
  double beta[MAX_N_PAR + 1];
  double r = mrqcof(g_ee, g_ee0, g_tim, g_brightness, g_sig, const_cast<double *>(cg), g_ia, ma, g_aalpha, beta);
  int mfit = 0;
  for (int j = 1; j <= ma; j++)
    if (g_ia[j]) mfit++;
  for (int j = 1; j <= mfit; j++) {
    beta_out[j] = beta[j];
    for (int k = 1; k <= mfit; k++) alpha_out[j * (MAX_N_PAR + 1) + k] = g_aalpha[j][k];
  }
  return r;
}
// a: (MAX_N_PAR+1)^2 row-major, 1-indexed.
int ref_gauss_errc(double *a, int n, double *b) {
  for (int j = 1; j <= n; j++)
    for (int k = 1; k <= n; k++) g_covar[j][k] = a[j * (MAX_N_PAR + 1) + k];
  int r = gauss_errc(g_covar, n, b);
  for (int j = 1; j <= n; j++)
    for (int k = 1; k <= n; k++) a[j * (MAX_N_PAR + 1) + k] = g_covar[j][k];
  return r;
}

double *ref_dptr(const char *nm) {
  if (!strcmp(nm, "Area")) return Area;
  if (!strcmp(nm, "Darea")) return Darea;
  if (!strcmp(nm, "Dsph")) return &Dsph[0][0];
  if (!strcmp(nm, "Dg")) return &Dg[0][0];
  if (!strcmp(nm, "Nor")) return &Nor[0][0];
  if (!strcmp(nm, "Pleg")) return &Pleg[0][0][0];
  if (!strcmp(nm, "Fc")) return &Fc[0][0];
  if (!strcmp(nm, "Fs")) return &Fs[0][0];
  if (!strcmp(nm, "Blmat")) return &Blmat[0][0];
  if (!strcmp(nm, "Dblm")) return &Dblm[0][0][0];
  if (!strcmp(nm, "at")) return g_at;
  if (!strcmp(nm, "af")) return g_af;
  if (!strcmp(nm, "cg_first")) return g_cg_first;
  return 0;
}
int *ref_ifp_row(int i) { return g_ifp[i]; }
double ref_get_scale() { return Scale; }
int ref_ncoef() { return Ncoef; }
int ref_numfac() { return Numfac; }
}
