/*
 * ps_oracle.c -- FP64 CPU oracle (test infrastructure only; see ps_oracle.h).
 *
 * Each function restates one reference function and cites it.  The
 * arithmetic keeps the reference's operation order (left-to-right sums that
 * start from 0, the same products and quotients), so that gcc on x86-64
 * (no FMA contraction) produces the same bits as the reference sources.
 * Reference citations are relative to the reference repository root.
 *
 * glibc's sincos() and separate sin()/cos() calls can differ in the last
 * bit.  The reference built with g++ -O3 fuses the pair only where the
 * compiler can see that no store in between may alias the argument: in
 * blmatrix, matrix and the first loop of ellfit.  This file is compiled with
 * -fno-builtin-sin -fno-builtin-cos (oracle/Makefile) and calls sincos()
 * explicitly at exactly those places, so it reproduces the same bits.
 */
#define _GNU_SOURCE
#include "ps_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NP (ORC_MAX_N_PAR + 1)
#define NF (ORC_MAX_N_FAC + 1)
#define NL (ORC_MAX_LM + 1)
#define NO (ORC_MAX_N_OBS + 4)
#define PI_ M_PI
#define DEG2RAD (PI_ / 180)
#define RAD2DEG (180 / PI_)

struct orc_ctx {
  /* globals of the reference (period_search/globals.h:5-19) */
  int Lmax, Mmax, Niter, Lastcall, Ncoef, Numfac, Lcurves, Nphpar, Deallocate;
  int Lpoints[ORC_MAX_LC + 2], Inrel[ORC_MAX_LC + 2];
  double Ochisq, Chisq, Alamda, Alamda_incr, Alamda_start, Phi_0, Scale;
  double Area[NF], Darea[NF];
  double Fc[NF][NL], Fs[NF][NL], Tc[NF][NL], Ts[NF][NL];
  double Dsph[NF][NP], Dg[NF][NP];
  double Nor[NF][4], Blmat[4][4], Pleg[NF][NL][NL], Dblm[3][4][4];
  double Weight[NO];

  /* mrqmin's statics (period_search/mrqmin.c:17-19) */
  int mfit;
  double atry[NP], beta[NP], da[NP];
  /* mrqcof's locals (period_search/mrqcof.c:20-24), kept off the stack */
  double ytemp[ORC_POINTS_MAX + 1], dytemp[ORC_POINTS_MAX + 1][NP], dave[NP];
  double covar[NP][NP], aalpha[NP][NP];

  /* driver data (period_search/period_search_BOINC.cpp:148-171) */
  int ndata, ndata_real, nrows, ndir, onlyrel, ia_prd, ia_par[4], ia_cl, n_iter_min_in;
  int ifp[NF][4];
  int ia[NP];
  double tim[NO], brightness[NO], sig[NO];
  double ee[NO][4], ee0[NO][4];
  double cg[NP], cg_first[NP];
  double t[NF], f[NF], at[NF], af[NF];
  double per_start, per_step_coef, per_end, jd_00, conw, par[4], cl, stop_condition;
  double jd_min, jd_max, escl, a_ax, b_ax, c_ax, conw_r;
  double freq_start, freq_end, freq_step;
  int new_conw;
  double chck[4];
  /* dev_new after iteration k of the last fit, k = 1..Niter (test record;
   * the reference keeps no such record) */
  double dev_trace[ORC_MAX_N_ITER + 2];
  /* test switch, not in the reference: 1 rounds every derivative bright()
   * and conv() return to float before mrqcof uses it (the precision of a
   * Jacobian held in FP32) */
  int dyda_f32;
};

orc_ctx *orc_create(void) { return (orc_ctx *)calloc(1, sizeof(orc_ctx)); }
void orc_destroy(orc_ctx *c) { free(c); }

/* ------------------------------------------------------------------ */
/* dot_product  (period_search/dot_product.c:8-17) */
static double dot3(const double a[4], const double b[4]) {
  double r;
  r = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  return r;
}

/* trifac  (period_search/trifac.c:10-82) */
static void trifac(orc_ctx *c, int nrows) {
  int nnod, i, j, j0, j1, j2, j3, ntri;
  int rows = 2 * nrows + 1, cols = 4 * nrows + 1;
  int *nod = (int *)calloc((size_t)rows * cols, sizeof(int));
#define NOD(a, b) nod[(a) * cols + (b)]
  nnod = 1;
  NOD(0, 0) = nnod;
  for (i = 1; i <= nrows; i++)
    for (j = 0; j <= 4 * i - 1; j++) {
      nnod++;
      NOD(i, j) = nnod;
      if (j == 0) NOD(i, 4 * i) = nnod;
    }
  for (i = nrows - 1; i >= 1; i--)
    for (j = 0; j <= 4 * i - 1; j++) {
      nnod++;
      NOD(2 * nrows - i, j) = nnod;
      if (j == 0) NOD(2 * nrows - i, 4 * i) = nnod;
    }
  NOD(2 * nrows, 0) = nnod + 1;
  ntri = 0;
  for (j1 = 1; j1 <= nrows; j1++)
    for (j3 = 1; j3 <= 4; j3++) {
      j0 = (j3 - 1) * j1;
      ntri++;
      c->ifp[ntri][1] = NOD(j1 - 1, j0 - (j3 - 1));
      c->ifp[ntri][2] = NOD(j1, j0);
      c->ifp[ntri][3] = NOD(j1, j0 + 1);
      for (j2 = j0 + 1; j2 <= j0 + j1 - 1; j2++) {
        ntri++;
        c->ifp[ntri][1] = NOD(j1, j2);
        c->ifp[ntri][2] = NOD(j1 - 1, j2 - (j3 - 1));
        c->ifp[ntri][3] = NOD(j1 - 1, j2 - 1 - (j3 - 1));
        ntri++;
        c->ifp[ntri][1] = NOD(j1 - 1, j2 - (j3 - 1));
        c->ifp[ntri][2] = NOD(j1, j2);
        c->ifp[ntri][3] = NOD(j1, j2 + 1);
      }
    }
  for (j1 = nrows + 1; j1 <= 2 * nrows; j1++)
    for (j3 = 1; j3 <= 4; j3++) {
      j0 = (j3 - 1) * (2 * nrows - j1);
      ntri++;
      c->ifp[ntri][1] = NOD(j1, j0);
      c->ifp[ntri][2] = NOD(j1 - 1, j0 + 1 + (j3 - 1));
      c->ifp[ntri][3] = NOD(j1 - 1, j0 + (j3 - 1));
      for (j2 = j0 + 1; j2 <= j0 + (2 * nrows - j1); j2++) {
        ntri++;
        c->ifp[ntri][1] = NOD(j1, j2);
        c->ifp[ntri][2] = NOD(j1 - 1, j2 + (j3 - 1));
        c->ifp[ntri][3] = NOD(j1, j2 - 1);
        ntri++;
        c->ifp[ntri][1] = NOD(j1, j2);
        c->ifp[ntri][2] = NOD(j1 - 1, j2 + 1 + (j3 - 1));
        c->ifp[ntri][3] = NOD(j1 - 1, j2 + (j3 - 1));
      }
    }
#undef NOD
  free(nod);
}

/* areanorm  (period_search/areanorm.c:10-61) */
static void areanorm(orc_ctx *c, int ndir, int nfac) {
  int i, j;
  double st, clen2, clen, cc[4], vx[4], vy[4], vz[4];
  double *x = (double *)malloc((ndir + 1) * sizeof(double));
  double *y = (double *)malloc((ndir + 1) * sizeof(double));
  double *z = (double *)malloc((ndir + 1) * sizeof(double));
  for (i = 1; i <= ndir; i++) {
    st = sin(c->t[i]);
    x[i] = st * cos(c->f[i]);
    y[i] = st * sin(c->f[i]);
    z[i] = cos(c->t[i]);
  }
  for (i = 1; i <= nfac; i++) {
    for (j = 2; j <= 3; j++) {
      vx[j] = x[c->ifp[i][j]] - x[c->ifp[i][1]];
      vy[j] = y[c->ifp[i][j]] - y[c->ifp[i][1]];
      vz[j] = z[c->ifp[i][j]] - z[c->ifp[i][1]];
    }
    cc[1] = vy[2] * vz[3] - vy[3] * vz[2];
    cc[2] = vz[2] * vx[3] - vz[3] * vx[2];
    cc[3] = vx[2] * vy[3] - vx[3] * vy[2];
    clen2 = cc[1] * cc[1] + cc[2] * cc[2] + cc[3] * cc[3];
    clen = sqrt(clen2);
    c->Nor[i][1] = cc[1] / clen;
    c->Nor[i][2] = cc[2] / clen;
    c->Nor[i][3] = cc[3] / clen;
    c->at[i] = acos(c->Nor[i][3]);
    c->af[i] = atan2(c->Nor[i][2], c->Nor[i][1]);
    c->Darea[i] = 0.5 * clen;
  }
  free(x);
  free(y);
  free(z);
}

/* sphfunc  (period_search/sphfunc.c:11-88) */
static void sphfunc(orc_ctx *c, int ndir) {
  int i, j, m, l, n, k, ibot;
  static double aleg[NL][NL][NL];
  const int Lmax = c->Lmax, Mmax = c->Mmax;
  for (i = 1; i <= ndir; i++) {
    c->Ts[i][0] = 1;
    c->Tc[i][0] = 1;
    c->Ts[i][1] = sin(c->at[i]);
    c->Tc[i][1] = cos(c->at[i]);
    for (j = 2; j <= Lmax; j++) {
      c->Ts[i][j] = c->Ts[i][1] * c->Ts[i][j - 1];
      c->Tc[i][j] = c->Tc[i][1] * c->Tc[i][j - 1];
    }
    c->Fs[i][0] = 0;
    c->Fc[i][0] = 1;
    for (j = 1; j <= Mmax; j++) {
      c->Fs[i][j] = sin(j * c->af[i]);
      c->Fc[i][j] = cos(j * c->af[i]);
    }
  }
  for (m = 0; m <= Lmax; m++)
    for (l = 0; l <= Lmax; l++)
      for (n = 0; n <= Lmax; n++) aleg[n][l][m] = 0;
  aleg[0][0][0] = 1;
  aleg[1][1][0] = 1;
  for (l = 1; l <= Lmax; l++) aleg[0][l][l] = aleg[0][l - 1][l - 1] * (2 * l - 1);
  for (m = 0; m <= Mmax; m++)
    for (l = m + 1; l <= Lmax; l++) {
      if ((2 * ((l - m) / 2)) == (l - m)) {
        aleg[0][l][m] = -(l + m - 1) * aleg[0][l - 2][m] / (1 * (l - m));
        ibot = 2;
      } else
        ibot = 1;
      if (l != 1)
        for (n = ibot; n <= l - m; n = n + 2)
          aleg[n][l][m] = ((2 * l - 1) * aleg[n - 1][l - 1][m] - (l + m - 1) * aleg[n][l - 2][m]) /
                          (1 * (l - m));
    }
  for (i = 1; i <= ndir; i++) {
    k = 0;
    for (m = 0; m <= Mmax; m++)
      for (l = m; l <= Lmax; l++) {
        c->Pleg[i][l][m] = 0;
        if ((2 * ((l - m) / 2)) == (l - m))
          ibot = 0;
        else
          ibot = 1;
        for (n = ibot; n <= l - m; n = n + 2)
          c->Pleg[i][l][m] = c->Pleg[i][l][m] + aleg[n][l][m] * c->Tc[i][n] * c->Ts[i][m];
        k++;
        c->Dsph[i][k] = c->Fc[i][m] * c->Pleg[i][l][m];
        if (m != 0) {
          k++;
          c->Dsph[i][k] = c->Fs[i][m] * c->Pleg[i][l][m];
        }
      }
  }
}

/* ludcmp  (period_search/ludcmp.c:9-69); a is (n+1)x(n+1), 1-indexed, row stride s */
static void ludcmp_s(double *a, int s, int n, int *indx, double *d) {
#define A(i, j) a[(i) * s + (j)]
  int i, imax = -999, j, k;
  double big, dum, sum, temp;
  double *v = (double *)malloc((n + 1) * sizeof(double));
  *d = 1.0;
  for (i = 1; i <= n; i++) {
    big = 0.0;
    for (j = 1; j <= n; j++)
      if ((temp = fabs(A(i, j))) > big) big = temp;
    if (big == 0.0) {
      fprintf(stderr, "Singular matrix in routine ludcmp\n");
      free(v);
      return;
    }
    v[i] = 1.0 / big;
  }
  for (j = 1; j <= n; j++) {
    for (i = 1; i < j; i++) {
      sum = A(i, j);
      for (k = 1; k < i; k++) sum -= A(i, k) * A(k, j);
      A(i, j) = sum;
    }
    big = 0.0;
    for (i = j; i <= n; i++) {
      sum = A(i, j);
      for (k = 1; k < j; k++) sum -= A(i, k) * A(k, j);
      A(i, j) = sum;
      if ((dum = v[i] * fabs(sum)) >= big) {
        big = dum;
        imax = i;
      }
    }
    if (j != imax) {
      for (k = 1; k <= n; k++) {
        dum = A(imax, k);
        A(imax, k) = A(j, k);
        A(j, k) = dum;
      }
      *d = -(*d);
      v[imax] = v[j];
    }
    indx[j] = imax;
    if (A(j, j) == 0.0) A(j, j) = 1.0e-20;
    if (j != n) {
      dum = 1.0 / (A(j, j));
      for (i = j + 1; i <= n; i++) A(i, j) *= dum;
    }
  }
  free(v);
#undef A
}

/* lubksb  (period_search/lubksb.c:3-26) */
static void lubksb_s(const double *a, int s, int n, const int *indx, double *b) {
#define A(i, j) a[(i) * s + (j)]
  int i, ii = 0, ip, j;
  double sum;
  for (i = 1; i <= n; i++) {
    ip = indx[i];
    sum = b[ip];
    b[ip] = b[i];
    if (ii)
      for (j = ii; j <= i - 1; j++) sum -= A(i, j) * b[j];
    else if (sum)
      ii = i;
    b[i] = sum;
  }
  for (i = n; i >= 1; i--) {
    sum = b[i];
    for (j = i + 1; j <= n; j++) sum -= A(i, j) * b[j];
    b[i] = sum / A(i, i);
  }
#undef A
}

/* ellfit  (period_search/ellfit.c:11-89) */
static void ellfit(orc_ctx *c, double *cg, double a, double b, double cc, int ndir, int ncoef) {
  int i, m, l, n, j, k;
  double sum, st, d;
  int s = ncoef + 1;
  int *indx = (int *)malloc((ncoef + 1) * sizeof(int));
  double *fitvec = (double *)malloc((ncoef + 1) * sizeof(double));
  double *er = (double *)malloc((ndir + 1) * sizeof(double));
  double *fmat = (double *)malloc((size_t)(ndir + 1) * s * sizeof(double));
  double *fitmat = (double *)malloc((size_t)s * s * sizeof(double));
  for (i = 1; i <= ndir; i++) {
    double ct, sa, ca;
    sincos(c->at[i], &st, &ct);
    sincos(c->af[i], &sa, &ca);
    sum = pow(a * st * ca, 2) + pow(b * st * sa, 2) + pow(cc * ct, 2);
    er[i] = 2 * (log(a * b * cc) - log(sum));
  }
  for (i = 1; i <= ndir; i++) {
    n = 0;
    for (m = 0; m <= c->Mmax; m++)
      for (l = m; l <= c->Lmax; l++) {
        n++;
        if (m != 0) {
          fmat[i * s + n] = c->Pleg[i][l][m] * cos(m * c->af[i]);
          n++;
          fmat[i * s + n] = c->Pleg[i][l][m] * sin(m * c->af[i]);
        } else
          fmat[i * s + n] = c->Pleg[i][l][m];
      }
  }
  for (i = 1; i <= ncoef; i++) {
    for (j = 1; j <= ncoef; j++) {
      fitmat[i * s + j] = 0;
      for (k = 1; k <= ndir; k++) fitmat[i * s + j] = fitmat[i * s + j] + fmat[k * s + i] * fmat[k * s + j];
    }
    fitvec[i] = 0;
    for (j = 1; j <= ndir; j++) fitvec[i] = fitvec[i] + fmat[j * s + i] * er[j];
  }
  ludcmp_s(fitmat, s, ncoef, indx, &d);
  lubksb_s(fitmat, s, ncoef, indx, fitvec);
  for (i = 1; i <= ncoef; i++) cg[i] = fitvec[i];
  free(indx);
  free(fitvec);
  free(er);
  free(fmat);
  free(fitmat);
}

/* ------------------------------------------------------------------ */
/* curv  (period_search/curv.c:10-38) */
static void curv(orc_ctx *c, const double cg[]) {
  int i, m, n, l, k;
  double fsum, g;
  for (i = 1; i <= c->Numfac; i++) {
    g = 0;
    n = 0;
    for (m = 0; m <= c->Mmax; m++)
      for (l = m; l <= c->Lmax; l++) {
        n++;
        fsum = cg[n] * c->Fc[i][m];
        if (m != 0) {
          n++;
          fsum = fsum + cg[n] * c->Fs[i][m];
        }
        g = g + c->Pleg[i][l][m] * fsum;
      }
    g = exp(g);
    c->Area[i] = c->Darea[i] * g;
    for (k = 1; k <= n; k++) c->Dg[i][k] = g * c->Dsph[i][k];
  }
}

/* blmatrix  (period_search/blmatrix.c:9-46) */
static void blmatrix(orc_ctx *c, double bet, double lam) {
  double cb, sb, cl, sl;
  sincos(bet, &sb, &cb);
  sincos(lam, &sl, &cl);
  c->Blmat[1][1] = cb * cl;
  c->Blmat[1][2] = cb * sl;
  c->Blmat[1][3] = -sb;
  c->Blmat[2][1] = -sl;
  c->Blmat[2][2] = cl;
  c->Blmat[2][3] = 0;
  c->Blmat[3][1] = sb * cl;
  c->Blmat[3][2] = sb * sl;
  c->Blmat[3][3] = cb;
  c->Dblm[1][1][1] = -sb * cl;
  c->Dblm[1][1][2] = -sb * sl;
  c->Dblm[1][1][3] = -cb;
  c->Dblm[1][2][1] = 0;
  c->Dblm[1][2][2] = 0;
  c->Dblm[1][2][3] = 0;
  c->Dblm[1][3][1] = cb * cl;
  c->Dblm[1][3][2] = cb * sl;
  c->Dblm[1][3][3] = -sb;
  c->Dblm[2][1][1] = -cb * sl;
  c->Dblm[2][1][2] = cb * cl;
  c->Dblm[2][1][3] = 0;
  c->Dblm[2][2][1] = -cl;
  c->Dblm[2][2][2] = -sl;
  c->Dblm[2][2][3] = 0;
  c->Dblm[2][3][1] = -sb * sl;
  c->Dblm[2][3][2] = sb * cl;
  c->Dblm[2][3][3] = 0;
}

/* matrix  (period_search/matrix.c:11-58) */
static void matrix(orc_ctx *c, double omg, double t, double tmat[4][4], double dtm[4][4][4]) {
  double f, cf, sf, dfm[4][4], fmat[4][4];
  int i, j, k;
  f = omg * t + c->Phi_0;
  f = fmod(f, 2 * PI_);
  sincos(f, &sf, &cf);
  fmat[1][1] = cf;
  fmat[1][2] = sf;
  fmat[1][3] = 0;
  fmat[2][1] = -sf;
  fmat[2][2] = cf;
  fmat[2][3] = 0;
  fmat[3][1] = 0;
  fmat[3][2] = 0;
  fmat[3][3] = 1;
  dfm[1][1] = -t * sf;
  dfm[1][2] = t * cf;
  dfm[1][3] = 0;
  dfm[2][1] = -t * cf;
  dfm[2][2] = -t * sf;
  dfm[2][3] = 0;
  dfm[3][1] = 0;
  dfm[3][2] = 0;
  dfm[3][3] = 0;
  for (i = 1; i <= 3; i++)
    for (j = 1; j <= 3; j++) {
      tmat[i][j] = 0;
      dtm[1][i][j] = 0;
      dtm[2][i][j] = 0;
      dtm[3][i][j] = 0;
      for (k = 1; k <= 3; k++) {
        tmat[i][j] = tmat[i][j] + fmat[i][k] * c->Blmat[k][j];
        dtm[1][i][j] = dtm[1][i][j] + fmat[i][k] * c->Dblm[1][k][j];
        dtm[2][i][j] = dtm[2][i][j] + fmat[i][k] * c->Dblm[2][k][j];
        dtm[3][i][j] = dtm[3][i][j] + dfm[i][k] * c->Blmat[k][j];
      }
    }
}

/* phasec  (period_search/phasec.c:11-24) */
static void phasec(orc_ctx *c, double dcdp[], double alpha, const double p[]) {
  double e, cc;
  e = exp(-alpha / p[2]);
  cc = 1 + p[1] * e + p[3] * alpha;
  dcdp[1] = e;
  dcdp[2] = p[1] * e * alpha / (p[2] * p[2]);
  dcdp[3] = alpha;
  c->Scale = cc;
}

/* bright  (period_search/bright.c:14-126) */
static double bright(orc_ctx *c, const double ee[4], const double ee0[4], double t, const double cg[],
                     double dyda[], int ncoef) {
  int ncoef0, i, j, k;
  static int incl[NF];
  static double mu[NF], mu0[NF], s[NF], dbr[NF], dsmu[NF], dsmu0[NF];
  double cos_alpha, br, cl, cls, alpha, sum, sum0, dnom;
  double e[4], e0[4], php[ORC_N_PHOT_PAR + 1], dphp[ORC_N_PHOT_PAR + 1];
  double de[4][4], de0[4][4], tmat[4][4], dtm[4][4][4];
  const int Numfac = c->Numfac, Nphpar = c->Nphpar;

  ncoef0 = ncoef - 2 - Nphpar;
  cl = exp(cg[ncoef - 1]);
  cls = cg[ncoef];
  cos_alpha = dot3(ee, ee0);
  alpha = acos(cos_alpha);
  for (i = 1; i <= Nphpar; i++) php[i] = cg[ncoef0 + i];
  phasec(c, dphp, alpha, php);
  matrix(c, cg[ncoef0], t, tmat, dtm);
  br = 0;
  for (i = 1; i <= 3; i++) {
    e[i] = 0;
    e0[i] = 0;
    for (j = 1; j <= 3; j++) {
      e[i] = e[i] + tmat[i][j] * ee[j];
      e0[i] = e0[i] + tmat[i][j] * ee0[j];
      de[i][j] = 0;
      de0[i][j] = 0;
      for (k = 1; k <= 3; k++) {
        de[i][j] = de[i][j] + dtm[j][i][k] * ee[k];
        de0[i][j] = de0[i][j] + dtm[j][i][k] * ee0[k];
      }
    }
  }
  for (i = 1; i <= Numfac; i++) {
    incl[i] = 0;
    mu[i] = e[1] * c->Nor[i][1] + e[2] * c->Nor[i][2] + e[3] * c->Nor[i][3];
    mu0[i] = e0[1] * c->Nor[i][1] + e0[2] * c->Nor[i][2] + e0[3] * c->Nor[i][3];
    if ((mu[i] > ORC_TINY) && (mu0[i] > ORC_TINY)) {
      incl[i] = 1;
      dnom = mu[i] + mu0[i];
      s[i] = mu[i] * mu0[i] * (cl + cls / dnom);
      br = br + c->Area[i] * s[i];
      dsmu[i] = cls * pow(mu0[i] / dnom, 2) + cl * mu0[i];
      dsmu0[i] = cls * pow(mu[i] / dnom, 2) + cl * mu[i];
      dbr[i] = c->Darea[i] * s[i];
    }
  }
  for (i = 1; i <= ncoef0 - 3; i++) {
    dyda[i] = 0;
    for (j = 1; j <= Numfac; j++)
      if (incl[j] == 1) dyda[i] = dyda[i] + dbr[j] * c->Dg[j][i];
    dyda[i] = c->Scale * dyda[i];
  }
  for (k = 1; k <= 3; k++) {
    dyda[ncoef0 - 3 + k] = 0;
    for (i = 1; i <= Numfac; i++)
      if (incl[i] == 1) {
        sum = 0;
        sum0 = 0;
        for (j = 1; j <= 3; j++) {
          sum = sum + c->Nor[i][j] * de[j][k];
          sum0 = sum0 + c->Nor[i][j] * de0[j][k];
        }
        dyda[ncoef0 - 3 + k] = dyda[ncoef0 - 3 + k] + c->Area[i] * (dsmu[i] * sum + dsmu0[i] * sum0);
      }
    dyda[ncoef0 - 3 + k] = c->Scale * dyda[ncoef0 - 3 + k];
  }
  for (i = 1; i <= Nphpar; i++) dyda[ncoef0 + i] = br * dphp[i];
  dyda[ncoef - 1] = 0;
  dyda[ncoef] = 0;
  for (i = 1; i <= Numfac; i++)
    if (incl[i] == 1) {
      dyda[ncoef - 1] = dyda[ncoef - 1] + mu[i] * mu0[i] * c->Area[i];
      dyda[ncoef] = dyda[ncoef] + c->Area[i] * mu[i] * mu0[i] / (mu[i] + mu0[i]);
    }
  dyda[ncoef - 1] = c->Scale * dyda[ncoef - 1] * cl;
  dyda[ncoef] = c->Scale * dyda[ncoef];
  br *= c->Scale;
  return br;
}

/* conv  (period_search/conv.c:12-29) */
static double conv(orc_ctx *c, int nc, double dres[], int ma) {
  int i, j;
  double res = 0;
  for (j = 1; j <= ma; j++) dres[j] = 0;
  for (i = 1; i <= c->Numfac; i++) {
    res += c->Area[i] * c->Nor[i][nc];
    for (j = 1; j <= c->Ncoef; j++) dres[j] += c->Darea[i] * c->Dg[i][j] * c->Nor[i][nc];
  }
  return res;
}

/* mrqcof  (period_search/mrqcof.c:16-148).  The Lastcall==1 branches
 * (light-curve output, never taken by the period search) are omitted. */
static double mrqcof(orc_ctx *c, const double a[], const int ia[], int ma, double alpha[][NP],
                     double beta[]) {
  int mfit, i, j, k, l, m, np, np1, np2, jp, ic;
  double xx1[4], xx2[4], dy, sig2i, wt, dyda[NP], ymod, coef, ave = 0, trial_chisq, wght;

  curv(c, a);
  blmatrix(c, a[ma - 4 - c->Nphpar], a[ma - 3 - c->Nphpar]);
  mfit = 0;
  for (j = 1; j <= ma; j++)
    if (ia[j]) mfit++;
  for (j = 1; j <= mfit; j++) {
    for (k = 1; k <= j; k++) alpha[j][k] = 0;
    beta[j] = 0;
  }
  trial_chisq = 0;
  np = 0;
  np1 = 0;
  np2 = 0;
  for (i = 1; i <= c->Lcurves; i++) {
    if (c->Inrel[i] == 1) {
      ave = 0;
      for (l = 1; l <= ma; l++) c->dave[l] = 0;
    }
    for (jp = 1; jp <= c->Lpoints[i]; jp++) {
      np++;
      for (ic = 1; ic <= 3; ic++) {
        xx1[ic] = c->ee[np][ic];
        xx2[ic] = c->ee0[np][ic];
      }
      if (i < c->Lcurves)
        ymod = bright(c, xx1, xx2, c->tim[np], a, dyda, ma);
      else
        ymod = conv(c, jp, dyda, ma);
      if (c->dyda_f32)
        for (l = 1; l <= ma; l++) dyda[l] = (float)dyda[l];
      c->ytemp[jp] = ymod;
      if (c->Inrel[i] == 1) ave = ave + ymod;
      for (l = 1; l <= ma; l++) {
        c->dytemp[jp][l] = dyda[l];
        if (c->Inrel[i] == 1) c->dave[l] = c->dave[l] + dyda[l];
      }
    }
    for (jp = 1; jp <= c->Lpoints[i]; jp++) {
      np1++;
      if (c->Inrel[i] == 1) {
        coef = c->sig[np1] * c->Lpoints[i] / ave;
        for (l = 1; l <= ma; l++) c->dytemp[jp][l] = coef * (c->dytemp[jp][l] - c->ytemp[jp] * c->dave[l] / ave);
        c->ytemp[jp] = coef * c->ytemp[jp];
        c->dytemp[jp][1] = 0;
      }
    }
    for (jp = 1; jp <= c->Lpoints[i]; jp++) {
      ymod = c->ytemp[jp];
      for (l = 1; l <= ma; l++) dyda[l] = c->dytemp[jp][l];
      np2++;
      sig2i = 1 / (c->sig[np2] * c->sig[np2]);
      wght = c->Weight[np2];
      dy = c->brightness[np2] - ymod;
      j = 0;
      for (l = 1; l <= ma; l++) {
        if (ia[l]) {
          j++;
          wt = dyda[l] * sig2i * wght;
          k = 0;
          for (m = 1; m <= l; m++) {
            if (ia[m]) {
              k++;
              alpha[j][k] = alpha[j][k] + wt * dyda[m];
            }
          }
          beta[j] = beta[j] + dy * wt;
        }
      }
      trial_chisq = trial_chisq + dy * dy * sig2i * wght;
    }
  }
  for (j = 2; j <= mfit; j++)
    for (k = 1; k <= j - 1; k++) alpha[k][j] = alpha[j][k];
  return trial_chisq;
}

/* gauss_errc  (period_search/gauss_errc.c:9-73).  NR gaussj with full
 * pivoting; 1 = singular (pivot column reused), 2 = zero pivot. */
static int gauss_errc(double a[][NP], int n, double b[]) {
  int indxc[NP], indxr[NP], ipiv[NP];
  int i, icol = 0, irow = 0, j, k, l, ll;
  double big, dum, pivinv, temp;
  for (j = 1; j <= n; j++) ipiv[j] = 0;
  for (i = 1; i <= n; i++) {
    big = 0.0;
    for (j = 1; j <= n; j++)
      if (ipiv[j] != 1)
        for (k = 1; k <= n; k++) {
          if (ipiv[k] == 0) {
            if (fabs(a[j][k]) >= big) {
              big = fabs(a[j][k]);
              irow = j;
              icol = k;
            }
          } else if (ipiv[k] > 1)
            return 1;
        }
    ++(ipiv[icol]);
    if (irow != icol) {
      for (l = 1; l <= n; l++) {
        temp = a[irow][l];
        a[irow][l] = a[icol][l];
        a[icol][l] = temp;
      }
      temp = b[irow];
      b[irow] = b[icol];
      b[icol] = temp;
    }
    indxr[i] = irow;
    indxc[i] = icol;
    if (a[icol][icol] == 0.0) return 2;
    pivinv = 1.0 / a[icol][icol];
    a[icol][icol] = 1.0;
    for (l = 1; l <= n; l++) a[icol][l] *= pivinv;
    b[icol] *= pivinv;
    for (ll = 1; ll <= n; ll++)
      if (ll != icol) {
        dum = a[ll][icol];
        a[ll][icol] = 0.0;
        for (l = 1; l <= n; l++) a[ll][l] -= a[icol][l] * dum;
        b[ll] -= b[icol] * dum;
      }
  }
  for (l = n; l >= 1; l--) {
    if (indxr[l] != indxc[l])
      for (k = 1; k <= n; k++) {
        temp = a[k][indxr[l]];
        a[k][indxr[l]] = a[k][indxc[l]];
        a[k][indxc[l]] = temp;
      }
  }
  return 0;
}

/* mrqmin  (period_search/mrqmin.c:11-110); Lastcall is always 0 in the
 * period search, so its branches are omitted. */
static int mrqmin(orc_ctx *c, double a[], const int ia[], int ma, double covar[][NP], double alpha[][NP]) {
  int j, k, l, err_code;
  double temp;
  if (c->Deallocate == 1) return 0;
  if (c->Alamda < 0) {
    c->mfit = 0;
    for (j = 1; j <= ma; j++)
      if (ia[j]) c->mfit++;
    c->Alamda = c->Alamda_start;
    temp = mrqcof(c, a, ia, ma, alpha, c->beta);
    c->Ochisq = temp;
    for (j = 1; j <= ma; j++) c->atry[j] = a[j];
  }
  for (j = 1; j <= c->mfit; j++) {
    for (k = 1; k <= c->mfit; k++) covar[j][k] = alpha[j][k];
    covar[j][j] = alpha[j][j] * (1 + c->Alamda);
    c->da[j] = c->beta[j];
  }
  err_code = gauss_errc(covar, c->mfit, c->da);
  if (err_code != 0) return err_code;
  j = 0;
  for (l = 1; l <= ma; l++)
    if (ia[l]) {
      j++;
      c->atry[l] = a[l] + c->da[j];
    }
  temp = mrqcof(c, c->atry, ia, ma, covar, c->da);
  c->Chisq = temp;
  if (temp < c->Ochisq) {
    c->Alamda = c->Alamda / c->Alamda_incr;
    for (j = 1; j <= c->mfit; j++) {
      for (k = 1; k <= c->mfit; k++) alpha[j][k] = covar[j][k];
      c->beta[j] = c->da[j];
    }
    for (l = 1; l <= ma; l++) a[l] = c->atry[l];
  } else {
    c->Alamda = c->Alamda_incr * c->Alamda;
    c->Chisq = c->Ochisq;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Driver pieces (period_search/period_search_BOINC.cpp) */

static const double kBetaPole[ORC_N_POLES + 1] = {0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
static const double kLambdaPole[ORC_N_POLES + 1] = {0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300};

/* period_search_BOINC.cpp:277-485 (weights parsing follows the checked
 * fscanf of period_search_optimization_cpu/period_search/period_search_BOINC.cpp:440-450,
 * which gives the same result as the base loop on well-formed files). */
int orc_read_input(orc_ctx *c, const char *path) {
  char line[1000];
  int i, j, k, k2, i_temp, ial0, ial0_abs, lc_number;
  double e[4], e0[4], e0len, elen, ave, cos_alpha, al0, al0_abs, jd_0, rfit, a, b, c_axis;
  double al[ORC_MAX_LC + 2], weight_lc[ORC_MAX_LC + 2];
  const double a0 = 1.05, b0 = 1.00, c0 = 0.95;
  FILE *in = fopen(path, "r");
  if (!in) return -1;
#define SKIP() (void)!fgets(line, sizeof line, in)
  if (fscanf(in, "%lf %lf %lf %d", &c->per_start, &c->per_step_coef, &c->per_end, &c->ia_prd) != 4) goto bad;
  SKIP();
  if (fscanf(in, "%lf", &c->jd_00) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%lf", &c->Phi_0) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%lf", &c->conw) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%d %d", &c->Lmax, &c->Mmax) != 2) goto bad;
  SKIP();
  if (fscanf(in, "%d", &c->nrows) != 1) goto bad;
  SKIP();
  for (i = 1; i <= 3; i++) {
    if (fscanf(in, "%lf %d", &c->par[i], &c->ia_par[i]) != 2) goto bad;
    SKIP();
  }
  if (fscanf(in, "%lf %d", &c->cl, &c->ia_cl) != 2) goto bad;
  SKIP();
  if (fscanf(in, "%lf", &c->stop_condition) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%d", &c->n_iter_min_in) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%lf", &c->Alamda_incr) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%lf", &c->Alamda_start) != 1) goto bad;
  SKIP();
  if (fscanf(in, "%d", &c->Lcurves) != 1) goto bad;
  if (c->Lcurves > ORC_MAX_LC) goto bad;

  c->ndata = 0;
  k2 = 0;
  al0 = al0_abs = PI_;
  ial0 = ial0_abs = -1;
  c->jd_min = 1e20;
  c->jd_max = -1e40;
  c->onlyrel = 1;
  jd_0 = c->jd_00;
  a = a0;
  b = b0;
  c_axis = c0;
  for (i = 1; i <= c->Lcurves; i++) {
    ave = 0;
    if (fscanf(in, "%d %d", &c->Lpoints[i], &i_temp) != 2) goto bad;
    SKIP();
    c->Inrel[i] = 1 - i_temp;
    if (c->Inrel[i] == 0) c->onlyrel = 0;
    if (c->Lpoints[i] > ORC_POINTS_MAX) goto bad;
    for (j = 1; j <= c->Lpoints[i]; j++) {
      c->ndata++;
      if (c->ndata > ORC_MAX_N_OBS) goto bad;
      if (fscanf(in, "%lf %lf", &c->tim[c->ndata], &c->brightness[c->ndata]) != 2) goto bad;
      if (fscanf(in, "%lf %lf %lf", &e0[1], &e0[2], &e0[3]) != 3) goto bad;
      if (fscanf(in, "%lf %lf %lf", &e[1], &e[2], &e[3]) != 3) goto bad;
      if (c->tim[c->ndata] < c->jd_min) c->jd_min = c->tim[c->ndata];
      if (c->tim[c->ndata] > c->jd_max) c->jd_max = c->tim[c->ndata];
      e0len = sqrt(e0[1] * e0[1] + e0[2] * e0[2] + e0[3] * e0[3]);
      elen = sqrt(e[1] * e[1] + e[2] * e[2] + e[3] * e[3]);
      ave += c->brightness[c->ndata];
      for (k = 1; k <= 3; k++) {
        c->ee[c->ndata][k] = e[k] / elen;
        c->ee0[c->ndata][k] = e0[k] / e0len;
      }
      if (j == 1) {
        cos_alpha = dot3(e, e0) / (elen * e0len);
        al[i] = acos(cos_alpha);
        if (al[i] < al0) {
          al0 = al[i];
          ial0 = c->ndata;
        }
        if ((al[i] < al0_abs) && (c->Inrel[i] == 0)) {
          al0_abs = al[i];
          ial0_abs = c->ndata;
        }
      }
    }
    ave /= c->Lpoints[i];
    for (j = 1; j <= c->Lpoints[i]; j++) {
      k2++;
      c->sig[k2] = ave;
    }
  }
  for (i = 1; i <= c->Lcurves; i++) weight_lc[i] = -1;
  while (fscanf(in, "%d", &lc_number) == 1) {
    double w;
    if (fscanf(in, "%lf", &w) != 1) break;
    if (lc_number >= 1 && lc_number <= c->Lcurves) weight_lc[lc_number] = w;
  }
  fclose(in);
  in = NULL;
  if (jd_0 <= 0) jd_0 = (int)c->jd_min;
  for (i = 1; i <= c->ndata; i++) c->tim[i] = c->tim[i] - jd_0;
  c->Phi_0 = c->Phi_0 * DEG2RAD;
  k = 0;
  for (i = 1; i <= c->Lcurves; i++)
    for (j = 1; j <= c->Lpoints[i]; j++) {
      k++;
      if (weight_lc[i] == -1)
        c->Weight[k] = 1;
      else
        c->Weight[k] = weight_lc[i];
    }
  for (i = 1; i <= 3; i++) c->Weight[k + i] = 1;
  if (c->onlyrel == 0) {
    al0 = al0_abs;
    ial0 = ial0_abs;
  }
  rfit = sqrt(2 * c->sig[ial0] / (0.5 * PI_ * (1 + cos(al0))));
  c->escl = rfit / sqrt((a * b + b * c_axis + a * c_axis) / 3);
  if (c->onlyrel == 0) c->escl *= 0.8;
  c->a_ax = a * c->escl;
  c->b_ax = b * c->escl;
  c->c_ax = c_axis * c->escl;
  c->Lcurves = c->Lcurves + 1;
  c->Lpoints[c->Lcurves] = 3;
  c->Inrel[c->Lcurves] = 0;
  c->conw_r = c->conw / c->escl / c->escl;
  c->ndata_real = c->ndata;
  c->new_conw = 0;
  c->Nphpar = 3;
  return 0;
bad:
  if (in) fclose(in);
  return -2;
#undef SKIP
}

/* period_search_BOINC.cpp:491-566 (and its copy :717-802): conv rows,
 * shape basis, frequency grid and ia mask. */
void orc_prepare_round(orc_ctx *c, double conw_r) {
  int i, j, k, l, m, nrows = c->nrows;
  double dth, dph;
  c->conw_r = conw_r;
  c->ndata = c->ndata_real;
  for (j = 1; j <= 3; j++) {
    c->ndata++;
    c->brightness[c->ndata] = 0;
    c->sig[c->ndata] = 1 / conw_r;
  }
  c->Ncoef = 0;
  for (m = 0; m <= c->Mmax; m++)
    for (l = m; l <= c->Lmax; l++) {
      c->Ncoef++;
      if (m != 0) c->Ncoef++;
    }
  dth = PI_ / (2 * nrows);
  k = 1;
  c->t[1] = 0;
  c->f[1] = 0;
  for (i = 1; i <= nrows; i++) {
    dph = PI_ / (2 * i);
    for (j = 0; j <= 4 * i - 1; j++) {
      k++;
      c->t[k] = i * dth;
      c->f[k] = j * dph;
    }
  }
  for (i = nrows - 1; i >= 1; i--) {
    dph = PI_ / (2 * i);
    for (j = 0; j <= 4 * i - 1; j++) {
      k++;
      c->t[k] = PI_ - i * dth;
      c->f[k] = j * dph;
    }
  }
  c->ndir = k + 1;
  c->t[c->ndir] = PI_;
  c->f[c->ndir] = 0;
  c->Numfac = 8 * nrows * nrows;
  trifac(c, nrows);
  areanorm(c, c->ndir, c->Numfac);
  sphfunc(c, c->Numfac);
  ellfit(c, c->cg_first, c->a_ax, c->b_ax, c->c_ax, c->Numfac, c->Ncoef);
  c->freq_start = 1 / c->per_start;
  c->freq_end = 1 / c->per_end;
  c->freq_step = 0.5 / (c->jd_max - c->jd_min) / 24 * c->per_step_coef;
  c->ia[c->Ncoef + 1] = 1;
  c->ia[c->Ncoef + 2] = 1;
  c->ia[c->Ncoef + 3] = c->ia_prd;
  c->Nphpar = 3;
  for (i = 1; i <= c->Ncoef; i++) c->ia[i] = 1;
  if (c->onlyrel == 1) c->ia[1] = 0;
  c->ia[c->Ncoef + 3 + c->Nphpar + 1] = c->ia_cl;
  c->ia[c->Ncoef + 3 + c->Nphpar + 2] = 0;
  /* The reference sets the phase-parameter flags inside every fit
   * (period_search_BOINC.cpp:838-840, always to ia_par); setting them here
   * as well lets a single mrqcof() be evaluated before any fit has run. */
  for (i = 1; i <= c->Nphpar; i++) c->ia[c->Ncoef + 3 + i] = c->ia_par[i];
}

void orc_end_round(orc_ctx *c) { c->ndata = c->ndata_real; }

int orc_n_periods(orc_ctx *c) { return (int)((c->freq_start - c->freq_end) / c->freq_step) + 1; }

/* One fit: period_search_BOINC.cpp:819-911 (scan copy). */
void orc_fit(orc_ctx *c, double freq, int m, orc_fit_result *out) {
  int i, j, n_iter_max, n_iter_min = c->n_iter_min_in;
  const int Ncoef = c->Ncoef, Nphpar = c->Nphpar, ma = Ncoef + 5 + Nphpar;
  double prd, dev_old, dev_new, iter_diff, iter_diff_max, rchisq, totarea, sum, dark;
  double *cg = c->cg;

  prd = 1 / freq;
  for (i = 1; i <= Ncoef; i++) cg[i] = c->cg_first[i];
  cg[Ncoef + 1] = kBetaPole[m];
  cg[Ncoef + 2] = kLambdaPole[m];
  cg[Ncoef + 1] = 90 - cg[Ncoef + 1];
  cg[Ncoef + 1] = DEG2RAD * cg[Ncoef + 1];
  cg[Ncoef + 2] = DEG2RAD * cg[Ncoef + 2];
  cg[Ncoef + 3] = 24 * 2 * PI_ / prd;
  for (i = 1; i <= Nphpar; i++) {
    cg[Ncoef + 3 + i] = c->par[i];
    c->ia[Ncoef + 3 + i] = c->ia_par[i];
  }
  cg[Ncoef + 3 + Nphpar + 2] = 1;
  cg[Ncoef + 3 + Nphpar + 1] = log(c->cl);

  n_iter_max = 0;
  iter_diff_max = -1;
  rchisq = -1;
  if (c->stop_condition > 1) {
    n_iter_max = (int)c->stop_condition;
    iter_diff_max = 0;
    n_iter_min = 0;
  }
  if (c->stop_condition < 1) {
    n_iter_max = ORC_MAX_N_ITER;
    iter_diff_max = c->stop_condition;
  }
  c->Alamda = -1;
  c->Niter = 0;
  iter_diff = 1e40;
  dev_old = 1e30;
  dev_new = 0;
  c->Lastcall = 0;
  while (((c->Niter < n_iter_max) && (iter_diff > iter_diff_max)) || (c->Niter < n_iter_min)) {
    mrqmin(c, cg, c->ia, ma, c->covar, c->aalpha);
    c->Niter++;
    if ((c->Niter == 1) || (c->Chisq < c->Ochisq)) {
      c->Ochisq = c->Chisq;
      curv(c, cg);
      for (i = 1; i <= 3; i++) {
        c->chck[i] = 0;
        for (j = 1; j <= c->Numfac; j++) c->chck[i] = c->chck[i] + c->Area[j] * c->Nor[j][i];
      }
      rchisq = c->Chisq - (pow(c->chck[1], 2) + pow(c->chck[2], 2) + pow(c->chck[3], 2)) * pow(c->conw_r, 2);
    }
    dev_new = sqrt(rchisq / (c->ndata - 3));
    if (c->Niter <= ORC_MAX_N_ITER) c->dev_trace[c->Niter] = dev_new;
    if (dev_old - dev_new > 1e-10) {
      iter_diff = dev_old - dev_new;
      dev_old = dev_new;
    }
  }
  totarea = 0;
  for (i = 1; i <= c->Numfac; i++) totarea = totarea + c->Area[i];
  sum = pow(c->chck[1], 2) + pow(c->chck[2], 2) + pow(c->chck[3], 2);
  dark = sqrt(sum);
  out->dev = dev_new;
  out->chisq = c->Chisq;
  out->rchisq = rchisq;
  out->prd = 2 * PI_ / cg[Ncoef + 3];
  out->la = RAD2DEG * cg[Ncoef + 2];
  out->be = 90 - RAD2DEG * cg[Ncoef + 1];
  out->dark = dark;
  out->totarea = totarea;
  out->alamda = c->Alamda;
  out->niter = c->Niter;
}

/* One trial period: period_search_BOINC.cpp:812-924 (precalc copy :585-707).
 * row = {24*per_best, dev_best, dev_best^2*(ndata-3), dark_best, la_best, be_best}. */
void orc_period(orc_ctx *c, int n, double row[6], int precalc) {
  int m;
  double freq = c->freq_start - (n - 1) * c->freq_step;
  double per_best = 0, dark_best = 0, la_best = 0, be_best = 0, dev_best = 1e40;
  orc_fit_result r;
  for (m = 1; m <= ORC_N_POLES; m++) {
    orc_fit(c, freq, m, &r);
    if (r.dev < dev_best) {
      dev_best = r.dev;
      per_best = r.prd;
      dark_best = r.dark / r.totarea * 100;
      la_best = r.la;
      be_best = r.be;
    }
  }
  if (la_best < 0) la_best += 360;
  if (precalc && isnan(dark_best)) dark_best = 1.0;
  row[0] = 24 * per_best;
  row[1] = dev_best;
  row[2] = dev_best * dev_best * (c->ndata - 3);
  row[3] = dark_best;
  row[4] = la_best;
  row[5] = be_best;
}

/* conw optimisation loop (period_search_BOINC.cpp:487-715). */
double orc_precalc(orc_ctx *c) {
  int n, max_test_periods, n_iter;
  double sum_dark_facet, ave_dark_facet, row[6];
  double conw_r = c->conw / c->escl / c->escl;
  c->new_conw = 0;
  while ((c->new_conw != 1) && ((conw_r * c->escl * c->escl) < 10.0)) {
    orc_prepare_round(c, conw_r);
    max_test_periods = 10;
    sum_dark_facet = 0.0;
    n_iter = orc_n_periods(c);
    if (n_iter < max_test_periods) max_test_periods = n_iter;
    for (n = 1; n <= max_test_periods; n++) {
      orc_period(c, n, row, 1);
      sum_dark_facet = sum_dark_facet + row[3];
    }
    ave_dark_facet = sum_dark_facet / max_test_periods;
    if (ave_dark_facet < 1.0) c->new_conw = 1;
    if (ave_dark_facet >= 1.0) conw_r = conw_r * 2;
    orc_end_round(c);
  }
  c->conw_r = conw_r;
  return conw_r;
}

/* Scan loop and row writer (period_search_BOINC.cpp:807-931). */
int orc_scan_to_file(orc_ctx *c, int n_from, int n_to, const char *out_path, const char *mode) {
  int n, rows = 0;
  double row[6];
  FILE *out = fopen(out_path, mode);
  if (!out) return -1;
  orc_prepare_round(c, c->conw_r);
  for (n = n_from; n <= n_to; n++) {
    orc_period(c, n, row, 0);
    if (n == 1)
      fprintf(out, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", row[0], row[1], row[2], c->conw_r * c->escl * c->escl,
              row[4], row[5]);
    else
      fprintf(out, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", row[0], row[1], row[2], row[3], row[4], row[5]);
    rows++;
  }
  fclose(out);
  return rows;
}

/* ------------------------------------------------------------------ */
/* Accessors and function-level entry points for tests. */

double *orc_dptr(orc_ctx *c, const char *nm) {
  if (!strcmp(nm, "Area")) return c->Area;
  if (!strcmp(nm, "Darea")) return c->Darea;
  if (!strcmp(nm, "Fc")) return &c->Fc[0][0];
  if (!strcmp(nm, "Fs")) return &c->Fs[0][0];
  if (!strcmp(nm, "Dsph")) return &c->Dsph[0][0];
  if (!strcmp(nm, "Dg")) return &c->Dg[0][0];
  if (!strcmp(nm, "Nor")) return &c->Nor[0][0];
  if (!strcmp(nm, "Blmat")) return &c->Blmat[0][0];
  if (!strcmp(nm, "Dblm")) return &c->Dblm[0][0][0];
  if (!strcmp(nm, "Pleg")) return &c->Pleg[0][0][0];
  if (!strcmp(nm, "Weight")) return c->Weight;
  if (!strcmp(nm, "tim")) return c->tim;
  if (!strcmp(nm, "brightness")) return c->brightness;
  if (!strcmp(nm, "sig")) return c->sig;
  if (!strcmp(nm, "ee")) return &c->ee[0][0];
  if (!strcmp(nm, "ee0")) return &c->ee0[0][0];
  if (!strcmp(nm, "cg_first")) return c->cg_first;
  if (!strcmp(nm, "cg")) return c->cg;
  if (!strcmp(nm, "t")) return c->t;
  if (!strcmp(nm, "f")) return c->f;
  if (!strcmp(nm, "at")) return c->at;
  if (!strcmp(nm, "af")) return c->af;
  if (!strcmp(nm, "par")) return c->par;
  if (!strcmp(nm, "chck")) return c->chck;
  if (!strcmp(nm, "dev_trace")) return c->dev_trace;
  if (!strcmp(nm, "aalpha")) return &c->aalpha[0][0];
  if (!strcmp(nm, "covar")) return &c->covar[0][0];
  return NULL;
}

int *orc_iptr(orc_ctx *c, const char *nm) {
  if (!strcmp(nm, "Lpoints")) return c->Lpoints;
  if (!strcmp(nm, "Inrel")) return c->Inrel;
  if (!strcmp(nm, "ia")) return c->ia;
  if (!strcmp(nm, "ifp")) return &c->ifp[0][0];
  if (!strcmp(nm, "ia_par")) return c->ia_par;
  return NULL;
}

#define GETSET(X)                                  \
  X(Lmax, int) X(Mmax, int) X(Ncoef, int) X(Numfac, int) X(Lcurves, int) X(Nphpar, int) \
  X(ndata, int) X(ndata_real, int) X(nrows, int) X(ndir, int) X(onlyrel, int) X(ia_prd, int) \
  X(ia_cl, int) X(n_iter_min_in, int) X(new_conw, int) X(Niter, int) X(dyda_f32, int) \
  X(Ochisq, double) X(Chisq, double) X(Alamda, double) X(Alamda_incr, double) \
  X(Alamda_start, double) X(Phi_0, double) X(Scale, double) X(per_start, double) \
  X(per_step_coef, double) X(per_end, double) X(jd_00, double) X(conw, double) X(cl, double) \
  X(stop_condition, double) X(jd_min, double) X(jd_max, double) X(escl, double) X(a_ax, double) \
  X(b_ax, double) X(c_ax, double) X(conw_r, double) X(freq_start, double) X(freq_end, double) \
  X(freq_step, double)

double orc_get(orc_ctx *c, const char *nm) {
#define G(f, T) if (!strcmp(nm, #f)) return (double)c->f;
  GETSET(G)
#undef G
  return NAN;
}

void orc_set(orc_ctx *c, const char *nm, double v) {
#define S(f, T) if (!strcmp(nm, #f)) { c->f = (T)v; return; }
  GETSET(S)
#undef S
}

void orc_trifac(orc_ctx *c, int nrows) { trifac(c, nrows); }
void orc_areanorm(orc_ctx *c, int ndir, int nfac) { areanorm(c, ndir, nfac); }
void orc_sphfunc(orc_ctx *c, int ndir) { sphfunc(c, ndir); }
void orc_ellfit(orc_ctx *c, double *cg, double a, double b, double cc, int ndir, int ncoef) {
  ellfit(c, cg, a, b, cc, ndir, ncoef);
}
void orc_ludcmp(double *a, int n, int *indx, double *d) { ludcmp_s(a, n + 1, n, indx, d); }
void orc_lubksb(const double *a, int n, const int *indx, double *b) { lubksb_s(a, n + 1, n, indx, b); }
void orc_curv(orc_ctx *c, const double *cg) { curv(c, cg); }
void orc_blmatrix(orc_ctx *c, double bet, double lam) { blmatrix(c, bet, lam); }
void orc_matrix(orc_ctx *c, double omg, double t, double *tmat16, double *dtm64) {
  matrix(c, omg, t, (double(*)[4])tmat16, (double(*)[4][4])dtm64);
}
void orc_phasec(orc_ctx *c, double *dcdp, double alpha, const double *p) { phasec(c, dcdp, alpha, p); }
double orc_bright(orc_ctx *c, const double *ee, const double *ee0, double t, const double *cg, double *dyda,
                  int ncoef) {
  return bright(c, ee, ee0, t, cg, dyda, ncoef);
}
double orc_conv(orc_ctx *c, int nc, double *dres, int ma) { return conv(c, nc, dres, ma); }
double orc_mrqcof(orc_ctx *c, const double *cg, const int *ia, int ma, double *alpha, double *beta) {
  return mrqcof(c, cg, ia, ma, (double(*)[NP])alpha, beta);
}
int orc_gauss_errc(double *a, int n, double *b) { return gauss_errc((double(*)[NP])a, n, b); }
