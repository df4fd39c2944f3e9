// ps_host.cpp -- host side of the period-search engine: period_search_in
// reader, geometry normalisation, Gaussian-sphere shape basis and the
// period_search_out row text.  This is driver work (it runs once per
// asteroid, outside the fit loop); the fit loop itself runs only in the HIP
// kernels of ps_engine.hip.
//
// The shape basis must be bit-identical to the reference's, because every
// fit starts from cg_first and samples Dsph; the arithmetic below keeps the
// reference's operation order (period_search/trifac.c, areanorm.c,
// sphfunc.c, ellfit.c, ludcmp.c, lubksb.c), on 0-indexed storage.
// glibc's sincos() and separate sin()/cos() calls can differ in the last
// bit; the reference build fuses the pair only in the first loop of ellfit,
// so this file is compiled with -fno-builtin-sin/-cos/-sincos and calls
// sincos() there explicitly.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "periodsearch.h"
#include "ps_internal.h"

namespace {

constexpr double kPi = M_PI;
constexpr int kMaxLc = 200;        // period_search/constants.h:3
constexpr int kMaxObs = 20000;     // period_search/constants.h:2
constexpr int kPointsMax = 2000;   // period_search_cuda/period_search/constants.h (POINTS_MAX)
constexpr int kMaxFac = 1000;      // period_search/constants.h:5
constexpr int kMaxLm = 10;         // period_search/constants.h:8

// Vertex node numbering of the standard triangulation (trifac.c:16-40),
// then the triangle table (trifac.c:42-79).  ifp: [F][3], 1-based vertex ids.
void triangulate(int nrows, std::vector<int> &ifp) {
  const int cols = 4 * nrows + 1;
  std::vector<int> nod((2 * nrows + 1) * cols, 0);
  auto N = [&](int r, int c) -> int & { return nod[r * cols + c]; };
  int nnod = 1;
  N(0, 0) = nnod;
  for (int i = 1; i <= nrows; i++)
    for (int j = 0; j < 4 * i; j++) {
      N(i, j) = ++nnod;
      if (j == 0) N(i, 4 * i) = nnod;
    }
  for (int i = nrows - 1; i >= 1; i--)
    for (int j = 0; j < 4 * i; j++) {
      N(2 * nrows - i, j) = ++nnod;
      if (j == 0) N(2 * nrows - i, 4 * i) = nnod;
    }
  N(2 * nrows, 0) = nnod + 1;
  ifp.assign(8 * nrows * nrows * 3, 0);
  int t = 0;
  auto put = [&](int a, int b, int c) {
    ifp[3 * t] = a;
    ifp[3 * t + 1] = b;
    ifp[3 * t + 2] = c;
    t++;
  };
  for (int r = 1; r <= nrows; r++)
    for (int q = 0; q < 4; q++) {
      int j0 = q * r;
      put(N(r - 1, j0 - q), N(r, j0), N(r, j0 + 1));
      for (int j2 = j0 + 1; j2 <= j0 + r - 1; j2++) {
        put(N(r, j2), N(r - 1, j2 - q), N(r - 1, j2 - 1 - q));
        put(N(r - 1, j2 - q), N(r, j2), N(r, j2 + 1));
      }
    }
  for (int r = nrows + 1; r <= 2 * nrows; r++)
    for (int q = 0; q < 4; q++) {
      int j0 = q * (2 * nrows - r);
      put(N(r, j0), N(r - 1, j0 + 1 + q), N(r - 1, j0 + q));
      for (int j2 = j0 + 1; j2 <= j0 + (2 * nrows - r); j2++) {
        put(N(r, j2), N(r - 1, j2 + q), N(r, j2 - 1));
        put(N(r, j2), N(r - 1, j2 + 1 + q), N(r - 1, j2 + q));
      }
    }
}

// Vertex directions of the Gaussian sphere (period_search_BOINC.cpp:507-542).
void vertex_dirs(int nrows, std::vector<double> &th, std::vector<double> &ph) {
  const double dth = kPi / (2 * nrows);
  th.clear();
  ph.clear();
  th.push_back(0);
  ph.push_back(0);
  for (int i = 1; i <= nrows; i++) {
    double dph = kPi / (2 * i);
    for (int j = 0; j < 4 * i; j++) {
      th.push_back(i * dth);
      ph.push_back(j * dph);
    }
  }
  for (int i = nrows - 1; i >= 1; i--) {
    double dph = kPi / (2 * i);
    for (int j = 0; j < 4 * i; j++) {
      th.push_back(kPi - i * dth);
      ph.push_back(j * dph);
    }
  }
  th.push_back(kPi);
  ph.push_back(0);
}

// Facet normals, areas and normal directions (areanorm.c:17-55).
void facet_normals(const std::vector<double> &th, const std::vector<double> &ph, const std::vector<int> &ifp,
                   int nfac, double *nor, double *darea, std::vector<double> &at, std::vector<double> &af) {
  const size_t nv = th.size();
  std::vector<double> x(nv), y(nv), z(nv);
  for (size_t i = 0; i < nv; i++) {
    double st = sin(th[i]);
    x[i] = st * cos(ph[i]);
    y[i] = st * sin(ph[i]);
    z[i] = cos(th[i]);
  }
  at.resize(nfac);
  af.resize(nfac);
  for (int i = 0; i < nfac; i++) {
    int v0 = ifp[3 * i] - 1, v1 = ifp[3 * i + 1] - 1, v2 = ifp[3 * i + 2] - 1;
    double ax = x[v1] - x[v0], ay = y[v1] - y[v0], az = z[v1] - z[v0];
    double bx = x[v2] - x[v0], by = y[v2] - y[v0], bz = z[v2] - z[v0];
    double c1 = ay * bz - by * az;
    double c2 = az * bx - bz * ax;
    double c3 = ax * by - bx * ay;
    double len = sqrt(c1 * c1 + c2 * c2 + c3 * c3);
    nor[3 * i] = c1 / len;
    nor[3 * i + 1] = c2 / len;
    nor[3 * i + 2] = c3 / len;
    at[i] = acos(nor[3 * i + 2]);
    af[i] = atan2(nor[3 * i + 1], nor[3 * i]);
    darea[i] = 0.5 * len;
  }
}

// Unnormalised associated Legendre functions (no Condon-Shortley phase) and
// the Laplace-series basis Dsph at the facet normals (sphfunc.c:17-86).
// Coefficients of P_l^m as a polynomial in cos(theta) times sin^m(theta):
// P_l^m = sin^m * sum_n aleg[n][l][m] cos^n.
void legendre_basis(int nfac, int lmax, int mmax, const std::vector<double> &at, const std::vector<double> &af,
                    double *pleg, double *fc, double *fs, double *dsph, int ncoef) {
  const int L = lmax + 1, M = mmax + 1;
  std::vector<double> ts(L), tc(L);
  std::vector<double> aleg(L * L * L, 0.0);
  auto AL = [&](int n, int l, int m) -> double & { return aleg[(n * L + l) * L + m]; };
  AL(0, 0, 0) = 1;
  if (lmax >= 1) AL(1, 1, 0) = 1;
  for (int l = 1; l <= lmax; l++) AL(0, l, l) = AL(0, l - 1, l - 1) * (2 * l - 1);
  for (int m = 0; m <= mmax; m++)
    for (int l = m + 1; l <= lmax; l++) {
      int ibot;
      if ((l - m) % 2 == 0) {
        AL(0, l, m) = -(l + m - 1) * AL(0, l - 2, m) / (1 * (l - m));
        ibot = 2;
      } else {
        ibot = 1;
      }
      if (l != 1)
        for (int n = ibot; n <= l - m; n += 2)
          AL(n, l, m) = ((2 * l - 1) * AL(n - 1, l - 1, m) - (l + m - 1) * AL(n, l - 2, m)) / (1 * (l - m));
    }
  for (int i = 0; i < nfac; i++) {
    ts[0] = 1;
    tc[0] = 1;
    if (lmax >= 1) {
      ts[1] = sin(at[i]);
      tc[1] = cos(at[i]);
    }
    for (int j = 2; j <= lmax; j++) {
      ts[j] = ts[1] * ts[j - 1];
      tc[j] = tc[1] * tc[j - 1];
    }
    fs[i * M] = 0;
    fc[i * M] = 1;
    for (int j = 1; j <= mmax; j++) {
      fs[i * M + j] = sin(j * af[i]);
      fc[i * M + j] = cos(j * af[i]);
    }
    for (int q = 0; q < L * L; q++) pleg[i * L * L + q] = 0;
    int k = 0;
    for (int m = 0; m <= mmax; m++)
      for (int l = m; l <= lmax; l++) {
        double p = 0;
        for (int n = (l - m) % 2; n <= l - m; n += 2) p = p + AL(n, l, m) * tc[n] * ts[m];
        pleg[(i * L + l) * L + m] = p;
        dsph[i * ncoef + k++] = fc[i * M + m] * p;
        if (m != 0) dsph[i * ncoef + k++] = fs[i * M + m] * p;
      }
  }
}

// Crout LU with partial pivoting and its back-substitution (ludcmp.c:9-69,
// lubksb.c:3-26), 0-indexed, row-major n x n.
void lu_solve(std::vector<double> &a, int n, std::vector<double> &b) {
  std::vector<double> v(n);
  std::vector<int> indx(n);
  for (int i = 0; i < n; i++) {
    double big = 0.0;
    for (int j = 0; j < n; j++) {
      double t = fabs(a[i * n + j]);
      if (t > big) big = t;
    }
    v[i] = 1.0 / big;
  }
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < j; i++) {
      double sum = a[i * n + j];
      for (int k = 0; k < i; k++) sum -= a[i * n + k] * a[k * n + j];
      a[i * n + j] = sum;
    }
    double big = 0.0;
    int imax = -1;
    for (int i = j; i < n; i++) {
      double sum = a[i * n + j];
      for (int k = 0; k < j; k++) sum -= a[i * n + k] * a[k * n + j];
      a[i * n + j] = sum;
      double dum = v[i] * fabs(sum);
      if (dum >= big) {
        big = dum;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 0; k < n; k++) std::swap(a[imax * n + k], a[j * n + k]);
      v[imax] = v[j];
    }
    indx[j] = imax;
    if (a[j * n + j] == 0.0) a[j * n + j] = 1.0e-20;
    if (j != n - 1) {
      double dum = 1.0 / a[j * n + j];
      for (int i = j + 1; i < n; i++) a[i * n + j] *= dum;
    }
  }
  int ii = -1;
  for (int i = 0; i < n; i++) {
    int ip = indx[i];
    double sum = b[ip];
    b[ip] = b[i];
    if (ii >= 0)
      for (int j = ii; j <= i - 1; j++) sum -= a[i * n + j] * b[j];
    else if (sum)
      ii = i;
    b[i] = sum;
  }
  for (int i = n - 1; i >= 0; i--) {
    double sum = b[i];
    for (int j = i + 1; j < n; j++) sum -= a[i * n + j] * b[j];
    b[i] = sum / a[i * n + i];
  }
}

// Least-squares fit of the ellipsoid's log-curvature with the Laplace
// series (ellfit.c:24-82).
void ellipsoid_coefs(int nfac, int lmax, int mmax, int ncoef, double a, double b, double c,
                     const std::vector<double> &at, const std::vector<double> &af, const double *pleg,
                     double *cg) {
  const int L = lmax + 1;
  std::vector<double> er(nfac), fm((size_t)nfac * ncoef);
  for (int i = 0; i < nfac; i++) {
    double st, ct, sa, ca;
    sincos(at[i], &st, &ct);
    sincos(af[i], &sa, &ca);
    double u = a * st * ca, v = b * st * sa, w = c * ct;
    double sum = u * u + v * v + w * w;
    er[i] = 2 * (log(a * b * c) - log(sum));
  }
  for (int i = 0; i < nfac; i++) {
    int n = 0;
    for (int m = 0; m <= mmax; m++)
      for (int l = m; l <= lmax; l++) {
        double p = pleg[(i * L + l) * L + m];
        if (m != 0) {
          fm[(size_t)i * ncoef + n++] = p * cos(m * af[i]);
          fm[(size_t)i * ncoef + n++] = p * sin(m * af[i]);
        } else {
          fm[(size_t)i * ncoef + n++] = p;
        }
      }
  }
  std::vector<double> nm((size_t)ncoef * ncoef), rhs(ncoef);
  for (int i = 0; i < ncoef; i++) {
    for (int j = 0; j < ncoef; j++) {
      double s = 0;
      for (int k = 0; k < nfac; k++) s = s + fm[(size_t)k * ncoef + i] * fm[(size_t)k * ncoef + j];
      nm[(size_t)i * ncoef + j] = s;
    }
    double s = 0;
    for (int k = 0; k < nfac; k++) s = s + fm[(size_t)k * ncoef + i] * er[k];
    rhs[i] = s;
  }
  lu_solve(nm, ncoef, rhs);
  for (int i = 0; i < ncoef; i++) cg[i] = rhs[i];
}

int count_coefs(int lmax, int mmax) {
  int n = 0;
  for (int m = 0; m <= mmax; m++)
    for (int l = m; l <= lmax; l++) n += (m != 0) ? 2 : 1;
  return n;
}

}  // namespace

struct ps_input {
  // header (period_search_BOINC.cpp:277-303)
  double per_start = 0, per_step_coef = 0, per_end = 0, jd00 = 0, phi0_deg = 0, conw = 0;
  int ia_prd = 1, lmax = 0, mmax = 0, nrows = 0, ia_cl = 0, n_iter_min = 0;
  double par[3] = {0, 0, 0};
  int ia_par[3] = {0, 0, 0};
  double cl = 0, stop_condition = 0, alamda_incr = 0, alamda_start = 0;
  // data
  std::vector<int> lc_points, lc_inrel;
  std::vector<double> tim, bright, sig, weight, ee, ee0;
  int onlyrel = 1;
  double jd_min = 1e20, jd_max = -1e40, escl = 0, a = 0, b = 0, c = 0;
  // shape basis
  int nfac = 0, ncoef = 0;
  std::vector<double> nor, darea, pleg, fc, fs, dsph, cg_first;
  std::vector<double> beta_pole, lambda_pole;
};

void psamd_triangulate(int nrows, int *ifp) {
  std::vector<int> t;
  triangulate(nrows, t);
  std::memcpy(ifp, t.data(), t.size() * sizeof(int));
}

namespace {
int read_input(const char *path, int flags, ps_input **out);
}

extern "C" {

int ps_shape_basis(int nrows, int lmax, int mmax, double a, double b, double c, double *nor, double *darea,
                   double *pleg, double *fc, double *fs, double *dsph, double *cg_first) {
  if (nrows < 1 || lmax < 0 || mmax < 0 || mmax > lmax || lmax > kMaxLm) return PS_ERR_ARG;
  const int nfac = 8 * nrows * nrows;
  if (nfac > kMaxFac) return PS_ERR_LIMIT;
  const int ncoef = count_coefs(lmax, mmax);
  std::vector<double> th, ph, at, af;
  std::vector<int> ifp;
  vertex_dirs(nrows, th, ph);
  triangulate(nrows, ifp);
  facet_normals(th, ph, ifp, nfac, nor, darea, at, af);
  legendre_basis(nfac, lmax, mmax, at, af, pleg, fc, fs, dsph, ncoef);
  ellipsoid_coefs(nfac, lmax, mmax, ncoef, a, b, c, at, af, pleg, cg_first);
  return nfac;
}

int ps_input_read(const char *path, ps_input **out) { return read_input(path, 0, out); }

int ps_input_read_deferred(const char *path, ps_input **out) { return read_input(path, PS_READ_DEFER_SHAPE, out); }

int ps_input_read_ex(const char *path, int flags, ps_input **out) {
  if (flags & ~(PS_READ_DEFER_SHAPE | PS_READ_CLAMP_MIN)) return PS_ERR_ARG;
  return read_input(path, flags, out);
}

int ps_input_shape_job(ps_input *in, ps_shape_job *job) {
  if (!in || !job) return PS_ERR_ARG;
  job->nrows = in->nrows;
  job->lmax = in->lmax;
  job->mmax = in->mmax;
  job->a = in->a;
  job->b = in->b;
  job->c = in->c;
  job->nor = in->nor.data();
  job->darea = in->darea.data();
  job->pleg = in->pleg.data();
  job->fc = in->fc.data();
  job->fs = in->fs.data();
  job->dsph = in->dsph.data();
  job->cg_first = in->cg_first.data();
  return PS_OK;
}

}  // extern "C"

namespace {

// period_search_BOINC.cpp:277-485.  Weights use the checked loop of
// period_search_optimization_cpu/period_search/period_search_BOINC.cpp:440-450.
// With PS_READ_DEFER_SHAPE the shape tables are sized and left zero, for
// ps_shape_basis_many to fill.  PS_READ_CLAMP_MIN raises each time and
// brightness below DBL_MIN to DBL_MIN as it is read, before the JD range is
// taken (period_search_optimization_cpu/period_search/period_search_BOINC.cpp:371-380).
int read_input(const char *path, int flags, ps_input **out) {
  if (!path || !out) return PS_ERR_ARG;
  const bool shape = !(flags & PS_READ_DEFER_SHAPE), clamp = (flags & PS_READ_CLAMP_MIN) != 0;
  FILE *f = fopen(path, "r");
  if (!f) return PS_ERR_IO;
  auto *in = new ps_input();
  char line[1024];
  auto skip = [&]() { (void)!fgets(line, sizeof line, f); };
  int ok = 1, nlc = 0;
  ok &= fscanf(f, "%lf %lf %lf %d", &in->per_start, &in->per_step_coef, &in->per_end, &in->ia_prd) == 4;
  skip();
  ok &= fscanf(f, "%lf", &in->jd00) == 1;
  skip();
  ok &= fscanf(f, "%lf", &in->phi0_deg) == 1;
  skip();
  ok &= fscanf(f, "%lf", &in->conw) == 1;
  skip();
  ok &= fscanf(f, "%d %d", &in->lmax, &in->mmax) == 2;
  skip();
  ok &= fscanf(f, "%d", &in->nrows) == 1;
  skip();
  for (int i = 0; i < 3; i++) {
    ok &= fscanf(f, "%lf %d", &in->par[i], &in->ia_par[i]) == 2;
    skip();
  }
  ok &= fscanf(f, "%lf %d", &in->cl, &in->ia_cl) == 2;
  skip();
  ok &= fscanf(f, "%lf", &in->stop_condition) == 1;
  skip();
  ok &= fscanf(f, "%d", &in->n_iter_min) == 1;
  skip();
  ok &= fscanf(f, "%lf", &in->alamda_incr) == 1;
  skip();
  ok &= fscanf(f, "%lf", &in->alamda_start) == 1;
  skip();
  ok &= fscanf(f, "%d", &nlc) == 1;
  if (!ok || nlc < 1 || nlc > kMaxLc || in->lmax > kMaxLm || in->mmax > in->lmax) {
    fclose(f);
    delete in;
    return PS_ERR_IO;
  }
  double al0 = kPi, al0_abs = kPi;
  int ial0 = -1, ial0_abs = -1;
  for (int i = 0; i < nlc; i++) {
    int np = 0, flag = 0;
    if (fscanf(f, "%d %d", &np, &flag) != 2 || np < 1 || np > kPointsMax ||
        (int)in->tim.size() + np > kMaxObs) {
      fclose(f);
      delete in;
      return PS_ERR_IO;
    }
    skip();
    in->lc_points.push_back(np);
    in->lc_inrel.push_back(1 - flag);
    if (flag != 0) in->onlyrel = 0;
    double ave = 0;
    for (int j = 0; j < np; j++) {
      double t, y, e0[3], e[3];
      if (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf", &t, &y, &e0[0], &e0[1], &e0[2], &e[0], &e[1], &e[2]) != 8) {
        fclose(f);
        delete in;
        return PS_ERR_IO;
      }
      if (clamp) {
        if (t < DBL_MIN) t = DBL_MIN;
        if (y < DBL_MIN) y = DBL_MIN;
      }
      if (t < in->jd_min) in->jd_min = t;
      if (t > in->jd_max) in->jd_max = t;
      double e0len = sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
      double elen = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      ave += y;
      in->tim.push_back(t);
      in->bright.push_back(y);
      for (int k = 0; k < 3; k++) {
        in->ee.push_back(e[k] / elen);
        in->ee0.push_back(e0[k] / e0len);
      }
      if (j == 0) {
        double cos_alpha = (e[0] * e0[0] + e[1] * e0[1] + e[2] * e0[2]) / (elen * e0len);
        double al = acos(cos_alpha);
        int idx = (int)in->tim.size() - 1;
        if (al < al0) {
          al0 = al;
          ial0 = idx;
        }
        if (al < al0_abs && flag != 0) {
          al0_abs = al;
          ial0_abs = idx;
        }
      }
    }
    ave /= np;
    for (int j = 0; j < np; j++) in->sig.push_back(ave);
  }
  std::vector<double> wlc(nlc, -1.0);
  int lcn;
  double w;
  while (fscanf(f, "%d", &lcn) == 1 && fscanf(f, "%lf", &w) == 1)
    if (lcn >= 1 && lcn <= nlc) wlc[lcn - 1] = w;
  fclose(f);

  double jd0 = in->jd00;
  if (jd0 <= 0) jd0 = (int)in->jd_min;
  for (double &t : in->tim) t = t - jd0;
  for (int i = 0; i < nlc; i++)
    for (int j = 0; j < in->lc_points[i]; j++) in->weight.push_back(wlc[i] == -1 ? 1.0 : wlc[i]);
  if (in->onlyrel == 0) {
    al0 = al0_abs;
    ial0 = ial0_abs;
  }
  const double a0 = 1.05, b0 = 1.00, c0 = 0.95;
  double rfit = sqrt(2 * in->sig[ial0] / (0.5 * kPi * (1 + cos(al0))));
  in->escl = rfit / sqrt((a0 * b0 + b0 * c0 + a0 * c0) / 3);
  if (in->onlyrel == 0) in->escl *= 0.8;
  in->a = a0 * in->escl;
  in->b = b0 * in->escl;
  in->c = c0 * in->escl;

  in->nfac = 8 * in->nrows * in->nrows;
  if (in->nfac > kMaxFac) {
    delete in;
    return PS_ERR_LIMIT;
  }
  in->ncoef = count_coefs(in->lmax, in->mmax);
  const int L = in->lmax + 1, M = in->mmax + 1;
  in->nor.resize(in->nfac * 3);
  in->darea.resize(in->nfac);
  in->pleg.resize((size_t)in->nfac * L * L);
  in->fc.resize(in->nfac * M);
  in->fs.resize(in->nfac * M);
  in->dsph.resize((size_t)in->nfac * in->ncoef);
  in->cg_first.resize(in->ncoef);
  if (shape)
    ps_shape_basis(in->nrows, in->lmax, in->mmax, in->a, in->b, in->c, in->nor.data(), in->darea.data(),
                   in->pleg.data(), in->fc.data(), in->fs.data(), in->dsph.data(), in->cg_first.data());
  // initial poles (period_search_BOINC.cpp:197-206)
  in->beta_pole = {0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
  in->lambda_pole = {0, 90, 180, 270, 60, 180, 300, 60, 180, 300};
  *out = in;
  return PS_OK;
}

}  // namespace

extern "C" {

void ps_input_free(ps_input *in) { delete in; }

int ps_input_problem(ps_input *in, ps_problem *p) {
  if (!in || !p) return PS_ERR_ARG;
  std::memset(p, 0, sizeof *p);
  p->n_points = (int)in->tim.size();
  p->n_lc = (int)in->lc_points.size();
  p->lc_points = in->lc_points.data();
  p->lc_inrel = in->lc_inrel.data();
  p->tim = in->tim.data();
  p->brightness = in->bright.data();
  p->sig = in->sig.data();
  p->weight = in->weight.data();
  p->ee = in->ee.data();
  p->ee0 = in->ee0.data();
  p->lmax = in->lmax;
  p->mmax = in->mmax;
  p->n_fac = in->nfac;
  p->nor = in->nor.data();
  p->darea = in->darea.data();
  p->pleg = in->pleg.data();
  p->fc = in->fc.data();
  p->fs = in->fs.data();
  p->dsph = in->dsph.data();
  p->cg_first = in->cg_first.data();
  p->phi0 = in->phi0_deg * (kPi / 180);
  for (int i = 0; i < 3; i++) {
    p->par[i] = in->par[i];
    p->ia_par[i] = in->ia_par[i];
  }
  p->cl = in->cl;
  p->ia_cl = in->ia_cl;
  p->ia_prd = in->ia_prd;
  p->onlyrel = in->onlyrel;
  p->alamda_start = in->alamda_start;
  p->alamda_incr = in->alamda_incr;
  p->stop_condition = in->stop_condition;
  p->n_iter_min = in->n_iter_min;
  p->n_poles = (int)in->beta_pole.size();
  p->beta_pole = in->beta_pole.data();
  p->lambda_pole = in->lambda_pole.data();
  return PS_OK;
}

int ps_input_scalars(const ps_input *in, double *o, int *ndata) {
  if (!in || !o) return PS_ERR_ARG;
  o[0] = in->per_start;
  o[1] = in->per_end;
  o[2] = in->per_step_coef;
  o[3] = in->jd_min;
  o[4] = in->jd_max;
  o[5] = in->escl;
  o[6] = in->conw;
  o[7] = 1 / in->per_start;
  o[8] = 1 / in->per_end;
  o[9] = 0.5 / (in->jd_max - in->jd_min) / 24 * in->per_step_coef;
  if (ndata) *ndata = (int)in->tim.size();
  return PS_OK;
}

int ps_best_row(const ps_row *rows, long long n, long long *index) {
  if (!rows || !index || n < 1) return PS_ERR_ARG;
  long long best = -1;
  for (long long i = 0; i < n; i++) {
    if (std::isnan(rows[i].chisq)) continue;
    if (best < 0 || rows[i].chisq < rows[best].chisq) best = i;
  }
  if (best < 0) return PS_ERR_ARG;
  *index = best;
  return PS_OK;
}

int ps_read_rows(const char *path, ps_row *rows, long long cap, long long *n) {
  if (!path || !n || (cap > 0 && !rows)) return PS_ERR_ARG;
  *n = 0;
  FILE *f = fopen(path, "r");
  if (!f) return PS_ERR_IO;
  char line[512];
  long long k = 0;
  int rc = PS_OK;
  while (fgets(line, sizeof line, f)) {
    ps_row r;
    char extra;
    int got = sscanf(line, "%lf %lf %lf %lf %lf %lf %c", &r.period_h, &r.rms, &r.chisq, &r.dark, &r.lambda,
                     &r.beta, &extra);
    if (got == EOF || got <= 0) {  // blank line
      bool blank = true;
      for (const char *c = line; *c; c++) blank = blank && (*c == ' ' || *c == '\t' || *c == '\n' || *c == '\r');
      if (blank) continue;
    }
    if (got != 6) {
      rc = PS_ERR_IO;
      break;
    }
    if (k >= cap) {
      rc = PS_ERR_LIMIT;
      break;
    }
    rows[k++] = r;
  }
  fclose(f);
  *n = k;
  return rc;
}

int ps_format_row(const ps_row *r, int round_pole, char *buf, int buflen) {
  if (!r || !buf) return PS_ERR_ARG;
  double la = round_pole ? std::round(r->lambda) : r->lambda;
  double be = round_pole ? std::round(r->beta) : r->beta;
  return snprintf(buf, buflen, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", r->period_h, r->rms, r->chisq, r->dark, la,
                  be);
}

// Checkpoint files.  Each reference app writes one format into a temp file
// and renames it over period_search_state, and on start accepts only its own
// field count (anything else starts afresh with the output truncated):
//   base  "%d"                   period_search/period_search_BOINC.cpp:85-101, 237-251
//   CUDA  "%d %d %.17g"          period_search_cuda/period_search/period_search_BOINC.cpp:96-112, 287-300
//   OC    "%d %d %.17g %.17g %d" period_search_optimization_cpu/period_search/period_search_BOINC.cpp:86-101, 238-253
// ps_state_read accepts all three: the fields read decide the format.
int ps_state_read(const char *path, ps_state *st) {
  if (!path || !st) return PS_ERR_ARG;
  *st = ps_state{};
  FILE *f = fopen(path, "r");
  if (!f) return PS_OK;  // no checkpoint: a fresh start
  int n = fscanf(f, "%d %d %lf %lf %d", &st->nlines, &st->new_conw, &st->conw_r, &st->sum_dark, &st->test_periods);
  fclose(f);
  // fields beyond the first complete format are ignored; 2 or 4 fields read
  // as the shorter format they contain
  st->format = n >= 5 ? PS_STATE_OC : (n >= 3 ? PS_STATE_CUDA : (n >= 1 ? PS_STATE_BASE : PS_STATE_NONE));
  if (st->format == PS_STATE_NONE || st->nlines < 0 ||
      (st->format >= PS_STATE_CUDA && !(st->conw_r > 0 && std::isfinite(st->conw_r))) ||
      (st->format == PS_STATE_OC && (st->test_periods < 0 || !std::isfinite(st->sum_dark)))) {
    *st = ps_state{};  // unreadable or inconsistent: a fresh start
    return PS_OK;
  }
  if (st->format < PS_STATE_OC) {
    st->sum_dark = 0;
    st->test_periods = 0;
  }
  if (st->format < PS_STATE_CUDA) {
    st->new_conw = 0;
    st->conw_r = 0;
  }
  return PS_OK;
}

int ps_state_write(const char *path, const ps_state *st) {
  if (!path || !st || st->format < PS_STATE_BASE || st->format > PS_STATE_OC) return PS_ERR_ARG;
  const std::string tmp = std::string(path) + ".temp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return PS_ERR_IO;
  int w;
  if (st->format == PS_STATE_BASE)
    w = fprintf(f, "%d", st->nlines);
  else if (st->format == PS_STATE_CUDA)
    w = fprintf(f, "%d %d %.17g", st->nlines, st->new_conw, st->conw_r);
  else
    w = fprintf(f, "%d %d %.17g %.17g %d", st->nlines, st->new_conw, st->conw_r, st->sum_dark, st->test_periods);
  if (fclose(f) != 0 || w < 0) return PS_ERR_IO;
  return rename(tmp.c_str(), path) == 0 ? PS_OK : PS_ERR_IO;
}

}  // extern "C"
