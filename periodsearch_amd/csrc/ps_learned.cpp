// ps_learned.cpp -- the learned scan: trial periods started from a
// network's proposal, screened, and the best candidates judged by the plain
// ten-pole fits.  Host orchestration only: the fits, the light-curve
// features and the network run in the engine's and the network's HIP kernels
// (ps_engine.hip, ps_mlp.hip) through their C ABI.
//
// The reference starts every trial period from the same ten poles and the
// ellipsoid shape (period_search/period_search_BOINC.cpp:819-844) and keeps
// the best of the ten fits (:900-924).  Here, for every trial period:
//   1. features: the phase-folded light curve at the period (ps_lc_features)
//      and the asteroid's amplitude-aspect features (ps_aspect_features);
//   2. the network maps them to a pole (doubled-longitude unit vector) and
//      the l <= 2 shape coefficients; one start (and optionally its mirror
//      pole, lambda + 180 deg with the m = 1 coefficients negated) is fitted
//      by the full LM loop under the problem's stop rule;
//   3. per block of trial periods, the n_candidates periods of lowest rms
//      (no two within halo of each other) and halo periods on either side of
//      each are refitted from the reference's ten poles by the judge engine
//      (any precision; FP64 makes the judged rows the plain FP64 rows), while
//      the next blocks' proposed fits run on the screening engine (the
//      blocks' judge batches queue behind each other on one thread, so the
//      screening never waits for the judge);
//   4. the best judged row is walked to a local minimum of the judged chi2
//      profile (its unjudged neighbours judged, a few at a time).
// Judged rows replace the screened rows and are flagged; only they are
// results in the reference's sense.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <string>
#include <thread>
#include <vector>

#include "periodsearch.h"
#include "ps_internal.h"
#include "ps_proposal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

int count_coefs(int lmax, int mmax) {
  int n = 0;
  for (int m = 0; m <= mmax; m++) n += (lmax - m + 1) * (m ? 2 : 1);
  return n;
}

// Positions of (l, m, cos/sin) with l <= 2 in the coefficient vector (for
// m = 0..mmax, l = m..lmax: one entry for m = 0, a cos/sin pair otherwise),
// in the order c00, c10, c20, c11c, c11s, c21c, c21s, c22c, c22s; m1 marks
// the m = 1 entries.  Needs lmax, mmax >= 2.
void low_order_index(int lmax, int mmax, int idx[9], bool m1[9]) {
  const int want[9][3] = {{0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {1, 1, 0}, {1, 1, 1},
                          {2, 1, 0}, {2, 1, 1}, {2, 2, 0}, {2, 2, 1}};
  for (int w = 0; w < 9; w++) {
    int k = 0;
    for (int m = 0; m <= mmax; m++)
      for (int l = m; l <= lmax; l++) {
        const int parts = m ? 2 : 1;
        for (int s = 0; s < parts; s++, k++)
          if (l == want[w][0] && m == want[w][1] && s == want[w][2]) idx[w] = k;
      }
    m1[w] = want[w][1] == 1;
  }
}

// The reference's pole reduction (period_search_BOINC.cpp:913-924) over
// per fits per trial period: the first fit whose dev is strictly below the
// best so far (from 1e40) wins, so a NaN never does; lambda wrapped to
// [0, 360); chi2 = dev^2 (ndata - 3).
void rows_from_fits(const ps_fit_out *f, int n_periods, int per, int n_points, ps_row *rows) {
  for (int i = 0; i < n_periods; i++) {
    double dev_best = 1e40, per_best = 0, dark_best = 0, la_best = 0, be_best = 0;
    for (int m = 0; m < per; m++) {
      const ps_fit_out &o = f[(size_t)i * per + m];
      if (o.dev < dev_best) {
        dev_best = o.dev;
        per_best = o.prd;
        dark_best = o.dark / o.totarea * 100;
        la_best = o.la;
        be_best = o.be;
      }
    }
    if (la_best < 0) la_best += 360;
    ps_row &r = rows[i];
    r.period_h = 24 * per_best;
    r.rms = dev_best;
    r.chisq = dev_best * dev_best * n_points;
    r.dark = dark_best;
    r.lambda = la_best;
    r.beta = be_best;
  }
}

// Positions (0-based within the block) of the n_candidates lowest-rms rows,
// no two within halo of each other, each widened by halo on either side.
std::vector<int> pick_candidates(const ps_row *rows, int n, int n_candidates, int halo) {
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  // NaN rms sort last
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const double ra = rows[a].rms, rb = rows[b].rms;
    if (std::isnan(ra)) return false;
    if (std::isnan(rb)) return true;
    return ra < rb;
  });
  std::vector<int> picked;
  for (int k : order) {
    if ((int)picked.size() >= n_candidates) break;
    bool near = false;
    for (int q : picked) near = near || std::abs(k - q) <= halo;
    if (!near) picked.push_back(k);
  }
  std::vector<char> mark(n, 0);
  for (int k : picked)
    for (int j = std::max(0, k - halo); j <= std::min(n - 1, k + halo); j++) mark[j] = 1;
  std::vector<int> out;
  for (int i = 0; i < n; i++)
    if (mark[i]) out.push_back(i);
  return out;
}

// ---- network file: uncompressed .npz of 1-D float32 .npy members --------
// A zip local header (signature 50 4b 03 04) is followed by the member's
// name, an extra field (zip64 sizes under tag 0x0001 when the 32-bit sizes
// read 0xffffffff) and the stored bytes; an .npy member is the magic
// \x93NUMPY, a version, a header length (2 bytes in version 1, 4 after) and
// a dict text with 'descr', 'fortran_order' and 'shape'.
bool read_file(const char *path, std::vector<unsigned char> &buf) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  unsigned char tmp[1 << 16];
  size_t n;
  while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
  const bool ok = !ferror(f);
  fclose(f);
  return ok;
}

unsigned long long le(const unsigned char *p, int n) {
  unsigned long long v = 0;
  for (int i = n - 1; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}

// The float32 values of a 1-D little-endian .npy member; false if not that.
bool npy_f32(const unsigned char *d, size_t size, std::vector<float> &out) {
  if (size < 10 || memcmp(d, "\x93NUMPY", 6) != 0) return false;
  const int major = d[6];
  const size_t hl = major == 1 ? le(d + 8, 2) : le(d + 8, 4);
  const size_t h0 = major == 1 ? 10 : 12;
  if (h0 + hl > size) return false;
  const std::string h((const char *)d + h0, hl);
  if (h.find("'descr': '<f4'") == std::string::npos) return false;
  const size_t sp = h.find("'shape': (");
  if (sp == std::string::npos) return false;
  const size_t close = h.find(')', sp);
  const std::string dims = h.substr(sp + 10, close - sp - 10);
  char *end = nullptr;
  const long long n = strtoll(dims.c_str(), &end, 10);
  for (const char *c = end; *c; c++)  // exactly one dimension: "N," or "N"
    if (*c != ',' && *c != ' ') return false;
  if (n < 0 || h0 + hl + (size_t)n * 4 != size) return false;
  out.resize(n);
  memcpy(out.data(), d + h0 + hl, (size_t)n * 4);
  return true;
}

bool npz_members(const std::vector<unsigned char> &z, std::vector<std::pair<std::string, std::vector<float>>> &m) {
  size_t off = 0;
  while (off + 30 <= z.size() && le(&z[off], 4) == 0x04034b50) {
    const unsigned char *h = &z[off];
    const int flags = (int)le(h + 6, 2), method = (int)le(h + 8, 2);
    unsigned long long csize = le(h + 18, 4), usize = le(h + 22, 4);
    const size_t nl = le(h + 26, 2), el = le(h + 28, 2);
    if (off + 30 + nl + el > z.size()) return false;
    std::string name((const char *)h + 30, nl);
    const unsigned char *ex = h + 30 + nl;
    for (size_t q = 0; q + 4 <= el;) {
      const size_t tag = le(ex + q, 2), len = le(ex + q + 2, 2);
      if (tag == 1 && len >= 16) {
        if (usize == 0xffffffffULL) usize = le(ex + q + 4, 8);
        if (csize == 0xffffffffULL) csize = le(ex + q + 12, 8);
      }
      q += 4 + len;
    }
    if (method != 0 || (flags & 8) || csize != usize) return false;  // stored members only
    const size_t d0 = off + 30 + nl + el;
    if (d0 + csize > z.size()) return false;
    if (name.size() > 4 && name.compare(name.size() - 4, 4, ".npy") == 0) name.resize(name.size() - 4);
    std::vector<float> v;
    if (npy_f32(&z[d0], csize, v)) m.emplace_back(name, std::move(v));
    off = d0 + csize;
  }
  return !m.empty();
}

}  // namespace

extern "C" {

int ps_rows_from_fits(const ps_fit_out *fits, int n_periods, int per, int n_points, ps_row *rows) {
  if (!fits || !rows || n_periods < 0 || per < 1 || n_points < 1) return PS_ERR_ARG;
  rows_from_fits(fits, n_periods, per, n_points, rows);
  return PS_OK;
}

int ps_aspect_features(const ps_problem *p, float out[5]) {
  if (!p || !out || p->n_points < 1 || p->n_lc < 1 || !p->tim || !p->brightness || !p->ee || !p->lc_points)
    return PS_ERR_ARG;
  std::vector<double> amp, lon;
  int q = 0;
  for (int i = 0; i < p->n_lc; i++) {
    const int k = p->lc_points[i];
    const int q0 = q;
    q += k;
    if (k < 10) continue;
    double tmin = p->tim[q0], tmax = tmin, ymin = p->brightness[q0], ymax = ymin, ysum = 0, e0 = 0, e1 = 0;
    for (int j = q0; j < q; j++) {
      tmin = std::min(tmin, p->tim[j]);
      tmax = std::max(tmax, p->tim[j]);
      ymin = std::min(ymin, p->brightness[j]);
      ymax = std::max(ymax, p->brightness[j]);
      ysum += p->brightness[j];
      e0 += p->ee[3 * j];
      e1 += p->ee[3 * j + 1];
    }
    if (tmax - tmin > 1.0) continue;
    const double m = ysum / k;
    if (!(m > 0)) continue;
    amp.push_back((ymax - ymin) / m);
    lon.push_back(std::atan2(e1 / k, e0 / k));
  }
  for (int i = 0; i < 5; i++) out[i] = 0.0f;
  const size_t n = amp.size();
  if (n >= 2) {
    double am = 0;
    for (double a : amp) am += a;
    am /= n;
    double c = 0, s = 0;
    for (size_t i = 0; i < n; i++) {
      c += (amp[i] - am) * std::cos(2 * lon[i]);
      s += (amp[i] - am) * std::sin(2 * lon[i]);
    }
    c /= n * am;
    s /= n * am;
    out[0] = (float)c;
    out[1] = (float)s;
    out[2] = (float)std::hypot(c, s);
    out[3] = (float)am;
    out[4] = (float)(n / 10.0);
  }
  return PS_OK;
}

int ps_proposal_open(const char *path, int device, int max_batch, ps_proposal_net *net) {
  if (!path || !net || max_batch < 1) return PS_ERR_ARG;
  memset(net, 0, sizeof *net);
  std::vector<unsigned char> z;
  std::vector<std::pair<std::string, std::vector<float>>> m;
  if (!read_file(path, z) || !npz_members(z, m)) return PS_ERR_IO;
  auto get = [&](const char *k) -> const std::vector<float> * {
    for (auto &kv : m)
      if (kv.first == k) return &kv.second;
    return nullptr;
  };
  const std::vector<float> *par = get("params"), *xm = get("x_mean"), *xs = get("x_std"), *tm = get("t_mean"),
                           *ts = get("t_std");
  if (!par || !xm || !xs || !tm || !ts || xm->size() != xs->size() || tm->size() != ts->size()) return PS_ERR_IO;
  const long long ni = (long long)xm->size(), no = (long long)tm->size(), np = (long long)par->size();
  if (ni <= 7 || no != 12) return PS_ERR_IO;
  // (ni + 1) H + (H + 1) H + (H + 1) no = np: H^2 + (ni + no + 2) H + no - np = 0
  const double b = (double)(ni + no + 2), c = (double)(no - np);
  const long long hid = llround((-b + std::sqrt(b * b - 4 * c)) / 2);
  if (hid < 1 || (ni + 1) * hid + (hid + 1) * hid + (hid + 1) * no != np) return PS_ERR_IO;
  ps_mlp *mlp = nullptr;
  int rc = ps_mlp_create(device, (int)ni, (int)hid, (int)no, max_batch, &mlp);
  if (rc != PS_OK) return rc;
  if ((rc = ps_mlp_set_params(mlp, par->data(), np)) != PS_OK) {
    ps_mlp_destroy(mlp);
    return rc;
  }
  float *sc = (float *)malloc(sizeof(float) * 2 * (ni + no));
  if (!sc) {
    ps_mlp_destroy(mlp);
    return PS_ERR_ARG;
  }
  memcpy(sc, xm->data(), ni * 4);
  memcpy(sc + ni, xs->data(), ni * 4);
  memcpy(sc + 2 * ni, tm->data(), no * 4);
  memcpy(sc + 2 * ni + no, ts->data(), no * 4);
  net->mlp = mlp;
  net->n_in = (int)ni;
  net->n_out = (int)no;
  net->nbins = (int)ni - 7;
  net->max_batch = max_batch;
  net->x_mean = sc;
  net->x_std = sc + ni;
  net->t_mean = sc + 2 * ni;
  net->t_std = sc + 2 * ni + no;
  return PS_OK;
}

void ps_proposal_close(ps_proposal_net *net) {
  if (!net) return;
  if (net->mlp) ps_mlp_destroy(net->mlp);
  free((void *)net->x_mean);  // the block that holds all four scalings
  memset(net, 0, sizeof *net);
}

int ps_scan_learned(ps_engine *e, const ps_problem *p, const ps_proposal_net *net, const ps_learned_opts *opt,
                    double freq_start, double freq_step, int n_from, int n_to, double conw_r, ps_row *rows,
                    int *judged, ps_row *screened, long long *n_fits) {
  if (!e || !p || !net || !net->mlp || !opt || !rows || !judged || n_from < 1 || n_to < n_from || opt->block < 1 ||
      opt->n_candidates < 0 || opt->halo < 0 || opt->refine < 0 || net->nbins < 1 || net->nbins > 256 || net->max_batch < 1 || !net->x_mean ||
      !net->x_std || !net->t_mean || !net->t_std)
    return PS_ERR_ARG;
  if (p->lmax < 2 || p->mmax < 2) {
    psamd_engine_error(e, "the learned start needs lmax, mmax >= 2");
    return PS_ERR_ARG;
  }
  int nc = 0, ma = 0, mfit = 0;
  int rc = ps_dims(e, &nc, &ma, &mfit);
  if (rc) return rc;
  if (nc != count_coefs(p->lmax, p->mmax)) {
    psamd_engine_error(e, "ps_scan_learned: the problem is not the engine's");
    return PS_ERR_ARG;
  }
  const int n_in = net->nbins + 2 + 5, n_out = 12;
  if (net->n_in != n_in || net->n_out != n_out) {
    psamd_engine_error(e, "ps_scan_learned: the network's shape does not match the features");
    return PS_ERR_ARG;
  }
  ps_engine *judge = opt->judge ? opt->judge : e;
  const int per = opt->mirror ? 2 : 1;
  const int npol = p->n_poles;
  int idx[9];
  bool m1[9];
  low_order_index(p->lmax, p->mmax, idx, m1);
  float aspect[5];
  if ((rc = ps_aspect_features(p, aspect))) return rc;
  const int N = n_to - n_from + 1;
  std::vector<double> freq(N);
  for (int i = 0; i < N; i++) freq[i] = freq_start - (n_from + i - 1) * freq_step;  // period_search_BOINC.cpp:812
  long long fits_run = 0;
  for (int i = 0; i < N; i++) judged[i] = 0;

  // judge batches.  When the judge is another engine each batch runs on a
  // thread of its own that first joins the previous batch's thread: the
  // batches run one after another on the judge engine, overlapped with the
  // screening fits of the following blocks, and the main thread never waits
  // for them until the screening is done.  Their rows are applied then, in
  // order.
  struct Judge {
    std::vector<int> pos;  // positions in [0, N)
    std::vector<double> f;
    std::vector<int> pole;
    std::vector<ps_fit_out> out;
    int rc = PS_OK;
    std::thread th;
  };
  std::list<Judge> jobs;  // stable addresses for the threads
  auto start_judge = [&](std::vector<int> pos) {
    Judge *prev = jobs.empty() ? nullptr : &jobs.back();
    jobs.emplace_back();
    Judge &J = jobs.back();
    J.pos = std::move(pos);
    for (int i : J.pos)
      for (int m = 0; m < npol; m++) {
        J.f.push_back(freq[i]);
        J.pole.push_back(m);
      }
    J.out.assign(J.f.size(), ps_fit_out{});
    auto run = [&J, judge, conw_r]() {
      J.rc = J.f.empty() ? PS_OK : ps_fit_batch(judge, J.f.data(), J.pole.data(), (int)J.f.size(), conw_r, J.out.data());
    };
    if (judge != e)
      J.th = std::thread([prev, run]() {
        if (prev && prev->th.joinable()) prev->th.join();
        run();
      });
    else
      run();
  };
  // waits for every started batch (the last thread joins the others) and
  // applies their rows; the first failure is returned
  auto finish_judges = [&]() -> int {
    if (!jobs.empty() && jobs.back().th.joinable()) jobs.back().th.join();
    int r = PS_OK;
    for (Judge &J : jobs) {
      if (J.rc) {
        if (!r) r = J.rc;
        continue;
      }
      fits_run += (long long)J.out.size();
      std::vector<ps_row> rr(J.pos.size());
      rows_from_fits(J.out.data(), (int)J.pos.size(), npol, p->n_points, rr.data());
      for (size_t k = 0; k < J.pos.size(); k++) {
        rows[J.pos[k]] = rr[k];
        judged[J.pos[k]] = 1;
      }
    }
    jobs.clear();
    return r;
  };

  std::vector<float> feat, x, y;
  std::vector<double> cg, cg_next;
  std::vector<ps_fit_out> out;
  std::vector<ps_row> brow;
  const double *cgf = p->cg_first;
  // the starts of the block [b0, b0 + nb): features, the network, and the
  // reference's start with the proposed pole and l <= 2 shape
  auto prepare = [&](int b0, int nb, std::vector<double> &cgb) -> int {
    // 1. features and the network
    feat.assign((size_t)nb * (net->nbins + 2), 0.0f);
    int r = ps_lc_features(e, freq.data() + b0, nb, net->nbins, feat.data());
    if (r) return r;
    x.assign((size_t)nb * n_in, 0.0f);
    for (int i = 0; i < nb; i++) {
      float *xi = x.data() + (size_t)i * n_in;
      for (int k = 0; k < net->nbins + 2; k++) xi[k] = feat[(size_t)i * (net->nbins + 2) + k];
      for (int k = 0; k < 5; k++) xi[net->nbins + 2 + k] = aspect[k];
      for (int k = 0; k < n_in; k++) xi[k] = (xi[k] - net->x_mean[k]) / net->x_std[k];
    }
    y.assign((size_t)nb * n_out, 0.0f);
    for (int i0 = 0; i0 < nb; i0 += net->max_batch) {
      const int nn = std::min(net->max_batch, nb - i0);
      if ((r = ps_mlp_forward(net->mlp, x.data() + (size_t)i0 * n_in, nn, y.data() + (size_t)i0 * n_out))) {
        psamd_engine_error(e, std::string("network: ") + ps_mlp_last_error(net->mlp));
        return r;
      }
    }
    // 2. starts: the reference's start (period_search_BOINC.cpp:819-844)
    //    with the proposed pole and l <= 2 coefficients (c00, the size, kept)
    cgb.assign((size_t)nb * per * ma, 0.0);
    for (int i = 0; i < nb; i++) {
      double yv[12];
      for (int k = 0; k < n_out; k++)
        yv[k] = (double)(y[(size_t)i * n_out + k] * net->t_std[k] + net->t_mean[k]);
      const double vn = std::max(std::sqrt(yv[0] * yv[0] + yv[1] * yv[1] + yv[2] * yv[2]), 1e-12);
      const double beta = std::asin(std::min(1.0, std::max(-1.0, yv[2] / vn)));
      const double lam = 0.5 * std::atan2(yv[1], yv[0]);
      for (int k = 0; k < per; k++) {
        double *a = cgb.data() + ((size_t)i * per + k) * ma;
        for (int c = 0; c < nc; c++) a[c] = cgf[c];
        for (int w = 1; w < 9; w++) a[idx[w]] = (k == 1 && m1[w]) ? -yv[3 + w] : yv[3 + w];
        a[nc] = kPi / 2 - beta;
        a[nc + 1] = std::fmod(lam + k * kPi, 2 * kPi);
        if (a[nc + 1] < 0) a[nc + 1] += 2 * kPi;
        a[nc + 2] = 24 * 2 * kPi / (1 / freq[b0 + i]);
        a[nc + 3] = p->par[0];
        a[nc + 4] = p->par[1];
        a[nc + 5] = p->par[2];
        a[nc + 6] = std::log(p->cl);
        a[nc + 7] = 1;
      }
    }
    return PS_OK;
  };
  rc = prepare(0, std::min(N, opt->block), cg);
  for (int b0 = 0; b0 < N && !rc; b0 += opt->block) {
    const int b1 = std::min(N, b0 + opt->block), nb = b1 - b0;
    // 3. screening fits (the earlier blocks' judge batches run meanwhile)
    out.assign((size_t)nb * per, ps_fit_out{});
    if ((rc = ps_fit_batch_from(e, cg.data(), nb * per, conw_r, out.data(), nullptr))) break;
    fits_run += (long long)nb * per;
    brow.assign(nb, ps_row{});
    rows_from_fits(out.data(), nb, per, p->n_points, brow.data());
    for (int i = 0; i < nb; i++) {
      rows[b0 + i] = brow[i];
      if (screened) screened[b0 + i] = brow[i];
    }
    // the next block's starts before this block's judge starts, so that the
    // features and the network never wait behind the judge's kernels
    if (b1 < N && (rc = prepare(b1, std::min(N, b1 + opt->block) - b1, cg_next))) break;
    // 4. candidates of this block go to the judge
    std::vector<int> cand = pick_candidates(brow.data(), nb, opt->n_candidates, opt->halo);
    for (int &i : cand) i += b0;
    start_judge(std::move(cand));
    cg.swap(cg_next);
  }
  {
    const int rc_j = finish_judges();
    if (rc) return rc;
    if ((rc = rc_j)) return rc;
  }
  // 5. the judged rows' chi2 over neighbouring trial periods is smooth near a
  //    minimum (neighbours converge to the same solution), so the best judged
  //    row is walked to a local minimum of the kept profile: while a
  //    neighbour of the best judged row is unjudged, judge it (at most
  //    opt->refine rounds)
  for (int round = 0; round < opt->refine; round++) {
    int best = -1;
    for (int i = 0; i < N; i++)
      if (judged[i] && !std::isnan(rows[i].chisq) && (best < 0 || rows[i].chisq < rows[best].chisq)) best = i;
    if (best < 0) break;
    std::vector<int> nbr;
    for (int j : {best - 1, best + 1})
      if (j >= 0 && j < N && !judged[j]) nbr.push_back(j);
    if (nbr.empty()) break;
    start_judge(std::move(nbr));
    if ((rc = finish_judges())) return rc;
  }
  if (n_fits) *n_fits = fits_run;
  return PS_OK;
}

}  // extern "C"
