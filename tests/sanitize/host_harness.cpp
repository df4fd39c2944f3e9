// host_harness.cpp -- the library's host-side code (csrc/ps_host.cpp,
// csrc/ps_learned.cpp) driven without a GPU, for a build under
// AddressSanitizer and UndefinedBehaviorSanitizer (test infrastructure).
//
//   host_harness <scratch dir> <network .npz> <rows file> <input>...
//
// Per input: parsing with each reader option, the problem view (every array
// read end to end), the scalars and the aspect features.  Then the row text
// (both pole modes, a short buffer), the rows reader and best-row pick, each
// checkpoint format written and read back plus damaged ones, the pole
// reduction with NaN fits, and the network-file reader on the given network
// and on damaged copies.  The engine and network entry points that
// ps_learned.cpp refers to are defined below to return PS_ERR_HIP: none is
// reached except ps_mlp_create, after the file has been parsed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "periodsearch.h"
#include "ps_internal.h"
#include "ps_proposal.h"

void psamd_engine_error(ps_engine *, const std::string &) {}

extern "C" {
int ps_dims(const ps_engine *, int *, int *, int *) { return PS_ERR_HIP; }
int ps_fit_batch(ps_engine *, const double *, const int *, int, double, ps_fit_out *) { return PS_ERR_HIP; }
int ps_fit_batch_from(ps_engine *, const double *, int, double, ps_fit_out *, double *) { return PS_ERR_HIP; }
int ps_lc_features(ps_engine *, const double *, int, int, float *) { return PS_ERR_HIP; }
int ps_mlp_create(int, int, int, int, int, ps_mlp **) { return PS_ERR_HIP; }
int ps_mlp_destroy(ps_mlp *) { return PS_ERR_HIP; }
int ps_mlp_forward(ps_mlp *, const float *, int, float *) { return PS_ERR_HIP; }
int ps_mlp_set_params(ps_mlp *, const float *, long long) { return PS_ERR_HIP; }
const char *ps_mlp_last_error(const ps_mlp *) { return "no device"; }
const char *ps_last_error(const ps_engine *) { return "no device"; }
}

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      failures++;                                                     \
    }                                                                 \
  } while (0)

static double sum(const double *a, long long n) {
  double s = 0;
  for (long long i = 0; i < n; i++) s += a[i];
  return s;
}

static void inputs(const char *path) {
  const int opts[3] = {0, PS_READ_CLAMP_MIN, PS_READ_DEFER_SHAPE};
  for (int flags : opts) {
    ps_input *in = nullptr;
    CHECK(ps_input_read_ex(path, flags, &in) == PS_OK);
    if (!in) continue;
    ps_problem p;
    CHECK(ps_input_problem(in, &p) == PS_OK);
    long long pts = 0;
    for (int k = 0; k < p.n_lc; k++) pts += p.lc_points[k] + p.lc_inrel[k] * 0;
    CHECK(pts == p.n_points);
    double s = sum(p.tim, p.n_points) + sum(p.brightness, p.n_points) + sum(p.sig, p.n_points) +
               sum(p.weight, p.n_points) + sum(p.ee, 3LL * p.n_points) + sum(p.ee0, 3LL * p.n_points);
    const int nc = [&] {
      int n = 0;
      for (int m = 0; m <= p.mmax; m++) n += (p.lmax - m + 1) * (m ? 2 : 1);
      return n;
    }();
    s += sum(p.nor, 3LL * p.n_fac) + sum(p.darea, p.n_fac) + sum(p.fc, (long long)p.n_fac * (p.mmax + 1)) +
         sum(p.fs, (long long)p.n_fac * (p.mmax + 1)) +
         sum(p.pleg, (long long)p.n_fac * (p.lmax + 1) * (p.lmax + 1)) + sum(p.dsph, (long long)p.n_fac * nc) +
         sum(p.cg_first, nc) + sum(p.beta_pole, p.n_poles) + sum(p.lambda_pole, p.n_poles);
    CHECK(std::isfinite(s));
    double sc[10];
    int ndata = 0;
    CHECK(ps_input_scalars(in, sc, &ndata) == PS_OK && ndata == p.n_points);
    float asp[5];
    CHECK(ps_aspect_features(&p, asp) == PS_OK);
    ps_input_free(in);
  }
}

static void truncated(const char *path, const std::string &dir) {
  // the first k lines of an input: a clean error, never a crash
  FILE *f = fopen(path, "r");
  if (!f) return;
  std::vector<std::string> lines;
  char buf[4096];
  while (fgets(buf, sizeof buf, f)) lines.emplace_back(buf);
  fclose(f);
  const std::string cut = dir + "/cut_in";
  for (size_t k : {size_t(0), size_t(3), size_t(12), lines.size() / 2, lines.size() - 1}) {
    FILE *g = fopen(cut.c_str(), "w");
    for (size_t i = 0; i < k && i < lines.size(); i++) fputs(lines[i].c_str(), g);
    fclose(g);
    ps_input *in = nullptr;
    const int rc = ps_input_read(cut.c_str(), &in);
    if (rc == PS_OK) ps_input_free(in);
  }
}

static void rows_and_text(const char *rows_path) {
  std::vector<ps_row> rows(100000);
  long long n = 0;
  CHECK(ps_read_rows(rows_path, rows.data(), (long long)rows.size(), &n) == PS_OK && n > 0);
  long long best = -1;
  CHECK(ps_best_row(rows.data(), n, &best) == PS_OK && best >= 0 && best < n);
  long long n2 = 0;
  ps_read_rows(rows_path, rows.data(), 2, &n2);  // capacity below the file's rows
  char buf[256], small[8];
  for (long long i = 0; i < n; i++)
    for (int rp = 0; rp < 2; rp++) {
      CHECK(ps_format_row(&rows[i], rp, buf, sizeof buf) > 0);
      ps_format_row(&rows[i], rp, small, sizeof small);
    }
}

static void states(const std::string &dir) {
  const std::string sp = dir + "/state";
  for (int fmt : {PS_STATE_BASE, PS_STATE_CUDA, PS_STATE_OC}) {
    ps_state w{}, r{};
    w.format = fmt;
    w.nlines = 17;
    w.new_conw = fmt == PS_STATE_OC ? 0 : 1;
    w.conw_r = 0.00026564558463277498;
    w.sum_dark = 3.25;
    w.test_periods = 4;
    CHECK(ps_state_write(sp.c_str(), &w) == PS_OK);
    CHECK(ps_state_read(sp.c_str(), &r) == PS_OK && r.format == fmt && r.nlines == 17);
  }
  const char *bad[] = {"", "garbage\n", "-4 1 0.1", "5 1 nan", "5x", "1 2 3 4 5 6 7 8 9 10 11 12",
                       "99999999999999999999 1 0.1"};
  for (const char *t : bad) {
    FILE *f = fopen(sp.c_str(), "w");
    fputs(t, f);
    fclose(f);
    ps_state r{};
    CHECK(ps_state_read(sp.c_str(), &r) == PS_OK);
  }
  ps_state r{};
  CHECK(ps_state_read((dir + "/missing").c_str(), &r) == PS_OK && r.format == PS_STATE_NONE);
}

static void pole_reduction() {
  const int per = 10, np = 7;
  std::vector<ps_fit_out> f(per * np);
  for (int i = 0; i < per * np; i++) {
    memset(&f[i], 0, sizeof f[i]);
    f[i].dev = (i % 3 == 0) ? NAN : 0.01 * (1 + (i * 7919) % 13);
    f[i].prd = 0.3;
    f[i].la = 400.0;
  }
  for (int k = 0; k < per; k++) f[3 * per + k].dev = NAN;  // a period with every pole NaN
  std::vector<ps_row> rows(np);
  CHECK(ps_rows_from_fits(f.data(), np, per, 1000, rows.data()) == PS_OK);
}

static void network(const char *net, const std::string &dir) {
  ps_proposal_net n{};
  CHECK(ps_proposal_open(net, 0, 64, &n) == PS_ERR_HIP);  // parsed, then no device
  CHECK(ps_proposal_open((dir + "/missing.npz").c_str(), 0, 64, &n) == PS_ERR_IO);
  // damaged copies: truncated at several points, and a flipped header byte
  FILE *f = fopen(net, "rb");
  std::vector<unsigned char> z;
  unsigned char tmp[1 << 16];
  size_t k;
  while ((k = fread(tmp, 1, sizeof tmp, f)) > 0) z.insert(z.end(), tmp, tmp + k);
  fclose(f);
  const std::string cp = dir + "/net_copy.npz";
  // (a cut inside the parameters fails; the last byte belongs to the zip's
  // directory, which the reader does not need)
  for (size_t cut : {size_t(10), size_t(31), size_t(200), z.size() / 2, z.size() - 1}) {
    FILE *g = fopen(cp.c_str(), "wb");
    fwrite(z.data(), 1, cut, g);
    fclose(g);
    const int rc = ps_proposal_open(cp.c_str(), 0, 64, &n);
    CHECK(cut == z.size() - 1 ? (rc == PS_ERR_IO || rc == PS_ERR_HIP) : rc == PS_ERR_IO);
  }
  for (size_t pos : {size_t(26), size_t(28), size_t(60), size_t(70)}) {
    std::vector<unsigned char> y = z;
    y[pos] ^= 0xff;
    FILE *g = fopen(cp.c_str(), "wb");
    fwrite(y.data(), 1, y.size(), g);
    fclose(g);
    const int rc = ps_proposal_open(cp.c_str(), 0, 64, &n);
    CHECK(rc == PS_ERR_IO || rc == PS_ERR_HIP);
  }
}

int main(int argc, char **argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s <scratch dir> <network .npz> <rows file> <input>...\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  for (int i = 4; i < argc; i++) {
    inputs(argv[i]);
    truncated(argv[i], dir);
  }
  rows_and_text(argv[3]);
  states(dir);
  pole_reduction();
  network(argv[2], dir);
  fprintf(stderr, "%d check(s) failed\n", failures);
  return failures ? 1 : 0;
}
