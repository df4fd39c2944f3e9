// ps_main.cpp -- command-line driver: a BOINC-free stand-in for the
// period_search_BOINC main() (period_search/period_search_BOINC.cpp:147-954)
// that runs the fit loop on the GPU through libpsamd.
//
//   period_search_amd [--device N] [--in F] [--out F] [--state F] [--batch P]
//                     [--round-pole] [--prec fp64|fp32]
//
// Files default to period_search_in / period_search_out / period_search_state
// in the working directory (period_search_BOINC.cpp:81-83).  Checkpoints are
// written atomically (temp file + rename, period_search_BOINC.cpp:85-101) in
// the CUDA app's format "nlines newConw conw_r"
// (period_search_cuda/period_search/period_search_BOINC.cpp:96-112); on start
// the base ("nlines"), CUDA and optimised-CPU ("nlines newConw conw_r sum
// testperiods") formats are all accepted.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "periodsearch.h"

namespace {

struct State {
  int nlines = 0, new_conw = 0;
  double conw_r = 0;
  int fields = 0;
};

State read_state(const std::string &path) {
  State s;
  FILE *f = fopen(path.c_str(), "r");
  if (!f) return s;
  s.fields = fscanf(f, "%d %d %lf", &s.nlines, &s.new_conw, &s.conw_r);
  fclose(f);
  if (s.fields < 1) s.fields = 0;
  return s;
}

int write_state(const std::string &path, int nlines, int new_conw, double conw_r) {
  std::string tmp = path + ".temp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return 1;
  fprintf(f, "%d %d %.17g", nlines, new_conw, conw_r);
  fclose(f);
  return rename(tmp.c_str(), path.c_str());
}

}  // namespace

int main(int argc, char **argv) {
  int device = 0, batch = 512, round_pole = 0, prec = PS_PREC_FP64;
  std::string in_path = "period_search_in", out_path = "period_search_out", state_path = "period_search_state";
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : ""; };
    if (a == "--device")
      device = atoi(next());
    else if (a == "--in")
      in_path = next();
    else if (a == "--out")
      out_path = next();
    else if (a == "--state")
      state_path = next();
    else if (a == "--batch")
      batch = atoi(next());
    else if (a == "--round-pole")
      round_pole = 1;
    else if (a == "--prec")
      prec = std::string(next()) == "fp32" ? PS_PREC_FP32 : PS_PREC_FP64;
    else {
      fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (batch < 1) batch = 1;
  ps_input *in = nullptr;
  if (ps_input_read(in_path.c_str(), &in) != PS_OK) {
    fprintf(stderr, "Couldn't read input file %s\n", in_path.c_str());
    return 1;
  }
  ps_problem prob;
  ps_input_problem(in, &prob);
  double sc[10];
  int ndata = 0;
  ps_input_scalars(in, sc, &ndata);
  const double escl = sc[5], conw = sc[6], freq_start = sc[7], freq_end = sc[8], freq_step = sc[9];
  const int n_periods = (int)((freq_start - freq_end) / freq_step) + 1;

  ps_engine *eng = nullptr;
  if (ps_create(device, &eng) != PS_OK) {
    fprintf(stderr, "cannot open GPU device %d\n", device);
    return 2;
  }
  if (ps_set_precision(eng, prec) != PS_OK || ps_set_problem(eng, &prob) != PS_OK) {
    fprintf(stderr, "engine setup failed: %s\n", ps_last_error(eng));
    return 2;
  }
  State st = read_state(state_path);
  const bool resume = st.fields >= 1;
  int n_start = resume ? st.nlines + 1 : 1;
  double conw_r = conw / escl / escl;
  int new_conw = 0;
  if (resume && st.fields >= 3) {
    conw_r = st.conw_r;
    new_conw = st.new_conw;
  }
  // convexity-weight precalc (period_search_BOINC.cpp:487-715); skipped when
  // the checkpoint carries the converged weight.
  if (!(resume && st.fields >= 3 && new_conw == 1)) {
    new_conw = 0;
    if (!(resume && st.fields >= 3)) conw_r = conw / escl / escl;
    while (new_conw != 1 && conw_r * escl * escl < 10.0) {
      if (ps_precalc(eng, freq_start, freq_end, freq_step, &conw_r, &new_conw) != PS_OK) {
        fprintf(stderr, "precalc failed: %s\n", ps_last_error(eng));
        return 2;
      }
      if (write_state(state_path, n_start - 1, new_conw, conw_r)) {
        fprintf(stderr, "checkpoint failed\n");
        return 1;
      }
    }
  }
  FILE *out = fopen(out_path.c_str(), n_start > 1 ? "a" : "w");
  if (!out) {
    fprintf(stderr, "output open failed: %s\n", out_path.c_str());
    return 1;
  }
  std::vector<ps_row> rows(batch);
  char buf[256];
  for (int n0 = n_start; n0 <= n_periods; n0 += batch) {
    int n1 = n0 + batch - 1;
    if (n1 > n_periods) n1 = n_periods;
    if (ps_scan(eng, freq_start, freq_step, n0, n1, conw_r, rows.data()) != PS_OK) {
      fprintf(stderr, "scan failed: %s\n", ps_last_error(eng));
      return 2;
    }
    for (int n = n0; n <= n1; n++) {
      ps_row r = rows[n - n0];
      if (n == 1) r.dark = conw_r * escl * escl;
      ps_format_row(&r, round_pole, buf, sizeof buf);
      fputs(buf, out);
    }
    fflush(out);
    if (write_state(state_path, n1, 1, conw_r)) {
      fprintf(stderr, "checkpoint failed\n");
      return 1;
    }
    fprintf(stderr, "rows %d..%d of %d done\n", n0, n1, n_periods);
  }
  fclose(out);
  ps_destroy(eng);
  ps_input_free(in);
  return 0;
}
