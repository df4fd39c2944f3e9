// ps_main.cpp -- command-line driver: a BOINC-free stand-in for the
// period_search_BOINC main() (period_search/period_search_BOINC.cpp:147-954)
// that runs the fit loop on the GPU through libpsamd.
//
//   period_search_amd [--device N] [--in F] [--out F] [--state F] [--batch P]
//                     [--round-pole] [--prec fp64|fp32|bf16|bf16x3|bf16x6]
//                     [--state-format base|cuda|oc] [--clamp-min] [--progress]
//
// Files default to period_search_in / period_search_out / period_search_state
// in the working directory (period_search_BOINC.cpp:81-83).  Checkpoints are
// written atomically (temp file + rename, period_search_BOINC.cpp:85-101) in
// one of the three reference formats (--state-format, default the CUDA app's
// "nlines newConw conw_r", period_search_cuda/period_search/period_search_BOINC.cpp:96-112);
// on start any of the three is accepted (ps_state_read), and a missing or
// unreadable checkpoint starts afresh with the output truncated
// (period_search_BOINC.cpp:238-251).  An optimised-CPU-app checkpoint taken
// inside a convexity-weight round resumes that round
// (period_search_optimization_cpu/period_search/period_search_BOINC.cpp:627-640).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "periodsearch.h"

namespace {

void print_progress(void *, long long done, long long total) {
  fprintf(stderr, "  fits %lld of %lld\n", done, total);
}

int parse_prec(const std::string &s) {
  if (s == "fp64") return PS_PREC_FP64;
  if (s == "fp32") return PS_PREC_FP32;
  if (s == "bf16") return PS_PREC_BF16;
  if (s == "bf16x3") return PS_PREC_BF16X3;
  if (s == "bf16x6") return PS_PREC_BF16X6;
  return -1;
}

int parse_state_format(const std::string &s) {
  if (s == "base") return PS_STATE_BASE;
  if (s == "cuda") return PS_STATE_CUDA;
  if (s == "oc") return PS_STATE_OC;
  return -1;
}

}  // namespace

int main(int argc, char **argv) {
  int device = 0, batch = 512, round_pole = 0, prec = PS_PREC_FP64, state_format = PS_STATE_CUDA, flags = 0;
  bool progress = false;
  std::string in_path = "period_search_in", out_path = "period_search_out", state_path = "period_search_state";
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) {
        fprintf(stderr, "%s needs a value\n", a.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (a == "--device") {
      device = atoi(next().c_str());
    } else if (a == "--in") {
      in_path = next();
    } else if (a == "--out") {
      out_path = next();
    } else if (a == "--state") {
      state_path = next();
    } else if (a == "--batch") {
      batch = atoi(next().c_str());
    } else if (a == "--round-pole") {
      round_pole = 1;
    } else if (a == "--clamp-min") {
      flags |= PS_READ_CLAMP_MIN;
    } else if (a == "--progress") {
      progress = true;
    } else if (a == "--prec") {
      const std::string v = next();
      prec = parse_prec(v);
      if (prec < 0) {
        fprintf(stderr, "unknown precision '%s' (fp64, fp32, bf16, bf16x3, bf16x6)\n", v.c_str());
        return 2;
      }
    } else if (a == "--state-format") {
      const std::string v = next();
      state_format = parse_state_format(v);
      if (state_format < 0) {
        fprintf(stderr, "unknown checkpoint format '%s' (base, cuda, oc)\n", v.c_str());
        return 2;
      }
    } else {
      fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (batch < 1) batch = 1;
  ps_input *in = nullptr;
  if (ps_input_read_ex(in_path.c_str(), flags, &in) != PS_OK) {
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
  if (progress) ps_set_progress(eng, print_progress, nullptr);
  ps_state st;
  ps_state_read(state_path.c_str(), &st);
  const bool resume = st.format != PS_STATE_NONE;
  const int n_start = resume ? st.nlines + 1 : 1;
  // The convexity weight: the base format does not carry it, so the precalc
  // runs again (period_search/period_search_BOINC.cpp:487); the CUDA and OC
  // formats carry it, and newConw = 1 marks it final.
  double conw_r = conw / escl / escl;
  int new_conw = 0, done_periods = 0;
  double sum_dark = 0;
  if (st.format >= PS_STATE_CUDA) {
    conw_r = st.conw_r;
    new_conw = st.new_conw;
    if (st.format == PS_STATE_OC && new_conw != 1) {
      done_periods = st.test_periods;
      sum_dark = st.sum_dark;
    }
  }
  auto checkpoint = [&](int nlines, int sum_periods) {
    ps_state w{};
    w.format = state_format;
    w.nlines = nlines;
    w.new_conw = new_conw;
    w.conw_r = conw_r;
    w.sum_dark = sum_periods ? sum_dark : 0.0;
    w.test_periods = sum_periods;
    if (ps_state_write(state_path.c_str(), &w) != PS_OK) {
      fprintf(stderr, "checkpoint failed\n");
      exit(1);
    }
  };
  // convexity-weight precalc (period_search_BOINC.cpp:487-715)
  while (new_conw != 1 && conw_r * escl * escl < 10.0) {
    if (ps_precalc_resume(eng, freq_start, freq_end, freq_step, done_periods, sum_dark, &conw_r, &new_conw) !=
        PS_OK) {
      fprintf(stderr, "precalc failed: %s\n", ps_last_error(eng));
      return 2;
    }
    done_periods = 0;
    sum_dark = 0;
    checkpoint(n_start - 1, 0);
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
    new_conw = 1;
    checkpoint(n1, 0);
    fprintf(stderr, "rows %d..%d of %d done\n", n0, n1, n_periods);
  }
  fclose(out);
  ps_destroy(eng);
  ps_input_free(in);
  return 0;
}
