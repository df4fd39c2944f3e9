// ps_main.cpp -- command-line driver: a BOINC-free stand-in for the
// period_search_BOINC main() (period_search/period_search_BOINC.cpp:147-954)
// that runs the fit loop on the GPU through libpsamd.
//
//   period_search_amd [--device N] [--in F] [--out F] [--state F] [--batch P]
//                     [--round-pole] [--prec fp64|fp32|bf16|bf16x3|bf16x6|bf16a]
//                     [--state-format base|cuda|oc] [--clamp-min] [--progress]
//                     [--learned [--net F] [--block B] [--candidates K] [--halo H]
//                      [--refine R] [--mirror] [--judged-out F]]
//
// --learned runs the learned scan (ps_scan_learned) instead of the ten-pole
// scan: every trial period fitted from the network's proposed start in the
// --prec arithmetic, the candidates of each block and their neighbours
// refitted from the ten poles by a second engine in FP64 (those rows are the
// reference's; --judged-out writes "n judged" per row); the convexity-weight
// precalc then runs on that FP64 engine.  The network defaults to
// proposal_net.npz next to this executable.  In this mode --batch (the
// rows per call and checkpoint) defaults to the whole period range.
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
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "periodsearch.h"
#include "ps_proposal.h"

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
  if (s == "bf16a") return PS_PREC_BF16A;
  return -1;
}

int parse_state_format(const std::string &s) {
  if (s == "base") return PS_STATE_BASE;
  if (s == "cuda") return PS_STATE_CUDA;
  if (s == "oc") return PS_STATE_OC;
  return -1;
}

// proposal_net.npz in the directory of this executable
std::string default_net() {
  char buf[4096];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
  if (n <= 0) return "proposal_net.npz";
  std::string p(buf, n);
  const size_t sl = p.rfind('/');
  return (sl == std::string::npos ? std::string(".") : p.substr(0, sl)) + "/proposal_net.npz";
}

}  // namespace

int main(int argc, char **argv) {
  int device = 0, batch = 512, round_pole = 0, prec = PS_PREC_FP64, state_format = PS_STATE_CUDA, flags = 0;
  bool progress = false, learned = false, batch_given = false;
  ps_learned_opts lopt{2240, 8, 1, 0, 8, nullptr};
  std::string in_path = "period_search_in", out_path = "period_search_out", state_path = "period_search_state";
  std::string net_path = default_net(), judged_path;
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
      batch_given = true;
    } else if (a == "--learned") {
      learned = true;
    } else if (a == "--net") {
      net_path = next();
    } else if (a == "--block") {
      lopt.block = atoi(next().c_str());
    } else if (a == "--candidates") {
      lopt.n_candidates = atoi(next().c_str());
    } else if (a == "--halo") {
      lopt.halo = atoi(next().c_str());
    } else if (a == "--refine") {
      lopt.refine = atoi(next().c_str());
    } else if (a == "--mirror") {
      lopt.mirror = 1;
    } else if (a == "--judged-out") {
      judged_path = next();
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
        fprintf(stderr, "unknown precision '%s' (fp64, fp32, bf16, bf16x3, bf16x6, bf16a)\n", v.c_str());
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
  // learned scan: the network and an FP64 engine for the ten-pole judge fits
  ps_proposal_net net{};
  ps_engine *judge = nullptr;
  if (learned) {
    if (!batch_given) batch = n_periods;
    const int rc = ps_proposal_open(net_path.c_str(), device, std::max(1, std::min(batch, lopt.block)), &net);
    if (rc != PS_OK) {
      fprintf(stderr, "cannot load the proposal network %s (%d)\n", net_path.c_str(), rc);
      return 2;
    }
    if (ps_create(device, &judge) != PS_OK || ps_set_precision(judge, PS_PREC_FP64) != PS_OK ||
        ps_set_problem(judge, &prob) != PS_OK) {
      fprintf(stderr, "judge engine setup failed: %s\n", judge ? ps_last_error(judge) : "no device");
      return 2;
    }
    lopt.judge = judge;
  }
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
  // convexity-weight precalc (period_search_BOINC.cpp:487-715), in FP64 on
  // the judge engine when the learned scan has one
  ps_engine *pre = judge ? judge : eng;
  while (new_conw != 1 && conw_r * escl * escl < 10.0) {
    if (ps_precalc_resume(pre, freq_start, freq_end, freq_step, done_periods, sum_dark, &conw_r, &new_conw) !=
        PS_OK) {
      fprintf(stderr, "precalc failed: %s\n", ps_last_error(pre));
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
  FILE *jout = nullptr;
  if (learned && !judged_path.empty() && !(jout = fopen(judged_path.c_str(), n_start > 1 ? "a" : "w"))) {
    fprintf(stderr, "output open failed: %s\n", judged_path.c_str());
    return 1;
  }
  std::vector<ps_row> rows(batch);
  std::vector<int> judged(batch, 1);
  char buf[256];
  for (int n0 = n_start; n0 <= n_periods; n0 += batch) {
    int n1 = n0 + batch - 1;
    if (n1 > n_periods) n1 = n_periods;
    long long n_fits = 0;
    const int rc = learned ? ps_scan_learned(eng, &prob, &net, &lopt, freq_start, freq_step, n0, n1, conw_r,
                                             rows.data(), judged.data(), nullptr, &n_fits)
                           : ps_scan(eng, freq_start, freq_step, n0, n1, conw_r, rows.data());
    if (rc != PS_OK) {
      fprintf(stderr, "scan failed: %s\n", ps_last_error(eng));
      return 2;
    }
    for (int n = n0; n <= n1; n++) {
      ps_row r = rows[n - n0];
      if (n == 1) r.dark = conw_r * escl * escl;
      ps_format_row(&r, round_pole, buf, sizeof buf);
      fputs(buf, out);
      if (jout) fprintf(jout, "%d %d\n", n, judged[n - n0]);
    }
    fflush(out);
    if (jout) fflush(jout);
    new_conw = 1;
    checkpoint(n1, 0);
    fprintf(stderr, "rows %d..%d of %d done\n", n0, n1, n_periods);
  }
  fclose(out);
  if (jout) fclose(jout);
  if (learned) {
    ps_proposal_close(&net);
    ps_destroy(judge);
  }
  ps_destroy(eng);
  ps_input_free(in);
  return 0;
}
