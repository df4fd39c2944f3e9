/*
 * ps_oracle_main.c -- BOINC-free command-line driver for the oracle
 * (test infrastructure only).  Reads period_search_in, runs the conw precalc
 * and the scan, writes period_search_out, like period_search/period_search_BOINC.cpp
 * in standalone mode.
 *
 *   ps_oracle_main <in> <out> [n_from n_to] [conw_r]
 *
 * n_from/n_to restrict the scan to a slice of the trial-period index (the
 * BOINC work-unit split); a given conw_r skips the precalc.
 */
#include <stdio.h>
#include <stdlib.h>

#include "ps_oracle.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <period_search_in> <period_search_out> [n_from n_to] [conw_r]\n", argv[0]);
    return 2;
  }
  orc_ctx *c = orc_create();
  if (orc_read_input(c, argv[1]) != 0) {
    fprintf(stderr, "cannot parse %s\n", argv[1]);
    return 2;
  }
  if (argc >= 6)
    orc_set(c, "conw_r", atof(argv[5]));
  else
    orc_precalc(c);
  orc_prepare_round(c, orc_get(c, "conw_r"));
  int n_total = orc_n_periods(c);
  int n_from = 1, n_to = n_total;
  if (argc >= 5) {
    n_from = atoi(argv[3]);
    n_to = atoi(argv[4]);
    if (n_to > n_total) n_to = n_total;
  }
  int rows = orc_scan_to_file(c, n_from, n_to, argv[2], "w");
  fprintf(stderr, "conw_r %.17g rows %d of %d\n", orc_get(c, "conw_r"), rows, n_total);
  orc_destroy(c);
  return rows < 0;
}
