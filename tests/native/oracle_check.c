/* Sanitizer driver for the FP64 oracle (test infrastructure): round trips of
 * the four switch combinations that exercise every oracle path (2 and 3
 * equations, CW and NS, EOM and GDD, saving and power history), built with
 * -fsanitize=address,undefined by tests/native/Makefile.  Exits 0 when every
 * result is finite and no sanitizer fired. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/opo_oracle.h"

static int run_case(int three, int regime, int crystal, int upm, double gdd) {
  orc_inputs in = {0};
  in.three_eqs = three; in.regime = regime; in.crystal = crystal;
  in.N = 256; in.NZ = 4; in.NRT = 6;
  in.lp_nm = 532; in.Lambda = crystal == ORC_PPLN ? 6.97 : 7.97; in.Nth = 4; in.R_pct = 98;
  in.delta = 0.05; in.gdd_pct = gdd; in.upm = upm; in.md = 0.25; in.fm = 4; in.lcr_mm = 5;
  orc_params p;
  if (orc_setup(&in, &p) != 0) return 1;
  const size_t N = (size_t)p.N, n_ap = regime == ORC_CW ? N : p.SIZEL;
  const size_t slots = regime == ORC_CW ? (size_t)p.Nrts : (size_t)p.NRT;
  double *T = malloc(N * sizeof *T), *w = malloc(N * sizeof *w);
  double *Tp = malloc(p.SIZEL * sizeof *Tp), *Fp = malloc(p.SIZEL * sizeof *Fp);
  double *Ap_in = malloc(2 * n_ap * sizeof *Ap_in);
  double *Ap = malloc(2 * N * sizeof *Ap), *As = malloc(2 * N * sizeof *As), *Ai = calloc(2 * N, sizeof *Ai);
  double *sp = malloc(2 * slots * N * sizeof *sp), *ss = malloc(2 * slots * N * sizeof *ss);
  double *si = malloc(2 * slots * N * sizeof *si), *pw = malloc((size_t)p.NRT * sizeof *pw);
  orc_grids(&p, T, w);
  orc_grids_long(&p, Tp, Fp);
  orc_pump_input(&p, Ap_in);
  orc_noise(3, 0, 1e-3, As, N);
  if (three) orc_noise(3, 1, 1e-3, Ai, N);
  orc_read_pump(&p, Ap_in, 0, Ap);
  orc_run(&p, w, T, Ap_in, Ap, As, three ? Ai : NULL, 0, p.NRT, sp, ss, three ? si : NULL, pw);
  int bad = 0;
  for (size_t i = 0; i < 2 * N; ++i) bad |= !isfinite(Ap[i]) || !isfinite(As[i]) || !isfinite(Ai[i]);
  for (int i = 0; i < p.NRT; ++i) bad |= !isfinite(pw[i]);
  free(T); free(w); free(Tp); free(Fp); free(Ap_in); free(Ap); free(As); free(Ai);
  free(sp); free(ss); free(si); free(pw);
  printf("three=%d regime=%d crystal=%d upm=%d gdd=%g: %s\n", three, regime, crystal, upm, gdd, bad ? "bad" : "ok");
  return bad;
}

int main(void) {
  int bad = 0;
  bad |= run_case(0, ORC_CW, ORC_PPLN, 0, 0);
  bad |= run_case(1, ORC_CW, ORC_SPPLT, 0, 0);
  bad |= run_case(1, ORC_NS, ORC_PPLN, 1, 100);
  bad |= run_case(0, ORC_NS, ORC_PPLN, 1, 50);
  return bad;
}
