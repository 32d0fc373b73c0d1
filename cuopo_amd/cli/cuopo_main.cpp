// cuopo_main.cpp -- drop-in replacement for the cuOPO executable
// (src/cuOPO.cu:52-688): the same compile-time switches, the same 12 positional
// arguments, the same stdout parameter dump and the same .dat outputs, with
// the round-trip loop running on one MI355X through libcuopo_amd.so.
//
//   hipcc cuopo_main.cpp -D<CW_OPO|NS_OPO> -D<PPLN|SPPLT> [-DTHREE_EQS] -lcuopo_amd -o cuOPO
//   ./cuOPO <LP_nm> <T_C> <Lambda_um> <N> <R_%> <delta_pi> <GDD_%> <TOD_%> <UPM> <beta_pi> <df> <Lcr_mm> [flags]
//
// Optional trailing flags (defaults reproduce the reference):
//   --N=<pow2> --nz=<n> --nrt=<n>     grid size, slices per pass, round trips
//   --seed=<u64> --noise=<sigma>      seeded Gaussian noise on signal/idler
//   --ls=<um>                         signal wavelength (reference: 2*lp)
//   --pm=3wave                        phase-match the three-wave mismatch
//   --out=<dir>                       output directory (default: cwd)
//   --fp32c                           FP32-complex variant
//   --no-inputs                       skip the *_input_*.dat files
//   --format=dat|bin|both             text .dat files (default), raw float64
//                                     .bin side-cars, or both
//   --gdd-fs2=<fs^2>                  intracavity GDD in fs^2 instead of argv7's
//                                     compensation percentage
//   --converge[=<tol>]                CW: stop the round trips once the signal
//                                     power changes by < tol (1e-6) relative
//                                     over 50 of them, then run the 64 saved
//                                     round trips
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/opo.h"
#include "opo_cli_common.h"

#if !defined(CW_OPO) && !defined(NS_OPO)
#error "define the regime: -DCW_OPO or -DNS_OPO (src/cuOPO.cu:65-70)"
#endif
#if !defined(PPLN) && !defined(SPPLT)
#error "define the crystal: -DPPLN or -DSPPLT (src/cuOPO.cu:36-41)"
#endif

#ifndef OPO_SIZE
#define OPO_SIZE (1 << 14)  // SIZE, src/cuOPO.cu:28
#endif
#ifndef OPO_NZ
#define OPO_NZ 150          // NZ, src/cuOPO.cu:29
#endif
#ifndef OPO_NRT
#define OPO_NRT 10000       // NRT, src/cuOPO.cu:30
#endif

// CHECK() of src/headers/common.cuh:6-16: print and exit(1)
#define CHECK(call)                                                                      \
  do {                                                                                   \
    const int st_ = (call);                                                              \
    if (st_ != OPO_OK) {                                                                 \
      std::fprintf(stderr, "Error: %s:%d, ", __FILE__, __LINE__);                        \
      std::fprintf(stderr, "code: %d, reason: %s (%s)\n", st_, opo_strerror(st_), opo_last_error()); \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

namespace {

double seconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void usage(const char *prog) {
  std::fprintf(stderr,
               "usage: %s <LP_nm> <T_C> <Lambda_um> <N> <R_%%> <delta_pi> <GDD_%%> <TOD_%%> <UPM> <beta_pi> <df> "
               "<Lcr_mm> [--N=] [--nz=] [--nrt=] [--seed=] [--noise=] [--ls=] [--pm=3wave] [--out=] [--fp32c] "
               "[--no-inputs] [--format=dat|bin|both] [--gdd-fs2=] [--converge[=tol]]\n",
               prog);
}

}  // namespace

int main(int argc, char *argv[]) {
#ifdef NS_OPO
  const int regime = OPO_NS;
#else
  const int regime = OPO_CW;
#endif
#ifdef THREE_EQS
  const int three = 1;
#else
  const int three = 0;
#endif
  opo_cli::print_banner(std::cout, regime, three);
  const double iStart = seconds();
  if (argc < 13) {
    usage(argv[0]);
    return 1;
  }

  opo_args a{};
  a.three_eqs = three;
  a.regime = regime;
#ifdef SPPLT
  a.crystal = OPO_SPPLT;
#else
  a.crystal = OPO_PPLN;
#endif
  a.N = OPO_SIZE; a.NZ = OPO_NZ; a.NRT = OPO_NRT;
  a.lp_nm = std::atof(argv[1]); a.temp_arg = std::atof(argv[2]); a.Lambda = std::atof(argv[3]);
  a.Nth = std::atof(argv[4]); a.R_pct = std::atof(argv[5]); a.delta = std::atof(argv[6]);
  a.gdd_pct = std::atof(argv[7]); a.tod_pct = std::atof(argv[8]); a.upm = std::atoi(argv[9]);
  a.md = std::atof(argv[10]); a.fm = std::atof(argv[11]); a.lcr_mm = std::atof(argv[12]);
  uint64_t seed = (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();  // functions.cuh:60
  double sigma = 1.0e-15;                                                                    // functions.cuh:62
  std::string out = ".";
  int precision = OPO_FP64;
  bool save_inputs = true;
  double gdd_fs2 = 0.0, converge = 0.0;
  bool have_gdd_fs2 = false;
  for (int k = 13; k < argc; ++k) {
    const std::string f = argv[k];
    auto val = [&](const char *key) -> const char * {
      const size_t n = std::strlen(key);
      return f.compare(0, n, key) == 0 ? f.c_str() + n : nullptr;
    };
    if (const char *v = val("--N=")) a.N = std::atoi(v);
    else if (const char *v = val("--nz=")) a.NZ = std::atoi(v);
    else if (const char *v = val("--nrt=")) a.NRT = std::atoi(v);
    else if (const char *v = val("--seed=")) seed = std::strtoull(v, nullptr, 10);
    else if (const char *v = val("--noise=")) sigma = std::atof(v);
    else if (const char *v = val("--ls=")) a.ls_um = std::atof(v);
    else if (const char *v = val("--out=")) out = v;
    else if (f == "--pm=3wave") a.pm_mode = OPO_PM_THREE_WAVE;
    else if (f == "--fp32c") precision = OPO_FP32C;
    else if (f == "--no-inputs") save_inputs = false;
    else if (const char *v = val("--format=")) {
      if (!opo_cli::set_format(v)) { usage(argv[0]); return 1; }
    } else if (const char *v = val("--gdd-fs2=")) { gdd_fs2 = std::atof(v); have_gdd_fs2 = true; }
    else if (const char *v = val("--converge=")) converge = std::atof(v);
    else if (f == "--converge") converge = 1e-6;
    else { usage(argv[0]); return 1; }
  }
  if (have_gdd_fs2) {
    // GDD = -epsilon * beta2_s * Lcr (src/cuOPO.cu:142): the percentage that gives it
    opo_derived d0{};
    CHECK(opo_derive(&a, &d0));
    a.gdd_pct = -100.0 * (gdd_fs2 * 1e-6) / (d0.b2s * d0.Lcr);
  }

  opo_derived d{};
  CHECK(opo_derive(&a, &d));
  const size_t N = (size_t)d.N;
  std::vector<double> T(N), w(N), Tp(d.SIZEL), Fp(d.SIZEL);
  opo_grids(&d, T.data(), w.data());
  opo_grids_long(&d, Tp.data(), Fp.data());
  const size_t n_ap = d.regime == OPO_CW ? N : d.SIZEL;
  std::vector<double> Ap_in(2 * n_ap), As(2 * N), Ai(2 * N, 0.0);
  opo_pump_input(&d, Ap_in.data());
  opo_noise(seed, 0, sigma, As.data(), N);
  if (d.three_eqs) opo_noise(seed, 1, sigma, Ai.data(), N);

  if (save_inputs) {  // src/cuOPO.cu:272-286
    opo_cli::write_complex(out + "/pump_input", Ap_in.data(), n_ap);
    opo_cli::write_complex(out + "/signal_input", As.data(), N);
    if (d.three_eqs) opo_cli::write_complex(out + "/idler_input", Ai.data(), N);
  }

  // parameter dump, src/cuOPO.cu:289-369
  opo_cli::print_params(std::cout, d, a, have_gdd_fs2 ? a.gdd_pct : std::atof(argv[7]), std::atof(argv[8]),
                        Fp[d.SIZEL - 1]);

  // solver (device allocation + uploads, src/cuOPO.cu:379-442)
  opo_params p{};
  opo_params_from_derived(&d, precision, &p);
  opo_solver *s = nullptr;
  CHECK(opo_create(&p, &s));
  std::vector<double> Ap(Ap_in.begin(), Ap_in.begin() + 2 * N);
  CHECK(opo_upload(s, w.data(), T.data(), Ap_in.data(), n_ap, Ap.data(), As.data(), d.three_eqs ? Ai.data() : nullptr));
  CHECK(opo_enable_save(s, 1));
  std::cout << "Starting main loop on CPU & GPU...\n" << std::endl;

  if (converge > 0 && d.regime == OPO_CW) {
    // chunks of 50 round trips until the power settles (or only the saved
    // tail of NRT is left), then the round trips the reference saves
    const int tail = std::min(d.NRT, d.Nrts);
    std::vector<double> hist(50);
    double prev = -1.0;
    int nn = 0;
    bool steady = false;
    while (nn < d.NRT - tail && !steady) {
      const int chunk = std::min(50, d.NRT - tail - nn);
      CHECK(opo_run(s, nn, chunk));
      CHECK(opo_power_history(s, hist.data(), chunk));
      nn += chunk;
      const double P = hist[chunk - 1];
      steady = chunk == 50 && prev > 0 && std::fabs(P - prev) <= converge * P;
      prev = P;
    }
    std::cout << (steady ? "Steady after " : "Not steady after ") << nn << " round trips" << std::endl;
    CHECK(opo_run(s, d.NRT - tail, tail));
  } else {
    // round-trip loop (src/cuOPO.cu:474-610) in chunks of 500 for the progress line
    for (int nn = 0; nn < d.NRT;) {
      std::cout << "#round trip: " << nn << " - Completed " << (long long)nn * 100 / d.NRT << "%" << "\t\r"
                << std::flush;
      const int chunk = std::min(500, d.NRT - nn);
      CHECK(opo_run(s, nn, chunk));
      nn += chunk;
    }
    std::cout << "#round trip: " << d.NRT - 1 << " - Completed " << (long long)(d.NRT - 1) * 100 / d.NRT << "%"
              << std::endl;
  }

  // outputs, src/cuOPO.cu:620-635
  std::cout << "\nSaving time and frequency vectors...\n" << std::endl;
  opo_cli::write_real(out + "/Tp.dat", Tp.data(), d.SIZEL);
  opo_cli::write_real(out + "/freq.dat", Fp.data(), d.SIZEL);
  opo_cli::write_real(out + "/T.dat", T.data(), N);
  // one field at a time, in the order written: each field's device-to-host
  // copy overlaps the formatting of the one before
  std::vector<double> sp(2 * d.SIZEL), ss(2 * d.SIZEL), si(d.three_eqs ? 2 * d.SIZEL : 0);
  const char *stems[3] = {"/signal_output", "/pump_output", "/idler_output"};
  double *bufs[3] = {ss.data(), sp.data(), si.data()};
  std::string fetch_err;
  const int rc = opo_cli::fetch_and_write(
      d.three_eqs ? 3 : 2,
      [&](int k) {
        const int st = opo_download_saved(s, k == 1 ? bufs[1] : nullptr, k == 0 ? bufs[0] : nullptr,
                                          k == 2 ? bufs[2] : nullptr);
        if (st != OPO_OK) fetch_err = opo_last_error();
        return st;
      },
      [&](int k) { opo_cli::write_complex(out + stems[k], bufs[k], d.SIZEL); });
  if (rc != OPO_OK) {
    std::fprintf(stderr, "Error: %s:%d, ", __FILE__, __LINE__);
    std::fprintf(stderr, "code: %d, reason: %s (%s)\n", rc, opo_strerror(rc), fetch_err.c_str());
    return 1;
  }
  CHECK(opo_destroy(s));

  const double iElaps = seconds() - iStart;
  if (iElaps > 60) std::cout << "\n\n...time elapsed " << iElaps / 60.0 << " min\n\n " << std::endl;
  else std::cout << "\n\n...time elapsed " << iElaps << " seconds\n\n " << std::endl;
  time_t current_time;
  time(&current_time);
  std::cout << ctime(&current_time) << std::endl;
  return 0;
}
