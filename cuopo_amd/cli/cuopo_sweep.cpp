// cuopo_sweep.cpp -- batched replacement for the sweep script src/cuOPO.sh.
//
// The script compiles cuOPO per crystal (cuOPO.sh:29-40), then runs it once
// per point of ten nested loops (cuOPO.sh:91-175) and moves each point's .dat
// files and its tee'd log into a folder named after the point
// (cuOPO.sh:147-164).  Here the whole grid runs as batches of instances on one
// GPU through libcuopo_amd.so, and every point gets the same folder, the same
// log file name and the same .dat files a single cuOPO run writes.
//
//   cuopo_sweep <spec> [--out=DIR] [--N=] [--nz=] [--nrt=] [--seed=] [--noise=]
//               [--ls=] [--pm=3wave] [--fp32c] [--no-inputs] [--max-batch=K]
//               [--mem-gb=G] [--format=dat|bin|both] [--dry-run]
//
// Spec file: one `KEY = v1 v2 ...` per line ('#' starts a comment), the
// arrays of cuOPO.sh:29,62-63,68-88 with their names:
//   CRYSTAL = sPPLT PPLN        REG = cw|ns        EQS = 2|3
//   LP TCR GP                   (paired by index, cuOPO.sh:69-75)
//   NN RE D GD TD UPM MD FM LCR (cuOPO.sh:80-88)
// A key may be given per crystal as `PPLN.LP = 1550` (the script's
// if/elif blocks, cuOPO.sh:68-78); the plain key is the default.
//
// Point k uses noise seed (seed + k) so a point can be re-run alone with
// cuOPO --seed=<seed+k> and give the same files.  Without --seed the base
// seed comes from the clock, as the reference's does (functions.cuh:60).
// --dry-run prints the expanded points (folder, arguments) and needs no GPU.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/opo.h"
#include "opo_cli_common.h"

namespace {

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "Error: %s\n", msg.c_str());
  std::exit(1);
}

#define CHECK(call)                                                                      \
  do {                                                                                   \
    const int st_ = (call);                                                              \
    if (st_ != OPO_OK) {                                                                 \
      std::fprintf(stderr, "Error: %s:%d, ", __FILE__, __LINE__);                        \
      std::fprintf(stderr, "code: %d, reason: %s (%s)\n", st_, opo_strerror(st_), opo_last_error()); \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

using Spec = std::map<std::string, std::vector<std::string>>;

Spec read_spec(const std::string &path) {
  std::ifstream in(path);
  if (!in) die("cannot read spec file " + path);
  Spec s;
  std::string line;
  int ln = 0;
  while (std::getline(in, line)) {
    ++ln;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) {
      if (line.find_first_not_of(" \t\r") != std::string::npos)
        die(path + ":" + std::to_string(ln) + ": expected KEY = values");
      continue;
    }
    std::istringstream k(line.substr(0, eq)), v(line.substr(eq + 1));
    std::string key, tok;
    k >> key;
    std::vector<std::string> vals;
    while (v >> tok) vals.push_back(tok);
    if (key.empty() || vals.empty()) die(path + ":" + std::to_string(ln) + ": empty key or value list");
    s[key] = vals;
  }
  return s;
}

const std::vector<std::string> &get(const Spec &s, const std::string &crystal, const std::string &key) {
  auto it = s.find(crystal + "." + key);
  if (it != s.end()) return it->second;
  it = s.find(key);
  if (it == s.end()) die("spec has no " + key + " (nor " + crystal + "." + key + ")");
  return it->second;
}

struct Point {
  std::string crystal, folder;
  std::vector<std::string> argv;  // the 12 positional arguments as written in the spec
};

// The loop nest of cuOPO.sh:29-30,91-110, innermost last.
std::vector<Point> expand(const Spec &s) {
  std::vector<Point> pts;
  for (const auto &cr : get(s, "", "CRYSTAL")) {
    if (cr != "PPLN" && cr != "sPPLT") die("CRYSTAL must be PPLN or sPPLT, got " + cr);
    const auto &LP = get(s, cr, "LP"), &TCR = get(s, cr, "TCR"), &GP = get(s, cr, "GP");
    if (TCR.size() != LP.size() || GP.size() != LP.size()) die("LP, TCR and GP must have the same length");
    const auto &NN = get(s, cr, "NN"), &RE = get(s, cr, "RE"), &D = get(s, cr, "D"), &GD = get(s, cr, "GD");
    const auto &TD = get(s, cr, "TD"), &UPM = get(s, cr, "UPM"), &MD = get(s, cr, "MD"), &FM = get(s, cr, "FM");
    const auto &LCR = get(s, cr, "LCR");
    for (const auto &lcry : LCR)
      for (size_t l = 0; l < LP.size(); ++l)
        for (const auto &u : UPM)
          for (const auto &n : NN)
            for (const auto &md : MD)
              for (const auto &fm : FM)
                for (const auto &dl : D)
                  for (const auto &re : RE)
                    for (const auto &gd : GD)
                      for (const auto &td : TD) {
                        Point p;
                        p.crystal = cr;
                        // cuOPO.sh:148-153 (the same name for both crystals)
                        p.folder = "MgO_" + cr + "_beta_" + md + "_N_" + n + "_GDDcomp_" + gd + "_TOD_comp" + td +
                                   "_LP_" + LP[l] + "nm_Lcr_" + lcry;
                        p.argv = {LP[l], TCR[l], GP[l], n, re, dl, gd, td, u, md, fm, lcry};
                        pts.push_back(p);
                      }
  }
  return pts;
}

struct Flags {
  std::string out = ".";
  int N = 1 << 14, NZ = 150, NRT = 10000;  // SIZE, NZ, NRT of src/cuOPO.cu:28-30
  uint64_t seed = 0;
  bool have_seed = false;
  double sigma = 1.0e-15;                    // functions.cuh:62
  double ls_um = 0;
  int pm_mode = OPO_PM_REFERENCE;
  int precision = OPO_FP64;
  bool inputs = true, dry = false;
  int max_batch = 64;
  double mem_gb = 160;
};

struct Prepared {
  opo_args a{};
  opo_derived d{};
  std::vector<double> T, w, Tp, Fp, Ap_in, As, Ai;
  size_t n_ap = 0;
  uint64_t seed = 0;
};

void make_dir(const std::string &path) {
  if (mkdir(path.c_str(), 0755) != 0 && errno != EEXIST) die("cannot create directory " + path);
}

double seconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::fprintf(stderr,
                 "usage: %s <spec> [--out=DIR] [--N=] [--nz=] [--nrt=] [--seed=] [--noise=] [--ls=] [--pm=3wave] "
                 "[--fp32c] [--no-inputs] [--max-batch=K] [--mem-gb=G] [--format=dat|bin|both] [--dry-run]\n",
                 argv[0]);
    return 1;
  }
  Flags fl;
  for (int k = 2; k < argc; ++k) {
    const std::string f = argv[k];
    auto val = [&](const char *key) -> const char * {
      const size_t n = std::strlen(key);
      return f.compare(0, n, key) == 0 ? f.c_str() + n : nullptr;
    };
    if (const char *v = val("--out=")) fl.out = v;
    else if (const char *v = val("--N=")) fl.N = std::atoi(v);
    else if (const char *v = val("--nz=")) fl.NZ = std::atoi(v);
    else if (const char *v = val("--nrt=")) fl.NRT = std::atoi(v);
    else if (const char *v = val("--seed=")) { fl.seed = std::strtoull(v, nullptr, 10); fl.have_seed = true; }
    else if (const char *v = val("--noise=")) fl.sigma = std::atof(v);
    else if (const char *v = val("--ls=")) fl.ls_um = std::atof(v);
    else if (const char *v = val("--max-batch=")) fl.max_batch = std::max(1, std::atoi(v));
    else if (const char *v = val("--mem-gb=")) fl.mem_gb = std::atof(v);
    else if (f == "--pm=3wave") fl.pm_mode = OPO_PM_THREE_WAVE;
    else if (f == "--fp32c") fl.precision = OPO_FP32C;
    else if (f == "--no-inputs") fl.inputs = false;
    else if (f == "--dry-run") fl.dry = true;
    else if (const char *v = val("--format=")) {
      if (!opo_cli::set_format(v)) die("--format is dat, bin or both");
    } else die("unknown flag " + f);
  }
  if (!fl.have_seed) fl.seed = (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();

  const Spec spec = read_spec(argv[1]);
  const std::string reg = get(spec, "", "REG").at(0), eqs = get(spec, "", "EQS").at(0);
  if (reg != "cw" && reg != "ns") die("REG must be cw or ns");
  if (eqs != "2" && eqs != "3") die("EQS must be 2 or 3");
  const int regime = reg == "ns" ? OPO_NS : OPO_CW, three = eqs == "3";
  const std::vector<Point> pts = expand(spec);
  std::cout << "cuopo_sweep: " << pts.size() << " points, " << (three ? 3 : 2) << " equations, "
            << (regime == OPO_NS ? "ns" : "cw") << ", N = " << fl.N << ", NZ = " << fl.NZ << ", NRT = " << fl.NRT
            << std::endl;
  if (fl.dry) {
    for (size_t k = 0; k < pts.size(); ++k) {
      std::cout << k << "\t" << pts[k].folder << "\t" << pts[k].crystal;
      for (const auto &v : pts[k].argv) std::cout << " " << v;
      std::cout << "\t--seed=" << fl.seed + k << std::endl;
    }
    return 0;
  }
  make_dir(fl.out);

  // the reference's positional arguments of point k (src/cuOPO.cu:92-267)
  auto args_of = [&](size_t k) {
    const auto &v = pts[k].argv;
    opo_args a{};
    a.three_eqs = three; a.regime = regime;
    a.crystal = pts[k].crystal == "PPLN" ? OPO_PPLN : OPO_SPPLT;
    a.N = fl.N; a.NZ = fl.NZ; a.NRT = fl.NRT;
    a.lp_nm = std::atof(v[0].c_str()); a.temp_arg = std::atof(v[1].c_str()); a.Lambda = std::atof(v[2].c_str());
    a.Nth = std::atof(v[3].c_str()); a.R_pct = std::atof(v[4].c_str()); a.delta = std::atof(v[5].c_str());
    a.gdd_pct = std::atof(v[6].c_str()); a.tod_pct = std::atof(v[7].c_str()); a.upm = std::atoi(v[8].c_str());
    a.md = std::atof(v[9].c_str()); a.fm = std::atof(v[10].c_str()); a.lcr_mm = std::atof(v[11].c_str());
    a.ls_um = fl.ls_um; a.pm_mode = fl.pm_mode;
    return a;
  };

  // per-point device bytes: state + tables (~4 NF N) + pump (N or N*NRT) +
  // save slots (NF * slots * N, slots = the derived Nrts for CW, NRT for
  // NS), complex of 16 or 8 bytes; the largest point sizes the chunks
  const double es = fl.precision == OPO_FP32C ? 8 : 16;
  double per_point = 1.0;
  for (size_t k = 0; k < pts.size(); ++k) {
    opo_args a = args_of(k);
    opo_derived d;
    CHECK(opo_derive(&a, &d));
    const double slots = regime == OPO_CW ? d.Nrts : d.NRT;
    const double pump = regime == OPO_CW ? 1.0 : d.NRT;
    per_point = std::max(per_point, es * d.N * ((three ? 3 : 2) * (4 + slots) + pump));
  }
  const int by_mem = std::max(1, (int)(fl.mem_gb * 1e9 / per_point));
  const int chunk = std::min(fl.max_batch, by_mem);

  const double t0 = seconds();
  for (size_t c0 = 0; c0 < pts.size(); c0 += (size_t)chunk) {
    const size_t c1 = std::min(pts.size(), c0 + (size_t)chunk);
    std::vector<Prepared> prep(c1 - c0);
    std::vector<opo_params> params(c1 - c0);
    for (size_t k = c0; k < c1; ++k) {
      Prepared &P = prep[k - c0];
      P.a = args_of(k);
      CHECK(opo_derive(&P.a, &P.d));
      const opo_derived &d = P.d;
      const size_t N = (size_t)d.N;
      P.T.resize(N); P.w.resize(N); P.Tp.resize(d.SIZEL); P.Fp.resize(d.SIZEL);
      opo_grids(&d, P.T.data(), P.w.data());
      opo_grids_long(&d, P.Tp.data(), P.Fp.data());
      P.n_ap = d.regime == OPO_CW ? N : d.SIZEL;
      P.Ap_in.resize(2 * P.n_ap); P.As.resize(2 * N); P.Ai.assign(2 * N, 0.0);
      opo_pump_input(&d, P.Ap_in.data());
      P.seed = fl.seed + k;
      opo_noise(P.seed, 0, fl.sigma, P.As.data(), N);
      if (three) opo_noise(P.seed, 1, fl.sigma, P.Ai.data(), N);
      opo_params_from_derived(&d, fl.precision, &params[k - c0]);
    }
    opo_batch *b = nullptr;
    CHECK(opo_batch_create(params.data(), (int)params.size(), &b));
    for (size_t k = c0; k < c1; ++k) {
      Prepared &P = prep[k - c0];
      std::vector<double> Ap(P.Ap_in.begin(), P.Ap_in.begin() + 2 * (size_t)P.d.N);
      CHECK(opo_batch_upload(b, (int)(k - c0), P.w.data(), P.T.data(), P.Ap_in.data(), P.n_ap, Ap.data(),
                             P.As.data(), three ? P.Ai.data() : nullptr));
    }
    CHECK(opo_batch_enable_save(b, 1));
    std::cout << "points " << c0 << ".." << c1 - 1 << ": running " << fl.NRT << " round trips" << std::endl;
    const double tb = seconds();
    std::vector<int> rts(c1 - c0);
    CHECK(opo_batch_run(b, fl.NRT, 0, 0.0, rts.data()));  // every round trip, as cuOPO.cu:474-610
    const double run_s = seconds() - tb;
    // outputs in loop order (a folder shared by several points ends with the
    // last one's files, as the script's mv does)
    for (size_t k = c0; k < c1; ++k) {
      Prepared &P = prep[k - c0];
      const opo_derived &d = P.d;
      const std::string dir = fl.out + "/" + pts[k].folder;
      make_dir(dir);
      // the point's log (cuOPO.sh:159 tees it, :164 moves it; a later point of
      // the same folder replaces it)
      std::ofstream log(dir + "/" + pts[k].folder + ".txt", std::ios::trunc);
      opo_cli::print_banner(log, regime, three);
      opo_cli::print_params(log, d, P.a, std::atof(pts[k].argv[6].c_str()), std::atof(pts[k].argv[7].c_str()),
                            P.Fp[d.SIZEL - 1]);
      log << "Noise seed              = " << P.seed << std::endl;
      log << "Batched run: " << c1 - c0 << " points, " << run_s << " s of GPU time for the batch" << std::endl;
      if (fl.inputs) {
        opo_cli::write_complex(dir + "/pump_input", P.Ap_in.data(), P.n_ap);
        opo_cli::write_complex(dir + "/signal_input", P.As.data(), d.N);
        if (three) opo_cli::write_complex(dir + "/idler_input", P.Ai.data(), d.N);
      }
      opo_cli::write_real(dir + "/Tp.dat", P.Tp.data(), d.SIZEL);
      opo_cli::write_real(dir + "/freq.dat", P.Fp.data(), d.SIZEL);
      opo_cli::write_real(dir + "/T.dat", P.T.data(), d.N);
      // one field at a time: each field's device-to-host copy overlaps the
      // formatting of the one before
      std::vector<double> sp(2 * d.SIZEL), ss(2 * d.SIZEL), si(three ? 2 * d.SIZEL : 0);
      const char *stems[3] = {"/signal_output", "/pump_output", "/idler_output"};
      double *bufs[3] = {ss.data(), sp.data(), si.data()};
      std::string fetch_err;
      const int rc = opo_cli::fetch_and_write(
          three ? 3 : 2,
          [&](int f) {
            const int st = opo_batch_download_saved(b, (int)(k - c0), f == 1 ? bufs[1] : nullptr,
                                                    f == 0 ? bufs[0] : nullptr, f == 2 ? bufs[2] : nullptr);
            if (st != OPO_OK) fetch_err = opo_last_error();
            return st;
          },
          [&](int f) { opo_cli::write_complex(dir + stems[f], bufs[f], d.SIZEL); });
      if (rc != OPO_OK) {
        std::fprintf(stderr, "Error: %s:%d, ", __FILE__, __LINE__);
        std::fprintf(stderr, "code: %d, reason: %s (%s)\n", rc, opo_strerror(rc), fetch_err.c_str());
        return 1;
      }
      std::cout << "  " << pts[k].folder << std::endl;
    }
    CHECK(opo_batch_destroy(b));
  }
  std::cout << "cuopo_sweep: done in " << seconds() - t0 << " s" << std::endl;
  return 0;
}
