// opo_cli_common.cpp -- .dat writer, banner and parameter dump shared by the
// cuOPO executable and the sweep driver.
#include "opo_cli_common.h"

#include <algorithm>
#include <cstdio>
#include <ctime>
#include <future>
#include <string>
#include <thread>
#include <vector>

namespace opo_cli {

namespace {
Format g_format = Format::Dat;

// raw little-endian float64, the .dat file's values in its order
bool write_bin(const std::string &path, const double *v, size_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = n == 0 || std::fwrite(v, sizeof(double), n, f) == n;  // v may be null when n == 0
  return std::fclose(f) == 0 && ok;
}

bool write_dat(const std::string &path, const double *v, size_t n, int threads);
}  // namespace

int fetch_and_write(int count, const std::function<int(int)> &fetch, const std::function<void(int)> &write) {
  std::vector<std::promise<int>> done(count > 0 ? count : 0);
  std::thread helper([&] {
    int rc = OPO_OK;
    for (int k = 0; k < count; ++k) {
      if (rc == OPO_OK) rc = fetch(k);
      done[k].set_value(rc);  // after a failure every later item reports it too
    }
  });
  int rc = OPO_OK;
  for (int k = 0; k < count && rc == OPO_OK; ++k) {
    rc = done[k].get_future().get();
    if (rc == OPO_OK) write(k);
  }
  helper.join();
  return rc;
}

bool set_format(const std::string &name) {
  if (name == "dat") g_format = Format::Dat;
  else if (name == "bin") g_format = Format::Bin;
  else if (name == "both") g_format = Format::Both;
  else return false;
  return true;
}

bool write_real(const std::string &path, const double *v, size_t n, int threads) {
  bool ok = true;
  if (g_format != Format::Bin) ok = write_dat(path, v, n, threads);
  if (g_format != Format::Dat) {
    const std::string stem = path.size() > 4 && path.compare(path.size() - 4, 4, ".dat") == 0
                                 ? path.substr(0, path.size() - 4) : path;
    ok = write_bin(stem + ".bin", v, n) && ok;
  }
  return ok;
}

namespace {
bool write_dat(const std::string &path, const double *v, size_t n, int threads) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  // format blocks of values in parallel, write the blocks in order
  const size_t block = 1 << 16;
  const size_t nblocks = (n + block - 1) / block;
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::min<size_t>((size_t)nt, std::max<size_t>(nblocks, 1));
  bool ok = true;
  for (size_t b0 = 0; b0 < nblocks && ok; b0 += (size_t)nt) {
    const size_t nb = std::min<size_t>((size_t)nt, nblocks - b0);
    std::vector<std::string> text(nb);
    auto work = [&](size_t j) {
      const size_t lo = (b0 + j) * block, hi = std::min(n, lo + block);
      std::string &s = text[j];
      s.reserve((hi - lo) * 26);
      char buf[40];
      for (size_t i = lo; i < hi; ++i) {
        const int len = std::snprintf(buf, sizeof buf, "%.20g\n", v[i]);
        s.append(buf, (size_t)len);
      }
    };
    std::vector<std::thread> pool;
    for (size_t j = 1; j < nb; ++j) pool.emplace_back(work, j);
    work(0);
    for (auto &t : pool) t.join();
    for (auto &s : text) ok = ok && std::fwrite(s.data(), 1, s.size(), f) == s.size();
  }
  return std::fclose(f) == 0 && ok;
}
}  // namespace

bool write_complex(const std::string &stem, const double *v, size_t n, int threads) {
  std::vector<double> re(n), im(n);
  for (size_t i = 0; i < n; ++i) { re[i] = v[2 * i]; im[i] = v[2 * i + 1]; }
  return write_real(stem + "_r.dat", re.data(), n, threads) && write_real(stem + "_i.dat", im.data(), n, threads);
}

void print_banner(std::ostream &o, int regime, int three_eqs) {
  o << "\n\n\n#######---Welcome to OPO simulator---#######\n\n\n" << std::endl;
  time_t now;
  time(&now);
  o << "Date: " << ctime(&now) << std::endl;
  o << (regime == OPO_NS ? "Regime: nanosecond" : "Regime: continuous wave") << std::endl;
  o << (three_eqs ? "Three equations" : "Two equations") << std::endl;
}

void print_params(std::ostream &o, const opo_derived &d, const opo_args &a, double gdd_arg, double tod_arg,
                  double max_freq) {
  const auto p0 = o.precision();
  o << "\n\nSimulation parameters:\n\n " << std::endl;
  o << "Number of round trips   = " << d.NRT << std::endl;
  o << "Pump wavelength         = " << d.lp * 1e3 << " nm" << std::endl;
  o << "Signal wavelength       = " << d.ls * 1e3 << " nm" << std::endl;
  if (d.three_eqs) o << "Idler wavelength        = " << d.li * 1e3 << " nm" << std::endl;
  o << "Temp                    = " << d.Temp << " ºC" << std::endl;
  o << "np                      = " << d.np << std::endl;
  o << "ns                      = " << d.ns << std::endl;
  if (d.three_eqs) o << "ni                      = " << d.ni << std::endl;
  o << "ν⁻¹ pump                = " << 1.0 / d.vp << " ps/μm" << std::endl;
  o << "ν⁻¹ signal              = " << 1.0 / d.vs << " ps/μm" << std::endl;
  if (d.three_eqs) o << "ν⁻¹ idler               = " << 1.0 / d.vi << " ps/μm" << std::endl;
  o << "Δk                      = " << d.dk << " μm⁻¹" << std::endl;
  o << "Δk'                     = " << d.dkp << " ps/μm" << std::endl;
  o << "GVD pump                = " << d.b2p << " ps²/μm" << std::endl;
  o << "GVD signal              = " << d.b2s << " ps²/μm" << std::endl;
  if (d.three_eqs) o << "GVD idler               = " << d.b2i << " ps²/μm" << std::endl;
  o << "GVD compensation        = " << (int)gdd_arg << " %" << std::endl;
  o << "Net GVD                 = " << (1 - d.epsilon) * d.b2s << " ps²/μm" << std::endl;
  o << "TOD pump                = " << d.b3p << " ps³/μm" << std::endl;
  o << "TOD signal              = " << d.b3s << " ps³/μm" << std::endl;
  o << "TOD compensation        = " << tod_arg << " %" << std::endl;
  o << "Net TOD                 = " << (1 - 0.01 * (int)tod_arg) * d.b3s << " ps³/μm" << std::endl;
  o << "dQ                      = " << d.dQ * 1e6 << " pm/V" << std::endl;
  o << "Λ                       = " << d.Lambda << " μm" << std::endl;
  o << "αcp                     = " << d.alpha_crp << " μm⁻¹" << std::endl;
  o << "αcs                     = " << d.alpha_crs << " μm⁻¹" << std::endl;
  o << "αs                      = " << d.alphas << std::endl;
  if (d.three_eqs) {
    o << "αci                      = " << d.alpha_cri << " μm⁻¹" << std::endl;
    o << "αi                      = " << d.alphai << std::endl;
  }
  o << "Crystal length          = " << d.Lcr * 1e-3 << " mm" << std::endl;
  o << "Cavity  length          = " << d.Lcav * 1e-3 << " mm" << std::endl;
  o << "Δz                      = " << d.dz << " μm" << std::endl;
  o << "Reflectivity (signal)   = " << d.Rs * 100 << " %" << std::endl;
  if (d.three_eqs) o << "Reflectivity (idler)    = " << d.Ri * 100 << " %" << std::endl;
  o << "Cavity Finesse          = " << d.finesse << std::endl;
  o << "Cavity lw (FWHM)        = " << d.lw << " MHz" << std::endl;
  o.precision(15);
  o << "Round-trip time         = " << d.t_rt << " ps" << std::endl;
  o << "FSR                     = " << d.FSR * 1e3 << " GHz" << std::endl;
  o << "Cavity detuning (δ)     = " << d.delta << "π" << std::endl;
  o << "Using N                 = " << d.N << " points" << std::endl;
  o << "dT                      = " << d.dT << " ps" << std::endl;
  o << "SIZEL                   = " << d.SIZEL << std::endl;
  o << "dFp                     = " << d.dFp << " THz" << std::endl;
  o << "Max frequency           = " << max_freq << " THz" << std::endl;
  o << "Ap0                     = " << d.Ap0 << " V/um" << std::endl;
  o << "waist                   = " << d.waist << " μm" << std::endl;
  o << "spot                    = " << d.spot << " μm²" << std::endl;
  o << "Power threshold         = " << d.Pth << " W" << std::endl;
  o << "Power                   = " << d.Power << " W" << std::endl;
  o << "Times above the thres.  = " << d.Nth << std::endl;
  o << "(ΓL)²                   = " << d.GammaL * d.GammaL << std::endl;
  if (d.use_eom) {
    o << "Using a phase modulator:" << std::endl;
    o << "Mod. depth (β)          = " << a.md << "π rad = " << d.mod_depth << " rad" << std::endl;
    o << "Freq. detuning (δf)     = " << d.df * 1e6 << " MHz" << std::endl;
    o << "Mod. frequency(fm)      = " << d.fpm * 1e3 << " GHz" << std::endl;
  } else {
    o << "No phase modulator" << std::endl;
  }
  o.precision(p0);
}

}  // namespace opo_cli
