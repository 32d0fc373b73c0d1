// opo_cli_common.cpp -- .dat writer, banner and parameter dump shared by the
// cuOPO executable and the sweep driver.
#include "opo_cli_common.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
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

using u128 = unsigned __int128;

// 5^q for q = 0..32: 5^32 < 2^75, so m 5^q < 2^128 for a 53-bit significand m
struct Pow5 {
  u128 p[33];
  Pow5() {
    p[0] = 1;
    for (int q = 1; q <= 32; ++q) p[q] = p[q - 1] * 5;
  }
};
const Pow5 kPow5;
const u128 kTen19 = (u128)10000000000000000000ULL, kTen20 = kTen19 * 10;

// 5^q for q = 33..343 as little-endian 64-bit limbs (5^343 < 2^797): the
// scale factors of magnitudes below 1e-13 down to the smallest normal double
constexpr int kBigQ = 343, kBigLimbs = 13;
struct BigPow5 {
  uint64_t p[kBigQ + 1][kBigLimbs];
  BigPow5() {
    uint64_t x[kBigLimbs] = {1};
    for (int q = 0; q <= kBigQ; ++q) {
      std::memcpy(p[q], x, sizeof x);
      u128 carry = 0;
      for (int i = 0; i < kBigLimbs; ++i) {
        const u128 y = (u128)x[i] * 5 + carry;
        x[i] = (uint64_t)y;
        carry = y >> 64;
      }
    }
  }
};
const BigPow5 kBigPow5;

// t = floor(m 2^e 10^q) and whether the exact rest rounds it up (to nearest,
// ties to even); false when that does not fit the integer arithmetic below
bool scale10(uint64_t m, int e, int q, u128 *t, bool *up) {
  if (q < 0 || q > kBigQ) return false;
  if (q <= 32) {
    const u128 n = (u128)m * kPow5.p[q];
    const int sh = e + q;  // m 2^e 10^q = n 2^sh
    if (sh >= 0) {
      if (sh >= 64 || (n >> (127 - sh)) != 0) return false;
      *t = n << sh;
      *up = false;
      return true;
    }
    const int s = -sh;
    if (s >= 127) return false;
    const u128 r = n & (((u128)1 << s) - 1), half = (u128)1 << (s - 1);
    *t = n >> s;
    *up = r > half || (r == half && (*t & 1));
    return true;
  }
  // q > 32: n = m 5^q in limbs, then t = n >> s with s = -(e + q) > 0
  const int s = -(e + q);
  if (s <= 0) return false;
  uint64_t n[kBigLimbs + 1];
  u128 carry = 0;
  for (int i = 0; i < kBigLimbs; ++i) {
    const u128 y = (u128)kBigPow5.p[q][i] * m + carry;
    n[i] = (uint64_t)y;
    carry = y >> 64;
  }
  n[kBigLimbs] = (uint64_t)carry;
  constexpr int kBits = 64 * (kBigLimbs + 1);
  auto word = [&](int pos) -> uint64_t {  // bits pos .. pos+63 of n
    if (pos >= kBits) return 0;
    const int li = pos >> 6, bo = pos & 63;
    uint64_t w = n[li] >> bo;
    if (bo && li + 1 <= kBigLimbs) w |= n[li + 1] << (64 - bo);
    return w;
  };
  if (s >= kBits || word(s + 128) != 0) return false;  // t would not fit 128 bits
  *t = (u128)word(s) | ((u128)word(s + 64) << 64);
  const int hb = s - 1;  // the half bit
  const bool half = (n[hb >> 6] >> (hb & 63)) & 1;
  bool below = (n[hb >> 6] & ((1ULL << (hb & 63)) - 1)) != 0;
  for (int i = 0; i < (hb >> 6) && !below; ++i) below = n[i] != 0;
  *up = half && (below || (*t & 1));
  return true;
}

// the 20 decimal digits of d (10^19 <= d < 10^20)
void digits20(u128 d, char *out) {
  uint64_t hi = (uint64_t)(d / 10000000000ULL), lo = (uint64_t)(d % 10000000000ULL);
  for (int i = 9; i >= 0; --i) { out[10 + i] = (char)('0' + lo % 10); lo /= 10; }
  for (int i = 9; i >= 0; --i) { out[i] = (char)('0' + hi % 10); hi /= 10; }
}
}  // namespace

int format_g20(double v, char *buf) {
  uint64_t bits;
  std::memcpy(&bits, &v, sizeof bits);
  const int E = (int)((bits >> 52) & 0x7ff);
  if (E == 0 || E == 0x7ff) return std::snprintf(buf, 32, "%.20g", v);  // zero, subnormal, inf, nan
  const uint64_t m = (bits & ((1ULL << 52) - 1)) | (1ULL << 52);
  const int e = E - 1075;  // |v| = m 2^e, 2^(e+52) <= |v| < 2^(e+53)
  // X = floor(log10 |v|): an estimate from the binary exponent, corrected by
  // the exact truncation (t must have exactly 20 digits)
  int X = (int)std::floor((e + 52) * 0.30102999566398120);
  u128 t = 0;
  bool up = false, found = false;
  for (int tries = 0; tries < 3 && !found; ++tries) {
    if (!scale10(m, e, 19 - X, &t, &up)) return std::snprintf(buf, 32, "%.20g", v);
    if (t < kTen19) --X;
    else if (t >= kTen20) ++X;
    else found = true;
  }
  if (!found) return std::snprintf(buf, 32, "%.20g", v);
  u128 d = t + (up ? 1 : 0);
  if (d == kTen20) { d = kTen19; ++X; }  // rounding carried into a new digit
  char dg[20];
  digits20(d, dg);
  int nd = 20;  // significant digits after dropping trailing zeros (%g without '#')
  while (nd > 1 && dg[nd - 1] == '0') --nd;
  char *o = buf;
  if (bits >> 63) *o++ = '-';
  if (X >= -4 && X < 20) {  // style f with 19 - X decimals
    if (X >= 0) {
      for (int i = 0; i <= X; ++i) *o++ = dg[i];
      if (nd > X + 1) {
        *o++ = '.';
        for (int i = X + 1; i < nd; ++i) *o++ = dg[i];
      }
    } else {
      *o++ = '0';
      *o++ = '.';
      for (int i = 0; i < -X - 1; ++i) *o++ = '0';
      for (int i = 0; i < nd; ++i) *o++ = dg[i];
    }
  } else {  // style e
    *o++ = dg[0];
    if (nd > 1) {
      *o++ = '.';
      for (int i = 1; i < nd; ++i) *o++ = dg[i];
    }
    *o++ = 'e';
    *o++ = X < 0 ? '-' : '+';
    const int ax = X < 0 ? -X : X;
    if (ax >= 100) *o++ = (char)('0' + ax / 100);
    *o++ = (char)('0' + (ax / 10) % 10);
    *o++ = (char)('0' + ax % 10);
  }
  *o = 0;
  return (int)(o - buf);
}

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
  bool ok = true, wrote = true;
  std::vector<std::string> text, out;  // the batch being formatted, the batch being written
  std::thread writer;                  // writes `out` while the next batch is formatted
  for (size_t b0 = 0; b0 < nblocks && ok; b0 += (size_t)nt) {
    const size_t nb = std::min<size_t>((size_t)nt, nblocks - b0);
    text.assign(nb, std::string());
    auto work = [&](size_t j) {
      const size_t lo = (b0 + j) * block, hi = std::min(n, lo + block);
      std::string &s = text[j];
      s.reserve((hi - lo) * 26);
      char buf[40];
      for (size_t i = lo; i < hi; ++i) {
        const int len = format_g20(v[i], buf);
        buf[len] = '\n';
        s.append(buf, (size_t)len + 1);
      }
    };
    std::vector<std::thread> pool;
    for (size_t j = 1; j < nb; ++j) pool.emplace_back(work, j);
    work(0);
    for (auto &t : pool) t.join();
    if (writer.joinable()) writer.join();
    ok = wrote;
    out.swap(text);
    writer = std::thread([&] {
      for (const auto &s : out) wrote = wrote && std::fwrite(s.data(), 1, s.size(), f) == s.size();
    });
  }
  if (writer.joinable()) writer.join();
  ok = ok && wrote;
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
