// Sanitizer driver for the CLI's host code (test infrastructure): the .dat
// writer's number formatting against printf, the writer in every output
// format, on sizes around its 64 Ki-value blocks, and
// the parameter dump, and the overlapped fetch/write helper, built with -fsanitize=address,undefined by
// tests/native/Makefile.  Reads every file back and exits 0 when all values
// round-trip and no sanitizer fired.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../cuopo_amd/cli/opo_cli_common.h"

static bool check(const std::string &dir, const char *fmt, size_t n) {
  if (!opo_cli::set_format(fmt)) return false;
  std::vector<double> v(2 * n);
  for (size_t i = 0; i < v.size(); ++i) v[i] = std::sin(0.37 * (double)i) * std::pow(10.0, (double)(i % 41) - 20);
  const std::string stem = dir + "/x" + fmt + std::to_string(n);
  if (!opo_cli::write_complex(stem, v.data(), n, 3)) return false;
  const std::string f(fmt);
  for (int part = 0; part < 2; ++part) {
    const std::string base = stem + (part ? "_i" : "_r");
    if (f != "bin") {
      std::ifstream in(base + ".dat");
      double x;
      size_t k = 0;
      while (in >> x) {
        if (k >= n || x != v[2 * k + part]) return false;
        ++k;
      }
      if (k != n) return false;
    }
    if (f != "dat") {
      std::ifstream in(base + ".bin", std::ios::binary);
      std::vector<double> b(n + 1);
      in.read(reinterpret_cast<char *>(b.data()), (std::streamsize)((n + 1) * sizeof(double)));
      if ((size_t)in.gcount() != n * sizeof(double)) return false;
      for (size_t k = 0; k < n; ++k)
        if (b[k] != v[2 * k + part]) return false;
    }
  }
  return true;
}

// format_g20 against printf("%.20g"): random bit patterns (every exponent,
// the snprintf fallback included), random significands at every decimal scale
// from the smallest normal double up to beyond its exact path, powers of ten
// and their neighbours, integers, halves and dyadic fractions (exact ties at
// the 20th digit)
static bool check_format() {
  long bad = 0;
  auto chk = [&](double v) {
    char a[64], b[64];
    opo_cli::format_g20(v, a);
    std::snprintf(b, sizeof b, "%.20g", v);
    if (std::strcmp(a, b) != 0 && bad++ < 5) std::printf("format_g20(%a) = %s, printf %s\n", v, a, b);
  };
  std::mt19937_64 g(7);
  for (int i = 0; i < 200000; ++i) {
    const uint64_t u = g();
    double v;
    std::memcpy(&v, &u, sizeof v);
    chk(v);
  }
  std::uniform_real_distribution<double> unit(1.0, 10.0);
  for (int k = -20; k <= 22; ++k)
    for (int i = 0; i < 5000; ++i) {
      const double v = unit(g) * std::pow(10.0, k);
      chk(v);
      chk(-v);
    }
  for (int k = -308; k < -20; ++k)
    for (int i = 0; i < 300; ++i) chk(unit(g) * std::pow(10.0, k));
  for (int k = -308; k <= 22; ++k) {
    double lo = std::pow(10.0, k), hi = lo;
    for (int j = 0; j < 4; ++j, lo = std::nextafter(lo, 0.0), hi = std::nextafter(hi, 1e300)) {
      chk(lo);
      chk(hi);
    }
  }
  for (int i = 0; i < 100000; ++i) {
    chk((double)i);
    chk(i + 0.5);
    chk((double)(g() >> 11));
  }
  for (int j = 0; j < 70; ++j)
    for (int i = 1; i < 300; ++i) chk(std::ldexp((double)i, -j));
  for (double v : {0.0, -0.0, 5e-324, 1e-13, 1e20, 99999999999999999999.0, 1.7976931348623157e308}) chk(v);
  return bad == 0;
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  bool ok = true;
  {
    const bool c = check_format();
    std::printf("format_g20 against printf: %s\n", c ? "ok" : "bad");
    ok = ok && c;
  }
  for (const char *fmt : {"dat", "bin", "both"})
    for (size_t n : {(size_t)0, (size_t)1, (size_t)65535, (size_t)65536, (size_t)200001}) {
      const bool c = check(dir, fmt, n);
      std::printf("format=%s n=%zu: %s\n", fmt, n, c ? "ok" : "bad");
      ok = ok && c;
    }
  ok = ok && !opo_cli::set_format("csv");
  opo_derived d{};
  d.NRT = 10; d.lp = 0.532; d.ls = 1.064; d.li = 1.064; d.three_eqs = 1; d.vp = d.vs = d.vi = 100.0;
  opo_args a{};
  std::ostringstream o;
  opo_cli::print_banner(o, OPO_CW, 1);
  opo_cli::print_params(o, d, a, 50.0, 0.0, 1.0);
  ok = ok && o.str().find("Number of round trips") != std::string::npos;
  // fetch_and_write: every item written once, in order, each after its own
  // fetch; a failing fetch stops the writes there and its status comes back
  for (int fail_at : {-1, 0, 2}) {
    std::vector<int> fetched(4, 0), written;
    const int rc = opo_cli::fetch_and_write(
        4, [&](int k) { fetched[k] = 1; return k == fail_at ? OPO_ERR_HIP : OPO_OK; },
        [&](int k) { if (fetched[k]) written.push_back(k); });
    const size_t expect = fail_at < 0 ? 4 : (size_t)fail_at;
    bool c = rc == (fail_at < 0 ? OPO_OK : OPO_ERR_HIP) && written.size() == expect;
    for (size_t i = 0; i < written.size(); ++i) c = c && written[i] == (int)i;
    for (int k = fail_at + 1; fail_at >= 0 && k < 4; ++k) c = c && !fetched[k];
    std::printf("fetch_and_write fail_at=%d: %s\n", fail_at, c ? "ok" : "bad");
    ok = ok && c;
  }
  return ok ? 0 : 1;
}
