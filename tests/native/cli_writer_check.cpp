// Sanitizer driver for the CLI's host code (test infrastructure): the .dat
// writer in every output format, on sizes around its 64 Ki-value blocks, and
// the parameter dump, and the overlapped fetch/write helper, built with -fsanitize=address,undefined by
// tests/native/Makefile.  Reads every file back and exits 0 when all values
// round-trip and no sanitizer fired.
#include <cmath>
#include <cstdio>
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

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  bool ok = true;
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
