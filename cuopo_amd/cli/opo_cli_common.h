// opo_cli_common.h -- pieces shared by the cuOPO drop-in executable and the
// batched sweep driver: the .dat writer (src/headers/files.cuh:21-75), the
// banner and the parameter dump (src/cuOPO.cu:52-64,289-369).
#pragma once

#include <cstddef>
#include <functional>
#include <ostream>
#include <string>

#include "../../include/opo.h"

namespace opo_cli {

// Output files: the reference's text .dat files, a binary side-car per file
// (X.bin next to X.dat: raw little-endian float64, the same values in the same
// order, e.g. numpy.fromfile(path, "<f8")), or both.  Process-wide; set once
// from the command line (--format=dat|bin|both; false for any other name).
enum class Format { Dat, Bin, Both };
bool set_format(const std::string &name);

// printf("%.20g") of v into buf (at least 32 bytes, NUL-terminated); returns
// the length.  Normal values below 1e20 in magnitude take an exact integer
// path (the value's 20 leading decimal digits, rounded half to even on its
// exact binary value, as glibc does); the rest go through snprintf.
int format_g20(double v, char *buf);

// One value per line at 20 significant digits (SaveVectorReal,
// files.cuh:21-40), as format_g20 writes them.  Formatting runs on `threads` host threads (0 = all
// hardware threads), the file is written in order.  `path` names the .dat
// file; the side-car, when selected, replaces its ".dat" by ".bin".
bool write_real(const std::string &path, const double *v, size_t n, int threads = 0);
// Interleaved complex -> stem_r.dat / stem_i.dat (SaveVectorComplex,
// files.cuh:43-62).
bool write_complex(const std::string &stem, const double *v, size_t n, int threads = 0);

// Device-to-host copies overlapped with the writes: fetch(0), fetch(1), ...
// run in order on a helper thread, and write(k) runs on the calling thread as
// soon as fetch(k) has returned OPO_OK, so each copy overlaps the formatting
// of the items before it.  Returns OPO_OK, or the first failing fetch's status
// (the items from that one on are not written).  opo_last_error() is per
// thread, so a fetch that fails keeps its own message for the caller; it is
// readable once this returns.
int fetch_and_write(int count, const std::function<int(int)> &fetch, const std::function<void(int)> &write);

// "Welcome", date, regime and equation lines of src/cuOPO.cu:52-64, 75-88.
void print_banner(std::ostream &o, int regime, int three_eqs);
// The parameter dump of src/cuOPO.cu:289-369.  gdd_arg / tod_arg are argv7 /
// argv8 as given (the reference prints them back).
void print_params(std::ostream &o, const opo_derived &d, const opo_args &a, double gdd_arg, double tod_arg,
                  double max_freq);

}  // namespace opo_cli
