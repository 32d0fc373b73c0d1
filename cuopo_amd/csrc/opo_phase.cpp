// opo_phase.cpp -- the linear step's phase (cwes3.cuh:112-125 /
// cwes2.cuh:108-118) as per-field coefficients, and its factorisation into
// small tables for the row stage.  Host code, FP64 (angles in long double).
//
// The linear operator multiplies frequency q by
//   e^{i z w_q (D + w_q b2/2 + w_q^2 b3/6)} e^{-alpha z/2} / N
// (the 1/N of CUFFTscale folded in, z = dz).  On the grid of cuOPO.cu:170-189,
// w = 2 pi fftshift(linspace(-N dF/2, N dF/2, N)), w is affine in q on each
// half: w_q = alpha_h + s q with h = (q >= N/2).  Writing q = k1 + N1 k2 (the
// four-step frequency index: row k1, position k2 within the row), with
// u = alpha_h + s k1 and v = s N1, and b3 = 0 (as everywhere in the reference,
// where the third-order terms are multiplied by zero):
//   theta = [z D u + z b2 u^2/2]                        row term    R_h[k1]
//         + [z (D + b2 alpha_h) v k2 + z b2 v^2 k2^2/2] column term S[k2]
//         + z b2 s v k1 k2                              cross term  X(k1 k2)
// and h is fixed by k2 (q >= N/2 <=> k2 >= N2/2).  The row stage's thread t
// holds k2 = t + m N2/E (m < E, E = 4 or 8 samples per thread), so
// h = (m >= E/2) and X(k1 k2) = X(k1 t) X(k1 m N2/E), which gives per field
// the tables
//   Z[E k1 + m] = R_{m>=E/2}[k1] X(k1 m N2/E)   (E N1; uniform per row)
//   S[k2]                                       (N2)
//   Xhi[e >> L], Xlo[e & (2^L-1)] = X(e), e = k1 t < N/E   (two short tables)
// and the phase of element m is Z[E k1 + m] S[t + m N2/E] X(k1 t): two
// independent complex products per element, E N1 + N2 + N/(E 2^L) + 2^L
// values per field instead of N.
#include <cfloat>
#include <cmath>

#include "../../include/opo.h"
#include "opo_internal.h"

namespace opo {

int lin_coeffs(const opo_params &p, double D[3], double b2[3], double b3[3], double al[3]) {
  if (p.three_eqs) {
    // cwes3.cuh:112-125: the pump uses 1/vp - 1/vi, the idler 1/vi - 1/vs
    D[0] = (1 / p.vp - 1 / p.vi); D[1] = (1 / p.vs - 1 / p.vs); D[2] = (1 / p.vi - 1 / p.vs);
    b2[0] = p.b2p; b2[1] = p.b2s; b2[2] = p.b2i; b3[0] = p.b3p; b3[1] = p.b3s; b3[2] = p.b3i;
    al[0] = p.alpha_crp; al[1] = p.alpha_crs; al[2] = p.alpha_cri;
    return 3;
  }
  // cwes2.cuh:108-118
  D[0] = (1 / p.vp - 1 / p.vs); D[1] = (1 / p.vs - 1 / p.vs);
  b2[0] = p.b2p; b2[1] = p.b2s; b3[0] = p.b3p; b3[1] = p.b3s;
  al[0] = p.alpha_crp; al[1] = p.alpha_crs;
  D[2] = b2[2] = b3[2] = al[2] = 0;
  return 2;
}

bool affine_grid(const double *w, int N, double &a0, double &a1, double &s) {
  if (N < 4) return false;
  const int H = N / 2;
  s = (w[H - 1] - w[0]) / (H - 1);
  a0 = w[0];
  a1 = w[H] - s * H;
  double wmax = 0;
  for (int q = 0; q < N; ++q) wmax = std::fmax(wmax, std::fabs(w[q]));
  const double tol = 64 * DBL_EPSILON * wmax;
  for (int q = 0; q < N; ++q) {
    const double e = (q < H ? a0 : a1) + s * q;
    if (!(std::fabs(w[q] - e) <= tol)) return false;
  }
  return true;
}

size_t lin_factor_len(int N, int N1, int xl_bits, int E) {
  return (size_t)E * N1 + (size_t)(N / N1) + (((size_t)N / E) >> xl_bits) + ((size_t)1 << xl_bits);
}

int lin_factors(const opo_params &p, const double *w, int N1, int xl_bits, int E, double *out) {
  const int N = p.N, N2 = N / N1;
  if ((E != 4 && E != 8) || N2 < E || N2 % E != 0 || (size_t)(N / E) % ((size_t)1 << xl_bits) != 0) return 0;
  const int T2 = N2 / E;
  double D[3], b2[3], b3[3], al[3];
  const int NF = lin_coeffs(p, D, b2, b3, al);
  for (int f = 0; f < NF; ++f)
    if (b3[f] != 0) return 0;
  double a0, a1, s;
  if (!affine_grid(w, N, a0, a1, s)) return 0;
  const size_t FB = lin_factor_len(N, N1, xl_bits, E);
  const long double z = p.dz, v = (long double)s * N1;
  auto put = [](double *o, long double th, long double amp) {
    o[0] = (double)(amp * cosl(th));
    o[1] = (double)(amp * sinl(th));
  };
  for (int f = 0; f < NF; ++f) {
    double *o = out + 2 * FB * f;
    const long double Df = D[f], bf = b2[f];
    const long double amp = expl(-0.5L * al[f] * z) / N;
    const long double beta = z * bf * (long double)s * v;  // cross term coefficient
    for (int k1 = 0; k1 < N1; ++k1)
      for (int m = 0; m < E; ++m) {
        const long double u = (2 * m >= E ? a1 : a0) + (long double)s * k1;
        const long double row = z * Df * u + 0.5L * z * bf * u * u;
        put(o + 2 * ((size_t)E * k1 + m), row + beta * (long double)k1 * m * T2, amp);
      }
    o += 2 * (size_t)E * N1;
    for (int k2 = 0; k2 < N2; ++k2) {
      const long double ah = k2 >= N2 / 2 ? a1 : a0;
      put(o + 2 * (size_t)k2, z * (Df + bf * ah) * v * k2 + 0.5L * z * bf * v * v * (long double)k2 * k2, 1.0L);
    }
    o += 2 * (size_t)N2;
    const size_t nh = ((size_t)N / E) >> xl_bits, nl = (size_t)1 << xl_bits;
    for (size_t j = 0; j < nh; ++j) put(o + 2 * j, beta * (long double)(j << xl_bits), 1.0L);
    o += 2 * nh;
    for (size_t l = 0; l < nl; ++l) put(o + 2 * l, beta * (long double)l, 1.0L);
  }
  return 1;
}

}  // namespace opo

extern "C" int opo_linear_factors(const opo_params *p, const double *w, int N1, int E, int xl_bits, double *out) {
  if (!p || !w || !out || N1 < 2 || p->N % N1 != 0 || (E != 4 && E != 8) || xl_bits < 0 || xl_bits > 30)
    return OPO_ERR_INVALID;
  return opo::lin_factors(*p, w, N1, xl_bits, E, out);
}
