// opo_setup.cpp -- host-side physics of the cuOPO setup block in FP64:
// Sellmeier materials, phase-matching scan, derived scalars, grids, pump and
// seeded noise.  Product code (the cuOPO executable and the Python package
// call it); the oracle restates the same reference lines independently.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/opo.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 299792458.0 * 1e6 / 1e12;         // um/ps, src/cuOPO.cu:25
constexpr double kEps0 = 8.8541878128e-12 * 1e12 / 1e6;  // src/cuOPO.cu:26
constexpr double kAlphaP = 0.025e-4, kAlphaS = 0.002e-4, kAlphaI = 0.002e-4;  // ppln.cuh:24-26

// Sellmeier model: n^2 = a + sum_k B_k/(L^2 - C_k^2) + d*L^2.  Both crystals
// of the reference have this shape; the derivatives below follow from it.
struct Sellmeier {
  double a, d;
  double B[3], C2[3];
  int nterms;
};

// MgO:PPLN, Gayer et al. (src/headers/ppln.cuh:33-53)
Sellmeier ppln(double T) {
  const double f = (T - 24.5) * (T + 570.82);
  Sellmeier s{};
  s.a = 5.756 + 2.860e-6 * f;
  const double g3 = 0.2020 + 6.113e-8 * f;
  s.B[0] = 0.0983 + 4.700e-8 * f; s.C2[0] = std::pow(g3, 2);
  s.B[1] = 189.32 + 1.516e-4 * f; s.C2[1] = std::pow(12.52, 2);
  s.nterms = 2;
  s.d = -1.32e-2;
  return s;
}
// MgO:sPPLT, Bruner et al. (src/headers/spplt.cuh:31-47)
Sellmeier spplt(double T) {
  const double tk2 = std::pow(T + 273.15, 2);
  Sellmeier s{};
  s.a = 4.502483;
  s.B[0] = 0.007294 + 3.483933e-8 * tk2; s.C2[0] = std::pow(0.185087 + 1.607839e-8 * tk2, 2);
  s.B[1] = 0.073423; s.C2[1] = std::pow(0.199595, 2);
  s.B[2] = 0.001; s.C2[2] = std::pow(7.99724, 2);
  s.nterms = 3;
  s.d = -0.02357;
  return s;
}
Sellmeier material(int crystal, double T) { return crystal == OPO_PPLN ? ppln(T) : spplt(T); }

double n_of(const Sellmeier &s, double L) {
  double acc = s.a;
  for (int k = 0; k < s.nterms; ++k) acc += s.B[k] / (std::pow(L, 2) - s.C2[k]);
  return std::sqrt(acc + s.d * L * L);
}
// sum_k B_k / (L^2 - C_k^2)^p  ("resonances", ppln.cuh:57-78)
double res_of(const Sellmeier &s, double L, int p) {
  double acc = 0;
  for (int k = 0; k < s.nterms; ++k) acc += s.B[k] / std::pow(std::pow(L, 2) - s.C2[k], p);
  return acc;
}
// dn/dL = -L (res2 - d) / n   (ppln.cuh:102, spplt.cuh:84)
double dn_of(const Sellmeier &s, double L) { return -L * (res_of(s, L, 2) - s.d) / n_of(s, L); }
// d2n/dL2 (ppln.cuh:129-132, spplt.cuh:105)
double d2n_of(const Sellmeier &s, double L) {
  const double n = n_of(s, L);
  return (L * dn_of(s, L) / (n * n) - 1 / n) * (res_of(s, L, 2) - s.d) + 4 * L * L / n * res_of(s, L, 3);
}
double vg_of(const Sellmeier &s, double L) { return kC / (n_of(s, L) - L * dn_of(s, L)); }      // ppln.cuh:171-175
double gvd_of(const Sellmeier &s, double L) { return std::pow(L, 3) * d2n_of(s, L) / (2 * kPi * kC * kC); }  // :179-182

void linspace(double *x, size_t n, double lo, double hi) {  // functions.cuh:12-18
  for (size_t i = 0; i < n; ++i) x[i] = lo + (double)i * (hi - lo) / (double)(n - 1);
}

}  // namespace

extern "C" {

double opo_n(int crystal, double L, double T) { return n_of(material(crystal, T), L); }
double opo_group_velocity(int crystal, double L, double T) { return vg_of(material(crystal, T), L); }
double opo_gvd(int crystal, double L, double T) { return gvd_of(material(crystal, T), L); }

double opo_find_pm_temp(int crystal, double lp, double ls, double Lambda, double Tmin, double Tmax,
                        int pm_mode) {
  // functions.cuh:23-54: 5000-point scan, first strict minimum of |dk| wins.
  const size_t m = 5000;
  std::vector<double> T(m);
  linspace(T.data(), m, Tmin, Tmax);
  const double li = lp * ls / (ls - lp);
  size_t best = 0;
  double best_abs = 0;
  for (size_t i = 0; i < m; ++i) {
    const Sellmeier s = material(crystal, T[i]);
    double dk;
    if (pm_mode == OPO_PM_REFERENCE)
      dk = 2 * kPi / lp * std::fabs(n_of(s, ls) - n_of(s, lp)) - 2 * kPi / Lambda;
    else
      dk = 2 * kPi * (n_of(s, lp) / lp - n_of(s, ls) / ls - n_of(s, li) / li - 1 / Lambda);
    if (i == 0 || std::fabs(dk) < best_abs) { best = i; best_abs = std::fabs(dk); }
  }
  return T[best];
}

int opo_derive(const opo_args *a, opo_derived *d) {
  if (!a || !d) return OPO_ERR_INVALID;
  std::memset(d, 0, sizeof(*d));
  d->three_eqs = a->three_eqs; d->regime = a->regime; d->crystal = a->crystal;
  d->N = a->N; d->NZ = a->NZ; d->NRT = a->NRT;
  d->deff = a->crystal == OPO_PPLN ? 25.2e-6 : 11.00e-6;  // ppln.cuh:21, spplt.cuh:19
  d->dQ = 2.0 * d->deff / kPi;
  d->alpha_crp = kAlphaP; d->alpha_crs = kAlphaS; d->alpha_cri = kAlphaI;
  // wavelengths, temperature, crystal (src/cuOPO.cu:92-101)
  d->lp = a->lp_nm * 1e-3;
  d->ls = a->ls_um > 0 ? a->ls_um : 2 * d->lp;
  d->li = d->lp * d->ls / (d->ls - d->lp);
  d->Lambda = a->Lambda;
  d->Temp = opo_find_pm_temp(a->crystal, d->lp, d->ls, d->Lambda, 10, 150, a->pm_mode);
  d->Lcr = a->lcr_mm * 1e3;
  const Sellmeier s = material(a->crystal, d->Temp);
  // per wave (src/cuOPO.cu:103-119); TOD is multiplied by 0 in the reference
  d->np = n_of(s, d->lp); d->vp = vg_of(s, d->lp); d->b2p = gvd_of(s, d->lp); d->b3p = 0;
  d->kp = 2 * kPi * d->dQ / (d->np * d->lp);
  d->ns = n_of(s, d->ls); d->vs = vg_of(s, d->ls); d->b2s = gvd_of(s, d->ls); d->b3s = 0;
  d->ks = 2 * kPi * d->dQ / (d->ns * d->ls);
  if (a->three_eqs) {
    d->ni = n_of(s, d->li); d->vi = vg_of(s, d->li); d->b2i = gvd_of(s, d->li); d->b3i = 0;
    d->ki = 2 * kPi * d->dQ / (d->ni * d->li);
    d->dk = 2 * kPi * (d->np / d->lp - d->ns / d->ls - d->ni / d->li - 1 / d->Lambda);  // :123
  } else {
    d->dk = 2 * kPi * (d->np / d->lp - 2 * d->ns / d->ls - 1 / d->Lambda);  // :125
  }
  d->dkp = 1 / d->vp - 1 / d->vs;
  // cavity (src/cuOPO.cu:129-146)
  d->Lcav = 100e3;
  d->Rs = a->R_pct * 0.01;
  d->alphas = 0.5 * ((1 - d->Rs) + d->alpha_crs * d->Lcr);
  d->Ri = 0.98;
  d->alphai = 0.5 * ((1 - d->Ri) + d->alpha_cri * d->Lcr);
  d->t_rt = (d->Lcav + d->Lcr * (d->ns - 1)) / kC;
  d->FSR = 1 / d->t_rt;
  d->finesse = 2 * kPi / (1 - d->Rs);
  d->lw = d->FSR / d->finesse * 1e6;
  d->delta = a->delta;
  d->epsilon = a->gdd_pct * 0.01;
  d->GDD = -d->epsilon * d->b2s * d->Lcr;
  d->TODscomp = -0.01 * a->tod_pct * d->b3s * d->Lcr;
  d->TODicomp = -0.01 * a->tod_pct * d->b3i * d->Lcr;
  // discretisation (src/cuOPO.cu:149-167)
  d->dz = d->Lcr / a->NZ;
  d->dT = d->t_rt / a->N;
  d->dF = 1 / d->t_rt;
  d->Nrts = 64;
  if (a->regime == OPO_CW) { d->SIZEL = (size_t)a->N * d->Nrts; d->T_width = d->t_rt * d->Nrts; }
  else { d->SIZEL = (size_t)a->N * a->NRT; d->T_width = d->t_rt * a->NRT; }
  d->dFp = 1 / d->T_width;
  // threshold and pump (src/cuOPO.cu:202-234)
  d->waist = 55;
  d->spot = kPi * d->waist * d->waist;
  if (a->three_eqs)
    d->Ith = kEps0 * kC * d->np * d->ns * d->ni * d->ls * d->li * std::pow((1 / d->dQ / d->Lcr / kPi), 2) *
             d->alphas * d->alphai / 8;
  else
    d->Ith = kEps0 * kC * d->np * std::pow((d->ns * d->ls * d->alphas / d->dQ / d->Lcr / kPi), 2) / 8;
  d->Pth = d->Ith * d->spot;
  d->Nth = a->Nth;
  d->Inten = a->Nth * d->Ith;
  d->Power = d->Inten * d->spot;
  d->Ap0 = std::sqrt(2 * d->Inten / (d->np * kEps0 * kC));
  d->GammaL = a->three_eqs ? 2 * kPi * d->dQ * d->Ap0 * d->Lcr / std::sqrt(d->ns * d->ni * d->ls * d->li)
                           : 2 * kPi * d->dQ * d->Ap0 * d->Lcr / (d->ns * d->ls);
  d->FWHM = 10000;
  d->sigmap = d->FWHM * std::sqrt(2) / (2 * std::sqrt(2 * std::log(2)));
  // intracavity EOM (src/cuOPO.cu:263-269)
  d->use_eom = a->upm;
  if (a->upm) {
    if (!(a->Nth >= 1)) return OPO_ERR_INVALID;
    d->mod_depth = a->md * kPi;
    d->df = a->fm * std::sqrt(a->Nth - 1) * d->alphas / (kPi * d->mod_depth) * d->FSR;
    d->fpm = d->FSR - d->df;
  }
  return OPO_OK;
}

void opo_params_from_derived(const opo_derived *d, int precision, opo_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->three_eqs = d->three_eqs; p->regime = d->regime; p->N = d->N; p->NZ = d->NZ; p->NRT = d->NRT;
  p->precision = precision;
  p->dz = d->dz; p->dk = d->dk; p->kp = d->kp; p->ks = d->ks; p->ki = d->ki;
  p->vp = d->vp; p->vs = d->vs; p->vi = d->vi;
  p->b2p = d->b2p; p->b2s = d->b2s; p->b2i = d->b2i; p->b3p = d->b3p; p->b3s = d->b3s; p->b3i = d->b3i;
  p->alpha_crp = d->alpha_crp; p->alpha_crs = d->alpha_crs; p->alpha_cri = d->alpha_cri;
  p->Rs = d->Rs; p->Ri = d->Ri; p->delta = d->delta;
  p->GDD = d->GDD; p->TODscomp = d->TODscomp; p->TODicomp = d->TODicomp;
  p->use_eom = d->use_eom; p->mod_depth = d->mod_depth; p->fpm = d->fpm;
  p->Nrts = d->Nrts;
  p->device = 0;
}

void opo_grids(const opo_derived *d, double *T, double *w) {
  const size_t N = (size_t)d->N;
  linspace(T, N, -0.5 * d->t_rt, 0.5 * d->t_rt);  // src/cuOPO.cu:170-171
  std::vector<double> F(N);
  linspace(F.data(), N, -0.5 * N * d->dF, +0.5 * N * d->dF);  // :183-184
  const size_t c = N / 2;  // fftshift, functions.cuh:113-122
  for (size_t i = 0; i < c; ++i) { w[i + c] = F[i]; w[i] = F[i + c]; }
  for (size_t i = 0; i < N; ++i) w[i] = 2 * kPi * w[i];  // :187-189
}

void opo_grids_long(const opo_derived *d, double *Tp, double *Fp) {
  linspace(Tp, d->SIZEL, -0.5 * d->T_width, 0.5 * d->T_width);                // :175
  linspace(Fp, d->SIZEL, -0.5 * d->SIZEL * d->dFp, +0.5 * d->SIZEL * d->dFp);  // :180
}

void opo_pump_input(const opo_derived *d, double *Ap_in) {
  if (d->regime == OPO_CW) {  // functions.cuh:91-100
    for (int i = 0; i < d->N; ++i) { Ap_in[2 * i] = d->Ap0; Ap_in[2 * i + 1] = 0; }
    return;
  }
  std::vector<double> Tp(d->SIZEL);  // functions.cuh:77-85, src/cuOPO.cu:244-247
  linspace(Tp.data(), d->SIZEL, -0.5 * d->T_width, 0.5 * d->T_width);
  const double T0 = d->sigmap;
  for (size_t i = 0; i < d->SIZEL; ++i) {
    Ap_in[2 * i] = d->Ap0 * std::exp(-(Tp[i] * Tp[i]) / (2 * T0 * T0));
    Ap_in[2 * i + 1] = 0;
  }
}

// Seeded complex Gaussian noise: splitmix64 (Steele, Lea & Flood 2014) feeds
// Box-Muller.  Stream s of seed k starts at k ^ (s * 0xD1B54A32D192ED03);
// sample j consumes two outputs mapped to (0,1] by ((x >> 11) + 1) * 2^-53.
void opo_noise(uint64_t seed, uint64_t stream, double sigma, double *A, size_t n) {
  uint64_t st = seed ^ (stream * 0xD1B54A32D192ED03ull);
  auto next = [&st]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const double inv53 = 1.0 / 9007199254740992.0;
  for (size_t j = 0; j < n; ++j) {
    const double u1 = (double)((next() >> 11) + 1) * inv53;
    const double u2 = (double)((next() >> 11) + 1) * inv53;
    const double r = std::sqrt(-2.0 * std::log(u1));
    A[2 * j] = sigma * r * std::cos(2 * kPi * u2);
    A[2 * j + 1] = sigma * r * std::sin(2 * kPi * u2);
  }
}

}  // extern "C"
