/*
 * opo_oracle.c -- FP64 CPU restatement of cuOPO's SSFM round-trip solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the package smoke
 * check and bench.py's cpu_baseline leg load this library; the product path
 * never does.  Each function cites the reference lines it restates; the
 * header lists what pins each one.  Expressions keep the reference's
 * operation order (left-to-right complex operator chains of
 * src/headers/operators.cuh:19-136) so that FP64 rounding is the only change.
 */
#include "opo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
/* src/cuOPO.cu:25-26 */
#define ORC_C (299792458.0 * 1e6 / 1e12)        /* um/ps */
#define ORC_EPS0 (8.8541878128e-12 * 1e12 / 1e6) /* W ps / (V^2 um) */

/* ------------------------------------------------------------------ */
/* complex helpers, same evaluation order as operators.cuh             */
/* ------------------------------------------------------------------ */
typedef struct { double x, y; } cx;

static inline cx cmul(cx a, cx b) { /* operators.cuh:99-106 */
  cx c; c.x = a.x * b.x - a.y * b.y; c.y = a.x * b.y + a.y * b.x; return c;
}
static inline cx cadd(cx a, cx b) { cx c; c.x = a.x + b.x; c.y = a.y + b.y; return c; }
static inline cx csub(cx a, cx b) { cx c; c.x = a.x - b.x; c.y = a.y - b.y; return c; }
static inline cx cscale(cx b, double a) { /* operators.cuh:79-96 */
  cx c; c.x = a * b.x; c.y = a * b.y; return c;
}
static inline cx cdivr(cx b, double a) { /* operators.cuh:109-116 */
  cx c; c.x = b.x / a; c.y = b.y / a; return c;
}
static inline cx cexpi(double a) { /* CpxExp, operators.cuh:120-126 */
  cx b; b.x = cos(a); b.y = sin(a); return b;
}
static inline cx cconj(cx a) { cx b; b.x = +a.x; b.y = -a.y; return b; }
static inline cx ld(const double *p, size_t i) { cx c; c.x = p[2 * i]; c.y = p[2 * i + 1]; return c; }
static inline void st(double *p, size_t i, cx c) { p[2 * i] = c.x; p[2 * i + 1] = c.y; }

/* ------------------------------------------------------------------ */
/* material: MgO:PPLN (Gayer) src/headers/ppln.cuh:21-189               */
/* ------------------------------------------------------------------ */
static const double PPLN_DEFF = 25.2e-6;    /* ppln.cuh:21 */
static const double SPPLT_DEFF = 11.00e-6;  /* spplt.cuh:19 */
static const double ALPHA_CRP = 0.025e-4;   /* ppln.cuh:24, spplt.cuh:23 */
static const double ALPHA_CRS = 0.002e-4;   /* ppln.cuh:25, spplt.cuh:24 */
static const double ALPHA_CRI = 0.002e-4;   /* ppln.cuh:26, spplt.cuh:25 */

typedef struct { double G1, G2, G3, G4, a5, a6; } ppln_coef;
static ppln_coef ppln_c(double T) { /* ppln.cuh:36-50 */
  ppln_coef c;
  double f = (T - 24.5) * (T + 570.82);
  double a1 = 5.756, a2 = 0.0983, a3 = 0.2020, a4 = 189.32;
  double b1 = 2.860e-6, b2 = 4.700e-8, b3 = 6.113e-8, b4 = 1.516e-4;
  c.a5 = 12.52; c.a6 = 1.32e-2;
  c.G1 = a1 + b1 * f; c.G2 = a2 + b2 * f; c.G3 = a3 + b3 * f; c.G4 = a4 + b4 * f;
  return c;
}
typedef struct { double A, B, C, D, E, F, G, H, b, c; } spplt_coef;
static spplt_coef spplt_c(double T) { /* spplt.cuh:34-43 */
  spplt_coef k;
  k.A = 4.502483; k.B = 0.007294; k.C = 0.185087; k.D = -0.02357;
  k.E = 0.073423; k.F = 0.199595; k.G = 0.001; k.H = 7.99724;
  k.b = 3.483933e-8 * pow(T + 273.15, 2);
  k.c = 1.607839e-8 * pow(T + 273.15, 2);
  return k;
}

double orc_n(int crystal, double L, double T) {
  if (crystal == ORC_PPLN) { /* ppln.cuh:33-53 */
    ppln_coef c = ppln_c(T);
    return sqrt(c.G1 + c.G2 / (pow(L, 2) - pow(c.G3, 2)) + c.G4 / (pow(L, 2) - pow(c.a5, 2)) - c.a6 * L * L);
  } else { /* spplt.cuh:31-47 */
    spplt_coef k = spplt_c(T);
    return sqrt(k.A + (k.B + k.b) / (pow(L, 2) - pow((k.C + k.c), 2)) + k.E / (pow(L, 2) - pow(k.F, 2)) +
                k.G / (pow(L, 2) - pow(k.H, 2)) + k.D * pow(L, 2));
  }
}

double orc_resonances(int crystal, double L, double T, int p) {
  if (crystal == ORC_PPLN) { /* ppln.cuh:57-78 */
    ppln_coef c = ppln_c(T);
    return c.G2 / pow((pow(L, 2) - pow(c.G3, 2)), p) + c.G4 / pow((pow(L, 2) - pow(c.a5, 2)), p);
  } else { /* spplt.cuh:51-66 */
    spplt_coef k = spplt_c(T);
    return (k.B + k.b) / pow((pow(L, 2) - pow((k.C + k.c), 2)), p) + k.E / pow((pow(L, 2) - pow(k.F, 2)), p) +
           k.G / pow((pow(L, 2) - pow(k.H, 2)), p);
  }
}

/* the "+a6" (PPLN) / "-D" (sPPLT) term of dn/dlambda */
static double lin_term(int crystal, double T) {
  if (crystal == ORC_PPLN) return ppln_c(T).a6;
  return -spplt_c(T).D;
}

double orc_dndl(int crystal, double L, double T) {
  /* ppln.cuh:83-104: -L*(res2 + a6)/n ; spplt.cuh:71-85: -L/n*(res2 - D) */
  if (crystal == ORC_PPLN)
    return -L * (orc_resonances(crystal, L, T, 2) + lin_term(crystal, T)) / orc_n(crystal, L, T);
  return -L / orc_n(crystal, L, T) * (orc_resonances(crystal, L, T, 2) + lin_term(crystal, T));
}

double orc_d2ndl2(int crystal, double L, double T) {
  /* ppln.cuh:109-134, spplt.cuh:90-106 */
  double n = orc_n(crystal, L, T);
  double A = (L * orc_dndl(crystal, L, T) / pow(n, 2) - 1 / n) * (orc_resonances(crystal, L, T, 2) + lin_term(crystal, T));
  double B = 4 * L * L / n * orc_resonances(crystal, L, T, 3);
  return A + B;
}

double orc_d3ndl3(int crystal, double L, double T) {
  /* ppln.cuh:139-167, spplt.cuh:111-134 (dead: multiplied by 0 in cuOPO.cu) */
  double n = orc_n(crystal, L, T);
  double d1 = orc_dndl(crystal, L, T), d2 = orc_d2ndl2(crystal, L, T);
  double A1 = (2 * d1 + L * d2) / pow(n, 2);
  double A2 = -2 * L * pow(d1, 2) / pow(n, 3);
  double AA = (A1 + A2) * (orc_resonances(crystal, L, T, 2) + lin_term(crystal, T));
  double B1 = 12 * L / n;
  double B2 = -4 * L * L / n * d2 * (1 - 1 / n);
  double BB = (B1 + B2) * orc_resonances(crystal, L, T, 3);
  double CC = -24 * L * L * L / n * orc_resonances(crystal, L, T, 4);
  return AA + BB + CC;
}

double orc_GV(int crystal, double L, double T) { /* ppln.cuh:171-175 */
  return ORC_C / (orc_n(crystal, L, T) - L * orc_dndl(crystal, L, T));
}
double orc_GVD(int crystal, double L, double T) { /* ppln.cuh:179-182 */
  return pow(L, 3) * orc_d2ndl2(crystal, L, T) / (2 * ORC_PI * ORC_C * ORC_C);
}
double orc_TOD(int crystal, double L, double T) { /* ppln.cuh:186-189 */
  return -pow(L, 4) / (4 * ORC_PI * ORC_PI * ORC_C * ORC_C * ORC_C) *
         (3 * orc_d2ndl2(crystal, L, T) + L * orc_d3ndl3(crystal, L, T));
}

/* ------------------------------------------------------------------ */
/* grids and helpers: src/headers/functions.cuh                        */
/* ------------------------------------------------------------------ */
void orc_linspace(double *x, size_t n, double xmin, double xmax) { /* functions.cuh:12-18 */
  for (size_t i = 0; i < n; i++) x[i] = xmin + (double)i * (xmax - xmin) / (double)(n - 1);
}

void orc_fftshift(double *out, const double *in, size_t n) { /* functions.cuh:113-122 */
  size_t c = n / 2;
  for (size_t i = 0; i < n / 2; i++) {
    out[i + c] = in[i];
    out[i] = in[i + c];
  }
}

double orc_find_pm_temp(int crystal, double lp, double ls, double Lambda, double Tmin, double Tmax,
                        int pm_mode) {
  /* functions.cuh:23-54 (pm_mode 0).  pm_mode 1 (opt-in extension) scans the
   * three-wave mismatch 2pi(np/lp - ns/ls - ni/li - 1/Lambda) instead. */
  const size_t m = 5000;
  double *T = (double *)malloc(m * sizeof(double));
  double *dk = (double *)malloc(m * sizeof(double));
  orc_linspace(T, m, Tmin, Tmax);
  double li = lp * ls / (ls - lp);
  for (size_t i = 0; i < m; i++) {
    if (pm_mode == ORC_PM_REFERENCE)
      dk[i] = 2 * ORC_PI / lp * fabs(orc_n(crystal, ls, T[i]) - orc_n(crystal, lp, T[i])) - 2 * ORC_PI / Lambda;
    else
      dk[i] = 2 * ORC_PI * (orc_n(crystal, lp, T[i]) / lp - orc_n(crystal, ls, T[i]) / ls -
                            orc_n(crystal, li, T[i]) / li - 1 / Lambda);
  }
  size_t pos = 0;
  for (size_t i = 1; i < m; i++)
    if (fabs(dk[i]) < fabs(dk[pos])) pos = i;
  double Temp = T[pos];
  free(T); free(dk);
  return Temp;
}

/* ------------------------------------------------------------------ */
/* setup: src/cuOPO.cu:92-269                                          */
/* ------------------------------------------------------------------ */
int orc_setup(const orc_inputs *in, orc_params *p) {
  memset(p, 0, sizeof(*p));
  p->three_eqs = in->three_eqs; p->regime = in->regime; p->crystal = in->crystal;
  p->N = in->N; p->NZ = in->NZ; p->NRT = in->NRT;
  int cr = in->crystal;
  p->deff = cr == ORC_PPLN ? PPLN_DEFF : SPPLT_DEFF;
  p->dQ = 2.0 * p->deff / ORC_PI; /* ppln.cuh:22 */
  p->alpha_crp = ALPHA_CRP; p->alpha_crs = ALPHA_CRS; p->alpha_cri = ALPHA_CRI;

  p->lp = in->lp_nm * 1e-3;                           /* cuOPO.cu:92 */
  p->ls = in->ls_um > 0 ? in->ls_um : 2 * p->lp;      /* cuOPO.cu:93 */
  p->li = p->lp * p->ls / (p->ls - p->lp);            /* cuOPO.cu:95 */
  p->Lambda = in->Lambda;                             /* cuOPO.cu:99 */
  p->Temp = orc_find_pm_temp(cr, p->lp, p->ls, p->Lambda, 10, 150, in->pm_mode); /* :100 */
  p->Lcr = in->lcr_mm * 1e3;                          /* :101 */
  double T = p->Temp;
  p->np = orc_n(cr, p->lp, T); p->vp = orc_GV(cr, p->lp, T); p->b2p = orc_GVD(cr, p->lp, T);
  p->b3p = 0. * orc_TOD(cr, p->lp, T); p->kp = 2 * ORC_PI * p->dQ / (p->np * p->lp);    /* :103-107 */
  p->ns = orc_n(cr, p->ls, T); p->vs = orc_GV(cr, p->ls, T); p->b2s = orc_GVD(cr, p->ls, T);
  p->b3s = 0. * orc_TOD(cr, p->ls, T); p->ks = 2 * ORC_PI * p->dQ / (p->ns * p->ls);    /* :109-113 */
  if (p->three_eqs) {
    p->ni = orc_n(cr, p->li, T); p->vi = orc_GV(cr, p->li, T); p->b2i = orc_GVD(cr, p->li, T);
    p->b3i = 0. * orc_TOD(cr, p->li, T); p->ki = 2 * ORC_PI * p->dQ / (p->ni * p->li);  /* :115-119 */
    p->dk = 2 * ORC_PI * (p->np / p->lp - p->ns / p->ls - p->ni / p->li - 1 / p->Lambda); /* :123 */
  } else {
    p->dk = 2 * ORC_PI * (p->np / p->lp - 2 * p->ns / p->ls - 1 / p->Lambda);             /* :125 */
  }
  p->dkp = (1 / p->vp - 1 / p->vs);                                 /* :128 */
  p->Lcav = 100e3;                                                  /* :129 */
  p->Rs = in->R_pct * 0.01;                                         /* :130 */
  p->alphas = 0.5 * ((1 - p->Rs) + p->alpha_crs * p->Lcr);          /* :131 */
  p->Ri = 0.98;                                                     /* :133 */
  p->alphai = 0.5 * ((1 - p->Ri) + p->alpha_cri * p->Lcr);          /* :134 */
  p->t_rt = (p->Lcav + p->Lcr * (p->ns - 1)) / ORC_C;               /* :136 */
  p->FSR = 1 / p->t_rt;                                             /* :137 */
  p->finesse = 2 * ORC_PI / (1 - p->Rs);                            /* :138 */
  p->lw = p->FSR / p->finesse * 1e6;                                /* :139 */
  p->delta = in->delta;                                             /* :140 */
  p->epsilon = in->gdd_pct * 0.01;                                  /* :141 */
  p->GDD = -p->epsilon * p->b2s * p->Lcr;                           /* :142 */
  p->TODscomp = -0.01 * in->tod_pct * p->b3s * p->Lcr;              /* :143 */
  p->TODicomp = -0.01 * in->tod_pct * p->b3i * p->Lcr;              /* :145 */
  p->dz = p->Lcr / p->NZ;                                           /* :149 */
  p->dT = p->t_rt / p->N;                                           /* :150 */
  p->dF = 1 / p->t_rt;                                              /* :151 */
  p->Nrts = 64;                                                     /* :154-157 */
  if (p->regime == ORC_CW) {                                        /* :160-167 */
    p->SIZEL = (size_t)p->N * p->Nrts; p->T_width = p->t_rt * p->Nrts;
  } else {
    p->SIZEL = (size_t)p->N * p->NRT; p->T_width = p->t_rt * p->NRT;
  }
  p->dFp = 1 / p->T_width;                                          /* :178 */
  p->waist = 55;                                                    /* :202 */
  p->spot = ORC_PI * p->waist * p->waist;                           /* :203 */
  if (p->three_eqs) /* DRO branch, :215-218 */
    p->Ith = ORC_EPS0 * ORC_C * p->np * p->ns * p->ni * p->ls * p->li * pow((1 / p->dQ / p->Lcr / ORC_PI), 2) *
             p->alphas * p->alphai / 8;
  else /* :222 */
    p->Ith = ORC_EPS0 * ORC_C * p->np * pow((p->ns * p->ls * p->alphas / p->dQ / p->Lcr / ORC_PI), 2) / 8;
  p->Pth = p->Ith * p->spot;                                        /* :219/223 */
  p->Nth = in->Nth;                                                 /* :226 */
  p->Inten = in->Nth * p->Ith;                                      /* :227 */
  p->Power = p->Inten * p->spot;                                    /* :228 */
  p->Ap0 = sqrt(2 * p->Inten / (p->np * ORC_EPS0 * ORC_C));         /* :229 */
  if (p->three_eqs)
    p->GammaL = 2 * ORC_PI * p->dQ * p->Ap0 * p->Lcr / sqrt(p->ns * p->ni * p->ls * p->li); /* :231 */
  else
    p->GammaL = 2 * ORC_PI * p->dQ * p->Ap0 * p->Lcr / (p->ns * p->ls);                     /* :233 */
  p->FWHM = 10000;                                                  /* :244 */
  p->sigmap = p->FWHM * sqrt(2) / (2 * sqrt(2 * log(2)));           /* :245 */
  p->use_eom = in->upm;                                             /* :263 */
  if (p->use_eom) {                                                 /* :265-269 */
    if (!(in->Nth >= 1)) return -1; /* sqrt(Nth-1) undefined: reference computes NaN */
    p->mod_depth = in->md * ORC_PI;
    p->df = in->fm * sqrt(in->Nth - 1) * p->alphas / (ORC_PI * p->mod_depth) * p->FSR;
    p->fpm = p->FSR - p->df;
  }
  return 0;
}

void orc_grids(const orc_params *p, double *T, double *w) {
  size_t N = (size_t)p->N;
  orc_linspace(T, N, -0.5 * p->t_rt, 0.5 * p->t_rt);                 /* cuOPO.cu:171 */
  double *F = (double *)malloc(N * sizeof(double));
  orc_linspace(F, N, -0.5 * N * p->dF, +0.5 * N * p->dF);            /* cuOPO.cu:184 */
  orc_fftshift(w, F, N);                                             /* cuOPO.cu:186 */
  for (size_t i = 0; i < N; i++) w[i] = 2 * ORC_PI * w[i];           /* cuOPO.cu:187-189 */
  free(F);
}

void orc_grids_long(const orc_params *p, double *Tp, double *Fp) {
  orc_linspace(Tp, p->SIZEL, -0.5 * p->T_width, 0.5 * p->T_width);              /* cuOPO.cu:175 */
  orc_linspace(Fp, p->SIZEL, -0.5 * p->SIZEL * p->dFp, +0.5 * p->SIZEL * p->dFp); /* cuOPO.cu:180 */
}

void orc_pump_input(const orc_params *p, double *Ap_in) {
  if (p->regime == ORC_CW) { /* functions.cuh:91-100 */
    for (int i = 0; i < p->N; i++) { Ap_in[2 * i] = p->Ap0; Ap_in[2 * i + 1] = 0; }
  } else { /* functions.cuh:77-85 with Tp (cuOPO.cu:244-247) */
    double *Tp = (double *)malloc(p->SIZEL * sizeof(double));
    orc_linspace(Tp, p->SIZEL, -0.5 * p->T_width, 0.5 * p->T_width);
    double T0 = p->sigmap;
    for (size_t i = 0; i < p->SIZEL; i++) {
      Ap_in[2 * i] = p->Ap0 * exp(-(Tp[i] * Tp[i]) / (2 * T0 * T0));
      Ap_in[2 * i + 1] = 0;
    }
    free(Tp);
  }
}

/* splitmix64 (Steele, Lea & Flood 2014) + Box-Muller.  Stream s of seed k
 * starts from state k ^ (s * 0xD1B54A32D192ED03); sample j takes two
 * successive outputs u1, u2 mapped to (0, 1] as ((x >> 11) + 1) * 2^-53 and
 * returns sigma*sqrt(-2 ln u1)*(cos 2 pi u2, sin 2 pi u2) as (re, im). */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
void orc_noise(uint64_t seed, uint64_t stream, double sigma, double *A, size_t n) {
  uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull);
  for (size_t j = 0; j < n; j++) {
    double u1 = (double)((splitmix64(&s) >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u2 = (double)((splitmix64(&s) >> 11) + 1) * (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(u1));
    A[2 * j] = sigma * r * cos(2 * ORC_PI * u2);
    A[2 * j + 1] = sigma * r * sin(2 * ORC_PI * u2);
  }
}

/* ------------------------------------------------------------------ */
/* FFT (replaces cufftExecC2C; cwes2.cuh:170-182).  Iterative radix-2,  */
/* twiddles e^{2 pi i k/N} from libm per k (no recurrence).            */
/* ------------------------------------------------------------------ */
static __thread double *tw_cache = NULL;
static __thread size_t tw_n = 0;

static const double *twiddles(size_t N) {
  if (tw_n != N) {
    free(tw_cache);
    tw_cache = (double *)malloc(N * sizeof(double)); /* N/2 complex */
    for (size_t k = 0; k < N / 2; k++) {
      double a = 2.0 * ORC_PI * (double)k / (double)N;
      tw_cache[2 * k] = cos(a);
      tw_cache[2 * k + 1] = sin(a);
    }
    tw_n = N;
  }
  return tw_cache;
}

void orc_fft(double *x, size_t N, int sign) {
  if (N < 2) return;
  const double *tw = twiddles(N);
  for (size_t i = 1, j = 0; i < N; i++) { /* bit reversal */
    size_t bit = N >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t0 = x[2 * i], t1 = x[2 * i + 1];
      x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
      x[2 * j] = t0; x[2 * j + 1] = t1;
    }
  }
  for (size_t len = 2; len <= N; len <<= 1) {
    size_t half = len >> 1, step = N / len;
    for (size_t i = 0; i < N; i += len) {
      for (size_t j = 0; j < half; j++) {
        double wr = tw[2 * j * step], wi = sign > 0 ? tw[2 * j * step + 1] : -tw[2 * j * step + 1];
        double *u = x + 2 * (i + j), *v = x + 2 * (i + j + half);
        double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
        double ur = u[0], ui = u[1];
        u[0] = ur + vr; u[1] = ui + vi;
        v[0] = ur - vr; v[1] = ui - vi;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* nonlinear half step: cwes2.cuh:30-87,147-166 ; cwes3.cuh:30-90,153-172 */
/* ------------------------------------------------------------------ */
typedef struct { cx p, s, i; } cx3;

static inline cx3 dAdz(const orc_params *p, cx Ap, cx As, cx Ai, double z) {
  cx Im; Im.x = 0; Im.y = 1;
  cx3 d;
  if (p->three_eqs) { /* cwes3.cuh:37-39 */
    /* operator chains Im*k*A*B*CpxExp(..) evaluate left to right */
    d.p = cmul(cmul(cmul(cscale(Im, p->kp), As), Ai), cexpi(-p->dk * z));
    d.s = cmul(cmul(cmul(cscale(Im, p->ks), Ap), cconj(Ai)), cexpi(+p->dk * z));
    d.i = cmul(cmul(cmul(cscale(Im, p->ki), Ap), cconj(As)), cexpi(+p->dk * z));
  } else { /* cwes2.cuh:37-38 */
    d.p = cmul(cmul(cmul(cscale(Im, p->kp), As), As), cexpi(+p->dk * z));
    d.s = cmul(cmul(cmul(cscale(Im, p->ks), Ap), cconj(As)), cexpi(-p->dk * z));
    d.i.x = 0; d.i.y = 0;
  }
  return d;
}

void orc_nl_half(const orc_params *p, double *Apv, double *Asv, double *Aiv, double z) {
  const double dz = p->dz, h = dz / 2; /* rk4(..., dz/2): cwes3.cuh:171 */
  const int three = p->three_eqs;
  for (int j = 0; j < p->N; j++) {
    cx Ap = ld(Apv, j), As = ld(Asv, j), Ai;
    if (three) Ai = ld(Aiv, j); else { Ai.x = 0; Ai.y = 0; }
    /* k1 = dAdz(A, z) */
    cx3 k1 = dAdz(p, Ap, As, Ai, z);
    /* aux = A + 0.5 k1 (LinealCombination); 2eq evaluates k2 on aux, 3eq on A (Q1) */
    cx3 a;
    a.p = cadd(Ap, cscale(k1.p, 0.5)); a.s = cadd(As, cscale(k1.s, 0.5)); a.i = cadd(Ai, cscale(k1.i, 0.5));
    cx3 k2 = three ? dAdz(p, Ap, As, Ai, z + dz / 4) : dAdz(p, a.p, a.s, a.i, z + dz / 4);
    a.p = cadd(Ap, cscale(k2.p, 0.5)); a.s = cadd(As, cscale(k2.s, 0.5)); a.i = cadd(Ai, cscale(k2.i, 0.5));
    cx3 k3 = three ? dAdz(p, Ap, As, Ai, z + dz / 4) : dAdz(p, a.p, a.s, a.i, z + dz / 4);
    a.p = cadd(Ap, cscale(k3.p, 1.0)); a.s = cadd(As, cscale(k3.s, 1.0)); a.i = cadd(Ai, cscale(k3.i, 1.0));
    cx3 k4 = three ? dAdz(p, Ap, As, Ai, z + dz / 2) : dAdz(p, a.p, a.s, a.i, z + dz / 2);
    /* rk4: A + (k1 + 2*k2 + 2*k3 + k4) * h / 6  (cwes3.cuh:84-86) */
    cx sp = cadd(cadd(cadd(k1.p, cscale(k2.p, 2)), cscale(k3.p, 2)), k4.p);
    cx ss = cadd(cadd(cadd(k1.s, cscale(k2.s, 2)), cscale(k3.s, 2)), k4.s);
    st(Apv, j, cadd(Ap, cdivr(cscale(sp, h), 6)));
    st(Asv, j, cadd(As, cdivr(cscale(ss, h), 6)));
    if (three) {
      cx si = cadd(cadd(cadd(k1.i, cscale(k2.i, 2)), cscale(k3.i, 2)), k4.i);
      st(Aiv, j, cadd(Ai, cdivr(cscale(si, h), 6)));
    }
  }
}

/* ------------------------------------------------------------------ */
/* linear step: cwes3.cuh:107-128,175-194 ; cwes2.cuh:103-121,170-183   */
/* ------------------------------------------------------------------ */
static void linear_field(const orc_params *p, const double *w, double *X, double D, double b2, double b3,
                         double alpha) {
  const int N = p->N;
  const double z = p->dz; /* LinearOperator(..., dz): Q8 */
  orc_fft(X, (size_t)N, +1);                                /* CUFFT_INVERSE */
  for (int j = 0; j < N; j++) st(X, j, cdivr(ld(X, j), (double)N)); /* CUFFTscale */
  double atten = exp(-0.5 * alpha * z);
  for (int j = 0; j < N; j++) {
    double wj = w[j];
    cx aux = cmul(ld(X, j), cexpi(z * wj * (D + 0.5 * wj * b2 + wj * wj * b3 / 6)));
    st(X, j, cscale(aux, atten));
  }
  orc_fft(X, (size_t)N, -1);                                /* CUFFT_FORWARD */
}

void orc_linear(const orc_params *p, const double *w, double *Ap, double *As, double *Ai) {
  if (p->three_eqs) { /* cwes3.cuh:117-119: pump uses 1/vp - 1/vi (Q5) */
    linear_field(p, w, As, (1 / p->vs - 1 / p->vs), p->b2s, p->b3s, p->alpha_crs);
    linear_field(p, w, Ai, (1 / p->vi - 1 / p->vs), p->b2i, p->b3i, p->alpha_cri);
    linear_field(p, w, Ap, (1 / p->vp - 1 / p->vi), p->b2p, p->b3p, p->alpha_crp);
  } else { /* cwes2.cuh:112-113 */
    linear_field(p, w, As, (1 / p->vs - 1 / p->vs), p->b2s, p->b3s, p->alpha_crs);
    linear_field(p, w, Ap, (1 / p->vp - 1 / p->vs), p->b2p, p->b3p, p->alpha_crp);
  }
}

void orc_crystal_pass(const orc_params *p, const double *w, double *Ap, double *As, double *Ai) {
  double z = 0; /* cwes3.cuh:149-220 */
  for (int s = 0; s < p->NZ; s++) {
    orc_nl_half(p, Ap, As, Ai, z);
    orc_linear(p, w, Ap, As, Ai);
    orc_nl_half(p, Ap, As, Ai, z); /* same z (Q3) */
    z += p->dz;
  }
}

/* ------------------------------------------------------------------ */
/* cavity boundary: functions.cuh:139-253 ; cuOPO.cu:478-608            */
/* ------------------------------------------------------------------ */
void orc_eom(const orc_params *p, const double *T, double *A) { /* functions.cuh:242-253 */
  for (int j = 0; j < p->N; j++)
    st(A, j, cmul(ld(A, j), cexpi(p->mod_depth * sin(2 * ORC_PI * p->fpm * T[j]))));
}

static void gdd_field(const orc_params *p, const double *w, double *A, double TODcomp, double b3) {
  const int N = p->N; /* cuOPO.cu:522-535 */
  orc_fft(A, (size_t)N, +1);
  for (int j = 0; j < N; j++) st(A, j, cdivr(ld(A, j), (double)N));
  for (int j = 0; j < N; j++) {
    double wj = w[j];
    if (TODcomp != 0 && b3 != 0) /* AddGDDTOD, functions.cuh:173-186 (dead: b3 = 0) */
      st(A, j, cmul(ld(A, j), cexpi(0.5 * wj * (wj * p->GDD + wj * wj * TODcomp / 3.0))));
    else /* AddGDD, functions.cuh:156-169 */
      st(A, j, cmul(ld(A, j), cexpi(0.5 * wj * wj * p->GDD)));
  }
  orc_fft(A, (size_t)N, -1);
}

void orc_gdd(const orc_params *p, const double *w, double *A) { gdd_field(p, w, A, p->TODscomp, p->b3s); }

void orc_mirror(const orc_params *p, double R, double *A) { /* functions.cuh:139-152 */
  for (int j = 0; j < p->N; j++) st(A, j, cmul(cscale(ld(A, j), sqrt(R)), cexpi(ORC_PI * p->delta)));
}

void orc_read_pump(const orc_params *p, const double *Ap_in, int nn, double *Ap) {
  const size_t N = (size_t)p->N;
  if (p->regime == ORC_CW) { memcpy(Ap, Ap_in, 2 * N * sizeof(double)); return; } /* cuOPO.cu:480 */
  size_t off; /* functions.cuh:196-210 */
  if (nn == 0) off = 0;
  else if (nn > 0 && nn < p->NRT - 1) off = (size_t)nn * N;
  else off = (N * (size_t)p->NRT - 1) - N; /* last round trip: one sample early (Q17) */
  memcpy(Ap, Ap_in + 2 * off, 2 * N * sizeof(double));
}

void orc_run(const orc_params *p, const double *w, const double *T, const double *Ap_in, double *Ap,
             double *As, double *Ai, int rt0, int nrt, double *save_p, double *save_s, double *save_i,
             double *power_out) {
  const size_t N = (size_t)p->N;
  for (int nn = rt0; nn < rt0 + nrt; nn++) { /* cuOPO.cu:474-610 */
    orc_read_pump(p, Ap_in, nn, Ap);
    orc_crystal_pass(p, w, Ap, As, Ai);
    if (p->use_eom) {                          /* :509-518 */
      orc_eom(p, T, As);
      if (p->three_eqs) orc_eom(p, T, Ai);     /* idler_pm = true (:441) */
    }
    if (p->GDD != 0) {                         /* :521-553, beta2_s for both (Q16) */
      gdd_field(p, w, As, p->TODscomp, p->b3s);
      if (p->three_eqs) gdd_field(p, w, Ai, p->TODicomp, p->b3i);
    }
    orc_mirror(p, p->Rs, As);                  /* :557-560 */
    if (p->three_eqs) orc_mirror(p, p->Ri, Ai); /* :562-565 */
    long slot = -1;                            /* :570-608 */
    if (p->regime == ORC_CW) { if (nn >= p->NRT - p->Nrts) slot = nn - (p->NRT - p->Nrts); }
    else slot = nn;
    if (slot >= 0) {
      if (save_s) memcpy(save_s + 2 * N * slot, As, 2 * N * sizeof(double));
      if (save_p) memcpy(save_p + 2 * N * slot, Ap, 2 * N * sizeof(double));
      if (save_i && p->three_eqs) memcpy(save_i + 2 * N * slot, Ai, 2 * N * sizeof(double));
    }
    if (power_out) {
      double P = 0;
      for (size_t j = 0; j < N; j++) P += As[2 * j] * As[2 * j] + As[2 * j + 1] * As[2 * j + 1];
      power_out[nn - rt0] = P;
    }
  }
}
