/*
 * opo_oracle.h -- FP64 CPU restatement of cuOPO's split-step Fourier round-trip
 * solver.  TEST INFRASTRUCTURE ONLY: tests/, the package smoke check and the
 * CPU-baseline leg of bench.py may call this library (to compare results
 * against, or to time as the CPU baseline); the product path (libcuopo_amd.so,
 * the cuOPO executable) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the cuOPO source tree).  The reference computes in FP32 with cuFFT; this
 * restatement evaluates the same expressions, in the same operation order, in
 * FP64 with a radix-2 FFT of its own.
 *
 * Pinning (what checks each function, all in tests/):
 *   orc_n/orc_dndl/orc_d2ndl2 ...... finite differences of orc_n; QPM periods the
 *                                     reference sweep script pairs with
 *                                     temperatures (src/cuOPO.sh:69-75)
 *   orc_find_pm_temp ............... same (Λ, T) pairs; brute scan in numpy
 *   orc_setup ...................... closed forms restated independently in numpy
 *   orc_linspace/orc_fftshift/grids  closed form of the N/(N-1)-stretched grid
 *   orc_fft ........................ brute O(N^2) DFT, numpy.fft, Parseval,
 *                                     impulse/tone known answers, round trip
 *   orc_nl_half (3 eqs) ............ closed form A + (h/6) i k P(A) (e0+4e1+e2)
 *   orc_nl_half (2 eqs) ............ brute-force stage chain in numpy
 *   orc_linear ..................... single-tone eigenvector closed form
 *   orc_eom/orc_gdd/orc_mirror ..... closed forms (tone for GDD)
 *   orc_read_pump/orc_save ......... index arithmetic incl. last-slot shift
 *   orc_crystal_pass/orc_run ....... independent numpy restatement, N<=256
 *   orc_noise ...................... moments + independent numpy splitmix64
 * d3ndl3/TOD are multiplied by zero in the reference (src/cuOPO.cu:106,112,118)
 * and are restated for completeness only: parity unpinned (dead code).
 */
#ifndef OPO_ORACLE_H
#define OPO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PPLN = 0, ORC_SPPLT = 1 };
enum { ORC_CW = 0, ORC_NS = 1 };
enum { ORC_PM_REFERENCE = 0, ORC_PM_THREE_WAVE = 1 };

/* The 12 positional arguments of the reference executable (src/cuOPO.cu:92-267)
 * plus the compile-time switches and the opt-in extensions. */
typedef struct {
  int three_eqs;   /* -DTHREE_EQS */
  int regime;      /* ORC_CW (-DCW_OPO) / ORC_NS (-DNS_OPO) */
  int crystal;     /* ORC_PPLN / ORC_SPPLT */
  int N, NZ, NRT;  /* SIZE, NZ, NRT (src/cuOPO.cu:28-30) */
  double lp_nm;    /* argv1 */
  double temp_arg; /* argv2, ignored by the reference */
  double Lambda;   /* argv3 [um] */
  double Nth;      /* argv4 */
  double R_pct;    /* argv5 */
  double delta;    /* argv6, units of pi */
  double gdd_pct;  /* argv7 */
  double tod_pct;  /* argv8 */
  int upm;         /* argv9 */
  double md;       /* argv10, units of pi */
  double fm;       /* argv11 */
  double lcr_mm;   /* argv12 */
  double ls_um;    /* 0 => reference ls = 2*lp (src/cuOPO.cu:93) */
  int pm_mode;     /* ORC_PM_REFERENCE: SHG-style scan (functions.cuh:23-54) */
} orc_inputs;

/* Derived scalars (src/cuOPO.cu:92-269). */
typedef struct {
  int three_eqs, regime, crystal, N, NZ, NRT;
  double lp, ls, li, Lambda, Temp, Lcr;
  double np, ns, ni, vp, vs, vi, b2p, b2s, b2i, b3p, b3s, b3i, kp, ks, ki;
  double dk, dkp, Lcav, Rs, Ri, alphas, alphai, t_rt, FSR, finesse, lw;
  double delta, epsilon, GDD, TODscomp, TODicomp, dz, dT, dF;
  double deff, dQ, alpha_crp, alpha_crs, alpha_cri;
  double waist, spot, Ith, Pth, Nth, Inten, Power, Ap0, GammaL;
  int use_eom; double mod_depth, df, fpm;
  int Nrts; size_t SIZEL; double T_width, dFp, FWHM, sigmap;
} orc_params;

/* ---- material (src/headers/ppln.cuh:33-189, spplt.cuh:31-154) ---- */
double orc_n(int crystal, double L, double T);
double orc_resonances(int crystal, double L, double T, int p);
double orc_dndl(int crystal, double L, double T);
double orc_d2ndl2(int crystal, double L, double T);
double orc_d3ndl3(int crystal, double L, double T);
double orc_GV(int crystal, double L, double T);
double orc_GVD(int crystal, double L, double T);
double orc_TOD(int crystal, double L, double T);
/* functions.cuh:23-54 (pm_mode 0); pm_mode 1 scans the three-wave mismatch. */
double orc_find_pm_temp(int crystal, double lp, double ls, double Lambda,
                        double Tmin, double Tmax, int pm_mode);

/* ---- setup ---- */
int orc_setup(const orc_inputs *in, orc_params *p);
void orc_linspace(double *x, size_t n, double xmin, double xmax);  /* functions.cuh:12-18 */
void orc_fftshift(double *out, const double *in, size_t n);        /* functions.cuh:113-122 */
/* T (N) and w (N): src/cuOPO.cu:170-189 */
void orc_grids(const orc_params *p, double *T, double *w);
/* Tp, Fp (SIZEL): src/cuOPO.cu:173-180 */
void orc_grids_long(const orc_params *p, double *Tp, double *Fp);
/* Pump: CW functions.cuh:91-100 (n = N); NS functions.cuh:77-85 (n = SIZEL). */
void orc_pump_input(const orc_params *p, double *Ap_in /* 2*n */);
/* Seeded replacement of NoiseGeneratorCPU (functions.cuh:58-71). */
void orc_noise(uint64_t seed, uint64_t stream, double sigma, double *A, size_t n);

/* ---- FFT: unnormalised, sign +1 = sum x e^{+2 pi i jk/N} (CUFFT_INVERSE),
 *      sign -1 = CUFFT_FORWARD.  N a power of two. ---- */
void orc_fft(double *x, size_t N, int sign);

/* ---- propagation (all fields interleaved complex128, length N) ---- */
/* One RK "half step" (cwes2.cuh:147-166 / cwes3.cuh:153-172); z as passed. */
void orc_nl_half(const orc_params *p, double *Ap, double *As, double *Ai, double z);
/* Linear step for dz: IFFT+/N, LinearOperator, FFT- (cwes*.cuh:170-194). */
void orc_linear(const orc_params *p, const double *w, double *Ap, double *As, double *Ai);
/* EvolutionInCrystal (cwes2.cuh:132-215, cwes3.cuh:139-225). */
void orc_crystal_pass(const orc_params *p, const double *w, double *Ap, double *As, double *Ai);
/* PhaseModulatorIntraCavity (functions.cuh:242-253). */
void orc_eom(const orc_params *p, const double *T, double *A);
/* GDD block (cuOPO.cu:521-553, functions.cuh:156-169). */
void orc_gdd(const orc_params *p, const double *w, double *A);
/* AddPhase (functions.cuh:139-152). */
void orc_mirror(const orc_params *p, double R, double *A);
/* ReadPump (functions.cuh:191-213): slot nn of Ap_in into Ap. */
void orc_read_pump(const orc_params *p, const double *Ap_in, int nn, double *Ap);

/* Round-trip loop rt0 .. rt0+nrt-1 (cuOPO.cu:474-610).  save_* may be NULL;
 * when given they have Nrts*N (CW) or NRT*N (NS) complex entries and receive
 * the reference's save slots.  power_out (may be NULL) receives sum |As|^2
 * after each round trip (nrt values). */
void orc_run(const orc_params *p, const double *w, const double *T,
             const double *Ap_in, double *Ap, double *As, double *Ai,
             int rt0, int nrt, double *save_p, double *save_s, double *save_i,
             double *power_out);

#ifdef __cplusplus
}
#endif
#endif
