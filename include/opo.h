/*
 * opo.h -- C ABI of libcuopo_amd.so, the MI355X (gfx950) split-step Fourier
 * round-trip solver for the three-wave-mixing coupled-wave equations.
 *
 * Plain C types only: sizes are int/size_t, fields are interleaved complex128
 * (double re, double im -- the layout of std::complex<double>), host buffers
 * are caller-owned and copied, device buffers are owned by the handle.  The
 * library never exits the process: every entry point returns an opo_status
 * (the cuOPO executable maps a non-zero status to the reference's
 * "Error: file:line" print and exit(1), src/headers/common.cuh:6-16).
 *
 * Threads: each handle (solver, batch, MLP) owns its own HIP stream and
 * device memory, and the library keeps no other mutable state, so distinct
 * handles may be driven from different host threads at the same time.  Calls
 * on one handle must not overlap.  opo_last_error() reports the calling
 * thread's last failure.  (The reference runs one host thread on the legacy
 * default stream of device 0, src/cuOPO.cu:78-82.)
 *
 * Reference interfaces replaced (paths relative to the cuOPO source tree):
 *   opo_derive ........... setup block of main(), src/cuOPO.cu:92-269
 *   opo_grids/_long ...... linspace+fftshift grids, src/cuOPO.cu:170-189
 *   opo_pump_input ....... input_field_T, src/headers/functions.cuh:77-100
 *   opo_noise ............ NoiseGeneratorCPU, src/headers/functions.cuh:58-71
 *   opo_create/upload .... device allocation + H2D copies, src/cuOPO.cu:379-442
 *   opo_crystal_pass ..... EvolutionInCrystal, src/headers/cwes3.cuh:139-225
 *                          and src/headers/cwes2.cuh:132-215
 *   opo_linear_step ...... IFFT/CUFFTscale/LinearOperator/FFT, cwes3.cuh:175-194
 *   opo_nl_half .......... dAdz/LinealCombination/rk4 chain, cwes3.cuh:153-172
 *   opo_boundary ......... EOM + GDD + AddPhase, src/cuOPO.cu:509-566
 *   opo_run .............. round-trip loop, src/cuOPO.cu:474-610
 *   opo_download* ........ SaveVectorComplexGPU D2H, src/headers/files.cuh:65-75
 *   opo_destroy .......... teardown, src/cuOPO.cu:645-668
 *   opo_batch_* .......... the serial sweep of src/cuOPO.sh:91-175, batched
 * Entry points marked "no reference counterpart" are additions.
 */
#ifndef OPO_H
#define OPO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OPO_OK = 0,
  OPO_ERR_INVALID = -1,  /* bad argument or unsupported size */
  OPO_ERR_HIP = -2,      /* HIP runtime error (message via opo_last_error) */
  OPO_ERR_NOMEM = -3,    /* device allocation failed */
  OPO_ERR_STATE = -4,    /* call out of order (e.g. run before upload) */
  OPO_ERR_DIVERGED = -5  /* non-finite intracavity power detected */
} opo_status;

enum { OPO_PPLN = 0, OPO_SPPLT = 1 };        /* -DPPLN / -DSPPLT   (src/cuOPO.cu:36-41) */
enum { OPO_CW = 0, OPO_NS = 1 };             /* -DCW_OPO / -DNS_OPO (src/cuOPO.cu:65-70) */
enum { OPO_FP64 = 0, OPO_FP32C = 1 };        /* FP32-complex variant: no reference counterpart */
enum { OPO_PM_REFERENCE = 0, OPO_PM_THREE_WAVE = 1 };

/* argv[1..12] of the reference executable plus its compile-time switches
 * (src/cuOPO.cu:36-47,92-267) and the opt-in extensions (ls, pm_mode). */
typedef struct {
  int three_eqs, regime, crystal;
  int N, NZ, NRT;
  double lp_nm, temp_arg, Lambda, Nth, R_pct, delta, gdd_pct, tod_pct;
  int upm;
  double md, fm, lcr_mm;
  double ls_um;  /* 0 => ls = 2*lp as in the reference (src/cuOPO.cu:93) */
  int pm_mode;   /* OPO_PM_REFERENCE => functions.cuh:23-54 scan */
} opo_args;

/* Derived physical scalars (src/cuOPO.cu:92-269). */
typedef struct {
  int three_eqs, regime, crystal, N, NZ, NRT;
  double lp, ls, li, Lambda, Temp, Lcr;
  double np, ns, ni, vp, vs, vi, b2p, b2s, b2i, b3p, b3s, b3i, kp, ks, ki;
  double dk, dkp, Lcav, Rs, Ri, alphas, alphai, t_rt, FSR, finesse, lw;
  double delta, epsilon, GDD, TODscomp, TODicomp, dz, dT, dF;
  double deff, dQ, alpha_crp, alpha_crs, alpha_cri;
  double waist, spot, Ith, Pth, Nth, Inten, Power, Ap0, GammaL;
  int use_eom;
  double mod_depth, df, fpm;
  int Nrts;
  size_t SIZEL;
  double T_width, dFp, FWHM, sigmap;
} opo_derived;

/* What the solver needs: the scalar arguments of EvolutionInCrystal
 * (cwes3.cuh:139) plus the boundary constants of the round-trip loop. */
typedef struct {
  int three_eqs, regime, N, NZ, NRT;
  int precision;              /* OPO_FP64 or OPO_FP32C */
  double dz, dk, kp, ks, ki;
  double vp, vs, vi, b2p, b2s, b2i, b3p, b3s, b3i;
  double alpha_crp, alpha_crs, alpha_cri;
  double Rs, Ri, delta;
  double GDD, TODscomp, TODicomp;
  int use_eom;
  double mod_depth, fpm;
  int Nrts;                   /* CW: number of trailing round trips saved (64) */
  int device;                 /* HIP device ordinal */
} opo_params;

/* ---------------- host-side setup (product code, FP64) ---------------- */
/* Returns OPO_ERR_INVALID when the EOM is on and Nth < 1 (the reference takes
 * sqrtf(Nth-1), src/cuOPO.cu:267). */
int opo_derive(const opo_args *a, opo_derived *d);
void opo_params_from_derived(const opo_derived *d, int precision, opo_params *p);
void opo_grids(const opo_derived *d, double *T /*N*/, double *w /*N*/);
void opo_grids_long(const opo_derived *d, double *Tp /*SIZEL*/, double *Fp /*SIZEL*/);
/* CW: N complex; NS: SIZEL complex */
void opo_pump_input(const opo_derived *d, double *Ap_in);
/* splitmix64 + Box-Muller, seeded (replaces the clock-seeded generator) */
void opo_noise(uint64_t seed, uint64_t stream, double sigma, double *A, size_t n);
/* Phase-matching temperature scan (functions.cuh:23-54). */
double opo_find_pm_temp(int crystal, double lp, double ls, double Lambda, double Tmin, double Tmax,
                        int pm_mode);
/* Factorised phase of the linear step (cwes3.cuh:112-125 / cwes2.cuh:108-118)
 * that the row stage uses on the reference's frequency grid, for a row FFT of
 * E = 4 or 8 samples per thread.  Per field: Z (E*N1), S (N/N1),
 * Xhi ((N/E) >> xl_bits), Xlo (1 << xl_bits) complex values with, for
 * frequency q = k1 + N1 k2, k2 = t + m N/(E N1) (m < E), e = k1 t:
 * e^{i phase_q} e^{-alpha dz/2}/N =
 * Z[E k1 + m] S[k2] Xhi[e >> xl_bits] Xlo[e & (2^xl_bits - 1)].
 * Returns 1 (out filled), 0 (w is not the affine two-half grid or a
 * third-order term is nonzero: the solver then uses the full phase table) or
 * OPO_ERR_INVALID.  No reference counterpart. */
int opo_linear_factors(const opo_params *p, const double *w, int N1, int E, int xl_bits, double *out);
/* Refractive index and dispersion of the crystal (ppln.cuh / spplt.cuh). */
double opo_n(int crystal, double L, double T);
double opo_group_velocity(int crystal, double L, double T);
double opo_gvd(int crystal, double L, double T);

/* ---------------- single-instance solver ---------------- */
typedef struct opo_solver opo_solver;

int opo_create(const opo_params *p, opo_solver **out);
/* w, T: N doubles (angular-frequency and time grids).  Ap_in: n_ap complex
 * (N for CW, N*NRT for NS).  Ap/As/Ai: N complex initial fields (Ai ignored
 * for two equations, Ap may be NULL => zeros). */
int opo_upload(opo_solver *s, const double *w, const double *T, const double *Ap_in, size_t n_ap,
               const double *Ap, const double *As, const double *Ai);
int opo_set_fields(opo_solver *s, const double *Ap, const double *As, const double *Ai);
int opo_crystal_pass(opo_solver *s);
int opo_linear_step(opo_solver *s);
/* n_times applications of the nonlinear half step of slice `slice` (z as in
 * the reference's accumulated z of that slice). */
int opo_nl_half(opo_solver *s, int slice, int n_times);
int opo_boundary(opo_solver *s);
/* Round trips rt0 .. rt0+n_rt-1.  Saves slots per the reference when the
 * solver has save buffers (opo_enable_save). */
int opo_run(opo_solver *s, int rt0, int n_rt);
int opo_enable_save(opo_solver *s, int enable);
int opo_download(opo_solver *s, double *Ap, double *As, double *Ai);
/* slots*N complex each (slots = Nrts for CW, NRT for NS); any may be NULL */
int opo_download_saved(opo_solver *s, double *save_p, double *save_s, double *save_i);
/* sum |As|^2 after each of the last run's round trips: n >= 0 values, those
 * past the last run's round-trip count are 0 */
int opo_power_history(opo_solver *s, double *out, int n);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): opo_run of one instance on caller-owned device buffers,
 * calling no hipMalloc/hipFree and creating no stream, event or graph.
 *   params        as for opo_create (validated the same way)
 *   w, T          host grids, N doubles each (T may be NULL without the EOM)
 *   d_pump        device pump, n_ap complex (N for CW, N*NRT for NS); read only
 *   d_fields      device fields [Ap | As (| Ai)], N complex each, time order,
 *                 updated in place (complex128 for FP64, complex64 for FP32C)
 *   rt0, n_rt     round trips rt0 .. rt0+n_rt-1 (n_rt >= 1)
 *   power_out     host, n_rt doubles (sum |As|^2 after each round trip), or NULL
 *   stream        hipStream_t to launch on (NULL: the null stream); the call
 *                 returns after the stream has finished the work
 * Every other device buffer (phase, GDD and EOM tables, twiddles, slice
 * coefficients, counters, power history) lives in one static 8 MiB array per
 * GPU, held until the process ends.  Requests that do not fit return
 * OPO_ERR_INVALID; FP64 three equations with the EOM and GDD tables fit up to
 * N = 2^16 (FP32C 2^17).  No convergence test and no save slots.  Calls that
 * overlap on one GPU (two host threads, or work still queued on another
 * stream) share that array and race: serialise them. */
int opo_run_noalloc(const opo_params *params, const double *w, const double *T, const void *d_pump, size_t n_ap,
                    void *d_fields, int rt0, int n_rt, double *power_out, void *stream);
/* Selects kernel tiles: col_threads (per field) / row_threads in
 * {32,64,128,256,512}; 0 = default (128-byte column tiles, 64-thread row tiles).
 * No reference counterpart (autotuner hook). */
int opo_set_tiles(opo_solver *s, int col_threads, int row_threads);
/* Launch settings of the instances of one grid size (no reference
 * counterpart).  Inputs: col_threads / row_threads (threads per field of a
 * column tile / per row tile; 0 = default), xcd_pack (-1 default, 0, 1),
 * phase_mode (-1 auto, 0 full phase table, 1 factorised tables),
 * col_spt / row_spt (samples per thread of the column / row FFTs: 8 =
 * radix-8 passes, 4 = radix-4 passes with twice the threads; 0 = 8).
 * Outputs of get: N, count, col_tile / row_tile (columns / rows per block),
 * factored (the factorised phase is in use) and slice_us (time of one slice,
 * both stages back to back, of the autotuner's choice; 0 if not tuned). */
typedef struct {
  int N, count;
  int col_threads, row_threads, xcd_pack, phase_mode, col_spt, row_spt;
  int col_tile, row_tile, factored;
  double slice_us;
} opo_tuning;
/* Times the candidate settings on the uploaded instances (iters back-to-back
 * slices per candidate) and keeps the fastest; the fields are restored and
 * the round-trip counters reset afterwards. */
int opo_autotune(opo_solver *s, int iters);
int opo_get_tuning(opo_solver *s, opo_tuning *out);
int opo_set_tuning(opo_solver *s, const opo_tuning *t);
/* Launch every stage on one XCD (8x grid, 7 of 8 blocks exit at once), so a
 * small instance's stage-to-stage data stays in that XCD's L2.  Autotuner
 * knob; no reference counterpart. */
int opo_set_xcd_pack(opo_solver *s, int enable);
/* Capture one steady round trip in a hipGraph and replay it (default on).
 * No reference counterpart. */
int opo_use_graph(opo_solver *s, int enable);
/* Device time of the last opo_run in milliseconds (hipEvents on the solver's
 * stream).  No reference counterpart. */
int opo_last_run_ms(opo_solver *s, double *ms);
/* Measurement hook: average device time (microseconds) of one column-stage
 * and one row-stage launch of a crystal slice, each timed alone with
 * start/stop events on the solver's stream (iters launches each).  Leaves the
 * fields undefined.  No reference counterpart. */
int opo_time_stages(opo_solver *s, int iters, double *col_us, double *row_us);
int opo_destroy(opo_solver *s);

/* ---------------- batched sweeps (no single reference counterpart: the
 * serial loop of src/cuOPO.sh:91-175 run as one launch) ---------------- */
typedef struct opo_batch opo_batch;

/* count instances; each params[k] may have its own N (grouped internally),
 * all must share three_eqs, regime, NZ and precision. */
int opo_batch_create(const opo_params *params, int count, opo_batch **out);
int opo_batch_upload(opo_batch *b, int idx, const double *w, const double *T, const double *Ap_in,
                     size_t n_ap, const double *Ap, const double *As, const double *Ai);
/* Runs round trips until every instance has converged (relative change of
 * sum |As|^2 over `check_every` round trips below tol) or max_rt is reached.
 * rts_out[k] receives the round trips instance k ran (its convergence point),
 * or -1 if it did not converge.  tol <= 0 runs exactly max_rt round trips. */
int opo_batch_run(opo_batch *b, int max_rt, int check_every, double tol, int *rts_out);
/* (opo_batch_run applies the off test below with off_ratio = 1e-6.)
 * As opo_batch_run, and an instance whose power has fallen below
 * off_ratio x its power at the first check (and is still falling) is stopped
 * as "off" (below threshold).  states_out[k]: 1 steady state, 2 off, 0 not
 * converged, -1 diverged; rts_out[k] the round trips it ran: its convergence
 * point, -1 if it did not converge, or for a diverged instance the round trip
 * whose power was non-finite. */
int opo_batch_run_ex(opo_batch *b, int max_rt, int check_every, double tol, double off_ratio, int *rts_out,
                     int *states_out);
int opo_batch_download(opo_batch *b, int idx, double *Ap, double *As, double *Ai);
/* sum |As|^2 after each round trip of instance idx's last run (n values) */
int opo_batch_power_history(opo_batch *b, int idx, double *out, int n);
/* final sum |As|^2 per instance (count values) */
int opo_batch_power(opo_batch *b, double *out);
int opo_batch_last_run_ms(opo_batch *b, double *ms);
/* Round-trip extrapolation for the convergence-tested runs (opo_batch_run,
 * opo_batch_run_ex; no reference counterpart: the reference runs a fixed
 * number of round trips).  every > 0: every `every` round trips (a multiple
 * of check_every) from `start` on, an active instance whose last three
 * snapshots show one geometric relaxation mode (ratio 0.05..0.995 per
 * `every` round trips, the last step within 10% of the ratio times the one
 * before) jumps to the series' limit and iterates on from there; its steady
 * state is the same fixed point of the round trip, reached in fewer round
 * trips.  every = 0 (the default) turns it off; results are then those of
 * plain iteration.  Device memory: two copies of the resonant fields per
 * instance that reaches `start`. */
int opo_batch_set_accel(opo_batch *b, int every, int start);
/* Per-group launch settings (autotuner; no reference counterpart).  The
 * instances of one grid size N form a group launched together; a batch of a
 * single grid size with 64 or more instances runs as two groups of half the
 * instances each, on two streams (OPO_SPLIT=0 in the environment at batch
 * creation keeps one group).  opo_batch_autotune tunes each group as
 * opo_autotune does; with several grid sizes it then also times every
 * group's pick applied to all groups, all groups running side by side as in
 * a round trip, and keeps the fastest assignment. */
int opo_batch_autotune(opo_batch *b, int iters);
/* fills min(max, groups) entries, returns the number of groups */
int opo_batch_get_tuning(opo_batch *b, opo_tuning *out, int max);
/* applies t to every group with grid size t->N */
int opo_batch_set_tuning(opo_batch *b, const opo_tuning *t);
/* opo_time_stages for the first group of grid size N (its count is in
 * opo_batch_get_tuning). */
int opo_batch_time_stages(opo_batch *b, int N, int iters, double *col_us, double *row_us);
int opo_batch_set_tiles(opo_batch *b, int col_threads, int row_threads);
/* Save slots for every instance (opo_enable_save per instance; the round
 * trips of src/cuOPO.cu:570-608) and their download (slots*N complex each,
 * any pointer may be NULL). */
int opo_batch_enable_save(opo_batch *b, int enable);
int opo_batch_download_saved(opo_batch *b, int idx, double *save_p, double *save_s, double *save_i);
int opo_batch_destroy(opo_batch *b);

/* ---------------- warm-start surrogate (no reference counterpart) ----------
 * MLP n_in -> width (GELU) x depth -> n_out with BF16 weights/activations and
 * FP32 accumulation on MFMA; FP32 master weights and Adam.  The flat parameter
 * vector holds, layer by layer, W [dout x din] row-major then b [dout].
 * Inputs/outputs are row-major FP32 host arrays [batch x n]. */
typedef struct opo_mlp opo_mlp;
int opo_mlp_create(int n_in, int width, int depth, int n_out, int max_batch, uint64_t seed, int device,
                   opo_mlp **out);
int opo_mlp_param_count(opo_mlp *m, size_t *n);
int opo_mlp_set_weights(opo_mlp *m, const float *w, size_t n);
int opo_mlp_get_weights(opo_mlp *m, float *w, size_t n);
int opo_mlp_forward(opo_mlp *m, const float *x, int batch, float *y);
/* gradient of the mean-squared error loss (grad: param_count floats, may be NULL) */
int opo_mlp_gradients(opo_mlp *m, const float *x, const float *target, int batch, float *grad, float *loss);
int opo_mlp_train_step(opo_mlp *m, const float *x, const float *target, int batch, float lr, float *loss);
int opo_mlp_destroy(opo_mlp *m);

/* ---------------- errors ---------------- */
/* Streaming copy of `bytes` between two device buffers, iters times:
 * measured bandwidth (GB/s, read+write bytes) and average kernel time.  Used
 * as the measured bandwidth peak and to scale HBM counters.  No reference
 * counterpart. */
int opo_copy_probe(size_t bytes, int iters, int device, double *gbps, double *us);
const char *opo_strerror(int status);
const char *opo_last_error(void);
int opo_device_count(int *n);

#ifdef __cplusplus
}
#endif
#endif
