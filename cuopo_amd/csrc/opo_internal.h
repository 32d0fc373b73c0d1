// opo_internal.h -- structures shared by the host driver and the gfx950
// kernels of libcuopo_amd.so.
//
// Data layout in HBM.  An instance holds NF fields (pump, signal[, idler]) of
// N = N1*N2 complex samples each, field-major.  Between kernel stages a field
// sits in the "four-step" layout: position N2*r + c holds column c, row r.
// In the time domain (at API boundaries and inside the column stage) position
// n = N2*n1 + n2 is time sample n.  After a forward column stage, position
// N2*k1 + n2 holds the column-transformed, twiddled value; after the row FFT
// position N2*k1 + k2 holds frequency k1 + N1*k2, so the linear-phase tables
// are stored in that permuted order.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/opo.h"

namespace opo {

// Column-stage pointwise program (applied in this order).
enum ColOp : unsigned {
  OP_NL_END = 1u << 0,    // nonlinear half step of slice s_end (end of that slice)
  OP_EOM = 1u << 1,       // intracavity phase modulator on signal/idler
  OP_MIRROR = 1u << 2,    // sqrt(R) e^{i pi delta} on signal/idler
  OP_SAVE = 1u << 3,      // copy the fields into the round trip's save slot
  OP_POWER = 1u << 4,     // per-tile sum |As|^2 -> partials
  OP_PUMP = 1u << 5,      // pump re-injection (CW copy / NS slot read)
  OP_NL_START = 1u << 6,  // nonlinear half step of slice s_start
  OP_SNAP = 1u << 7,      // at convergence-check round trips, copy the fields to snap
};

// Per-instance descriptor (device memory, one per instance of a group).
struct OpoInst {
  void *state;             // NF*N complex
  const void *lin;         // NF*N complex: e^{i phase}*atten/N, permuted (p, s, i); null when lfac is used
  const void *lfac;        // factorised phase for the row FFT's samples per thread E: per field
                           // Z (E*N1), S (N2), Xhi (N/E >> L), Xlo (2^L); or null
  const void *gdd;         // 2N complex (signal, idler): e^{i w^2 GDD/2}/N, permuted; null if unused
  const void *eom;         // N complex: e^{i m sin(2 pi fpm T)}, time order; null if unused
  const double *nl;        // NZ*16 doubles: 8 complex FP64 coefficients per slice
  const void *pump;        // pump_len complex (Ap_in)
  long long pump_len;
  void *save[3];           // pump, signal, idler save buffers (slots*N complex) or null
  int save_first;          // first saved round trip (CW: NRT-Nrts, NS: 0)
  int save_slots;
  int nrt;                 // NRT (NS slot arithmetic)
  int regime;              // 0 CW, 1 NS
  double h;                // dz/2, the RK step of the two-equation chain
  double kp, ks, ki;       // kappas (two-equation chain)
  double sqrt_rs, sqrt_ri; // mirror amplitude factors
  double mir[2];           // e^{i pi delta}
  int *rt;                 // round-trip counter
  int *active;             // 1 while the instance is iterated
  double *partials;        // per column-tile sum |As|^2 of the last round trip
  double *power_hist;      // per round trip sum |As|^2 (hist_len entries) or null
  int hist_len;
  int *conv_rt;            // round trip at which convergence was declared (-1)
  double *p_ref;           // power at the previous convergence check
  double *p_first;         // power at the first convergence check (off-state test)
  double *r_prev;          // P(check k-1) / P(check k-2)
  int *status;             // 0 ok, -5 non-finite power
  void *snap;              // NF*N complex: time-domain fields at the last check
  int active_v;            // copy of *active read by the stage kernels with the other fields
  int pad_;
};

struct ColArgs {
  const void *tw1;   // N1-point table e^{2 pi i e/N1}
  const void *twlo;  // e^{2 pi i l/N}, l < 2^lobits
  const void *twhi;  // e^{2 pi i h 2^lobits/N}
  int lobits;
  int N2, log2ct;
  unsigned ops;
  int load_mask, inv_mask, fwd_mask, store_mask;
  int s_end, s_start, pump_off;
  int three_eqs;
  int snap_every;    // OP_SNAP fires when (rt + 1) % snap_every == 0
  int xcd_pack;      // 1: grid is 8x wider and only blocks with blockIdx.x % 8 == 0 work
  int swz;           // 1: XCD-contiguous tile order (tiles narrower than a cache line)
  int spt;           // samples per thread of the column FFT (4 or 8)
  // single always-active instance (solver): the slice-interior kernel takes
  // these instead of reading the instance descriptor first
  int direct;
  void *d_state;
  const double *d_nl;
  double d_kp, d_ks, d_h;
};

struct RowArgs {
  const void *tw2;   // N2-point table
  int N1, log2rt;
  int nf, f0, f1, f2;
  int use_gdd;
  int xcd_pack;      // as in ColArgs
  int xl_bits;       // L of the cross-term tables of OpoInst::lfac
  int spt;           // samples per thread of the row FFT (4 or 8)
  int direct;        // as in ColArgs: pointers below instead of the descriptor
  void *d_state;
  const void *d_lin, *d_lfac, *d_gdd;
};

struct TickArgs {
  int count;
  int tiles;         // column tiles per instance (partials per instance)
  int check_every;   // 0 = no convergence test
  double tol;
  double off_ratio;  // "off" test: P < off_ratio * P(first check), see tick_kernel
};

// Launchers (defined in opo_kernels.hip).  Return hipError_t as int.
// ev0/ev1 (hipEvent_t, may be null): recorded at kernel start/stop through
// hipExtLaunchKernelGGL, so their interval is the kernel alone.
int launch_col(int fp32, int N1, int NF, int threads_ct, int count, const OpoInst *insts, const ColArgs &a,
               void *stream, void *ev0 = nullptr, void *ev1 = nullptr);
int launch_row(int fp32, int N2, int NF, int rt, int count, int tiles_y, const OpoInst *insts,
               const RowArgs &a, void *stream, void *ev0 = nullptr, void *ev1 = nullptr);
// streaming copy of n 16-byte elements (bandwidth probe / counter calibration)
int launch_copy(const void *src, void *dst, long long n, void *stream, void *ev0, void *ev1);
// linear-step phase coefficients (returns NF) and its factorisation (host)
int lin_coeffs(const opo_params &p, double D[3], double b2[3], double b3[3], double al[3]);
bool affine_grid(const double *w, int N, double &a0, double &a1, double &s);
size_t lin_factor_len(int N, int N1, int xl_bits, int E);
int lin_factors(const opo_params &p, const double *w, int N1, int xl_bits, int E, double *out);

int launch_tick(OpoInst *insts, const TickArgs &a, void *stream);

// Extrapolation of slowly converging round trips (batched runs).  For one
// instance: the resonant fields (signal, and idler with three equations,
// stored one after the other, time domain) as the convergence snapshot held
// them at three checks M round trips apart (s0 oldest, s2 the latest), and the
// instance's field state (pump, signal, idler at 0, N, 2N).
struct AccelDesc {
  const void *s0, *s1, *s2;
  void *state;
  double step;  // accel_apply: state <- s2 + step (s2 - R s1); 0 copies s2
  double c, s;  // R = e^{+i theta} on the signal, e^{-i theta} on the idler
};
// Nine FP64 sums per instance and resonant field (out[(k * nres + f) * 9 + j]):
// |s0|^2, |s1|^2, |s2|^2, then <s0,s1>, <s0,s2>, <s1,s2> (re, im), <a,b> = sum conj(a) b.
int launch_accel_dots(int fp32, const AccelDesc *d, int count, int nres, long long N, double *out, void *stream);
int launch_accel_apply(int fp32, const AccelDesc *d, int count, int nres, long long N, void *stream);
// zero fill of `words` 8-byte words on `stream`
int launch_zero(void *dst, long long words, void *stream);

// The no-allocation entry point's device memory: a static array of
// kScratchBytes per GPU (opo_kernels.hip).  scratch_region returns the
// current device's copy (address looked up once per device).
constexpr size_t kScratchBytes = size_t(8) << 20;
int scratch_region(void **ptr, size_t *bytes);

}  // namespace opo
