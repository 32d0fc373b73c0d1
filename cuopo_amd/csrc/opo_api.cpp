// opo_api.cpp -- host driver and C ABI of libcuopo_amd.so.
//
// An Engine owns one HIP stream and a set of Groups; a Group is every
// instance that shares the grid size N (so one launch of each stage covers the
// whole group).  opo_solver is an Engine with one instance, opo_batch one
// with many.  The round trip of src/cuOPO.cu:474-610 is emitted as
//   [row LIN, col MID] x (NZ-1), row LIN, col BOUNDARY, tick
// (plus col A / row GDD / col B when GDD compensation is on) and captured once
// into a hipGraph that is replayed for every steady round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/opo.h"
#include "opo_internal.h"

using opo::OpoInst;
using cd = std::complex<double>;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(OPO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define RCHK(x)                    \
  do {                             \
    int r_ = (x);                  \
    if (r_ != OPO_OK) return r_;   \
  } while (0)

constexpr long double kPiL = 3.141592653589793238462643383279502884L;
constexpr double kPi = 3.14159265358979323846;

int ilog2ll(long long x) {
  int l = 0;
  while ((1LL << l) < x) ++l;
  return l;
}
bool is_pow2(long long x) { return x > 0 && (x & (x - 1)) == 0; }

// e^{2 pi i e / M} for e = e0 + j*step, j < count, in long double then rounded
std::vector<cd> roots(long long count, long long step, long long M) {
  std::vector<cd> r((size_t)count);
  for (long long j = 0; j < count; ++j) {
    const long long e = (j * step) % M;
    const long double a = 2.0L * kPiL * (long double)e / (long double)M;
    r[(size_t)j] = cd((double)cosl(a), (double)sinl(a));
  }
  return r;
}

struct DevMem {
  void *p = nullptr;
  ~DevMem() { if (p) (void)hipFree(p); }
};

struct Instance {
  opo_params prm{};
  void *state = nullptr, *lin = nullptr, *gdd = nullptr, *eom = nullptr, *pump = nullptr, *snap = nullptr;
  void *lfac[2] = {nullptr, nullptr};  // factorised linear phase for 8 / 4 samples per thread
  bool use_fac = false;
  void *save[3] = {nullptr, nullptr, nullptr};
  double *nl = nullptr;
  long long pump_len = 0;
  bool uploaded = false;
  // round-trip extrapolation: the resonant fields at the last two snapshots
  // (acc_hist[acc_new] the newer), acc_n of them valid
  void *acc_hist[2] = {nullptr, nullptr};
  int acc_n = 0, acc_new = 0;
};

struct Group {
  int N = 0, N1 = 0, N2 = 0, NF = 0, fp32 = 0, lobits = 0;
  int xl_bits[2] = {0, 0};  // cross-term table split of the factorised phase, 8 / 4 samples per thread
  int ce = 8, re = 8;        // samples per thread of the column / row FFTs (4 or 8)
  std::vector<int> members;
  std::vector<Instance> inst;
  void *tw1 = nullptr, *tw2 = nullptr, *twlo = nullptr, *twhi = nullptr;
  OpoInst *d_insts = nullptr;
  int *d_rt = nullptr, *d_active = nullptr, *d_conv = nullptr, *d_status = nullptr;
  double *d_partials = nullptr, *d_pref = nullptr, *d_pfirst = nullptr, *d_rprev = nullptr, *d_hist = nullptr;
  int hist_len = 0;
  // instances the stage and tick launches cover: the first launch_n descriptors of d_insts (all
  // members after sync_insts; the still-active ones after a compaction, see check_active)
  int launch_n = 0;
  int ct = 0, rt = 0;  // column tile (columns per block), row tile (rows per block)
  int rt_tuned = 0;    // the row tile to restore after a compacted run narrowed it (0: not narrowed)
  std::vector<int> act;                // active flags (by member) at the last convergence check
  opo::AccelDesc *d_acc = nullptr;     // extrapolation descriptors (one per member at most)
  double *d_dots = nullptr;            // their sums: 9 per member and resonant field
  int col_threads = 0, row_threads = 0;  // per-group overrides (0 = engine setting)
  int xcd_pack = -1;                      // per-group override (-1 = engine setting)
  int phase_mode = -1;                    // linear phase: -1 auto, 0 full table, 1 factorised
  double tuned_us = 0;                    // slice time of the autotuner's choice (0 = not tuned)
  bool any_gdd = false;
  std::vector<void *> allocs;
  // no-allocation mode: device memory comes from the caller's buffers (fields,
  // pump) and a bump pool over the static scratch array instead of hipMalloc
  char *pool = nullptr;
  size_t pool_cap = 0, pool_used = 0;
  void *ext_state = nullptr;
  const void *ext_pump = nullptr;
  hipStream_t gs = nullptr;      // the group's own stream (batches with several grid sizes)
  hipEvent_t done_ev = nullptr;  // join point of the group's chain
  hipStream_t cur = nullptr;     // stream the stage emitters launch on
};

struct Engine {
  std::vector<Group> groups;
  int count = 0, NZ = 0, three_eqs = 0, regime = 0, fp32 = 0, device = 0, NRT = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool use_graph = true;
  bool compact = true;  // shrink the launches to the still-active instances during a converging run
  int accel_every = 0;     // round-trip extrapolation every this many round trips (0: off), see accel_step
  int accel_start = 1000;  // ... from this many round trips of the run on
  bool save_enabled = false;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_key = -1;  // 0: the captured graph matches the settings below; -1: rebuild
  int graph_check_every = 0;
  double graph_tol = 0;
  double graph_off_ratio = 0;
  bool graph_save = false;
  float last_ms = 0;
  int last_rt0 = 0, last_nrt = 0;
  int col_threads = 0, row_threads = 0;
  int xcd_pack = 0;
  bool solver = false;  // one instance that stays active: stage kernels take its pointers directly
  double off_ratio = 1e-6;
  bool inst_dirty = true;
  bool own_stream = true;  // false: `stream` is the caller's (no-allocation entry point)
  std::vector<std::pair<int, int>> where;  // batch index -> (group, slot)

  hipEvent_t fork_ev = nullptr;

  ~Engine() {
    // release on the engine's device (the caller's current device may differ),
    // then give the caller its device back
    int prev = -1;
    const bool switch_dev = stream && hipGetDevice(&prev) == hipSuccess && prev != device;
    if (switch_dev) (void)hipSetDevice(device);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &g : groups) {
      for (void *p : g.allocs) (void)hipFree(p);
      if (g.done_ev) (void)hipEventDestroy(g.done_ev);
      if (g.gs) (void)hipStreamDestroy(g.gs);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
    if (switch_dev) (void)hipSetDevice(prev);
  }
};

size_t esize(const Group &g) { return g.fp32 ? 8 : 16; }
int eidx(int spt) { return spt == 4 ? 1 : 0; }
// the kernels' launch bounds (column stage; row stage of N2 points)
int max_threads(int spt) { return spt == 4 ? 1024 : 512; }
int max_row_threads(int spt, int N2) { return (spt == 4 && N2 >= 4096) ? 1024 : 512; }

// Linear-phase source of a group.  The factorised tables (opo_phase.cpp) save
// the N-point table read per field and slice but cost extra loads and
// products per sample: measured on gfx950 they win once a group carries about
// 2^19 samples per field (bandwidth-bound launches) and lose on smaller,
// latency-bound ones.  The autotuner overrides the choice per group.
bool group_factored(const Group &g) {
  if (g.phase_mode >= 0) return g.phase_mode == 1;
  return (long long)g.members.size() * g.N >= (1LL << 19);
}

int dalloc(Group &g, void **p, size_t bytes) {
  if (bytes == 0) { *p = nullptr; return OPO_OK; }
  if (g.pool) {
    const size_t at = (g.pool_used + 255) & ~size_t(255);
    if (at > g.pool_cap || bytes > g.pool_cap - at)
      return fail(OPO_ERR_INVALID, "the instance needs more device memory than the no-allocation scratch holds (" +
                                       std::to_string(g.pool_cap >> 20) + " MiB per GPU)");
    *p = g.pool + at;
    g.pool_used = at + bytes;
    return OPO_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(OPO_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  g.allocs.push_back(*p);
  return OPO_OK;
}

// copy complex host data to a device buffer of the group's element type
int upload_c(Engine &E, Group &g, void *dst, const cd *src, size_t n) {
  if (g.fp32) {
    std::vector<float> tmp(2 * n);
    for (size_t i = 0; i < n; ++i) { tmp[2 * i] = (float)src[i].real(); tmp[2 * i + 1] = (float)src[i].imag(); }
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  } else {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(cd), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  }
  return OPO_OK;
}

int download_c(Engine &E, Group &g, cd *dst, const void *src, size_t n) {
  if (g.fp32) {
    std::vector<float> tmp(2 * n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    for (size_t i = 0; i < n; ++i) dst[i] = cd(tmp[2 * i], tmp[2 * i + 1]);
  } else {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(cd), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  }
  return OPO_OK;
}

int validate(const opo_params *p) {
  if (!p) return fail(OPO_ERR_INVALID, "null params");
  if (!is_pow2(p->N) || p->N < 256 || p->N > (1 << 18))
    return fail(OPO_ERR_INVALID, "N must be a power of two in [256, 262144]");
  if (p->NZ < 1 || p->NRT < 1) return fail(OPO_ERR_INVALID, "NZ and NRT must be >= 1");
  if (p->precision != OPO_FP64 && p->precision != OPO_FP32C) return fail(OPO_ERR_INVALID, "bad precision");
  if (p->regime != OPO_CW && p->regime != OPO_NS) return fail(OPO_ERR_INVALID, "bad regime");
  if (p->Nrts < 0) return fail(OPO_ERR_INVALID, "Nrts must be >= 0");
  if (!(p->dz > 0) || !std::isfinite(p->dk)) return fail(OPO_ERR_INVALID, "dz must be > 0, dk finite");
  return OPO_OK;
}

// LDS bytes of one column-stage / row-stage block (must match the launchers)
size_t col_lds(const Group &g, int ct) { return ((size_t)g.N1 + 2 * (size_t)g.NF * ct * g.N1) * esize(g) + 512; }
size_t row_lds(const Group &g, int rt) { return 2 * (size_t)rt * (g.N2 + g.N2 / 8) * esize(g); }
constexpr size_t kMaxLds = 160 * 1024;

int floor_pow2(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

// Column tile ct (columns per block, threads NF*ct*N1/ce) and row tile rt
// (rows per block, threads rt*N2/re): powers of two dividing N2 / N1, within
// the kernels' launch bounds and the LDS of one CU.
void pick_tiles(Engine &E, Group &g) {
  const int T1 = g.N1 / g.ce, T2 = g.N2 / g.re;
  // default column tile: one 128-byte line of columns (8 FP64 / 16 FP32)
  const int colt = g.col_threads ? g.col_threads : E.col_threads;
  int ct = colt ? std::max(1, colt / T1) : (int)(128 / esize(g));
  // default row tile: 64 threads (the autotuner's choice on gfx950 for almost every group)
  int rtn = g.row_threads ? g.row_threads : (E.row_threads ? E.row_threads : 64);
  ct = std::min(floor_pow2(ct), g.N2);
  while (ct > 1 && g.NF * ct * T1 > max_threads(g.ce)) ct >>= 1;
  g.ct = ct;
  g.rt_tuned = 0;
  g.rt = std::max(1, std::min(floor_pow2(std::max(1, rtn / T2)), g.N1));
  while (g.rt > 1 && g.rt * T2 > max_row_threads(g.re, g.N2)) g.rt >>= 1;
  while (g.ct > 1 && col_lds(g, g.ct) > kMaxLds) g.ct >>= 1;
  while (g.rt > 1 && row_lds(g, g.rt) > kMaxLds) g.rt >>= 1;
}

// host-side check of a launch shape against what the kernels assume
int check_shapes(const Group &g) {
  if ((g.ce != 4 && g.ce != 8) || (g.re != 4 && g.re != 8))
    return fail(OPO_ERR_INVALID, "internal: samples per thread must be 4 or 8");
  const int T1 = g.N1 / g.ce, T2 = g.N2 / g.re;
  if (!is_pow2(g.ct) || g.N2 % g.ct != 0 || g.NF * g.ct * T1 > max_threads(g.ce) || col_lds(g, g.ct) > kMaxLds)
    return fail(OPO_ERR_INVALID, "internal: column tile does not fit the kernel");
  if (!is_pow2(g.rt) || g.N1 % g.rt != 0 || g.rt * T2 > max_row_threads(g.re, g.N2) || row_lds(g, g.rt) > kMaxLds)
    return fail(OPO_ERR_INVALID, "internal: row tile does not fit the kernel");
  if (g.N1 < 16 || g.N2 < 16 || g.N1 > 128 || g.N2 > 4096 || (long long)g.N1 * g.N2 != g.N)
    return fail(OPO_ERR_INVALID, "internal: four-step split out of range");
  return OPO_OK;
}

int make_group(Engine &E, Group &g) {
  g.NF = E.three_eqs ? 3 : 2;
  g.fp32 = E.fp32;
  // four-step split: short columns (N1 <= 64 where possible) so a column
  // tile is a full cache line, rows up to 2048 (FP64) / 4096 (FP32) points
  const int L = ilog2ll(g.N);
  g.N1 = std::min(1 << (L / 2), 64);
  g.N2 = g.N / g.N1;
  const int row_max = g.fp32 ? 4096 : 2048;
  while (g.N2 > row_max) { g.N2 >>= 1; g.N1 <<= 1; }
  g.lobits = (L + 1) / 2;
  g.xl_bits[0] = std::max(0, (L - 3 + 1) / 2);  // cross-term tables cover e < N/8
  g.xl_bits[1] = std::max(0, (L - 2 + 1) / 2);  // ... and e < N/4
  const size_t es = esize(g);
  RCHK(dalloc(g, &g.tw1, g.N1 * es));
  RCHK(dalloc(g, &g.tw2, g.N2 * es));
  RCHK(dalloc(g, &g.twlo, ((size_t)1 << g.lobits) * es));
  RCHK(dalloc(g, &g.twhi, ((size_t)g.N >> g.lobits) * es));
  auto r1 = roots(g.N1, 1, g.N1), r2 = roots(g.N2, 1, g.N2);
  auto rlo = roots(1LL << g.lobits, 1, g.N), rhi = roots((long long)g.N >> g.lobits, 1LL << g.lobits, g.N);
  RCHK(upload_c(E, g, g.tw1, r1.data(), r1.size()));
  RCHK(upload_c(E, g, g.tw2, r2.data(), r2.size()));
  RCHK(upload_c(E, g, g.twlo, rlo.data(), rlo.size()));
  RCHK(upload_c(E, g, g.twhi, rhi.data(), rhi.size()));
  const int n = (int)g.members.size();
  RCHK(dalloc(g, (void **)&g.d_insts, n * sizeof(OpoInst)));
  RCHK(dalloc(g, (void **)&g.d_rt, n * sizeof(int)));
  RCHK(dalloc(g, (void **)&g.d_active, n * sizeof(int)));
  RCHK(dalloc(g, (void **)&g.d_conv, n * sizeof(int)));
  RCHK(dalloc(g, (void **)&g.d_status, n * sizeof(int)));
  RCHK(dalloc(g, (void **)&g.d_pref, n * sizeof(double)));
  RCHK(dalloc(g, (void **)&g.d_pfirst, n * sizeof(double)));
  RCHK(dalloc(g, (void **)&g.d_rprev, n * sizeof(double)));
  RCHK(dalloc(g, (void **)&g.d_partials, (size_t)n * g.N2 * sizeof(double)));
  g.hist_len = std::max(E.NRT, 1);
  RCHK(dalloc(g, (void **)&g.d_hist, (size_t)n * g.hist_len * sizeof(double)));
  if (g.pool) {
    const int e = opo::launch_zero(g.d_hist, (long long)n * g.hist_len, E.stream);
    if (e != 0) return fail(OPO_ERR_HIP, std::string("zero fill: ") + hipGetErrorString((hipError_t)e));
  } else {
    HIPCHK(hipMemsetAsync(g.d_hist, 0, (size_t)n * g.hist_len * sizeof(double), E.stream));
  }
  g.inst.resize(n);
  for (int k = 0; k < n; ++k) {
    Instance &I = g.inst[k];
    if (g.ext_state) I.state = g.ext_state;  // the caller's field buffer (one instance)
    else RCHK(dalloc(g, &I.state, (size_t)g.NF * g.N * es));
    RCHK(dalloc(g, (void **)&I.nl, (size_t)E.NZ * 16 * sizeof(double)));
  }
  pick_tiles(E, g);
  return OPO_OK;
}

int grow_hist(Engine &E, Group &g, int need) {
  if (need <= g.hist_len) return OPO_OK;
  const int n = (int)g.members.size();
  void *p = nullptr;
  RCHK(dalloc(g, &p, (size_t)n * need * sizeof(double)));
  HIPCHK(hipMemsetAsync(p, 0, (size_t)n * need * sizeof(double), E.stream));
  g.d_hist = (double *)p;
  g.hist_len = need;
  E.inst_dirty = true;
  return OPO_OK;
}

// Host tables of one instance.  Every expression follows the reference line
// it replaces, in FP64.
int upload_instance(Engine &E, Group &g, int k, const double *w, const double *T, const double *Ap_in,
                    size_t n_ap, const double *Ap, const double *As, const double *Ai) {
  Instance &I = g.inst[k];
  const opo_params &p = I.prm;
  const int N = g.N, N1 = g.N1, NF = g.NF;
  const size_t es = esize(g);
  if (!w || (!Ap_in && !g.ext_pump)) return fail(OPO_ERR_INVALID, "w and Ap_in are required");
  const size_t need_ap = p.regime == OPO_CW ? (size_t)N : (size_t)N * p.NRT;
  if (n_ap < need_ap) return fail(OPO_ERR_INVALID, "Ap_in too short (N for CW, N*NRT for NS)");
  // position in the four-step frequency layout of natural frequency index q
  auto perm = [&](int q) { return (size_t)(g.N2 * (q % N1) + q / N1); };

  // linear operator, cwes3.cuh:112-125 / cwes2.cuh:108-118 with z = dz (Q8); 1/N of CUFFTscale folded in.
  // The full N-point table per field always; on the reference's grid also the
  // factorised tables (opo_phase.cpp).  group_factored picks one per group.
  {
    I.use_fac = !g.pool;  // no-allocation mode reads the full table only
    for (int k = 0; k < 2 && I.use_fac; ++k) {
      const int spt = k ? 4 : 8;
      std::vector<double> fac((size_t)2 * NF * opo::lin_factor_len(N, N1, g.xl_bits[k], spt));
      I.use_fac = opo::lin_factors(p, w, N1, g.xl_bits[k], spt, fac.data()) == 1;
      if (I.use_fac) {
        if (!I.lfac[k]) RCHK(dalloc(g, &I.lfac[k], fac.size() / 2 * es));
        RCHK(upload_c(E, g, I.lfac[k], reinterpret_cast<const cd *>(fac.data()), fac.size() / 2));
      }
    }
    std::vector<cd> lin((size_t)NF * N);
    double D[3], b2[3], b3[3], al[3];
    opo::lin_coeffs(p, D, b2, b3, al);
    const double z = p.dz;
    for (int f = 0; f < NF; ++f) {
      const double atten = std::exp(-0.5 * al[f] * z);
      for (int q = 0; q < N; ++q) {
        const double wq = w[q];
        const double ph = z * wq * (D[f] + 0.5 * wq * b2[f] + wq * wq * b3[f] / 6);
        lin[(size_t)f * N + perm(q)] = cd(std::cos(ph), std::sin(ph)) * (atten / N);
      }
    }
    if (!I.lin) RCHK(dalloc(g, &I.lin, lin.size() * es));
    RCHK(upload_c(E, g, I.lin, lin.data(), lin.size()));
    E.inst_dirty = true;
    E.graph_key = -1;  // launch arguments may carry this instance's pointers
  }

  // GDD compensation, cuOPO.cu:521-553 (AddGDD functions.cuh:156-169; AddGDDTOD :173-186)
  if (g.any_gdd) {
    if (!I.gdd) RCHK(dalloc(g, &I.gdd, (size_t)2 * N * es));
    std::vector<cd> gt((size_t)2 * N);
    const double tod[2] = {p.TODscomp, p.TODicomp}, bb3[2] = {p.b3s, p.b3i};
    for (int f = 0; f < 2; ++f)
      for (int q = 0; q < N; ++q) {
        const double wq = w[q];
        double ph;
        if (p.GDD == 0) ph = 0;
        else if (tod[f] != 0 && bb3[f] != 0) ph = 0.5 * wq * (wq * p.GDD + wq * wq * tod[f] / 3.0);
        else ph = 0.5 * wq * wq * p.GDD;
        gt[(size_t)f * N + perm(q)] = cd(std::cos(ph), std::sin(ph)) * (1.0 / N);
      }
    RCHK(upload_c(E, g, I.gdd, gt.data(), gt.size()));
  }
  // EOM, functions.cuh:242-253
  if (p.use_eom) {
    if (!T) return fail(OPO_ERR_INVALID, "T is required when the EOM is on");
    if (!I.eom) RCHK(dalloc(g, &I.eom, (size_t)N * es));
    std::vector<cd> et(N);
    for (int j = 0; j < N; ++j) {
      const double a = p.mod_depth * std::sin(2 * kPi * p.fpm * T[j]);
      et[j] = cd(std::cos(a), std::sin(a));
    }
    RCHK(upload_c(E, g, I.eom, et.data(), et.size()));
  }
  // nonlinear coefficients per slice, z accumulated as cwes3.cuh:149,219
  std::vector<double> nl((size_t)E.NZ * 16, 0.0);
  {
    double zz = 0;
    const double h = p.dz / 2;
    for (int s = 0; s < E.NZ; ++s) {
      double *q = nl.data() + 16 * s;
      const double za = zz, zb = zz + p.dz / 4, zc = zz + p.dz / 2;
      auto ex = [](double a) { return cd(std::cos(a), std::sin(a)); };
      if (E.three_eqs) {
        // dAdz (cwes3.cuh:37-39): pump carries e^{-i dk z}, signal/idler e^{+i dk z}
        const cd em = (ex(-p.dk * za) + 2.0 * ex(-p.dk * zb)) + 2.0 * ex(-p.dk * zb) + ex(-p.dk * zc);
        const cd ep = (ex(+p.dk * za) + 2.0 * ex(+p.dk * zb)) + 2.0 * ex(+p.dk * zb) + ex(+p.dk * zc);
        const cd C[3] = {cd(0, p.kp) * em * h / 6.0, cd(0, p.ks) * ep * h / 6.0, cd(0, p.ki) * ep * h / 6.0};
        for (int j = 0; j < 3; ++j) { q[2 * j] = C[j].real(); q[2 * j + 1] = C[j].imag(); }
      } else {
        // dAdz (cwes2.cuh:37-38): pump e^{+i dk z}, signal e^{-i dk z}
        const double zs[3] = {za, zb, zc};
        for (int j = 0; j < 3; ++j) {
          const cd a = ex(+p.dk * zs[j]), b = ex(-p.dk * zs[j]);
          q[2 * j] = a.real(); q[2 * j + 1] = a.imag();
          q[6 + 2 * j] = b.real(); q[6 + 2 * j + 1] = b.imag();
        }
      }
      zz += p.dz;
    }
  }
  HIPCHK(hipMemcpyAsync(I.nl, nl.data(), nl.size() * sizeof(double), hipMemcpyHostToDevice, E.stream));
  HIPCHK(hipStreamSynchronize(E.stream));
  // pump (Ap_in): N (CW) or N*NRT (NS) complex
  if (g.ext_pump) {
    I.pump = const_cast<void *>(g.ext_pump);  // the caller's device copy, read only
    I.pump_len = (long long)need_ap;
    I.uploaded = true;
    E.inst_dirty = true;
    return OPO_OK;  // the fields are already in the caller's buffer
  }
  if (!I.pump || I.pump_len != (long long)need_ap) {
    RCHK(dalloc(g, &I.pump, need_ap * es));
    I.pump_len = (long long)need_ap;
  }
  RCHK(upload_c(E, g, I.pump, reinterpret_cast<const cd *>(Ap_in), need_ap));
  // fields
  std::vector<cd> st((size_t)NF * N, cd(0, 0));
  for (int j = 0; j < N; ++j) {
    if (Ap) st[j] = cd(Ap[2 * j], Ap[2 * j + 1]);
    if (As) st[(size_t)N + j] = cd(As[2 * j], As[2 * j + 1]);
    if (NF == 3 && Ai) st[(size_t)2 * N + j] = cd(Ai[2 * j], Ai[2 * j + 1]);
  }
  RCHK(upload_c(E, g, I.state, st.data(), st.size()));
  I.uploaded = true;
  E.inst_dirty = true;
  return OPO_OK;
}

// Device descriptor of instance k of group g (pointers into the group's
// per-instance arrays, so its place in d_insts does not matter).
OpoInst make_desc(const Engine &E, const Group &g, int k) {
  const Instance &I = g.inst[k];
  const opo_params &p = I.prm;
  OpoInst d;
  std::memset(&d, 0, sizeof(d));
  d.state = I.state; d.lin = I.lin; d.lfac = I.use_fac && group_factored(g) ? I.lfac[eidx(g.re)] : nullptr; d.gdd = I.gdd; d.eom = p.use_eom ? I.eom : nullptr; d.nl = I.nl;
  d.pump = I.pump; d.pump_len = I.pump_len;
  for (int f = 0; f < 3; ++f) d.save[f] = E.save_enabled ? I.save[f] : nullptr;
  d.save_first = p.regime == OPO_CW ? p.NRT - p.Nrts : 0;
  d.save_slots = p.regime == OPO_CW ? p.Nrts : p.NRT;
  d.nrt = p.NRT;
  d.regime = p.regime;
  d.h = p.dz / 2;
  d.kp = p.kp; d.ks = p.ks; d.ki = p.ki;
  d.sqrt_rs = std::sqrt(p.Rs); d.sqrt_ri = std::sqrt(p.Ri);  // AddPhase, functions.cuh:145
  d.mir[0] = std::cos(kPi * p.delta); d.mir[1] = std::sin(kPi * p.delta);
  d.rt = g.d_rt + k; d.active = g.d_active + k; d.partials = g.d_partials + (size_t)k * g.N2;
  d.power_hist = g.d_hist + (size_t)k * g.hist_len; d.hist_len = g.hist_len;
  d.conv_rt = g.d_conv + k; d.p_ref = g.d_pref + k; d.p_first = g.d_pfirst + k; d.r_prev = g.d_rprev + k; d.status = g.d_status + k;
  d.snap = I.snap;
  d.active_v = 1;
  return d;
}

int sync_insts(Engine &E) {
  if (!E.inst_dirty) return OPO_OK;
  for (auto &g : E.groups) {
    const int n = (int)g.members.size();
    std::vector<OpoInst> h(n);
    for (int k = 0; k < n; ++k) h[k] = make_desc(E, g, k);
    HIPCHK(hipMemcpyAsync(g.d_insts, h.data(), n * sizeof(OpoInst), hipMemcpyHostToDevice, E.stream));
    g.launch_n = n;
  }
  HIPCHK(hipStreamSynchronize(E.stream));
  E.inst_dirty = false;
  return OPO_OK;
}

// ------------------------------------------------------------ stage emitters

int emit_col(Engine &E, Group &g, unsigned ops, int load, int inv, int fwd, int store, int s_end, int s_start,
             int pump_off, int snap_every, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  RCHK(check_shapes(g));
  opo::ColArgs a{};
  a.tw1 = g.tw1; a.twlo = g.twlo; a.twhi = g.twhi; a.lobits = g.lobits;
  a.N2 = g.N2; a.log2ct = ilog2ll(g.ct);
  a.ops = ops; a.load_mask = load; a.inv_mask = inv; a.fwd_mask = fwd; a.store_mask = store;
  a.s_end = s_end; a.s_start = s_start; a.pump_off = pump_off; a.three_eqs = E.three_eqs;
  a.snap_every = snap_every;
  a.xcd_pack = g.xcd_pack >= 0 ? g.xcd_pack : E.xcd_pack;
  a.swz = !a.xcd_pack && g.ct * esize(g) < 128 && (g.N2 / g.ct) % 8 == 0;
  a.spt = g.ce;
  if (E.solver) {
    const Instance &I = g.inst[0];
    a.direct = 1;
    a.d_state = I.state; a.d_nl = I.nl;
    a.d_kp = I.prm.kp; a.d_ks = I.prm.ks; a.d_h = I.prm.dz / 2;
  }
  const int threads_ct = g.ct;
  int e = opo::launch_col(g.fp32, g.N1, g.NF, threads_ct, g.launch_n, g.d_insts, a, g.cur, e0, e1);
  if (e != 0) return fail(OPO_ERR_HIP, std::string("col_stage launch: ") + hipGetErrorString((hipError_t)e));
  return OPO_OK;
}

int emit_row(Engine &E, Group &g, bool gdd, int mask, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  RCHK(check_shapes(g));
  opo::RowArgs a{};
  a.tw2 = g.tw2; a.N1 = g.N1; a.log2rt = ilog2ll(g.rt); a.use_gdd = gdd ? 1 : 0;
  int fl[3] = {0, 0, 0}, nf = 0;
  for (int f = 0; f < g.NF; ++f)
    if ((mask >> f) & 1) fl[nf++] = f;
  a.nf = nf; a.f0 = fl[0]; a.f1 = fl[1]; a.f2 = fl[2];
  a.xcd_pack = g.xcd_pack >= 0 ? g.xcd_pack : E.xcd_pack;
  a.xl_bits = g.xl_bits[eidx(g.re)];
  a.spt = g.re;
  if (E.solver) {
    const Instance &I = g.inst[0];
    a.direct = 1;
    a.d_state = I.state; a.d_lin = I.lin; a.d_gdd = I.gdd;
    a.d_lfac = I.use_fac && group_factored(g) ? I.lfac[eidx(g.re)] : nullptr;
  }
  int e = opo::launch_row(g.fp32, g.N2, g.NF, g.rt, g.launch_n, 0, g.d_insts, a, g.cur, e0, e1);
  if (e != 0) return fail(OPO_ERR_HIP, std::string("row_stage launch: ") + hipGetErrorString((hipError_t)e));
  return OPO_OK;
}

int emit_tick(Engine &E, Group &g, int check_every, double tol) {
  opo::TickArgs a{};
  a.count = g.launch_n; a.tiles = g.N2 / g.ct; a.check_every = check_every; a.tol = tol;
  a.off_ratio = E.off_ratio;
  int e = opo::launch_tick(g.d_insts, a, g.cur);
  if (e != 0) return fail(OPO_ERR_HIP, std::string("tick launch: ") + hipGetErrorString((hipError_t)e));
  return OPO_OK;
}

int all_mask(const Group &g) { return (1 << g.NF) - 1; }
int si_mask(const Group &g) { return g.NF == 3 ? 6 : 2; }

// start of a run: time-domain fields -> pump re-injection, NL(z_0), forward columns
int emit_start(Engine &E, Group &g) {
  const int A = all_mask(g);
  return emit_col(E, g, opo::OP_PUMP | opo::OP_NL_START, A, 0, A, A, 0, 0, 0, 0);
}

// one round trip after emit_start (or after a previous non-final round trip)
int emit_rt_body(Engine &E, Group &g, bool last, int check_every, double tol) {
  using namespace opo;
  if (g.launch_n == 0) return OPO_OK;  // every instance of the group has finished (compaction)
  const int A = all_mask(g), SI = si_mask(g), NZ = E.NZ;
  for (int s = 0; s < NZ; ++s) {
    RCHK(emit_row(E, g, false, A));
    if (s < NZ - 1) RCHK(emit_col(E, g, OP_NL_END | OP_NL_START, A, A, A, A, s, s + 1, 0, 0));
  }
  const unsigned tail = OP_MIRROR | OP_SAVE | OP_POWER | OP_SNAP;
  const unsigned cont = last ? 0u : (OP_PUMP | OP_NL_START);
  const int fwd = last ? 0 : A;
  if (!g.any_gdd) {
    RCHK(emit_col(E, g, OP_NL_END | OP_EOM | tail | cont, A, A, fwd, A, NZ - 1, 0, 1, check_every));
  } else {
    RCHK(emit_col(E, g, OP_NL_END | OP_EOM, A, A, SI, A, NZ - 1, 0, 0, 0));
    RCHK(emit_row(E, g, true, SI));
    RCHK(emit_col(E, g, tail | cont, A, SI, fwd, A, 0, 0, 1, check_every));
  }
  return emit_tick(E, g, check_every, tol);
}

int reset_counters(Engine &E, int rt0) {
  E.inst_dirty = true;
  RCHK(sync_insts(E));  // descriptors carry active_v = 1
  for (auto &g : E.groups) {
    for (auto &I : g.inst) I.acc_n = 0;  // extrapolation histories belong to one run
    g.act.clear();
    const int n = (int)g.members.size();
    std::vector<int> rt(n, rt0), act(n, 1), conv(n, -1), stat(n, 0);
    std::vector<double> pref(n, 0.0);
    HIPCHK(hipMemcpyAsync(g.d_rt, rt.data(), n * sizeof(int), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(g.d_active, act.data(), n * sizeof(int), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(g.d_conv, conv.data(), n * sizeof(int), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(g.d_status, stat.data(), n * sizeof(int), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(g.d_pref, pref.data(), n * sizeof(double), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(g.d_pfirst, pref.data(), n * sizeof(double), hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(g.d_rprev, pref.data(), n * sizeof(double), hipMemcpyHostToDevice, E.stream));
  }
  HIPCHK(hipStreamSynchronize(E.stream));
  return OPO_OK;
}

// Batches with several grid sizes run each group's chain on the group's own
// stream (forked from and joined back to the engine stream), so the hardware
// queues overlap the groups' kernels; one group uses the engine stream.
int fork_groups(Engine &E) {
  if (E.groups.size() == 1) {
    E.groups[0].cur = E.stream;
    return OPO_OK;
  }
  HIPCHK(hipEventRecord(E.fork_ev, E.stream));
  for (auto &g : E.groups) {
    HIPCHK(hipStreamWaitEvent(g.gs, E.fork_ev, 0));
    g.cur = g.gs;
  }
  return OPO_OK;
}

int join_groups(Engine &E) {
  if (E.groups.size() == 1) return OPO_OK;
  for (auto &g : E.groups) {
    HIPCHK(hipEventRecord(g.done_ev, g.gs));
    HIPCHK(hipStreamWaitEvent(E.stream, g.done_ev, 0));
    g.cur = E.stream;
  }
  return OPO_OK;
}

int build_graph(Engine &E, int check_every, double tol) {
  // the captured kernel arguments carry check_every (column stages, tick),
  // tol and off_ratio (tick); a change of these or of the save switch means a
  // new capture
  if (E.gexec && E.graph_key == 0 && E.graph_check_every == check_every && E.graph_tol == tol &&
      E.graph_off_ratio == E.off_ratio && E.graph_save == E.save_enabled)
    return OPO_OK;
  if (E.gexec) { (void)hipGraphExecDestroy(E.gexec); E.gexec = nullptr; }
  if (E.graph) { (void)hipGraphDestroy(E.graph); E.graph = nullptr; }
  HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
  int rc = fork_groups(E);
  for (auto &g : E.groups) {
    if (rc != OPO_OK) break;
    rc = emit_rt_body(E, g, false, check_every, tol);
  }
  if (rc == OPO_OK) rc = join_groups(E);
  hipGraph_t gr = nullptr;
  hipError_t ee = hipStreamEndCapture(E.stream, &gr);
  if (rc != OPO_OK) { if (gr) (void)hipGraphDestroy(gr); return rc; }
  if (ee != hipSuccess) return fail(OPO_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
  E.graph = gr;
  HIPCHK(hipGraphInstantiate(&E.gexec, E.graph, nullptr, nullptr, 0));
  E.graph_key = 0;
  E.graph_check_every = check_every;
  E.graph_tol = tol;
  E.graph_off_ratio = E.off_ratio;
  E.graph_save = E.save_enabled;
  return OPO_OK;
}

// After a convergence check: *alive = some instance still iterates.  With
// compaction on, a group in which an instance has finished since its last
// compaction gets a descriptor array of the active ones, and its launches
// shrink to them (*changed; the caller captures the round trip again, a few
// ms against the 50 round trips between checks); a group with none left
// launches nothing (its chain of a two-group batch no longer runs beside the
// other's).  Finished instances were skipped by every block anyway, so the
// arithmetic and the results do not change.
constexpr int kNarrowRowsBelow = 64;  // launches of this many instances or fewer get one-wavefront row blocks

int check_active(Engine &E, bool *alive, bool *changed) {
  *alive = false;
  *changed = false;
  for (auto &g : E.groups) {
    const int n = (int)g.members.size();
    std::vector<int> act(n);
    HIPCHK(hipMemcpyAsync(act.data(), g.d_active, n * sizeof(int), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    g.act = act;
    std::vector<OpoInst> h;
    for (int k = 0; k < n; ++k)
      if (act[k]) h.push_back(make_desc(E, g, k));
    if (!h.empty()) *alive = true;
    const int m = (int)h.size();
    if (!E.compact || m >= g.launch_n) continue;
    if (m > 0) {
      HIPCHK(hipMemcpyAsync(g.d_insts, h.data(), h.size() * sizeof(OpoInst), hipMemcpyHostToDevice, E.stream));
      HIPCHK(hipStreamSynchronize(E.stream));
    }
    g.launch_n = m;  // 0: the group is done and launches nothing until the next run
    *changed = true;
    if (m == 0) continue;
    // A few instances leave most CUs idle and every stage latency-bound; rows
    // of at most one wavefront per block then spread them over more CUs (the
    // tail of a sweep at 2^14 samples: 2-7% less time per round trip at 1-64
    // instances).
    // Only rows are regrouped: every row is transformed exactly as before,
    // while the column tile also fixes the power partials' summation order.
    const int one_wave_rt = std::max(1, 64 / (g.N2 / g.re));
    if (g.launch_n <= kNarrowRowsBelow && g.rt > one_wave_rt) {
      if (!g.rt_tuned) g.rt_tuned = g.rt;
      g.rt = one_wave_rt;
    }
  }
  return OPO_OK;
}

// the tuned row tile again after a run that narrowed it
void restore_row_tiles(Engine &E) {
  for (auto &g : E.groups)
    if (g.rt_tuned) { g.rt = g.rt_tuned; g.rt_tuned = 0; }
}

// Extrapolation of slowly converging round trips (opo_batch_set_accel).  Near
// its steady state an instance relaxes like x_n = x* + c lambda^n v once the
// faster transients have died out; a slow one (lambda^400 of 0.6-0.95 was
// measured on the headline sweep's stragglers) needs 10^4 round trips to pass
// the steady test.  Every accel_every round trips from accel_start on, each
// active instance's resonant fields (the convergence snapshot, time domain)
// are kept; with three of them, s0, s1, s2, aligned by the phase that the pump
// leaves free (signal e^{i theta}, idler e^{-i theta}), d1 = s1 - s0 and
// d2 = s2 - s1 give rho = Re<d1,d2>/|d1|^2.  When d2 is rho d1 to within 10%
// (one relaxation mode dominates) and 0.05 < rho < 0.995, the instance jumps to
// the limit of the geometric series, s2 + d2 rho/(1 - rho) (at most 50 d2), and
// starts its next round trip from there; the steady test then runs on as
// before.  Every other active instance of the group restarts from its own
// snapshot unchanged (the same pump re-injection, NL(z_0) and forward columns
// the round trip would have run, so its results are unchanged), and its
// snapshot joins its history.  Sums on the device in FP64, decisions on the host.
constexpr double kAccelMaxResidual = 0.1, kAccelRhoMin = 0.05, kAccelRhoMax = 0.995, kAccelMaxStep = 50.0;

int accel_step(Engine &E) {
  using cplx = std::complex<double>;
  for (auto &g : E.groups) {
    const int n = (int)g.members.size();
    const int nres = g.NF - 1;
    const size_t es = esize(g), hbytes = (size_t)nres * g.N * es;
    if ((int)g.act.size() != n) continue;
    std::vector<int> live;
    for (int k = 0; k < n; ++k)
      if (g.act[k] && g.inst[k].snap) live.push_back(k);
    if (live.empty()) continue;
    if (!g.d_acc) {
      RCHK(dalloc(g, (void **)&g.d_acc, (size_t)n * sizeof(opo::AccelDesc)));
      RCHK(dalloc(g, (void **)&g.d_dots, (size_t)n * nres * 9 * sizeof(double)));
    }
    auto res_of = [&](const Instance &I) { return (const void *)((const char *)I.snap + (size_t)g.N * es); };
    // 1. the instances with two earlier snapshots: sums, then the decision
    std::vector<opo::AccelDesc> dd;
    std::vector<int> who;
    for (int k : live) {
      const Instance &I = g.inst[k];
      if (I.acc_n < 2) continue;
      opo::AccelDesc a{};
      a.s0 = I.acc_hist[1 - I.acc_new];
      a.s1 = I.acc_hist[I.acc_new];
      a.s2 = res_of(I);
      dd.push_back(a);
      who.push_back(k);
    }
    std::vector<double> step(n, 0.0), theta(n, 0.0);
    bool any = false;
    if (!dd.empty()) {
      HIPCHK(hipMemcpyAsync(g.d_acc, dd.data(), dd.size() * sizeof(opo::AccelDesc), hipMemcpyHostToDevice,
                            E.stream));
      int e = opo::launch_accel_dots(g.fp32, g.d_acc, (int)dd.size(), nres, g.N, g.d_dots, E.stream);
      if (e != 0) return fail(OPO_ERR_HIP, std::string("accel_dots launch: ") + hipGetErrorString((hipError_t)e));
      std::vector<double> h(dd.size() * nres * 9);
      HIPCHK(hipMemcpyAsync(h.data(), g.d_dots, h.size() * sizeof(double), hipMemcpyDeviceToHost, E.stream));
      HIPCHK(hipStreamSynchronize(E.stream));
      for (size_t j = 0; j < dd.size(); ++j) {
        const double *q = h.data() + j * nres * 9;  // signal first
        const double th0 = std::arg(cplx(q[5], q[6])), th1 = std::arg(cplx(q[7], q[8]));
        double d11 = 0, d22 = 0, d12 = 0;
        for (int f = 0; f < nres; ++f) {
          const double *r = q + f * 9, sg = f == 0 ? 1.0 : -1.0;
          const cplx e0 = std::polar(1.0, sg * th0), e1 = std::polar(1.0, sg * th1);
          const cplx s10 = std::conj(e1) * e0 * std::conj(cplx(r[3], r[4]));  // <s1', s0'>
          const cplx s21 = e1 * std::conj(cplx(r[7], r[8]));                   // <s2, s1'>
          const cplx s20 = e0 * std::conj(cplx(r[5], r[6]));                   // <s2, s0'>
          d11 += r[1] + r[0] - 2 * s10.real();
          d22 += r[2] + r[1] - 2 * s21.real();
          d12 += s21.real() - r[1] - s20.real() + s10.real();
        }
        if (!(d11 > 0) || !(d22 > 0)) continue;
        const double rho = d12 / d11;
        const double res2 = (d22 - d12 * d12 / d11) / d22;
        if (rho > kAccelRhoMin && rho < kAccelRhoMax && res2 < kAccelMaxResidual * kAccelMaxResidual) {
          step[who[j]] = std::min(rho / (1 - rho), kAccelMaxStep);
          theta[who[j]] = th1;
          any = true;
        }
      }
    }
    // 2. the jump: every active instance restarts its next round trip from its
    //    snapshot, moved by its step (0 for most)
    if (any) {
      std::vector<opo::AccelDesc> aa;
      for (int k : live) {
        const Instance &I = g.inst[k];
        opo::AccelDesc a{};
        a.s2 = res_of(I);
        a.s1 = step[k] != 0 ? I.acc_hist[I.acc_new] : a.s2;
        a.state = I.state;
        a.step = step[k];
        a.c = std::cos(theta[k]);
        a.s = std::sin(theta[k]);
        aa.push_back(a);
      }
      HIPCHK(hipMemcpyAsync(g.d_acc, aa.data(), aa.size() * sizeof(opo::AccelDesc), hipMemcpyHostToDevice,
                            E.stream));
      int e = opo::launch_accel_apply(g.fp32, g.d_acc, (int)aa.size(), nres, g.N, E.stream);
      if (e != 0) return fail(OPO_ERR_HIP, std::string("accel_apply launch: ") + hipGetErrorString((hipError_t)e));
      RCHK(emit_start(E, g));
      const double zero = 0.0;
      for (int k : live)
        if (step[k] != 0) {
          g.inst[k].acc_n = 0;
          // the off test's decay ratio starts again (its two windows straddled the jump)
          HIPCHK(hipMemcpyAsync(g.d_rprev + k, &zero, sizeof(double), hipMemcpyHostToDevice, E.stream));
        }
      HIPCHK(hipStreamSynchronize(E.stream));  // `zero` and `aa` outlive the copies
    }
    // 3. the snapshot joins the history of every active instance that did not jump
    for (int k : live) {
      if (step[k] != 0) continue;
      Instance &I = g.inst[k];
      for (auto &b : I.acc_hist)
        if (!b) RCHK(dalloc(g, &b, hbytes));
      const int dst = I.acc_n == 0 ? 0 : 1 - I.acc_new;
      HIPCHK(hipMemcpyAsync(I.acc_hist[dst], res_of(I), hbytes, hipMemcpyDeviceToDevice, E.stream));
      I.acc_new = dst;
      I.acc_n = std::min(2, I.acc_n + 1);
    }
  }
  return OPO_OK;
}

// n_rt round trips starting at rt0 (every instance), ending in the time domain.
int engine_run(Engine &E, int rt0, int n_rt, int check_every, double tol) {
  if (n_rt <= 0) return OPO_OK;
  restore_row_tiles(E);  // after a compacted run that ended early on an error
  for (auto &g : E.groups) {
    for (auto &I : g.inst)
      if (!I.uploaded) return fail(OPO_ERR_STATE, "run before upload");
    RCHK(grow_hist(E, g, rt0 + n_rt));
    if (check_every > 0)
      for (auto &I : g.inst)
        if (!I.snap) { RCHK(dalloc(g, &I.snap, (size_t)g.NF * g.N * esize(g))); E.inst_dirty = true; }
  }
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, rt0));
  const bool graph = E.use_graph && n_rt > 2;
  if (graph) RCHK(build_graph(E, check_every, tol));
  if (E.ev0) HIPCHK(hipEventRecord(E.ev0, E.stream));
  RCHK(fork_groups(E));
  for (auto &g : E.groups) RCHK(emit_start(E, g));
  RCHK(join_groups(E));
  int done = 0;
  bool alive = true, compacted = false;
  while (done < n_rt - 1 && alive) {
    const int seg = check_every > 0 ? std::min(check_every, n_rt - 1 - done) : n_rt - 1 - done;
    for (int r = 0; r < seg; ++r) {
      if (graph) {
        HIPCHK(hipGraphLaunch(E.gexec, E.stream));
      } else {
        RCHK(fork_groups(E));
        for (auto &g : E.groups) RCHK(emit_rt_body(E, g, false, check_every, tol));
        RCHK(join_groups(E));
      }
    }
    done += seg;
    if (check_every > 0) {
      bool changed = false;
      RCHK(check_active(E, &alive, &changed));
      if (E.accel_every > 0 && alive && done >= E.accel_start && done % E.accel_every == 0 && done < n_rt - 1)
        RCHK(accel_step(E));
      if (changed) {
        compacted = true;
        E.graph_key = -1;  // the captured launches cover the old instance count
        if (graph && alive && done < n_rt - 1) RCHK(build_graph(E, check_every, tol));
      }
    }
  }
  if (alive) {
    RCHK(fork_groups(E));
    for (auto &g : E.groups) RCHK(emit_rt_body(E, g, true, check_every, tol));
    RCHK(join_groups(E));
  }
  if (E.ev1) {
    HIPCHK(hipEventRecord(E.ev1, E.stream));
    HIPCHK(hipEventSynchronize(E.ev1));
    HIPCHK(hipEventElapsedTime(&E.last_ms, E.ev0, E.ev1));
  } else {
    HIPCHK(hipStreamSynchronize(E.stream));
  }
  E.last_rt0 = rt0; E.last_nrt = n_rt;
  if (compacted) {
    // the next run starts from every instance again: full descriptors, the
    // tuned row tiles, new capture
    E.inst_dirty = true;
    E.graph_key = -1;
    restore_row_tiles(E);
  }
  // surface a launch-time fault immediately
  HIPCHK(hipGetLastError());
  return OPO_OK;
}

// A batch of one grid size with at least kSplitMin instances runs as two
// groups of half the instances each, on two streams (forked and joined inside
// every round trip, as groups of different sizes are).  The two chains then
// overlap: once compaction has shrunk the launches to a few instances, every
// stage is latency-bound and one chain's launches fill the other's gaps
// (headline sweep batch: 52.4 -> 46.9 s, results bit-identical; four groups
// took 59.8 s).  Descriptors hold pointers, so the split changes no arithmetic.
constexpr int kSplitMin = 64;

void split_large_group(Engine &E) {
  if (E.groups.size() != 1 || (int)E.groups[0].members.size() < kSplitMin) return;
  const std::vector<int> all = E.groups[0].members;
  const int N = E.groups[0].N, n = (int)all.size(), h = n / 2;
  E.groups.clear();
  E.groups.resize(2);
  for (int q = 0; q < 2; ++q) {
    Group &g = E.groups[q];
    g.N = N;
    g.members.assign(all.begin() + (q ? h : 0), q ? all.end() : all.begin() + h);
    for (size_t j = 0; j < g.members.size(); ++j) E.where[g.members[j]] = {q, (int)j};
  }
}

int engine_init(Engine &E, const opo_params *params, int count) {
  if (count < 1) return fail(OPO_ERR_INVALID, "count must be >= 1");
  for (int k = 0; k < count; ++k) RCHK(validate(&params[k]));
  const opo_params &p0 = params[0];
  for (int k = 1; k < count; ++k) {
    const opo_params &p = params[k];
    if (p.three_eqs != p0.three_eqs || p.regime != p0.regime || p.NZ != p0.NZ || p.precision != p0.precision ||
        p.device != p0.device)
      return fail(OPO_ERR_INVALID, "batch instances must share three_eqs, regime, NZ, precision and device");
  }
  // OPO_USE_GRAPH=0 launches every stage eagerly (profilers that cannot
  // trace graph-launched kernels); results are identical either way.
  if (const char *ug = std::getenv("OPO_USE_GRAPH")) E.use_graph = ug[0] != '0';
  // OPO_COMPACT=0 keeps every instance in every launch of a converging batch run
  if (const char *cp = std::getenv("OPO_COMPACT")) E.compact = cp[0] != '0';
  // OPO_SPLIT=0 keeps a large one-grid batch in one group (see split_large_group)
  bool split = true;
  if (const char *sp = std::getenv("OPO_SPLIT")) split = sp[0] != '0';
  E.count = count; E.NZ = p0.NZ; E.three_eqs = p0.three_eqs ? 1 : 0; E.regime = p0.regime;
  E.fp32 = p0.precision == OPO_FP32C; E.device = p0.device;
  E.NRT = 1;
  for (int k = 0; k < count; ++k) E.NRT = std::max(E.NRT, params[k].NRT);
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (E.device < 0 || E.device >= ndev) return fail(OPO_ERR_INVALID, "device ordinal out of range");
  HIPCHK(hipSetDevice(E.device));
  HIPCHK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&E.ev0));
  HIPCHK(hipEventCreate(&E.ev1));
  // group by N, preserving order
  E.where.assign(count, {-1, -1});
  for (int k = 0; k < count; ++k) {
    int gi = -1;
    for (size_t j = 0; j < E.groups.size(); ++j)
      if (E.groups[j].N == params[k].N) gi = (int)j;
    if (gi < 0) { E.groups.emplace_back(); gi = (int)E.groups.size() - 1; E.groups[gi].N = params[k].N; }
    E.where[k] = {gi, (int)E.groups[gi].members.size()};
    E.groups[gi].members.push_back(k);
  }
  if (split) split_large_group(E);
  HIPCHK(hipEventCreateWithFlags(&E.fork_ev, hipEventDisableTiming));
  for (auto &g : E.groups) {
    for (int k : g.members) if (params[k].GDD != 0) g.any_gdd = true;
    g.cur = E.stream;
    if (E.groups.size() > 1) {
      HIPCHK(hipStreamCreateWithFlags(&g.gs, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&g.done_ev, hipEventDisableTiming));
    }
    RCHK(make_group(E, g));
    for (size_t j = 0; j < g.members.size(); ++j) g.inst[j].prm = params[g.members[j]];
  }
  return OPO_OK;
}

}  // namespace

namespace opo {
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace opo

// ================================================================== C ABI

struct opo_solver {
  Engine E;
};
struct opo_batch {
  Engine E;
};

extern "C" {

const char *opo_strerror(int status) {
  switch (status) {
    case OPO_OK: return "success";
    case OPO_ERR_INVALID: return "invalid argument";
    case OPO_ERR_HIP: return "HIP runtime error";
    case OPO_ERR_NOMEM: return "device out of memory";
    case OPO_ERR_STATE: return "call out of order";
    case OPO_ERR_DIVERGED: return "non-finite intracavity power";
  }
  return "unknown status";
}

const char *opo_last_error(void) { return g_err.c_str(); }

int opo_device_count(int *n) {
  if (!n) return fail(OPO_ERR_INVALID, "null");
  HIPCHK(hipGetDeviceCount(n));
  return OPO_OK;
}

int opo_create(const opo_params *p, opo_solver **out) {
  if (!out) return fail(OPO_ERR_INVALID, "null out");
  *out = nullptr;
  std::unique_ptr<opo_solver> s(new opo_solver());
  RCHK(engine_init(s->E, p, 1));
  s->E.solver = true;
  *out = s.release();
  return OPO_OK;
}

int opo_destroy(opo_solver *s) {
  delete s;
  return OPO_OK;
}

int opo_upload(opo_solver *s, const double *w, const double *T, const double *Ap_in, size_t n_ap, const double *Ap,
               const double *As, const double *Ai) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  Engine &E = s->E;
  HIPCHK(hipSetDevice(E.device));
  return upload_instance(E, E.groups[0], 0, w, T, Ap_in, n_ap, Ap, As, Ai);
}

int opo_set_fields(opo_solver *s, const double *Ap, const double *As, const double *Ai) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  if (!g.inst[0].uploaded) return fail(OPO_ERR_STATE, "set_fields before upload");
  std::vector<cd> st((size_t)g.NF * g.N, cd(0, 0));
  for (int j = 0; j < g.N; ++j) {
    if (Ap) st[j] = cd(Ap[2 * j], Ap[2 * j + 1]);
    if (As) st[(size_t)g.N + j] = cd(As[2 * j], As[2 * j + 1]);
    if (g.NF == 3 && Ai) st[(size_t)2 * g.N + j] = cd(Ai[2 * j], Ai[2 * j + 1]);
  }
  return upload_c(E, g, g.inst[0].state, st.data(), st.size());
}

namespace {

// save slots (CW: the last Nrts round trips; NS: every round trip) for every
// instance of the engine, src/cuOPO.cu:570-608
int engine_enable_save(Engine &E, int enable) {
  if (enable) {
    for (auto &g : E.groups)
      for (auto &I : g.inst) {
        const size_t slots = I.prm.regime == OPO_CW ? (size_t)I.prm.Nrts : (size_t)I.prm.NRT;
        for (int f = 0; f < g.NF; ++f)
          if (!I.save[f]) {
            RCHK(dalloc(g, &I.save[f], slots * g.N * esize(g)));
            HIPCHK(hipMemsetAsync(I.save[f], 0, slots * g.N * esize(g), E.stream));
          }
      }
    HIPCHK(hipStreamSynchronize(E.stream));
  }
  E.save_enabled = enable != 0;
  E.inst_dirty = true;
  E.graph_key = -1;
  return OPO_OK;
}

int engine_download_saved(Engine &E, Group &g, Instance &I, double *sp, double *ss, double *si) {
  const size_t slots = I.prm.regime == OPO_CW ? (size_t)I.prm.Nrts : (size_t)I.prm.NRT;
  double *dst[3] = {sp, ss, si};
  for (int f = 0; f < g.NF; ++f) {
    if (!dst[f]) continue;
    if (!I.save[f]) return fail(OPO_ERR_STATE, "saving was not enabled");
    RCHK(download_c(E, g, reinterpret_cast<cd *>(dst[f]), I.save[f], slots * g.N));
  }
  return OPO_OK;
}

}  // namespace

int opo_enable_save(opo_solver *s, int enable) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  return engine_enable_save(s->E, enable);
}

int opo_set_tiles(opo_solver *s, int col_threads, int row_threads) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  for (int v : {col_threads, row_threads})
    if (v != 0 && v != 32 && v != 64 && v != 128 && v != 256 && v != 512)
      return fail(OPO_ERR_INVALID, "tile threads in {32,64,128,256,512}");
  s->E.col_threads = col_threads; s->E.row_threads = row_threads;
  for (auto &g : s->E.groups) pick_tiles(s->E, g);
  s->E.graph_key = -1;
  return OPO_OK;
}

int opo_set_xcd_pack(opo_solver *s, int enable) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  s->E.xcd_pack = enable ? 1 : 0;
  s->E.graph_key = -1;
  return OPO_OK;
}

int opo_use_graph(opo_solver *s, int enable) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  s->E.use_graph = enable != 0;
  return OPO_OK;
}

int opo_crystal_pass(opo_solver *s) {
  using namespace opo;
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  if (!g.inst[0].uploaded) return fail(OPO_ERR_STATE, "crystal_pass before upload");
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  const int A = all_mask(g);
  RCHK(emit_col(E, g, OP_NL_START, A, 0, A, A, 0, 0, 0, 0));
  for (int sl = 0; sl < E.NZ; ++sl) {
    RCHK(emit_row(E, g, false, A));
    if (sl < E.NZ - 1) RCHK(emit_col(E, g, OP_NL_END | OP_NL_START, A, A, A, A, sl, sl + 1, 0, 0));
  }
  RCHK(emit_col(E, g, OP_NL_END, A, A, 0, A, E.NZ - 1, 0, 0, 0));
  HIPCHK(hipStreamSynchronize(E.stream));
  return OPO_OK;
}

int opo_linear_step(opo_solver *s) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  if (!g.inst[0].uploaded) return fail(OPO_ERR_STATE, "linear_step before upload");
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  const int A = all_mask(g);
  RCHK(emit_col(E, g, 0, A, 0, A, A, 0, 0, 0, 0));
  RCHK(emit_row(E, g, false, A));
  RCHK(emit_col(E, g, 0, A, A, 0, A, 0, 0, 0, 0));
  HIPCHK(hipStreamSynchronize(E.stream));
  return OPO_OK;
}

int opo_nl_half(opo_solver *s, int slice, int n_times) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  if (!g.inst[0].uploaded) return fail(OPO_ERR_STATE, "nl_half before upload");
  if (slice < 0 || slice >= E.NZ) return fail(OPO_ERR_INVALID, "slice out of range");
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  const int A = all_mask(g);
  for (int r = 0; r < n_times; ++r) RCHK(emit_col(E, g, opo::OP_NL_START, A, 0, 0, A, 0, slice, 0, 0));
  HIPCHK(hipStreamSynchronize(E.stream));
  return OPO_OK;
}

int opo_boundary(opo_solver *s) {
  using namespace opo;
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  if (!g.inst[0].uploaded) return fail(OPO_ERR_STATE, "boundary before upload");
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  const int A = all_mask(g), SI = si_mask(g);
  if (g.any_gdd) {
    RCHK(emit_col(E, g, OP_EOM, A, 0, SI, A, 0, 0, 0, 0));
    RCHK(emit_row(E, g, true, SI));
    RCHK(emit_col(E, g, OP_MIRROR, A, SI, 0, A, 0, 0, 0, 0));
  } else {
    RCHK(emit_col(E, g, OP_EOM | OP_MIRROR, A, 0, 0, A, 0, 0, 0, 0));
  }
  HIPCHK(hipStreamSynchronize(E.stream));
  return OPO_OK;
}

int opo_run(opo_solver *s, int rt0, int n_rt) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  if (rt0 < 0 || n_rt < 0) return fail(OPO_ERR_INVALID, "negative round-trip index");
  if (n_rt > INT_MAX - rt0) return fail(OPO_ERR_INVALID, "rt0 + n_rt overflows int");
  HIPCHK(hipSetDevice(s->E.device));
  return engine_run(s->E, rt0, n_rt, 0, 0.0);
}

// opo_run on caller-owned device buffers, with every other device buffer
// carved from the static per-GPU scratch array: no hipMalloc, hipFree, stream,
// event or graph is created.  Same stage kernels and default launch settings
// as a fresh solver of the same parameters, launched eagerly on `stream`.
int opo_run_noalloc(const opo_params *params, const double *w, const double *T, const void *d_pump, size_t n_ap,
                    void *d_fields, int rt0, int n_rt, double *power_out, void *stream) {
  RCHK(validate(params));
  if (!w || !d_pump || !d_fields) return fail(OPO_ERR_INVALID, "w, d_pump and d_fields are required");
  if (rt0 < 0 || n_rt < 1) return fail(OPO_ERR_INVALID, "rt0 must be >= 0 and n_rt >= 1");
  if (n_rt > INT_MAX - rt0) return fail(OPO_ERR_INVALID, "rt0 + n_rt overflows int");
  const size_t align = params->precision == OPO_FP32C ? 8 : 16;
  if ((reinterpret_cast<uintptr_t>(d_fields) | reinterpret_cast<uintptr_t>(d_pump)) % align != 0)
    return fail(OPO_ERR_INVALID, "d_fields and d_pump must be aligned to the complex element size");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (params->device < 0 || params->device >= ndev) return fail(OPO_ERR_INVALID, "device ordinal out of range");
  HIPCHK(hipSetDevice(params->device));
  void *base = nullptr;
  size_t cap = 0;
  if (const int e = opo::scratch_region(&base, &cap))
    return fail(OPO_ERR_HIP, std::string("scratch address: ") + hipGetErrorString((hipError_t)e));

  Engine E;
  E.own_stream = false;
  E.stream = reinterpret_cast<hipStream_t>(stream);
  E.use_graph = false;
  E.solver = true;
  E.count = 1; E.NZ = params->NZ; E.three_eqs = params->three_eqs ? 1 : 0; E.regime = params->regime;
  E.fp32 = params->precision == OPO_FP32C; E.device = params->device;
  E.NRT = std::max(params->NRT, rt0 + n_rt);  // the power history covers every round trip run
  E.where.assign(1, {0, 0});
  E.groups.emplace_back();
  Group &g = E.groups[0];
  g.N = params->N;
  g.members.push_back(0);
  g.any_gdd = params->GDD != 0;
  g.cur = E.stream;
  g.pool = static_cast<char *>(base);
  g.pool_cap = cap;
  g.ext_state = d_fields;
  g.ext_pump = d_pump;
  g.phase_mode = 0;  // the full phase table (the factorised tables are not built here)
  RCHK(make_group(E, g));
  g.inst[0].prm = *params;
  RCHK(upload_instance(E, g, 0, w, T, nullptr, n_ap, nullptr, nullptr, nullptr));
  RCHK(engine_run(E, rt0, n_rt, 0, 0.0));
  if (power_out) {
    HIPCHK(hipMemcpyAsync(power_out, g.d_hist + rt0, (size_t)n_rt * sizeof(double), hipMemcpyDeviceToHost,
                          E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  }
  return OPO_OK;
}

int opo_download(opo_solver *s, double *Ap, double *As, double *Ai) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  std::vector<cd> st((size_t)g.NF * g.N);
  RCHK(download_c(E, g, st.data(), g.inst[0].state, st.size()));
  if (Ap) std::memcpy(Ap, st.data(), g.N * sizeof(cd));
  if (As) std::memcpy(As, st.data() + g.N, g.N * sizeof(cd));
  if (Ai && g.NF == 3) std::memcpy(Ai, st.data() + 2 * (size_t)g.N, g.N * sizeof(cd));
  return OPO_OK;
}

int opo_download_saved(opo_solver *s, double *sp, double *ss, double *si) {
  if (!s) return fail(OPO_ERR_INVALID, "null solver");
  HIPCHK(hipSetDevice(s->E.device));
  return engine_download_saved(s->E, s->E.groups[0], s->E.groups[0].inst[0], sp, ss, si);
}

int opo_power_history(opo_solver *s, double *out, int n) {
  if (!s || !out) return fail(OPO_ERR_INVALID, "null");
  HIPCHK(hipSetDevice(s->E.device));
  Engine &E = s->E;
  Group &g = E.groups[0];
  if (n < 0) return fail(OPO_ERR_INVALID, "n must be >= 0");
  // the last run holds last_nrt values; the rest of out is zero-filled
  const int have = std::min(n, E.last_nrt);
  if (have > 0) HIPCHK(hipMemcpy(out, g.d_hist + E.last_rt0, have * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = std::max(have, 0); k < n; ++k) out[k] = 0.0;
  return OPO_OK;
}

// average kernel time of the slice-interior column stage and the LIN row stage
// Average device time of one column-stage and one row-stage launch, each
// timed alone between events on the group's stream (what a kernel profiler
// reports per dispatch).
int engine_time_stages(Engine &E, Group &g, int iters, double *col_us, double *row_us) {
  using namespace opo;
  for (auto &I : g.inst)
    if (!I.uploaded) return fail(OPO_ERR_STATE, "time_stages before upload");
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  const int A = all_mask(g);
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  double tc = 0, tr = 0;
  int rc = OPO_OK;
  for (int it = 0; it < iters + 1 && rc == OPO_OK; ++it) {  // first iteration warms up
    float ms = 0;
    rc = emit_col(E, g, OP_NL_END | OP_NL_START, A, A, A, A, 0, 1 % E.NZ, 0, 0, a, b);
    if (rc != OPO_OK) break;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) { rc = fail(OPO_ERR_HIP, "event"); break; }
    if (it) tc += ms;
    rc = emit_row(E, g, false, A, a, b);
    if (rc != OPO_OK) break;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) { rc = fail(OPO_ERR_HIP, "event"); break; }
    if (it) tr += ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != OPO_OK) return rc;
  if (col_us) *col_us = 1e3 * tc / iters;
  if (row_us) *row_us = 1e3 * tr / iters;
  return OPO_OK;
}

// Average device time of one crystal slice as a round trip runs it: after a
// warm-up slice, iters back-to-back (column stage, row stage) pairs between
// two events on the group's stream.  The autotuner's objective.
int engine_time_slices(Engine &E, Group &g, int iters, double *slice_us) {
  using namespace opo;
  for (auto &I : g.inst)
    if (!I.uploaded) return fail(OPO_ERR_STATE, "time_slices before upload");
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  const int A = all_mask(g);
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = OPO_OK;
  for (auto &e : ev)
    if (rc == OPO_OK && hipEventCreate(&e) != hipSuccess) rc = fail(OPO_ERR_HIP, "event create");
  auto slice = [&] {
    const int r = emit_col(E, g, OP_NL_END | OP_NL_START, A, A, A, A, 0, 1 % E.NZ, 0, 0);
    return r != OPO_OK ? r : emit_row(E, g, false, A);
  };
  if (rc == OPO_OK) rc = slice();
  if (rc == OPO_OK && hipEventRecord(ev[0], g.cur) != hipSuccess) rc = fail(OPO_ERR_HIP, "event record");
  for (int it = 0; it < iters && rc == OPO_OK; ++it) rc = slice();
  if (rc == OPO_OK && hipEventRecord(ev[1], g.cur) != hipSuccess) rc = fail(OPO_ERR_HIP, "event record");
  float ms = 0;
  if (rc == OPO_OK && (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
    rc = fail(OPO_ERR_HIP, "event");
  if (rc != OPO_OK) (void)hipStreamSynchronize(g.cur);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc != OPO_OK) return rc;
  *slice_us = 1e3 * ms / iters;
  return OPO_OK;
}

// Every group's slices side by side, as a round trip runs them: between one
// fork and one join each group's stream runs `iters` back-to-back (column,
// row) pairs; *slice_us is the time from the fork to the join per slice.
int engine_time_joint(Engine &E, int iters, double *slice_us) {
  using namespace opo;
  RCHK(sync_insts(E));
  RCHK(reset_counters(E, 0));
  auto slices = [&](int k) -> int {
    RCHK(fork_groups(E));
    int r = OPO_OK;
    for (auto &g : E.groups) {
      const int A = all_mask(g);
      for (int it = 0; it < k && r == OPO_OK; ++it) {
        r = emit_col(E, g, OP_NL_END | OP_NL_START, A, A, A, A, 0, 1 % E.NZ, 0, 0);
        if (r == OPO_OK) r = emit_row(E, g, false, A);
      }
    }
    const int j = join_groups(E);
    return r != OPO_OK ? r : j;
  };
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = OPO_OK;
  for (auto &e : ev)
    if (rc == OPO_OK && hipEventCreate(&e) != hipSuccess) rc = fail(OPO_ERR_HIP, "event create");
  if (rc == OPO_OK) rc = slices(1);
  if (rc == OPO_OK && hipEventRecord(ev[0], E.stream) != hipSuccess) rc = fail(OPO_ERR_HIP, "event record");
  if (rc == OPO_OK) rc = slices(iters);
  if (rc == OPO_OK && hipEventRecord(ev[1], E.stream) != hipSuccess) rc = fail(OPO_ERR_HIP, "event record");
  float ms = 0;
  if (rc == OPO_OK && (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
    rc = fail(OPO_ERR_HIP, "event");
  if (rc != OPO_OK) {
    for (auto &g : E.groups)
      if (g.gs) (void)hipStreamSynchronize(g.gs);
    (void)hipStreamSynchronize(E.stream);
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc != OPO_OK) return rc;
  *slice_us = 1e3 * ms / iters;
  return OPO_OK;
}

int opo_time_stages(opo_solver *s, int iters, double *col_us, double *row_us) {
  if (!s || iters < 1) return fail(OPO_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->E.device));
  return engine_time_stages(s->E, s->E.groups[0], iters, col_us, row_us);
}

int opo_copy_probe(size_t bytes, int iters, int device, double *gbps, double *us) {
  if (bytes < 16 || iters < 1) return fail(OPO_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(device));
  void *x = nullptr, *y = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  auto cleanup = [&]() {
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st) (void)hipStreamDestroy(st);
  };
  const long long n = (long long)(bytes / 16);
  if (hipMalloc(&x, n * 16) != hipSuccess || hipMalloc(&y, n * 16) != hipSuccess) {
    cleanup();
    return fail(OPO_ERR_NOMEM, "copy probe allocation");
  }
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&a) != hipSuccess ||
      hipEventCreate(&b) != hipSuccess || hipMemsetAsync(x, 0, n * 16, st) != hipSuccess) {
    cleanup();
    return fail(OPO_ERR_HIP, "copy probe setup");
  }
  double tot = 0;
  for (int it = 0; it < iters + 1; ++it) {
    float ms = 0;
    if (opo::launch_copy(x, y, n, st, a, b) != 0 || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&ms, a, b) != hipSuccess) {
      cleanup();
      return fail(OPO_ERR_HIP, "copy probe launch");
    }
    if (it) tot += ms;
  }
  cleanup();
  const double sec = tot / iters * 1e-3;
  if (gbps) *gbps = 2.0 * n * 16 / sec / 1e9;
  if (us) *us = sec * 1e6;
  return OPO_OK;
}

int opo_last_run_ms(opo_solver *s, double *ms) {
  if (!s || !ms) return fail(OPO_ERR_INVALID, "null");
  *ms = s->E.last_ms;
  return OPO_OK;
}

// ---------------------------------------------------------------- batches

int opo_batch_create(const opo_params *params, int count, opo_batch **out) {
  if (!out || !params) return fail(OPO_ERR_INVALID, "null");
  *out = nullptr;
  std::unique_ptr<opo_batch> b(new opo_batch());
  RCHK(engine_init(b->E, params, count));
  *out = b.release();
  return OPO_OK;
}

int opo_batch_enable_save(opo_batch *b, int enable) {
  if (!b) return fail(OPO_ERR_INVALID, "null batch");
  HIPCHK(hipSetDevice(b->E.device));
  return engine_enable_save(b->E, enable);
}

int opo_batch_download_saved(opo_batch *b, int idx, double *sp, double *ss, double *si) {
  if (!b || idx < 0 || idx >= b->E.count) return fail(OPO_ERR_INVALID, "bad batch index");
  HIPCHK(hipSetDevice(b->E.device));
  const auto gs = b->E.where[idx];
  Group &g = b->E.groups[gs.first];
  return engine_download_saved(b->E, g, g.inst[gs.second], sp, ss, si);
}

int opo_batch_destroy(opo_batch *b) {
  delete b;
  return OPO_OK;
}

int opo_batch_upload(opo_batch *b, int idx, const double *w, const double *T, const double *Ap_in, size_t n_ap,
                     const double *Ap, const double *As, const double *Ai) {
  if (!b || idx < 0 || idx >= b->E.count) return fail(OPO_ERR_INVALID, "bad batch index");
  Engine &E = b->E;
  HIPCHK(hipSetDevice(E.device));
  auto [gi, k] = E.where[idx];
  return upload_instance(E, E.groups[gi], k, w, T, Ap_in, n_ap, Ap, As, Ai);
}

int opo_batch_set_tiles(opo_batch *b, int col_threads, int row_threads) {
  if (!b) return fail(OPO_ERR_INVALID, "null batch");
  for (int v : {col_threads, row_threads})
    if (v != 0 && v != 32 && v != 64 && v != 128 && v != 256 && v != 512)
      return fail(OPO_ERR_INVALID, "tile threads in {32,64,128,256,512}");
  b->E.col_threads = col_threads; b->E.row_threads = row_threads;
  for (auto &g : b->E.groups) pick_tiles(b->E, g);
  b->E.graph_key = -1;
  return OPO_OK;
}

int opo_batch_run_ex(opo_batch *b, int max_rt, int check_every, double tol, double off_ratio, int *rts_out,
                     int *states_out) {
  if (!b) return fail(OPO_ERR_INVALID, "null batch");
  if (max_rt < 1) return fail(OPO_ERR_INVALID, "max_rt must be >= 1");
  Engine &E = b->E;
  HIPCHK(hipSetDevice(E.device));
  if (tol <= 0) check_every = 0;
  E.off_ratio = off_ratio;  // part of the round-trip graph's cache key
  RCHK(engine_run(E, 0, max_rt, check_every, tol));
  int rc = OPO_OK;
  for (auto &g : E.groups) {
    const int n = (int)g.members.size();
    std::vector<int> conv(n), stat(n), ran(n);
    HIPCHK(hipMemcpy(conv.data(), g.d_conv, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(stat.data(), g.d_status, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ran.data(), g.d_rt, n * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) {
      // a diverged instance stopped at the round trip whose power was non-finite
      if (rts_out) rts_out[g.members[k]] = stat[k] < 0 ? ran[k] : conv[k];
      if (stat[k] < 0) rc = OPO_ERR_DIVERGED;
      if (states_out) states_out[g.members[k]] = stat[k] < 0 ? -1 : (conv[k] < 0 ? 0 : (stat[k] == 1 ? 2 : 1));
    }
  }
  if (rc != OPO_OK) return fail(rc, "non-finite intracavity power in at least one instance");
  return OPO_OK;
}

int opo_batch_run(opo_batch *b, int max_rt, int check_every, double tol, int *rts_out) {
  return opo_batch_run_ex(b, max_rt, check_every, tol, 1e-6, rts_out, nullptr);
}

int opo_batch_download(opo_batch *b, int idx, double *Ap, double *As, double *Ai) {
  if (!b || idx < 0 || idx >= b->E.count) return fail(OPO_ERR_INVALID, "bad batch index");
  HIPCHK(hipSetDevice(b->E.device));
  Engine &E = b->E;
  auto [gi, k] = E.where[idx];
  Group &g = E.groups[gi];
  int conv = -1;
  HIPCHK(hipMemcpy(&conv, g.d_conv + k, sizeof(int), hipMemcpyDeviceToHost));
  const void *src = (conv >= 0 && g.inst[k].snap) ? g.inst[k].snap : g.inst[k].state;
  std::vector<cd> st((size_t)g.NF * g.N);
  RCHK(download_c(E, g, st.data(), src, st.size()));
  if (Ap) std::memcpy(Ap, st.data(), g.N * sizeof(cd));
  if (As) std::memcpy(As, st.data() + g.N, g.N * sizeof(cd));
  if (Ai && g.NF == 3) std::memcpy(Ai, st.data() + 2 * (size_t)g.N, g.N * sizeof(cd));
  return OPO_OK;
}

int opo_batch_power_history(opo_batch *b, int idx, double *out, int n) {
  if (!b || !out || idx < 0 || idx >= b->E.count || n < 0) return fail(OPO_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(b->E.device));
  Engine &E = b->E;
  auto [gi, k] = E.where[idx];
  Group &g = E.groups[gi];
  n = std::min(n, g.hist_len);
  HIPCHK(hipMemcpy(out, g.d_hist + (size_t)k * g.hist_len, n * sizeof(double), hipMemcpyDeviceToHost));
  return OPO_OK;
}

int opo_batch_power(opo_batch *b, double *out) {
  if (!b || !out) return fail(OPO_ERR_INVALID, "null");
  HIPCHK(hipSetDevice(b->E.device));
  Engine &E = b->E;
  for (auto &g : E.groups) {
    const int n = (int)g.members.size();
    std::vector<int> rt(n);
    HIPCHK(hipMemcpy(rt.data(), g.d_rt, n * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) {
      const int last = std::max(0, std::min(rt[k], g.hist_len) - 1);
      HIPCHK(hipMemcpy(out + g.members[k], g.d_hist + (size_t)k * g.hist_len + last, sizeof(double),
                       hipMemcpyDeviceToHost));
    }
  }
  return OPO_OK;
}

namespace {

void get_tuning(const Engine &E, const Group &g, opo_tuning *t) {
  std::memset(t, 0, sizeof(*t));
  t->N = g.N;
  t->count = (int)g.members.size();
  t->col_threads = g.col_threads ? g.col_threads : E.col_threads;
  t->row_threads = g.row_threads ? g.row_threads : E.row_threads;
  t->xcd_pack = g.xcd_pack >= 0 ? g.xcd_pack : E.xcd_pack;
  t->phase_mode = g.phase_mode;
  t->col_spt = g.ce;
  t->row_spt = g.re;
  t->col_tile = g.ct;
  t->row_tile = g.rt;
  t->factored = 0;
  for (const auto &I : g.inst) t->factored |= (I.use_fac && group_factored(g)) ? 1 : 0;
  t->slice_us = g.tuned_us;
}

int set_tuning(Engine &E, Group &g, const opo_tuning *t) {
  for (int v : {t->col_threads, t->row_threads})
    if (v != 0 && v != 32 && v != 64 && v != 128 && v != 256 && v != 512)
      return fail(OPO_ERR_INVALID, "tile threads in {0,32,64,128,256,512}");
  if (t->phase_mode < -1 || t->phase_mode > 1) return fail(OPO_ERR_INVALID, "phase_mode in {-1,0,1}");
  for (int v : {t->col_spt, t->row_spt})
    if (v != 0 && v != 4 && v != 8) return fail(OPO_ERR_INVALID, "samples per thread in {0,4,8}");
  const Group old = g;  // restored if the settings do not fit the kernels
  g.col_threads = t->col_threads;
  g.row_threads = t->row_threads;
  g.xcd_pack = t->xcd_pack < 0 ? -1 : (t->xcd_pack ? 1 : 0);
  g.phase_mode = t->phase_mode;
  g.ce = t->col_spt ? t->col_spt : 8;
  g.re = t->row_spt ? t->row_spt : 8;
  g.tuned_us = 0;
  pick_tiles(E, g);
  const int rc = check_shapes(g);
  if (rc != OPO_OK) {
    g.col_threads = old.col_threads; g.row_threads = old.row_threads; g.xcd_pack = old.xcd_pack;
    g.phase_mode = old.phase_mode; g.ce = old.ce; g.re = old.re; g.tuned_us = old.tuned_us;
    g.ct = old.ct; g.rt = old.rt;
    return rc;
  }
  E.graph_key = -1;
  E.inst_dirty = true;
  return OPO_OK;
}

bool same_setting(const opo_tuning &a, const opo_tuning &b) {
  return a.col_threads == b.col_threads && a.row_threads == b.row_threads && a.xcd_pack == b.xcd_pack &&
         a.phase_mode == b.phase_mode && a.col_spt == b.col_spt && a.row_spt == b.row_spt;
}

// A round trip runs the groups of different grid sizes side by side on their
// streams, and a setting timed for its group alone need not be the best beside
// the others (on the 256-point batch of four grid sizes, the N = 2^15 group's
// pick on every group ran the round trip 2-6% faster than each group's own
// pick).  So each group's pick is also tried on every group, with all groups
// timed together (engine_time_joint, the better of two timings), and the
// fastest assignment is kept.  The fields are saved before and restored after.
int autotune_joint(Engine &E, int iters) {
  const size_t G = E.groups.size();
  std::vector<opo_tuning> own(G);
  std::vector<double> own_us(G);
  for (size_t i = 0; i < G; ++i) {
    get_tuning(E, E.groups[i], &own[i]);
    own_us[i] = E.groups[i].tuned_us;
  }
  std::vector<std::vector<opo_tuning>> cands;
  for (size_t j = 0; j < G; ++j) {
    std::vector<opo_tuning> c(G);
    bool differs = false;
    for (size_t i = 0; i < G; ++i) {
      c[i] = own[j];
      c[i].N = own[i].N;
      c[i].count = own[i].count;
      differs = differs || !same_setting(c[i], own[i]);
    }
    bool dup = false;
    for (const auto &x : cands) {
      bool eq = true;
      for (size_t i = 0; i < G; ++i) eq = eq && same_setting(x[i], c[i]);
      dup = dup || eq;
    }
    if (differs && !dup) cands.push_back(c);
  }
  if (cands.empty()) return OPO_OK;  // one grid size, or every group took the same setting

  std::vector<void *> bak(G, nullptr);
  struct Frees {
    std::vector<void *> &v;
    ~Frees() {
      for (void *p : v)
        if (p) (void)hipFree(p);
    }
  } frees{bak};
  for (size_t i = 0; i < G; ++i) {
    Group &g = E.groups[i];
    const size_t bytes = (size_t)g.NF * g.N * esize(g);
    HIPCHK(hipMalloc(&bak[i], bytes * g.members.size()));
    for (size_t k = 0; k < g.members.size(); ++k)
      HIPCHK(hipMemcpyAsync((char *)bak[i] + bytes * k, g.inst[k].state, bytes, hipMemcpyDeviceToDevice, E.stream));
  }
  HIPCHK(hipStreamSynchronize(E.stream));
  auto apply = [&](const std::vector<opo_tuning> &ts) -> bool {
    for (size_t i = 0; i < G; ++i) {
      opo_tuning t = ts[i];
      if (set_tuning(E, E.groups[i], &t) != OPO_OK) return false;
    }
    return true;
  };
  auto timed = [&](double *us) -> int {
    double a = 0, b = 0;
    RCHK(engine_time_joint(E, iters, &a));
    RCHK(engine_time_joint(E, iters, &b));
    *us = std::min(a, b);
    return OPO_OK;
  };
  double best_us = 0;
  int rc = timed(&best_us);  // every group's own pick is in place
  int best = -1;             // -1: the own picks
  for (size_t c = 0; c < cands.size() && rc == OPO_OK; ++c) {
    if (!apply(cands[c])) continue;  // the setting does not fit some group's kernels
    double us = 0;
    rc = timed(&us);
    if (rc == OPO_OK && us < best_us) { best_us = us; best = (int)c; }
  }
  // the winner in place, with each group's own slice time under it
  if (!apply(best < 0 ? own : cands[(size_t)best]))
    return fail(OPO_ERR_INVALID, "internal: a tuned setting no longer applies");
  for (size_t i = 0; i < G; ++i) {
    Group &g = E.groups[i];
    if (best < 0 || same_setting(cands[(size_t)best][i], own[i])) {
      g.tuned_us = own_us[i];
    } else if (rc == OPO_OK) {
      double u = 0;
      rc = engine_time_slices(E, g, iters, &u);
      g.tuned_us = u;
    }
  }
  for (size_t i = 0; i < G; ++i) {
    Group &g = E.groups[i];
    const size_t bytes = (size_t)g.NF * g.N * esize(g);
    for (size_t k = 0; k < g.members.size(); ++k)
      HIPCHK(hipMemcpyAsync(g.inst[k].state, (char *)bak[i] + bytes * k, bytes, hipMemcpyDeviceToDevice, E.stream));
  }
  HIPCHK(hipStreamSynchronize(E.stream));
  return rc;
}

// Time the candidates of each group with engine_time_slices (back-to-back
// slices, as a round trip runs them; the better of two timings) and keep the
// fastest: for each XCD packing, the column settings (tile of 1/2, 1 or 2
// cache lines, and 1/4 for small groups, x 4 or 8 samples per thread) with the
// row settings fixed, then the row settings (64..512-thread tiles, and 32 for
// small groups, x 4 or 8 samples per thread x phase source) with the best
// column settings, then the column settings once more beside those rows.
// With several grid sizes, autotune_joint then weighs the picks side by side.
// The timed slices mutate the fields, so each group's state is saved before
// and restored after.
int engine_autotune(Engine &E, int iters) {
  for (size_t gi = 0; gi < E.groups.size(); ++gi) {
    Group &g = E.groups[gi];
    for (auto &I : g.inst)
      if (!I.uploaded) return fail(OPO_ERR_STATE, "autotune before upload");
    // the groups of one grid size (a large one-size batch split in two) share
    // the settings tuned on the first of them
    bool shared = false;
    for (size_t j = 0; j < gi && !shared; ++j) {
      if (E.groups[j].N != g.N) continue;
      opo_tuning t;
      get_tuning(E, E.groups[j], &t);
      RCHK(set_tuning(E, g, &t));
      g.tuned_us = E.groups[j].tuned_us;
      shared = true;
    }
    if (shared) continue;
    const size_t bytes = (size_t)g.NF * g.N * esize(g);
    const int n = (int)g.members.size();
    void *bak = nullptr;
    HIPCHK(hipMalloc(&bak, bytes * n));
    std::unique_ptr<void, void (*)(void *)> guard(bak, [](void *p) { (void)hipFree(p); });
    for (int k = 0; k < n; ++k)
      HIPCHK(hipMemcpyAsync((char *)bak + bytes * k, g.inst[k].state, bytes, hipMemcpyDeviceToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    opo_tuning base;
    get_tuning(E, g, &base);
    const int ct0 = (int)(128 / esize(g));
    std::vector<int> xcds = {0};
    if ((long long)n * g.N <= (1LL << 16)) xcds.push_back(1);
    // latency-bound groups (one launch is a single wave of blocks) also try
    // quarter-line column tiles and one-row row tiles: more, smaller blocks
    const bool small = (long long)n * g.N <= (1LL << 17);
    std::vector<int> cts = {ct0 / 2, ct0, ct0 * 2}, rts = {64, 128, 256, 512};
    if (small) { cts.insert(cts.begin(), ct0 / 4); rts.insert(rts.begin(), 32); }
    bool can_fac = true;
    for (auto &I : g.inst) can_fac &= I.use_fac;
    std::vector<int> phases = {0};
    if (can_fac) phases.push_back(1);
    // the better of two timings: one pass can catch a clock or cache transient
    auto time_it = [&](const opo_tuning &t, double *us) -> int {
      opo_tuning tt = t;
      if (set_tuning(E, g, &tt) != OPO_OK) return 1;  // shape does not fit: skip
      RCHK(sync_insts(E));
      double u0 = 0, u1 = 0;
      RCHK(engine_time_slices(E, g, iters, &u0));
      RCHK(engine_time_slices(E, g, iters, &u1));
      *us = std::min(u0, u1);
      return OPO_OK;
    };
    opo_tuning best = base;
    double best_us = 1e30;
    // column settings with the row settings of `t`: the best (and its time) into tc, bc
    auto col_pass = [&](const opo_tuning &t, opo_tuning *tc, double *bc) -> int {
      for (int spt : {8, 4})
        for (int ct : cts) {
          const int th = ct * (g.N1 / spt);
          if (th < 32 || th > 512 || (th & (th - 1))) continue;
          opo_tuning c = t;
          c.col_threads = th; c.col_spt = spt;
          double us = 0;
          const int rc = time_it(c, &us);
          if (rc < 0) return rc;
          if (rc == OPO_OK && us < *bc) { *bc = us; *tc = c; }
        }
      return OPO_OK;
    };
    for (int x : xcds) {
      opo_tuning t = base;
      t.xcd_pack = x;
      t.row_threads = 64; t.row_spt = 8; t.phase_mode = -1;
      double bc = 1e30;
      opo_tuning tc = t;
      RCHK(col_pass(t, &tc, &bc));
      double br = 1e30;
      opo_tuning tr = tc;
      for (int spt : {8, 4})
        for (int r : rts)
          for (int ph : phases) {
            opo_tuning c = tc;
            c.row_threads = r; c.row_spt = spt; c.phase_mode = ph;
            double us = 0;
            const int rc = time_it(c, &us);
            if (rc < 0) return rc;
            if (rc == OPO_OK && us < br) { br = us; tr = c; }
          }
      // the column settings again, now beside the chosen rows: the first pass
      // ranked them beside default rows, and a column tile's neighbour in the
      // slice can change its rank
      const opo_tuning rows = tr;
      RCHK(col_pass(rows, &tr, &br));
      if (br < best_us) { best_us = br; best = tr; }
    }
    RCHK(set_tuning(E, g, &best));
    g.tuned_us = best_us;
    for (int k = 0; k < n; ++k)
      HIPCHK(hipMemcpyAsync(g.inst[k].state, (char *)bak + bytes * k, bytes, hipMemcpyDeviceToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
  }
  RCHK(autotune_joint(E, iters));
  RCHK(sync_insts(E));
  return reset_counters(E, 0);
}

}  // namespace

int opo_autotune(opo_solver *s, int iters) {
  if (!s || iters < 1) return fail(OPO_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(s->E.device));
  return engine_autotune(s->E, iters);
}

int opo_get_tuning(opo_solver *s, opo_tuning *out) {
  if (!s || !out) return fail(OPO_ERR_INVALID, "null");
  get_tuning(s->E, s->E.groups[0], out);
  return OPO_OK;
}

int opo_set_tuning(opo_solver *s, const opo_tuning *t) {
  if (!s || !t) return fail(OPO_ERR_INVALID, "null");
  return set_tuning(s->E, s->E.groups[0], t);
}

int opo_batch_autotune(opo_batch *b, int iters) {
  if (!b || iters < 1) return fail(OPO_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(b->E.device));
  return engine_autotune(b->E, iters);
}

int opo_batch_get_tuning(opo_batch *b, opo_tuning *out, int max) {
  if (!b || (!out && max > 0)) return fail(OPO_ERR_INVALID, "null");
  const int n = (int)b->E.groups.size();
  for (int k = 0; k < n && k < max; ++k) get_tuning(b->E, b->E.groups[k], out + k);
  return n;
}

int opo_batch_set_tuning(opo_batch *b, const opo_tuning *t) {
  if (!b || !t) return fail(OPO_ERR_INVALID, "null");
  int found = 0;  // every group of that grid size (a large batch of one size runs as two)
  for (auto &g : b->E.groups)
    if (g.N == t->N) { ++found; RCHK(set_tuning(b->E, g, t)); }
  if (!found) return fail(OPO_ERR_INVALID, "no group with that N");
  return OPO_OK;
}

int opo_batch_time_stages(opo_batch *b, int N, int iters, double *col_us, double *row_us) {
  if (!b || iters < 1) return fail(OPO_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(b->E.device));
  for (auto &g : b->E.groups)
    if (g.N == N) return engine_time_stages(b->E, g, iters, col_us, row_us);
  return fail(OPO_ERR_INVALID, "no group with that N");
}

int opo_batch_set_accel(opo_batch *b, int every, int start) {
  if (!b || every < 0 || start < 0) return fail(OPO_ERR_INVALID, "bad arguments");
  b->E.accel_every = every;
  b->E.accel_start = start;
  return OPO_OK;
}

int opo_batch_last_run_ms(opo_batch *b, double *ms) {
  if (!b || !ms) return fail(OPO_ERR_INVALID, "null");
  *ms = b->E.last_ms;
  return OPO_OK;
}

}  // extern "C"
