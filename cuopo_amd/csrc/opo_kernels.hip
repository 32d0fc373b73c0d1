// opo_kernels.hip -- gfx950 kernels of the SSFM round trip.
//
// A crystal slice of cuOPO (src/headers/cwes3.cuh:150-220) is
//   NL_half(z) ; IFFT+/N -> LinearOperator -> FFT- ; NL_half(z)
// and consecutive slices meet as NL_half(z_s) NL_half(z_{s+1}), both pointwise.
// With the four-step factorisation N = N1*N2 every slice becomes two passes
// over HBM:
//   col_stage  -- per column: inverse column FFT of the previous slice, the
//                 fused pointwise program (NL end/start, EOM, mirror, save,
//                 power, pump re-injection), forward column FFT of this slice;
//   row_stage  -- per row: row FFT, linear phase (GVD, walk-off, absorption,
//                 1/N), inverse row FFT.
// Each stage reads and writes every field once with coalesced 16-byte
// accesses; the sub-FFTs run in registers + LDS (team_fft).
#include <hip/hip_runtime.h>

#include "opo_device.h"
#include "opo_internal.h"

namespace opo {

template <typename V> __device__ __forceinline__ V ld_c(const double *p) {
  return mk<V>((decltype(V::x))p[0], (decltype(V::x))p[1]);
}

// four-step twiddle e^{S 2 pi i e/N} from two small tables
template <int S, typename V>
__device__ __forceinline__ V twiddle_n(unsigned e, const V *__restrict__ lo, const V *__restrict__ hi, int lobits) {
  const V w = cmul(hi[e >> lobits], lo[e & ((1u << lobits) - 1u)]);
  return tw_sign<S>(w);
}

template <typename T, int N1, int NF>
__global__ __launch_bounds__(256) void col_stage(const OpoInst *__restrict__ insts, ColArgs a) {
  using V = typename vec2<T>::type;
  constexpr int T1 = N1 / 8;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);

  const OpoInst *I = insts + blockIdx.y;
  if (*I->active == 0) return;  // block-uniform

  const int CT = 1 << a.log2ct;
  const int tid = threadIdx.x;
  const int c = tid & (CT - 1);
  const int t = tid >> a.log2ct;
  const int N2 = a.N2;
  const int n2 = blockIdx.x * CT + c;
  const long long N = (long long)N1 * N2;
  const unsigned nmask = (unsigned)(N - 1);
  V *st = reinterpret_cast<V *>(I->state);
  const V *tw1 = reinterpret_cast<const V *>(a.tw1);
  const V *twlo = reinterpret_cast<const V *>(a.twlo);
  const V *twhi = reinterpret_cast<const V *>(a.twhi);
  const int bufstride = CT * N1;

  V v[NF][8];
#pragma unroll
  for (int f = 0; f < NF; ++f)
    if ((a.load_mask >> f) & 1) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[f][m] = st[f * N + (long long)N2 * (t + m * T1) + n2];
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[f][m] = mk<V>(0, 0);
    }

  int ph = 0;
  // ---- inverse column transform (back to time domain) ----
#pragma unroll
  for (int f = 0; f < NF; ++f)
    if ((a.inv_mask >> f) & 1) {
#pragma unroll
      for (int m = 0; m < 8; ++m)
        v[f][m] = cmul(v[f][m], twiddle_n<-1>((unsigned)(n2 * (t + m * T1)) & nmask, twlo, twhi, a.lobits));
      team_fft<N1, -1>(v[f], lds, ph, t, c, CT, bufstride, tw1);
    }

  // ---- pointwise program on time samples n = N2*(t + m*T1) + n2 ----
  const unsigned ops = a.ops;
  auto nl_half = [&](int s) {
    const double *q = I->nl + 16 * s;
    if (a.three_eqs) {
      // Closed form of cwes3.cuh:153-172: every stage evaluates dAdz on the
      // unchanged fields, so A += (h/6)(k1+2k2+2k3+k4) = A + C*P(A) with
      // C = i kappa (h/6)(E(z)+4E(z+dz/4)+E(z+dz/2)) precomputed in FP64.
      const V Cp = ld_c<V>(q), Cs = ld_c<V>(q + 2), Ci = ld_c<V>(q + 4);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const V ap = v[0][m], as = v[1][m], ai = v[NF - 1][m];
        v[0][m] = cadd(ap, cmul(Cp, cmul(as, ai)));
        v[1][m] = cadd(as, cmul(Cs, cmul(ap, cconj(ai))));
        if constexpr (NF == 3) v[2][m] = cadd(ai, cmul(Ci, cmul(ap, cconj(as))));
      }
    } else {
      // cwes2.cuh:147-166: stages 2-4 evaluate dAdz on A + s*k (s = 0.5, 0.5, 1).
      using R = decltype(V::x);
      const V ep0 = ld_c<V>(q), ep1 = ld_c<V>(q + 2), ep2 = ld_c<V>(q + 4);
      const V em0 = ld_c<V>(q + 6), em1 = ld_c<V>(q + 8), em2 = ld_c<V>(q + 10);
      const V ikp = mk<V>((R)0, (R)I->kp), iks = mk<V>((R)0, (R)I->ks);
      const R h = (R)I->h;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const V ap = v[0][m], as = v[1][m];
        V k1p = cmul(cmul(cmul(ikp, as), as), ep0), k1s = cmul(cmul(cmul(iks, ap), cconj(as)), em0);
        V bp = cadd(ap, cscale(k1p, (R)0.5)), bs = cadd(as, cscale(k1s, (R)0.5));
        V k2p = cmul(cmul(cmul(ikp, bs), bs), ep1), k2s = cmul(cmul(cmul(iks, bp), cconj(bs)), em1);
        bp = cadd(ap, cscale(k2p, (R)0.5)); bs = cadd(as, cscale(k2s, (R)0.5));
        V k3p = cmul(cmul(cmul(ikp, bs), bs), ep1), k3s = cmul(cmul(cmul(iks, bp), cconj(bs)), em1);
        bp = cadd(ap, k3p); bs = cadd(as, k3s);
        V k4p = cmul(cmul(cmul(ikp, bs), bs), ep2), k4s = cmul(cmul(cmul(iks, bp), cconj(bs)), em2);
        V sp = cadd(cadd(cadd(k1p, cscale(k2p, (R)2)), cscale(k3p, (R)2)), k4p);
        V ss = cadd(cadd(cadd(k1s, cscale(k2s, (R)2)), cscale(k3s, (R)2)), k4s);
        sp = cscale(sp, h); ss = cscale(ss, h);
        v[0][m] = cadd(ap, mk<V>(sp.x / (R)6, sp.y / (R)6));
        v[1][m] = cadd(as, mk<V>(ss.x / (R)6, ss.y / (R)6));
      }
    }
  };

  if (ops & OP_NL_END) nl_half(a.s_end);
  if ((ops & OP_EOM) && I->eom) {
    const V *eom = reinterpret_cast<const V *>(I->eom);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const V e = eom[(long long)N2 * (t + m * T1) + n2];
      v[1][m] = cmul(v[1][m], e);
      if constexpr (NF == 3) v[2][m] = cmul(v[2][m], e);
    }
  }
  if (ops & OP_MIRROR) {
    using R = decltype(V::x);
    const V e = mk<V>((R)I->mir[0], (R)I->mir[1]);
    const R rs = (R)I->sqrt_rs, ri = (R)I->sqrt_ri;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      v[1][m] = cmul(cscale(v[1][m], rs), e);
      if constexpr (NF == 3) v[2][m] = cmul(cscale(v[2][m], ri), e);
    }
  }
  if (ops & OP_SAVE) {
    const int slot = *I->rt - I->save_first;
    if (slot >= 0 && slot < I->save_slots) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        V *sv = reinterpret_cast<V *>(I->save[f]);
        if (sv) {
#pragma unroll
          for (int m = 0; m < 8; ++m) sv[(long long)slot * N + (long long)N2 * (t + m * T1) + n2] = v[f][m];
        }
      }
    }
  }
  if ((ops & OP_SNAP) && a.snap_every > 0 && ((*I->rt + 1) % a.snap_every) == 0) {
    V *sn = reinterpret_cast<V *>(I->snap);
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int m = 0; m < 8; ++m) sn[f * N + (long long)N2 * (t + m * T1) + n2] = v[f][m];
  }
  if (ops & OP_POWER) {
    double acc = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m) acc += (double)v[1][m].x * (double)v[1][m].x + (double)v[1][m].y * (double)v[1][m].y;
    // deterministic block reduction: wave shuffle tree, then waves in order
    // blocks of 32 threads occupy half a wave: start the tree inside the block
    for (int off = (blockDim.x < 64 ? (int)blockDim.x : 64) >> 1; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    double *red = reinterpret_cast<double *>(lds + 2 * bufstride);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
      double s = 0;
      const int nw = (blockDim.x + 63) >> 6;
      for (int w = 0; w < nw; ++w) s += red[w];
      I->partials[blockIdx.x] = s;
    }
  }
  if (ops & OP_PUMP) {
    const V *pin = reinterpret_cast<const V *>(I->pump);
    const int rp = *I->rt + a.pump_off;
    long long off = 0;  // functions.cuh:196-210 (NS); CW copies Ap_in (cuOPO.cu:480)
    if (I->regime == 1) {
      if (rp == 0) off = 0;
      else if (rp > 0 && rp < I->nrt - 1) off = (long long)rp * N;
      else off = N * (long long)I->nrt - 1 - N;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) v[0][m] = pin[off + (long long)N2 * (t + m * T1) + n2];
  }
  if (ops & OP_NL_START) nl_half(a.s_start);

  // ---- forward column transform (start of the next linear step) ----
#pragma unroll
  for (int f = 0; f < NF; ++f)
    if ((a.fwd_mask >> f) & 1) {
      team_fft<N1, +1>(v[f], lds, ph, t, c, CT, bufstride, tw1);
#pragma unroll
      for (int m = 0; m < 8; ++m)
        v[f][m] = cmul(v[f][m], twiddle_n<+1>((unsigned)(n2 * (t + m * T1)) & nmask, twlo, twhi, a.lobits));
    }

#pragma unroll
  for (int f = 0; f < NF; ++f)
    if ((a.store_mask >> f) & 1) {
#pragma unroll
      for (int m = 0; m < 8; ++m) st[f * N + (long long)N2 * (t + m * T1) + n2] = v[f][m];
    }
}

template <typename T, int N2, int NF>
__global__ __launch_bounds__(256) void row_stage(const OpoInst *__restrict__ insts, RowArgs a) {
  using V = typename vec2<T>::type;
  constexpr int T2 = N2 / 8;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);

  const OpoInst *I = insts + blockIdx.y;
  if (*I->active == 0) return;

  const int RT = 1 << a.log2rt;
  const int tiles_per_field = a.N1 >> a.log2rt;
  const int fi = blockIdx.x / tiles_per_field;
  const int f = fi == 0 ? a.f0 : (fi == 1 ? a.f1 : a.f2);
  const int tid = threadIdx.x;
  const int t = tid & (T2 - 1);
  const int r = tid / T2;
  const int k1 = (blockIdx.x - fi * tiles_per_field) * RT + r;
  const long long N = (long long)a.N1 * N2;
  V *row = reinterpret_cast<V *>(I->state) + f * N + (long long)N2 * k1;
  const V *tab = a.use_gdd ? reinterpret_cast<const V *>(I->gdd) + (f - 1) * N
                          : reinterpret_cast<const V *>(I->lin) + f * N;
  tab += (long long)N2 * k1;
  const V *tw2 = reinterpret_cast<const V *>(a.tw2);

  V v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = row[t + m * T2];
  int ph = 0;
  team_fft<N2, +1>(v, lds, ph, t, r * N2, 1, RT * N2, tw2);
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = cmul(v[m], tab[t + m * T2]);
  team_fft<N2, -1>(v, lds, ph, t, r * N2, 1, RT * N2, tw2);
#pragma unroll
  for (int m = 0; m < 8; ++m) row[t + m * T2] = v[m];
}

// One thread per instance: reduce the column-tile partials in order, record
// the power, advance the round-trip counter and, every check_every round
// trips, test |P - P_prev| <= tol * P.
__global__ void tick_kernel(const OpoInst *__restrict__ insts, TickArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.count) return;
  const OpoInst *I = insts + b;
  if (*I->active == 0) return;
  double P = 0;
  for (int j = 0; j < a.tiles; ++j) P += I->partials[j];
  const int rt = *I->rt;
  if (I->power_hist && rt >= 0 && rt < I->hist_len) I->power_hist[rt] = P;
  const int done = rt + 1;
  *I->rt = done;
  if (!(P == P) || P > 1.7e308) {  // NaN or Inf
    *I->status = -5;
    *I->active = 0;
    return;
  }
  if (a.check_every > 0 && done % a.check_every == 0) {
    const double prev = *I->p_ref;
    if (prev > 0 && P > 0 && fabs(P - prev) <= a.tol * P) {
      *I->conv_rt = done;
      *I->active = 0;
    }
    *I->p_ref = P;
  }
}

// ------------------------------------------------------------------ launchers

template <typename T, int N1, int NF>
static int launch_col_t(int ct, int count, const OpoInst *insts, const ColArgs &a, hipStream_t s) {
  using V = typename vec2<T>::type;
  const int threads = ct * (N1 / 8);
  const size_t lds = 2 * (size_t)ct * N1 * sizeof(V) + 64 * sizeof(double);
  dim3 grid(a.N2 / ct, count);
  static bool big_lds = false;  // gfx950 has 160 KiB of LDS; opt in above 64 KiB once
  if (lds > 65536 && !big_lds) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&col_stage<T, N1, NF>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    big_lds = true;
  }
  hipLaunchKernelGGL((col_stage<T, N1, NF>), grid, dim3(threads), lds, s, insts, a);
  return (int)hipGetLastError();
}

template <typename T, int N2, int NF>
static int launch_row_t(int rt, int count, int tiles_y, const OpoInst *insts, const RowArgs &a, hipStream_t s) {
  using V = typename vec2<T>::type;
  const int threads = rt * (N2 / 8);
  const size_t lds = 2 * (size_t)rt * N2 * sizeof(V);
  dim3 grid((a.N1 / rt) * a.nf, count);
  (void)tiles_y;
  hipLaunchKernelGGL((row_stage<T, N2, NF>), grid, dim3(threads), lds, s, insts, a);
  return (int)hipGetLastError();
}

#define OPO_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512)

template <typename T, int NF>
static int dispatch_col(int N1, int ct, int count, const OpoInst *insts, const ColArgs &a, hipStream_t s) {
  switch (N1) {
#define CASE(M) \
  case M: return launch_col_t<T, M, NF>(ct, count, insts, a, s);
    OPO_SIZES(CASE)
#undef CASE
  }
  return (int)hipErrorInvalidValue;
}

template <typename T, int NF>
static int dispatch_row(int N2, int rt, int count, int ty, const OpoInst *insts, const RowArgs &a, hipStream_t s) {
  switch (N2) {
#define CASE(M) \
  case M: return launch_row_t<T, M, NF>(rt, count, ty, insts, a, s);
    OPO_SIZES(CASE)
#undef CASE
  }
  return (int)hipErrorInvalidValue;
}

int launch_col(int fp32, int N1, int NF, int ct, int count, const OpoInst *insts, const ColArgs &a, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (fp32) return NF == 3 ? dispatch_col<float, 3>(N1, ct, count, insts, a, s) : dispatch_col<float, 2>(N1, ct, count, insts, a, s);
  return NF == 3 ? dispatch_col<double, 3>(N1, ct, count, insts, a, s) : dispatch_col<double, 2>(N1, ct, count, insts, a, s);
}

int launch_row(int fp32, int N2, int NF, int rt, int count, int ty, const OpoInst *insts, const RowArgs &a,
               void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (fp32) return NF == 3 ? dispatch_row<float, 3>(N2, rt, count, ty, insts, a, s) : dispatch_row<float, 2>(N2, rt, count, ty, insts, a, s);
  return NF == 3 ? dispatch_row<double, 3>(N2, rt, count, ty, insts, a, s) : dispatch_row<double, 2>(N2, rt, count, ty, insts, a, s);
}

int launch_tick(const OpoInst *insts, const TickArgs &a, void *stream) {
  dim3 grid((a.count + 63) / 64);
  hipLaunchKernelGGL(tick_kernel, grid, dim3(64), 0, (hipStream_t)stream, insts, a);
  return (int)hipGetLastError();
}

}  // namespace opo
