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
#include <hip/hip_ext.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "opo_device.h"
#include "opo_internal.h"

namespace opo {

template <typename V, typename P> __device__ __forceinline__ V ld_c(P p) {
  return mk<V>((decltype(V::x))p[0], (decltype(V::x))p[1]);
}

// four-step twiddle e^{S 2 pi i e/N} from two small tables
template <int S, typename V>
__device__ __forceinline__ V twiddle_n(unsigned e, const V *__restrict__ lo, const V *__restrict__ hi, int lobits) {
  const V w = cmul(hi[e >> lobits], lo[e & ((1u << lobits) - 1u)]);
  return tw_sign<S>(w);
}

// Field f of three values.  The operands are taken by value so the choice is a
// register select, never a load through a selected address (which would keep
// the field arrays out of registers).
template <typename V> __device__ __forceinline__ V pick(int f, V a, V b, V c) {
  const V r = f == 1 ? b : c;
  return f == 0 ? a : r;
}

// Coefficients of one nonlinear half step (slice s), read once per block at
// kernel start so their latency overlaps the field loads.
template <int NF, typename V> struct NlCoef {
  V c[NF == 3 ? 3 : 6];
  decltype(V::x) kp, ks, h;
};

template <int NF, typename V>
__device__ __forceinline__ NlCoef<NF, V> load_nl(const double *nl, double kp, double ks, double h, int s) {
  NlCoef<NF, V> r;
  const auto q = gp_scalar(nl) + 16 * s;
#pragma unroll
  for (int j = 0; j < (NF == 3 ? 3 : 6); ++j) r.c[j] = ld_c<V>(q + 2 * j);
  using R = decltype(V::x);
  r.kp = (R)kp; r.ks = (R)ks; r.h = (R)h;
  return r;
}

// Pointwise nonlinear half step on all fields of K samples.
template <int NF, int K, typename V>
__device__ __forceinline__ void nl_half_pt(V *xp, V *xs, V *xi, const NlCoef<NF, V> &q) {
  if constexpr (NF == 3) {
    // Closed form of cwes3.cuh:153-172: every stage evaluates dAdz on the
    // unchanged fields, so A += (h/6)(k1+2k2+2k3+k4) = A + C*P(A) with
    // C = i kappa (h/6)(E(z)+4E(z+dz/4)+E(z+dz/2)) precomputed in FP64.
    const V Cp = q.c[0], Cs = q.c[1], Ci = q.c[2];
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const V ap = xp[m], as = xs[m], ai = xi[m];
      xp[m] = cadd(ap, cmul(Cp, cmul(as, ai)));
      xs[m] = cadd(as, cmul(Cs, cmul(ap, cconj(ai))));
      xi[m] = cadd(ai, cmul(Ci, cmul(ap, cconj(as))));
    }
  } else {
    // cwes2.cuh:147-166: stages 2-4 evaluate dAdz on A + s*k (s = 0.5, 0.5, 1).
    using R = decltype(V::x);
    const V ep0 = q.c[0], ep1 = q.c[1], ep2 = q.c[2];
    const V em0 = q.c[3], em1 = q.c[4], em2 = q.c[5];
    const V ikp = mk<V>((R)0, q.kp), iks = mk<V>((R)0, q.ks);
    const R h = q.h;
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const V ap = xp[m], as = xs[m];
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
      xp[m] = cadd(ap, mk<V>(sp.x / (R)6, sp.y / (R)6));
      xs[m] = cadd(as, mk<V>(ss.x / (R)6, ss.y / (R)6));
    }
  }
}

// Column stage.  The block holds CT columns of every field; the threads are
// split into NF groups of FT = CT*N1/E (E = 4 or 8 samples per thread), and
// group f runs field f's column FFTs, so the fields' transforms proceed on
// different SIMDs at once.  The time-domain samples of all fields meet in LDS
// for the pointwise program; each group keeps its own field.  MID compiles
// the slice-interior variant only (inverse -> NL end -> NL start -> forward).
template <typename T, int N1, int NF, bool MID, bool DIRECT, int E>
__global__ __launch_bounds__(E == 4 ? 1024 : 512) void col_stage(const OpoInst *__restrict__ insts, ColArgs a) {
  static_assert(MID || !DIRECT, "direct arguments serve the slice-interior kernel");
  static_assert(E == 4 || E == 8, "4 or 8 samples per thread");
  using V = typename vec2<T>::type;
  constexpr int T1 = N1 / E;
  extern __shared__ unsigned char smem_raw[];
  V *ltw = reinterpret_cast<V *>(smem_raw);  // N1-point root table
  V *lds = ltw + N1;                         // 2 buffers (1 for wave) x NF fields x CT*N1

  const OpoInst *I = insts + blockIdx.y;
  if constexpr (!DIRECT) {
    if (I->active_v == 0) return;  // block-uniform
  }
  // XCD packing: work-groups are dispatched round-robin over the 8 XCDs, so
  // keeping only every 8th block puts the whole stage on one XCD (and its L2)
  if (a.xcd_pack && (blockIdx.x & 7)) return;
  int bx = a.xcd_pack ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  // tile swizzle: block b runs on XCD b % 8, so tile (b % 8)*(tiles/8) + b/8
  // gives each XCD a contiguous range of columns and the tiles that share a
  // cache line (tiles narrower than 128 bytes) the same L2
  if (a.swz) bx = (bx & 7) * (int)(gridDim.x >> 3) + (bx >> 3);

  const int CT = 1 << a.log2ct;
  const int FT = CT * T1;
  // slice interior with one wavefront (or less) per field: single-buffer
  // wave-private column exchanges (the launcher sizes the LDS to match)
  const bool wave = MID && team_in_wave(FT);
  const int tid = threadIdx.x;
  const int f = tid / FT;  // this thread's field
  const int lt = tid - f * FT;
  const int c = lt & (CT - 1);
  const int t = lt >> a.log2ct;
  const int N2 = a.N2;
  const int n2 = bx * CT + c;
  const long long N = (long long)N1 * N2;
  const unsigned nmask = (unsigned)(N - 1);
  const auto st = gp(reinterpret_cast<V *>(DIRECT ? a.d_state : I->state));
  const int fstride = CT * N1;
  const int bufstride = NF * fstride;
  const int inv_mask = MID ? (1 << NF) - 1 : a.inv_mask;
  const int fwd_mask = MID ? (1 << NF) - 1 : a.fwd_mask;
  const int load_mask = MID ? (1 << NF) - 1 : a.load_mask;
  const int store_mask = MID ? (1 << NF) - 1 : a.store_mask;
  const bool any_fft = (inv_mask | fwd_mask) != 0;

  const unsigned ops = MID ? (OP_NL_END | OP_NL_START) : a.ops;
  // Issue order of the loads: the field samples first, then the four-step
  // twiddle gathers, then the column root table and the NL coefficients.  A
  // wait for any of the later ones (the table's LDS copy waits for its load;
  // a scalar-load wait covers every scalar load in flight) then no longer
  // holds back the field loads, so the kernel pays one memory latency.
  V v[E];
  const bool ld = (load_mask >> f) & 1;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = ld ? V(st[f * N + (long long)N2 * (t + m * T1) + n2]) : mk<V>(0, 0);
  // four-step twiddles W_m = e^{2 pi i n2 k1 / N} for this thread's
  // k1 = t + m*T1, from two-level root tables: X(e) = hi[e >> lobits] lo[e & mask]
  const auto twlo = gp(reinterpret_cast<const V *>(a.twlo));
  const auto twhi = gp(reinterpret_cast<const V *>(a.twhi));
  const unsigned lm = (1u << a.lobits) - 1u;
  constexpr bool F64 = std::is_same<T, double>::value;
  // gathered roots: FP64 x0 and the base b1 (hi, lo each); FP32 one pair per m
  V gh[F64 ? 1 : E], gl[F64 ? 1 : E];
  V x0h{}, x0l{}, b1h{}, b1l{};
  if (any_fft) {
    if constexpr (F64) {
      const unsigned e0 = (unsigned)(n2 * t) & nmask, e1 = (unsigned)(n2 * T1) & nmask;
      x0h = twhi[e0 >> a.lobits]; x0l = twlo[e0 & lm];
      b1h = twhi[e1 >> a.lobits]; b1l = twlo[e1 & lm];
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const unsigned e = (unsigned)(n2 * (t + m * T1)) & nmask;
        gh[m] = twhi[e >> a.lobits]; gl[m] = twlo[e & lm];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  NlCoef<NF, V> nle{}, nls{};
  if constexpr (DIRECT) {
    nle = load_nl<NF, V>(a.d_nl, a.d_kp, a.d_ks, a.d_h, a.s_end);
    nls = load_nl<NF, V>(a.d_nl, a.d_kp, a.d_ks, a.d_h, a.s_start);
  } else {
    if (ops & OP_NL_END) nle = load_nl<NF, V>(I->nl, I->kp, I->ks, I->h, a.s_end);
    if (ops & OP_NL_START) nls = load_nl<NF, V>(I->nl, I->kp, I->ks, I->h, a.s_start);
  }
  V W[E];
  V x0r{}, b1r{};  // FP64: the four-step roots' base, kept across the pointwise program instead of W
  if (any_fft) {
    // column root table into LDS (the block has NF*FT threads; blockDim
    // would cost a scalar load whose wait covers the NL loads)
    const auto tw1 = gp(reinterpret_cast<const V *>(a.tw1));
    for (int e = tid; e < N1; e += NF * FT) ltw[e] = tw1[e];
    if (wave) __syncthreads();  // the exchanges no longer order the table's writes
    if constexpr (F64) {
      // W_m = X(n2 t) X(n2 T1)^m, the powers from b, b^2, b^4 by the bits of m
      // (at most three complex products per twiddle; the base's rounding
      // grows at most 7x, ~1e-15 in FP64): 4 gathers per thread instead of 2E
      const V x0 = cmul(x0h, x0l);
      const V b1 = cmul(b1h, b1l);
      x0r = x0; b1r = b1;
      const V b2 = cmul(b1, b1), b4 = cmul(b2, b2);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        V w = x0;
        if (m & 4) w = cmul(w, b4);
        if (m & 2) w = cmul(w, b2);
        if (m & 1) w = cmul(w, b1);
        W[m] = w;
      }
    } else {
      // FP32: the same growth would cost accuracy, so each twiddle is looked up
#pragma unroll
      for (int m = 0; m < E; ++m) W[m] = cmul(gh[m], gl[m]);
    }
  }

  int ph = 0;
  V *mine0 = lds + f * fstride;
  // ---- inverse column transform of this group's field ----
  const bool inv = (inv_mask >> f) & 1;
  if (inv_mask) {
    if (inv) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(v[m], cconj(W[m]));
    }
    team_fft<N1, -1, false, false, E>(v, inv, mine0, ph, t, c, CT, bufstride, ltw, wave);
  }
  // ---- all fields meet in the next LDS buffer ----
  V *xb = lds + (wave ? 0 : (ph & 1) * bufstride);
#pragma unroll
  for (int m = 0; m < E; ++m) xb[f * fstride + c + (t + m * T1) * CT] = v[m];
  __syncthreads();
  ++ph;
  if constexpr (MID) {
    // Slice interior: the thread of field f updates every field at its
    // samples m = f (mod NF), in place (each sample belongs to one thread),
    // then every thread reads its own field back.
#pragma unroll
    for (int m = 0; m < E; ++m) {
      if (m % NF == f) {
        const int o = c + (t + m * T1) * CT;
        V xp[1] = {xb[o]}, xs[1] = {xb[fstride + o]};
        V xi[1] = {NF == 3 ? xb[2 * fstride + o] : mk<V>(0, 0)};
        nl_half_pt<NF, 1>(xp, xs, xi, nle);
        nl_half_pt<NF, 1>(xp, xs, xi, nls);
        xb[o] = xp[0];
        xb[fstride + o] = xs[0];
        if constexpr (NF == 3) xb[2 * fstride + o] = xi[0];
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = xb[f * fstride + c + (t + m * T1) * CT];
  } else {
  // ---- pointwise program on time samples n = N2*(t + m*T1) + n2, in
  //      chunks of 4 samples to bound the live registers ----
  int slot = -1, rp = 0;
  bool snap = false;
  if constexpr (!MID) {
    if (ops & OP_SAVE) slot = *I->rt - I->save_first;
    if (ops & OP_PUMP) rp = *I->rt + a.pump_off;
    snap = (ops & OP_SNAP) && a.snap_every > 0 && ((*I->rt + 1) % a.snap_every) == 0;
  }
  double pacc = 0;
  // chunk of 4 samples m = hb .. hb+3 (hb a compile-time constant, so every
  // register array below is indexed statically)
  auto chunk = [&](auto hbc) {
    constexpr int hb = decltype(hbc)::value;
    V xp[4], xs[4], xi[4];  // pump, signal, idler
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int o = c + (t + (hb + q) * T1) * CT;
      xp[q] = xb[o];
      xs[q] = xb[fstride + o];
      xi[q] = NF == 3 ? xb[2 * fstride + o] : mk<V>(0, 0);
    }
    if (ops & OP_NL_END) nl_half_pt<NF, 4>(xp, xs, xi, nle);
    if constexpr (!MID) {
      if ((ops & OP_EOM) && I->eom) {
        const auto eom = gp(reinterpret_cast<const V *>(I->eom));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const V e = eom[(long long)N2 * (t + (hb + q) * T1) + n2];
          xs[q] = cmul(xs[q], e);
          if constexpr (NF == 3) xi[q] = cmul(xi[q], e);
        }
      }
      if (ops & OP_MIRROR) {
        using R = decltype(V::x);
        const V e = mk<V>((R)I->mir[0], (R)I->mir[1]);
        const R rs = (R)I->sqrt_rs, ri = (R)I->sqrt_ri;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          xs[q] = cmul(cscale(xs[q], rs), e);
          if constexpr (NF == 3) xi[q] = cmul(cscale(xi[q], ri), e);
        }
      }
      if ((ops & OP_SAVE) || snap || (ops & OP_POWER)) {
        V w4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) w4[q] = pick(f, xp[q], xs[q], xi[q]);  // this group's field
        const auto sv = gp(reinterpret_cast<V *>(I->save[f]));
        if ((ops & OP_SAVE) && slot >= 0 && slot < I->save_slots && sv) {
#pragma unroll
          for (int q = 0; q < 4; ++q) sv[(long long)slot * N + (long long)N2 * (t + (hb + q) * T1) + n2] = w4[q];
        }
        if (snap) {
          const auto sn = gp(reinterpret_cast<V *>(I->snap));
#pragma unroll
          for (int q = 0; q < 4; ++q) sn[f * N + (long long)N2 * (t + (hb + q) * T1) + n2] = w4[q];
        }
        if ((ops & OP_POWER) && f == 1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) pacc += (double)w4[q].x * (double)w4[q].x + (double)w4[q].y * (double)w4[q].y;
        }
      }
      if (ops & OP_PUMP) {
        const auto pin = gp(reinterpret_cast<const V *>(I->pump));
        long long off = 0;  // functions.cuh:196-210 (NS); CW copies Ap_in (cuOPO.cu:480)
        if (I->regime == 1) {
          if (rp == 0) off = 0;
          else if (rp > 0 && rp < I->nrt - 1) off = (long long)rp * N;
          else off = N * (long long)I->nrt - 1 - N;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) xp[q] = pin[off + (long long)N2 * (t + (hb + q) * T1) + n2];
      }
    }
    if (ops & OP_NL_START) nl_half_pt<NF, 4>(xp, xs, xi, nls);
#pragma unroll
    for (int q = 0; q < 4; ++q) v[hb + q] = pick(f, xp[q], xs[q], xi[q]);
  };
  chunk(std::integral_constant<int, 0>{});
  if constexpr (E == 8) chunk(std::integral_constant<int, 4>{});
  {
    if (ops & OP_POWER) {
      // deterministic block reduction: wave shuffle tree, then waves in order;
      // blocks of 32 threads occupy half a wave, so start the tree inside the block
      double acc = pacc;
      for (int off = (blockDim.x < 64 ? (int)blockDim.x : 64) >> 1; off > 0; off >>= 1)
        acc += __shfl_down(acc, off, 64);
      double *red = reinterpret_cast<double *>(lds + 2 * bufstride);
      if ((tid & 63) == 0) red[tid >> 6] = acc;
      __syncthreads();
      if (tid == 0) {
        double sum = 0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) sum += red[w];
        I->partials[bx] = sum;
      }
    }
  }
  }  // boundary / general path

  // ---- forward column transform of this group's field ----
  const bool fw = (fwd_mask >> f) & 1;
  if (fwd_mask) {
    team_fft<N1, +1, false, false, E>(v, fw, mine0, ph, t, c, CT, bufstride, ltw, wave);
    if (fw) {
      if constexpr (F64) {
        // the same W_m again, from the two kept roots: holding 2 complex
        // values instead of E across the middle of the kernel takes 28 VGPRs
        // off the E = 8 slice-interior build (154 -> 126, 3 -> 4 waves per
        // SIMD).  The empty asm hides the recomputation from common
        // subexpression elimination, which would otherwise keep W alive.
        asm volatile("" : "+v"(x0r.x), "+v"(x0r.y), "+v"(b1r.x), "+v"(b1r.y));
        const V b2 = cmul(b1r, b1r), b4 = cmul(b2, b2);
#pragma unroll
        for (int m = 0; m < E; ++m) {
          V w = x0r;
          if (m & 4) w = cmul(w, b4);
          if (m & 2) w = cmul(w, b2);
          if (m & 1) w = cmul(w, b1r);
          W[m] = w;
        }
      }
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(v[m], W[m]);
    }
  }
  if ((store_mask >> f) & 1) {
#pragma unroll
    for (int m = 0; m < E; ++m) st[f * N + (long long)N2 * (t + m * T1) + n2] = v[m];
  }
}

// threads per row-stage block: a 4096-point row at 4 samples per thread needs 1024
constexpr int row_max_threads(int N2, int E) { return (E == 4 && N2 >= 4096) ? 1024 : 512; }

template <typename T, int N2, int NF, bool DIRECT, int E>
__global__ __launch_bounds__(row_max_threads(N2, E)) void row_stage(const OpoInst *__restrict__ insts, RowArgs a) {
  static_assert(E == 4 || E == 8, "4 or 8 samples per thread");
  using V = typename vec2<T>::type;
  constexpr int T2 = N2 / E;
  constexpr int NP2 = N2 + N2 / 8;  // padded row in LDS
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);  // 2 buffers (1 when T2 <= 64) x RT padded rows

  const OpoInst *I = insts + blockIdx.y;
  if constexpr (!DIRECT) {
    if (I->active_v == 0) return;
  }
  if (a.xcd_pack && (blockIdx.x & 7)) return;  // see col_stage
  const int bx = a.xcd_pack ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;

  const int RT = 1 << a.log2rt;
  const int tiles_per_field = a.N1 >> a.log2rt;
  const int fi = bx / tiles_per_field;
  const int f = fi == 0 ? a.f0 : (fi == 1 ? a.f1 : a.f2);
  const int tid = threadIdx.x;
  const int t = tid & (T2 - 1);
  const int r = tid / T2;
  const int k1 = (bx - fi * tiles_per_field) * RT + r;
  const long long N = (long long)a.N1 * N2;
  const auto row = gp(reinterpret_cast<V *>(DIRECT ? a.d_state : I->state)) + f * N + (long long)N2 * k1;

  constexpr int NTW = tw_count(N2, E, tw_pow<V>());
  V v[E], ph_tab[E], treg[NTW > 0 ? NTW : 1];
  const auto fac = gp(reinterpret_cast<const V *>(DIRECT ? a.d_lfac : I->lfac));
  const bool use_fac = !a.use_gdd && fac;
  auto tab = a.use_gdd ? gp(reinterpret_cast<const V *>(DIRECT ? a.d_gdd : I->gdd)) + (f - 1) * N
                     : gp(reinterpret_cast<const V *>(DIRECT ? a.d_lin : I->lin)) + f * N;
  // phase loads first (the factorised products only wait for these), then
  // the row, then the twiddles
  V xr = mk<V>(0, 0);
  if (use_fac) {
    // factorised phase Z[E k1 + m] S[t + m T2] X(k1 t), X = Xhi Xlo
    const int L = a.xl_bits;
    const long long FB = (long long)E * a.N1 + N2 + ((N / E) >> L) + (1LL << L);
    const auto F = fac + f * FB;
    const auto S = F + E * a.N1;
    const auto XH = S + N2;
    const auto XL = XH + ((N / E) >> L);
    const unsigned e = (unsigned)k1 * (unsigned)t;
    xr = cmul(V(XH[e >> L]), V(XL[e & ((1u << L) - 1u)]));
#pragma unroll
    for (int m = 0; m < E; ++m) ph_tab[m] = cmul(V(F[E * k1 + m]), V(S[t + m * T2]));
  } else {
    tab = tab + (long long)N2 * k1;
#pragma unroll
    for (int m = 0; m < E; ++m) ph_tab[m] = tab[t + m * T2];
  }
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = row[t + m * T2];
  load_twiddles<N2, E>(treg, t, gp(reinterpret_cast<const V *>(a.tw2)));
  if (use_fac) {
#pragma unroll
    for (int m = 0; m < E; ++m) ph_tab[m] = cmul(ph_tab[m], xr);
  }
  int ph = 0;
  const int bufstride = RT * NP2;
  constexpr bool wave = team_in_wave(T2);  // one row per wavefront or less: see fft_pass
  team_fft<N2, +1, true, true, E>(v, true, lds, ph, t, r * NP2, 1, bufstride, treg, wave);
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cmul(v[m], ph_tab[m]);
  team_fft<N2, -1, true, true, E>(v, true, lds, ph, t, r * NP2, 1, bufstride, treg, wave);
#pragma unroll
  for (int m = 0; m < E; ++m) row[t + m * T2] = v[m];
}

// One wavefront per instance: reduce the column-tile partials in a fixed
// order, record the power, advance the round-trip counter and, every
// check_every round trips, test |P - P_prev| <= tol * P.
__global__ __launch_bounds__(64) void tick_kernel(OpoInst *__restrict__ insts, TickArgs a) {
  const int b = blockIdx.x;  // one wavefront per instance
  if (b >= a.count) return;
  OpoInst *I = insts + b;
  if (I->active_v == 0) return;  // uniform over the wave
  // the partials in a fixed order: lane l sums tiles l, l+64, ... in order,
  // then a fixed shuffle tree over the lanes (deterministic, loads in parallel)
  const int lane = threadIdx.x;
  double P = 0;
  for (int j = lane; j < a.tiles; j += 64) P += I->partials[j];
  for (int off = 32; off > 0; off >>= 1) P += __shfl_down(P, off, 64);
  if (lane != 0) return;
  const int rt = *I->rt;
  if (I->power_hist && rt >= 0 && rt < I->hist_len) I->power_hist[rt] = P;
  const int done = rt + 1;
  *I->rt = done;
  if (!(P == P) || P > 1.7e308) {  // NaN or Inf
    *I->status = -5;
    *I->active = 0;
    I->active_v = 0;
    return;
  }
  if (a.check_every > 0 && done % a.check_every == 0) {
    const double prev = *I->p_ref;
    if (prev > 0 && P > 0 && fabs(P - prev) <= a.tol * P) {
      *I->conv_rt = done;  // steady state
      *I->active = 0;
      I->active_v = 0;
    } else if (prev > 0 && *I->p_first > 0) {
      // Below threshold: in the small-signal regime P is a sum of decaying or
      // growing exponentials, so log P is convex and the window ratio r never
      // decreases; once r < 1 has stopped changing the slowest mode dominates
      // and decays.  Require also a collapse below off_ratio x P(first check).
      const double r = P / prev, rp = *I->r_prev;
      if (r < 1.0 && rp > 0 && fabs(r - rp) <= 1e-3 * r && P < a.off_ratio * *I->p_first) {
        *I->conv_rt = done;
        *I->status = 1;
        *I->active = 0;
        I->active_v = 0;
      }
    }
    if (prev > 0) *I->r_prev = P / prev;
    if (*I->p_first == 0) *I->p_first = P;
    *I->p_ref = P;
  }
}

// ------------------------------------------------------------------ launchers

// gfx950 has 160 KiB of LDS per CU; a kernel opts in above 64 KiB once per
// device (the attribute belongs to the device's copy of the function).
constexpr int kMaxDevices = 64;
static int opt_in_big_lds(const void *kern, std::once_flag (&flags)[kMaxDevices]) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  // a failed opt-in leaves the flag unset (call_once does not complete when
  // its callable throws), so the next launch on this device tries again
  try {
    std::call_once(flags[dev], [&] {
      const hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      if (e != hipSuccess) throw e;
    });
  } catch (hipError_t e) {
    return (int)e;
  }
  return 0;
}

template <int NF> static bool col_is_mid(const ColArgs &a) {
  return a.ops == (OP_NL_END | OP_NL_START) && a.inv_mask == (1 << NF) - 1 && a.fwd_mask == (1 << NF) - 1 &&
         a.load_mask == (1 << NF) - 1 && a.store_mask == (1 << NF) - 1;
}

template <typename T, int N1, int NF, int E>
static int launch_col_e(int ct, int count, const OpoInst *insts, const ColArgs &a, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
  using V = typename vec2<T>::type;
  dim3 grid((a.N2 / ct) * (a.xcd_pack ? 8 : 1), count);
  const bool mid = col_is_mid<NF>(a);
  // one exchange buffer when the slice-interior kernel's field teams are
  // single wavefronts (the rule col_stage applies), two otherwise
  const int nbuf = (mid && team_in_wave(ct * (N1 / E))) ? 1 : 2;
  const size_t lds = ((size_t)N1 + nbuf * (size_t)NF * ct * N1) * sizeof(V) + 64 * sizeof(double);
  const bool direct = mid && a.direct;
  static std::once_flag big_lds[3][kMaxDevices];
  const int kv = direct ? 2 : (mid ? 1 : 0);
  // the general (boundary) variant always runs 8 samples per thread
  auto kern = direct ? &col_stage<T, N1, NF, true, true, E>
                     : (mid ? &col_stage<T, N1, NF, true, false, E> : &col_stage<T, N1, NF, false, false, 8>);
  const int threads = NF * ct * (N1 / (mid ? E : 8));
  if (lds > 65536) {
    const int rc = opt_in_big_lds(reinterpret_cast<const void *>(kern), big_lds[kv]);
    if (rc != 0) return rc;
  }
  if (e0 || e1) hipExtLaunchKernelGGL(kern, grid, dim3(threads), lds, s, e0, e1, 0, insts, a);
  else hipLaunchKernelGGL(kern, grid, dim3(threads), lds, s, insts, a);
  return (int)hipGetLastError();
}

template <typename T, int N1, int NF>
static int launch_col_t(int ct, int count, const OpoInst *insts, const ColArgs &a, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
  if (a.spt == 4) return launch_col_e<T, N1, NF, 4>(ct, count, insts, a, s, e0, e1);
  if (a.spt == 8) return launch_col_e<T, N1, NF, 8>(ct, count, insts, a, s, e0, e1);
  return (int)hipErrorInvalidValue;
}

template <typename T, int N2, int NF, int E>
static int launch_row_e(int rt, int count, const OpoInst *insts, const RowArgs &a, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
  using V = typename vec2<T>::type;
  const int threads = rt * (N2 / E);
  const size_t lds = (team_in_wave(N2 / E) ? 1 : 2) * (size_t)rt * (N2 + N2 / 8) * sizeof(V);  // see row_stage
  dim3 grid((a.N1 / rt) * a.nf * (a.xcd_pack ? 8 : 1), count);
  static std::once_flag big_lds[2][kMaxDevices];
  auto kern = a.direct ? &row_stage<T, N2, NF, true, E> : &row_stage<T, N2, NF, false, E>;
  if (lds > 65536) {
    const int rc = opt_in_big_lds(reinterpret_cast<const void *>(kern), big_lds[a.direct ? 1 : 0]);
    if (rc != 0) return rc;
  }
  if (e0 || e1) hipExtLaunchKernelGGL(kern, grid, dim3(threads), lds, s, e0, e1, 0, insts, a);
  else hipLaunchKernelGGL(kern, grid, dim3(threads), lds, s, insts, a);
  return (int)hipGetLastError();
}

template <typename T, int N2, int NF>
static int launch_row_t(int rt, int count, int tiles_y, const OpoInst *insts, const RowArgs &a, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1) {
  (void)tiles_y;
  if (a.spt == 4) return launch_row_e<T, N2, NF, 4>(rt, count, insts, a, s, e0, e1);
  if (a.spt == 8) return launch_row_e<T, N2, NF, 8>(rt, count, insts, a, s, e0, e1);
  return (int)hipErrorInvalidValue;
}

#define OPO_COL_SIZES(X) X(16) X(32) X(64) X(128)
#define OPO_ROW_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

template <typename T, int NF>
static int dispatch_col(int N1, int ct, int count, const OpoInst *insts, const ColArgs &a, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1) {
  switch (N1) {
#define CASE(M) \
  case M: return launch_col_t<T, M, NF>(ct, count, insts, a, s, e0, e1);
    OPO_COL_SIZES(CASE)
#undef CASE
  }
  return (int)hipErrorInvalidValue;
}

template <typename T, int NF>
static int dispatch_row(int N2, int rt, int count, int ty, const OpoInst *insts, const RowArgs &a, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1) {
  switch (N2) {
#define CASE(M) \
  case M: return launch_row_t<T, M, NF>(rt, count, ty, insts, a, s, e0, e1);
    OPO_ROW_SIZES(CASE)
#undef CASE
  }
  return (int)hipErrorInvalidValue;
}

int launch_col(int fp32, int N1, int NF, int ct, int count, const OpoInst *insts, const ColArgs &a, void *stream,
               void *ev0, void *ev1) {
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = (hipEvent_t)ev0, e1 = (hipEvent_t)ev1;
  if (fp32)
    return NF == 3 ? dispatch_col<float, 3>(N1, ct, count, insts, a, s, e0, e1)
                   : dispatch_col<float, 2>(N1, ct, count, insts, a, s, e0, e1);
  return NF == 3 ? dispatch_col<double, 3>(N1, ct, count, insts, a, s, e0, e1)
                 : dispatch_col<double, 2>(N1, ct, count, insts, a, s, e0, e1);
}

int launch_row(int fp32, int N2, int NF, int rt, int count, int ty, const OpoInst *insts, const RowArgs &a,
               void *stream, void *ev0, void *ev1) {
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = (hipEvent_t)ev0, e1 = (hipEvent_t)ev1;
  if (fp32)
    return NF == 3 ? dispatch_row<float, 3>(N2, rt, count, ty, insts, a, s, e0, e1)
                   : dispatch_row<float, 2>(N2, rt, count, ty, insts, a, s, e0, e1);
  return NF == 3 ? dispatch_row<double, 3>(N2, rt, count, ty, insts, a, s, e0, e1)
                 : dispatch_row<double, 2>(N2, rt, count, ty, insts, a, s, e0, e1);
}

// plain streaming copy, 16 bytes per thread per iteration (grid-stride)
__global__ __launch_bounds__(256) void copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int launch_copy(const void *src, void *dst, long long n, void *stream, void *ev0, void *ev1) {
  const long long blocks = std::min<long long>((n + 255) / 256, 256LL * 64);
  hipExtLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (hipEvent_t)ev0,
                        (hipEvent_t)ev1, 0, (const double2 *)src, (double2 *)dst, n);
  return (int)hipGetLastError();
}

int launch_tick(OpoInst *insts, const TickArgs &a, void *stream) {
  hipLaunchKernelGGL(tick_kernel, dim3(a.count), dim3(64), 0, (hipStream_t)stream, insts, a);
  return (int)hipGetLastError();
}

// ------------------------------------------------ round-trip extrapolation
// accel_dots: one block per (instance, resonant field); each thread sums a
// strided subset in order, then a fixed tree over the block (deterministic).
template <typename T>
__global__ __launch_bounds__(256) void accel_dots(const AccelDesc *__restrict__ d, int nres, long long N,
                                                  double *__restrict__ out) {
  using V = typename vec2<T>::type;
  const AccelDesc D = d[blockIdx.x];
  const long long off = (long long)blockIdx.y * N;
  const auto a = gp(reinterpret_cast<const V *>(D.s0)) + off;
  const auto b = gp(reinterpret_cast<const V *>(D.s1)) + off;
  const auto c = gp(reinterpret_cast<const V *>(D.s2)) + off;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long long i = threadIdx.x; i < N; i += blockDim.x) {
    const V x = a[i], y = b[i], z = c[i];
    const double xr = x.x, xi = x.y, yr = y.x, yi = y.y, zr = z.x, zi = z.y;
    acc[0] += xr * xr + xi * xi;
    acc[1] += yr * yr + yi * yi;
    acc[2] += zr * zr + zi * zi;
    acc[3] += xr * yr + xi * yi; acc[4] += xr * yi - xi * yr;  // conj(x) y
    acc[5] += xr * zr + xi * zi; acc[6] += xr * zi - xi * zr;  // conj(x) z
    acc[7] += yr * zr + yi * zi; acc[8] += yr * zi - yi * zr;  // conj(y) z
  }
  __shared__ double red[9][256];
#pragma unroll
  for (int j = 0; j < 9; ++j) red[j][threadIdx.x] = acc[j];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
#pragma unroll
      for (int j = 0; j < 9; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x < 9) out[((long long)blockIdx.x * nres + blockIdx.y) * 9 + threadIdx.x] = red[threadIdx.x][0];
  (void)nres;
}

// accel_apply: the instance's resonant fields <- s2 + step (s2 - R s1)
template <typename T>
__global__ __launch_bounds__(256) void accel_apply(const AccelDesc *__restrict__ d, int nres, long long N) {
  using V = typename vec2<T>::type;
  const AccelDesc D = d[blockIdx.y];
  const auto b = gp(reinterpret_cast<const V *>(D.s1));
  const auto c = gp(reinterpret_cast<const V *>(D.s2));
  const auto st = gp(reinterpret_cast<V *>(D.state)) + N;  // signal at N, idler at 2N
  const long long n = (long long)nres * N;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const V y = b[i], z = c[i];
    const double sg = i < N ? 1.0 : -1.0;  // idler rotates the other way
    const double rr = (double)y.x * D.c - sg * (double)y.y * D.s, ri = (double)y.y * D.c + sg * (double)y.x * D.s;
    V o;
    o.x = (decltype(o.x))((double)z.x + D.step * ((double)z.x - rr));
    o.y = (decltype(o.y))((double)z.y + D.step * ((double)z.y - ri));
    st[i] = o;
  }
}

int launch_accel_dots(int fp32, const AccelDesc *d, int count, int nres, long long N, double *out, void *stream) {
  if (count <= 0) return 0;
  dim3 grid((unsigned)count, (unsigned)nres);
  if (fp32) hipLaunchKernelGGL(accel_dots<float>, grid, dim3(256), 0, (hipStream_t)stream, d, nres, N, out);
  else hipLaunchKernelGGL(accel_dots<double>, grid, dim3(256), 0, (hipStream_t)stream, d, nres, N, out);
  return (int)hipGetLastError();
}

int launch_accel_apply(int fp32, const AccelDesc *d, int count, int nres, long long N, void *stream) {
  if (count <= 0) return 0;
  const long long per = std::min<long long>(((long long)nres * N + 255) / 256, 64);
  dim3 grid((unsigned)per, (unsigned)count);
  if (fp32) hipLaunchKernelGGL(accel_apply<float>, grid, dim3(256), 0, (hipStream_t)stream, d, nres, N);
  else hipLaunchKernelGGL(accel_apply<double>, grid, dim3(256), 0, (hipStream_t)stream, d, nres, N);
  return (int)hipGetLastError();
}

// zero fill of n 8-byte words (grid-stride; a kernel rather than a memset so
// the clear is an ordinary launch on the caller's stream)
__global__ __launch_bounds__(256) void zero_kernel(unsigned long long *__restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = 0ull;
}

int launch_zero(void *dst, long long words, void *stream) {
  if (words <= 0) return 0;
  const long long blocks = std::min<long long>((words + 255) / 256, 1024);
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (unsigned long long *)dst, words);
  return (int)hipGetLastError();
}

// ------------------------------------------------ no-allocation scratch
// Device memory of the no-allocation entry point: one array of kScratchBytes
// per GPU, loaded with the code object and held until the process ends.  In
// host code the array's name is not a device address; the address of this
// device's copy comes from hipGetSymbolAddress, looked up once per device
// (with that device current) and kept.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int scratch_region(void **ptr, size_t *bytes) {
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  *bytes = kScratchBytes;
  return 0;
}

}  // namespace opo
