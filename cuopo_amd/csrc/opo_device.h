// opo_device.h -- device-side building blocks for the gfx950 SSFM kernels:
// complex arithmetic on double2/float2, radix-2/4/8 DFT butterflies and a
// Stockham FFT executed by a "team" of M/E threads that each hold E = 4 or 8
// samples in registers and exchange through LDS.
//
// Sign convention: S = +1 computes sum_j x_j e^{+2 pi i jk/M} (the transform
// cuOPO reaches with CUFFT_INVERSE, src/headers/cwes3.cuh:175), S = -1 the
// CUFFT_FORWARD one (cwes3.cuh:189).  Transforms are unnormalised.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace opo {

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

// A device-memory pointer in the global address space.  Pointers read from
// instance descriptors are generic, and loads and stores through a generic
// pointer are FLAT operations, which the hardware also counts on the LDS /
// scalar-memory counter: every later LDS wait of the wave would then also wait
// for them.  Through a global pointer they are GLOBAL operations (uniform
// reads become scalar loads), so field loads in flight overlap the transform's
// LDS exchanges.  Only device-memory (hipMalloc) pointers are cast.
// Scalars: a plain global pointer.  Complex values (double2 / float2, whose
// class assignment cannot bind to a global-space reference): a pointer to the
// native two-element vector of the same layout, indexed through a proxy that
// loads and stores one 16- (8-) byte element.
template <typename S> __device__ __forceinline__ __attribute__((address_space(1))) S *gp_scalar(S *p) {
  return (__attribute__((address_space(1))) S *)p;
}
template <typename V> struct native2;
template <> struct native2<double2> { using type = double __attribute__((ext_vector_type(2))); };
template <> struct native2<float2> { using type = float __attribute__((ext_vector_type(2))); };
template <typename V, bool W> struct GPtr {
  using N = typename native2<V>::type;
  using P = typename std::conditional<W, __attribute__((address_space(1))) N *,
                                      const __attribute__((address_space(1))) N *>::type;
  P p;
  struct Ref {
    P q;
    __device__ __forceinline__ operator V() const { const N r = *q; V v; v.x = r.x; v.y = r.y; return v; }
    __device__ __forceinline__ const Ref &operator=(V v) const {
      static_assert(W, "store through a read-only pointer");
      N r; r.x = v.x; r.y = v.y; *q = r; return *this;
    }
  };
  __device__ __forceinline__ Ref operator[](long long i) const { return Ref{p + i}; }
  __device__ __forceinline__ GPtr operator+(long long i) const { return GPtr{p + i}; }
  __device__ __forceinline__ explicit operator bool() const { return p != nullptr; }
};
template <typename V> __device__ __forceinline__ GPtr<V, true> gp(V *p) {
  return GPtr<V, true>{(typename GPtr<V, true>::P)p};
}
template <typename V> __device__ __forceinline__ GPtr<V, false> gp(const V *p) {
  return GPtr<V, false>{(typename GPtr<V, false>::P)p};
}

template <typename V> __device__ __forceinline__ V mk(decltype(V::x) a, decltype(V::x) b) {
  V r; r.x = a; r.y = b; return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return mk<V>(a.x, -a.y); }
template <typename V> __device__ __forceinline__ V cscale(V a, decltype(V::x) s) { return mk<V>(a.x * s, a.y * s); }
// multiply by S*i
template <int S, typename V> __device__ __forceinline__ V mul_si(V a) {
  return S > 0 ? mk<V>(-a.y, a.x) : mk<V>(a.y, -a.x);
}
// twiddle lookup from a table of e^{+2 pi i e/M}; conjugated for S < 0
template <int S, typename V> __device__ __forceinline__ V tw_sign(V w) { return S > 0 ? w : cconj(w); }

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }
// passes of radix E (E = 4 or 8 samples per thread), the last one of radix
// 2^(log2 M mod log2 E) when log2 M is not a multiple of log2 E
constexpr int npasses(int M, int E) { return (ilog2c(M) + ilog2c(E) - 1) / ilog2c(E); }
constexpr int radix_at(int M, int q, int E) {
  return q < ilog2c(M) / ilog2c(E) ? E : (1 << (ilog2c(M) - ilog2c(E) * (ilog2c(M) / ilog2c(E))));
}
constexpr int p_at(int q, int E) { return 1 << (ilog2c(E) * q); }

// ---- butterflies (natural-order in, natural-order out) ----
template <int S, typename V> __device__ __forceinline__ void dft2(V* u) {
  V a = u[0], b = u[1];
  u[0] = cadd(a, b); u[1] = csub(a, b);
}
template <int S, typename V> __device__ __forceinline__ void dft4(V* u) {
  V c0 = cadd(u[0], u[2]), c1 = cadd(u[1], u[3]);
  V d0 = csub(u[0], u[2]), d1 = mul_si<S>(csub(u[1], u[3]));
  u[0] = cadd(c0, c1); u[1] = cadd(d0, d1); u[2] = csub(c0, c1); u[3] = csub(d0, d1);
}
template <int S, typename V> __device__ __forceinline__ void dft8(V* u) {
  using T = decltype(V::x);
  const T r = (T)0.70710678118654752440;
  V a[4], b[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) { a[n] = cadd(u[n], u[n + 4]); b[n] = csub(u[n], u[n + 4]); }
  // b1 *= e^{S i pi/4}, b2 *= S i, b3 *= e^{3 S i pi/4}
  b[1] = mk<V>((b[1].x - (T)S * b[1].y) * r, (b[1].y + (T)S * b[1].x) * r);
  b[2] = mul_si<S>(b[2]);
  b[3] = mk<V>((-b[3].x - (T)S * b[3].y) * r, (-b[3].y + (T)S * b[3].x) * r);
  dft4<S>(a); dft4<S>(b);
#pragma unroll
  for (int k = 0; k < 4; ++k) { u[2 * k] = a[k]; u[2 * k + 1] = b[k]; }
}
template <int R, int S, typename V> __device__ __forceinline__ void dftr(V* u) {
  if constexpr (R == 8) dft8<S>(u);
  else if constexpr (R == 4) dft4<S>(u);
  else if constexpr (R == 2) dft2<S>(u);
}

// Team FFT of length M: a team of M/E threads, each holding E samples in
// registers (E = 8: radix-8 passes; E = 4: radix-4 passes, twice the threads
// for latency-bound shapes), with a radix-2 or -4 pass last when log2 M is
// not a multiple of log2 E.  On entry v[m] = x[t + m*M/E]; on exit
// v[m] = X[t + m*M/E].  Passes exchange through LDS; the last pass's outputs
// already sit in the natural distribution (output i + j*M/R belongs to the
// thread that read input i + j*M/R), so it runs in registers and an FFT of
// npasses(M, E) passes costs npasses(M, E) - 1 exchanges.
//
// Team element e lives at buf[gbase + lidx(e)], lidx(e) = e*es (PAD false,
// column tiles: es interleaved columns) or e + e/8 (PAD true, contiguous rows:
// one pad slot per 8 samples spreads the strided writes of the first pass over
// the banks).  The LDS region holds two ping-pong buffers `bufstride`
// elements apart; `ph` (block-uniform) alternates them across every exchange
// of the kernel, so one barrier per exchange suffices: a buffer is rewritten
// only after the barrier that follows the next exchange's writes.  `on` (per
// thread) skips the arithmetic of threads whose data is not transformed while
// every thread still reaches the barriers.
//
// `wave` (block-uniform): the team is one wavefront or part of one, and no
// other team shares its LDS region.  LDS operations of a wavefront execute in
// issue order, so the team then exchanges through a single buffer in place
// (every lane's reads of an exchange are issued before any lane's writes of
// the next) with wave-scope ordering only: half the LDS and no block barrier.
//
// Twiddles e^{+2 pi i e/M} come from `tws`: with TREG false an M-point root
// table (LDS) indexed by e; with TREG true this thread's register set filled
// by load_twiddles, indexed by (pass, butterfly, leg).

template <bool PAD> __device__ __forceinline__ int lidx(int e, int es) { return PAD ? e + (e >> 3) : e * es; }

// Per-thread twiddle registers of passes 1 .. Q-1 (pass 0 has none).  FP64
// (`pow`) keeps one root per butterfly, w = e^{2 pi i k/(P R)}, and forms the
// other legs' roots w^2 .. w^(R-1) by products (at most three deep, so a few
// ulp, ~1e-15): fewer live registers, and the row stage holds more waves per
// SIMD.  FP32 keeps every leg's root as looked up, since the same product
// rounding would be 1e-7 there.
constexpr int tw_off(int M, int Q, int E, bool pow) {
  return Q <= 1 ? 0
                : tw_off(M, Q - 1, E, pow) +
                      (pow ? 1 : radix_at(M, Q - 1, E) - 1) * (E / radix_at(M, Q - 1, E));
}
constexpr int tw_count(int M, int E, bool pow) { return tw_off(M, npasses(M, E), E, pow); }
template <typename V> constexpr bool tw_pow() { return std::is_same<V, double2>::value; }

// u[j] *= w^j for j = 1 .. R-1, the powers by squaring
template <int R, typename V> __device__ __forceinline__ void mul_root_powers(V *u, V w) {
  V p[R > 1 ? R : 2];
  p[1] = w;
  if constexpr (R > 2) { p[2] = cmul(w, w); p[3] = cmul(p[2], w); }
  if constexpr (R > 4) { p[4] = cmul(p[2], p[2]); p[5] = cmul(p[4], w); p[6] = cmul(p[4], p[2]); p[7] = cmul(p[4], p[3]); }
#pragma unroll
  for (int j = 1; j < R; ++j) u[j] = cmul(u[j], p[j]);
}

// The single-wavefront rule, shared by the kernels and by the launchers that
// size their LDS (one exchange buffer when it holds, two otherwise): a team
// of `team` threads (a power of two) fits in one wavefront.
__host__ __device__ constexpr bool team_in_wave(int team) { return team <= 64; }

// orders this wavefront's LDS accesses (no instruction: a wave's LDS
// operations already execute in order; the fences keep the compiler from
// moving them across)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int M, int Q, int S, bool PAD, bool TREG, int E, typename V>
__device__ __forceinline__ void fft_pass(V *v, bool on, V *buf0, int &ph, int t, int gbase, int es, int bufstride,
                                         const V *tws, bool wave) {
  constexpr int R = radix_at(M, Q, E);
  constexpr int P = p_at(Q, E);
  constexpr int T = M / E;
  constexpr int NB = E / R;
  constexpr bool LAST = Q == npasses(M, E) - 1;
  constexpr bool POW = TREG && tw_pow<V>();
  constexpr int TOFF = tw_off(M, Q, E, POW);  // constant: keeps the register set statically indexed
  V *buf = wave ? buf0 : buf0 + (ph & 1) * bufstride;
  if (on) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int i = t + b * T;
      const int k = i & (P - 1);
      V u[R];
#pragma unroll
      for (int j = 0; j < R; ++j) u[j] = v[b + j * NB];
      if constexpr (P > 1) {
        if constexpr (POW) {
          mul_root_powers<R>(u, tw_sign<S>(tws[TOFF + b]));
        } else {
#pragma unroll
          for (int j = 1; j < R; ++j) {
            if constexpr (TREG) u[j] = cmul(u[j], tw_sign<S>(tws[TOFF + b * (R - 1) + j - 1]));
            else u[j] = cmul(u[j], tw_sign<S>(tws[j * k * (M / (P * R))]));
          }
        }
      }
      dftr<R, S>(u);
      if constexpr (LAST) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[b + j * NB] = u[j];
      } else {
        const int base = (i / P) * P * R + k;
#pragma unroll
        for (int j = 0; j < R; ++j) buf[gbase + lidx<PAD>(base + j * P, es)] = u[j];
      }
    }
  }
  if constexpr (!LAST) {
    if (wave) wave_sync();
    else __syncthreads();
    if (on) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = buf[gbase + lidx<PAD>(t + m * T, es)];
    }
    if (wave) wave_sync();
    ++ph;
  }
}

template <int M, int S, bool PAD, bool TREG, int E, int Q = 0, typename V>
__device__ __forceinline__ void team_fft(V *v, bool on, V *buf0, int &ph, int t, int gbase, int es, int bufstride,
                                         const V *tws, bool wave) {
  if constexpr (Q < npasses(M, E)) {
    fft_pass<M, Q, S, PAD, TREG, E>(v, on, buf0, ph, t, gbase, es, bufstride, tws, wave);
    team_fft<M, S, PAD, TREG, E, Q + 1>(v, on, buf0, ph, t, gbase, es, bufstride, tws, wave);
  }
}

// This thread's twiddles of every pass, from the M-point root table `g`
// (global memory), into treg[tw_count(M, E, tw_pow<V>())]; both transform
// directions use them.
template <int M, int E, int Q = 1, typename V, typename G>
__device__ __forceinline__ void load_twiddles(V *treg, int t, G g) {
  if constexpr (Q < npasses(M, E)) {
    constexpr int R = radix_at(M, Q, E);
    constexpr int P = p_at(Q, E);
    constexpr int T = M / E;
    constexpr int NB = E / R;
    constexpr bool POW = tw_pow<V>();
    constexpr int TOFF = tw_off(M, Q, E, POW);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = (t + b * T) & (P - 1);
      if constexpr (POW) {
        treg[TOFF + b] = V(g[k * (M / (P * R))]);
      } else {
#pragma unroll
        for (int j = 1; j < R; ++j) treg[TOFF + b * (R - 1) + j - 1] = V(g[j * k * (M / (P * R))]);
      }
    }
    load_twiddles<M, E, Q + 1>(treg, t, g);
  }
}

}  // namespace opo
