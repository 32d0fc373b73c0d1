// opo_device.h -- device-side building blocks for the gfx950 SSFM kernels:
// complex arithmetic on double2/float2, radix-2/4/8 DFT butterflies and a
// Stockham FFT executed by a "team" of M/8 threads that each hold 8 samples in
// registers and exchange through LDS.
//
// Sign convention: S = +1 computes sum_j x_j e^{+2 pi i jk/M} (the transform
// cuOPO reaches with CUFFT_INVERSE, src/headers/cwes3.cuh:175), S = -1 the
// CUFFT_FORWARD one (cwes3.cuh:189).  Transforms are unnormalised.
#pragma once

#include <hip/hip_runtime.h>

namespace opo {

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

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
constexpr int npasses(int M) { return (ilog2c(M) + 2) / 3; }
constexpr int radix_at(int M, int q) { return q < ilog2c(M) / 3 ? 8 : (1 << (ilog2c(M) - 3 * (ilog2c(M) / 3))); }
constexpr int p_at(int q) { return 1 << (3 * q); }

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

// One Stockham pass (radix R, stride P) for a team of M/8 threads.  On entry
// v[m] holds x[t + m*M/8]; outputs go to LDS element (i/P)*P*R + k + j*P.
// Team element e lives at buf[gbase + e*es].
template <int M, int R, int P, int S, typename V>
__device__ __forceinline__ void stockham_pass(const V* v, V* buf, int t, int gbase, int es, const V* tw) {
  constexpr int T = M / 8;
  constexpr int NB = 8 / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int i = t + b * T;
    const int k = i & (P - 1);
    V u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) u[j] = v[b + j * NB];
    if constexpr (P > 1) {
#pragma unroll
      for (int j = 1; j < R; ++j) u[j] = cmul(u[j], tw_sign<S>(tw[j * k * (M / (P * R))]));
    }
    dftr<R, S>(u);
    const int base = (i / P) * P * R + k;
#pragma unroll
    for (int j = 0; j < R; ++j) buf[gbase + (base + j * P) * es] = u[j];
  }
}

// Team FFT of length M over NFX fields at once (field f transformed when bit
// f of `mask` is set; the mask is block-uniform).  v[f][m] = x_f[t + m*M/8]
// on entry, X_f[t + m*M/8] on exit.  Field f's element e lives at
// lds[f*fstride + gbase + e*es]; lds holds two ping-pong buffers `bufstride`
// elements apart, and `ph` (block-uniform) alternates them across every pass
// of the kernel, so one barrier per pass suffices: a buffer is rewritten only
// after the barrier that follows the next pass's writes.  `tw` points to the
// M-point root table in LDS.  Every thread of the block must call this.
template <int M, int S, int NFX, int Q = 0, typename V>
__device__ __forceinline__ void team_fft(V (*v)[8], int mask, V* lds, int& ph, int t, int gbase, int es,
                                         int fstride, int bufstride, const V* tw) {
  if constexpr (Q < npasses(M)) {
    constexpr int R = radix_at(M, Q);
    constexpr int P = p_at(Q);
    constexpr int T = M / 8;
    V* buf = lds + (ph & 1) * bufstride;
#pragma unroll
    for (int f = 0; f < NFX; ++f)
      if ((mask >> f) & 1) stockham_pass<M, R, P, S>(v[f], buf + f * fstride, t, gbase, es, tw);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NFX; ++f)
      if ((mask >> f) & 1) {
#pragma unroll
        for (int m = 0; m < 8; ++m) v[f][m] = buf[f * fstride + gbase + (t + m * T) * es];
      }
    ++ph;
    team_fft<M, S, NFX, Q + 1>(v, mask, lds, ph, t, gbase, es, fstride, bufstride, tw);
  }
}

// Single-field variant: each thread owns one field (its wave's field); `on`
// (per thread) skips the work of threads whose field is not transformed,
// while every thread still reaches the barriers.  `buf0` is this thread's
// field region in buffer 0; buffer 1 is `bufstride` elements further.
template <int M, int S, int Q = 0, typename V>
__device__ __forceinline__ void team_fft1(V* v, bool on, V* buf0, int& ph, int t, int gbase, int es, int bufstride,
                                          const V* tw) {
  if constexpr (Q < npasses(M)) {
    constexpr int R = radix_at(M, Q);
    constexpr int P = p_at(Q);
    constexpr int T = M / 8;
    V* buf = buf0 + (ph & 1) * bufstride;
    if (on) stockham_pass<M, R, P, S>(v, buf, t, gbase, es, tw);
    __syncthreads();
    if (on) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = buf[gbase + (t + m * T) * es];
    }
    ++ph;
    team_fft1<M, S, Q + 1>(v, on, buf0, ph, t, gbase, es, bufstride, tw);
  }
}

}  // namespace opo
