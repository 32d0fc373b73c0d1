// opo_mlp.hip -- the warm-start surrogate: a small MLP (n_in -> width x depth,
// GELU -> n_out) trained and evaluated on the GPU with BF16 weights and
// activations and FP32 accumulation.  Every matrix product (forward, weight
// gradient, input gradient) runs through one hand-written MFMA kernel
// (v_mfma_f32_16x16x16_bf16, 64x64 block tiles); bias/GELU/GELU' are fused into
// its epilogue.  FP32 master weights, Adam in FP32.  No reference
// counterpart: cuOPO has no learned component.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/opo.h"

namespace opo {
void set_error(const std::string &msg);  // opo_api.cpp: message returned by opo_last_error()
}

namespace {

int mfail(int code, const std::string &m) {
  opo::set_error(m);
  return code;
}
#define MCHK(x)                                                                                 \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) return mfail(OPO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define MRCHK(x)                  \
  do {                            \
    int r_ = (x);                 \
    if (r_ != OPO_OK) return r_;  \
  } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__host__ __device__ inline unsigned short f2bf(float f) {
  unsigned int u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h) {
  unsigned int u = (unsigned int)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum Epi : int {
  EPI_STORE = 0,       // C = acc (fp32)
  EPI_BIAS = 1,        // C = acc + bias[n] (fp32), no activation (output layer)
  EPI_BIAS_GELU = 2,   // Z = acc + bias[n] (fp32, optional), H = gelu(Z) (bf16)
  EPI_GELU_GRAD = 3,   // C = acc * gelu'(Zprev[m, n]) (fp32)
};

// Element (r, c) of a matrix stored with row stride sr and column stride sc.
template <typename T> __device__ inline float ldv(const T *p, long long i);
template <> __device__ inline float ldv<float>(const float *p, long long i) { return p[i]; }
template <> __device__ inline float ldv<unsigned short>(const unsigned short *p, long long i) { return bf2f(p[i]); }

struct GemmArgs {
  int M, N, K;
  long long a_sm, a_sk;  // A(m, k) = A[m*a_sm + k*a_sk]
  long long b_sk, b_sn;  // B(k, n) = B[k*b_sk + n*b_sn]
  long long ldc;         // C(m, n) = C[m*ldc + n]
  const float *bias;     // [N]
  const float *zprev;    // EPI_GELU_GRAD: Z of the layer below, same shape as C
  float *zout;           // EPI_BIAS_GELU: pre-activation out (may be null)
  unsigned short *hout;  // EPI_BIAS_GELU: bf16 activation out
  float *cout;           // fp32 out
  int epi;
};

// C[M x N] = A[M x K] * B[K x N]: 256 threads, 64 x 64 tile, each wave a 32 x 32
// quarter as 2 x 2 MFMA 16x16x16 blocks; K in steps of 16 through LDS.
// Operand lanes of v_mfma_f32_16x16x16_bf16: lane l carries A(l%16, 4(l/16)+q)
// and B(4(l/16)+q, l%16), q = 0..3; result register q holds D(4(l/16)+q, l%16).
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void gemm_bf16(const TA *__restrict__ A, const TB *__restrict__ B, GemmArgs g) {
  __shared__ unsigned short As[64][16 + 2];
  __shared__ unsigned short Bs[16][64 + 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m_base = blockIdx.y * 64, n_base = blockIdx.x * 64;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    for (int e = tid; e < 64 * 16; e += 256) {
      const int r = e / 16, kk = e % 16;  // A tile (64 x 16)
      const int m = m_base + r, k = k0 + kk;
      float v = (m < g.M && k < g.K) ? ldv(A, (long long)m * g.a_sm + (long long)k * g.a_sk) : 0.f;
      As[r][kk] = f2bf(v);
      const int kb = e / 64, c = e % 64;  // B tile (16 x 64)
      const int kB = k0 + kb, n = n_base + c;
      float u = (kB < g.K && n < g.N) ? ldv(B, (long long)kB * g.b_sk + (long long)n * g.b_sn) : 0.f;
      Bs[kb][c] = f2bf(u);
    }
    __syncthreads();
    bf16x4 a[2], b[2];
    for (int i = 0; i < 2; ++i) {
      const int r = wm * 32 + i * 16 + (lane & 15);
      for (int q = 0; q < 4; ++q) a[i][q] = (short)As[r][4 * (lane >> 4) + q];
    }
    for (int j = 0; j < 2; ++j) {
      const int c = wn * 32 + j * 16 + (lane & 15);
      for (int q = 0; q < 4; ++q) b[j][q] = (short)Bs[4 * (lane >> 4) + q][c];
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) {
        const int m = m_base + wm * 32 + i * 16 + 4 * (lane >> 4) + q;
        const int n = n_base + wn * 32 + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const long long o = (long long)m * g.ldc + n;
        float v = acc[i][j][q];
        if (g.epi == EPI_STORE) {
          g.cout[o] = v;
        } else if (g.epi == EPI_BIAS) {
          g.cout[o] = v + g.bias[n];
        } else if (g.epi == EPI_BIAS_GELU) {
          const float z = v + g.bias[n];
          if (g.zout) g.zout[o] = z;
          g.hout[o] = f2bf(gelu(z));
        } else {
          g.cout[o] = v * gelu_grad(g.zprev[o]);
        }
      }
}

// dY = 2 (Y - T) / (B * n_out); per-block partial sums of (Y - T)^2
__global__ void mse_grad(const float *Y, const float *T, float *dY, float *partial, long long n, float scale) {
  __shared__ float red[256];
  float s = 0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * scale;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// db[n] = sum_b dZ[b, n]
__global__ void bias_grad(const float *dZ, float *db, int Bn, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0;
  for (int b = 0; b < Bn; ++b) s += dZ[(long long)b * N + n];
  db[n] = s;
}

__global__ void adam_step(float *w, unsigned short *wbf, const float *gr, float *m, float *v, long long n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float g = gr[i];
    const float mi = b1 * m[i] + (1 - b1) * g;
    const float vi = b2 * v[i] + (1 - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = f2bf(wi);
  }
}

__global__ void to_bf16(const float *x, unsigned short *y, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) y[i] = f2bf(x[i]);
}

}  // namespace

// ---------------------------------------------------------------- host side

struct opo_mlp {
  int n_in = 0, width = 0, depth = 0, n_out = 0, max_batch = 0, device = 0;
  int nl = 0;                                // layers = depth + 1
  std::vector<int> din, dout;                // per layer
  std::vector<size_t> woff, boff;            // offsets into the flat parameter vector
  size_t nparams = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 master, grads, Adam moments
  unsigned short *wbf = nullptr;             // BF16 copy of the flat vector
  unsigned short *x_bf = nullptr;            // input batch (bf16)
  std::vector<float *> Z;                    // pre-activations per hidden layer (fp32)
  std::vector<unsigned short *> H;           // activations per hidden layer (bf16)
  float *Y = nullptr, *Tt = nullptr, *dA = nullptr, *dB = nullptr, *partial = nullptr, *xf = nullptr;
  long long step = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  ~opo_mlp() {
    for (void *p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int malloc_m(opo_mlp *M, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return mfail(OPO_ERR_NOMEM, "mlp allocation");
  M->allocs.push_back(*p);
  return OPO_OK;
}

template <typename TA, typename TB>
int gemm(opo_mlp *M, const TA *A, const TB *B, GemmArgs g) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0) return mfail(OPO_ERR_INVALID, "empty gemm");
  dim3 grid((g.N + 63) / 64, (g.M + 63) / 64);
  hipLaunchKernelGGL((gemm_bf16<TA, TB>), grid, dim3(256), 0, M->stream, A, B, g);
  MCHK(hipGetLastError());
  return OPO_OK;
}

// forward pass for B rows already in M->x_bf; store_z keeps pre-activations for backward
int forward(opo_mlp *M, int Bn, bool store_z) {
  const unsigned short *in = M->x_bf;
  for (int l = 0; l < M->nl; ++l) {
    GemmArgs g{};
    g.M = Bn; g.N = M->dout[l]; g.K = M->din[l];
    g.a_sm = M->din[l]; g.a_sk = 1;        // activations [B x din]
    g.b_sk = 1; g.b_sn = M->din[l];        // W^T: W is [dout x din]
    g.ldc = M->dout[l];
    g.bias = M->w + M->boff[l];
    const unsigned short *W = M->wbf + M->woff[l];
    if (l < M->nl - 1) {
      g.epi = EPI_BIAS_GELU;
      g.zout = store_z ? M->Z[l] : nullptr;
      g.hout = M->H[l];
      MRCHK(gemm(M, in, W, g));
      in = M->H[l];
    } else {
      g.epi = EPI_BIAS;
      g.cout = M->Y;
      MRCHK(gemm(M, in, W, g));
    }
  }
  return OPO_OK;
}

// gradients of the MSE loss for the batch in x_bf / Tt; returns the loss
int backward(opo_mlp *M, int Bn, float *loss) {
  const long long ny = (long long)Bn * M->n_out;
  const int nb = 256;
  hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, M->stream, M->Y, M->Tt, M->dA, M->partial, ny,
                     1.0f / (float)ny);
  MCHK(hipGetLastError());
  float *dZ = M->dA, *dNext = M->dB;
  for (int l = M->nl - 1; l >= 0; --l) {
    const unsigned short *Hin = l == 0 ? M->x_bf : M->H[l - 1];
    // dW_l = dZ^T (B x dout -> dout x B) * Hin (B x din)
    GemmArgs gw{};
    gw.M = M->dout[l]; gw.N = M->din[l]; gw.K = Bn;
    gw.a_sm = 1; gw.a_sk = M->dout[l];
    gw.b_sk = M->din[l]; gw.b_sn = 1;
    gw.ldc = M->din[l];
    gw.epi = EPI_STORE;
    gw.cout = M->g + M->woff[l];
    MRCHK(gemm(M, dZ, Hin, gw));
    hipLaunchKernelGGL(bias_grad, dim3((M->dout[l] + 255) / 256), dim3(256), 0, M->stream, dZ, M->g + M->boff[l], Bn,
                       M->dout[l]);
    MCHK(hipGetLastError());
    if (l == 0) break;
    // dZ_{l-1} = (dZ * W_l) .* gelu'(Z_{l-1});  dZ [B x dout], W [dout x din]
    GemmArgs gh{};
    gh.M = Bn; gh.N = M->din[l]; gh.K = M->dout[l];
    gh.a_sm = M->dout[l]; gh.a_sk = 1;
    gh.b_sk = M->din[l]; gh.b_sn = 1;
    gh.ldc = M->din[l];
    gh.epi = EPI_GELU_GRAD;
    gh.zprev = M->Z[l - 1];
    gh.cout = dNext;
    MRCHK(gemm(M, dZ, M->wbf + M->woff[l], gh));
    std::swap(dZ, dNext);
  }
  std::vector<float> part(nb);
  MCHK(hipMemcpyAsync(part.data(), M->partial, nb * sizeof(float), hipMemcpyDeviceToHost, M->stream));
  MCHK(hipStreamSynchronize(M->stream));
  double s = 0;
  for (float p : part) s += p;
  if (loss) *loss = (float)(s / (double)ny);
  return OPO_OK;
}

int upload_batch(opo_mlp *M, const float *x, const float *t, int Bn) {
  if (Bn < 1 || Bn > M->max_batch) return mfail(OPO_ERR_INVALID, "batch size out of range");
  MCHK(hipMemcpyAsync(M->xf, x, (size_t)Bn * M->n_in * sizeof(float), hipMemcpyHostToDevice, M->stream));
  hipLaunchKernelGGL(to_bf16, dim3(64), dim3(256), 0, M->stream, M->xf, M->x_bf, (long long)Bn * M->n_in);
  MCHK(hipGetLastError());
  if (t) MCHK(hipMemcpyAsync(M->Tt, t, (size_t)Bn * M->n_out * sizeof(float), hipMemcpyHostToDevice, M->stream));
  return OPO_OK;
}

}  // namespace

extern "C" {

int opo_mlp_create(int n_in, int width, int depth, int n_out, int max_batch, uint64_t seed, int device,
                   opo_mlp **out) {
  if (!out || n_in < 1 || width < 1 || depth < 1 || n_out < 1 || max_batch < 1)
    return mfail(OPO_ERR_INVALID, "bad MLP shape");
  *out = nullptr;
  std::unique_ptr<opo_mlp> M(new opo_mlp());
  M->n_in = n_in; M->width = width; M->depth = depth; M->n_out = n_out; M->max_batch = max_batch;
  M->device = device;
  MCHK(hipSetDevice(device));
  MCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
  M->nl = depth + 1;
  for (int l = 0; l < M->nl; ++l) {
    M->din.push_back(l == 0 ? n_in : width);
    M->dout.push_back(l == M->nl - 1 ? n_out : width);
  }
  size_t off = 0;
  for (int l = 0; l < M->nl; ++l) {
    M->woff.push_back(off); off += (size_t)M->din[l] * M->dout[l];
    M->boff.push_back(off); off += (size_t)M->dout[l];
  }
  M->nparams = off;
  const size_t np = M->nparams;
  MRCHK(malloc_m(M.get(), (void **)&M->w, np * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->g, np * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->m, np * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->v, np * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->wbf, np * 2));
  MRCHK(malloc_m(M.get(), (void **)&M->xf, (size_t)max_batch * n_in * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->x_bf, (size_t)max_batch * n_in * 2));
  for (int l = 0; l < depth; ++l) {
    float *z; unsigned short *h;
    MRCHK(malloc_m(M.get(), (void **)&z, (size_t)max_batch * width * 4));
    MRCHK(malloc_m(M.get(), (void **)&h, (size_t)max_batch * width * 2));
    M->Z.push_back(z); M->H.push_back(h);
  }
  const size_t big = (size_t)max_batch * std::max(n_out, width);
  MRCHK(malloc_m(M.get(), (void **)&M->Y, (size_t)max_batch * n_out * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->Tt, (size_t)max_batch * n_out * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->dA, big * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->dB, big * 4));
  MRCHK(malloc_m(M.get(), (void **)&M->partial, 256 * 4));
  // He-style initialisation from a seeded generator; zero biases
  std::vector<float> h(np, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < M->nl; ++l) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / (float)M->din[l]));
    for (size_t i = 0; i < (size_t)M->din[l] * M->dout[l]; ++i) h[M->woff[l] + i] = nd(rng);
  }
  MCHK(hipMemcpy(M->w, h.data(), np * 4, hipMemcpyHostToDevice));
  MCHK(hipMemset(M->m, 0, np * 4));
  MCHK(hipMemset(M->v, 0, np * 4));
  hipLaunchKernelGGL(to_bf16, dim3(256), dim3(256), 0, M->stream, M->w, M->wbf, (long long)np);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(M->stream));
  *out = M.release();
  return OPO_OK;
}

int opo_mlp_param_count(opo_mlp *M, size_t *n) {
  if (!M || !n) return mfail(OPO_ERR_INVALID, "null");
  *n = M->nparams;
  return OPO_OK;
}

int opo_mlp_set_weights(opo_mlp *M, const float *w, size_t n) {
  if (!M || !w || n != M->nparams) return mfail(OPO_ERR_INVALID, "weight count mismatch");
  MCHK(hipSetDevice(M->device));
  MCHK(hipMemcpyAsync(M->w, w, n * 4, hipMemcpyHostToDevice, M->stream));
  hipLaunchKernelGGL(to_bf16, dim3(256), dim3(256), 0, M->stream, M->w, M->wbf, (long long)n);
  MCHK(hipGetLastError());
  MCHK(hipMemsetAsync(M->m, 0, n * 4, M->stream));
  MCHK(hipMemsetAsync(M->v, 0, n * 4, M->stream));
  M->step = 0;
  MCHK(hipStreamSynchronize(M->stream));
  return OPO_OK;
}

int opo_mlp_get_weights(opo_mlp *M, float *w, size_t n) {
  if (!M || !w || n != M->nparams) return mfail(OPO_ERR_INVALID, "weight count mismatch");
  MCHK(hipMemcpy(w, M->w, n * 4, hipMemcpyDeviceToHost));
  return OPO_OK;
}

int opo_mlp_forward(opo_mlp *M, const float *x, int Bn, float *y) {
  if (!M || !x || !y) return mfail(OPO_ERR_INVALID, "null");
  MCHK(hipSetDevice(M->device));
  MRCHK(upload_batch(M, x, nullptr, Bn));
  MRCHK(forward(M, Bn, false));
  MCHK(hipMemcpyAsync(y, M->Y, (size_t)Bn * M->n_out * 4, hipMemcpyDeviceToHost, M->stream));
  MCHK(hipStreamSynchronize(M->stream));
  return OPO_OK;
}

int opo_mlp_gradients(opo_mlp *M, const float *x, const float *t, int Bn, float *grad, float *loss) {
  if (!M || !x || !t) return mfail(OPO_ERR_INVALID, "null");
  MCHK(hipSetDevice(M->device));
  MRCHK(upload_batch(M, x, t, Bn));
  MRCHK(forward(M, Bn, true));
  MRCHK(backward(M, Bn, loss));
  if (grad) MCHK(hipMemcpy(grad, M->g, M->nparams * 4, hipMemcpyDeviceToHost));
  return OPO_OK;
}

int opo_mlp_train_step(opo_mlp *M, const float *x, const float *t, int Bn, float lr, float *loss) {
  if (!M || !x || !t) return mfail(OPO_ERR_INVALID, "null");
  MCHK(hipSetDevice(M->device));
  MRCHK(upload_batch(M, x, t, Bn));
  MRCHK(forward(M, Bn, true));
  MRCHK(backward(M, Bn, loss));
  M->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)M->step), c2 = 1.0f - std::pow(b2, (float)M->step);
  hipLaunchKernelGGL(adam_step, dim3(512), dim3(256), 0, M->stream, M->w, M->wbf, M->g, M->m, M->v,
                     (long long)M->nparams, lr, b1, b2, 1e-8f, c1, c2);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(M->stream));
  return OPO_OK;
}

int opo_mlp_destroy(opo_mlp *M) {
  delete M;
  return OPO_OK;
}

}  // extern "C"
