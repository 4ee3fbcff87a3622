// Warm-start MLP on MI355X: BF16-operand / FP32-accumulate MFMA GEMMs,
// GELU, MSE, Adam -- forward, backward and training loop on the device.
//
// GEMM tile: 64x64 outputs per workgroup (4 wavefronts x 32x32, each a 2x2
// block of v_mfma_f32_16x16x16_bf16 tiles), K staged through LDS in chunks of
// 16 and rounded to BF16 (round to nearest even) on the way in.  Operand
// access is by strides, so the same kernel computes X W, H' dZ and dZ W'.
//
// GELU is the exact form x * Phi(x) of Hendrycks & Gimpel (2016),
// Phi(x) = (1 + erf(x / sqrt 2)) / 2, with derivative Phi(x) + x phi(x);
// erf and exp are the HIP device math library's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "qtqp_mlp.h"

namespace {

thread_local std::string g_err;
constexpr int WG = 256;
constexpr int TM = 64, TN = 64, TK = 16;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ inline float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ inline float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// C(i, j) = sum_k A(i,k) B(k,j) (+ bias[j]); A(i,k) = A[i*sai + k*sak],
// B(k,j) = B[k*sbk + j*sbj].  Writes Z (ldc) and, if act, H = gelu(Z).
// v_mfma_f32_16x16x16_bf16: A lane l holds A[l%16][4(l/16)..+3], B lane l
// holds B[4(l/16)..+3][l%16], D register r holds D[4(l/16)+r][l%16].
__global__ __launch_bounds__(WG) void k_gemm(const float* __restrict__ A, int64_t sai, int64_t sak,
                                             const float* __restrict__ B, int64_t sbk, int64_t sbj,
                                             int M, int N, int K, const float* __restrict__ bias,
                                             float* __restrict__ Z, float* __restrict__ H,
                                             int64_t ldc, int act) {
  __shared__ short As[TM][TK + 4];
  __shared__ short Bs[TN][TK + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  const int wi = 32 * (wave >> 1), wj = 32 * (wave & 1);
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int e = tid; e < TM * TK; e += WG) {
      const int i = e / TK, k = e % TK;
      const float v = (i0 + i < M && k0 + k < K) ? A[(int64_t)(i0 + i) * sai + (int64_t)(k0 + k) * sak] : 0.f;
      As[i][k] = to_bf16(v);
    }
    for (int e = tid; e < TN * TK; e += WG) {
      const int j = e % TN, k = e / TN;
      const float v = (j0 + j < N && k0 + k < K) ? B[(int64_t)(k0 + k) * sbk + (int64_t)(j0 + j) * sbj] : 0.f;
      Bs[j][k] = to_bf16(v);
    }
    __syncthreads();
    const int r16 = lane & 15, kq = 4 * (lane >> 4);
    s4 av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const short* p = &As[wi + 16 * a + r16][kq];
      av[a] = s4{p[0], p[1], p[2], p[3]};
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const short* p = &Bs[wj + 16 * b + r16][kq];
      bv[b] = s4{p[0], p[1], p[2], p[3]};
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[a], bv[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi + 16 * a + 4 * (lane >> 4) + r;
        const int j = j0 + wj + 16 * b + (lane & 15);
        if (i < M && j < N) {
          const float v = acc[a][b][r] + (bias ? bias[j] : 0.f);
          Z[(int64_t)i * ldc + j] = v;
          if (act) H[(int64_t)i * ldc + j] = gelu(v);
        }
      }
}

// dZ = dH * gelu'(Z)
__global__ void k_gelu_back(const float* __restrict__ dH, const float* __restrict__ Z,
                            float* __restrict__ dZ, int64_t count) {
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  if (i < count) dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// dO = 2 (O - T) / count; per-block partial sums of (O - T)^2.
__global__ void k_mse(const float* __restrict__ O, const float* __restrict__ T,
                      float* __restrict__ dO, int64_t count, double* partial) {
  __shared__ double red[WG / 64];
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  double sq = 0.0;
  if (i < count) {
    const float d = O[i] - T[i];
    dO[i] = 2.f * d / (float)count;
    sq = (double)d * d;
  }
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// loss = sum(partial) / count (one workgroup, fixed order).
__global__ void k_loss(const double* partial, int nparts, int64_t count, float* out) {
  __shared__ double red[WG];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += WG) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = WG / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (float)(red[0] / (double)count);
}

// db[j] = sum_i dZ[i][j]
__global__ void k_colsum(const float* __restrict__ dZ, int rows, int cols, float* __restrict__ db) {
  const int j = blockIdx.x * WG + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int i = 0; i < rows; ++i) s += dZ[(int64_t)i * cols + j];
  db[j] = s;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Gathers rows idx[0..b) of X (n cols) and T (k cols) into Xb, Tb.
__global__ void k_gather_rows(const float* __restrict__ X, const float* __restrict__ T,
                              const int64_t* __restrict__ idx, int b, int n, int k,
                              float* __restrict__ Xb, float* __restrict__ Tb) {
  const int r = blockIdx.x;
  if (r >= b) return;
  const int64_t src = idx[r];
  for (int c = threadIdx.x; c < n; c += WG) Xb[(int64_t)r * n + c] = X[src * n + c];
  for (int c = threadIdx.x; c < k; c += WG) Tb[(int64_t)r * k + c] = T[src * k + c];
}

// A starting point of the interior-point method from network outputs O
// (rows x n_out, n_out = n + 2 m + 2; one workgroup per row): raw = O ts + tm
// in FP64 (product and sum rounded separately), then x = raw[0:n],
// y = exp(clip(raw[n:n+m], -700, 700)) and s = exp(clip(raw[n+m:n+2m])) on
// the inequality rows, y = raw and s = 0 on the z equality rows, and
// tau = exp(clip(raw[n+2m])); raw[n+2m+1] is log kappa.  A row with any
// non-finite raw value gets the default start (x = 0, y = s = 1 on the
// inequality rows, 0 on the equality rows, tau = 1), and valid[r] = 0.
__global__ __launch_bounds__(WG) void k_decode_start(const float* __restrict__ O, int n_out,
                                                     const double* __restrict__ tm,
                                                     const double* __restrict__ ts, int n, int m,
                                                     int z, double* __restrict__ x,
                                                     double* __restrict__ y,
                                                     double* __restrict__ s,
                                                     double* __restrict__ tau,
                                                     int* __restrict__ valid) {
  const int64_t r = blockIdx.x;
  const float* o = O + r * n_out;
  auto raw = [&](int k) { return __dadd_rn(__dmul_rn((double)o[k], ts[k]), tm[k]); };
  int fin = 1;
  for (int k = threadIdx.x; k < n_out; k += WG) fin &= isfinite(raw(k)) ? 1 : 0;
  const int ok = __syncthreads_and(fin);
  auto ex = [](double v) { return exp(fmin(fmax(v, -700.0), 700.0)); };
  double* xr = x + r * n;
  double* yr = y + r * m;
  double* sr = s + r * m;
  for (int j = threadIdx.x; j < n; j += WG) xr[j] = ok ? raw(j) : 0.0;
  for (int i = threadIdx.x; i < m; i += WG) {
    if (i < z) {
      yr[i] = ok ? raw(n + i) : 0.0;
      sr[i] = 0.0;
    } else {
      yr[i] = ok ? ex(raw(n + i)) : 1.0;
      sr[i] = ok ? ex(raw(n + m + i)) : 1.0;
    }
  }
  if (threadIdx.x == 0) {
    tau[r] = ok ? ex(raw(n + 2 * m)) : 1.0;
    valid[r] = ok;
  }
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + WG - 1) / WG); }

}  // namespace

struct qtqp_mlp {
  int n_in = 0, hid = 0, n_out = 0, max_batch = 0;
  hipStream_t stream = nullptr;
  int64_t np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float *X = nullptr, *T = nullptr;                         // current batch
  float *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *O = nullptr;
  float *dO = nullptr, *dH = nullptr, *dZ = nullptr;
  double* partial = nullptr;
  float* dloss = nullptr;
  double *dtm = nullptr, *dts = nullptr;  // output standardisation of qtqp_mlp_propose
  int* dvalid = nullptr;                  // max_batch flags of qtqp_mlp_propose
  int64_t step = 0;
  // offsets into p/g
  int64_t oW1, ob1, oW2, ob2, oW3, ob3;
};

namespace {

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPC(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void gemm(qtqp_mlp* h, const float* A, int64_t sai, int64_t sak, const float* B, int64_t sbk,
          int64_t sbj, int M, int N, int K, const float* bias, float* Z, float* H, int act) {
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  hipLaunchKernelGGL(k_gemm, grid, dim3(WG), 0, h->stream, A, sai, sak, B, sbk, sbj, M, N, K,
                     bias, Z, H, (int64_t)N, act);
}

// Forward on the batch in h->X (b rows).
void forward(qtqp_mlp* h, int b) {
  const float* p = h->p;
  gemm(h, h->X, h->n_in, 1, p + h->oW1, h->hid, 1, b, h->hid, h->n_in, p + h->ob1, h->Z1, h->H1, 1);
  gemm(h, h->H1, h->hid, 1, p + h->oW2, h->hid, 1, b, h->hid, h->hid, p + h->ob2, h->Z2, h->H2, 1);
  gemm(h, h->H2, h->hid, 1, p + h->oW3, h->n_out, 1, b, h->n_out, h->hid, p + h->ob3, h->O, nullptr, 0);
}

// Loss and gradient for the batch in h->X / h->T (b rows) into h->g, h->dloss.
void backward(qtqp_mlp* h, int b) {
  const float* p = h->p;
  float* g = h->g;
  const int64_t no = (int64_t)b * h->n_out;
  const int nparts = (int)nblk(no);
  hipLaunchKernelGGL(k_mse, dim3(nparts), dim3(WG), 0, h->stream, h->O, h->T, h->dO, no, h->partial);
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(WG), 0, h->stream, h->partial, nparts, no, h->dloss);
  // layer 3: dW3 = H2' dO, db3 = sum dO, dH2 = dO W3'
  gemm(h, h->H2, 1, h->hid, h->dO, h->n_out, 1, h->hid, h->n_out, b, nullptr, g + h->oW3, nullptr, 0);
  hipLaunchKernelGGL(k_colsum, dim3(nblk(h->n_out)), dim3(WG), 0, h->stream, h->dO, b, h->n_out, g + h->ob3);
  gemm(h, h->dO, h->n_out, 1, p + h->oW3, 1, h->n_out, b, h->hid, h->n_out, nullptr, h->dH, nullptr, 0);
  const int64_t nh = (int64_t)b * h->hid;
  hipLaunchKernelGGL(k_gelu_back, dim3(nblk(nh)), dim3(WG), 0, h->stream, h->dH, h->Z2, h->dZ, nh);
  // layer 2
  gemm(h, h->H1, 1, h->hid, h->dZ, h->hid, 1, h->hid, h->hid, b, nullptr, g + h->oW2, nullptr, 0);
  hipLaunchKernelGGL(k_colsum, dim3(nblk(h->hid)), dim3(WG), 0, h->stream, h->dZ, b, h->hid, g + h->ob2);
  gemm(h, h->dZ, h->hid, 1, p + h->oW2, 1, h->hid, b, h->hid, h->hid, nullptr, h->dH, nullptr, 0);
  hipLaunchKernelGGL(k_gelu_back, dim3(nblk(nh)), dim3(WG), 0, h->stream, h->dH, h->Z1, h->dZ, nh);
  // layer 1
  gemm(h, h->X, 1, h->n_in, h->dZ, h->hid, 1, h->n_in, h->hid, b, nullptr, g + h->oW1, nullptr, 0);
  hipLaunchKernelGGL(k_colsum, dim3(nblk(h->hid)), dim3(WG), 0, h->stream, h->dZ, b, h->hid, g + h->ob1);
}

void release(qtqp_mlp* h) {
  void* bufs[] = {h->p, h->g, h->m, h->v, h->X, h->T, h->Z1, h->H1, h->Z2, h->H2, h->O, h->dO,
                  h->dH, h->dZ, h->partial, h->dloss, h->dtm, h->dts, h->dvalid};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

}  // namespace

extern "C" {

const char* qtqp_mlp_last_error(void) { return g_err.c_str(); }

int qtqp_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, int32_t max_batch,
                    qtqp_mlp** out) {
  if (!out) return fail(1, "out is NULL");
  *out = nullptr;
  if (n_in < 1 || hidden < 1 || n_out < 1 || max_batch < 1) return fail(1, "sizes must be >= 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(2, "no HIP device available");
  auto* h = new qtqp_mlp();
  h->n_in = n_in;
  h->hid = hidden;
  h->n_out = n_out;
  h->max_batch = max_batch;
  h->oW1 = 0;
  h->ob1 = h->oW1 + (int64_t)n_in * hidden;
  h->oW2 = h->ob1 + hidden;
  h->ob2 = h->oW2 + (int64_t)hidden * hidden;
  h->oW3 = h->ob2 + hidden;
  h->ob3 = h->oW3 + (int64_t)hidden * n_out;
  h->np = h->ob3 + n_out;
  const int64_t b = max_batch;
  hipError_t e = hipStreamCreate(&h->stream);
  auto A = [&](float** ptr, int64_t count) {
    if (e == hipSuccess) e = hipMalloc(ptr, sizeof(float) * std::max<int64_t>(count, 1));
  };
  A(&h->p, h->np);
  A(&h->g, h->np);
  A(&h->m, h->np);
  A(&h->v, h->np);
  A(&h->X, b * n_in);
  A(&h->T, b * n_out);
  A(&h->Z1, b * hidden);
  A(&h->H1, b * hidden);
  A(&h->Z2, b * hidden);
  A(&h->H2, b * hidden);
  A(&h->O, b * n_out);
  A(&h->dO, b * n_out);
  A(&h->dH, b * hidden);
  A(&h->dZ, b * hidden);
  if (e == hipSuccess) e = hipMalloc(&h->partial, sizeof(double) * (b * std::max(n_out, hidden) / WG + 2));
  A(&h->dloss, 1);
  if (e == hipSuccess) e = hipMalloc(&h->dtm, sizeof(double) * n_out);
  if (e == hipSuccess) e = hipMalloc(&h->dts, sizeof(double) * n_out);
  if (e == hipSuccess) e = hipMalloc(&h->dvalid, sizeof(int) * b);
  if (e == hipSuccess) e = hipMemset(h->m, 0, sizeof(float) * h->np);
  if (e == hipSuccess) e = hipMemset(h->v, 0, sizeof(float) * h->np);
  if (e != hipSuccess) {
    std::string msg = std::string("allocation failed: ") + hipGetErrorString(e);
    release(h);
    delete h;
    return fail(2, msg);
  }
  *out = h;
  return 0;
}

int64_t qtqp_mlp_num_params(qtqp_mlp* h) { return h ? h->np : -1; }

int qtqp_mlp_set_params(qtqp_mlp* h, const float* params) {
  if (!h || !params) return fail(1, "NULL argument");
  HIPC(hipMemcpy(h->p, params, sizeof(float) * h->np, hipMemcpyHostToDevice));
  HIPC(hipMemset(h->m, 0, sizeof(float) * h->np));
  HIPC(hipMemset(h->v, 0, sizeof(float) * h->np));
  h->step = 0;
  return 0;
}

int qtqp_mlp_get_params(qtqp_mlp* h, float* params) {
  if (!h || !params) return fail(1, "NULL argument");
  HIPC(hipMemcpy(params, h->p, sizeof(float) * h->np, hipMemcpyDeviceToHost));
  return 0;
}

int qtqp_mlp_forward(qtqp_mlp* h, const float* x, int64_t batch, float* y) {
  if (!h || !x || !y || batch < 0) return fail(1, "bad argument");
  for (int64_t r0 = 0; r0 < batch; r0 += h->max_batch) {
    const int b = (int)std::min<int64_t>(h->max_batch, batch - r0);
    HIPC(hipMemcpyAsync(h->X, x + r0 * h->n_in, sizeof(float) * b * h->n_in, hipMemcpyHostToDevice,
                        h->stream));
    forward(h, b);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(y + r0 * h->n_out, h->O, sizeof(float) * b * h->n_out,
                        hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
  }
  return 0;
}

int qtqp_mlp_propose(qtqp_mlp* h, const float* x, int64_t batch, const double* tm,
                     const double* ts, int32_t n, int32_t m, int32_t z, double* x_start,
                     double* y_start, double* s_start, double* tau_start, int32_t* valid) {
  if (!h || !x || !tm || !ts || !x_start || !y_start || !s_start || !tau_start || batch < 0)
    return fail(1, "bad argument");
  if (n < 1 || m < 1 || z < 0 || z > m || (int64_t)n + 2 * (int64_t)m + 2 != h->n_out)
    return fail(1, "n_out must be n + 2 m + 2 with 0 <= z <= m");
  HIPC(hipMemcpyAsync(h->dtm, tm, sizeof(double) * h->n_out, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->dts, ts, sizeof(double) * h->n_out, hipMemcpyHostToDevice, h->stream));
  for (int64_t r0 = 0; r0 < batch; r0 += h->max_batch) {
    const int b = (int)std::min<int64_t>(h->max_batch, batch - r0);
    HIPC(hipMemcpyAsync(h->X, x + r0 * h->n_in, sizeof(float) * b * h->n_in, hipMemcpyHostToDevice,
                        h->stream));
    forward(h, b);
    hipLaunchKernelGGL(k_decode_start, dim3(b), dim3(WG), 0, h->stream, h->O, h->n_out, h->dtm,
                       h->dts, n, m, z, x_start + r0 * n, y_start + r0 * m, s_start + r0 * m,
                       tau_start + r0, h->dvalid);
    HIPC(hipGetLastError());
    if (valid)
      HIPC(hipMemcpyAsync(valid + r0, h->dvalid, sizeof(int) * b, hipMemcpyDeviceToHost,
                          h->stream));
  }
  HIPC(hipStreamSynchronize(h->stream));
  return 0;
}

int qtqp_mlp_loss_grad(qtqp_mlp* h, const float* x, const float* t, int32_t batch, float* grad,
                       double* loss) {
  if (!h || !x || !t || batch < 1 || batch > h->max_batch) return fail(1, "bad argument");
  HIPC(hipMemcpyAsync(h->X, x, sizeof(float) * batch * h->n_in, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->T, t, sizeof(float) * batch * h->n_out, hipMemcpyHostToDevice, h->stream));
  forward(h, batch);
  backward(h, batch);
  HIPC(hipGetLastError());
  float l = 0.f;
  if (grad) HIPC(hipMemcpyAsync(grad, h->g, sizeof(float) * h->np, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipMemcpyAsync(&l, h->dloss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  if (loss) *loss = l;
  return 0;
}

int qtqp_mlp_train(qtqp_mlp* h, const float* x, const float* t, int64_t count, int32_t batch,
                   int32_t steps, float lr, float beta1, float beta2, float eps, uint64_t seed,
                   float* loss_per_step) {
  if (!h || !x || !t || count < 1 || batch < 1 || batch > h->max_batch || steps < 0)
    return fail(1, "bad argument");
  float *dX = nullptr, *dT = nullptr, *dl = nullptr;
  int64_t* didx = nullptr;
  const int64_t nsteps = steps;
  HIPC(hipMalloc(&dX, sizeof(float) * count * h->n_in));
  HIPC(hipMalloc(&dT, sizeof(float) * count * h->n_out));
  HIPC(hipMalloc(&dl, sizeof(float) * std::max<int64_t>(nsteps, 1)));
  HIPC(hipMalloc(&didx, sizeof(int64_t) * std::max<int64_t>(nsteps * batch, 1)));
  HIPC(hipMemcpy(dX, x, sizeof(float) * count * h->n_in, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dT, t, sizeof(float) * count * h->n_out, hipMemcpyHostToDevice));
  // mini-batch indices: a fresh permutation per epoch
  std::vector<int64_t> idx((size_t)(nsteps * batch));
  std::vector<int64_t> perm((size_t)count);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng(seed);
  int64_t pos = count;
  for (int64_t k = 0; k < nsteps * batch; ++k) {
    if (pos >= count) {
      std::shuffle(perm.begin(), perm.end(), rng);
      pos = 0;
    }
    idx[(size_t)k] = perm[(size_t)pos++];
  }
  HIPC(hipMemcpy(didx, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice));
  for (int64_t s = 0; s < nsteps; ++s) {
    hipLaunchKernelGGL(k_gather_rows, dim3(batch), dim3(WG), 0, h->stream, dX, dT,
                       didx + s * batch, batch, h->n_in, h->n_out, h->X, h->T);
    forward(h, batch);
    backward(h, batch);
    ++h->step;
    const float c1 = 1.f - std::pow(beta1, (float)h->step);
    const float c2 = 1.f - std::pow(beta2, (float)h->step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(h->np)), dim3(WG), 0, h->stream, h->p, h->g, h->m, h->v,
                       h->np, lr, beta1, beta2, eps, c1, c2);
    HIPC(hipMemcpyAsync(dl + s, h->dloss, sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPC(hipGetLastError());
  }
  HIPC(hipStreamSynchronize(h->stream));
  if (loss_per_step && nsteps > 0)
    HIPC(hipMemcpy(loss_per_step, dl, sizeof(float) * nsteps, hipMemcpyDeviceToHost));
  (void)hipFree(dX);
  (void)hipFree(dT);
  (void)hipFree(dl);
  (void)hipFree(didx);
  return 0;
}

void qtqp_mlp_free(qtqp_mlp* h) {
  if (!h) return;
  release(h);
  delete h;
}

}  // extern "C"
