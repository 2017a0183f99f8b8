// Rank-predictor MLP on the GPU: forward, backward and Adam, complex-free
// FP32 (the learned part of the chosen form; no counterpart in the
// reference).
//
// Network: y = W3 gelu(W2 gelu(W1 a0 + b1) + b2) + b3 with a0 the input
// column (nin features; the first min(64, nin), the previous bond spectrum,
// enter as log10(max(x, 1e-16)) / 8, i.e. in [-2, 0]).  Activations are column-major (features x
// batch).  Parameters are one flat float array: W1 (width x nin), b1, W2
// (width x width), b2, W3 (1 x width), b3, each block column-major.
// GELU is the exact form 0.5 z (1 + erf(z / sqrt 2)).
#include <algorithm>
#include <cmath>
#include <vector>

#include "qtn_internal.hpp"

struct qtn_mlp {
  qtn_ctx* ctx = nullptr;
  int nin = 0, width = 0;
  int64_t np = 0;          // number of parameters
  float* params = nullptr; // device
  float* grads = nullptr;
  float* m1 = nullptr;     // Adam moments
  float* m2 = nullptr;
  int64_t step = 0;
  void* arena = nullptr;
};

namespace qtn {
namespace {

constexpr int kSlotA0 = 100, kSlotZ1 = 101, kSlotA1 = 102, kSlotZ2 = 103,
              kSlotA2 = 104, kSlotY = 105, kSlotD1 = 106, kSlotD2 = 107,
              kSlotIn = 108, kSlotT = 109, kSlotLoss = 110;
constexpr int kNLog = 64;

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                          8192));
}

// C (M x N) = op(A) op(B) [+ bias broadcast over columns] [+ C], FP32,
// column-major.  op(A) is A (M x K, lda) or A^T (A is K x M); op(B) is B
// (K x N) or B^T.  32 x 32 output tile per 256-thread block, 4 per thread,
// K staged through LDS in slices of 16.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void sgemm(int64_t M, int64_t N, int64_t K,
                                             const float* __restrict__ A,
                                             int64_t lda,
                                             const float* __restrict__ B,
                                             int64_t ldb, float* __restrict__ C,
                                             int64_t ldc,
                                             const float* __restrict__ bias,
                                             int accumulate) {
  __shared__ float As[16][33];  // [k][m]
  __shared__ float Bs[16][33];  // [k][n]
  const int tid = threadIdx.x;
  const int64_t m0 = (int64_t)blockIdx.x * 32, n0 = (int64_t)blockIdx.y * 32;
  const int tm = tid % 32, tn = tid / 32;  // tn 0..7, each 4 columns
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    for (int idx = tid; idx < 16 * 32; idx += 256) {
      const int kk = idx / 32, i = idx % 32;
      const int64_t gm = m0 + i, gk = k0 + kk;
      float a = 0.f;
      if (gm < M && gk < K) a = TA ? A[gk + gm * lda] : A[gm + gk * lda];
      As[kk][i] = a;
      const int64_t gn = n0 + i;
      float b = 0.f;
      if (gn < N && gk < K) b = TB ? B[gn + gk * ldb] : B[gk + gn * ldb];
      Bs[kk][i] = b;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const float a = As[kk][tm];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += a * Bs[kk][tn * 4 + j];
    }
    __syncthreads();
  }
  const int64_t gm = m0 + tm;
  if (gm >= M) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t gn = n0 + tn * 4 + j;
    if (gn >= N) continue;
    float v = acc[j];
    if (bias) v += bias[gm];
    if (accumulate) v += C[gm + gn * ldc];
    C[gm + gn * ldc] = v;
  }
}

template <bool TA, bool TB>
int32_t launch_sgemm(hipStream_t s, int64_t M, int64_t N, int64_t K,
                     const float* A, int64_t lda, const float* B, int64_t ldb,
                     float* C, int64_t ldc, const float* bias,
                     int accumulate) {
  dim3 grid((unsigned)((M + 31) / 32), (unsigned)((N + 31) / 32));
  hipLaunchKernelGGL((sgemm<TA, TB>), grid, dim3(256), 0, s, M, N, K, A, lda,
                     B, ldb, C, ldc, bias, accumulate);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

__global__ void prep_input(const float* __restrict__ X, int nin, int64_t nb,
                           float* __restrict__ A0) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       i < (int64_t)nin * nb; i += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(i % nin);
    const float x = X[i];
    A0[i] = f < kNLog ? 0.125f * log10f(fmaxf(x, 1e-16f)) : x;
  }
}

__device__ __forceinline__ float gelu_f(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__global__ void gelu_fwd(const float* __restrict__ Z, float* __restrict__ A,
                         int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    A[i] = gelu_f(Z[i]);
}

// D <- D * gelu'(Z)
__global__ void gelu_bwd(const float* __restrict__ Z, float* __restrict__ D,
                         int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    D[i] *= gelu_d(Z[i]);
}

// dY = 2 (Y - T) / nb; loss[0] = sum (Y - T)^2 / nb (one workgroup).
__global__ __launch_bounds__(256) void mse_grad(const float* __restrict__ Y,
                                                const float* __restrict__ T,
                                                int64_t nb,
                                                float* __restrict__ dY,
                                                float* __restrict__ loss) {
  float s = 0.f;
  const float inv = 1.0f / (float)nb;
  for (int64_t i = threadIdx.x; i < nb; i += 256) {
    const float r = Y[i] - T[i];
    dY[i] = 2.0f * r * inv;
    s += r * r;
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] * inv;
}

// db[r] = sum_b D[r + rows * b] (one wave per row).  grid ceil(rows/4).
__global__ __launch_bounds__(256) void row_sum(const float* __restrict__ D,
                                               int64_t rows, int64_t nb,
                                               float* __restrict__ db) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  float s = 0.f;
  for (int64_t b = lane; b < nb; b += 64) s += D[r + rows * b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) db[r] = s;
}

__global__ void adam_step(float* __restrict__ p, const float* __restrict__ g,
                          float* __restrict__ m1, float* __restrict__ m2,
                          int64_t n, float lr, float b1, float b2, float eps,
                          float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
  }
}

struct Offsets {
  int64_t W1, b1, W2, b2, W3, b3;
};

Offsets offsets(int nin, int w) {
  Offsets o;
  o.W1 = 0;
  o.b1 = o.W1 + (int64_t)w * nin;
  o.W2 = o.b1 + w;
  o.b2 = o.W2 + (int64_t)w * w;
  o.W3 = o.b2 + w;
  o.b3 = o.W3 + w;
  return o;
}

struct Acts {
  float *A0, *Z1, *A1, *Z2, *A2, *Y;
};

// Forward pass on device inputs X (nin x nb), activations kept for the
// backward pass.
int32_t forward(qtn_mlp* m, int64_t nb, const float* X, Acts& a) {
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  const int w = m->width, nin = m->nin;
  void* p;
  QTN_TRY(ctx->buffer(kSlotA0, (size_t)nin * nb * sizeof(float), &p));
  a.A0 = (float*)p;
  QTN_TRY(ctx->buffer(kSlotZ1, (size_t)w * nb * sizeof(float), &p));
  a.Z1 = (float*)p;
  QTN_TRY(ctx->buffer(kSlotA1, (size_t)w * nb * sizeof(float), &p));
  a.A1 = (float*)p;
  QTN_TRY(ctx->buffer(kSlotZ2, (size_t)w * nb * sizeof(float), &p));
  a.Z2 = (float*)p;
  QTN_TRY(ctx->buffer(kSlotA2, (size_t)w * nb * sizeof(float), &p));
  a.A2 = (float*)p;
  QTN_TRY(ctx->buffer(kSlotY, (size_t)nb * sizeof(float), &p));
  a.Y = (float*)p;
  const Offsets o = offsets(nin, w);
  const float* P = m->params;
  hipLaunchKernelGGL(prep_input, dim3(grid_for((int64_t)nin * nb)),
                     dim3(256), 0, s, X, nin, nb, a.A0);
  QTN_CHECK_HIP(hipGetLastError());
  QTN_TRY((launch_sgemm<false, false>(s, w, nb, nin, P + o.W1, w, a.A0, nin,
                                      a.Z1, w, P + o.b1, 0)));
  hipLaunchKernelGGL(gelu_fwd, dim3(grid_for((int64_t)w * nb)), dim3(256), 0,
                     s, a.Z1, a.A1, (int64_t)w * nb);
  QTN_TRY((launch_sgemm<false, false>(s, w, nb, w, P + o.W2, w, a.A1, w,
                                      a.Z2, w, P + o.b2, 0)));
  hipLaunchKernelGGL(gelu_fwd, dim3(grid_for((int64_t)w * nb)), dim3(256), 0,
                     s, a.Z2, a.A2, (int64_t)w * nb);
  QTN_TRY((launch_sgemm<false, false>(s, 1, nb, w, P + o.W3, 1, a.A2, w, a.Y,
                                      1, P + o.b3, 0)));
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

// Loss and gradient into m->grads for device inputs X and targets T.
int32_t loss_grad(qtn_mlp* m, int64_t nb, const float* X, const float* T,
                  float* loss_dev) {
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  const int w = m->width, nin = m->nin;
  Acts a;
  QTN_TRY(forward(m, nb, X, a));
  const Offsets o = offsets(nin, w);
  const float* P = m->params;
  float* G = m->grads;
  void *pd1, *pd2, *pdy;
  QTN_TRY(ctx->buffer(kSlotD1, (size_t)w * nb * sizeof(float), &pd1));
  QTN_TRY(ctx->buffer(kSlotD2, (size_t)w * nb * sizeof(float), &pd2));
  QTN_TRY(ctx->buffer(kSlotT, (size_t)nb * sizeof(float) * 2, &pdy));
  float* dZ1 = (float*)pd1;
  float* dZ2 = (float*)pd2;
  float* dY = (float*)pdy + nb;
  hipLaunchKernelGGL(mse_grad, dim3(1), dim3(256), 0, s, a.Y, T, nb, dY,
                     loss_dev);
  // dW3 = dY A2^T (1 x w), db3 = sum dY
  QTN_TRY((launch_sgemm<false, true>(s, 1, w, nb, dY, 1, a.A2, w, G + o.W3,
                                     1, nullptr, 0)));
  hipLaunchKernelGGL(row_sum, dim3(1), dim3(256), 0, s, dY, 1, nb, G + o.b3);
  // dZ2 = W3^T dY * gelu'(Z2)
  QTN_TRY((launch_sgemm<true, false>(s, w, nb, 1, P + o.W3, 1, dY, 1, dZ2, w,
                                     nullptr, 0)));
  hipLaunchKernelGGL(gelu_bwd, dim3(grid_for((int64_t)w * nb)), dim3(256), 0,
                     s, a.Z2, dZ2, (int64_t)w * nb);
  // dW2 = dZ2 A1^T, db2
  QTN_TRY((launch_sgemm<false, true>(s, w, w, nb, dZ2, w, a.A1, w, G + o.W2,
                                     w, nullptr, 0)));
  hipLaunchKernelGGL(row_sum, dim3((w + 3) / 4), dim3(256), 0, s, dZ2, w, nb,
                     G + o.b2);
  // dZ1 = W2^T dZ2 * gelu'(Z1)
  QTN_TRY((launch_sgemm<true, false>(s, w, nb, w, P + o.W2, w, dZ2, w, dZ1,
                                     w, nullptr, 0)));
  hipLaunchKernelGGL(gelu_bwd, dim3(grid_for((int64_t)w * nb)), dim3(256), 0,
                     s, a.Z1, dZ1, (int64_t)w * nb);
  // dW1 = dZ1 A0^T, db1
  QTN_TRY((launch_sgemm<false, true>(s, w, nin, nb, dZ1, w, a.A0, nin,
                                     G + o.W1, w, nullptr, 0)));
  hipLaunchKernelGGL(row_sum, dim3((w + 3) / 4), dim3(256), 0, s, dZ1, w, nb,
                     G + o.b1);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

int32_t upload(qtn_ctx* ctx, int slot, const float* host, size_t n,
               float** dev) {
  void* p;
  QTN_TRY(ctx->buffer(slot, std::max<size_t>(n, 1) * sizeof(float), &p));
  QTN_CHECK_HIP(hipMemcpyAsync(p, host, n * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
  *dev = (float*)p;
  return QTN_OK;
}

}  // namespace

// Prediction for device features (used by the learned two-site update):
// Y (nb floats, device) = mlp(X) for X (nin x nb, device).
int32_t mlp_predict_device(qtn_mlp* m, int64_t nb, const float* X,
                           float** Y) {
  Acts a;
  QTN_TRY(forward(m, nb, X, a));
  *Y = a.Y;
  return QTN_OK;
}

int mlp_nin(const qtn_mlp* m) { return m->nin; }
qtn_ctx* mlp_ctx(const qtn_mlp* m) { return m->ctx; }

}  // namespace qtn

using namespace qtn;

extern "C" {

int32_t qtn_mlp_create(qtn_ctx* ctx, int32_t nin, int32_t width,
                       const float* params, qtn_mlp** out) try {
  clear_error();
  QTN_ARG(ctx && params && out, "qtn_mlp_create: NULL argument");
  QTN_ARG(nin >= 1 && nin <= 4096 && width >= 1 && width <= 4096,
          "qtn_mlp_create: bad sizes");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  qtn_mlp* m = new qtn_mlp();
  m->ctx = ctx;
  m->nin = nin;
  m->width = width;
  m->np = (int64_t)width * nin + width + (int64_t)width * width + width +
          width + 1;
  const size_t bytes = (size_t)m->np * 4 * sizeof(float);
  hipError_t e = hipMalloc(&m->arena, bytes);
  if (e != hipSuccess) {
    delete m;
    set_error("qtn_mlp_create: hipMalloc failed");
    return QTN_ERR_OOM;
  }
  m->params = (float*)m->arena;
  m->grads = m->params + m->np;
  m->m1 = m->grads + m->np;
  m->m2 = m->m1 + m->np;
  QTN_CHECK_HIP(hipMemsetAsync(m->arena, 0, bytes, ctx->stream));
  QTN_CHECK_HIP(hipMemcpyAsync(m->params, params, m->np * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = m;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mlp_destroy(qtn_mlp* m) try {
  clear_error();
  if (!m) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  QTN_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  if (m->arena) QTN_CHECK_HIP(hipFree(m->arena));
  delete m;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mlp_get_params(const qtn_mlp* m, float* params) try {
  clear_error();
  QTN_ARG(m && params, "qtn_mlp_get_params: NULL argument");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  QTN_CHECK_HIP(hipMemcpyAsync(params, m->params, m->np * sizeof(float),
                               hipMemcpyDeviceToHost, m->ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mlp_forward(qtn_mlp* m, int64_t nb, const float* X, float* Y) try {
  clear_error();
  QTN_ARG(m && X && Y, "qtn_mlp_forward: NULL argument");
  QTN_ARG(nb >= 1, "qtn_mlp_forward: empty batch");
  qtn_ctx* ctx = m->ctx;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  float* dX;
  QTN_TRY(upload(ctx, kSlotIn, X, (size_t)m->nin * nb, &dX));
  float* dY;
  QTN_TRY(mlp_predict_device(m, nb, dX, &dY));
  QTN_CHECK_HIP(hipMemcpyAsync(Y, dY, nb * sizeof(float),
                               hipMemcpyDeviceToHost, ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mlp_loss_grad(qtn_mlp* m, int64_t nb, const float* X,
                          const float* T, float* loss, float* grad) try {
  clear_error();
  QTN_ARG(m && X && T && loss && grad, "qtn_mlp_loss_grad: NULL argument");
  QTN_ARG(nb >= 1, "qtn_mlp_loss_grad: empty batch");
  qtn_ctx* ctx = m->ctx;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  float *dX, *dT;
  QTN_TRY(upload(ctx, kSlotIn, X, (size_t)m->nin * nb, &dX));
  void* pt;
  QTN_TRY(ctx->buffer(kSlotT, (size_t)nb * sizeof(float) * 2, &pt));
  dT = (float*)pt;
  QTN_CHECK_HIP(hipMemcpyAsync(dT, T, nb * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
  void* pl;
  QTN_TRY(ctx->buffer(kSlotLoss, sizeof(float), &pl));
  QTN_TRY(loss_grad(m, nb, dX, dT, (float*)pl));
  QTN_CHECK_HIP(hipMemcpyAsync(loss, pl, sizeof(float),
                               hipMemcpyDeviceToHost, ctx->stream));
  QTN_CHECK_HIP(hipMemcpyAsync(grad, m->grads, m->np * sizeof(float),
                               hipMemcpyDeviceToHost, ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mlp_train(qtn_mlp* m, int64_t nb, const float* X, const float* T,
                      int32_t steps, int64_t minibatch, float lr,
                      float* loss_hist) try {
  clear_error();
  QTN_ARG(m && X && T, "qtn_mlp_train: NULL argument");
  QTN_ARG(nb >= 1 && steps >= 0, "qtn_mlp_train: bad sizes");
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  const int64_t mb = (minibatch > 0 && minibatch < nb) ? minibatch : nb;
  // the whole training set stays on the device; step t trains on the
  // contiguous slice starting at (t * mb) mod nb (the caller shuffles)
  float *dX, *dT;
  QTN_TRY(upload(ctx, kSlotIn, X, (size_t)m->nin * nb, &dX));
  void* pt;
  QTN_TRY(ctx->buffer(kSlotT + 10, (size_t)nb * sizeof(float), &pt));
  dT = (float*)pt;
  QTN_CHECK_HIP(hipMemcpyAsync(dT, T, nb * sizeof(float),
                               hipMemcpyHostToDevice, s));
  void* pl;
  QTN_TRY(ctx->buffer(kSlotLoss, (size_t)std::max(1, steps) * sizeof(float),
                      &pl));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int t = 0; t < steps; ++t) {
    const int64_t start = ((int64_t)t * mb) % nb;
    const int64_t cnt = std::min(mb, nb - start);
    QTN_TRY(loss_grad(m, cnt, dX + start * m->nin, dT + start,
                      (float*)pl + t));
    m->step += 1;
    const float c1 = 1.f - powf(b1, (float)m->step);
    const float c2 = 1.f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(adam_step, dim3(grid_for(m->np)), dim3(256), 0, s,
                       m->params, (const float*)m->grads, m->m1, m->m2, m->np,
                       lr, b1, b2, eps, c1, c2);
    QTN_CHECK_HIP(hipGetLastError());
  }
  if (loss_hist && steps > 0)
    QTN_CHECK_HIP(hipMemcpyAsync(loss_hist, pl, steps * sizeof(float),
                                 hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
