// BF16x3 split-precision complex GEMM on the gfx950 bf16 matrix cores.
//
// C (M x N, complex FP64 out) = op(A) B with op(A) = A (M x K) or A^H (A is
// K x M), B (K x N), all column-major complex FP64.  The product runs as
// one real GEMM on the real embedding
//
//   [C_re; C_im] = [[Re opA, -Im opA], [Im opA, Re opA]] [Re B; Im B]
//
// (2M x 2K times 2K x N), whose operands are split into three bf16 terms,
// x = x0 + x1 + x2 with x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 -
// x1) (24 significand bits, the FP32 input's), and whose product keeps the
// six terms a_i b_j with i + j <= 2 (the dropped ones are below 2^-24
// relative), accumulated in FP32 by v_mfma_f32_16x16x32_bf16.  The result
// carries FP32-level error (about 1e-7 relative to |opA| |B| per entry,
// growing like sqrt(K)); it replaces FP64 only where that is the stated
// precision (the large contractions of the MPO expectation sweep).
#include <algorithm>
#include <cstdint>

#include "qtn_internal.hpp"

namespace qtn {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kLdsK = kBK + 8;  // padded row (bf16 elements) of a staged tile

// Round-to-nearest-even float -> bf16 bits (finite inputs; NaN stays NaN).
__device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_float(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}

// x -> three bf16 terms.
__device__ __forceinline__ void split3(double x, uint16_t& b0, uint16_t& b1,
                                       uint16_t& b2) {
  const float f = (float)x;
  b0 = bf16_bits(f);
  const float r1 = f - bf16_float(b0);
  b1 = bf16_bits(r1);
  const float r2 = r1 - bf16_float(b1);
  b2 = bf16_bits(r2);
}

// Embedded A: row r of the 2M x 2K real matrix, k-contiguous, three planes
// of Mp x Kp bf16 (plane stride Mp * Kp), zero in the padding.
//   r < M:  [Re opA[r, :], -Im opA[r, :]]
//   r >= M: [Im opA[r-M, :], Re opA[r-M, :]]
// opA[r, k] = A[r + k lda] (opA = A) or conj(A[k + r lda]) (opA = A^H).
__global__ void split_a(const double2* __restrict__ A, int64_t lda, int conjT,
                        int64_t M, int64_t K, int64_t Mp, int64_t Kp,
                        uint16_t* __restrict__ out) {
  const int64_t total = Mp * Kp;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / Kp, k2 = i % Kp;
    double v = 0.0;
    if (r < 2 * M && k2 < 2 * K) {
      const int64_t rr = r < M ? r : r - M;
      const int64_t k = k2 < K ? k2 : k2 - K;
      const double2 a = conjT ? A[k + rr * lda] : A[rr + k * lda];
      const double re = a.x, im = conjT ? -a.y : a.y;
      if (r < M)
        v = k2 < K ? re : -im;
      else
        v = k2 < K ? im : re;
    }
    uint16_t b0, b1, b2;
    split3(v, b0, b1, b2);
    out[i] = b0;
    out[i + total] = b1;
    out[i + 2 * total] = b2;
  }
}

// Embedded B: column n = [Re B[:, n]; Im B[:, n]], k-contiguous, three
// planes of Np x Kp bf16, zero in the padding.
__global__ void split_b(const double2* __restrict__ B, int64_t ldb, int64_t K,
                        int64_t N, int64_t Np, int64_t Kp,
                        uint16_t* __restrict__ out) {
  const int64_t total = Np * Kp;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = i / Kp, k2 = i % Kp;
    double v = 0.0;
    if (n < N && k2 < 2 * K) {
      const double2 b = B[(k2 < K ? k2 : k2 - K) + n * ldb];
      v = k2 < K ? b.x : b.y;
    }
    uint16_t b0, b1, b2;
    split3(v, b0, b1, b2);
    out[i] = b0;
    out[i + total] = b1;
    out[i + 2 * total] = b2;
  }
}

// One 128 x 128 tile of the embedded product per workgroup (4 waves, each
// 64 x 64 = 4 x 4 MFMA tiles of 16 x 16), K in steps of 32 staged through
// LDS (three planes of A and of B, rows padded to 40 bf16).  MFMA operand
// layout (16x16x32): lane l holds row (A) / column (B) l % 16 and the 8
// consecutive k of group l / 16; the accumulator holds rows 4 (l / 16) + e,
// column l % 16.  Output: rows r < M go to C[r, n].x, rows M <= r < 2M to
// C[r - M, n].y.
__global__ __launch_bounds__(256) void gemm_bf16x3(
    const uint16_t* __restrict__ Ap, const uint16_t* __restrict__ Bp,
    int64_t M, int64_t N, int64_t Mp, int64_t Np, int64_t Kp,
    double2* __restrict__ C, int64_t ldc) {
  __shared__ __attribute__((aligned(16))) uint16_t As[3][kBM][kLdsK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[3][kBN][kLdsK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * kBM, n0 = (int64_t)blockIdx.y * kBN;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int64_t planeA = Mp * Kp, planeB = Np * Kp;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int64_t k0 = 0; k0 < Kp; k0 += kBK) {
    // stage: per plane 128 rows x 32 k = 512 chunks of 8 bf16 (16 B)
#pragma unroll
    for (int pl = 0; pl < 3; ++pl)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int c = tid + 256 * u;
        const int row = c >> 2, kq = (c & 3) * 8;
        const uint4 va = *(const uint4*)(Ap + pl * planeA +
                                         (m0 + row) * Kp + k0 + kq);
        const uint4 vb = *(const uint4*)(Bp + pl * planeB +
                                         (n0 + row) * Kp + k0 + kq);
        *(uint4*)&As[pl][row][kq] = va;
        *(uint4*)&Bs[pl][row][kq] = vb;
      }
    __syncthreads();
    const int kg = (lane >> 4) * 8;
    bf16x8 a[3][4], b[3][4];
#pragma unroll
    for (int pl = 0; pl < 3; ++pl)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[pl][i] = *(const bf16x8*)&As[pl][wm + i * 16 + (lane & 15)][kg];
        b[pl][i] = *(const bf16x8*)&Bs[pl][wn + i * 16 + (lane & 15)][kg];
      }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // smallest terms first
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a[2][i], b[0][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a[1][i], b[1][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a[0][i], b[2][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a[1][i], b[0][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a[0][i], b[1][j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a[0][i], b[0][j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t n = n0 + wn + j * 16 + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = m0 + wm + i * 16 + 4 * (lane >> 4) + e;
        if (r < M)
          reinterpret_cast<double*>(C + r + n * ldc)[0] = (double)acc[i][j][e];
        else if (r < 2 * M)
          reinterpret_cast<double*>(C + (r - M) + n * ldc)[1] =
              (double)acc[i][j][e];
      }
    }
}

constexpr int kSlotSplitA = 190, kSlotSplitB = 191;

unsigned fill_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                          16384));
}

}  // namespace

int32_t launch_zgemm_bf16x3(qtn_ctx* ctx, int opA, int64_t m, int64_t n,
                            int64_t k, const double2* A, int64_t lda,
                            const double2* B, int64_t ldb, double2* C,
                            int64_t ldc) {
  if (opA != 0 && opA != 2) {
    set_error("bf16x3 gemm: op(A) must be N or C");
    return QTN_ERR_ARG;
  }
  if (m <= 0 || n <= 0 || k <= 0) return QTN_OK;
  const int64_t Mp = ((2 * m + kBM - 1) / kBM) * kBM;
  const int64_t Np = ((n + kBN - 1) / kBN) * kBN;
  const int64_t Kp = ((2 * k + kBK - 1) / kBK) * kBK;
  if (Mp / kBM > 65535 || Np / kBN > 65535) {
    set_error("bf16x3 gemm: dimension too large for the launch grid");
    return QTN_ERR_ARG;
  }
  void *pa, *pb;
  QTN_TRY(ctx->buffer(kSlotSplitA, (size_t)3 * Mp * Kp * sizeof(uint16_t),
                      &pa));
  QTN_TRY(ctx->buffer(kSlotSplitB, (size_t)3 * Np * Kp * sizeof(uint16_t),
                      &pb));
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(split_a, dim3(fill_grid(Mp * Kp)), dim3(256), 0, s, A,
                     lda, opA == 2 ? 1 : 0, m, k, Mp, Kp, (uint16_t*)pa);
  hipLaunchKernelGGL(split_b, dim3(fill_grid(Np * Kp)), dim3(256), 0, s, B,
                     ldb, k, n, Np, Kp, (uint16_t*)pb);
  hipLaunchKernelGGL(gemm_bf16x3, dim3((unsigned)(Mp / kBM),
                                       (unsigned)(Np / kBN)),
                     dim3(256), 0, s, (const uint16_t*)pa,
                     (const uint16_t*)pb, m, n, Mp, Np, Kp, C, ldc);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

}  // namespace qtn

using namespace qtn;

extern "C" {

int32_t qtn_zgemm_bf16x3(qtn_ctx* ctx, char opA, int64_t m, int64_t n,
                         int64_t k, const double* A, int64_t lda,
                         const double* B, int64_t ldb, double* C,
                         int64_t ldc) try {
  clear_error();
  QTN_ARG(ctx && A && B && C, "qtn_zgemm_bf16x3: NULL argument");
  QTN_ARG(m >= 0 && n >= 0 && k >= 0, "qtn_zgemm_bf16x3: negative size");
  const int op = (opA == 'N' || opA == 'n') ? 0
                 : (opA == 'C' || opA == 'c') ? 2
                                              : -1;
  QTN_ARG(op >= 0, "qtn_zgemm_bf16x3: opA must be 'N' or 'C'");
  QTN_ARG(lda >= std::max<int64_t>(1, op == 0 ? m : k) &&
              ldb >= std::max<int64_t>(1, k) && ldc >= std::max<int64_t>(1, m),
          "qtn_zgemm_bf16x3: leading dimension too small");
  if (m == 0 || n == 0) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // The host arrays are read and written only inside their rows x cols
  // extents (a BLAS caller need not allocate past ld*(cols-1)+rows): they
  // are staged by pitched copies into packed device buffers, and only the
  // m rows of each column of C are copied back.
  const int64_t rowsA = op == 0 ? m : k, colsA = op == 0 ? k : m;
  void *dA, *dB, *dC;
  QTN_TRY(ctx->buffer(192, (size_t)rowsA * colsA * sizeof(double2), &dA));
  QTN_TRY(ctx->buffer(193, (size_t)k * n * sizeof(double2), &dB));
  QTN_TRY(ctx->buffer(194, (size_t)m * n * sizeof(double2), &dC));
  if (k == 0) {
    QTN_CHECK_HIP(hipMemsetAsync(dC, 0, (size_t)m * n * sizeof(double2), s));
  } else {
    QTN_CHECK_HIP(hipMemcpy2DAsync(
        dA, rowsA * sizeof(double2), A, lda * sizeof(double2),
        rowsA * sizeof(double2), colsA, hipMemcpyHostToDevice, s));
    QTN_CHECK_HIP(hipMemcpy2DAsync(dB, k * sizeof(double2), B,
                                   ldb * sizeof(double2), k * sizeof(double2),
                                   n, hipMemcpyHostToDevice, s));
    QTN_TRY(launch_zgemm_bf16x3(ctx, op, m, n, k, (const double2*)dA, rowsA,
                                (const double2*)dB, k, (double2*)dC, m));
  }
  QTN_CHECK_HIP(hipMemcpy2DAsync(C, ldc * sizeof(double2), dC,
                                 m * sizeof(double2), m * sizeof(double2), n,
                                 hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
