// Complex GEMM on CDNA4 matrix cores.
//
// C = alpha * op(A) * op(B) + beta * C, column-major, interleaved complex.
// A complex product is four real MFMA products on split re/im planes:
//   Cr += Ar*Br - Ai*Bi,  Ci += Ar*Bi + Ai*Br
// FP64 uses v_mfma_f64_16x16x4f64, FP32 uses v_mfma_f32_16x16x4f32 (FP32
// accumulation).  Both take A[i][k] from lane i + 16k and B[k][j] from lane
// j + 16k; the f32 result element e of lane l is D[4*(l/16) + e][l%16] while
// the f64 one is D[4*e + l/16][l%16] (measured on gfx950).  A 256-thread workgroup (4 wave64) computes a BM x BN tile;
// each wave owns a (TM*16) x (TN*16) sub-tile.  Operands are staged through
// LDS as separate real / imaginary planes laid out [k][m] and [k][n] so the
// MFMA operand fetch (lane l reads row l%16 of k-slice l/16) is a unit-stride
// LDS read.  op = N / T / C is resolved while staging, so transposed and
// conjugated operands cost nothing extra.
#include <algorithm>
#include "qtn_internal.hpp"

namespace qtn {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename R>
struct Traits;
template <>
struct Traits<double> {
  typedef double2 c_t;
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result element e of lane l is row 4*e + l/16, column l%16
  static __device__ __forceinline__ int row(int lane, int e) {
    return 4 * e + (lane >> 4);
  }
};
template <>
struct Traits<float> {
  typedef float2 c_t;
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // result element e of lane l is row 4*(l/16) + e, column l%16
  static __device__ __forceinline__ int row(int lane, int e) {
    return 4 * (lane >> 4) + e;
  }
};

constexpr int kBK = 16;

template <typename R, int OPA, int OPB, int TM, int TN>
__device__ __forceinline__ void gemm_tile(
    int64_t M, int64_t N, int64_t K, typename Traits<R>::c_t alpha,
    const typename Traits<R>::c_t* __restrict__ A, int64_t lda,
    const typename Traits<R>::c_t* __restrict__ B, int64_t ldb,
    typename Traits<R>::c_t beta, typename Traits<R>::c_t* __restrict__ C,
    int64_t ldc, int64_t m0, int64_t n0) {
  using c_t = typename Traits<R>::c_t;
  using acc_t = typename Traits<R>::acc_t;
  constexpr int BM = 2 * TM * 16;
  constexpr int BN = 2 * TN * 16;
  constexpr int BK = kBK;
  __shared__ R As_re[BK][BM + 1];
  __shared__ R As_im[BK][BM + 1];
  __shared__ R Bs_re[BK][BN + 1];
  __shared__ R Bs_im[BK][BN + 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * TM * 16;
  const int wn = (wave >> 1) * TN * 16;

  acc_t acc_re[TM][TN];
  acc_t acc_im[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        acc_re[i][j][e] = R(0);
        acc_im[i][j][e] = R(0);
      }

  const R sgnA = (OPA == 2) ? R(-1) : R(1);
  const R sgnB = (OPB == 2) ? R(-1) : R(1);

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    // ---- stage op(A)[m0:m0+BM, k0:k0+BK] ----
    if (OPA == 0) {
      // element (i,k) at A[i + k*lda]; i fastest -> coalesced
      constexpr int KSTEP = 256 / BM;
      const int i = tid % BM;
#pragma unroll
      for (int kk = tid / BM; kk < BK; kk += KSTEP) {
        const int64_t gi = m0 + i, gk = k0 + kk;
        c_t v;
        if (gi < M && gk < K) {
          v = A[gi + gk * lda];
        } else {
          v.x = R(0);
          v.y = R(0);
        }
        As_re[kk][i] = v.x;
        As_im[kk][i] = v.y;
      }
    } else {
      // element (i,k) at A[k + i*lda]; k fastest -> coalesced
      constexpr int ISTEP = 256 / BK;
      const int kk = tid % BK;
#pragma unroll
      for (int i = tid / BK; i < BM; i += ISTEP) {
        const int64_t gi = m0 + i, gk = k0 + kk;
        c_t v;
        if (gi < M && gk < K) {
          v = A[gk + gi * lda];
        } else {
          v.x = R(0);
          v.y = R(0);
        }
        As_re[kk][i] = v.x;
        As_im[kk][i] = sgnA * v.y;
      }
    }
    // ---- stage op(B)[k0:k0+BK, n0:n0+BN] ----
    if (OPB == 0) {
      // element (k,j) at B[k + j*ldb]; k fastest
      constexpr int JSTEP = 256 / BK;
      const int kk = tid % BK;
#pragma unroll
      for (int j = tid / BK; j < BN; j += JSTEP) {
        const int64_t gj = n0 + j, gk = k0 + kk;
        c_t v;
        if (gj < N && gk < K) {
          v = B[gk + gj * ldb];
        } else {
          v.x = R(0);
          v.y = R(0);
        }
        Bs_re[kk][j] = v.x;
        Bs_im[kk][j] = v.y;
      }
    } else {
      // element (k,j) at B[j + k*ldb]; j fastest
      constexpr int KSTEP = 256 / BN;
      const int j = tid % BN;
#pragma unroll
      for (int kk = tid / BN; kk < BK; kk += KSTEP) {
        const int64_t gj = n0 + j, gk = k0 + kk;
        c_t v;
        if (gj < N && gk < K) {
          v = B[gj + gk * ldb];
        } else {
          v.x = R(0);
          v.y = R(0);
        }
        Bs_re[kk][j] = v.x;
        Bs_im[kk][j] = sgnB * v.y;
      }
    }
    __syncthreads();

#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      R a_re[TM], a_im[TM], b_re[TN], b_im[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        a_re[i] = As_re[kr][wm + i * 16 + (lane & 15)];
        a_im[i] = As_im[kr][wm + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        b_re[j] = Bs_re[kr][wn + j * 16 + (lane & 15)];
        b_im[j] = Bs_im[kr][wn + j * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc_re[i][j] = Traits<R>::mfma(a_re[i], b_re[j], acc_re[i][j]);
          acc_re[i][j] = Traits<R>::mfma(-a_im[i], b_im[j], acc_re[i][j]);
          acc_im[i][j] = Traits<R>::mfma(a_re[i], b_im[j], acc_im[i][j]);
          acc_im[i][j] = Traits<R>::mfma(a_im[i], b_re[j], acc_im[i][j]);
        }
    }
    __syncthreads();
  }

  // ---- epilogue: lane holds column lane%16, rows Traits<R>::row(lane, e) ----
  const bool use_beta = (beta.x != R(0) || beta.y != R(0));
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t col = n0 + wn + j * 16 + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = m0 + wm + i * 16 + Traits<R>::row(lane, e);
        if (row >= M) continue;
        const R pr = acc_re[i][j][e], pi = acc_im[i][j][e];
        c_t out;
        out.x = alpha.x * pr - alpha.y * pi;
        out.y = alpha.x * pi + alpha.y * pr;
        c_t* cp = C + row + col * ldc;
        if (use_beta) {
          const c_t old = *cp;
          out.x += beta.x * old.x - beta.y * old.y;
          out.y += beta.x * old.y + beta.y * old.x;
        }
        *cp = out;
      }
    }
}

template <typename R, int OPA, int OPB, int TM, int TN>
__global__ __launch_bounds__(256) void gemm_kernel(
    int64_t M, int64_t N, int64_t K, typename Traits<R>::c_t alpha,
    const typename Traits<R>::c_t* __restrict__ A, int64_t lda, int64_t sA,
    const typename Traits<R>::c_t* __restrict__ B, int64_t ldb, int64_t sB,
    typename Traits<R>::c_t beta, typename Traits<R>::c_t* __restrict__ C,
    int64_t ldc, int64_t sC) {
  const int64_t bz = blockIdx.z;
  gemm_tile<R, OPA, OPB, TM, TN>(M, N, K, alpha, A + bz * sA, lda, B + bz * sB,
                                 ldb, beta, C + bz * sC, ldc,
                                 (int64_t)blockIdx.x * (32 * TM),
                                 (int64_t)blockIdx.y * (32 * TN));
}

// Grouped form: problem blockIdx.z described by desc[blockIdx.z]; tiles
// outside that problem's M x N exit at once.
template <typename R, int OPA, int OPB, int TM, int TN>
__global__ __launch_bounds__(256) void gemm_grouped_kernel(
    const GemmDesc<R>* __restrict__ desc) {
  const GemmDesc<R> d = desc[blockIdx.z];
  const int64_t m0 = (int64_t)blockIdx.x * (32 * TM);
  const int64_t n0 = (int64_t)blockIdx.y * (32 * TN);
  if (m0 >= d.M || n0 >= d.N) return;
  gemm_tile<R, OPA, OPB, TM, TN>(d.M, d.N, d.K, d.alpha, d.A, d.lda, d.B,
                                 d.ldb, d.beta, d.C, d.ldc, m0, n0);
}

template <typename R, int OPA, int OPB>
int32_t launch_ops(hipStream_t s, int64_t m, int64_t n, int64_t k,
                   typename Traits<R>::c_t alpha,
                   const typename Traits<R>::c_t* A, int64_t lda, int64_t sA,
                   const typename Traits<R>::c_t* B, int64_t ldb, int64_t sB,
                   typename Traits<R>::c_t beta, typename Traits<R>::c_t* C,
                   int64_t ldc, int64_t sC, int64_t batch) {
  // 64x64 tiles when they alone give a few hundred workgroups, else 32x32.
  const int64_t big_tiles = ((m + 63) / 64) * ((n + 63) / 64) * batch;
  if (big_tiles >= 256) {
    dim3 grid((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64),
              (unsigned)batch);
    hipLaunchKernelGGL((gemm_kernel<R, OPA, OPB, 2, 2>), grid, dim3(256), 0, s,
                       m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                       sC);
  } else {
    dim3 grid((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32),
              (unsigned)batch);
    hipLaunchKernelGGL((gemm_kernel<R, OPA, OPB, 1, 1>), grid, dim3(256), 0, s,
                       m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                       sC);
  }
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

template <typename R>
int32_t launch_gemm(hipStream_t s, int opA, int opB, int64_t m, int64_t n,
                    int64_t k, typename Traits<R>::c_t alpha,
                    const typename Traits<R>::c_t* A, int64_t lda, int64_t sA,
                    const typename Traits<R>::c_t* B, int64_t ldb, int64_t sB,
                    typename Traits<R>::c_t beta, typename Traits<R>::c_t* C,
                    int64_t ldc, int64_t sC, int64_t batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return QTN_OK;
  if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || batch > 65535) {
    set_error("gemm: dimension too large for the launch grid");
    return QTN_ERR_ARG;
  }
  if (k <= 0) {
    // C = beta*C: one GEMM with K = 0 reduces to the epilogue only.
    k = 0;
  }
#define QTN_GEMM_CASE(a, b)                                                  \
  if (opA == a && opB == b)                                                  \
    return launch_ops<R, a, b>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB,    \
                               beta, C, ldc, sC, batch);
  QTN_GEMM_CASE(0, 0)
  QTN_GEMM_CASE(0, 1)
  QTN_GEMM_CASE(0, 2)
  QTN_GEMM_CASE(1, 0)
  QTN_GEMM_CASE(1, 1)
  QTN_GEMM_CASE(1, 2)
  QTN_GEMM_CASE(2, 0)
  QTN_GEMM_CASE(2, 1)
  QTN_GEMM_CASE(2, 2)
#undef QTN_GEMM_CASE
  set_error("gemm: bad op code");
  return QTN_ERR_ARG;
}

template <typename R, int OPA, int OPB>
int32_t launch_grouped_ops(hipStream_t s, const GemmDesc<R>* d, int count,
                           int64_t maxM, int64_t maxN) {
  const int64_t big_tiles = ((maxM + 63) / 64) * ((maxN + 63) / 64) * count;
  if (big_tiles >= 256) {
    dim3 grid((unsigned)((maxM + 63) / 64), (unsigned)((maxN + 63) / 64),
              (unsigned)count);
    hipLaunchKernelGGL((gemm_grouped_kernel<R, OPA, OPB, 2, 2>), grid,
                       dim3(256), 0, s, d);
  } else {
    dim3 grid((unsigned)((maxM + 31) / 32), (unsigned)((maxN + 31) / 32),
              (unsigned)count);
    hipLaunchKernelGGL((gemm_grouped_kernel<R, OPA, OPB, 1, 1>), grid,
                       dim3(256), 0, s, d);
  }
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

template <typename R>
int32_t launch_grouped(hipStream_t s, int opA, int opB, const GemmDesc<R>* d,
                       int count, int64_t maxM, int64_t maxN) {
  if (count <= 0 || maxM <= 0 || maxN <= 0) return QTN_OK;
  if (count > 65535 || maxM > INT32_MAX / 2 || maxN > INT32_MAX / 2) {
    set_error("grouped gemm: dimension too large for the launch grid");
    return QTN_ERR_ARG;
  }
#define QTN_GG_CASE(a, b)                                                    \
  if (opA == a && opB == b)                                                  \
    return launch_grouped_ops<R, a, b>(s, d, count, maxM, maxN);
  QTN_GG_CASE(0, 0)
  QTN_GG_CASE(0, 1)
  QTN_GG_CASE(0, 2)
  QTN_GG_CASE(1, 0)
  QTN_GG_CASE(2, 0)
  QTN_GG_CASE(2, 2)
#undef QTN_GG_CASE
  set_error("grouped gemm: unsupported op combination");
  return QTN_ERR_ARG;
}

int op_code(char c) {
  switch (c) {
    case 'N':
    case 'n':
      return 0;
    case 'T':
    case 't':
      return 1;
    case 'C':
    case 'c':
      return 2;
    default:
      return -1;
  }
}

}  // namespace

int32_t launch_zgemm(hipStream_t s, int opA, int opB, int64_t m, int64_t n,
                     int64_t k, double2 alpha, const double2* A, int64_t lda,
                     int64_t sA, const double2* B, int64_t ldb, int64_t sB,
                     double2 beta, double2* C, int64_t ldc, int64_t sC,
                     int64_t batch) {
  return launch_gemm<double>(s, opA, opB, m, n, k, alpha, A, lda, sA, B, ldb,
                             sB, beta, C, ldc, sC, batch);
}

int32_t launch_cgemm(hipStream_t s, int opA, int opB, int64_t m, int64_t n,
                     int64_t k, float2 alpha, const float2* A, int64_t lda,
                     int64_t sA, const float2* B, int64_t ldb, int64_t sB,
                     float2 beta, float2* C, int64_t ldc, int64_t sC,
                     int64_t batch) {
  return launch_gemm<float>(s, opA, opB, m, n, k, alpha, A, lda, sA, B, ldb,
                            sB, beta, C, ldc, sC, batch);
}

int32_t launch_zgemm_grouped(hipStream_t s, int opA, int opB,
                             const GemmDesc<double>* d_desc, int count,
                             int64_t maxM, int64_t maxN) {
  return launch_grouped<double>(s, opA, opB, d_desc, count, maxM, maxN);
}

int32_t launch_cgemm_grouped(hipStream_t s, int opA, int opB,
                             const GemmDesc<float>* d_desc, int count,
                             int64_t maxM, int64_t maxN) {
  return launch_grouped<float>(s, opA, opB, d_desc, count, maxM, maxN);
}

}  // namespace qtn

extern "C" {

int32_t qtn_zgemm_dev(qtn_ctx* ctx, char opA, char opB, int64_t m, int64_t n,
                      int64_t k, const double* alpha, const double* A,
                      int64_t lda, int64_t strideA, const double* B,
                      int64_t ldb, int64_t strideB, const double* beta,
                      double* C, int64_t ldc, int64_t strideC, int64_t batch) try {
  qtn::clear_error();
  QTN_ARG(ctx && alpha && beta, "qtn_zgemm_dev: NULL argument");
  const int oa = qtn::op_code(opA), ob = qtn::op_code(opB);
  QTN_ARG(oa >= 0 && ob >= 0, "qtn_zgemm_dev: op must be 'N', 'T' or 'C'");
  QTN_ARG(m >= 0 && n >= 0 && k >= 0 && batch >= 0,
          "qtn_zgemm_dev: negative dimension");
  // as BLAS: A is stored m x k (op 'N') or k x m, B k x n or n x k
  QTN_ARG(lda >= std::max<int64_t>(1, oa == 0 ? m : k) &&
              ldb >= std::max<int64_t>(1, ob == 0 ? k : n) &&
              ldc >= std::max<int64_t>(1, m),
          "qtn_zgemm_dev: bad leading dimension");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return qtn::launch_zgemm(ctx->stream, oa, ob, m, n, k,
                           make_double2(alpha[0], alpha[1]),
                           (const double2*)A, lda, strideA,
                           (const double2*)B, ldb, strideB,
                           make_double2(beta[0], beta[1]), (double2*)C, ldc,
                           strideC, batch);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_cgemm_dev(qtn_ctx* ctx, char opA, char opB, int64_t m, int64_t n,
                      int64_t k, const float* alpha, const float* A,
                      int64_t lda, int64_t strideA, const float* B,
                      int64_t ldb, int64_t strideB, const float* beta,
                      float* C, int64_t ldc, int64_t strideC, int64_t batch) try {
  qtn::clear_error();
  QTN_ARG(ctx && alpha && beta, "qtn_cgemm_dev: NULL argument");
  const int oa = qtn::op_code(opA), ob = qtn::op_code(opB);
  QTN_ARG(oa >= 0 && ob >= 0, "qtn_cgemm_dev: op must be 'N', 'T' or 'C'");
  QTN_ARG(m >= 0 && n >= 0 && k >= 0 && batch >= 0,
          "qtn_cgemm_dev: negative dimension");
  // as BLAS: A is stored m x k (op 'N') or k x m, B k x n or n x k
  QTN_ARG(lda >= std::max<int64_t>(1, oa == 0 ? m : k) &&
              ldb >= std::max<int64_t>(1, ob == 0 ? k : n) &&
              ldc >= std::max<int64_t>(1, m),
          "qtn_cgemm_dev: bad leading dimension");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return qtn::launch_cgemm(ctx->stream, oa, ob, m, n, k,
                           make_float2(alpha[0], alpha[1]), (const float2*)A,
                           lda, strideA, (const float2*)B, ldb, strideB,
                           make_float2(beta[0], beta[1]), (float2*)C, ldc,
                           strideC, batch);
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
