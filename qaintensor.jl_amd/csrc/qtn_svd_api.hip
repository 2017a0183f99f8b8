// Host-pointer truncated SVD entry points (qtn_svd_trunc, qtn_svd_values)
// on top of the batched Jacobi engine.  They replace LinearAlgebra.svd plus
// the truncation of contract_svd (src/svd.jl:26-33).
#include <algorithm>
#include <cmath>
#include <vector>

#include "qtn_internal.hpp"
#include "qtn_qr.hpp"
#include "qtn_svd.hpp"

namespace qtn {
namespace {

// X (rows_p x cols_p, zero padded) from A (m x n, ld lda):
// trans == 0: X = A; trans == 1: X = A^H.
__global__ void build_x(const double2* __restrict__ A, int64_t m, int64_t n,
                        int64_t lda, int trans, double2* __restrict__ X,
                        int64_t rows_p, int64_t cols_p) {
  const int64_t total = rows_p * cols_p;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % rows_p, c = idx / rows_p;
    double2 v = make_double2(0.0, 0.0);
    if (!trans) {
      if (r < m && c < n) v = A[r + c * lda];
    } else {
      if (r < n && c < m) {
        const double2 a = A[c + r * lda];
        v = make_double2(a.x, -a.y);
      }
    }
    X[idx] = v;
  }
}

// dst column j (j < k) = src column perm[j] (first `rows` rows) scaled by
// 1/sig[j] when scale != 0.  Column-major dst with ld = rows.
__global__ void gather_cols(const double2* __restrict__ src, int64_t ld_src,
                            int64_t rows, const int* __restrict__ perm,
                            const double* __restrict__ sig, int scale,
                            int64_t k, double2* __restrict__ dst) {
  const int64_t total = rows * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % rows, j = idx / rows;
    double2 v = src[r + (int64_t)perm[j] * ld_src];
    if (scale) {
      const double s = sig[j];
      const double inv = s > 0.0 ? 1.0 / s : 0.0;
      v.x *= inv;
      v.y *= inv;
    }
    dst[idx] = v;
  }
}

// dst (k x cols, ld k) row j = conj(src column perm[j])^T; scale 1 divides
// by sig[j], scale 2 multiplies by it.
__global__ void gather_rows_conj(const double2* __restrict__ src,
                                 int64_t ld_src, int64_t cols,
                                 const int* __restrict__ perm,
                                 const double* __restrict__ sig, int scale,
                                 int64_t k, double2* __restrict__ dst) {
  const int64_t total = cols * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx % k, c = idx / k;
    double2 v = src[c + (int64_t)perm[j] * ld_src];
    v.y = -v.y;
    if (scale) {
      const double s = sig[j];
      const double f = scale == 2 ? s : (s > 0.0 ? 1.0 / s : 0.0);
      v.x *= f;
      v.y *= f;
    }
    dst[idx] = v;
  }
}

// build_x into complex FP32.
__global__ void build_x32(const double2* __restrict__ A, int64_t m, int64_t n,
                          int64_t lda, int trans, float2* __restrict__ X,
                          int64_t rows_p, int64_t cols_p) {
  const int64_t total = rows_p * cols_p;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % rows_p, c = idx / rows_p;
    float2 v = make_float2(0.f, 0.f);
    if (!trans) {
      if (r < m && c < n) {
        const double2 a = A[r + c * lda];
        v = make_float2((float)a.x, (float)a.y);
      }
    } else {
      if (r < n && c < m) {
        const double2 a = A[c + r * lda];
        v = make_float2((float)a.x, (float)-a.y);
      }
    }
    X[idx] = v;
  }
}

// dst (rows x k, ld rows, FP64) column j = src (FP32) column perm[j].
__global__ void gather_cols32(const float2* __restrict__ src, int64_t ld_src,
                              int64_t rows, const int* __restrict__ perm,
                              int64_t k, double2* __restrict__ dst) {
  const int64_t total = rows * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % rows, j = idx / rows;
    const float2 v = src[r + (int64_t)perm[j] * ld_src];
    dst[idx] = make_double2((double)v.x, (double)v.y);
  }
}

// Q <- 1.5 Q - 0.5 T (one Newton-Schulz step after T = Q (Q^H Q)).
__global__ void ns_step(double2* __restrict__ Q, const double2* __restrict__ T,
                        int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 q = Q[i], t = T[i];
    Q[i] = make_double2(1.5 * q.x - 0.5 * t.x, 1.5 * q.y - 0.5 * t.y);
  }
}

// Partial sums of |x|^2 over a column-major rows x cols block (ld ld):
// block b writes part[b]; the host adds the parts in order.
constexpr int kSumsqBlocks = 512;
__global__ __launch_bounds__(256) void sumsq_partial(
    const double2* __restrict__ X, int64_t rows, int64_t cols, int64_t ld,
    double* __restrict__ part) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < rows * cols;
       i += (int64_t)gridDim.x * 256) {
    const double2 v = X[(i % rows) + (i / rows) * ld];
    acc += v.x * v.x + v.y * v.y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

// Partial maxima of |re| and |im| over a column-major rows x cols block
// (ld ld): block b writes part[b]; NaN propagates.
__global__ __launch_bounds__(256) void absmax_partial(
    const double2* __restrict__ X, int64_t rows, int64_t cols, int64_t ld,
    double* __restrict__ part) {
  double acc = 0.0;
  bool nan = false;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < rows * cols;
       i += (int64_t)gridDim.x * 256) {
    const double2 v = X[(i % rows) + (i / rows) * ld];
    nan = nan || v.x != v.x || v.y != v.y;
    acc = fmax(acc, fmax(fabs(v.x), fabs(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc = fmax(acc, __shfl_xor(acc, off, 64));
  }
  nan = __any(nan);
  __shared__ double w[4];
  __shared__ int wn[4];
  if ((threadIdx.x & 63) == 0) {
    w[threadIdx.x >> 6] = acc;
    wn[threadIdx.x >> 6] = nan ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    part[blockIdx.x] = (wn[0] | wn[1] | wn[2] | wn[3])
                           ? __longlong_as_double(0x7ff8000000000000LL)
                           : fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
}

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                          8192));
}

// dst (rows_dst x k, ld ld_dst) column j = src column perm[j] in its first
// n_src rows (scaled by 1/sig[j] when scale != 0), zero below.
__global__ void gather_cols_pad(const double2* __restrict__ src,
                                int64_t ld_src, int64_t n_src,
                                const int* __restrict__ perm,
                                const double* __restrict__ sig, int scale,
                                int64_t k, double2* __restrict__ dst,
                                int64_t ld_dst, int64_t rows_dst) {
  const int64_t total = rows_dst * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % rows_dst, j = idx / rows_dst;
    double2 v = make_double2(0.0, 0.0);
    if (r < n_src) {
      v = src[r + (int64_t)perm[j] * ld_src];
      if (scale) {
        const double sj = sig[j];
        const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
        v.x *= inv;
        v.y *= inv;
      }
    }
    dst[r + j * ld_dst] = v;
  }
}

// dst (k x cols, ld k) = src^H for src cols x k (ld ld_src).
__global__ void adjoint_k(const double2* __restrict__ src, int64_t ld_src,
                          int64_t cols, int64_t k, double2* __restrict__ dst) {
  const int64_t total = cols * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx % k, c = idx / k;
    const double2 v = src[c + j * ld_src];
    dst[idx] = make_double2(v.x, -v.y);
  }
}

// Rows of Z (rows x k, ld ldz) moved to their places before the column
// sort: out row qperm[i] = Z row i.  adj == 0: out (rows x k, ld rows);
// adj != 0: out = that matrix's adjoint (k x rows, ld k), rows scaled by
// sig[j] when adj == 2.
__global__ void unsort_rows(const double2* __restrict__ Z, int64_t ldz,
                            int64_t rows, int64_t k,
                            const int* __restrict__ qperm,
                            const double* __restrict__ sig, int adj,
                            double2* __restrict__ out) {
  const int64_t total = rows * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % rows, j = idx / rows;
    double2 v = Z[i + j * ldz];
    const int64_t r = qperm[i];
    if (!adj) {
      out[r + j * rows] = v;
    } else {
      const double f = adj == 2 ? sig[j] : 1.0;
      out[j + r * k] = make_double2(v.x * f, -v.y * f);
    }
  }
}

constexpr int kSlotA = 11, kSlotX = 12, kSlotW = 13, kSlotU = 14, kSlotV = 15;
// QR-preconditioned path: the magnitude check, the QR dims, two reflector
// sets (8 slots each), the sorted X, R1^H, column norms and the sort
// permutation.
constexpr int kSlotPreMax = 118, kSlotPreDims = 119, kSlotPreQr1 = 120,
              kSlotPreQr2 = 128, kSlotPreA1 = 136, kSlotPreY1 = 137,
              kSlotPreNrm = 138, kSlotPrePerm = 139;
// Splits with at least this many columns (of X) are QR-preconditioned.
constexpr int64_t kPreMinCols = 512;
// The Householder norms are plain sums of squares: the preconditioner runs
// only when every entry is at most 2^450 in magnitude (a sum of 2^12 squares
// stays finite) and the largest is at least 2^-450; other inputs take the
// rescaling engine directly.
constexpr double kPreMaxAbs = 0x1p450, kPreMinAbs = 0x1p-450;

}  // namespace

namespace {

// QR preconditioning switch, QTN_SVD_PRECOND (read at every call): unset,
// the default of the caller; 0 off; any other value on.
bool precond_enabled(bool dflt) {
  const char* e = getenv("QTN_SVD_PRECOND");
  if (!e || !*e) return dflt;
  return atoi(e) != 0;
}

// Truncation results of a finished engine run (host copies).
int32_t read_trunc(qtn_ctx* ctx, const JacobiEngine& eng, bool truncate,
                   int64_t* k, double* disc) {
  if (!truncate) return QTN_OK;
  hipStream_t s = ctx->stream;
  int st = 0;
  QTN_CHECK_HIP(hipMemcpyAsync(k, eng.kept, sizeof(int64_t),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(&st, eng.status, sizeof(int),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(disc, eng.disc, sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  if (st != 0) {
    set_error("svd: every singular value is within the truncation error");
    return QTN_ERR_TRUNC;
  }
  return QTN_OK;
}

// Whether A (m x n, ld lda) may take the QR preconditioner: every entry
// finite and the largest magnitude within [kPreMinAbs, kPreMaxAbs].
int32_t precond_range_ok(qtn_ctx* ctx, const double2* dA, int64_t m,
                         int64_t n, int64_t lda, bool* ok) {
  hipStream_t s = ctx->stream;
  void* dP;
  QTN_TRY(ctx->buffer(kSlotPreMax, kSumsqBlocks * sizeof(double), &dP));
  std::vector<double> part(kSumsqBlocks);
  hipLaunchKernelGGL(absmax_partial, dim3(kSumsqBlocks), dim3(256), 0, s, dA,
                     m, n, lda, (double*)dP);
  QTN_CHECK_HIP(hipGetLastError());
  QTN_CHECK_HIP(hipMemcpyAsync(part.data(), dP, kSumsqBlocks * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  double amax = 0.0;
  bool finite = true;
  for (double v : part) {
    finite = finite && std::isfinite(v);
    amax = std::max(amax, v);
  }
  *ok = finite && amax <= kPreMaxAbs && amax >= kPreMinAbs;
  return QTN_OK;
}

// svd_device through the preconditioner of Drmac & Veselic, "New fast and
// accurate Jacobi SVD algorithm I", SIAM J. Matrix Anal. Appl. 29 (2008):
// the columns of X (A, or A^H when m < n; xr x xc, xr >= xc) sorted by
// decreasing norm, X P = Q1 R1, R1^H = Q2 R2, and the one-sided Jacobi
// sweeps run on Y = R2^H (xc x xc), with W accumulated: a tall X loses its
// extra rows and Y's columns start much closer to orthogonal, so the
// sweeps are cheaper and fewer.  With Y W = U_Y S,
//   X = (Q1 [U_Y; 0]) S (P Q2 W)^H.
// Contract as svd_device.  Uses slots 12, 13, 118-139 and the engine's.
int32_t svd_device_pre(qtn_ctx* ctx, const double2* dA, int64_t m,
                       int64_t n, int64_t lda, bool truncate, double cutoff,
                       int relative, int64_t chimax, int absorb, double2* dU,
                       double* dS, double2* dVh, int64_t* k_out,
                       double* disc_out) {
  hipStream_t s = ctx->stream;
  const int trans = (m < n) ? 1 : 0;
  const int64_t xr = trans ? n : m, xc = trans ? m : n;
  const int64_t rows_p = round_up64(xr), cols_p = round_up64(xc);
  const bool want_vectors = dU != nullptr || dVh != nullptr;
  void *dX, *dW = nullptr, *dA1, *dY1, *pd;
  QTN_TRY(ctx->buffer(kSlotX, (size_t)rows_p * cols_p * sizeof(double2),
                      &dX));
  if (want_vectors)
    QTN_TRY(ctx->buffer(kSlotW, (size_t)cols_p * cols_p * sizeof(double2),
                        &dW));
  QTN_TRY(ctx->buffer(kSlotPreA1, (size_t)rows_p * cols_p * sizeof(double2),
                      &dA1));
  QTN_TRY(ctx->buffer(kSlotPreY1, (size_t)cols_p * cols_p * sizeof(double2),
                      &dY1));
  QTN_TRY(ctx->buffer(kSlotPreDims, 2 * sizeof(QrDims), &pd));
  const std::vector<QrDims> d1{QrDims{xr, xc, 1}}, d2{QrDims{xc, xc, 1}};
  const QrDims hd[2] = {d1[0], d2[0]};
  QrDims* d_dims = (QrDims*)pd;
  QTN_CHECK_HIP(hipMemcpyAsync(d_dims, hd, sizeof(hd), hipMemcpyHostToDevice,
                               s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));  // hd lives on this stack
  hipLaunchKernelGGL(build_x, dim3(grid_for(rows_p * cols_p)), dim3(256), 0,
                     s, dA, m, n, lda, trans, (double2*)dX, rows_p, cols_p);
  QTN_CHECK_HIP(hipGetLastError());
  int* qperm = nullptr;
  QTN_TRY(qr_sort_columns(ctx, (const double2*)dX, (double2*)dA1, rows_p,
                          cols_p, rows_p * cols_p, 1, d_dims, kSlotPreNrm,
                          kSlotPrePerm, &qperm));
  BatchedQR qr1, qr2;
  qr1.slot_base = kSlotPreQr1;
  qr2.slot_base = kSlotPreQr2;
  QTN_TRY(qr1.factor(ctx, (double2*)dA1, rows_p, rows_p * cols_p, cols_p, d1,
                     want_vectors));
  QTN_TRY(qr_r_to_adjoint(ctx, (const double2*)dA1, rows_p, rows_p * cols_p,
                          (double2*)dY1, cols_p, cols_p * cols_p, cols_p,
                          cols_p, 1, d_dims));
  QTN_TRY(qr2.factor(ctx, (double2*)dY1, cols_p, cols_p * cols_p, cols_p, d2,
                     want_vectors));
  double2* Y = (double2*)dX;  // Y = R2^H, cols_p x cols_p, ld cols_p
  QTN_TRY(qr_r_to_adjoint(ctx, (const double2*)dY1, cols_p, cols_p * cols_p,
                          Y, cols_p, cols_p * cols_p, cols_p, cols_p, 1,
                          d_dims + 1));
  JacobiEngine eng;
  QTN_TRY(eng.reserve(ctx, 1, cols_p, cols_p));
  const int64_t nsv = xc;
  int sweeps = 0;
  QTN_TRY(eng.run(ctx, Y, cols_p, cols_p * cols_p, (double2*)dW,
                  cols_p * cols_p, nullptr, &nsv, truncate ? cutoff : 0.0,
                  truncate ? relative : 0, truncate ? chimax : 0, 60,
                  &sweeps));
  int64_t k = nsv;
  double disc = 0.0;
  QTN_TRY(read_trunc(ctx, eng, truncate, &k, &disc));
  if (k_out) *k_out = k;
  if (disc_out) *disc_out = disc;
  if (dS && k > 0)
    QTN_CHECK_HIP(hipMemcpyAsync(dS, eng.sig, k * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
  if (k == 0 || !want_vectors) return QTN_OK;
  // left vectors of X: Q1 [Y_k; 0], Y's columns normalised (or kept as
  // U_Y S where they become the absorbed rows S V^H)
  const bool left_needed = trans ? dVh != nullptr : dU != nullptr;
  const bool right_needed = trans ? dU != nullptr : dVh != nullptr;
  if (left_needed) {
    double2* L = trans ? (double2*)dA1 : dU;  // xr x k, ld xr
    const int scale = (trans && absorb) ? 0 : 1;
    hipLaunchKernelGGL(gather_cols_pad, dim3(grid_for(xr * k)), dim3(256), 0,
                       s, (const double2*)Y, cols_p, xc, eng.perm, eng.sig,
                       scale, k, L, xr, xr);
    QTN_CHECK_HIP(hipGetLastError());
    QTN_TRY(qr1.apply_q(ctx, L, {xr}, 0, {k}, d1));
    if (trans)
      hipLaunchKernelGGL(adjoint_k, dim3(grid_for(xr * k)), dim3(256), 0, s,
                         (const double2*)L, xr, xr, k, dVh);
  }
  // right vectors of X: P Q2 W_k
  if (right_needed) {
    double2* Z = (double2*)dY1;  // xc x k, ld cols_p
    hipLaunchKernelGGL(gather_cols_pad, dim3(grid_for(xc * k)), dim3(256), 0,
                       s, (const double2*)dW, cols_p, xc, eng.perm, eng.sig,
                       0, k, Z, cols_p, xc);
    QTN_CHECK_HIP(hipGetLastError());
    QTN_TRY(qr2.apply_q(ctx, Z, {cols_p}, 0, {k}, d2));
    hipLaunchKernelGGL(unsort_rows, dim3(grid_for(xc * k)), dim3(256), 0, s,
                       (const double2*)Z, cols_p, xc, k, (const int*)qperm,
                       eng.sig, trans ? 0 : (absorb ? 2 : 1),
                       trans ? dU : dVh);
  }
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

}  // namespace

// SVD of a device matrix A (m x n, ld lda) by the batched Jacobi engine.
//   truncate: tail rule of src/svd.jl:29-33 (er = cutoff, or cutoff*||s||
//             when relative), then k <= chimax (chimax > 0); otherwise
//             k = min(m, n) (exact zeros kept, as LinearAlgebra.svd does).
//   absorb:   Vh rows are scaled by S (the S*V^H of src/switch.jl:82-89,
//             src/mps.jl:152 and src/mpo.jl:56).
// Device outputs (any may be null): dU (m x k, ld m), dS (k), dVh (k x n,
// ld k); their capacity must be min(m, n) columns / values / rows.  For an
// exactly zero singular value the matching vector is zero (the Jacobi
// engine produces no orthonormal completion).  Uses slots 12 and 13.  With
// QTN_SVD_PRECOND=1, splits with min(m, n) >= 512 take the QR-preconditioned
// sweeps (svd_device_pre) unless allow_pre is false: fewer and cheaper
// sweeps, but two panel-sequential QRs and their Q applications, which cost
// more than they save on the splits measured so far.
int32_t svd_device(qtn_ctx* ctx, const double2* dA, int64_t m, int64_t n,
                   int64_t lda, bool truncate, double cutoff, int relative,
                   int64_t chimax, int absorb, double2* dU, double* dS,
                   double2* dVh, int64_t* k_out, double* disc_out,
                   bool allow_pre) {
  hipStream_t s = ctx->stream;
  const int trans = (m < n) ? 1 : 0;  // columns of X = min(m, n)
  const int64_t xr = trans ? n : m;
  const int64_t xc = trans ? m : n;
  const int64_t rows_p = round_up64(xr), cols_p = round_up64(xc);
  if (cols_p > 4096) {
    set_error("svd: min(m, n) above 4096 is not supported");
    return QTN_ERR_ARG;
  }
  if (allow_pre && xc >= kPreMinCols && precond_enabled(false)) {
    bool ok = false;
    QTN_TRY(precond_range_ok(ctx, dA, m, n, lda, &ok));
    if (ok)
      return svd_device_pre(ctx, dA, m, n, lda, truncate, cutoff, relative,
                            chimax, absorb, dU, dS, dVh, k_out, disc_out);
  }
  const bool want_vectors = dU != nullptr || dVh != nullptr;
  void *dX, *dW = nullptr;
  QTN_TRY(ctx->buffer(kSlotX, (size_t)rows_p * cols_p * sizeof(double2),
                      &dX));
  if (want_vectors)
    QTN_TRY(ctx->buffer(kSlotW, (size_t)cols_p * cols_p * sizeof(double2),
                        &dW));
  hipLaunchKernelGGL(build_x, dim3(grid_for(rows_p * cols_p)), dim3(256), 0,
                     s, dA, m, n, lda, trans, (double2*)dX, rows_p, cols_p);
  QTN_CHECK_HIP(hipGetLastError());
  JacobiEngine eng;
  QTN_TRY(eng.reserve(ctx, 1, rows_p, cols_p));
  const int64_t nsv = std::min(m, n);
  int sweeps = 0;
  QTN_TRY(eng.run(ctx, (double2*)dX, rows_p, rows_p * cols_p, (double2*)dW,
                  cols_p * cols_p, nullptr, &nsv, truncate ? cutoff : 0.0,
                  truncate ? relative : 0, truncate ? chimax : 0, 60,
                  &sweeps));
  int64_t k = nsv;
  double disc = 0.0;
  QTN_TRY(read_trunc(ctx, eng, truncate, &k, &disc));
  if (k_out) *k_out = k;
  if (disc_out) *disc_out = disc;
  if (dS && k > 0)
    QTN_CHECK_HIP(hipMemcpyAsync(dS, eng.sig, k * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
  if (k == 0 || !want_vectors) return QTN_OK;
  if (!trans) {
    // X = A: A W = U diag(s), V = W
    if (dU)
      hipLaunchKernelGGL(gather_cols, dim3(grid_for(m * k)), dim3(256), 0, s,
                         (const double2*)dX, rows_p, m, eng.perm, eng.sig, 1,
                         k, dU);
    if (dVh)
      hipLaunchKernelGGL(gather_rows_conj, dim3(grid_for(n * k)), dim3(256),
                         0, s, (const double2*)dW, cols_p, n, eng.perm,
                         eng.sig, absorb ? 2 : 0, k, dVh);
  } else {
    // X = A^H: A^H W = V diag(s), U = W
    if (dU)
      hipLaunchKernelGGL(gather_cols, dim3(grid_for(m * k)), dim3(256), 0, s,
                         (const double2*)dW, cols_p, m, eng.perm, eng.sig, 0,
                         k, dU);
    if (dVh)
      hipLaunchKernelGGL(gather_rows_conj, dim3(grid_for(n * k)), dim3(256),
                         0, s, (const double2*)dX, rows_p, n, eng.perm,
                         eng.sig, absorb ? 0 : 1, k, dVh);
  }
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

// Left singular vectors and values of B = A (adj == 0) or B = A^H
// (adj != 0; the right singular vectors of A) only: U (mB x k, ld mB) and
// S, truncated as svd_device; the caller forms the other factor,
// S V^H = U^H B, with one GEMM, which is exact to rounding because U's
// columns are orthonormal.  No rotations are accumulated.  Tall or square
// B (mB >= nB): one-sided Jacobi on X = B, U its normalised columns.  Wide
// B: B^H P = Q1 R1 (columns sorted by norm), so B = P R1^H Q1^H, and the
// sweeps run on L = R1^H (mB x mB); then U = P U_L, and Q1 is never formed.
// Uses slots 12, 15, 118-139 and the engine's.
int32_t svd_device_left(qtn_ctx* ctx, const double2* dA, int64_t m,
                        int64_t n, int64_t lda, int adj, double cutoff,
                        int relative, int64_t chimax, double2* dU, double* dS,
                        int64_t* k_out, double* disc_out) {
  hipStream_t s = ctx->stream;
  const int64_t mb = adj ? n : m, nb = adj ? m : n;  // B is mb x nb
  const int64_t kmin = std::min(m, n);
  if (round_up64(kmin) > 4096) {
    set_error("svd: min(m, n) above 4096 is not supported");
    return QTN_ERR_ARG;
  }
  bool pre = mb < nb && precond_enabled(true);
  if (pre) QTN_TRY(precond_range_ok(ctx, dA, m, n, lda, &pre));
  if (mb < nb && !pre) {
    // U of B = A: svd_device's U; of B = A^H: svd_device's V (V^H turned)
    if (!adj)
      return svd_device(ctx, dA, m, n, lda, true, cutoff, relative, chimax, 0,
                        dU, dS, nullptr, k_out, disc_out, false);
    void* dV;
    QTN_TRY(ctx->buffer(kSlotV, (size_t)kmin * n * sizeof(double2), &dV));
    int64_t k = 0;
    QTN_TRY(svd_device(ctx, dA, m, n, lda, true, cutoff, relative, chimax, 0,
                       nullptr, dS, (double2*)dV, &k, disc_out, false));
    if (k_out) *k_out = k;
    if (k > 0 && dU) {
      hipLaunchKernelGGL(adjoint_k, dim3(grid_for(n * k)), dim3(256), 0, s,
                         (const double2*)dV, k, k, n, dU);
      QTN_CHECK_HIP(hipGetLastError());
    }
    return QTN_OK;
  }
  const int64_t xr = std::max(m, n);  // rows of the matrix factored
  const int64_t xc = kmin;
  const int64_t rows_p = round_up64(xr), cols_p = round_up64(xc);
  void *dX, *dA1, *pd;
  QTN_TRY(ctx->buffer(kSlotX, (size_t)rows_p * cols_p * sizeof(double2),
                      &dX));
  // X = B when B is tall, else B^H (the matrix the QR factors)
  const int x_adj = (mb >= nb) ? adj : !adj;  // X = A^H
  hipLaunchKernelGGL(build_x, dim3(grid_for(rows_p * cols_p)), dim3(256), 0,
                     s, dA, m, n, lda, x_adj, (double2*)dX, rows_p, cols_p);
  QTN_CHECK_HIP(hipGetLastError());
  double2* L = (double2*)dX;  // the matrix the sweeps run on
  int64_t lrows_p = rows_p;
  int* qperm = nullptr;
  if (pre) {
    QTN_TRY(ctx->buffer(kSlotPreA1, (size_t)rows_p * cols_p * sizeof(double2),
                        &dA1));
    QTN_TRY(ctx->buffer(kSlotPreDims, sizeof(QrDims), &pd));
    const std::vector<QrDims> d1{QrDims{xr, xc, 1}};
    QrDims* d_dims = (QrDims*)pd;
    QTN_CHECK_HIP(hipMemcpyAsync(d_dims, d1.data(), sizeof(QrDims),
                                 hipMemcpyHostToDevice, s));
    QTN_CHECK_HIP(hipStreamSynchronize(s));  // d1 lives on this stack
    QTN_TRY(qr_sort_columns(ctx, (const double2*)dX, (double2*)dA1, rows_p,
                            cols_p, rows_p * cols_p, 1, d_dims, kSlotPreNrm,
                            kSlotPrePerm, &qperm));
    BatchedQR qr1;
    qr1.slot_base = kSlotPreQr1;
    QTN_TRY(qr1.factor(ctx, (double2*)dA1, rows_p, rows_p * cols_p, cols_p,
                       d1, false));
    // L = R1^H (cols_p x cols_p, ld cols_p) over X, whose rows are no
    // longer needed
    QTN_TRY(qr_r_to_adjoint(ctx, (const double2*)dA1, rows_p,
                            rows_p * cols_p, L, cols_p, cols_p * cols_p,
                            cols_p, cols_p, 1, d_dims));
    lrows_p = cols_p;
  }
  JacobiEngine eng;
  QTN_TRY(eng.reserve(ctx, 1, lrows_p, cols_p));
  const int64_t nsv = xc;
  int sweeps = 0;
  QTN_TRY(eng.run(ctx, L, lrows_p, lrows_p * cols_p, nullptr, 0, nullptr,
                  &nsv, cutoff, relative, chimax, 60, &sweeps));
  int64_t k = nsv;
  double disc = 0.0;
  QTN_TRY(read_trunc(ctx, eng, true, &k, &disc));
  if (k_out) *k_out = k;
  if (disc_out) *disc_out = disc;
  if (dS && k > 0)
    QTN_CHECK_HIP(hipMemcpyAsync(dS, eng.sig, k * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
  if (k == 0 || !dU) return QTN_OK;
  if (!pre) {
    hipLaunchKernelGGL(gather_cols, dim3(grid_for(mb * k)), dim3(256), 0, s,
                       (const double2*)L, lrows_p, mb, eng.perm, eng.sig, 1,
                       k, dU);
  } else {
    // U_L (normalised columns) into the free sorted-X buffer, then its rows
    // back to their places before the column sort
    double2* Z = (double2*)dA1;
    hipLaunchKernelGGL(gather_cols_pad, dim3(grid_for(mb * k)), dim3(256), 0,
                       s, (const double2*)L, lrows_p, mb, eng.perm, eng.sig,
                       1, k, Z, mb, mb);
    hipLaunchKernelGGL(unsort_rows, dim3(grid_for(mb * k)), dim3(256), 0, s,
                       (const double2*)Z, mb, mb, k, (const int*)qperm,
                       eng.sig, 0, dU);
  }
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

namespace {
int32_t fp64_fallback(bool* fallback) {
  *fallback = true;
  return QTN_OK;
}
}  // namespace

// The chosen form of svd_device for a truncation the cap may decide
// (truncate with chimax > 0 and min(m, n) > chimax): the Jacobi sweeps run
// on a complex FP32 copy; the result is kept when the cap binds inside the
// spectrum FP32 resolves (kept == chimax, clean status).  Then the leading
// chimax + 64 singular vectors (U when m < n, V otherwise) are taken to
// FP64 and re-orthonormalised by three Newton-Schulz steps,
// Q <- Q (3 I - Q^H Q) / 2, and A's projection on span(Q) is split by the
// FP64 engine (Rayleigh-Ritz), whose factors give U, S and V^H.
// A-posteriori check, in FP64: the weight of A outside span(Q),
// w_out = |A - Q Q^H A|_F^2 (the residual formed explicitly and summed
// entry by entry), must not exceed the weight db the split discards inside
// span(Q) (the 64 guard directions past the cap).  When it does, the
// spectrum does not decay past the cap (a flat, Marchenko-Pastur-like
// spectrum): the kept subspace is then not resolved by the guard
// directions, and *fallback is set: the caller splits A in FP64 instead.
// The same happens when the cap does not bind (min(m, n) <= chimax) and
// when FP32 or the projection keeps another rank.  disc_out is the FP64
// discarded weight of the returned split, w_out + db.
// Outputs as svd_device (dU and dVh required), unless *fallback.  Uses
// slots 200-213 and svd_device's.
int32_t svd_device_chosen(qtn_ctx* ctx, const double2* dA, int64_t m,
                          int64_t n, int64_t lda, double cutoff, int relative,
                          int64_t chimax, int absorb, double2* dU, double* dS,
                          double2* dVh, int64_t* k_out, double* disc_out,
                          bool* fallback) {
  *fallback = false;
  const int64_t nsv = std::min(m, n);
  if (chimax <= 0 || nsv <= chimax || !dU || !dVh)
    return fp64_fallback(fallback);
  hipStream_t s = ctx->stream;
  const int trans = (m < n) ? 1 : 0;
  const int64_t xr = trans ? n : m, xc = trans ? m : n;
  const int64_t rows_p = round_up64(xr), cols_p = round_up64(xc);
  if (cols_p > 4096) {
    set_error("svd: min(m, n) above 4096 is not supported");
    return QTN_ERR_ARG;
  }
  void *dX, *dW;
  QTN_TRY(ctx->buffer(205, (size_t)rows_p * cols_p * sizeof(float2), &dX));
  QTN_TRY(ctx->buffer(206, (size_t)cols_p * cols_p * sizeof(float2), &dW));
  hipLaunchKernelGGL(build_x32, dim3(grid_for(rows_p * cols_p)), dim3(256), 0,
                     s, dA, m, n, lda, trans, (float2*)dX, rows_p, cols_p);
  QTN_CHECK_HIP(hipGetLastError());
  JacobiEngine eng;
  eng.slot_base = 200;
  QTN_TRY(eng.reserve(ctx, 1, rows_p, cols_p, true));
  int sweeps = 0;
  QTN_TRY(eng.run32(ctx, (float2*)dX, rows_p, rows_p * cols_p, (float2*)dW,
                    cols_p * cols_p, nullptr, &nsv, cutoff, relative, chimax,
                    60, &sweeps));
  int64_t k = 0;
  int st = 0;
  double disc = 0.0;
  QTN_CHECK_HIP(hipMemcpyAsync(&k, eng.kept, sizeof(int64_t),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(&st, eng.status, sizeof(int),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(&disc, eng.disc, sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  if (st != 0 || k != chimax) {
    ctx->n_rr_redone += 1;
    return fp64_fallback(fallback);
  }
  // Q: the leading kext = k + 64 vectors from W (U when trans, V
  // otherwise), FP64, re-orthonormalised by Newton-Schulz steps.
  const int64_t kext = std::min(nsv, k + 64);
  const int64_t r = trans ? m : n;
  void *dQ, *dG, *dT, *dB, *dF;
  QTN_TRY(ctx->buffer(207, (size_t)r * kext * sizeof(double2), &dQ));
  QTN_TRY(ctx->buffer(208, (size_t)kext * kext * sizeof(double2), &dG));
  QTN_TRY(ctx->buffer(209, (size_t)r * kext * sizeof(double2), &dT));
  QTN_TRY(ctx->buffer(210, (size_t)(trans ? n : m) * kext * sizeof(double2),
                      &dB));
  QTN_TRY(ctx->buffer(211, (size_t)kext * k * sizeof(double2), &dF));
  double2* Q = (double2*)dQ;
  double2* G = (double2*)dG;
  double2* T = (double2*)dT;
  double2* B = (double2*)dB;
  double2* F = (double2*)dF;
  hipLaunchKernelGGL(gather_cols32, dim3(grid_for(r * kext)), dim3(256), 0,
                     s, (const float2*)dW, cols_p, r, eng.perm, kext, Q);
  QTN_CHECK_HIP(hipGetLastError());
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  for (int it = 0; it < 3; ++it) {
    QTN_TRY(launch_zgemm(s, 2, 0, kext, kext, r, one, Q, r, 0, Q, r, 0, zero,
                         G, kext, 0, 1));
    QTN_TRY(launch_zgemm(s, 0, 0, r, kext, kext, one, Q, r, 0, G, kext, 0,
                         zero, T, r, 0, 1));
    hipLaunchKernelGGL(ns_step, dim3(grid_for(r * kext)), dim3(256), 0, s, Q,
                       T, r * kext);
    QTN_CHECK_HIP(hipGetLastError());
  }
  // Rayleigh-Ritz in FP64: the projection of A on span(Q) is split by the
  // FP64 engine (its columns are orthogonal to FP32 accuracy, so it takes
  // few sweeps and skips the QR preconditioner), which settles the order
  // of the values near the cap that FP32 cannot resolve.  The split must
  // keep exactly the cap again.
  int64_t kb = 0;
  double db = 0.0;
  if (trans) {
    // B = Q^H A (kext x n) = Ub S Vh: U = Q Ub
    QTN_TRY(launch_zgemm(s, 2, 0, kext, n, m, one, Q, m, 0, dA, lda, 0, zero,
                         B, kext, 0, 1));
    QTN_TRY(svd_device(ctx, B, kext, n, kext, true, cutoff, relative, chimax,
                       absorb, F, dS, dVh, &kb, &db, false));
  } else {
    // B = A Q (m x kext) = U S Vb: V^H = Vb Q^H
    QTN_TRY(launch_zgemm(s, 0, 0, m, kext, n, one, dA, lda, 0, Q, n, 0, zero,
                         B, m, 0, 1));
    QTN_TRY(svd_device(ctx, B, m, kext, m, true, cutoff, relative, chimax,
                       absorb, dU, dS, F, &kb, &db, false));
  }
  if (kb != chimax) {
    ctx->n_rr_redone += 1;
    return fp64_fallback(fallback);
  }
  // residual R = A - Q B (trans) or A - B Q^H, and its squared norm
  void *dR, *dP;
  QTN_TRY(ctx->buffer(212, (size_t)m * n * sizeof(double2), &dR));
  QTN_TRY(ctx->buffer(213, kSumsqBlocks * sizeof(double), &dP));
  double2* Rm = (double2*)dR;
  QTN_CHECK_HIP(hipMemcpy2DAsync(Rm, m * sizeof(double2), dA,
                                 lda * sizeof(double2), m * sizeof(double2),
                                 n, hipMemcpyDeviceToDevice, s));
  const double2 mone = make_double2(-1.0, 0.0);
  if (trans)
    QTN_TRY(launch_zgemm(s, 0, 0, m, n, kext, mone, Q, m, 0, B, kext, 0, one,
                         Rm, m, 0, 1));
  else
    QTN_TRY(launch_zgemm(s, 0, 2, m, n, kext, mone, B, m, 0, Q, n, 0, one,
                         Rm, m, 0, 1));
  hipLaunchKernelGGL(sumsq_partial, dim3(kSumsqBlocks), dim3(256), 0, s,
                     (const double2*)Rm, m, n, m, (double*)dP);
  QTN_CHECK_HIP(hipGetLastError());
  std::vector<double> part(kSumsqBlocks);
  QTN_CHECK_HIP(hipMemcpyAsync(part.data(), dP, kSumsqBlocks * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  double w_out = 0.0;
  for (double v : part) w_out += v;
  if (!(w_out <= db)) {
    ctx->n_rr_redone += 1;
    return fp64_fallback(fallback);
  }
  ctx->n_rr_accepted += 1;
  if (trans)
    QTN_TRY(launch_zgemm(s, 0, 0, m, kb, kext, one, Q, m, 0, F, kext, 0, zero,
                         dU, m, 0, 1));
  else
    QTN_TRY(launch_zgemm(s, 0, 2, kb, n, kext, one, F, kb, 0, Q, n, 0, zero,
                         dVh, kb, 0, 1));
  if (k_out) *k_out = kb;
  if (disc_out) *disc_out = w_out + db;
  (void)disc;
  return QTN_OK;
}

namespace {

// Host matrix in, host factors out (the qtn_svd* entry points).
int32_t svd_host(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                 int64_t lda, bool truncate, double cutoff, int relative,
                 int64_t chimax, int absorb, int64_t* k_out, double* U,
                 double* S, double* Vh, double* disc_out) {
  hipStream_t s = ctx->stream;
  const int64_t kk = std::min(m, n);
  void *dA, *dU = nullptr, *dV = nullptr, *dS = nullptr;
  QTN_TRY(ctx->buffer(kSlotA, (size_t)m * n * sizeof(double2), &dA));
  if (U) QTN_TRY(ctx->buffer(kSlotU, (size_t)m * kk * sizeof(double2), &dU));
  if (Vh) QTN_TRY(ctx->buffer(kSlotV, (size_t)kk * n * sizeof(double2), &dV));
  if (S) QTN_TRY(ctx->buffer(16, (size_t)kk * sizeof(double), &dS));
  QTN_CHECK_HIP(hipMemcpy2DAsync(dA, m * sizeof(double2), A,
                                 lda * sizeof(double2), m * sizeof(double2), n,
                                 hipMemcpyHostToDevice, s));
  int64_t k = 0;
  QTN_TRY(svd_device(ctx, (const double2*)dA, m, n, m, truncate, cutoff,
                     relative, chimax, absorb, (double2*)dU, (double*)dS,
                     (double2*)dV, &k, disc_out));
  if (k_out) *k_out = k;
  if (k > 0) {
    if (S)
      QTN_CHECK_HIP(hipMemcpyAsync(S, dS, k * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
    if (U)
      QTN_CHECK_HIP(hipMemcpyAsync(U, dU, (size_t)m * k * sizeof(double2),
                                   hipMemcpyDeviceToHost, s));
    if (Vh)
      QTN_CHECK_HIP(hipMemcpyAsync(Vh, dV, (size_t)k * n * sizeof(double2),
                                   hipMemcpyDeviceToHost, s));
  }
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
}

}  // namespace
}  // namespace qtn

extern "C" {

int32_t qtn_svd_trunc(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                      int64_t lda, double cutoff, int32_t relative,
                      int64_t chimax, int64_t* k, double* U, double* S,
                      double* Vh, double* disc_weight) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr && A != nullptr, "qtn_svd_trunc: NULL argument");
  QTN_ARG(m >= 1 && n >= 1, "qtn_svd_trunc: empty matrix");
  QTN_ARG(lda >= m, "qtn_svd_trunc: lda < m");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return qtn::svd_host(ctx, m, n, A, lda, true, cutoff, relative, chimax, 0,
                       k, U, S, Vh, disc_weight);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_svd(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                int64_t lda, int32_t absorb, double* U, double* S,
                double* Vh) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr && A != nullptr, "qtn_svd: NULL argument");
  QTN_ARG(m >= 1 && n >= 1, "qtn_svd: empty matrix");
  QTN_ARG(lda >= m, "qtn_svd: lda < m");
  QTN_ARG(absorb == 0 || absorb == 1, "qtn_svd: absorb must be 0 or 1");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return qtn::svd_host(ctx, m, n, A, lda, false, 0.0, 0, 0, absorb, nullptr,
                       U, S, Vh, nullptr);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_svd_values(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                       int64_t lda, double* S) try {
  qtn::clear_error();
  QTN_ARG(ctx != nullptr && A != nullptr && S != nullptr,
          "qtn_svd_values: NULL argument");
  QTN_ARG(m >= 1 && n >= 1, "qtn_svd_values: empty matrix");
  QTN_ARG(lda >= m, "qtn_svd_values: lda < m");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return qtn::svd_host(ctx, m, n, A, lda, false, 0.0, 0, 0, 0, nullptr,
                       nullptr, S, nullptr, nullptr);
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
