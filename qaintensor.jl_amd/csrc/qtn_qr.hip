// Batched blocked Householder QR for the Jacobi preconditioner.
//
// Algorithm: LAPACK's zgeqrf -- panels of NB columns factorised with
// zlarfg-style reflectors H = I - tau v v^H (v_0 = 1, beta real), the panel
// product written in compact WY form Q_p = H_1..H_NB = I - V T V^H (zlarft,
// forward / columnwise), the trailing matrix updated as A <- Q_p^H A =
// A - (V T^H)(V^H A) with two grouped MFMA GEMMs.  Used as in Drmac &
// Veselic, "New fast and accurate Jacobi SVD algorithm I/II", SIAM J.
// Matrix Anal. Appl. 29 (2008): X P = Q1 R1, R1^H = Q2 R2, one-sided Jacobi
// on R2^H.  The columns of X are sorted by decreasing norm first, which for
// the column-graded X of a two-site update gives the rank-revealing effect
// of column pivoting without a sequential pivot search.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qtn_internal.hpp"
#include "qtn_qr.hpp"

namespace qtn {
namespace {

constexpr int NB = 16;  // panel width

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a)*b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Column norms^2 of every column of A (rows x cols used, ld lda).
// grid (cols_p/4, batch); one wave per column.
__global__ __launch_bounds__(256) void qr_colnorm2(
    const double2* __restrict__ A, int64_t lda, int64_t sA, int64_t cols_p,
    const QrDims* __restrict__ dims, double* __restrict__ out) {
  const int bz = blockIdx.y;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= cols_p) return;
  const QrDims d = dims[bz];
  double s = 0.0;
  if (d.active && c < d.cols) {
    const double2* x = A + (int64_t)bz * sA + c * lda;
    for (int64_t r = lane; r < d.rows; r += 64) {
      const double2 v = x[r];
      s += v.x * v.x + v.y * v.y;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[(int64_t)bz * cols_p + c] = (d.active && c < d.cols)
                                                     ? s : -1.0;
}

// Sort the columns of each problem by decreasing norm: perm[bz][j] is the
// source column of destination column j.  grid (batch), 1024 threads.
__global__ __launch_bounds__(1024) void qr_sort_perm(
    const double* __restrict__ nrm2, int64_t cols_p, int* __restrict__ perm) {
  __shared__ double sv[4096];
  __shared__ int si[4096];
  const int bz = blockIdx.x;
  int n2 = 1;
  while (n2 < cols_p) n2 <<= 1;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    sv[i] = i < cols_p ? nrm2[(int64_t)bz * cols_p + i] : -2.0;
    si[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool desc = (i & k) == 0;
          const double a = sv[i], b = sv[ixj];
          const bool swap = desc ? (a < b || (a == b && si[i] > si[ixj]))
                                 : (a > b || (a == b && si[i] < si[ixj]));
          if (swap) {
            sv[i] = b;
            sv[ixj] = a;
            const int t = si[i];
            si[i] = si[ixj];
            si[ixj] = t;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < cols_p; i += blockDim.x)
    perm[(int64_t)bz * cols_p + i] = si[i];
}

// dst[:, j] = src[:, perm[j]] (rows_p x cols_p per problem).
__global__ void qr_gather_cols(const double2* __restrict__ src,
                               double2* __restrict__ dst, int64_t rows_p,
                               int64_t cols_p, int64_t s,
                               const int* __restrict__ perm) {
  const int bz = blockIdx.y;
  const int64_t total = rows_p * cols_p;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % rows_p, j = idx / rows_p;
    dst[bz * s + idx] = src[bz * s + r + rows_p * perm[(int64_t)bz * cols_p + j]];
  }
}

// Wave-shuffle then LDS sum of N complex partials per thread; every thread
// gets the totals in out[].  scratch: NW waves x N double2 in LDS.
template <int N, int NW = 4>
__device__ __forceinline__ void block_sum_fast(double2 (&part)[N],
                                               double2 (&out)[N],
                                               double2* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // one butterfly step for all N values at a time, so the 2N shuffles of a
  // step are independent and their latencies overlap
  double x[N], y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    x[i] = part[i].x;
    y[i] = part[i].y;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double sx[N], sy[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      sx[i] = __shfl_xor(x[i], off, 64);
      sy[i] = __shfl_xor(y[i], off, 64);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      x[i] += sx[i];
      y[i] += sy[i];
    }
  }
  if (lane < N) {
    double vx = 0.0, vy = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i == lane) {
        vx = x[i];
        vy = y[i];
      }
    scratch[wave * N + lane] = make_double2(vx, vy);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double2 t = scratch[i];
    for (int w = 1; w < NW; ++w) {
      t.x += scratch[w * N + i].x;
      t.y += scratch[w * N + i].y;
    }
    out[i] = t;
  }
  __syncthreads();
}

// Factor panel k of every active problem (one workgroup of NT threads per
// problem).  Writes the reflectors below the diagonal of A, R on and above
// it, the explicit V (rows x NB, unit diagonal, zeros above) into Vb, V T^H
// into VTh and V T into VTn (both rows x NB), all with ld = lda.
// kLds: the panel (rows - r0 <= kLdsRows rows x NB) is held in LDS;
// otherwise it is updated in place in A (the workgroup's own stores are
// visible to it after each barrier).  Per column one norm reduction and one
// 16-wide reduction that yields both w_c = v_j^H x_c (c > j, for the
// reflector application) and Gv[i][j] = v_i^H v_j (i < j, for T).
constexpr int kLdsRows = 512;
template <bool kLds, int NT>
__global__ __launch_bounds__(NT) void qr_panel_t(
    double2* __restrict__ A, int64_t lda, int64_t sA, int k,
    const QrDims* __restrict__ dims, double2* __restrict__ Vb,
    double2* __restrict__ VTh, double2* __restrict__ VTn, int64_t sV,
    int64_t sVh) {
  constexpr int NW = NT / 64;
  const int bz = blockIdx.x;
  const QrDims d = dims[bz];
  const int64_t r0 = (int64_t)k * NB;
  if (!d.active || r0 >= d.cols) return;
  const int rows = (int)(d.rows - r0);  // rows of the panel
  const int ncp = (int)min((int64_t)NB, d.cols - r0);
  A += (int64_t)bz * sA + r0 + r0 * lda;  // panel origin
  Vb += (int64_t)bz * sV;
  VTh += (int64_t)bz * sVh;
  VTn += (int64_t)bz * sV;
  const int tid = threadIdx.x;
  __shared__ double2 P[kLds ? NB : 1][kLds ? kLdsRows + 1 : 1];
  __shared__ double2 scratch[NW * NB];
  __shared__ double2 tau_s[NB];
  __shared__ double2 T[NB][NB + 1];
  __shared__ double2 Gv[NB][NB + 1];
  // panel entry (column c, panel row r)
  auto at = [&](int c, int r) -> double2& {
    if constexpr (kLds)
      return P[c][r];
    else
      return A[r + (int64_t)c * lda];
  };
  if constexpr (kLds) {
    for (int idx = tid; idx < NB * rows; idx += NT) {
      const int r = idx % rows, c = idx / rows;
      P[c][r] = c < ncp ? A[r + (int64_t)c * lda] : make_double2(0.0, 0.0);
    }
  }
  if (tid < NB)
    for (int c = 0; c < NB; ++c) {
      Gv[tid][c] = make_double2(0.0, 0.0);
      T[tid][c] = make_double2(0.0, 0.0);
    }
  __syncthreads();
  for (int j = 0; j < ncp; ++j) {
    double2 part[1] = {make_double2(0.0, 0.0)};
    for (int r = j + 1 + tid; r < rows; r += NT) {
      const double2 v = at(j, r);
      part[0].x += v.x * v.x + v.y * v.y;
    }
    double2 tot[1];
    block_sum_fast<1, NW>(part, tot, scratch);
    const double xn2 = tot[0].x;
    const double2 alpha = at(j, j);
    double2 tau, scl;
    double beta;
    if (xn2 == 0.0 && alpha.y == 0.0) {
      tau = make_double2(0.0, 0.0);
      scl = make_double2(0.0, 0.0);
      beta = alpha.x;
    } else {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      beta = alpha.x >= 0.0 ? -nrm : nrm;
      tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 den = make_double2(alpha.x - beta, alpha.y);
      const double dd = den.x * den.x + den.y * den.y;
      scl = make_double2(den.x / dd, -den.y / dd);
    }
    if (tid == 0) tau_s[j] = tau;
    for (int r = j + 1 + tid; r < rows; r += NT) at(j, r) = cmul(at(j, r), scl);
    __syncthreads();
    if (tid == 0) at(j, j) = make_double2(beta, 0.0);
    // c > j: w_c = v_j^H x_c ; c < j: g_c = v_c^H v_j (v_c lives in column
    // c below its diagonal, and v_j is zero above row j)
    double2 wp[NB], w[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) wp[c] = make_double2(0.0, 0.0);
    for (int r = j + tid; r < rows; r += NT) {
      const double2 v = (r == j) ? make_double2(1.0, 0.0) : at(j, r);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c > j && c < ncp) {
          const double2 pr = cmulc(v, at(c, r));
          wp[c].x += pr.x;
          wp[c].y += pr.y;
        } else if (c < j) {
          const double2 pr = cmulc(at(c, r), v);
          wp[c].x += pr.x;
          wp[c].y += pr.y;
        }
      }
    }
    block_sum_fast<NB, NW>(wp, w, scratch);
    if (tid < j) {
      double2 g = make_double2(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c == tid) g = w[c];
      Gv[tid][j] = g;
    }
    if (tau.x != 0.0 || tau.y != 0.0) {
      for (int r = j + tid; r < rows; r += NT) {
        const double2 v = (r == j) ? make_double2(1.0, 0.0) : at(j, r);
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c > j && c < ncp) {
            const double2 u = cmul(v, cmulc(tau, w[c]));
            double2& x = at(c, r);
            x.x -= u.x;
            x.y -= u.y;
          }
      }
    }
    __syncthreads();
  }
  if constexpr (kLds) {
    // write R / reflectors back
    for (int idx = tid; idx < NB * rows; idx += NT) {
      const int r = idx % rows, c = idx / rows;
      if (c < ncp) A[r + (int64_t)c * lda] = P[c][r];
    }
  }
  // T (zlarft): column j = -tau_j T[0:j,0:j] Gv[0:j, j], T[j][j] = tau_j.
  // Row tid of column j only needs T[tid][l] for tid <= l < j.  Barriers
  // are unconditional: every thread takes both of them each step.
  for (int j = 0; j < ncp; ++j) {
    double2 q = make_double2(0.0, 0.0);
    const double2 tj = tau_s[j];
    if (tid < j) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < NB; ++l)
        if (l >= tid && l < j) {
          const double2 pr = cmul(T[tid][l], Gv[l][j]);
          acc.x += pr.x;
          acc.y += pr.y;
        }
      q = cmul(tj, acc);
    }
    __syncthreads();
    if (tid < j) T[tid][j] = make_double2(-q.x, -q.y);
    if (tid == j) T[j][j] = tj;
    __syncthreads();
  }
  // explicit V, V T^H and V T for every row of the problem (zero above r0).
  // Thread (j = tid % NB, part = tid / NB) owns output column j and keeps
  // column j of T and row j of T in registers.
  {
    const int j = tid % NB, part = tid / NB;
    double2 tcol[NB], trow[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      tcol[i] = T[i][j];
      const double2 t = T[j][i];
      trow[i] = make_double2(t.x, -t.y);  // conj(T[j][i])
    }
    for (int64_t r = part; r < d.rows; r += NT / NB) {
      const int64_t rl = r - r0;
      double2 ah = make_double2(0.0, 0.0), an = make_double2(0.0, 0.0);
      double2 vj = make_double2(0.0, 0.0);
      if (rl >= 0) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          double2 v = make_double2(0.0, 0.0);
          if (i < ncp && rl >= i)
            v = (rl == i) ? make_double2(1.0, 0.0) : at(i, (int)rl);
          if (i == j) vj = v;
          const double2 h = cmul(v, trow[i]);
          const double2 n = cmul(v, tcol[i]);
          ah.x += h.x;
          ah.y += h.y;
          an.x += n.x;
          an.y += n.y;
        }
      }
      Vb[r + j * lda] = vj;
      VTh[r + j * lda] = ah;
      VTn[r + j * lda] = an;
    }
  }
}

// Y = R^H (lower triangular n x n, ld ldy) from the upper triangle of the
// factored A (ld lda); n = cols per problem; rest zero.
__global__ void qr_r_adjoint(const double2* __restrict__ A, int64_t lda,
                             int64_t sA, double2* __restrict__ Y, int64_t ldy,
                             int64_t sY, int64_t ny_rows, int64_t ny_cols,
                             const QrDims* __restrict__ dims) {
  const int bz = blockIdx.y;
  const QrDims d = dims[bz];
  if (!d.active) return;  // leave the destination of other problems alone
  const int64_t total = ny_rows * ny_cols;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx % ny_rows, c = idx / ny_rows;
    double2 v = make_double2(0.0, 0.0);
    if (r < d.cols && c < d.cols && c <= r) {
      const double2 a = A[(int64_t)bz * sA + c + r * lda];  // R[c][r]
      v = make_double2(a.x, -a.y);
    }
    Y[(int64_t)bz * sY + r + c * ldy] = v;
  }
}

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                          8192));
}

}  // namespace

// ---------------------------------------------------------------------------
int32_t BatchedQR::factor(qtn_ctx* ctx, double2* A, int64_t lda, int64_t sA,
                          int64_t cols_p, const std::vector<QrDims>& hd,
                          bool keep_q) {
  hipStream_t s = ctx->stream;
  batch = (int)hd.size();
  ld = lda;
  stride = sA;
  cols_pad = cols_p;
  kept_q = keep_q;
  int64_t maxc = 0, maxr = 0;
  for (const auto& d : hd)
    if (d.active) {
      maxc = std::max(maxc, d.cols);
      maxr = std::max(maxr, d.rows);
    }
  npanel = (int)((maxc + NB - 1) / NB);
  void* p;
  QTN_TRY(ctx->buffer(slot_base + 0, (size_t)batch * sizeof(QrDims), &p));
  ddims = (QrDims*)p;
  QTN_CHECK_HIP(hipMemcpyAsync(ddims, hd.data(), batch * sizeof(QrDims),
                               hipMemcpyHostToDevice, s));
  // V and V T^H of the current panel; all panels' V and V T when kept
  const int64_t vs = lda * NB;
  const int nkeep = keep_q ? npanel : 1;
  QTN_TRY(ctx->buffer(slot_base + 1,
                      (size_t)batch * nkeep * vs * sizeof(double2), &p));
  V = (double2*)p;
  QTN_TRY(ctx->buffer(slot_base + 2, (size_t)batch * vs * sizeof(double2),
                      &p));
  VTh = (double2*)p;
  QTN_TRY(ctx->buffer(slot_base + 3,
                      (size_t)batch * nkeep * vs * sizeof(double2), &p));
  VTn = (double2*)p;
  QTN_TRY(ctx->buffer(slot_base + 4,
                      (size_t)batch * NB * cols_p * sizeof(double2), &p));
  Wb = (double2*)p;
  const int64_t vstride = keep_q ? (int64_t)npanel * vs : vs;
  std::vector<GemmDesc<double>> g((size_t)npanel * 2 * batch);
  std::vector<int64_t> maxn(npanel, 0), maxm(npanel, 0);
  for (int k = 0; k < npanel; ++k) {
    const int64_t r0 = (int64_t)k * NB;
    double2* Vk = V + (keep_q ? (int64_t)k * vs : 0);
    GemmDesc<double>* gk = g.data() + (size_t)k * 2 * batch;
    for (int b = 0; b < batch; ++b) {
      const QrDims& d = hd[b];
      GemmDesc<double>& e1 = gk[b];
      GemmDesc<double>& e2 = gk[batch + b];
      const int64_t c0 = r0 + NB;
      const bool on = d.active && c0 < d.cols && r0 < d.rows;
      const int64_t nt = on ? d.cols - c0 : 0;
      const int64_t mt = on ? d.rows - r0 : 0;
      maxn[k] = std::max(maxn[k], nt);
      maxm[k] = std::max(maxm[k], mt);
      const int64_t vb = (int64_t)b * vstride;
      e1.M = on ? NB : 0;
      e1.N = nt;
      e1.K = mt;
      e1.alpha = make_double2(1.0, 0.0);
      e1.beta = make_double2(0.0, 0.0);
      e1.A = Vk + vb + r0;
      e1.lda = lda;
      e1.B = A + (int64_t)b * sA + r0 + c0 * lda;
      e1.ldb = lda;
      e1.C = Wb + (int64_t)b * NB * cols_p;
      e1.ldc = NB;
      e2.M = mt;
      e2.N = nt;
      e2.K = on ? NB : 0;
      e2.alpha = make_double2(-1.0, 0.0);
      e2.beta = make_double2(1.0, 0.0);
      e2.A = VTh + (int64_t)b * vs + r0;
      e2.lda = lda;
      e2.B = Wb + (int64_t)b * NB * cols_p;
      e2.ldb = NB;
      e2.C = A + (int64_t)b * sA + r0 + c0 * lda;
      e2.ldc = lda;
    }
  }
  QTN_TRY(ctx->buffer(slot_base + 6, g.size() * sizeof(GemmDesc<double>),
                      &p));
  GemmDesc<double>* dall = (GemmDesc<double>*)p;
  QTN_CHECK_HIP(hipMemcpyAsync(dall, g.data(),
                               g.size() * sizeof(GemmDesc<double>),
                               hipMemcpyHostToDevice, s));
  for (int k = 0; k < npanel; ++k) {
    double2* Vk = V + (keep_q ? (int64_t)k * vs : 0);
    double2* VTnk = VTn + (keep_q ? (int64_t)k * vs : 0);
    if (maxr - (int64_t)k * NB <= kLdsRows)
      hipLaunchKernelGGL((qr_panel_t<true, 256>), dim3(batch), dim3(256), 0,
                         s, A, lda, sA, k, (const QrDims*)ddims, Vk, VTh,
                         VTnk, vstride, vs);
    else
      hipLaunchKernelGGL((qr_panel_t<false, 512>), dim3(batch), dim3(512), 0,
                         s, A, lda, sA, k, (const QrDims*)ddims, Vk, VTh,
                         VTnk, vstride, vs);
    QTN_CHECK_HIP(hipGetLastError());
    if (maxn[k] == 0) continue;
    GemmDesc<double>* dk = dall + (size_t)k * 2 * batch;
    QTN_TRY(launch_zgemm_grouped(s, 2, 0, dk, batch, NB, maxn[k]));
    QTN_TRY(launch_zgemm_grouped(s, 0, 0, dk + batch, batch, maxm[k],
                                 maxn[k]));
  }
  // the host descriptor vector dies here: make sure the copy finished
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
}

int32_t BatchedQR::apply_q(qtn_ctx* ctx, double2* U,
                           const std::vector<int64_t>& ldu, int64_t sU,
                           const std::vector<int64_t>& ucols,
                           const std::vector<QrDims>& hd) {
  // U <- Q U with Q = H_1 .. H_n: panels applied last to first,
  // U -= (V T) (V^H U)
  if (!kept_q) {
    set_error("qr: reflectors were not kept");
    return QTN_ERR_ARG;
  }
  hipStream_t s = ctx->stream;
  const int64_t vs = ld * NB;
  const int64_t vstride = (int64_t)npanel * vs;
  void* p;
  int64_t maxu = 0;
  for (auto c : ucols) maxu = std::max(maxu, c);
  if (maxu == 0) return QTN_OK;
  QTN_TRY(ctx->buffer(slot_base + 4, (size_t)batch * NB * std::max(maxu, cols_pad) *
                                         sizeof(double2),
                      &p));
  Wb = (double2*)p;
  QTN_TRY(ctx->buffer(slot_base + 7,
                      (size_t)npanel * 2 * batch * sizeof(GemmDesc<double>),
                      &p));
  GemmDesc<double>* dall = (GemmDesc<double>*)p;
  std::vector<GemmDesc<double>> g((size_t)npanel * 2 * batch);
  for (int k = npanel - 1; k >= 0; --k) {
    const int64_t r0 = (int64_t)k * NB;
    GemmDesc<double>* gk = g.data() + (size_t)k * 2 * batch;
    for (int b = 0; b < batch; ++b) {
      const QrDims& d = hd[b];
      const bool on = d.active && r0 < d.cols && ucols[b] > 0;
      const int64_t mt = on ? d.rows - r0 : 0;
      GemmDesc<double>& e1 = gk[b];
      GemmDesc<double>& e2 = gk[batch + b];
      e1.M = on ? NB : 0;
      e1.N = on ? ucols[b] : 0;
      e1.K = mt;
      e1.alpha = make_double2(1.0, 0.0);
      e1.beta = make_double2(0.0, 0.0);
      e1.A = V + (int64_t)b * vstride + (int64_t)k * vs + r0;
      e1.lda = ld;
      e1.B = U + (int64_t)b * sU + r0;
      e1.ldb = ldu[b];
      e1.C = Wb + (int64_t)b * NB * std::max(maxu, cols_pad);
      e1.ldc = NB;
      e2.M = mt;
      e2.N = on ? ucols[b] : 0;
      e2.K = on ? NB : 0;
      e2.alpha = make_double2(-1.0, 0.0);
      e2.beta = make_double2(1.0, 0.0);
      e2.A = VTn + (int64_t)b * vstride + (int64_t)k * vs + r0;
      e2.lda = ld;
      e2.B = e1.C;
      e2.ldb = NB;
      e2.C = U + (int64_t)b * sU + r0;
      e2.ldc = ldu[b];
    }
  }
  QTN_CHECK_HIP(hipMemcpyAsync(dall, g.data(),
                               g.size() * sizeof(GemmDesc<double>),
                               hipMemcpyHostToDevice, s));
  int64_t maxr = 0;
  for (const auto& d : hd)
    if (d.active) maxr = std::max(maxr, d.rows);
  for (int k = npanel - 1; k >= 0; --k) {
    GemmDesc<double>* dk = dall + (size_t)k * 2 * batch;
    QTN_TRY(launch_zgemm_grouped(s, 2, 0, dk, batch, NB, maxu));
    QTN_TRY(launch_zgemm_grouped(s, 0, 0, dk + batch, batch, maxr, maxu));
  }
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
}

int32_t qr_sort_columns(qtn_ctx* ctx, const double2* src, double2* dst,
                        int64_t rows_p, int64_t cols_p, int64_t s, int batch,
                        const QrDims* d_dims, int slot_norm, int slot_perm,
                        int** perm_out) {
  hipStream_t st = ctx->stream;
  if (cols_p > 4096) {
    set_error("qr: more than 4096 columns are not supported");
    return QTN_ERR_ARG;
  }
  void *pn, *pp;
  QTN_TRY(ctx->buffer(slot_norm, (size_t)batch * cols_p * sizeof(double),
                      &pn));
  QTN_TRY(ctx->buffer(slot_perm, (size_t)batch * cols_p * sizeof(int), &pp));
  hipLaunchKernelGGL(qr_colnorm2, dim3((unsigned)((cols_p + 3) / 4), batch),
                     dim3(256), 0, st, src, rows_p, s, cols_p, d_dims,
                     (double*)pn);
  hipLaunchKernelGGL(qr_sort_perm, dim3(batch), dim3(1024), 0, st,
                     (const double*)pn, cols_p, (int*)pp);
  hipLaunchKernelGGL(qr_gather_cols, dim3(grid_for(rows_p * cols_p), batch),
                     dim3(256), 0, st, src, dst, rows_p, cols_p, s,
                     (const int*)pp);
  QTN_CHECK_HIP(hipGetLastError());
  if (perm_out) *perm_out = (int*)pp;
  return QTN_OK;
}

int32_t qr_r_to_adjoint(qtn_ctx* ctx, const double2* A, int64_t lda,
                        int64_t sA, double2* Y, int64_t ldy, int64_t sY,
                        int64_t ny_rows, int64_t ny_cols, int batch,
                        const QrDims* d_dims) {
  hipLaunchKernelGGL(qr_r_adjoint, dim3(grid_for(ny_rows * ny_cols), batch),
                     dim3(256), 0, ctx->stream, A, lda, sA, Y, ldy, sY,
                     ny_rows, ny_cols, d_dims);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

}  // namespace qtn
