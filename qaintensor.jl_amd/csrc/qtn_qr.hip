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
    int64_t sVh, int write_vtn) {
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
      if (write_vtn) VTn[r + j * lda] = an;
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

// One reduce-scatter step of a wave: the lanes with bit `H` set keep the
// upper half of their 2H values, the others the lower half, each adding the
// half its partner lane sends.
template <int H>
__device__ __forceinline__ void rs_step(const double (&in)[2 * H],
                                        double (&out)[H]) {
  const bool up = (threadIdx.x & H) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double keep = up ? in[i + H] : in[i];
    const double send = up ? in[i] : in[i + H];
    out[i] = keep + __shfl_xor(send, H, 64);
  }
}

// Wave sums of 64 values at once by reduce-scatter (63 shuffles instead of
// 6 x 64): lane L returns the sum of v[L] over the wave's 64 lanes.
__device__ __forceinline__ double wave_reduce_scatter64(const double (&v)[64]) {
  double a[32], b[16], c[8], d[4], e[2], f[1];
  rs_step<32>(v, a);
  rs_step<16>(a, b);
  rs_step<8>(b, c);
  rs_step<4>(c, d);
  rs_step<2>(d, e);
  rs_step<1>(e, f);
  return f[0];
}

// Sum of one complex value over the workgroup, wave by wave: lane 0 of
// every wave leaves its wave's total in red[wave].  (Callers barrier and
// add the NT / 64 wave totals.)
__device__ __forceinline__ void wave_sum_to(double2 v, double2* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
}

// Panel k of every active problem, left-looking, for panels too tall for
// LDS (one workgroup of NT threads per problem; the panel stays in A).
// Column j first receives the reflectors of columns 0..j-1 in block form,
//   x_j <- x_j - V T^H (V^H x_j),
// then its own reflector (zlarfg).  Per column that reads the j earlier
// reflector columns twice instead of streaming the whole panel twice, as a
// right-looking panel does; v_j^H v_l for T rides in the next column's
// first reduction.  Outputs as qr_panel_lds.
template <int NT>
__global__ __launch_bounds__(NT) void qr_panel_left(
    double2* __restrict__ A, int64_t lda, int64_t sA, int k,
    const QrDims* __restrict__ dims, double2* __restrict__ Vb,
    double2* __restrict__ VTh, double2* __restrict__ VTn, int64_t sV,
    int64_t sVh, int write_vtn) {
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
  static_assert(4 * NB == 64, "one reduce-scatter value per lane");
  __shared__ double redd[64][NW];  // wave totals of w, g (re / im)
  __shared__ double2 red[NW];      // wave totals of the norm
  __shared__ double2 tot[2 * NB];
  __shared__ double2 tau_s[NB];
  __shared__ double2 T[NB][NB + 1];
  __shared__ double2 y_s[NB];
  __shared__ double2 alpha_s;
  if (tid < NB)
    for (int c = 0; c < NB; ++c) T[tid][c] = make_double2(0.0, 0.0);
  // v_l at panel row r (l < the current column): unit at r == l, the
  // stored reflector below it, zero above.  The load is unconditional (R
  // sits above the reflector), so a row's loads issue together.
  auto vcol = [&](int l, int r) -> double2 {
    const double2 a = A[r + (int64_t)l * lda];
    const double one = r == l ? 1.0 : 0.0;
    return r > l ? a : make_double2(one, 0.0);
  };
  // Column j of T from g_l = v_l^H v_j (l < j) in tot[NB + l]:
  // T[0:j, j] = -tau_j T[0:j, 0:j] g, T[j][j] = tau_j.  Every thread takes
  // both barriers.
  auto t_column = [&](int j) {
    double2 q = make_double2(0.0, 0.0);
    const double2 tj = tau_s[j];
    if (tid < j) {
      double2 acc = make_double2(0.0, 0.0);
      for (int l = tid; l < j; ++l) {
        const double2 pr = cmul(T[tid][l], tot[NB + l]);
        acc.x += pr.x;
        acc.y += pr.y;
      }
      q = cmul(tj, acc);
    }
    __syncthreads();
    if (tid < j) T[tid][j] = make_double2(-q.x, -q.y);
    if (tid == j) T[j][j] = tj;
    __syncthreads();
  };
  __syncthreads();
  for (int j = 0; j <= ncp; ++j) {
    // 1. w_l = v_l^H x_j (l < j) and g_l = v_l^H v_{j-1} (l < j - 1); for
    //    j == ncp only g (the last column of T)
    {
      double2 w[NB], g[NB];
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        w[l] = make_double2(0.0, 0.0);
        g[l] = make_double2(0.0, 0.0);
      }
      if (j > 0) {
        for (int r = tid; r < rows; r += NT) {
          const double2 x = A[r + (int64_t)min(j, ncp - 1) * lda];
          const double2 xj = j < ncp ? x : make_double2(0.0, 0.0);
          const double2 vp = vcol(j - 1, r);
#pragma unroll
          for (int l = 0; l < NB; ++l)
            if (l < j) {  // uniform
              const double2 v = vcol(l, r);
              const double2 a = cmulc(v, xj);
              w[l].x += a.x;
              w[l].y += a.y;
              if (l < j - 1) {
                const double2 b = cmulc(v, vp);
                g[l].x += b.x;
                g[l].y += b.y;
              }
            }
        }
      }
      // wave sums of the 64 doubles (w then g, re / im interleaved): lane L
      // of each wave gets value L
      double vals[64];
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        vals[2 * l] = w[l].x;
        vals[2 * l + 1] = w[l].y;
        vals[2 * NB + 2 * l] = g[l].x;
        vals[2 * NB + 2 * l + 1] = g[l].y;
      }
      redd[tid & 63][tid >> 6] = wave_reduce_scatter64(vals);
      __syncthreads();
      if (tid < 2 * NB) {
        double tx = 0.0, ty = 0.0;
        for (int wv = 0; wv < NW; ++wv) {
          tx += redd[2 * tid][wv];
          ty += redd[2 * tid + 1][wv];
        }
        tot[tid] = make_double2(tx, ty);
      }
      __syncthreads();
    }
    if (j > 0) t_column(j - 1);
    if (j == ncp) break;
    // 2. y = T^H w, then x_j -= V y and the norm of its tail
    if (tid < j) {
      double2 acc = make_double2(0.0, 0.0);
      for (int l = 0; l <= tid; ++l) {
        const double2 pr = cmulc(T[l][tid], tot[l]);  // conj(T[l][a]) w_l
        acc.x += pr.x;
        acc.y += pr.y;
      }
      y_s[tid] = acc;
    }
    __syncthreads();
    double2 nrm = make_double2(0.0, 0.0);
#pragma unroll 3
    for (int r = tid; r < rows; r += NT) {
      double2 x = A[r + (int64_t)j * lda];
      if (j > 0) {
#pragma unroll
        for (int l = 0; l < NB; ++l)
          if (l < j) {  // uniform
            const double2 u = cmul(vcol(l, r), y_s[l]);
            x.x -= u.x;
            x.y -= u.y;
          }
        A[r + (int64_t)j * lda] = x;
      }
      if (r == j) alpha_s = x;
      if (r > j) nrm.x += x.x * x.x + x.y * x.y;
    }
    wave_sum_to(nrm, red);
    __syncthreads();
    double xn2 = 0.0;
    for (int wv = 0; wv < NW; ++wv) xn2 += red[wv].x;
    // 3. reflector (zlarfg): beta, tau, v = x_tail / (alpha - beta)
    const double2 alpha = alpha_s;
    double2 tau, scl;
    double beta;
    if (xn2 == 0.0 && alpha.y == 0.0) {
      tau = make_double2(0.0, 0.0);
      scl = make_double2(0.0, 0.0);
      beta = alpha.x;
    } else {
      const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      beta = alpha.x >= 0.0 ? -nr : nr;
      tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 den = make_double2(alpha.x - beta, alpha.y);
      const double dd = den.x * den.x + den.y * den.y;
      scl = make_double2(den.x / dd, -den.y / dd);
    }
    if (tid == 0) tau_s[j] = tau;
    for (int r = j + 1 + tid; r < rows; r += NT)
      A[r + (int64_t)j * lda] = cmul(A[r + (int64_t)j * lda], scl);
    if (tid == 0) A[j + (int64_t)j * lda] = make_double2(beta, 0.0);
    __syncthreads();
  }
  // explicit V, V T^H and V T, one row per thread (zero above r0)
  for (int64_t r = tid; r < d.rows; r += NT) {
    const int64_t rl = r - r0;
    double2 v[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int ic = min(i, ncp - 1);
      const double2 vi = vcol(ic, (int)max<int64_t>(rl, 0));
      v[i] = (rl >= 0 && i < ncp) ? vi : make_double2(0.0, 0.0);
    }
#pragma unroll 1
    for (int jj = 0; jj < NB; ++jj) {
      double2 ah = make_double2(0.0, 0.0), an = make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const double2 t = T[i][jj], th = T[jj][i];
        const double2 n = cmul(v[i], t);
        const double2 h = cmul(v[i], make_double2(th.x, -th.y));
        an.x += n.x;
        an.y += n.y;
        ah.x += h.x;
        ah.y += h.y;
      }
      VTh[r + jj * lda] = ah;
      if (write_vtn) VTn[r + jj * lda] = an;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) Vb[r + i * lda] = v[i];
  }
}

// out[b][i] = sum over s < S of parts[b][s][i] for i < count; parts and
// out hold `stride` entries per (problem, slice) and per problem (one
// problem per grid row).  The partials of a split-K reflector product.
__global__ void qr_sum_parts(const double2* __restrict__ parts, int S,
                             int64_t stride, int64_t count,
                             double2* __restrict__ out) {
  const int64_t b = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int sl = 0; sl < S; ++sl) {
      const double2 v = parts[(b * S + sl) * stride + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    out[b * stride + i] = acc;
  }
}

// Slices of the row (K) dimension for W = V^H A (NB x N, K = rows): its
// few output tiles alone would leave most CUs idle, so the rows (when
// there are at least 1024) are split until about 512 workgroups run, each
// slice at least 128 rows long.
int split_k(int64_t maxn, int batch, int64_t maxk) {
  if (maxk < 1024) return 1;  // short reductions: the extra launch costs more
  const int64_t tiles = std::max<int64_t>(1, ((maxn + 31) / 32) * batch);
  int64_t sk = std::max<int64_t>(1, 512 / tiles);
  sk = std::min<int64_t>(sk, std::max<int64_t>(1, maxk / 128));
  return (int)std::min<int64_t>(sk, 16);
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
  // V of the last two panels (all panels, with V T, when kept) and V T^H
  // of the last two: panel k+1 is factored while the update of panel k
  // still reads its V and V T^H
  const int64_t vs = lda * NB;
  const int nkeep = keep_q ? npanel : 2;
  QTN_TRY(ctx->buffer(slot_base + 1,
                      (size_t)batch * nkeep * vs * sizeof(double2), &p));
  V = (double2*)p;
  QTN_TRY(ctx->buffer(slot_base + 2, (size_t)batch * 2 * vs * sizeof(double2),
                      &p));
  VTh = (double2*)p;
  QTN_TRY(ctx->buffer(slot_base + 3,
                      (size_t)batch * (keep_q ? npanel : 1) * vs *
                          sizeof(double2),
                      &p));
  VTn = (double2*)p;
  const int64_t vstride = (int64_t)nkeep * vs;  // per problem
  auto vslot = [&](int k) { return keep_q ? k : (k & 1); };
  // Look-ahead (panels too tall for LDS, whose one-workgroup factorisation
  // dominates): the update of panel k is split into the next panel's
  // columns (main stream, before panel k+1) and the rest (side stream,
  // behind panel k+1).  Otherwise one update of every trailing column on
  // the main stream.  Each part forms W = V^H A_part in `sk` row slices
  // (partials summed into its own W), then A_part -= (V T^H) W.
  const bool lookahead = maxr > kLdsRows && maxc > 2 * NB;
  const int sk_a = lookahead ? split_k(NB, batch, maxr)
                             : split_k(maxc, batch, maxr);
  const int sk_b = split_k(maxc, batch, maxr);
  const int64_t wsz = (int64_t)NB * cols_p;
  QTN_TRY(ctx->buffer(slot_base + 4, (size_t)2 * batch * wsz * sizeof(double2),
                      &p));
  Wb = (double2*)p;
  double2* const Wpart[2] = {Wb, Wb + (int64_t)batch * wsz};
  const int skmax = std::max(sk_a, sk_b);
  double2* Wp = nullptr;
  if (skmax > 1) {
    QTN_TRY(ctx->buffer(slot_base + 5,
                        (size_t)2 * batch * skmax * wsz * sizeof(double2),
                        &p));
    Wp = (double2*)p;
  }
  double2* const Wpp[2] = {Wp, Wp ? Wp + (int64_t)batch * skmax * wsz
                                  : nullptr};
  const int sks[2] = {sk_a, sk_b};
  const int per = sk_a + sk_b + 2;  // descriptors per problem and panel
  std::vector<GemmDesc<double>> g((size_t)npanel * per * batch);
  std::vector<int64_t> maxn(2 * npanel, 0), maxm(npanel, 0);
  for (int k = 0; k < npanel; ++k) {
    const int64_t r0 = (int64_t)k * NB;
    double2* Vk = V + (int64_t)vslot(k) * vs;
    double2* VThk = VTh + (int64_t)(k & 1) * vs;
    GemmDesc<double>* gk = g.data() + (size_t)k * per * batch;
    for (int b = 0; b < batch; ++b) {
      const QrDims& d = hd[b];
      const int64_t c0 = r0 + NB;
      const bool on = d.active && c0 < d.cols && r0 < d.rows;
      const int64_t nt = on ? d.cols - c0 : 0;
      const int64_t mt = on ? d.rows - r0 : 0;
      maxm[k] = std::max(maxm[k], mt);
      const int64_t ks_all = mt;
      for (int part = 0; part < 2; ++part) {
        // part 0: columns c0 .. c0 + NB (the next panel), or every trailing
        // column without look-ahead; part 1: the rest
        const int64_t cs = part == 0 ? c0 : c0 + NB;
        const int64_t n = !lookahead ? (part == 0 ? nt : 0)
                          : part == 0 ? std::min<int64_t>(NB, nt)
                                      : std::max<int64_t>(0, nt - NB);
        maxn[2 * k + part] = std::max(maxn[2 * k + part], n);
        const int sk = sks[part];
        GemmDesc<double>* gp = gk + (size_t)(part == 0 ? 0 : sk_a + 1) * batch;
        const int64_t ks = (ks_all + sk - 1) / sk;
        for (int sl = 0; sl < sk; ++sl) {
          const int64_t k0 = std::min<int64_t>(mt, sl * ks);
          const int64_t k1 = std::min<int64_t>(mt, k0 + ks);
          GemmDesc<double>& e1 = gp[(size_t)sl * batch + b];
          e1.M = (on && n > 0) ? NB : 0;
          e1.N = n;
          e1.K = k1 - k0;  // an empty slice still writes its zero partial
          e1.alpha = make_double2(1.0, 0.0);
          e1.beta = make_double2(0.0, 0.0);
          e1.A = Vk + (int64_t)b * vstride + r0 + k0;
          e1.lda = lda;
          e1.B = A + (int64_t)b * sA + r0 + k0 + cs * lda;
          e1.ldb = lda;
          e1.C = sk > 1 ? Wpp[part] + ((int64_t)b * sk + sl) * wsz
                        : Wpart[part] + (int64_t)b * wsz;
          e1.ldc = NB;
        }
        GemmDesc<double>& e2 = gp[(size_t)sk * batch + b];
        e2.M = n > 0 ? mt : 0;
        e2.N = n;
        e2.K = (on && n > 0) ? NB : 0;
        e2.alpha = make_double2(-1.0, 0.0);
        e2.beta = make_double2(1.0, 0.0);
        e2.A = VThk + (int64_t)b * 2 * vs + r0;
        e2.lda = lda;
        e2.B = Wpart[part] + (int64_t)b * wsz;
        e2.ldb = NB;
        e2.C = A + (int64_t)b * sA + r0 + cs * lda;
        e2.ldc = lda;
      }
    }
  }
  QTN_TRY(ctx->buffer(slot_base + 6, g.size() * sizeof(GemmDesc<double>),
                      &p));
  GemmDesc<double>* dall = (GemmDesc<double>*)p;
  QTN_CHECK_HIP(hipMemcpyAsync(dall, g.data(),
                               g.size() * sizeof(GemmDesc<double>),
                               hipMemcpyHostToDevice, s));
  hipStream_t side = nullptr;
  if (lookahead) QTN_TRY(ctx->side_stream(&side));
  // side_ev[0]: panel k factored (main); side_ev[2 + (k & 1)]: the rest of
  // panel k's update done (side)
  hipEvent_t ev_panel = ctx->side_ev[0];
  auto ev_rest = [&](int k) { return ctx->side_ev[2 + (k & 1)]; };
  // one part of panel k's update on stream st
  auto update = [&](hipStream_t st, int k, int part) -> int32_t {
    const int64_t n = maxn[2 * k + part];
    if (n == 0) return QTN_OK;
    const int sk = sks[part];
    GemmDesc<double>* dk =
        dall + (size_t)k * per * batch + (size_t)(part == 0 ? 0 : sk_a + 1) * batch;
    QTN_TRY(launch_zgemm_grouped(st, 2, 0, dk, sk * batch, NB, n));
    if (sk > 1) {
      hipLaunchKernelGGL(qr_sum_parts,
                         dim3(std::max(1u, (unsigned)std::min<int64_t>(
                                               (NB * n + 255) / 256, 4096)),
                              batch),
                         dim3(256), 0, st, (const double2*)Wpp[part], sk,
                         wsz, (int64_t)NB * n, Wpart[part]);
      QTN_CHECK_HIP(hipGetLastError());
    }
    return launch_zgemm_grouped(st, 0, 0, dk + (size_t)sk * batch, batch,
                                maxm[k], n);
  };
  for (int k = 0; k < npanel; ++k) {
    double2* Vk = V + (int64_t)vslot(k) * vs;
    double2* VThk = VTh + (int64_t)(k & 1) * vs;
    double2* VTnk = VTn + (keep_q ? (int64_t)k * vs : 0);
    if (maxr - (int64_t)k * NB <= kLdsRows)
      hipLaunchKernelGGL((qr_panel_t<true, 256>), dim3(batch), dim3(256), 0,
                         s, A, lda, sA, k, (const QrDims*)ddims, Vk, VThk,
                         VTnk, vstride, 2 * vs, keep_q ? 1 : 0);
    else
      hipLaunchKernelGGL((qr_panel_left<512>), dim3(batch), dim3(512), 0, s,
                         A, lda, sA, k, (const QrDims*)ddims, Vk, VThk, VTnk,
                         vstride, 2 * vs, keep_q ? 1 : 0);
    QTN_CHECK_HIP(hipGetLastError());
    if (!lookahead) {
      QTN_TRY(update(s, k, 0));
      continue;
    }
    // the rest of the update, behind the next panel
    QTN_CHECK_HIP(hipEventRecord(ev_panel, s));
    QTN_CHECK_HIP(hipStreamWaitEvent(side, ev_panel, 0));
    QTN_TRY(update(side, k, 1));
    QTN_CHECK_HIP(hipEventRecord(ev_rest(k), side));
    // the next panel's columns: after the rest of panel k-1's update has
    // reached them (and panel k-1's V / V T^H slots are free again)
    if (k >= 1) QTN_CHECK_HIP(hipStreamWaitEvent(s, ev_rest(k - 1), 0));
    QTN_TRY(update(s, k, 0));
  }
  if (lookahead && npanel >= 1)
    QTN_CHECK_HIP(hipStreamWaitEvent(s, ev_rest(npanel - 1), 0));
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
  int64_t maxr = 0;
  for (const auto& d : hd)
    if (d.active) maxr = std::max(maxr, d.rows);
  const int64_t wsz = (int64_t)NB * std::max(maxu, cols_pad);
  QTN_TRY(ctx->buffer(slot_base + 4, (size_t)batch * wsz * sizeof(double2),
                      &p));
  Wb = (double2*)p;
  // W = V^H U in `sk` row slices (partials summed into Wb), then
  // U -= (V T) W
  const int sk = split_k(maxu, batch, maxr);
  double2* Wp = nullptr;
  if (sk > 1) {
    QTN_TRY(ctx->buffer(slot_base + 5,
                        (size_t)batch * sk * wsz * sizeof(double2), &p));
    Wp = (double2*)p;
  }
  const int per = sk + 1;
  QTN_TRY(ctx->buffer(slot_base + 7,
                      (size_t)npanel * per * batch * sizeof(GemmDesc<double>),
                      &p));
  GemmDesc<double>* dall = (GemmDesc<double>*)p;
  std::vector<GemmDesc<double>> g((size_t)npanel * per * batch);
  for (int k = npanel - 1; k >= 0; --k) {
    const int64_t r0 = (int64_t)k * NB;
    GemmDesc<double>* gk = g.data() + (size_t)k * per * batch;
    for (int b = 0; b < batch; ++b) {
      const QrDims& d = hd[b];
      const bool on = d.active && r0 < d.cols && ucols[b] > 0;
      const int64_t mt = on ? d.rows - r0 : 0;
      const int64_t ks = (mt + sk - 1) / sk;
      for (int sl = 0; sl < sk; ++sl) {
        const int64_t k0 = std::min<int64_t>(mt, sl * ks);
        const int64_t k1 = std::min<int64_t>(mt, k0 + ks);
        GemmDesc<double>& e1 = gk[(size_t)sl * batch + b];
        e1.M = on ? NB : 0;
        e1.N = on ? ucols[b] : 0;
        e1.K = k1 - k0;  // an empty slice still writes its zero partial
        e1.alpha = make_double2(1.0, 0.0);
        e1.beta = make_double2(0.0, 0.0);
        e1.A = V + (int64_t)b * vstride + (int64_t)k * vs + r0 + k0;
        e1.lda = ld;
        e1.B = U + (int64_t)b * sU + r0 + k0;
        e1.ldb = ldu[b];
        e1.C = sk > 1 ? Wp + ((int64_t)b * sk + sl) * wsz
                      : Wb + (int64_t)b * wsz;
        e1.ldc = NB;
      }
      GemmDesc<double>& e2 = gk[(size_t)sk * batch + b];
      e2.M = mt;
      e2.N = on ? ucols[b] : 0;
      e2.K = on ? NB : 0;
      e2.alpha = make_double2(-1.0, 0.0);
      e2.beta = make_double2(1.0, 0.0);
      e2.A = VTn + (int64_t)b * vstride + (int64_t)k * vs + r0;
      e2.lda = ld;
      e2.B = Wb + (int64_t)b * wsz;
      e2.ldb = NB;
      e2.C = U + (int64_t)b * sU + r0;
      e2.ldc = ldu[b];
    }
  }
  QTN_CHECK_HIP(hipMemcpyAsync(dall, g.data(),
                               g.size() * sizeof(GemmDesc<double>),
                               hipMemcpyHostToDevice, s));
  for (int k = npanel - 1; k >= 0; --k) {
    GemmDesc<double>* dk = dall + (size_t)k * per * batch;
    QTN_TRY(launch_zgemm_grouped(s, 2, 0, dk, sk * batch, NB, maxu));
    if (sk > 1) {
      hipLaunchKernelGGL(qr_sum_parts,
                         dim3(std::max(1u, (unsigned)std::min<int64_t>(
                                               (NB * maxu + 255) / 256, 4096)),
                              batch),
                         dim3(256), 0, s, (const double2*)Wp, sk, wsz,
                         (int64_t)NB * maxu, Wb);
      QTN_CHECK_HIP(hipGetLastError());
    }
    QTN_TRY(launch_zgemm_grouped(s, 0, 0, dk + (size_t)sk * batch, batch,
                                 maxr, maxu));
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
