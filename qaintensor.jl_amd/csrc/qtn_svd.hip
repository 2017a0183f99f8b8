// Batched block one-sided Jacobi SVD (Hestenes) for complex FP64, with the
// truncation rule of Qaintensor's contract_svd (src/svd.jl:29-33).
//
// Method (one-sided Jacobi: Hestenes 1958; block form with Gram-matrix
// rotations as in Boukaram, Turkiyyah, Ltaief, Keyes, "Batched QR and SVD
// algorithms on GPUs with applications in hierarchical matrix compression",
// Parallel Computing 74 (2018)):
//   X (rows x cols, zero padded to multiples of 64) is orthogonalised by
//   right-multiplication with unitary rotations until every pair of columns
//   satisfies |x_i^H x_j| <= tol * |x_i| |x_j|.  Columns are grouped into
//   2p blocks of 32; a sweep is 2p-1 rounds of the round-robin tournament,
//   each round handling p disjoint block pairs.  For a block pair (64
//   columns):
//     1. jac_gram   : G = X_IJ^H X_IJ (64x64) on MFMA, split over row chunks
//     2. jac_inner  : one cyclic two-sided Jacobi sweep on G in LDS (Hermitian
//                     2x2 rotations, Golub & Van Loan 8.5), accumulating Q
//     3. jac_update : X_IJ <- X_IJ Q on MFMA (and W_IJ <- W_IJ Q if vectors
//                     of the other side are needed)
//   The pair is skipped when G already meets the tolerance, so convergence
//   is always decided on a Gram matrix freshly computed from the columns.
// At convergence sigma_j = |x_j|; X W = U_X diag(sigma) with W unitary.
// All updates are unitary rotations applied to X, so the singular values are
// backward stable (absolute error O(eps |X|_F)), as for LAPACK's gesdd.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include <hip/hip_ext.h>

#include "qtn_internal.hpp"
#include "qtn_svd.hpp"

namespace qtn {
namespace {

// Every value the launches of one sweep take (JacobiEngine::run_t): the
// key of its cached graph.  Zero-filled before use, so padding compares.
struct SweepKey {
  int dbl, w, fused, cross, batch, p, n_inner, nchunk_x, nchunk_w,
      per_problem_w;
  int64_t ldx, sx, sw, ldw;
  double tol, conv2;
  const void* ptr[11];
};

// QTN_SWEEP_GRAPHS=0 launches every sweep directly (read at every call).
bool sweep_graphs_enabled() {
  const char* e = getenv("QTN_SWEEP_GRAPHS");
  return !(e && *e && atoi(e) == 0);
}

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// Precision traits of the sweep kernels.  Both MFMA forms take A[i][k] from
// lane i + 16k and B[k][j] from lane j + 16k; the f64 result element e of
// lane l is D[4*e + l/16][l%16], the f32 one D[4*(l/16) + e][l%16].
template <typename R>
struct JT;
template <>
struct JT<double> {
  typedef double2 c_t;
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) {
    return 4 * e + (lane >> 4);
  }
  static __device__ __forceinline__ c_t mk(double re, double im) {
    return make_double2(re, im);
  }
  static __device__ __forceinline__ double rsq(double x) { return rsqrt(x); }
  static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
  static __device__ __forceinline__ double ab(double x) { return fabs(x); }
  static __device__ __forceinline__ double dv(double a, double b) {
    return a / b;
  }
  // Rotation parameters (make_rot): the hardware v_rsq_f64 / v_rcp_f64
  // estimates refined by two Newton steps each (quadratic convergence from
  // their ~2^-26 relative error to below FP64 rounding), which shortens the
  // dependent chain of the inner sweep's rounds against the correctly
  // rounded library sequences.  Arguments are positive and normal (the
  // rotation guards them).
  static __device__ __forceinline__ double rsq_r(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
  }
  static __device__ __forceinline__ double sq_r(double x) {
    return x * rsq_r(x);
  }
  static __device__ __forceinline__ double dv_r(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return a * r;
  }
};
template <>
struct JT<float> {
  typedef float2 c_t;
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) {
    return 4 * (lane >> 4) + e;
  }
  static __device__ __forceinline__ c_t mk(float re, float im) {
    return make_float2(re, im);
  }
  // The FP32 sweeps only produce candidates that are re-orthogonalised and
  // re-measured in FP64, so the rotation parameters use the one-instruction
  // hardware forms (v_rsq_f32, v_sqrt_f32, v_rcp_f32, about 1 ulp) instead
  // of the correctly rounded library sequences: the inner sweep is bound by
  // its per-round instruction count.
  static __device__ __forceinline__ float rsq(float x) {
    return __builtin_amdgcn_rsqf(x);
  }
  static __device__ __forceinline__ float sq(float x) {
    return __builtin_amdgcn_sqrtf(x);
  }
  static __device__ __forceinline__ float ab(float x) { return fabsf(x); }
  static __device__ __forceinline__ float dv(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float rsq_r(float x) { return rsq(x); }
  static __device__ __forceinline__ float sq_r(float x) { return sq(x); }
  static __device__ __forceinline__ float dv_r(float a, float b) {
    return dv(a, b);
  }
};

// Round-robin (circle method) position -> block for 2p blocks.
__device__ __forceinline__ int rr_block(int pos, int round, int nb) {
  return pos == 0 ? 0 : 1 + (pos - 1 + round) % (nb - 1);
}

// Global column of local column c (0..W-1) of the block pair (I, J) for
// blocks of W/2 columns.
template <int W>
__device__ __forceinline__ int64_t pair_col(int c, int I, int J) {
  return c < W / 2 ? (int64_t)I * (W / 2) + c
                   : (int64_t)J * (W / 2) + (c - W / 2);
}

// ---------------------------------------------------------------------------
// 1. Partial Gram of a block pair (W columns) over cpw consecutive 64-row
// chunks (the last workgroup may have fewer).  grid (npart, p, batch),
// npart = ceil(nchunk_x / cpw); Gp layout [batch][p][npart][W*W]
// col-major.  Each of waves 0, 2, 3 accumulates one (W/2) x (W/2) quadrant
// over the workgroup's chunks; tall problems take several chunks per
// workgroup so that the inner solve sums at most 8 partials.
template <typename R, int W>
__global__ __launch_bounds__(256) void jac_gram(
    const typename JT<R>::c_t* __restrict__ X, int64_t ldx, int64_t sX,
    int p, int round, int nchunk_x, int cpw,
    typename JT<R>::c_t* __restrict__ Gp, const int* __restrict__ active) {
  using c_t = typename JT<R>::c_t;
  using acc_t = typename JT<R>::acc_t;
  constexpr int T = W / 32;  // 16x16 MFMA tiles per wave per dimension
  const int bz = blockIdx.z;
  if (!active[bz]) return;
  const int nb = 2 * p;
  const int pi = blockIdx.y;
  const int part = blockIdx.x;
  const int npart = gridDim.x;
  const int b1 = rr_block(pi, round, nb);
  const int b2 = rr_block(nb - 1 - pi, round, nb);
  const int I = min(b1, b2), J = max(b1, b2);
  X += (int64_t)bz * sX;

  __shared__ R Xr[W][65];  // [local col][row]
  __shared__ R Xi[W][65];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // G is Hermitian: the lower-left quadrant (wave 1) is conj(upper-right)^T
  // and is filled in by jac_inner.
  const int wm = (wave & 1) * (W / 2), wn = (wave >> 1) * (W / 2);
  acc_t gr[T][T], gi[T][T];
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        gr[i][j][e] = (R)0;
        gi[i][j][e] = (R)0;
      }
  const int c0 = part * cpw, c1 = min(c0 + cpw, nchunk_x);
  for (int chunk = c0; chunk < c1; ++chunk) {
    if (chunk > c0) __syncthreads();  // the previous chunk's reads are done
    for (int idx = tid; idx < 64 * W; idx += 256) {
      const int r = idx & 63, c = idx >> 6;
      const c_t v =
          X[(int64_t)chunk * 64 + r + pair_col<W>(c, I, J) * ldx];
      Xr[c][r] = v.x;
      Xi[c][r] = v.y;
    }
    __syncthreads();
    if (wave == 1) continue;
    // G[c1][c2] = sum_r conj(X[r][c1]) X[r][c2]
#pragma unroll 4
    for (int ks = 0; ks < 64; ks += 4) {
      const int kr = ks + (lane >> 4);
      R ar[T], ai[T], br[T], bi[T];
#pragma unroll
      for (int i = 0; i < T; ++i) {
        ar[i] = Xr[wm + i * 16 + (lane & 15)][kr];
        ai[i] = -Xi[wm + i * 16 + (lane & 15)][kr];
        br[i] = Xr[wn + i * 16 + (lane & 15)][kr];
        bi[i] = Xi[wn + i * 16 + (lane & 15)][kr];
      }
#pragma unroll
      for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) {
          gr[i][j] = JT<R>::mfma(ar[i], br[j], gr[i][j]);
          gr[i][j] = JT<R>::mfma(-ai[i], bi[j], gr[i][j]);
          gi[i][j] = JT<R>::mfma(ar[i], bi[j], gi[i][j]);
          gi[i][j] = JT<R>::mfma(ai[i], br[j], gi[i][j]);
        }
    }
  }
  if (wave == 1) return;
  c_t* out =
      Gp + (((int64_t)bz * p + pi) * npart + part) * (int64_t)(W * W);
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j) {
      const int col = wn + j * 16 + (lane & 15);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = wm + i * 16 + JT<R>::row(lane, e);
        out[row + W * col] = JT<R>::mk(gr[i][j][e], gi[i][j][e]);
      }
    }
}

// Chunks per jac_gram workgroup: at most 8 partial Grams per pair.
inline int gram_cpw(int nchunk) { return nchunk > 8 ? (nchunk + 7) / 8 : 1; }
inline int gram_parts(int nchunk) {
  const int c = gram_cpw(nchunk);
  return (nchunk + c - 1) / c;
}

// Rotation of the pair (p, q) from the Gram entries a = G_pp, d = G_qq and
// g = G_pq (Golub & Van Loan 8.5, complex Hermitian form):
//   J = [[cs, sn], [-sn e, cs e]], e = e^{-i phi} = conj(g) / |g|,
// J^H [[a, g], [conj g, d]] J = diag(ap, dq).  With zeta = (d - a) / (2|g|)
// and tau = |zeta| + sqrt(1 + zeta^2) (so t = tan = sign(zeta) / tau):
// cs = tau / sqrt(1 + tau^2), sn = sign(zeta) / sqrt(1 + tau^2).  Computed
// without branches; a pair that already meets the tolerance (or has a
// column below the zero threshold) gets cs = 1, sn = 0, e = 1, which
// leaves every entry it touches unchanged bit for bit.
template <typename R>
struct Rot {
  R cs, sn, er, ei, ap, dq;
  bool on;
};

template <typename R>
__device__ __forceinline__ Rot<R> make_rot(R a, R d,
                                           typename JT<R>::c_t g, R zt,
                                           R tol2) {
  Rot<R> o;
  const R absc2 = g.x * g.x + g.y * g.y;
  o.on = (a > zt && d > zt && absc2 > tol2 * a * d);
  const R s2 = o.on ? absc2 : (R)1;
  const R inv = JT<R>::rsq_r(s2);
  const R absc = s2 * inv;
  const R zeta = (R)0.5 * (d - a) * inv;
  const R tau = JT<R>::ab(zeta) + JT<R>::sq_r((R)1 + zeta * zeta);
  const R rr = JT<R>::rsq_r((R)1 + tau * tau);
  const R sgn = zeta >= (R)0 ? (R)1 : (R)-1;
  const R tt = JT<R>::dv_r(sgn, tau);
  o.cs = o.on ? tau * rr : (R)1;
  o.sn = o.on ? sgn * rr : (R)0;
  o.er = o.on ? g.x * inv : (R)1;
  o.ei = o.on ? -g.y * inv : (R)0;
  o.ap = o.on ? a - tt * absc : a;
  o.dq = o.on ? d + tt * absc : d;
  return o;
}

// [u v] <- [u v] J (a row of two entries in the columns p, q of the pair)
template <typename R>
__device__ __forceinline__ void rot_right(const Rot<R>& j,
                                          typename JT<R>::c_t& u,
                                          typename JT<R>::c_t& v) {
  const R yr = j.er * v.x - j.ei * v.y;
  const R yi = j.er * v.y + j.ei * v.x;
  const typename JT<R>::c_t nu =
      JT<R>::mk(j.cs * u.x - j.sn * yr, j.cs * u.y - j.sn * yi);
  v = JT<R>::mk(j.sn * u.x + j.cs * yr, j.sn * u.y + j.cs * yi);
  u = nu;
}

// [u; v] <- J^H [u; v], J^H = [[cs, -sn conj(e)], [sn, cs conj(e)]]
template <typename R>
__device__ __forceinline__ void rot_left(const Rot<R>& j,
                                         typename JT<R>::c_t& u,
                                         typename JT<R>::c_t& v) {
  const R yr = j.er * v.x + j.ei * v.y;
  const R yi = j.er * v.y - j.ei * v.x;
  const typename JT<R>::c_t nu =
      JT<R>::mk(j.cs * u.x - j.sn * yr, j.cs * u.y - j.sn * yi);
  v = JT<R>::mk(j.sn * u.x + j.cs * yr, j.sn * u.y + j.cs * yi);
  u = nu;
}

// Round-robin slots of the inner sweep.  The W indices of the block pair
// sit in NP = W/2 pairs of slots (pair P: side 0 and side 1); a round
// rotates every pair at once, then every index moves one step along the
// ring side0(1) -> side0(2) -> ... -> side0(NP-1) -> side1(NP-1) -> ... ->
// side1(0) -> side0(1) while side0(0) stays (the circle method: all
// W(W-1)/2 index pairs meet once in W-1 rounds, after which every index is
// back in its first slot).  Round 0 holds index P at side0(P) and W-1-P at
// side1(P).
template <int NP>
__device__ __forceinline__ int slot_index0(int P, int side) {
  return side ? 2 * NP - 1 - P : P;
}
// slot (P, side) -> slot of the next round, as P' * 2 + side'.  cross:
// side0(P) holds index P of the first half for good and side1 shifts
// cyclically, so the NP rounds pair every first-half index with every
// second-half index once (the cross pairs of the two column blocks only).
template <int NP>
__device__ __forceinline__ int slot_next(int P, int side, bool cross) {
  if (cross) return side == 0 ? 2 * P : 2 * (P == 0 ? NP - 1 : P - 1) + 1;
  if (side == 0) {
    if (P == 0) return 0;
    if (P == NP - 1) return 2 * (NP - 1) + 1;
    return 2 * (P + 1);
  }
  if (P == 0) return 2 * 1;
  return 2 * (P - 1) + 1;
}

// Plane positions (round-0 slots) of the Gram entry (r, c) and of the
// rotation entry (r, c): the four planes are side x side in block order,
// index s NP^2 + P NP + P' with s = 2 (row side) + (column side).  Gram
// rows and columns both sit in slots (index i < NP: pair i, side 0; else
// pair 2 NP - 1 - i, side 1); rotation rows are natural (row 2P + side)
// and only its columns sit in slots.
template <int NP>
__device__ __forceinline__ int gpos(int r, int c) {
  const int sr = r >= NP ? 1 : 0, sc = c >= NP ? 1 : 0;
  const int pr = sr ? 2 * NP - 1 - r : r, pc = sc ? 2 * NP - 1 - c : c;
  return (2 * sr + sc) * NP * NP + pr * NP + pc;
}
template <int NP>
__device__ __forceinline__ int qpos(int r, int c) {
  const int sc = c >= NP ? 1 : 0, pc = sc ? 2 * NP - 1 - c : c;
  return (2 * (r & 1) + sc) * NP * NP + (r >> 1) * NP + pc;
}

// Check and n_inner cyclic sweeps on the Hermitian Gram block, given as
// planes in round-0 slots in Gp (gpos), accumulating the rotations in
// planes (the identity on entry, set here) whose final buffer is returned
// in qout (qpos).  Returns the number of rotations (0 when the block
// already meets the tolerance).  Called by all 256 threads; any_off / nrot
// are zeroed __shared__ ints (visible after a barrier) and offmax_bz the
// problem's sweep maximum.
//
// A round rotates the NP = W/2 slot pairs at once, i.e. G <- J^H G J with J
// block diagonal over the pairs, so the (pair P, pair P') 2x2 block of G
// becomes J_P^H G_PP' J_P': it depends on that block and the two pairs'
// rotations only.  Thread t owns the blocks (P, P') with P' = t mod NP and
// P = t div NP + k 256/NP, and rows 2P, 2P+1 of Q in the columns of pair P'.
// The sweep keeps G and Q as four planes (side x side) in block order, so a
// thread's own block is read at consecutive addresses across its wave and
// the diagonal block of pair P' at stride NP + 1 (no LDS bank conflicts);
// after the update each entry is written straight to the slot its index
// moves to, i.e. where its owner of the next round reads it.  Per round a
// thread reads its blocks and the diagonal block of its column pair,
// derives J_P' -- J_P comes from the lane of its wave that derived it as
// its own J_P' (a shuffle) -- transforms in registers and writes.  Gb / Qb (four planes each, 4 NP^2 entries) double-buffer the
// planes: the round reads one buffer and writes the other, one barrier per
// round.  Gb == Gp (no second G buffer): a barrier separates the reads from
// the writes.  Qb == Qp with a second G buffer (Q written in place): a
// thread's Q entries only move between the threads of its row pair, which
// share its wave, and a wave issues its LDS reads before its writes, so the
// round keeps its single barrier.  The planes of the first buffer live in
// G's and Q's own storage (W (W+1) >= 4 NP^2 entries).  The diagonal block of a rotated
// pair is set to diag(ap, dq) exactly.
template <typename R, int W>
__device__ __forceinline__ int inner_solve(
    typename JT<R>::c_t* Gp, typename JT<R>::c_t* Qp,
    typename JT<R>::c_t* Gb, typename JT<R>::c_t* Qb, R zt, R tol2,
    int n_inner, double* offmax_bz, int& any_off, int& nrot, bool cross,
    typename JT<R>::c_t*& qout) {
  using c_t = typename JT<R>::c_t;
  constexpr int NP = W / 2;             // rotation pairs per round
  constexpr int NBLK = NP * NP;         // 2x2 blocks
  constexpr int NBT = NBLK / 256;       // blocks per thread
  constexpr int RS = 256 / NP;          // row-pair stride between them
  static_assert(NBT >= 1 && NBT * 256 == NBLK, "W must be 32 or 64");
  const int tid = threadIdx.x;
  qout = Qp;
  // Largest scaled off-diagonal |G_pq|^2 / (G_pp G_qq) of the block; nothing
  // to do when it meets the tolerance (the common case in late sweeps).
  // Reduced per wave (shuffles), then across the 4 waves in LDS, so the
  // problem-wide maximum takes one global atomic per workgroup.
  __shared__ double wmax[4];
  __shared__ int woff[4];
  double omax = 0.0;
  bool off = false;
  for (int idx = tid; idx < W * W; idx += 256) {
    const int r = idx % W, c = idx / W;
    if (r < c) {
      const R a = Gp[gpos<NP>(r, r)].x, d = Gp[gpos<NP>(c, c)].x;
      if (a > zt && d > zt) {
        const c_t g = Gp[gpos<NP>(r, c)];
        const R o2 = (g.x * g.x + g.y * g.y) / (a * d);
        omax = fmax(omax, (double)o2);
        // the sweep can only reduce the pairs it rotates
        off = off || (o2 > tol2 && (!cross || (r < NP && c >= NP)));
      }
    }
  }
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1)
    omax = fmax(omax, __shfl_xor(omax, sh, 64));
  const bool woff_any = __any(off);
  if ((tid & 63) == 0) {
    wmax[tid >> 6] = omax;
    woff[tid >> 6] = woff_any ? 1 : 0;
  }
  __syncthreads();
  if (tid == 0) {
    const double bm = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    // positive doubles order like their bit patterns
    if (bm > 0.0)
      atomicMax((unsigned long long*)offmax_bz,
                (unsigned long long)__double_as_longlong(bm));
    any_off = woff[0] | woff[1] | woff[2] | woff[3];
  }
  __syncthreads();
  if (!any_off) return 0;

  const int lane = tid & 63;
  const int pc_i = tid % NP;
  const int pr0 = tid / NP;
  const bool dbuf = Gb != Gp;
  // in-place Q planes rely on a row pair's threads sharing one wave
  static_assert(NP <= 64 && 64 % NP == 0, "a row pair's threads in one wave");
  // Q planes = the identity (each thread its own blocks)
#pragma unroll
  for (int k = 0; k < NBT; ++k) {
    const int P = pr0 + RS * k;
    const int blk = P * NP + pc_i;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const bool one = 2 * P + (s >> 1) == slot_index0<NP>(pc_i, s & 1);
      Qp[s * NBLK + blk] = JT<R>::mk(one ? (R)1 : (R)0, (R)0);
    }
  }
  __syncthreads();
  // next-round slots of this thread's column sides and row sides (fixed)
  int dc[2], dr[NBT][2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    dc[s] = slot_next<NP>(pc_i, s, cross);
#pragma unroll
    for (int k = 0; k < NBT; ++k)
      dr[k][s] = slot_next<NP>(pr0 + RS * k, s, cross);
  }
  c_t* Gc = Gp;
  c_t* Gn = Gb;
  c_t* Qc = Qp;
  c_t* Qn = Qb;
  for (int sweep = 0; sweep < n_inner; ++sweep) {
    int sweep_rot = 0;
    const int nrounds = cross ? NP : W - 1;
    for (int round = 0; round < nrounds; ++round) {
      // --- every LDS read of the round
      const int dcol = pc_i * (NP + 1);
      const R ga = Gc[0 * NBLK + dcol].x, gd = Gc[3 * NBLK + dcol].x;
      const c_t gg = Gc[1 * NBLK + dcol];
      c_t b[NBT][4], q[NBT][4];
#pragma unroll
      for (int k = 0; k < NBT; ++k) {
        const int blk = (pr0 + RS * k) * NP + pc_i;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          b[k][s] = Gc[s * NBLK + blk];
          q[k][s] = Qc[s * NBLK + blk];
        }
      }
      if (!dbuf) __syncthreads();  // every read precedes every write
      // (in-place Q: this wave's reads precede its writes)
      else if (Qn == Qc) __builtin_amdgcn_wave_barrier();
      // --- rotations
      const Rot<R> jc = make_rot<R>(ga, gd, gg, zt, tol2);
      Rot<R> jr[NBT];
#pragma unroll
      for (int k = 0; k < NBT; ++k) {
        // the lane of this wave whose column pair is this row pair
        const int src = (lane & ~(NP - 1)) | (pr0 + RS * k);
        jr[k].cs = __shfl(jc.cs, src, 64);
        jr[k].sn = __shfl(jc.sn, src, 64);
        jr[k].er = __shfl(jc.er, src, 64);
        jr[k].ei = __shfl(jc.ei, src, 64);
      }
      // --- transform, then write each entry to its next-round slot
#pragma unroll
      for (int k = 0; k < NBT; ++k) {
        // b[s]: s = 2 row side + column side
        rot_right<R>(jc, b[k][0], b[k][1]);
        rot_right<R>(jc, b[k][2], b[k][3]);
        rot_left<R>(jr[k], b[k][0], b[k][2]);
        rot_left<R>(jr[k], b[k][1], b[k][3]);
        if (pr0 + RS * k == pc_i && jc.on) {
          b[k][0] = JT<R>::mk(jc.ap, (R)0);
          b[k][1] = JT<R>::mk((R)0, (R)0);
          b[k][2] = JT<R>::mk((R)0, (R)0);
          b[k][3] = JT<R>::mk(jc.dq, (R)0);
          ++sweep_rot;
        }
        rot_right<R>(jc, q[k][0], q[k][1]);
        rot_right<R>(jc, q[k][2], q[k][3]);
        const int P = pr0 + RS * k;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int rs = dr[k][s >> 1], cs = dc[s & 1];
          Gn[((rs & 1) * 2 + (cs & 1)) * NBLK + (rs >> 1) * NP + (cs >> 1)] =
              b[k][s];
          // Q rows stay with the thread; only the column moves
          Qn[((s >> 1) * 2 + (cs & 1)) * NBLK + P * NP + (cs >> 1)] = q[k][s];
        }
      }
      __syncthreads();
      c_t* t = Gc;
      Gc = Gn;
      Gn = t;
      t = Qc;
      Qc = Qn;
      Qn = t;
    }
    // rotations of this sweep (one count per pair per round)
    if (sweep_rot) atomicAdd(&nrot, sweep_rot);
    __syncthreads();
    const int total = nrot;
    __syncthreads();
    if (tid == 0) nrot = 0;
    if (total == 0) break;
    __syncthreads();
    if (tid == 0) any_off += total;  // running total (started at 1)
    __syncthreads();
  }
  // every index is back in its round-0 slot; the last barrier of the loop
  // made the final planes visible
  qout = Qc;
  return any_off - 1;
}

// ---------------------------------------------------------------------------
// 2. Sum the partial Grams, run inner cyclic Jacobi sweep(s), emit Q.
// grid (p, batch), 256 threads.  flags[bz*p + pi] = 1 if Q must be applied.
// Each of the W/2 rotation pairs of a round is owned by 512/W threads, which
// all derive the rotation from G_pp, G_qq, G_pq themselves; thread t of a
// pair updates rows / columns t, t+TP, ...  Two barriers per round: after
// the column update (G J, Q J) and after the row update (J^H G).  Operands
// are loaded into registers before any store so LDS latencies overlap.
template <typename R, int W>
__global__ __launch_bounds__(256) void jac_inner(
    const typename JT<R>::c_t* __restrict__ Gp, int nchunk, int p,
    double tol, const double* __restrict__ zthr, int n_inner,
    typename JT<R>::c_t* __restrict__ Qout, int* __restrict__ flags,
    int* __restrict__ counters, const int* __restrict__ active,
    double* __restrict__ offmax, int* __restrict__ qpairs, int cross,
    const int* __restrict__ want, int w_mode) {
  using c_t = typename JT<R>::c_t;
  constexpr int NP = W / 2;       // rotation pairs per round
  constexpr int TP = 256 / NP;    // threads per pair
  constexpr int RPT = W / TP;     // rows (columns) per thread
  const int bz = blockIdx.y;
  const int pi = blockIdx.x;
  if (!active[bz]) return;
  // The inner sweeps double-buffer their G and Q planes where the second
  // buffers (2 x W^2 entries) keep the kernel within 72 KiB of LDS (W =
  // 32; not W = 64)
  constexpr bool kDB =
      (2 * W * (W + 1) + 2 * W * W) * sizeof(c_t) <= 72 * 1024;
  __shared__ c_t G[W][W + 1];  // G[row][col]
  __shared__ c_t Q[W][W + 1];  // Q[row][col]
  __shared__ c_t G2[kDB ? W * W : 1];
  __shared__ c_t Q2[kDB ? W * W : 1];
  __shared__ int any_off;
  __shared__ int nrot;
  const int tid = threadIdx.x;
  const c_t* src =
      Gp + ((int64_t)bz * p + pi) * nchunk * (int64_t)(W * W);
  // Sum the partial Grams chunk by chunk (in chunk order), with CG chunks'
  // loads in flight per thread; the lower-left quadrant is filled by
  // symmetry below.
  constexpr int IPT = W * W / 256;       // entries per thread
  constexpr int CG = IPT >= 16 ? 1 : 16 / IPT;
  R sr[IPT], si[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    sr[i] = 0;
    si[i] = 0;
  }
  for (int ch0 = 0; ch0 < nchunk; ch0 += CG) {
    c_t v[CG][IPT];
#pragma unroll
    for (int u = 0; u < CG; ++u)
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const int idx = tid + 256 * i;
        const int r = idx % W, c = idx / W;
        const bool ll = r >= W / 2 && c < W / 2;
        v[u][i] = (ch0 + u < nchunk && !ll)
                      ? src[(int64_t)(ch0 + u) * (W * W) + idx]
                      : JT<R>::mk((R)0, (R)0);
      }
#pragma unroll
    for (int u = 0; u < CG; ++u)
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        if (ch0 + u < nchunk) {
          sr[i] += v[u][i].x;
          si[i] += v[u][i].y;
        }
      }
  }
  // G straight into the planes of the inner sweep (round-0 slots)
  c_t* Gpl = &G[0][0];
  c_t* Qpl = &Q[0][0];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const int idx = tid + 256 * i;
    const int r = idx % W, c = idx / W;
    if (!(r >= W / 2 && c < W / 2))
      Gpl[gpos<W / 2>(r, c)] = JT<R>::mk(sr[i], si[i]);
  }
  if (tid == 0) {
    any_off = 0;
    nrot = 0;
  }
  const R zt = (R)zthr[bz];
  const R tol2 = (R)(tol * tol);
  __syncthreads();
  for (int idx = tid; idx < (W / 2) * (W / 2); idx += 256) {
    const int r = W / 2 + idx % (W / 2), c = idx / (W / 2);
    const c_t u = Gpl[gpos<W / 2>(c, r)];
    Gpl[gpos<W / 2>(r, c)] = JT<R>::mk(u.x, -u.y);
  }
  __syncthreads();
  c_t* qf;
  const int total = inner_solve<R, W>(
      Gpl, Qpl, kDB ? G2 : Gpl, kDB ? Q2 : Qpl, zt, tol2, n_inner,
      &offmax[bz], any_off, nrot, cross != 0, qf);
  if (total == 0) {
    if (tid == 0) flags[bz * p + pi] = 0;
    return;
  }
  if (tid == 0) {
    flags[bz * p + pi] = total > 0 ? 1 : 0;
    if (total > 0) {
      atomicAdd(&counters[bz], total);
      atomicAdd(qpairs, 1);
      // rotated pairs whose W columns the update launch also rotates
      if (w_mode == 1 || (w_mode == 2 && want[bz])) atomicAdd(qpairs + 1, 1);
    }
  }
  if (total > 0) {
    c_t* q = Qout + ((int64_t)bz * p + pi) * (int64_t)(W * W);
    for (int idx = tid; idx < W * W; idx += 256) {
      const int r = idx % W, c = idx / W;
      q[idx] = qf[qpos<W / 2>(r, c)];
    }
  }
}

// ---------------------------------------------------------------------------
// 3. X_IJ <- X_IJ Q (and W_IJ <- W_IJ Q) for one 64-row chunk.  grid
// (nchunk_x + nchunk_w, p, batch): chunks below nchunk_x are X's, the rest
// W's (Wm non-null; only for the problems with want[bz] != 0 when want is
// given), so one launch updates both.
template <typename R, int W>
__global__ __launch_bounds__(256) void jac_update(
    typename JT<R>::c_t* __restrict__ X, int64_t ldx, int64_t sX,
    int nchunk_x, typename JT<R>::c_t* __restrict__ Wm, int64_t ldw,
    int64_t sW, int p, int round, const typename JT<R>::c_t* __restrict__ Q,
    const int* __restrict__ flags, const int* __restrict__ active,
    const int* __restrict__ want) {
  using c_t = typename JT<R>::c_t;
  using acc_t = typename JT<R>::acc_t;
  constexpr int T = W / 32;
  const int bz = blockIdx.z;
  if (!active[bz]) return;
  const int pi = blockIdx.y;
  if (!flags[bz * p + pi]) return;
  const int nb = 2 * p;
  int chunk = blockIdx.x;
  int64_t ldx_ = ldx;
  if (chunk >= nchunk_x) {
    if (want && !want[bz]) return;
    chunk -= nchunk_x;
    X = Wm + (int64_t)bz * sW;
    ldx_ = ldw;
  } else {
    X += (int64_t)bz * sX;
  }
  ldx = ldx_;
  const int b1 = rr_block(pi, round, nb);
  const int b2 = rr_block(nb - 1 - pi, round, nb);
  const int I = min(b1, b2), J = max(b1, b2);
  const c_t* Qp = Q + ((int64_t)bz * p + pi) * (int64_t)(W * W);

  __shared__ R Xr[W][65];  // [local col][row]
  __shared__ R Xi[W][65];
  __shared__ R Qr[W][W + 1];  // [k][n]
  __shared__ R Qi[W][W + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * W; idx += 256) {
    const int r = idx & 63, c = idx >> 6;
    const c_t v = X[(int64_t)chunk * 64 + r + pair_col<W>(c, I, J) * ldx];
    Xr[c][r] = v.x;
    Xi[c][r] = v.y;
  }
  for (int idx = tid; idx < W * W; idx += 256) {
    const int r = idx % W, c = idx / W;
    const c_t qv = Qp[idx];  // Q[r][c] col-major
    Qr[r][c] = qv.x;
    Qi[r][c] = qv.y;
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  // output 64 rows x W cols; wave tile (32) x (W/2)
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * (W / 2);
  acc_t yr[2][T], yi[2][T];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        yr[i][j][e] = (R)0;
        yi[i][j][e] = (R)0;
      }
  // Y[row][n] = sum_k X[row][k] Q[k][n]
#pragma unroll 4
  for (int ks = 0; ks < W; ks += 4) {
    const int kr = ks + (lane >> 4);
    R ar[2], ai[2], br[T], bi[T];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ar[i] = Xr[kr][wm + i * 16 + (lane & 15)];
      ai[i] = Xi[kr][wm + i * 16 + (lane & 15)];
    }
#pragma unroll
    for (int j = 0; j < T; ++j) {
      br[j] = Qr[kr][wn + j * 16 + (lane & 15)];
      bi[j] = Qi[kr][wn + j * 16 + (lane & 15)];
    }
    // Y^T = Q^T X^T: the accumulators hold 16 consecutive rows per column,
    // so each store instruction writes whole 16-row runs
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < T; ++j) {
        yr[i][j] = JT<R>::mfma(br[j], ar[i], yr[i][j]);
        yr[i][j] = JT<R>::mfma(bi[j], -ai[i], yr[i][j]);
        yi[i][j] = JT<R>::mfma(bi[j], ar[i], yi[i][j]);
        yi[i][j] = JT<R>::mfma(br[j], ai[i], yi[i][j]);
      }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j) {
      const int r = wm + i * 16 + (lane & 15);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = wn + j * 16 + JT<R>::row(lane, e);
        X[(int64_t)chunk * 64 + r + pair_col<W>(c, I, J) * ldx] =
            JT<R>::mk(yr[i][j][e], yi[i][j][e]);
      }
    }
}

// ---------------------------------------------------------------------------
// Fused round: one workgroup per block pair does all three steps of the
// pair -- G = X_IJ^H X_IJ on MFMA streaming X_IJ through LDS in 64-row
// chunks (register-prefetched one chunk ahead), the inner sweeps on G in
// LDS, then X_IJ <- X_IJ Q (and W_IJ <- W_IJ Q when wanted) on MFMA chunk
// by chunk in place.  One launch per round, no partial Grams in memory.
// grid (p, batch), 256 threads.  W = 32 (blocks of 16 columns).
template <typename R, int W>
__device__ __forceinline__ void load_tile(
    const typename JT<R>::c_t* __restrict__ X, int64_t ldx, int chunk, int I,
    int J, typename JT<R>::c_t (&v)[64 * W / 256]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 64 * W / 256; ++u) {
    const int idx = tid + 256 * u;
    const int r = idx & 63, c = idx >> 6;
    v[u] = X[(int64_t)chunk * 64 + r + pair_col<W>(c, I, J) * ldx];
  }
}

template <typename R, int W>
__device__ __forceinline__ void store_tile_lds(
    const typename JT<R>::c_t (&v)[64 * W / 256], R (*Xr)[65], R (*Xi)[65]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 64 * W / 256; ++u) {
    const int idx = tid + 256 * u;
    const int r = idx & 63, c = idx >> 6;
    Xr[c][r] = v[u].x;
    Xi[c][r] = v[u].y;
  }
}

// Y = X_IJ Q over every 64-row chunk of one matrix, in place.
template <typename R, int W>
__device__ __forceinline__ void apply_q_chunks(
    typename JT<R>::c_t* __restrict__ X, int64_t ldx, int nchunk, int I,
    int J, R (*Xr)[65], R (*Xi)[65], R (*Qr)[W + 1], R (*Qi)[W + 1]) {
  using c_t = typename JT<R>::c_t;
  using acc_t = typename JT<R>::acc_t;
  constexpr int T = W / 32;
  constexpr int EPT = 64 * W / 256;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * (W / 2);
  c_t nxt[EPT];
  load_tile<R, W>(X, ldx, 0, I, J, nxt);
  for (int ch = 0; ch < nchunk; ++ch) {
    __syncthreads();  // the previous chunk's MFMA reads of Xr / Xi are done
    store_tile_lds<R, W>(nxt, Xr, Xi);
    __syncthreads();
    if (ch + 1 < nchunk) load_tile<R, W>(X, ldx, ch + 1, I, J, nxt);
    acc_t yr[2][T], yi[2][T];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < T; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          yr[i][j][e] = (R)0;
          yi[i][j][e] = (R)0;
        }
#pragma unroll 4
    for (int ks = 0; ks < W; ks += 4) {
      const int kr = ks + (lane >> 4);
      R ar[2], ai[2], br[T], bi[T];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Xr[kr][wm + i * 16 + (lane & 15)];
        ai[i] = Xi[kr][wm + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int j = 0; j < T; ++j) {
        br[j] = Qr[kr][wn + j * 16 + (lane & 15)];
        bi[j] = Qi[kr][wn + j * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) {
          yr[i][j] = JT<R>::mfma(ar[i], br[j], yr[i][j]);
          yr[i][j] = JT<R>::mfma(-ai[i], bi[j], yr[i][j]);
          yi[i][j] = JT<R>::mfma(ar[i], bi[j], yi[i][j]);
          yi[i][j] = JT<R>::mfma(ai[i], br[j], yi[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < T; ++j) {
        const int c = wn + j * 16 + (lane & 15);
        const int64_t gc = pair_col<W>(c, I, J);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = wm + i * 16 + JT<R>::row(lane, e);
          X[(int64_t)ch * 64 + r + gc * ldx] =
              JT<R>::mk(yr[i][j][e], yi[i][j][e]);
        }
      }
  }
}

template <typename R, int W>
__global__ __launch_bounds__(256) void jac_round(
    typename JT<R>::c_t* __restrict__ X, int64_t ldx, int64_t sX,
    int nchunk_x, typename JT<R>::c_t* __restrict__ Wm, int64_t ldw,
    int64_t sW, int nchunk_w, const int* __restrict__ want, int p, int round,
    double tol, const double* __restrict__ zthr, int n_inner,
    int* __restrict__ counters, const int* __restrict__ active,
    double* __restrict__ offmax, int* __restrict__ qpairs) {
  using c_t = typename JT<R>::c_t;
  using acc_t = typename JT<R>::acc_t;
  constexpr int T = W / 32;
  constexpr int EPT = 64 * W / 256;
  const int bz = blockIdx.y;
  const int pi = blockIdx.x;
  if (!active[bz]) return;
  const int nb = 2 * p;
  const int b1 = rr_block(pi, round, nb);
  const int b2 = rr_block(nb - 1 - pi, round, nb);
  const int I = min(b1, b2), J = max(b1, b2);
  X += (int64_t)bz * sX;

  __shared__ R Xr[W][65];  // [local col][row] of the current chunk
  __shared__ R Xi[W][65];
  __shared__ c_t G[W][W + 1];
  __shared__ c_t Q[W][W + 1];
  // [k][n] planes of Q for the MFMA update live in G's storage (G is dead
  // once the inner sweeps are done), so two workgroups fit in one CU's LDS
  R (*Qr)[W + 1] = reinterpret_cast<R (*)[W + 1]>(&G[0][0]);
  R (*Qi)[W + 1] = Qr + W;
  __shared__ int any_off;
  __shared__ int nrot;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  // 1. Gram: waves 0, 2, 3 accumulate the upper-left, lower-right and
  // upper-right (W/2)^2 quadrants over all chunks.
  const int wm = (wave & 1) * (W / 2), wn = (wave >> 1) * (W / 2);
  acc_t gr[T][T], gi[T][T];
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        gr[i][j][e] = (R)0;
        gi[i][j][e] = (R)0;
      }
  {
    c_t nxt[EPT];
    load_tile<R, W>(X, ldx, 0, I, J, nxt);
    for (int ch = 0; ch < nchunk_x; ++ch) {
      __syncthreads();
      store_tile_lds<R, W>(nxt, Xr, Xi);
      __syncthreads();
      if (ch + 1 < nchunk_x) load_tile<R, W>(X, ldx, ch + 1, I, J, nxt);
      if (wave != 1) {
#pragma unroll 4
        for (int ks = 0; ks < 64; ks += 4) {
          const int kr = ks + (lane >> 4);
          R ar[T], ai[T], br[T], bi[T];
#pragma unroll
          for (int i = 0; i < T; ++i) {
            ar[i] = Xr[wm + i * 16 + (lane & 15)][kr];
            ai[i] = -Xi[wm + i * 16 + (lane & 15)][kr];
            br[i] = Xr[wn + i * 16 + (lane & 15)][kr];
            bi[i] = Xi[wn + i * 16 + (lane & 15)][kr];
          }
#pragma unroll
          for (int i = 0; i < T; ++i)
#pragma unroll
            for (int j = 0; j < T; ++j) {
              gr[i][j] = JT<R>::mfma(ar[i], br[j], gr[i][j]);
              gr[i][j] = JT<R>::mfma(-ai[i], bi[j], gr[i][j]);
              gi[i][j] = JT<R>::mfma(ar[i], bi[j], gi[i][j]);
              gi[i][j] = JT<R>::mfma(ai[i], br[j], gi[i][j]);
            }
        }
      }
    }
  }
  // G straight into the planes of the inner sweep (round-0 slots)
  c_t* Gpl = &G[0][0];
  if (wave != 1) {
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
      for (int j = 0; j < T; ++j) {
        const int col = wn + j * 16 + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = wm + i * 16 + JT<R>::row(lane, e);
          Gpl[gpos<W / 2>(row, col)] = JT<R>::mk(gr[i][j][e], gi[i][j][e]);
        }
      }
  }
  if (tid == 0) {
    any_off = 0;
    nrot = 0;
  }
  __syncthreads();
  for (int idx = tid; idx < (W / 2) * (W / 2); idx += 256) {
    const int r = W / 2 + idx % (W / 2), c = idx / (W / 2);
    const c_t u = Gpl[gpos<W / 2>(c, r)];
    Gpl[gpos<W / 2>(r, c)] = JT<R>::mk(u.x, -u.y);
  }
  __syncthreads();

  // 2. inner sweeps
  const R zt = (R)zthr[bz];
  const R tol2 = (R)(tol * tol);
  c_t* qf;
  const int total = inner_solve<R, W>(Gpl, &Q[0][0], Gpl, &Q[0][0], zt,
                                      tol2, n_inner, &offmax[bz], any_off,
                                      nrot, false, qf);
  if (total == 0) return;
  if (tid == 0) {
    atomicAdd(&counters[bz], total);
    atomicAdd(qpairs, 1);
  }
  // 3. X_IJ <- X_IJ Q (and W_IJ <- W_IJ Q)
  for (int idx = tid; idx < W * W; idx += 256) {
    const int r = idx % W, c = idx / W;
    const c_t q = qf[qpos<W / 2>(r, c)];
    Qr[r][c] = q.x;
    Qi[r][c] = q.y;
  }
  apply_q_chunks<R, W>(X, ldx, nchunk_x, I, J, Xr, Xi, Qr, Qi);
  if (Wm && (!want || want[bz]))
    apply_q_chunks<R, W>(Wm + (int64_t)bz * sW, ldw, nchunk_w, I, J, Xr, Xi,
                         Qr, Qi);
}

// ---------------------------------------------------------------------------
// The fused round with the block pair resident in LDS (complex FP32, at
// most kResRows rows, W = 32).  X_IJ is read from HBM once into LDS
// planes; its Gram (MFMA, the rows split over the four waves), the inner
// sweeps and X_IJ Q (MFMA, written back over the tile) all work on that
// copy, and X_IJ Q leaves in one coalesced pass (W_IJ Q follows through
// the same tile when wanted).  One launch per round instead of three, X_IJ
// crosses HBM once each way, and the only serial memory latencies are the
// tile load and store.  The tile takes most of a CU's LDS (one workgroup
// per CU).  grid (p, batch), 256 threads.
constexpr int kResRows = 512;
constexpr int kResLd = kResRows + 4;  // column stride of the tile planes

// rows x 32 tile of a matrix (columns of the block pair I, J) -> LDS planes
// Tr / Ti [col][row]; thread t moves rows 2t, 2t + 1 of every column as one
// 16-byte access (rows, ld and the batch stride are multiples of 64), and
// all 32 columns' loads are in flight together.
__device__ __forceinline__ void res_load(const float2* __restrict__ X,
                                         int64_t ldx, int rows, int I, int J,
                                         float (*Tr)[kResLd],
                                         float (*Ti)[kResLd]) {
  const int r = 2 * threadIdx.x;
  static_assert(kResRows <= 2 * 256, "two rows per thread and column");
  if (r >= rows) return;
  float4 v[32];
#pragma unroll
  for (int c = 0; c < 32; ++c)
    v[c] = *reinterpret_cast<const float4*>(X + pair_col<32>(c, I, J) * ldx +
                                            r);
#pragma unroll
  for (int c = 0; c < 32; ++c) {
    *reinterpret_cast<float2*>(&Tr[c][r]) = make_float2(v[c].x, v[c].z);
    *reinterpret_cast<float2*>(&Ti[c][r]) = make_float2(v[c].y, v[c].w);
  }
}

// Y = T Q for the tile in LDS, stored straight from the MFMA results into
// the rows x 32 tile of Y (global; it may be the matrix T was read from).
// Wave w owns rows [w rows/4, (w+1) rows/4) in 16-row tiles.  The MFMAs
// form Y^T = Q^T T^T, so each accumulator element holds one row of a
// 16-row run and a store instruction writes 16 consecutive rows of four
// columns (128 contiguous bytes each).  The complex product takes three
// real products (Gauss): with t = a + ib and q = c + id, K1 = c (a + b),
// K2 = a (d - c), K3 = b (c + d), so that Re = K1 - K3 and Im = K1 + K2;
// P0 / P1 / P2 hold c, d - c and c + d in [k][n] planes.
__device__ __forceinline__ void res_update(float2* __restrict__ Y,
                                           int64_t ldy, int rows, int I,
                                           int J, const float (*Tr)[kResLd],
                                           const float (*Ti)[kResLd],
                                           const float (*P0)[33],
                                           const float (*P1)[33],
                                           const float (*P2)[33]) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int per = rows / 4;
  // element e of column tile j: column j 16 + 4 (lane / 16) + e
  int64_t col[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      col[j][e] = pair_col<32>(j * 16 + 4 * (lane >> 4) + e, I, J) * ldy;
  for (int t0 = wave * per; t0 < (wave + 1) * per; t0 += 16) {
    f4_t k1[2], k2[2], k3[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        k1[j][e] = 0.f;
        k2[j][e] = 0.f;
        k3[j][e] = 0.f;
      }
#pragma unroll
    for (int ks = 0; ks < 32; ks += 4) {
      const int kr = ks + (lane >> 4);
      const float ar = Tr[kr][t0 + (lane & 15)];
      const float ai = Ti[kr][t0 + (lane & 15)];
      const float as = ar + ai;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int n = j * 16 + (lane & 15);
        k1[j] = JT<float>::mfma(P0[kr][n], as, k1[j]);
        k2[j] = JT<float>::mfma(P1[kr][n], ar, k2[j]);
        k3[j] = JT<float>::mfma(P2[kr][n], ai, k3[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        Y[col[j][e] + t0 + (lane & 15)] =
            make_float2(k1[j][e] - k3[j][e], k1[j][e] + k2[j][e]);
  }
}

__global__ __launch_bounds__(256) void jac_round_res(
    float2* __restrict__ X, int64_t ldx, int64_t sX, int nchunk_x,
    float2* __restrict__ Wm, int64_t ldw, int64_t sW, int nchunk_w,
    const int* __restrict__ want, int p, int round, double tol,
    const double* __restrict__ zthr, int n_inner, int* __restrict__ counters,
    const int* __restrict__ active, double* __restrict__ offmax,
    int* __restrict__ qpairs, int cross) {
  constexpr int W = 32;
  const int bz = blockIdx.y;
  const int pi = blockIdx.x;
  if (!active[bz]) return;
  const int nb = 2 * p;
  const int b1 = rr_block(pi, round, nb);
  const int b2 = rr_block(nb - 1 - pi, round, nb);
  const int I = min(b1, b2), J = max(b1, b2);
  X += (int64_t)bz * sX;
  const int rows = nchunk_x * 64;

  __shared__ float Tr[W][kResLd];
  __shared__ float Ti[W][kResLd];
  __shared__ float2 G[W][W + 1];
  __shared__ float2 Q[W][W + 1];
  __shared__ float2 G2[W * W];  // second buffer of the G planes
  __shared__ int any_off;
  __shared__ int nrot;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  res_load(X, ldx, rows, I, J, Tr, Ti);
  __syncthreads();

  // 1. Gram, G[c1][c2] = sum_r conj(X[r][c1]) X[r][c2]: every wave takes the
  // three (W/2)^2 quadrants UL, LR, UR over its quarter of the rows; the
  // partials are added in the fixed order (w0 + w2) + (w1 + w3).  For the
  // diagonal quadrants g[q][1] accumulates M = Xr^T Xi only (Im = M - M^T,
  // formed below), three products per step instead of four.
  f4_t g[3][2];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      g[q][0][e] = 0.f;
      g[q][1][e] = 0.f;
    }
  {
    const int per = rows / 4;
    for (int k0 = wave * per; k0 < (wave + 1) * per; k0 += 4) {
      const int kr = k0 + (lane >> 4);
      float xr[2], xi[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        xr[b] = Tr[b * 16 + (lane & 15)][kr];
        xi[b] = Ti[b * 16 + (lane & 15)][kr];
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        g[q][0] = JT<float>::mfma(xr[q], xr[q], g[q][0]);
        g[q][0] = JT<float>::mfma(xi[q], xi[q], g[q][0]);
        g[q][1] = JT<float>::mfma(xr[q], xi[q], g[q][1]);
      }
      g[2][0] = JT<float>::mfma(xr[0], xr[1], g[2][0]);
      g[2][0] = JT<float>::mfma(xi[0], xi[1], g[2][0]);
      g[2][1] = JT<float>::mfma(xr[0], xi[1], g[2][1]);
      g[2][1] = JT<float>::mfma(-xi[0], xr[1], g[2][1]);
    }
  }
  float* sg = reinterpret_cast<float*>(&G[0][0]);  // 1536 floats per wave
  float* sq = reinterpret_cast<float*>(&Q[0][0]);
  static_assert(24 * 64 <= W * (W + 1) * 2, "a wave's partial fits");
  auto put = [&](float* s) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) s[((q * 2 + c) * 4 + e) * 64 + lane] = g[q][c][e];
  };
  auto add = [&](const float* s) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) g[q][c][e] += s[((q * 2 + c) * 4 + e) * 64 + lane];
  };
  if (wave == 2) put(sg);
  if (wave == 3) put(sq);
  __syncthreads();
  if (wave == 0) add(sg);
  if (wave == 1) add(sq);
  __syncthreads();
  if (wave == 1) put(sq);
  __syncthreads();
  // G straight into the planes of the inner sweep (round-0 slots)
  float2* Gpl = &G[0][0];
  if (wave == 0) {
    add(sq);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int wm = q == 1 ? 16 : 0, wn = q == 0 ? 0 : 16;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        Gpl[gpos<W / 2>(wm + JT<float>::row(lane, e), wn + (lane & 15))] =
            make_float2(g[q][0][e], g[q][1][e]);
    }
  }
  __syncthreads();
  // diagonal quadrants: Im[r][c] = M[r][c] - M[c][r], one thread per
  // unordered pair {r, c} (it reads both before writing either); the
  // lower-left quadrant is the adjoint of the upper-right one
  for (int idx = tid; idx < W * W; idx += 256) {
    const int r = idx % W, c = idx / W;
    if ((r < W / 2) == (c < W / 2) && r <= c) {
      float2& grc = Gpl[gpos<W / 2>(r, c)];
      float2& gcr = Gpl[gpos<W / 2>(c, r)];
      const float m = grc.y - gcr.y;
      grc.y = m;
      gcr.y = -m;
    } else if (r >= W / 2 && c < W / 2) {
      const float2 u = Gpl[gpos<W / 2>(c, r)];
      Gpl[gpos<W / 2>(r, c)] = make_float2(u.x, -u.y);
    }
  }
  if (tid == 0) {
    any_off = 0;
    nrot = 0;
  }
  __syncthreads();

  // 2. inner sweeps
  const float zt = (float)zthr[bz];
  const float tol2 = (float)(tol * tol);
  float2* qf;
  const int total = inner_solve<float, W>(Gpl, &Q[0][0], G2, &Q[0][0], zt,
                                          tol2, n_inner, &offmax[bz],
                                          any_off, nrot, cross != 0, qf);
  if (total == 0) return;
  const bool do_w = Wm && (!want || want[bz]);
  if (tid == 0) {
    atomicAdd(&counters[bz], total);
    atomicAdd(qpairs, 1);
    if (do_w) atomicAdd(qpairs + 1, 1);
  }
  // 3. X_IJ <- X_IJ Q (and W_IJ <- W_IJ Q); the planes of Q for the three
  // products: P0 and P1 in G's storage, P2 in Q's once Q has been read
  float (*P0)[W + 1] = reinterpret_cast<float (*)[W + 1]>(&G[0][0]);
  float (*P1)[W + 1] = P0 + W;
  float (*P2)[W + 1] = reinterpret_cast<float (*)[W + 1]>(&Q[0][0]);
  constexpr int QPT = W * W / 256;
  float2 qv[QPT];
#pragma unroll
  for (int u = 0; u < QPT; ++u) {
    const int idx = tid + 256 * u;
    qv[u] = qf[qpos<W / 2>(idx % W, idx / W)];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < QPT; ++u) {
    const int idx = tid + 256 * u;
    const int r = idx % W, c = idx / W;
    P0[r][c] = qv[u].x;
    P1[r][c] = qv[u].y - qv[u].x;
    P2[r][c] = qv[u].x + qv[u].y;
  }
  __syncthreads();
  res_update(X, ldx, rows, I, J, Tr, Ti, P0, P1, P2);
  if (do_w) {
    float2* Wb = Wm + (int64_t)bz * sW;
    const int wrows = nchunk_w * 64;
    __syncthreads();  // every wave has read its rows of the X tile
    res_load(Wb, ldw, wrows, I, J, Tr, Ti);
    __syncthreads();
    res_update(Wb, ldw, wrows, I, J, Tr, Ti, P0, P1, P2);
  }
}

// After a sweep: a problem stays active only if it rotated something and
// the largest scaled off-diagonal |x_p^H x_q|^2 / (|x_p|^2 |x_q|^2) met in
// the sweep was above conv2.  Below it the sweep's own rotations leave the
// columns orthogonal to O(conv2) (cyclic Jacobi converges quadratically),
// so the confirming sweep is skipped.
// misc: [0] problems still active, [1] non-finite input seen, [2] / [3]
// the sweep's rotated pairs (of X / of X and W); [4..7] receive a snapshot
// of [0..3] for the host, and [2], [3] restart at zero, so the sweep needs
// no memset (and can be replayed from a graph).
__global__ void jac_sweep_end(int* __restrict__ active,
                              int* __restrict__ counters,
                              double* __restrict__ offmax, double conv2,
                              int batch, int* __restrict__ misc) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int b = threadIdx.x; b < batch; b += blockDim.x) {
    const int a = active[b] && counters[b] > 0 && offmax[b] > conv2;
    active[b] = a;
    counters[b] = 0;
    offmax[b] = 0.0;
    if (a) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    misc[0] = cnt;
    misc[4] = cnt;
    misc[5] = misc[1];
    misc[6] = misc[2];
    misc[7] = misc[3];
    misc[2] = 0;
    misc[3] = 0;
  }
}

// Frobenius norm, finiteness check, scaling exponent and zero-column
// threshold per problem, in two steps.  jac_prep_partial: every block of
// grid (kPrepBlocks, batch) reduces its slice to the scaled sum of squares
// (scale = max |re|, |im| seen, ssq = sum (x / scale)^2, combined with the
// rescaling of LAPACK's nrm2), so nothing overflows or underflows.
constexpr int kPrepBlocks = 32;

__device__ __forceinline__ void ssq_add(double a, double& scale, double& ssq) {
  if (a > 0.0) {
    if (scale < a) {
      const double r = scale / a;
      ssq = 1.0 + ssq * r * r;
      scale = a;
    } else {
      const double r = a / scale;
      ssq += r * r;
    }
  }
}

__device__ __forceinline__ void ssq_merge(double s2, double q2, double& scale,
                                          double& ssq) {
  if (s2 <= 0.0) return;
  if (scale < s2) {
    const double r = scale / s2;
    ssq = q2 + ssq * r * r;
    scale = s2;
  } else {
    const double r = s2 / scale;
    ssq += q2 * r * r;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void jac_prep_partial(
    const typename JT<R>::c_t* __restrict__ X, int64_t rows, int64_t ldx,
    int64_t cols, int64_t sX, const int* __restrict__ mask,
    double* __restrict__ part_scale, double* __restrict__ part_ssq,
    int* __restrict__ part_bad) {
  const int bz = blockIdx.y;
  const int slot = bz * kPrepBlocks + blockIdx.x;
  if (mask && !mask[bz]) {
    if (threadIdx.x == 0) {
      part_scale[slot] = 0.0;
      part_ssq[slot] = 0.0;
      part_bad[slot] = 0;
    }
    return;
  }
  X += bz * sX;
  double scale = 0.0, ssq = 0.0;
  int nonfinite = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < rows * cols;
       i += (int64_t)kPrepBlocks * 256) {
    const typename JT<R>::c_t v = X[(i % rows) + (i / rows) * ldx];
    if (!isfinite(v.x) || !isfinite(v.y)) {
      nonfinite = 1;
      continue;
    }
    ssq_add(fabs((double)v.x), scale, ssq);
    ssq_add(fabs((double)v.y), scale, ssq);
  }
  __shared__ double rs[256], rq[256];
  __shared__ int nf;
  if (threadIdx.x == 0) nf = 0;
  rs[threadIdx.x] = scale;
  rq[threadIdx.x] = ssq;
  __syncthreads();
  if (nonfinite) atomicOr(&nf, 1);
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double sc = rs[threadIdx.x], q = rq[threadIdx.x];
      ssq_merge(rs[threadIdx.x + w], rq[threadIdx.x + w], sc, q);
      rs[threadIdx.x] = sc;
      rq[threadIdx.x] = q;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_scale[slot] = rs[0];
    part_ssq[slot] = rq[0];
    part_bad[slot] = nf;
  }
}

// jac_prep_final: one thread per problem merges its partials in order.
// The sweeps run on X * 2^-e with 2^e <= max |x| < 2^(e+1) (an exact
// scaling), so the squared Gram entries stay far from overflow and
// underflow whatever the input's magnitude; jac_scale applies it and undoes
// it after the sweeps.  Columns with norm below rows*eps*|X|_F are rounding
// noise (the backward error of the rotations themselves): never rotated.
__global__ void jac_prep_final(int batch, double zrel,
                               const double* __restrict__ part_scale,
                               const double* __restrict__ part_ssq,
                               const int* __restrict__ part_bad,
                               const int* __restrict__ mask,
                               double* __restrict__ normF2,
                               double* __restrict__ zthr,
                               int* __restrict__ active,
                               int* __restrict__ counters,
                               int* __restrict__ bad,
                               int* __restrict__ scale_e) {
  const int bz = blockIdx.x * blockDim.x + threadIdx.x;
  if (bz >= batch) return;
  double scale = 0.0, ssq = 0.0;
  int nf = 0;
  for (int b = 0; b < kPrepBlocks; ++b) {
    ssq_merge(part_scale[bz * kPrepBlocks + b], part_ssq[bz * kPrepBlocks + b],
              scale, ssq);
    nf |= part_bad[bz * kPrepBlocks + b];
  }
  const bool on = !(mask && !mask[bz]);
  const int e = (on && scale > 0.0 && isfinite(scale)) ? ilogb(scale) : 0;
  // |X 2^-e|_F^2 = ssq (scale 2^-e)^2 with scale 2^-e in [1, 2)
  const double r = ldexp(scale, -e);
  const double f2s = on ? ssq * r * r : 0.0;
  scale_e[bz] = e;
  normF2[bz] = f2s;
  zthr[bz] = zrel * zrel * f2s;
  active[bz] = (on && f2s > 0.0 && !nf) ? 1 : 0;
  counters[bz] = 0;
  if (on && nf) atomicOr(bad, 1);
}

// X *= 2^(sgn * e[bz]) over the rows x cols block (exact).
// grid (tiles, batch).
template <typename R>
__global__ void jac_scale(typename JT<R>::c_t* __restrict__ X, int64_t rows,
                          int64_t ldx, int64_t cols, int64_t sX,
                          const int* __restrict__ scale_e, int sgn) {
  const int bz = blockIdx.y;
  const int e = scale_e[bz];
  if (e == 0) return;
  const R f = (R)ldexp(1.0, sgn * e);
  typename JT<R>::c_t* x = X + (int64_t)bz * sX;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       i < rows * cols; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i % rows, c = i / rows;
    typename JT<R>::c_t v = x[r + c * ldx];
    v.x *= f;
    v.y *= f;
    x[r + c * ldx] = v;
  }
}

// Column norms: one wave per column.  grid (cols/4, batch), 256 threads.
template <typename R>
__global__ __launch_bounds__(256) void jac_colnorms(
    const typename JT<R>::c_t* __restrict__ X, int64_t rows, int64_t ldx,
    int64_t cols, int64_t sX, double* __restrict__ sig,
    const double* __restrict__ floor2, const int* __restrict__ scale_e) {
  const int bz = blockIdx.y;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= cols) return;
  const typename JT<R>::c_t* x = X + (int64_t)bz * sX + c * ldx;
  double s = 0.0;
  for (int64_t r = lane; r < rows; r += 64) {
    const typename JT<R>::c_t v = x[r];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  // Columns at or below the zero threshold (the rotations' rounding floor)
  // count as exact zeros: the ones the sweeps left alone, and the ones a
  // rank-deficient X ends with, whose direction is rounding noise and
  // would otherwise be kept (at a zero cutoff) as a singular vector.
  if (floor2 && s <= floor2[bz]) s = 0.0;
  // X is still scaled by 2^-e here
  if (lane == 0) sig[(int64_t)bz * cols + c] = ldexp(sqrt(s), scale_e[bz]);
}

// Sort sigma descending (bitonic in LDS) and apply the truncation rule of
// src/svd.jl:29-33 exactly: tail norms are accumulated from the smallest
// value upwards in sequence, k = #values kept.  grid (batch), 1024 threads.
// nsv[bz]: number of meaningful singular values (min of the true dims).
constexpr int kSortMax = 4096;
__global__ __launch_bounds__(1024) void jac_sort_trunc(
    double* __restrict__ sig, int* __restrict__ perm, int64_t cols,
    const int64_t* __restrict__ nsv, double cutoff, int relative,
    int64_t chimax, int64_t* __restrict__ kept, double* __restrict__ disc,
    double* __restrict__ total, double* __restrict__ keptw,
    int* __restrict__ status, const int* __restrict__ scale_e,
    const int64_t* __restrict__ chimax_arr) {
  const int bz = blockIdx.x;
  __shared__ double sv[kSortMax];
  __shared__ int si[kSortMax];
  int n2 = 1;
  while (n2 < cols) n2 <<= 1;
  double* s = sig + (int64_t)bz * cols;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    sv[i] = i < cols ? s[i] : -1.0;
    si[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool desc = (i & k) == 0;
          const double a = sv[i], b = sv[ixj];
          // descending within desc blocks; ties broken by index for
          // determinism
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
  }
  for (int i = threadIdx.x; i < cols; i += blockDim.x) {
    s[i] = sv[i];
    perm[(int64_t)bz * cols + i] = si[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t n = nsv[bz];
    // The sums run on sigma * 2^-e (the sweeps' exact scaling) so that
    // squares of very large or small values neither overflow nor vanish;
    // the rule is unchanged (er scales with the values), the weights are
    // reported unscaled.
    const int e = scale_e ? scale_e[bz] : 0;
    // cum[j] = sum_{l >= j} s_l^2, accumulated from the smallest upwards
    double c2 = 0.0;
    for (int64_t j = n - 1; j >= 0; --j) {
      const double v = ldexp(sv[j], -e);
      c2 += v * v;
    }
    const double tot = c2;
    const double er = relative ? cutoff * sqrt(tot) : ldexp(cutoff, -e);
    c2 = 0.0;
    int64_t k = -1;
    double disc_w = 0.0;
    for (int64_t j = n - 1; j >= 0; --j) {
      const double v = ldexp(sv[j], -e);
      c2 += v * v;
      if (sqrt(c2) > er) {
        k = j + 1;
        break;
      }
    }
    int st = 0;
    if (k < 0) {
      st = QTN_ERR_TRUNC;
      k = 0;
    }
    const int64_t cm = chimax_arr ? chimax_arr[bz] : chimax;
    if (cm > 0 && k > cm) k = cm;
    // discarded weight: sum of squares of s[k..n-1], smallest first
    disc_w = 0.0;
    for (int64_t j = n - 1; j >= k; --j) {
      const double v = ldexp(sv[j], -e);
      disc_w += v * v;
    }
    double kw = 0.0;
    for (int64_t j = k - 1; j >= 0; --j) {
      const double v = ldexp(sv[j], -e);
      kw += v * v;
    }
    kept[bz] = k;
    disc[bz] = ldexp(disc_w, 2 * e);
    total[bz] = ldexp(tot, 2 * e);
    keptw[bz] = ldexp(kw, 2 * e);
    status[bz] = st;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
int32_t JacobiEngine::reserve(qtn_ctx* ctx, int batch_, int64_t rows_p_,
                              int64_t cols_p_, bool single_, int n_active) {
  batch = batch_;
  single = single_;
  // The fused round (one workgroup per block pair) leaves each 64-row
  // chunk's load latency exposed with only 1-2 workgroups per CU; the
  // three-launch round overlaps them across all chunks and is faster at
  // chi = 256 in FP64.  FP32 problems of 320 to 512 rows fuse by default:
  // their round keeps the whole block pair in LDS (jac_round_res), so only
  // the tile load and store wait on memory.  The tile's LDS allows one
  // workgroup per CU, which costs throughput on large batches of smaller
  // problems (a 256-circuit batch of chi <= 128 ran 8% slower fused), so
  // those keep the three launches.  QTN_JACOBI_FUSED=1 / 0 forces fused /
  // three-launch rounds everywhere.
  const size_t csz = single ? sizeof(float2) : sizeof(double2);
  rows_p = rows_p_;
  cols_p = cols_p_;
  // Block-pair width: one 64-column pair solved entirely in LDS for narrow
  // problems; 32-column pairs otherwise (the inner sweep's LDS traffic per
  // outer sweep grows linearly with W while the MFMA work does not change).
  // The context's tuned tiling for this column bucket (qtn_tune_jacobi)
  // replaces that default; QTN_JACOBI_W / QTN_JACOBI_FUSED override both.
  fused = single && std::max(rows_p_, cols_p_) > 256 &&
          std::max(rows_p_, cols_p_) <= 512;
  int wdef = 32;
  const auto it = ctx->jac_tiling.find(
      {cols_p, single ? 1 : 0,
       batch_class(n_active >= 0 ? n_active : batch_)});
  if (it != ctx->jac_tiling.end()) {
    wdef = it->second == QTN_TILING_W64 ? 64 : 32;
    fused = it->second == QTN_TILING_W32_FUSED;
  }
  if (const char* e = getenv("QTN_JACOBI_W")) wdef = atoi(e) == 64 ? 64 : 32;
  if (const char* e = getenv("QTN_JACOBI_FUSED")) fused = atoi(e) != 0;
  // Outer rounds 1.. of a sweep rotate only the cross pairs of their two
  // column blocks (16 inner rounds instead of 31 at W = 32); round 0, where
  // every block appears once, also rotates the pairs inside the blocks, so
  // each column pair is still visited once per sweep.  Measured on the
  // 40-qubit chi-256 circuit: about 8% more outer sweeps, 13-15% more gates
  // per second.  QTN_JACOBI_CROSS=0 restores full inner sweeps.
  cross_rounds = true;
  if (const char* e = getenv("QTN_JACOBI_CROSS")) cross_rounds = atoi(e) != 0;
  W = (cols_p <= 64) ? 64 : wdef;
  p = (int)(cols_p / W);
  const int64_t nchunk = std::max(rows_p, cols_p) / 64;
  void* ptr;
  QTN_TRY(ctx->buffer(slot_base + 0,
                      (size_t)batch * p * nchunk * W * W * csz, &ptr));
  Gp = ptr;
  QTN_TRY(ctx->buffer(slot_base + 1, (size_t)batch * p * W * W * csz, &ptr));
  Q = ptr;
  // small per-problem arrays packed in one slot
  const size_t small =
      (size_t)batch * (p * sizeof(int) + 7 * sizeof(int) + 6 * sizeof(double) +
                       3 * sizeof(int64_t) +
                       kPrepBlocks * (2 * sizeof(double) + sizeof(int))) +
      1024;
  QTN_TRY(ctx->buffer(slot_base + 2, small, &ptr));
  char* c = (char*)ptr;
  normF2 = (double*)c;
  c += batch * sizeof(double);
  zthr = (double*)c;
  c += batch * sizeof(double);
  disc = (double*)c;
  c += batch * sizeof(double);
  total = (double*)c;
  c += batch * sizeof(double);
  keptw = (double*)c;
  c += batch * sizeof(double);
  offmax = (double*)c;
  c += batch * sizeof(double);
  prep_scale = (double*)c;
  c += (size_t)batch * kPrepBlocks * sizeof(double);
  prep_ssq = (double*)c;
  c += (size_t)batch * kPrepBlocks * sizeof(double);
  kept = (int64_t*)c;
  c += batch * sizeof(int64_t);
  chimax_dev = (int64_t*)c;
  c += batch * sizeof(int64_t);
  nsv = (int64_t*)c;
  c += batch * sizeof(int64_t);
  flags = (int*)c;
  c += (size_t)batch * p * sizeof(int);
  counters = (int*)c;
  c += batch * sizeof(int);
  active = (int*)c;
  c += batch * sizeof(int);
  status = (int*)c;
  c += batch * sizeof(int);
  want_w = (int*)c;
  c += batch * sizeof(int);
  mask = (int*)c;
  c += batch * sizeof(int);
  scale_e = (int*)c;
  c += batch * sizeof(int);
  prep_bad = (int*)c;
  c += (size_t)batch * kPrepBlocks * sizeof(int);
  // [0] n_active, [1] nonfinite flag, [2] rotated pairs of the sweep,
  // [3] those of them whose W columns are rotated too
  misc = (int*)c;
  QTN_TRY(ctx->buffer(slot_base + 3, (size_t)batch * cols_p * sizeof(double),
                      &ptr));
  sig = (double*)ptr;
  QTN_TRY(ctx->buffer(slot_base + 4, (size_t)batch * cols_p * sizeof(int),
                      &ptr));
  perm = (int*)ptr;
  return QTN_OK;
}

template <typename R, int WW>
int32_t JacobiEngine::launch_round(hipStream_t s, void* Xv, int64_t ldx,
                                   int64_t sX, void* Wv, int64_t sW, int r,
                                   double tol, int n_inner, int nchunk_x,
                                   int nchunk_w, bool per_problem_w) {
  using c_t = typename JT<R>::c_t;
  c_t* X = (c_t*)Xv;
  c_t* Wm = (c_t*)Wv;
  c_t* G = (c_t*)Gp;
  c_t* Qc = (c_t*)Q;
  // While profiling, every launch carries its own start / stop events,
  // stamped by the dispatch itself (hipExtLaunchKernelGGL), so the times
  // are kernel execution times without the gaps between launches.
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (ctx_ && ctx_->prof) {
    for (int i = 0; i < 6; ++i) {
      if (ctx_->ev_pool.empty()) {
        hipEvent_t e;
        QTN_CHECK_HIP(hipEventCreate(&e));
        ctx_->ev_pool.push_back(e);
      }
      ev[i] = ctx_->ev_pool.back();
      ctx_->ev_pool.pop_back();
      prof_events.push_back(ev[i]);
    }
  }
#define QTN_TIMED_LAUNCH(e0, e1, kern, grid, ...)                           \
  do {                                                                      \
    if (e0)                                                                 \
      hipExtLaunchKernelGGL(kern, grid, dim3(256), 0, s, e0, e1, 0,         \
                            __VA_ARGS__);                                   \
    else                                                                    \
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, __VA_ARGS__);         \
  } while (0)
  if (fused && WW == 32) {
    if constexpr (std::is_same<R, float>::value) {
      if (res_round(nchunk_x, Wm ? nchunk_w : 0)) {
        QTN_TIMED_LAUNCH(ev[0], ev[1], jac_round_res, dim3(p, batch), X, ldx,
                         sX, nchunk_x, Wm, ldw ? ldw : cols_p, sW, nchunk_w,
                         (const int*)(per_problem_w ? want_w : nullptr), p, r,
                         tol, zthr, n_inner, counters, active, offmax,
                         misc + 2, (cross_rounds && p > 1 && r > 0) ? 1 : 0);
        return QTN_OK;
      }
    }
    QTN_TIMED_LAUNCH(ev[0], ev[1], (jac_round<R, 32>), dim3(p, batch), X, ldx,
                     sX, nchunk_x, Wm, ldw ? ldw : cols_p, sW, nchunk_w,
                     (const int*)(per_problem_w ? want_w : nullptr), p, r,
                     tol, zthr, n_inner, counters, active, offmax, misc + 2);
    return QTN_OK;
  }
  const int npart = gram_parts(nchunk_x);
  QTN_TIMED_LAUNCH(ev[0], ev[1], (jac_gram<R, WW>), dim3(npart, p, batch),
                   (const c_t*)X, ldx, sX, p, r, nchunk_x, gram_cpw(nchunk_x),
                   G, active);
  QTN_TIMED_LAUNCH(ev[2], ev[3], (jac_inner<R, WW>), dim3(p, batch),
                   (const c_t*)G, npart, p, tol, zthr, n_inner, Qc, flags,
                   counters, active, offmax, misc + 2,
                   (cross_rounds && p > 1 && r > 0) ? 1 : 0,
                   (const int*)want_w, Wm ? (per_problem_w ? 2 : 1) : 0);
  QTN_TIMED_LAUNCH(ev[4], ev[5], (jac_update<R, WW>),
                   dim3(nchunk_x + (Wm ? nchunk_w : 0), p, batch), X, ldx,
                   sX, nchunk_x, Wm, ldw ? ldw : cols_p, sW, p, r,
                   (const c_t*)Qc, flags, active,
                   (const int*)(per_problem_w ? want_w : nullptr));
#undef QTN_TIMED_LAUNCH
  return QTN_OK;
}

// Adds the event times of the sweep just finished to the context's
// per-kernel accumulators with the algorithmic work of its launches:
//   jac_gram   flops 3/4 * 8*64*W^2 per (problem, pair, chunk); bytes: the
//              64 x W X tile read per chunk + the 3/4 W^2 partial Gram
//              written per workgroup (gram_parts of them per pair)
//   jac_inner  bytes: the partial Grams read + Q written per (problem, pair)
//   jac_update flops 8*64*W^2 per (pair with a rotation, chunk of X, and
//              chunk of W where the problem accumulates W); bytes: the tile
//              read and written + Q read
int32_t JacobiEngine::account_sweep(int n_active, int qpairs, int qpairs_w,
                                    int nchunk, int nchunk_w) {
  qtn_ctx* ctx = ctx_;
  if (fused && W == 32) {
    // jac_round (id 3 FP64, 7 FP32): the Gram of every active pair plus
    // X_IJ <- X_IJ Q for every rotated pair; bytes: X_IJ read once for the
    // Gram, read and written once more for the update.  jac_round_res reads X_IJ once and
    // writes it once, and reads and writes W_IJ once for the pairs that
    // rotate W.
    const double Wd = W, c16 = single ? 8.0 : 16.0;
    const double rows = (double)nchunk * 64.0;
    const bool res = res_round(nchunk, nchunk_w);
    const int kr = single ? 7 : 3;
    for (size_t i = 0; i + 5 < prof_events.size(); i += 6) {
      float t01;
      QTN_CHECK_HIP(hipEventElapsedTime(&t01, prof_events[i],
                                        prof_events[i + 1]));
      ctx->prof_count[kr] += 1;
      ctx->prof_ms[kr] += t01;
      ctx->prof_flops[kr] += (double)n_active * p * 0.75 * 8.0 * rows * Wd * Wd;
      ctx->prof_bytes[kr] += (double)n_active * p * rows * Wd * c16;
    }
    ctx->prof_flops[kr] += (double)qpairs * 8.0 * rows * Wd * Wd;
    ctx->prof_bytes[kr] += (double)qpairs * (res ? 1.0 : 2.0) * rows * Wd * c16;
    if (res) {
      const double wrows = (double)nchunk_w * 64.0;
      ctx->prof_flops[kr] += (double)qpairs_w * 8.0 * wrows * Wd * Wd;
      ctx->prof_bytes[kr] += (double)qpairs_w * 2.0 * wrows * Wd * c16;
    }
    for (auto e : prof_events) ctx->ev_pool.push_back(e);
    prof_events.clear();
    return QTN_OK;
  }
  const double Wd = W, c16 = single ? 8.0 : 16.0;
  const double rounds = 2.0 * p - 1.0;
  const int ki = single ? 4 : 0, kg = single ? 5 : 1, ku = single ? 6 : 2;
  // jac_gram reads each 64-row chunk of X_IJ and writes one partial Gram
  // per workgroup (npart of them); jac_inner reads those partials
  const double npart = gram_parts(nchunk);
  const double gram_f = 0.75 * 8.0 * 64.0 * Wd * Wd;
  const double gram_b =
      (64.0 * Wd + 0.75 * Wd * Wd * npart / (double)nchunk) * c16;
  const double inner_b = (npart * 0.75 * Wd * Wd + Wd * Wd) * c16;
  const double upd_f = 8.0 * 64.0 * Wd * Wd;
  const double upd_b = (2.0 * 64.0 * Wd + Wd * Wd) * c16;
  const double pairs_per_launch = (double)n_active * p;
  // per round: (start, stop) of jac_gram, jac_inner and the update (X and,
  // when there is one, W in the same launch)
  for (size_t i = 0; i + 5 < prof_events.size(); i += 6) {
    float t01, t12, t23;
    QTN_CHECK_HIP(hipEventElapsedTime(&t01, prof_events[i], prof_events[i + 1]));
    QTN_CHECK_HIP(hipEventElapsedTime(&t12, prof_events[i + 2], prof_events[i + 3]));
    QTN_CHECK_HIP(hipEventElapsedTime(&t23, prof_events[i + 4], prof_events[i + 5]));
    ctx->prof_count[kg] += 1;
    ctx->prof_ms[kg] += t01;
    ctx->prof_flops[kg] += pairs_per_launch * nchunk * gram_f;
    ctx->prof_bytes[kg] += pairs_per_launch * nchunk * gram_b;
    ctx->prof_count[ki] += 1;
    ctx->prof_ms[ki] += t12;
    ctx->prof_bytes[ki] += pairs_per_launch * inner_b;
    ctx->prof_count[ku] += 1;
    ctx->prof_ms[ku] += t23;
  }
  // rotated pairs are only known per sweep: spread over its launches
  ctx->prof_flops[ku] += ((double)qpairs * nchunk +
                          (double)qpairs_w * nchunk_w) * upd_f;
  ctx->prof_bytes[ku] += ((double)qpairs * nchunk +
                          (double)qpairs_w * nchunk_w) * upd_b;
  (void)rounds;
  for (auto e : prof_events) ctx->ev_pool.push_back(e);
  prof_events.clear();
  return QTN_OK;
}

int32_t JacobiEngine::run(qtn_ctx* ctx, double2* X, int64_t ldx, int64_t sX,
                          double2* Wm, int64_t sW, const int* want_w_host,
                          const int64_t* nsv_host, double cutoff,
                          int relative, int64_t chimax, int max_sweeps,
                          int* sweeps_done, const int* mask_host) {
  if (single) {
    set_error("svd: engine reserved for FP32 sweeps");
    return QTN_ERR_ARG;
  }
  return run_t<double>(ctx, X, ldx, sX, Wm, sW, want_w_host, nsv_host, cutoff,
                       relative, chimax, max_sweeps, sweeps_done, mask_host);
}

int32_t JacobiEngine::run32(qtn_ctx* ctx, float2* X, int64_t ldx, int64_t sX,
                            float2* Wm, int64_t sW, const int* want_w_host,
                            const int64_t* nsv_host, double cutoff,
                            int relative, int64_t chimax, int max_sweeps,
                            int* sweeps_done, const int* mask_host) {
  if (!single) {
    set_error("svd: engine reserved for FP64 sweeps");
    return QTN_ERR_ARG;
  }
  return run_t<float>(ctx, X, ldx, sX, Wm, sW, want_w_host, nsv_host, cutoff,
                      relative, chimax, max_sweeps, sweeps_done, mask_host);
}

// FP64: tol = sqrt(rows) eps / 2 and zero columns below rows eps |X|_F
// (before the sweeps: never rotated; after them: exact zeros in sig).
// FP32 (the chosen form's sweeps): tol = sqrt(rows) eps32 and columns below
// 8 sqrt(rows) eps32 |X|_F (the rounding noise a column gathers over the
// sweeps) are left alone.
template <typename R>
int32_t JacobiEngine::run_t(qtn_ctx* ctx, typename CType<R>::type* X,
                            int64_t ldx, int64_t sX,
                            typename CType<R>::type* Wm, int64_t sW,
                            const int* want_w_host, const int64_t* nsv_host,
                            double cutoff, int relative, int64_t chimax,
                            int max_sweeps, int* sweeps_done,
                            const int* mask_host) {
  using c_t = typename JT<R>::c_t;
  hipStream_t s = ctx->stream;
  const bool dbl = sizeof(R) == sizeof(double);
  const double tol = dbl ? std::sqrt((double)rows_p) * (DBL_EPSILON * 0.5)
                         : std::sqrt((double)rows_p) * FLT_EPSILON;
  const double zrel = dbl ? (double)rows_p * DBL_EPSILON
                          : 8.0 * std::sqrt((double)rows_p) * FLT_EPSILON;
  const int nb = 2 * p;
  const int rounds = nb - 1;
  // inner sweeps per block pair: the whole problem when it is one block
  // pair, otherwise QTN_JACOBI_INNER (default 1)
  int n_inner_big = 1;
  if (const char* e = getenv("QTN_JACOBI_INNER")) n_inner_big = std::max(1, atoi(e));
  const int n_inner = (p == 1) ? 40 : n_inner_big;
  // stop once a sweep's largest scaled off-diagonal is below sqrt(tol)
  const double conv2 = tol;
  const int nchunk_x = (int)(rows_p / 64);
  const int nchunk_w = (int)(cols_p / 64);

  QTN_CHECK_HIP(hipMemcpyAsync(nsv, nsv_host, batch * sizeof(int64_t),
                               hipMemcpyHostToDevice, s));
  if (want_w_host)
    QTN_CHECK_HIP(hipMemcpyAsync(want_w, want_w_host, batch * sizeof(int),
                                 hipMemcpyHostToDevice, s));
  if (mask_host)
    QTN_CHECK_HIP(hipMemcpyAsync(mask, mask_host, batch * sizeof(int),
                                 hipMemcpyHostToDevice, s));
  if (chimax_per)
    QTN_CHECK_HIP(hipMemcpyAsync(chimax_dev, chimax_per,
                                 batch * sizeof(int64_t),
                                 hipMemcpyHostToDevice, s));
  QTN_CHECK_HIP(hipMemsetAsync(misc, 0, 4 * sizeof(int), s));
  ctx_ = ctx;
  prof_events.clear();
  int n_act = batch;
  if (mask_host) {
    n_act = 0;
    for (int b = 0; b < batch; ++b) n_act += mask_host[b] ? 1 : 0;
  }
  const int n_solved = n_act;
  QTN_CHECK_HIP(hipMemsetAsync(offmax, 0, batch * sizeof(double), s));
  hipLaunchKernelGGL(jac_prep_partial<R>, dim3(kPrepBlocks, batch),
                     dim3(256), 0, s, (const c_t*)X, rows_p, ldx, cols_p, sX,
                     (const int*)(mask_host ? mask : nullptr), prep_scale,
                     prep_ssq, prep_bad);
  hipLaunchKernelGGL(jac_prep_final, dim3((batch + 63) / 64), dim3(64), 0, s,
                     batch, zrel, (const double*)prep_scale,
                     (const double*)prep_ssq, (const int*)prep_bad,
                     (const int*)(mask_host ? mask : nullptr), normF2, zthr,
                     active, counters, misc + 1, scale_e);
  QTN_CHECK_HIP(hipGetLastError());
  const unsigned sgrid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((rows_p * cols_p + 255) / 256, 2048));
  hipLaunchKernelGGL(jac_scale<R>, dim3(sgrid, batch), dim3(256), 0, s, X,
                     rows_p, ldx, cols_p, sX, (const int*)scale_e, -1);
  QTN_CHECK_HIP(hipGetLastError());
  if (Wm) {
    // W starts as the identity on the padded columns.
    QTN_TRY(launch_identity_t<R>(s, Wm, cols_p, ldw ? ldw : cols_p, sW,
                                 batch));
  }
  int* host_flags = nullptr;
  QTN_TRY(ctx->pinned_buffer(4 * sizeof(int), (void**)&host_flags));
  // One sweep: every round, then the convergence bookkeeping.
  auto enqueue_sweep = [&]() -> int32_t {
    for (int r = 0; r < rounds; ++r) {
      if (this->W == 64)
        QTN_TRY((launch_round<R, 64>(s, X, ldx, sX, Wm, sW, r, tol, n_inner,
                                     nchunk_x, nchunk_w,
                                     want_w_host != nullptr)));
      else
        QTN_TRY((launch_round<R, 32>(s, X, ldx, sX, Wm, sW, r, tol, n_inner,
                                     nchunk_x, nchunk_w,
                                     want_w_host != nullptr)));
    }
    QTN_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(jac_sweep_end, dim3(1), dim3(256), 0, s, active,
                       counters, offmax, conv2, batch, misc);
    QTN_CHECK_HIP(hipGetLastError());
    return QTN_OK;
  };
  // The sweep's launches are captured once into a graph (kernel nodes
  // only) and replayed; graphs are cached on the context by every value
  // the launches take, so a later call with the same buffers and shapes
  // replays without capturing.  Not while profiling (per-launch events).
  hipGraphExec_t exec = nullptr;
  if (!ctx->prof && sweep_graphs_enabled()) {
    SweepKey key;
    std::memset(&key, 0, sizeof(key));
    key.dbl = dbl ? 1 : 0;
    key.w = this->W;
    key.fused = fused ? 1 : 0;
    key.cross = cross_rounds ? 1 : 0;
    key.batch = batch;
    key.p = p;
    key.n_inner = n_inner;
    key.nchunk_x = nchunk_x;
    key.nchunk_w = nchunk_w;
    key.per_problem_w = want_w_host != nullptr ? 1 : 0;
    key.ldx = ldx;
    key.sx = sX;
    key.sw = sW;
    key.ldw = ldw;
    key.tol = tol;
    key.conv2 = conv2;
    const void* ptrs[] = {X, Wm, want_w, zthr, counters, active, offmax,
                          misc, Gp, Q, flags};
    static_assert(sizeof(ptrs) == sizeof(key.ptr), "SweepKey pointers");
    std::memcpy(key.ptr, ptrs, sizeof(ptrs));
    QTN_TRY(ctx->sweep_graph(&key, sizeof(key), s, enqueue_sweep, &exec));
  }
  int sweep = 0;
  bool converged = false;
  for (; sweep < max_sweeps; ++sweep) {
    if (exec)
      QTN_CHECK_HIP(hipGraphLaunch(exec, s));
    else
      QTN_TRY(enqueue_sweep());
    QTN_CHECK_HIP(hipMemcpyAsync(host_flags, misc + 4, 4 * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    QTN_CHECK_HIP(hipStreamSynchronize(s));
    if (ctx->prof)
      QTN_TRY(account_sweep(n_act, host_flags[2], host_flags[3], nchunk_x,
                            nchunk_w));
    n_act = host_flags[0];
    if (host_flags[1]) {
      set_error("svd: input contains NaN or Inf");
      return QTN_ERR_NOTFINITE;
    }
    if (host_flags[0] == 0) {
      converged = true;
      ++sweep;
      break;
    }
  }
  if (sweeps_done) *sweeps_done = sweep;
  ctx->n_svd += n_solved;
  ctx->n_sweeps += (int64_t)sweep * n_solved;
  ctx->max_sweeps = std::max<int64_t>(ctx->max_sweeps, sweep);
  ctx->n_batches += 1;
  if (!converged) {
    set_error("svd: one-sided Jacobi did not converge in " +
              std::to_string(max_sweeps) + " sweeps");
    return QTN_ERR_NOCONV;
  }
  hipLaunchKernelGGL(jac_colnorms<R>,
                     dim3((unsigned)((cols_p + 3) / 4), batch), dim3(256), 0,
                     s, (const c_t*)X, rows_p, ldx, cols_p, sX, sig,
                     (const double*)zthr, (const int*)scale_e);
  hipLaunchKernelGGL(jac_scale<R>, dim3(sgrid, batch), dim3(256), 0, s, X,
                     rows_p, ldx, cols_p, sX, (const int*)scale_e, 1);
  QTN_CHECK_HIP(hipGetLastError());
  if (cols_p > kSortMax) {
    set_error("svd: more than 4096 columns are not supported");
    return QTN_ERR_ARG;
  }
  hipLaunchKernelGGL(jac_sort_trunc, dim3(batch), dim3(1024), 0, s, sig, perm,
                     cols_p, (const int64_t*)nsv, cutoff, relative, chimax,
                     kept, disc, total, keptw, status, (const int*)scale_e,
                     (const int64_t*)(chimax_per ? chimax_dev : nullptr));
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}

namespace {
template <typename R>
__global__ void identity_kernel(typename JT<R>::c_t* W, int64_t n,
                                int64_t ld, int64_t sW) {
  const int64_t bz = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i % n, c = i / n;
    W[bz * sW + r + ld * c] = JT<R>::mk(r == c ? (R)1 : (R)0, (R)0);
  }
}
}  // namespace

template <typename R>
int32_t launch_identity_t(hipStream_t s, typename CType<R>::type* W,
                          int64_t n, int64_t ld, int64_t sW, int batch) {
  const int64_t blocks = std::min<int64_t>((n * n + 255) / 256, 4096);
  hipLaunchKernelGGL(identity_kernel<R>, dim3((unsigned)blocks, batch),
                     dim3(256), 0, s, W, n, ld, sW);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
}
template int32_t launch_identity_t<double>(hipStream_t, double2*, int64_t,
                                           int64_t, int64_t, int);
template int32_t launch_identity_t<float>(hipStream_t, float2*, int64_t,
                                          int64_t, int64_t, int);

int32_t launch_identity(hipStream_t s, double2* W, int64_t n, int64_t sW,
                        int batch) {
  return launch_identity_t<double>(s, W, n, n, sW, batch);
}

namespace {

// batch problems of rows_p x cols_p: complex Gaussian entries (counter
// hash of the flat index, Box-Muller) in the leading rows x cols block,
// zeros in the padding.
template <typename R>
__global__ void tune_fill(typename JT<R>::c_t* __restrict__ X, int64_t rows,
                          int64_t cols, int64_t rows_p, int64_t cols_p,
                          int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i % rows_p, c = (i / rows_p) % cols_p;
    if (r >= rows || c >= cols) {
      X[i] = JT<R>::mk((R)0, (R)0);
      continue;
    }
    uint64_t z = 0x7a11ULL ^ ((uint64_t)i * 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    const double u1 = ((double)(z >> 40) + 0.5) * (1.0 / 16777216.0);
    const double u2 = (double)(z & 0xffffff) * (1.0 / 16777216.0);
    const double rad = sqrt(-log(u1));
    X[i] = JT<R>::mk((R)(rad * cos(6.283185307179586 * u2)),
                     (R)(rad * sin(6.283185307179586 * u2)));
  }
}

constexpr int kSlotTuneX = 180, kSlotTuneX0 = 181, kSlotTuneW = 182;

// Time of one batched SVD (ms, best of `reps`) with tiling `code` forced
// for the bucket; X is restored from X0 before every run.
template <typename R>
int32_t tune_time(qtn_ctx* ctx, int code, int batch, int64_t rows_p,
                  int64_t cols_p, int64_t nsv, typename CType<R>::type* X,
                  const typename CType<R>::type* X0,
                  typename CType<R>::type* Wm, int reps, double* ms) {
  const bool single = sizeof(R) == sizeof(float);
  ctx->jac_tiling[{cols_p, single ? 1 : 0, batch_class(batch)}] = code;
  hipStream_t s = ctx->stream;
  const size_t xbytes =
      (size_t)batch * rows_p * cols_p * sizeof(typename CType<R>::type);
  std::vector<int64_t> nsvv(batch, nsv);
  hipEvent_t e0, e1;
  QTN_CHECK_HIP(hipEventCreate(&e0));
  QTN_CHECK_HIP(hipEventCreate(&e1));
  double best = -1.0;
  for (int r = 0; r <= reps; ++r) {  // run 0 warms up (buffers, code)
    QTN_CHECK_HIP(hipMemcpyAsync(X, X0, xbytes, hipMemcpyDeviceToDevice, s));
    JacobiEngine eng;
    eng.slot_base = 185;
    QTN_TRY(eng.reserve(ctx, batch, rows_p, cols_p, single));
    int sweeps = 0;
    QTN_CHECK_HIP(hipEventRecord(e0, s));
    const int32_t st = eng.run_t<R>(ctx, X, rows_p, rows_p * cols_p, Wm,
                                    cols_p * cols_p, nullptr, nsvv.data(),
                                    1e-8, 1, 0, 60, &sweeps, nullptr);
    QTN_CHECK_HIP(hipEventRecord(e1, s));
    QTN_CHECK_HIP(hipEventSynchronize(e1));
    if (st != QTN_OK) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return st;
    }
    float t = 0.f;
    QTN_CHECK_HIP(hipEventElapsedTime(&t, e0, e1));
    if (r > 0 && (best < 0 || t < best)) best = t;
  }
  QTN_CHECK_HIP(hipEventDestroy(e0));
  QTN_CHECK_HIP(hipEventDestroy(e1));
  *ms = best;
  return QTN_OK;
}

template <typename R>
int32_t tune_bucket(qtn_ctx* ctx, int64_t rows, int64_t cols, int batch,
                    int* best_code, double* us) {
  using C = typename CType<R>::type;
  const bool single = sizeof(R) == sizeof(float);
  const int64_t rows_p = round_up64(std::max(rows, cols));
  const int64_t cols_p = round_up64(cols);
  const int64_t nsv = std::min(rows, cols);
  const size_t xn = (size_t)batch * rows_p * cols_p;
  void *px, *px0, *pw;
  QTN_TRY(ctx->buffer(kSlotTuneX, xn * sizeof(C), &px));
  QTN_TRY(ctx->buffer(kSlotTuneX0, xn * sizeof(C), &px0));
  QTN_TRY(ctx->buffer(kSlotTuneW, (size_t)batch * cols_p * cols_p * sizeof(C),
                      &pw));
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(tune_fill<R>,
                     dim3((unsigned)std::min<int64_t>((xn + 255) / 256, 8192)),
                     dim3(256), 0, s, (C*)px0, rows, cols, rows_p, cols_p,
                     (int64_t)xn);
  QTN_CHECK_HIP(hipGetLastError());
  const int codes[3] = {QTN_TILING_W32, QTN_TILING_W32_FUSED,
                        QTN_TILING_W64};
  double best = -1.0;
  int bc = QTN_TILING_W32;
  for (int i = 0; i < 3; ++i) {
    double ms = -1.0;
    // one 64-column bucket has a single tiling (one block pair)
    const bool applies = cols_p > 64 || codes[i] == QTN_TILING_W64;
    if (applies)
      QTN_TRY(tune_time<R>(ctx, codes[i], batch, rows_p, cols_p, nsv, (C*)px,
                           (const C*)px0, (C*)pw, 2, &ms));
    if (us) us[i] = ms >= 0 ? ms * 1e3 / batch : -1.0;
    if (ms >= 0 && (best < 0 || ms < best)) {
      best = ms;
      bc = codes[i];
    }
  }
  ctx->jac_tiling[{cols_p, single ? 1 : 0, batch_class(batch)}] = bc;
  *best_code = bc;
  return QTN_OK;
}

}  // namespace

}  // namespace qtn

using namespace qtn;

extern "C" {

int32_t qtn_tune_jacobi(qtn_ctx* ctx, int64_t rows, int64_t cols,
                        int32_t batch, int32_t single, int32_t* best,
                        double* us) try {
  clear_error();
  QTN_ARG(ctx && best, "qtn_tune_jacobi: NULL argument");
  QTN_ARG(rows >= 1 && cols >= 1 && rows <= 4096 && cols <= 4096,
          "qtn_tune_jacobi: sizes must lie in [1, 4096]");
  QTN_ARG(batch >= 1 && batch <= 1024, "qtn_tune_jacobi: batch in [1, 1024]");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  int code = 0;
  if (single)
    QTN_TRY(tune_bucket<float>(ctx, rows, cols, batch, &code, us));
  else
    QTN_TRY(tune_bucket<double>(ctx, rows, cols, batch, &code, us));
  *best = code;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_set_jacobi_tiling(qtn_ctx* ctx, int64_t cols, int32_t single,
                              int32_t batch, int32_t tiling) try {
  clear_error();
  QTN_ARG(ctx, "qtn_set_jacobi_tiling: NULL context");
  QTN_ARG(cols >= 1 && cols <= 4096,
          "qtn_set_jacobi_tiling: cols must lie in [1, 4096]");
  QTN_ARG(tiling == QTN_TILING_W32 || tiling == QTN_TILING_W32_FUSED ||
              tiling == QTN_TILING_W64,
          "qtn_set_jacobi_tiling: unknown tiling");
  QTN_ARG(batch >= 1, "qtn_set_jacobi_tiling: batch must be positive");
  ctx->jac_tiling[{round_up64(cols), single ? 1 : 0, batch_class(batch)}] =
      tiling;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
