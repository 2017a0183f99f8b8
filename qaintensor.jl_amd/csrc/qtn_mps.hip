// Device-resident matrix-product states and the two-site truncated update.
//
// Gauge: right-canonical site tensors B_j (chi_{j-1}, 2, chi_j) with the bond
// spectra lambda_j kept beside them (the "Vidal-free" TEBD form of Hastings,
// J. Math. Phys. 50, 095207 (2009)).  A gate on sites (i, i+1) is
//   Phi   = g . (B_i B_{i+1})                    rows (a,s'), cols (t',b)
//   theta = diag(lambda_{i-1} (x) 1_2) Phi
//   theta = U S V^H, k = truncation rule (src/svd.jl:29-33, relative er)
//   B_{i+1} <- V_k^H,  B_i <- Phi V_k / |S_k|,  lambda_i <- S_k / |S_k|
// so no division by small Schmidt values is ever needed and the gates of a
// brick-wall layer are independent (one batched pass per layer).  The SVD is
// taken of X = theta^H: its orthogonalised columns are V_k S_k directly.
// Site-index conventions follow switch! (src/switch.jl:26-55): theta rows are
// (a, s) with a fastest, columns (t, b) with t fastest, column-major.
#include <algorithm>
#include <set>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "qtn_internal.hpp"
#include "qtn_qr.hpp"
#include "qtn_svd.hpp"


namespace qtn {
namespace {

constexpr int kSlot1q = 47;
constexpr int kSlotPhi = 30, kSlotXm = 31, kSlotVk = 32, kSlotDesc = 33,
              kSlotGates = 34, kSlotEnv = 35, kSlotT1 = 36, kSlotT2 = 37,
              kSlotRed = 38, kSlotBits = 39, kSlotAmp = 40, kSlotW2 = 41,
              kSlotA1 = 42, kSlotA2 = 43, kSlotQd = 44, kSlotNrm = 45,
              kSlotPerm2 = 46, kSlotX32 = 70, kSlotW32 = 71, kSlotNsG = 72,
              kSlotNsT = 73, kSlotS64 = 74, kSlotMix = 75, kSlotEng32 = 80;

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                          8192));
}

int64_t pow2_cap(int e, int64_t cap) {
  if (e >= 62) return cap;
  return std::min<int64_t>(cap, (int64_t)1 << e);
}

// Per-gate data of a batched two-site update.
struct GateDesc {
  double2* Bl;        // site i tensor
  double2* Br;        // site i+1 tensor
  const double* lamL; // lambda_{i-1}
  double* lamM;       // lambda_i (written)
  double2* phi;       // Phi (M x N)
  double2* X;         // theta^H (trans) or theta, zero padded
  double2* W;         // accumulated rotations when !trans (V = W)
  double2* Vk;        // V_k (N x k)
  const double2* g;   // 4x4 gate
  int64_t chiL, chiM, chiR;
  int trans;          // 1: X = theta^H (M <= N), 0: X = theta (M > N)
  int mode;           // V_k from: 0 X columns, 1 Q1 [X columns; 0], 2 W
  int64_t ldx;        // leading dimension of X
};

// Phi <- g . Phi in place and X = (lambda_a Phi)^H (or lambda_a Phi when the
// gate's X is not transposed); one thread per (a, b) handles the four (s, t)
// entries.  grid (tiles, batch).
__global__ void gate_epilogue(const GateDesc* __restrict__ gd,
                              int64_t rows_p) {
  const GateDesc d = gd[blockIdx.y];
  const int64_t chiL = d.chiL, chiR = d.chiR;
  const int64_t M = 2 * chiL;
  const int64_t total = chiL * chiR;
  double2 g[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) g[i] = d.g[i];
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = idx % chiL, b = idx / chiL;
    double2 in[4], out[4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 2; ++s)
        in[s + 2 * t] = d.phi[(a + chiL * s) + M * (t + 2 * b)];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double2 gv = g[r + 4 * c];
        re += gv.x * in[c].x - gv.y * in[c].y;
        im += gv.x * in[c].y + gv.y * in[c].x;
      }
      out[r] = make_double2(re, im);
    }
    const double la = d.lamL[a];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const double2 v = out[s + 2 * t];
        const int64_t row = a + chiL * s, col = t + 2 * b;
        d.phi[row + M * col] = v;
        if (d.trans)
          d.X[col + d.ldx * row] = make_double2(la * v.x, -la * v.y);
        else
          d.X[row + d.ldx * col] = make_double2(la * v.x, la * v.y);
      }
  }
}

// V_k (N x k, ld N): mode 0: X[:, perm[j]] / sig[j]; mode 1: the first M
// rows from X[0:M, perm[j]] / sig[j] and zeros below (Q1 is applied
// afterwards); mode 2: W[:, perm[j]].  grid (tiles, batch).
__global__ void extract_vk(const GateDesc* __restrict__ gd,
                           const double* __restrict__ sig,
                           const int* __restrict__ perm,
                           const int64_t* __restrict__ kept, int64_t cols_p) {
  const int bz = blockIdx.y;
  const GateDesc d = gd[bz];
  const int64_t k = kept[bz];
  const int64_t N = 2 * d.chiR, M = 2 * d.chiL;
  const double* s = sig + (int64_t)bz * cols_p;
  const int* pm = perm + (int64_t)bz * cols_p;
  const int64_t total = N * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx % N, j = idx / N;
    double2 v = make_double2(0.0, 0.0);
    if (d.mode == 2) {
      v = d.W[c + cols_p * (int64_t)pm[j]];
    } else if (d.mode == 0 || c < M) {
      const double sj = s[j];
      const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
      const double2 x = d.X[c + d.ldx * (int64_t)pm[j]];
      v = make_double2(x.x * inv, x.y * inv);
    }
    d.Vk[c + N * j] = v;
  }
}

// B_{i+1}[j, c] = conj(V_k[c, j]) (k x N, ld k); lambda_i = sig / |S_k|.
__global__ void write_right(const GateDesc* __restrict__ gd,
                            const double* __restrict__ sig,
                            const int64_t* __restrict__ kept,
                            const double* __restrict__ keptw,
                            int64_t cols_p, int write_lambda) {
  const int bz = blockIdx.y;
  const GateDesc d = gd[bz];
  const int64_t k = kept[bz];
  const int64_t N = 2 * d.chiR;
  const int64_t total = N * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx % N, j = idx / N;
    const double2 v = d.Vk[c + N * j];
    d.Br[j + k * c] = make_double2(v.x, -v.y);
  }
  if (write_lambda && blockIdx.x == 0) {
    const double* s = sig + (int64_t)bz * cols_p;
    const double nrm = sqrt(keptw[bz]);
    for (int64_t j = threadIdx.x; j < k; j += blockDim.x)
      d.lamM[j] = s[j] / nrm;
  }
}

// A 2x2 complex matrix passed by value (column-major).
struct Mat2 {
  double2 u[4];
};

// Single-qubit gate on a site tensor in place: B[a, s', b] = u[s', s] B[a,s,b]
__global__ void apply_1q_kernel(double2* __restrict__ B, int64_t chiL,
                                int64_t chiR, Mat2 g) {
  const double2 u00 = g.u[0], u10 = g.u[1], u01 = g.u[2], u11 = g.u[3];
  const int64_t total = chiL * chiR;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = idx % chiL, b = idx / chiL;
    double2* p0 = B + a + chiL * (0 + 2 * b);
    double2* p1 = B + a + chiL * (1 + 2 * b);
    const double2 x0 = *p0, x1 = *p1;
    *p0 = make_double2(u00.x * x0.x - u00.y * x0.y + u01.x * x1.x - u01.y * x1.y,
                       u00.x * x0.y + u00.y * x0.x + u01.x * x1.y + u01.y * x1.x);
    *p1 = make_double2(u10.x * x0.x - u10.y * x0.y + u11.x * x1.x - u11.y * x1.y,
                       u10.x * x0.y + u10.y * x0.x + u11.x * x1.y + u11.y * x1.x);
  }
}

// A batch of single-qubit gates, one per descriptor (grid (tiles, batch)).
struct OneQDesc {
  double2* B;
  int64_t chiL, chiR;
  double2 u[4];
};

__global__ void apply_1q_multi_kernel(const OneQDesc* __restrict__ dd) {
  const OneQDesc d = dd[blockIdx.y];
  const double2 u00 = d.u[0], u10 = d.u[1], u01 = d.u[2], u11 = d.u[3];
  const int64_t total = d.chiL * d.chiR;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = idx % d.chiL, b = idx / d.chiL;
    double2* p0 = d.B + a + d.chiL * (0 + 2 * b);
    double2* p1 = d.B + a + d.chiL * (1 + 2 * b);
    const double2 x0 = *p0, x1 = *p1;
    *p0 = make_double2(u00.x * x0.x - u00.y * x0.y + u01.x * x1.x - u01.y * x1.y,
                       u00.x * x0.y + u00.y * x0.x + u01.x * x1.y + u01.y * x1.x);
    *p1 = make_double2(u10.x * x0.x - u10.y * x0.y + u11.x * x1.x - u11.y * x1.y,
                       u10.x * x0.y + u10.y * x0.x + u11.x * x1.y + u11.y * x1.x);
  }
}

// Amplitude contraction step: V'[i][c] = sum_a V[i][a] B[a, s_i, c].
// V is nb x chiL (ld nb), V' is nb x chiR.
__global__ void amp_step(const double2* __restrict__ V, int64_t nb,
                         int64_t chiL, int64_t chiR,
                         const double2* __restrict__ B,
                         const uint8_t* __restrict__ bits, int nsites,
                         int site0, double2* __restrict__ Vn) {
  const int64_t total = nb * chiR;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % nb, c = idx / nb;
    const int s = bits[i * nsites + site0] ? 1 : 0;
    double re = 0.0, im = 0.0;
    const double2* col = B + chiL * (s + 2 * c);
    for (int64_t a = 0; a < chiL; ++a) {
      const double2 v = V[i + nb * a];
      const double2 b = col[a];
      re += v.x * b.x - v.y * b.y;
      im += v.x * b.y + v.y * b.x;
    }
    Vn[idx] = make_double2(re, im);
  }
}

// sum over (a, s, b) of w_s conj(B[(a,s),b]) Y[(a,s),b] for w = (1,1) and
// w = (1,-1): out[0..1] = norm (complex), out[2..3] = Z moment (complex).
__global__ __launch_bounds__(256) void env_reduce(
    const double2* __restrict__ B, const double2* __restrict__ Y,
    int64_t chiL, int64_t chiR, double* __restrict__ out) {
  const int64_t total = 2 * chiL * chiR;
  double nr = 0, ni = 0, zr = 0, zi = 0;
  for (int64_t idx = threadIdx.x; idx < total; idx += 256) {
    const int64_t row = idx % (2 * chiL);
    const int s = row >= chiL ? 1 : 0;
    const double2 b = B[idx], y = Y[idx];
    const double pr = b.x * y.x + b.y * y.y;  // conj(b) * y
    const double pi = b.x * y.y - b.y * y.x;
    nr += pr;
    ni += pi;
    zr += s ? -pr : pr;
    zi += s ? -pi : pi;
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = nr;
  red[1][threadIdx.x] = ni;
  red[2][threadIdx.x] = zr;
  red[3][threadIdx.x] = zi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

// ---- chosen form: FP32 sweeps, FP64 re-orthogonalisation ----------------
// Y = (complex float) X over n elements.
__global__ void to_single(const double2* __restrict__ X,
                          float2* __restrict__ Y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = X[i];
    Y[i] = make_float2((float)v.x, (float)v.y);
  }
}

// V_k (N x k, ld N, FP64) per gate from the engine that solved it:
// src 1 (FP32 sweeps): mode 0 X32[:, perm32[j]] / sig32[j], mode 2
// W32[:, perm32[j]]; src 0 (FP64 engine): mode 0 X[:, perm[j]] / sig[j],
// mode 2 W[:, perm[j]].  X32 / W32 of gate bz start at bz * x_sz / w_sz;
// X and W have the batch layout (ld d.ldx / cols_p).  grid (tiles, batch).
__global__ void extract_vk_mixed(const GateDesc* __restrict__ gd,
                                 const int* __restrict__ src,
                                 const float2* __restrict__ X32,
                                 int64_t x_sz,
                                 const float2* __restrict__ W32,
                                 int64_t w_sz,
                                 const double* __restrict__ sig32,
                                 const int* __restrict__ perm32,
                                 const double* __restrict__ sig64,
                                 const int* __restrict__ perm64,
                                 const int64_t* __restrict__ kept,
                                 int64_t cols_p, int64_t cols32,
                                 int64_t cols64) {
  const int bz = blockIdx.y;
  if (src[bz] == 2) return;  // V_k written by the randomized range finder
  const GateDesc d = gd[bz];
  const int64_t k = kept[bz];
  const int64_t N = 2 * d.chiR;
  const bool single = src[bz] != 0;
  // sig / perm rows are as wide as the engine that solved the gate
  const int64_t sc = single ? cols32 : cols64;
  const double* s = (single ? sig32 : sig64) + (int64_t)bz * sc;
  const int* pm = (single ? perm32 : perm64) + (int64_t)bz * sc;
  const int64_t total = N * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx % N, j = idx / N;
    const int64_t col = pm[j];
    double2 v;
    if (d.mode == 2) {
      if (single) {
        const float2 w = W32[bz * w_sz + c + cols_p * col];
        v = make_double2(w.x, w.y);
      } else {
        v = d.W[c + cols_p * col];
      }
    } else {
      const double sj = s[j];
      const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
      double2 x;
      if (single) {
        const float2 xf = X32[bz * x_sz + c + d.ldx * col];
        x = make_double2(xf.x, xf.y);
      } else {
        x = d.X[c + d.ldx * col];
      }
      v = make_double2(x.x * inv, x.y * inv);
    }
    d.Vk[c + N * j] = v;
  }
}

// Newton-Schulz step V <- 1.5 V - 0.5 T with T = V (V^H V) (N x k, ld N).
__global__ void ns_combine(const GateDesc* __restrict__ gd,
                           const double2* __restrict__ T, int64_t t_sz,
                           const int64_t* __restrict__ kept) {
  const int bz = blockIdx.y;
  const GateDesc d = gd[bz];
  const int64_t total = 2 * d.chiR * kept[bz];
  const double2* t = T + bz * t_sz;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = d.Vk[idx], w = t[idx];
    d.Vk[idx] = make_double2(1.5 * v.x - 0.5 * w.x, 1.5 * v.y - 0.5 * w.y);
  }
}

// sig64[bz][j] = |diag(lambda_{i-1} (x) 1) C[:, j]| for C = Phi V_k, held in
// B_i (2 chiL x k, ld 2 chiL).  One wave per column; grid (ceil(k/4), batch).
__global__ __launch_bounds__(256) void left_colnorms(
    const GateDesc* __restrict__ gd, const int64_t* __restrict__ kept,
    double* __restrict__ sig64, int64_t cols_p) {
  const int bz = blockIdx.y;
  const GateDesc d = gd[bz];
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= kept[bz]) return;
  const int64_t M = 2 * d.chiL;
  const double2* c = d.Bl + M * j;
  double acc = 0.0;
  for (int64_t r = lane; r < M; r += 64) {
    const double2 v = c[r];
    const double la = d.lamL[r % d.chiL];
    acc += la * la * (v.x * v.x + v.y * v.y);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) sig64[bz * cols_p + j] = sqrt(acc);
}

// B_i <- C / |S_k| and lambda_i = S_k / |S_k| with S_k from left_colnorms;
// the kept weight is summed smallest-first.  grid (tiles, batch).
__global__ __launch_bounds__(256) void left_finish(
    const GateDesc* __restrict__ gd, const int64_t* __restrict__ kept,
    const double* __restrict__ sig64, int64_t cols_p) {
  const int bz = blockIdx.y;
  const GateDesc d = gd[bz];
  const int64_t k = kept[bz];
  const double* s = sig64 + bz * cols_p;
  __shared__ double nrm;
  if (threadIdx.x == 0) {
    double w = 0.0;
    for (int64_t j = k - 1; j >= 0; --j) w += s[j] * s[j];
    nrm = sqrt(w);
  }
  __syncthreads();
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  const int64_t total = 2 * d.chiL * k;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = d.Bl[idx];
    d.Bl[idx] = make_double2(v.x * inv, v.y * inv);
  }
  if (blockIdx.x == 0)
    for (int64_t j = threadIdx.x; j < k; j += blockDim.x)
      d.lamM[j] = s[j] * inv;
}

// ---- learned form: randomized range finder ------------------------------
// Per-gate view of the range finder's operands: theta is M x N; its X is
// theta^H (trans) or theta, ld ldx.
struct RandDesc {
  const double2* X;  // the gate's X
  int64_t ldx;
  int trans;
  int64_t M, N, l;   // theta dims, sketch width
  double2* Om;       // N x l sketch (ld: the batch's largest N)
  double2* Y;        // M x l (ld M)
  double2* Z;        // N x l (ld N)
  double2* Q;        // rows x l (ld: the batch's largest max(M, N))
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Om (N x l of the gate, ld ldo: the batch-wide leading dimension the
// products read it with) = complex Gaussian (N(0,1/2) + i N(0,1/2)) from a
// counter hash and the Box-Muller transform; deterministic in (seed, gate,
// element).
__global__ void gauss_fill(const RandDesc* __restrict__ rd, uint64_t seed,
                           int64_t ldo) {
  const int bz = blockIdx.y;
  const RandDesc d = rd[bz];
  const int64_t total = d.N * d.l;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h1 = mix64(seed ^ mix64(((uint64_t)bz << 40) ^ (uint64_t)i));
    const uint64_t h2 = mix64(h1);
    const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-log(u1));  // sqrt(-2 ln u) / sqrt(2)
    const double t = 6.283185307179586 * u2;
    d.Om[(i % d.N) + ldo * (i / d.N)] = make_double2(r * cos(t), r * sin(t));
  }
}

// Q = [I_l; 0] (rows x l, ld ldq, the batch-wide leading dimension of Q)
// for every gate (rows = M or N by `which`).
__global__ void eye_fill(const RandDesc* __restrict__ rd, int which,
                         int64_t ld) {
  const int bz = blockIdx.y;
  const RandDesc d = rd[bz];
  const int64_t rows = which ? d.N : d.M;
  const int64_t total = rows * d.l;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i % rows, c = i / rows;
    d.Q[r + ld * c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
}

// |theta|_F^2 per gate (one workgroup per gate) from X.
__global__ __launch_bounds__(256) void theta_norm2(
    const RandDesc* __restrict__ rd, double* __restrict__ out) {
  const RandDesc d = rd[blockIdx.x];
  const int64_t rows = d.trans ? d.N : d.M, cols = d.trans ? d.M : d.N;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < rows * cols; i += 256) {
    const double2 v = d.X[(i % rows) + (i / rows) * d.ldx];
    acc += v.x * v.x + v.y * v.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// |X - P|_F^2 per gate (one workgroup per gate): the weight of theta
// outside the range of Q, P = Q Q^H theta in X's orientation (rows_x x
// cols_x, ld ldp, batch stride sP).  Summed entry by entry, so it keeps its
// relative accuracy however small it is (|theta|^2 - |Q^H theta|^2 would
// cancel to rounding noise at the 1e-16 level a cutoff of 1e-8 asks for).
__global__ __launch_bounds__(256) void resid_norm2(
    const RandDesc* __restrict__ rd, const double2* __restrict__ P,
    int64_t ldp, int64_t sP, double* __restrict__ out) {
  const RandDesc d = rd[blockIdx.x];
  const int64_t rows = d.trans ? d.N : d.M, cols = d.trans ? d.M : d.N;
  const double2* p = P + (int64_t)blockIdx.x * sP;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < rows * cols; i += 256) {
    const int64_t r = i % rows, c = i / rows;
    const double2 x = d.X[r + c * d.ldx], y = p[r + c * ldp];
    const double dx = x.x - y.x, dy = x.y - y.y;
    acc += dx * dx + dy * dy;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// The exact check.  With sigma_j the singular values of B = Q^H theta
// (sorted) and r = |theta - Q B|_F^2, truncating theta to the top k
// directions of Q B loses exactly tail(k) = r + sum_{j >= k} sigma_j^2.
// k = the fewest values with tail(k) <= (cutoff |theta|_F)^2 (the rule of
// the full SVD, sums taken from the smallest term up); accepted when such k
// exists below the sketch width l and within the cap chimax, else the gate
// falls back to the full SVD.  One thread per gate.
__global__ void rand_decide(int ng, const double* __restrict__ sig,
                            int64_t cols_p, const double* __restrict__ total,
                            const double* __restrict__ resid,
                            const int64_t* __restrict__ lw, double cutoff,
                            const int64_t* __restrict__ chimax,
                            const int* __restrict__ attempt,
                            int64_t* __restrict__ kept,
                            int* __restrict__ accepted,
                            double* __restrict__ keptw,
                            double* __restrict__ discw) {
  const int bz = blockIdx.x * blockDim.x + threadIdx.x;
  if (bz >= ng) return;
  accepted[bz] = 0;
  if (!attempt[bz]) return;
  const double er2 = cutoff * cutoff * total[bz];
  const double* s = sig + (int64_t)bz * cols_p;
  const int64_t l = lw[bz];
  double tail = resid[bz];
  int64_t k = 0;
  for (int64_t j = l - 1; j >= 0; --j) {
    const double t = tail + s[j] * s[j];
    if (t > er2) {
      k = j + 1;
      break;
    }
    tail = t;
  }
  if (k > 0 && k < l && k <= chimax[bz]) {
    double kw = 0.0;
    for (int64_t j = k - 1; j >= 0; --j) kw += s[j] * s[j];
    accepted[bz] = 1;
    kept[bz] = k;
    keptw[bz] = kw;
    discw[bz] = tail;
  }
}

// V_k (N x k, ld N) = X_B[:, perm[j]] / sig[j] for accepted gates, X_B =
// B^H = theta^H Q (N x l, ld ldb).
__global__ void extract_vk_rand(const GateDesc* __restrict__ gd,
                                const int* __restrict__ accepted,
                                const double2* __restrict__ XB, int64_t ldb,
                                int64_t sB, const double* __restrict__ sig,
                                const int* __restrict__ perm, int64_t cols_p,
                                const int64_t* __restrict__ kept) {
  const int bz = blockIdx.y;
  if (!accepted[bz]) return;
  const GateDesc d = gd[bz];
  const int64_t k = kept[bz];
  const int64_t N = 2 * d.chiR;
  const double* s = sig + (int64_t)bz * cols_p;
  const int* pm = perm + (int64_t)bz * cols_p;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < N * k; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx % N, j = idx / N;
    const double sj = s[j];
    const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
    const double2 x = XB[bz * sB + c + ldb * (int64_t)pm[j]];
    d.Vk[c + N * j] = make_double2(x.x * inv, x.y * inv);
  }
}

// Features of the learned rank predictor, one column of kGateFeatures
// floats per gate: the bond spectrum before the gate (top 64 values, zero
// padded; the MLP takes their logarithm), the gate's four operator-Schmidt
// coefficients, the bond position i / n and the bond dimensions chi_{i-1},
// chi_i, chi_{i+1} / chimax.
constexpr int kGateFeatures = 72;
struct FeatDesc {
  const double* lam;
  int64_t len;
  float sch[4];
  float pos;
  float dims[3];
};

__global__ __launch_bounds__(128) void gate_features(
    const FeatDesc* __restrict__ fd, float* __restrict__ X) {
  const FeatDesc d = fd[blockIdx.x];
  const int t = threadIdx.x;
  float v = 0.f;
  if (t < 64)
    v = t < d.len ? (float)d.lam[t] : 0.f;
  else if (t < 68)
    v = d.sch[t - 64];
  else if (t == 68)
    v = d.pos;
  else if (t < kGateFeatures)
    v = d.dims[t - 69];
  if (t < kGateFeatures) X[(int64_t)blockIdx.x * kGateFeatures + t] = v;
}

// Operator-Schmidt coefficients of 4x4 gates (index s_w + 2 s_{w+1}), one
// thread per gate: the singular values of M[a + 2c, b + 2d] = g[a + 2b,
// c + 2d] (the split of the gate into (s_w, s_w') x (s_{w+1}, s_{w+1}')
// factors) by one-sided Jacobi (Hestenes) on M's four columns in FP64
// registers, sorted descending.
__global__ void gate_schmidt(int ng, const double2* __restrict__ gates,
                             double* __restrict__ out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const double2* G = gates + (int64_t)g * 16;
  double2 a[4][4];  // a[col][row]
  for (int b = 0; b < 2; ++b)
    for (int d = 0; d < 2; ++d)
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
          a[b + 2 * d][r + 2 * c] = G[(r + 2 * b) + 4 * (c + 2 * d)];
  for (int sweep = 0; sweep < 40; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 4; ++q) {
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int i = 0; i < 4; ++i) {
          al += a[p][i].x * a[p][i].x + a[p][i].y * a[p][i].y;
          be += a[q][i].x * a[q][i].x + a[q][i].y * a[q][i].y;
          // conj(a_p) a_q
          gr += a[p][i].x * a[q][i].x + a[p][i].y * a[q][i].y;
          gi += a[p][i].x * a[q][i].y - a[p][i].y * a[q][i].x;
        }
        const double gm = sqrt(gr * gr + gi * gi);
        if (gm <= 1e-15 * sqrt(al * be) || gm == 0.0) continue;
        rotated = true;
        // a_q <- e^{-i phi} a_q makes the inner product real (= gm)
        const double er = gr / gm, ei = -gi / gm;
        const double zeta = (be - al) / (2.0 * gm);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                         (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < 4; ++i) {
          const double2 x = a[p][i];
          const double2 y = make_double2(a[q][i].x * er - a[q][i].y * ei,
                                         a[q][i].x * ei + a[q][i].y * er);
          a[p][i] = make_double2(cs * x.x - sn * y.x, cs * x.y - sn * y.y);
          a[q][i] = make_double2(sn * x.x + cs * y.x, sn * x.y + cs * y.y);
        }
      }
    if (!rotated) break;
  }
  double sv[4];
  for (int c = 0; c < 4; ++c) {
    double s2 = 0.0;
    for (int i = 0; i < 4; ++i)
      s2 += a[c][i].x * a[c][i].x + a[c][i].y * a[c][i].y;
    sv[c] = sqrt(s2);
  }
  for (int i = 1; i < 4; ++i)
    for (int j = i; j > 0 && sv[j] > sv[j - 1]; --j) {
      const double t = sv[j];
      sv[j] = sv[j - 1];
      sv[j - 1] = t;
    }
  for (int c = 0; c < 4; ++c) out[(int64_t)g * 4 + c] = sv[c];
}

__global__ void fill_one(double2* p) { *p = make_double2(1.0, 0.0); }
__global__ void fill_lambda_one(double* p) { *p = 1.0; }

}  // namespace

// ---------------------------------------------------------------------------
// Batched two-site update of disjoint gates (host orchestration).
// Randomized range finder (Halko, Martinsson & Tropp, SIAM Rev. 53, 217
// (2011), algorithm 4.4: q = 2 power iterations with re-orthonormalisation)
// for the gates with attempt[g]: Q (M x l) spans theta Omega, B^H = theta^H
// Q is split by the Jacobi engine, the residual |theta - Q B|^2 is summed
// explicitly, and rand_decide applies the exact discarded-weight check.  Accepted gates get V_k written into their
// descriptor's Vk; kept / keptw / disc / total are filled for them.
constexpr int kSlotROm = 140, kSlotRY = 141, kSlotRZ = 142, kSlotRQ = 143,
              kSlotRDesc = 144, kSlotRSmall = 145, kSlotRDesc2 = 146,
              kSlotRP = 147, kSlotRRes = 148;

int32_t rand_range_finder(qtn_ctx* ctx, int ng,
                          const std::vector<GateDesc>& gd,
                          const GateDesc* d_gd, int64_t ldx,
                          const std::vector<int>& attempt,
                          const std::vector<int64_t>& lw, const int64_t* chis,
                          double cutoff, uint64_t seed,
                          std::vector<int>& accepted,
                          std::vector<int64_t>& kept,
                          std::vector<double>& keptw,
                          std::vector<double>& disc,
                          std::vector<double>& total) {
  hipStream_t s = ctx->stream;
  int64_t maxM = 1, maxN = 1, maxl = 1;
  for (int g = 0; g < ng; ++g) {
    if (!attempt[g]) continue;
    maxM = std::max(maxM, 2 * gd[g].chiL);
    maxN = std::max(maxN, 2 * gd[g].chiR);
    maxl = std::max(maxl, lw[g]);
  }
  const int64_t ldq = std::max(maxM, maxN);
  const int64_t rowsB = round_up64(maxN), colsB = round_up64(maxl);
  const int64_t om_sz = maxN * maxl, y_sz = maxM * maxl,
                z_sz = rowsB * colsB, q_sz = ldq * maxl;
  void *pOm, *pY, *pZ, *pQ, *pD, *pS;
  QTN_TRY(ctx->buffer(kSlotROm, (size_t)ng * om_sz * sizeof(double2), &pOm));
  QTN_TRY(ctx->buffer(kSlotRY, (size_t)ng * y_sz * sizeof(double2), &pY));
  QTN_TRY(ctx->buffer(kSlotRZ, (size_t)ng * z_sz * sizeof(double2), &pZ));
  QTN_TRY(ctx->buffer(kSlotRQ, (size_t)ng * q_sz * sizeof(double2), &pQ));
  QTN_TRY(ctx->buffer(kSlotRDesc, (size_t)ng * sizeof(RandDesc), &pD));
  QTN_TRY(ctx->buffer(kSlotRSmall,
                      (size_t)ng * (4 * sizeof(int64_t) + 4 * sizeof(double) +
                                    2 * sizeof(int)),
                      &pS));
  std::vector<RandDesc> rd(ng);
  std::vector<QrDims> qdM(ng), qdN(ng);
  for (int g = 0; g < ng; ++g) {
    RandDesc& r = rd[g];
    r.X = gd[g].X;
    r.ldx = ldx;
    r.trans = gd[g].trans;
    r.M = attempt[g] ? 2 * gd[g].chiL : 0;
    r.N = attempt[g] ? 2 * gd[g].chiR : 0;
    r.l = attempt[g] ? lw[g] : 0;
    r.Om = (double2*)pOm + (size_t)g * om_sz;
    r.Y = (double2*)pY + (size_t)g * y_sz;
    r.Z = (double2*)pZ + (size_t)g * z_sz;
    r.Q = (double2*)pQ + (size_t)g * q_sz;
    qdM[g] = QrDims{r.M, r.l, attempt[g] ? 1 : 0};
    qdN[g] = QrDims{r.N, r.l, attempt[g] ? 1 : 0};
  }
  RandDesc* d_rd = (RandDesc*)pD;
  QTN_CHECK_HIP(hipMemcpyAsync(d_rd, rd.data(), ng * sizeof(RandDesc),
                               hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(gauss_fill, dim3(grid_for(maxN * maxl), (unsigned)ng),
                     dim3(256), 0, s, (const RandDesc*)d_rd, seed, maxN);
  QTN_CHECK_HIP(hipGetLastError());
  // descriptors of the four products, split by the orientation of X:
  //   Y = theta Om | theta Q   (M x l): trans X^H B, else X B
  //   Z = theta^H Q            (N x l): trans X Q,   else X^H Q
  void* pG;
  QTN_TRY(ctx->buffer(kSlotRDesc2, (size_t)11 * ng * sizeof(GemmDesc<double>),
                      &pG));
  GemmDesc<double>* dG = (GemmDesc<double>*)pG;
  // host descriptor arrays live until the stream is synchronised below
  std::vector<std::vector<GemmDesc<double>>> keep;
  keep.reserve(16);
  auto run_products = [&](bool toY, const double2* Bbase, int64_t bstride,
                          int64_t ldb, double2* Cbase, int64_t cstride,
                          int64_t ldc, int slot_off) -> int32_t {
    keep.emplace_back();
    std::vector<GemmDesc<double>>& gt = keep.back();
    keep.emplace_back();
    std::vector<GemmDesc<double>>& gn = keep.back();
    for (int g = 0; g < ng; ++g) {
      if (!attempt[g]) continue;
      const RandDesc& r = rd[g];
      GemmDesc<double> e;
      e.M = toY ? r.M : r.N;
      e.N = r.l;
      e.K = toY ? r.N : r.M;
      e.alpha = make_double2(1.0, 0.0);
      e.beta = make_double2(0.0, 0.0);
      e.A = r.X;
      e.lda = ldx;
      e.B = Bbase + (size_t)g * bstride;
      e.ldb = ldb;
      e.C = Cbase + (size_t)g * cstride;
      e.ldc = ldc;
      // toY: trans -> X^H (op C); Z: non-trans -> X^H (op C)
      const bool conj = toY ? r.trans : !r.trans;
      (conj ? gt : gn).push_back(e);
    }
    GemmDesc<double>* d1 = dG + (size_t)slot_off * 2 * ng;
    GemmDesc<double>* d2 = d1 + ng;
    const int64_t mm = toY ? maxM : maxN;
    if (!gt.empty()) {
      QTN_CHECK_HIP(hipMemcpyAsync(d1, gt.data(),
                                   gt.size() * sizeof(GemmDesc<double>),
                                   hipMemcpyHostToDevice, s));
      QTN_TRY(launch_zgemm_grouped(s, 2, 0, d1, (int)gt.size(), mm, maxl));
    }
    if (!gn.empty()) {
      QTN_CHECK_HIP(hipMemcpyAsync(d2, gn.data(),
                                   gn.size() * sizeof(GemmDesc<double>),
                                   hipMemcpyHostToDevice, s));
      QTN_TRY(launch_zgemm_grouped(s, 0, 0, d2, (int)gn.size(), mm, maxl));
    }
    return QTN_OK;
  };
  // orthonormal basis of the first l columns of A (rows x l, ld lda) into
  // Q (ld ldq): Householder QR, then Q [I; 0]
  auto orth = [&](double2* A, int64_t lda, int64_t stride, int which,
                  const std::vector<QrDims>& dims) -> int32_t {
    BatchedQR qr;
    qr.slot_base = 150;
    QTN_TRY(qr.factor(ctx, A, lda, stride, maxl, dims, true));
    hipLaunchKernelGGL(eye_fill, dim3(grid_for(ldq * maxl), (unsigned)ng),
                       dim3(256), 0, s, (const RandDesc*)d_rd, which, ldq);
    QTN_CHECK_HIP(hipGetLastError());
    std::vector<int64_t> ldu(ng, ldq), uc(ng);
    for (int g = 0; g < ng; ++g) uc[g] = attempt[g] ? lw[g] : 0;
    return qr.apply_q(ctx, (double2*)pQ, ldu, q_sz, uc, dims);
  };
  // Y = theta Om
  QTN_TRY(run_products(true, (const double2*)pOm, om_sz, maxN, (double2*)pY,
                       y_sz, maxM, 0));
  for (int it = 0; it < 2; ++it) {
    QTN_TRY(orth((double2*)pY, maxM, y_sz, 0, qdM));           // Q = orth(Y)
    QTN_TRY(run_products(false, (const double2*)pQ, q_sz, ldq, (double2*)pZ,
                         z_sz, rowsB, 1 + 2 * it));             // Z = th^H Q
    QTN_TRY(orth((double2*)pZ, rowsB, z_sz, 1, qdN));          // Q = orth(Z)
    QTN_TRY(run_products(true, (const double2*)pQ, q_sz, ldq, (double2*)pY,
                         y_sz, maxM, 2 + 2 * it));              // Y = th Q
  }
  QTN_TRY(orth((double2*)pY, maxM, y_sz, 0, qdM));
  // B^H = theta^H Q (N x l) into a zeroed rowsB x colsB block per gate
  QTN_CHECK_HIP(hipMemsetAsync(pZ, 0, (size_t)ng * z_sz * sizeof(double2), s));
  QTN_TRY(run_products(false, (const double2*)pQ, q_sz, ldq, (double2*)pZ,
                       z_sz, rowsB, 3));
  // P = Q Q^H theta in X's orientation (non-trans: Q Z^H, trans: Z Q^H),
  // before the sweeps rotate Z; the residual |X - P|^2 feeds the check
  int64_t maxRX = 1, maxCX = 1;
  for (int g = 0; g < ng; ++g) {
    if (!attempt[g]) continue;
    maxRX = std::max(maxRX, rd[g].trans ? rd[g].N : rd[g].M);
    maxCX = std::max(maxCX, rd[g].trans ? rd[g].M : rd[g].N);
  }
  const int64_t p_sz = maxRX * maxCX;
  void* pP;
  QTN_TRY(ctx->buffer(kSlotRP, (size_t)ng * p_sz * sizeof(double2), &pP));
  {
    keep.emplace_back();
    std::vector<GemmDesc<double>>& gp = keep.back();
    for (int g = 0; g < ng; ++g) {
      if (!attempt[g]) continue;
      const RandDesc& r = rd[g];
      const double2* Qg = (const double2*)pQ + (size_t)g * q_sz;
      const double2* Zg = (const double2*)pZ + (size_t)g * z_sz;
      GemmDesc<double> e;
      e.M = r.trans ? r.N : r.M;
      e.N = r.trans ? r.M : r.N;
      e.K = r.l;
      e.alpha = make_double2(1.0, 0.0);
      e.beta = make_double2(0.0, 0.0);
      e.A = r.trans ? Zg : Qg;
      e.lda = r.trans ? rowsB : ldq;
      e.B = r.trans ? Qg : Zg;
      e.ldb = r.trans ? ldq : rowsB;
      e.C = (double2*)pP + (size_t)g * p_sz;
      e.ldc = maxRX;
      gp.push_back(e);
    }
    GemmDesc<double>* dp = dG + (size_t)10 * ng;
    QTN_CHECK_HIP(hipMemcpyAsync(dp, gp.data(),
                                 gp.size() * sizeof(GemmDesc<double>),
                                 hipMemcpyHostToDevice, s));
    QTN_TRY(launch_zgemm_grouped(s, 0, 2, dp, (int)gp.size(), maxRX, maxCX));
  }
  void* pR;
  QTN_TRY(ctx->buffer(kSlotRRes, (size_t)ng * sizeof(double), &pR));
  hipLaunchKernelGGL(resid_norm2, dim3(ng), dim3(256), 0, s,
                     (const RandDesc*)d_rd, (const double2*)pP, maxRX, p_sz,
                     (double*)pR);
  QTN_CHECK_HIP(hipGetLastError());
  JacobiEngine eng;
  eng.slot_base = 160;
  int nat = 0;
  for (int g = 0; g < ng; ++g) nat += attempt[g] ? 1 : 0;
  QTN_TRY(eng.reserve(ctx, ng, rowsB, colsB, false, nat));
  std::vector<int64_t> nsv(ng);
  for (int g = 0; g < ng; ++g) nsv[g] = attempt[g] ? lw[g] : 1;
  int sweeps = 0;
  QTN_TRY(eng.run(ctx, (double2*)pZ, rowsB, z_sz, nullptr, 0, nullptr,
                  nsv.data(), 0.0, 1, 0, 60, &sweeps, attempt.data()));
  int64_t* d_lw = (int64_t*)pS;
  int64_t* d_ch = d_lw + ng;
  int64_t* d_kept = d_ch + ng;
  double* d_tot = (double*)(d_kept + ng);
  double* d_kw = d_tot + ng;
  double* d_dw = d_kw + ng;
  int* d_att = (int*)(d_dw + ng);
  int* d_acc = d_att + ng;
  QTN_CHECK_HIP(hipMemcpyAsync(d_lw, lw.data(), ng * sizeof(int64_t),
                               hipMemcpyHostToDevice, s));
  QTN_CHECK_HIP(hipMemcpyAsync(d_ch, chis, ng * sizeof(int64_t),
                               hipMemcpyHostToDevice, s));
  QTN_CHECK_HIP(hipMemcpyAsync(d_att, attempt.data(), ng * sizeof(int),
                               hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(theta_norm2, dim3(ng), dim3(256), 0, s,
                     (const RandDesc*)d_rd, d_tot);
  hipLaunchKernelGGL(rand_decide, dim3((ng + 63) / 64), dim3(64), 0, s, ng,
                     (const double*)eng.sig, eng.cols_p,
                     (const double*)d_tot, (const double*)pR,
                     (const int64_t*)d_lw, cutoff,
                     (const int64_t*)d_ch, (const int*)d_att, d_kept, d_acc,
                     d_kw, d_dw);
  hipLaunchKernelGGL(extract_vk_rand,
                     dim3(grid_for(maxN * maxl), (unsigned)ng), dim3(256), 0,
                     s, d_gd, (const int*)d_acc, (const double2*)pZ, rowsB,
                     z_sz, (const double*)eng.sig, (const int*)eng.perm,
                     eng.cols_p, (const int64_t*)d_kept);
  QTN_CHECK_HIP(hipGetLastError());
  accepted.assign(ng, 0);
  std::vector<int64_t> k(ng, 0);
  std::vector<double> kw(ng, 0.0), dw(ng, 0.0), tt(ng, 0.0);
  QTN_CHECK_HIP(hipMemcpyAsync(accepted.data(), d_acc, ng * sizeof(int),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(k.data(), d_kept, ng * sizeof(int64_t),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(kw.data(), d_kw, ng * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(dw.data(), d_dw, ng * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipMemcpyAsync(tt.data(), d_tot, ng * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  for (int g = 0; g < ng; ++g) {
    if (!accepted[g]) continue;
    kept[g] = k[g];
    keptw[g] = kw[g];
    disc[g] = dw[g];
    total[g] = tt[g];
  }
  return QTN_OK;
}

// Gate g acts on sites (sites[g], sites[g] + 1) of state ms[g] with the
// bond cap chis[g]; the states may differ (one batch over many circuits)
// but share ctx.
// hint (learned form, may be NULL): predicted kept rank per gate; gates
// with 0 < hint < cap try the randomized range finder first (sketch width
// hint + 16) and fall back to the full SVD when its exact check fails;
// rand_accepted (may be NULL) receives 1 for the gates it decided, -1 for
// those whose sketch failed the check, 0 where no sketch was tried.
int32_t apply_gates_core(qtn_ctx* ctx, int ng, qtn_mps* const* ms,
                         const int32_t* sites, const double* gates,
                         double cutoff, const int64_t* chis, int form,
                         int64_t* kept_out, double* disc_out,
                         double* spec_out, int64_t* spec_len,
                         const int64_t* hint = nullptr,
                         int* rand_accepted = nullptr) {
  const bool chosen = form == QTN_FORM_CHOSEN;
  const bool learned = form == QTN_FORM_LEARNED;
  hipStream_t s = ctx->stream;
  std::map<const qtn_mps*, std::vector<char>> used;
  int64_t maxM = 0, maxN = 0, maxXr = 0, maxXc = 0;
  bool any_w = false;
  std::vector<int> trans(ng);
  for (int g = 0; g < ng; ++g) {
    const qtn_mps* m = ms[g];
    const int i = sites[g];
    QTN_ARG(m != nullptr, "apply_layer: NULL state");
    QTN_ARG(m->ctx == ctx || ctx->parent == m->ctx,
            "apply_layer: states on different contexts");
    QTN_ARG(i >= 1 && i < m->n, "apply_2q: site out of range");
    QTN_ARG(chis[g] >= 1 && chis[g] <= m->cap,
            "apply_layer: chimax above capacity");
    std::vector<char>& u = used[m];
    if (u.empty()) u.assign(m->n + 1, 0);
    QTN_ARG(!u[i] && !u[i + 1], "apply_layer: gates overlap");
    u[i] = u[i + 1] = 1;
    const int64_t M = 2 * m->dims[i - 1], N = 2 * m->dims[i + 1];
    maxM = std::max(maxM, M);
    maxN = std::max(maxN, N);
    // orthogonalise min(M, N) columns: theta^H when M <= N, else theta
    // with the right vectors accumulated in W
    trans[g] = (M <= N) ? 1 : 0;
    maxXr = std::max(maxXr, trans[g] ? N : M);
    maxXc = std::max(maxXc, trans[g] ? M : N);
    if (!trans[g]) any_w = true;
  }
  const int64_t rows_p = round_up64(maxXr), cols_p = round_up64(maxXc);
  const int64_t phi_sz = maxM * maxN, x_sz = rows_p * cols_p;
  // QR preconditioning halves the Jacobi sweeps but its panel chain costs
  // about what it saves at chi = 256; off unless QTN_JACOBI_PRECOND=1.
  bool precond = false;
  if (const char* e = getenv("QTN_JACOBI_PRECOND")) precond = atoi(e) != 0;
  if (chosen || learned) precond = false;
  const int64_t w_sz = cols_p * cols_p;
  void *dphi, *dX, *dVk, *ddesc, *dg, *dW = nullptr;
  if (any_w)
    QTN_TRY(ctx->buffer(kSlotW2, (size_t)ng * w_sz * sizeof(double2), &dW));
  QTN_TRY(ctx->buffer(kSlotPhi, (size_t)ng * phi_sz * sizeof(double2), &dphi));
  QTN_TRY(ctx->buffer(kSlotXm, (size_t)ng * x_sz * sizeof(double2), &dX));
  QTN_TRY(ctx->buffer(kSlotVk, (size_t)ng * maxN * std::max(maxM, maxN) *
                                   sizeof(double2),
                      &dVk));
  const size_t desc_bytes = (size_t)ng * (sizeof(GateDesc) +
                                          4 * sizeof(GemmDesc<double>));
  QTN_TRY(ctx->buffer(kSlotDesc, desc_bytes, &ddesc));
  QTN_TRY(ctx->buffer(kSlotGates, (size_t)ng * 16 * sizeof(double2), &dg));
  std::vector<GateDesc> gd(ng);
  std::vector<GemmDesc<double>> g1(ng);
  std::vector<int64_t> nsv(ng);
  for (int g = 0; g < ng; ++g) {
    const int i = sites[g];
    const qtn_mps* m = ms[g];
    GateDesc& d = gd[g];
    d.chiL = m->dims[i - 1];
    d.chiM = m->dims[i];
    d.chiR = m->dims[i + 1];
    d.Bl = m->site[i - 1];
    d.Br = m->site[i];
    d.lamL = m->lam[i - 1];
    d.lamM = m->lam[i];
    d.phi = (double2*)dphi + (size_t)g * phi_sz;
    d.X = (double2*)dX + (size_t)g * x_sz;
    d.W = dW ? (double2*)dW + (size_t)g * w_sz : nullptr;
    d.trans = trans[g];
    d.ldx = rows_p;
    // QR-preconditioned Jacobi for theta^H wider than one block pair
    d.mode = !trans[g] ? 2 : (precond && 2 * d.chiL > 64 ? 1 : 0);
    d.Vk = (double2*)dVk + (size_t)g * maxN * std::max(maxM, maxN);
    d.g = (const double2*)dg + 16 * g;
    const int64_t M = 2 * d.chiL, N = 2 * d.chiR;
    nsv[g] = std::min(M, N);
    GemmDesc<double>& e = g1[g];
    e.M = M;
    e.N = N;
    e.K = d.chiM;
    e.alpha = make_double2(1.0, 0.0);
    e.beta = make_double2(0.0, 0.0);
    e.A = d.Bl;
    e.lda = M;
    e.B = d.Br;
    e.ldb = d.chiM;
    e.C = d.phi;
    e.ldc = M;
  }
  GateDesc* d_gd = (GateDesc*)ddesc;
  GemmDesc<double>* d_g1 = (GemmDesc<double>*)(d_gd + ng);
  GemmDesc<double>* d_g2 = d_g1 + ng;
  QTN_CHECK_HIP(hipMemcpyAsync(dg, gates, (size_t)ng * 16 * sizeof(double2),
                               hipMemcpyHostToDevice, s));
  QTN_CHECK_HIP(hipMemcpyAsync(d_gd, gd.data(), ng * sizeof(GateDesc),
                               hipMemcpyHostToDevice, s));
  QTN_CHECK_HIP(hipMemcpyAsync(d_g1, g1.data(), ng * sizeof(GemmDesc<double>),
                               hipMemcpyHostToDevice, s));
  // 1. theta0 = B_i B_{i+1}
  QTN_TRY(launch_zgemm_grouped(s, 0, 0, d_g1, ng, maxM, maxN));
  // 2. gate, lambda scaling, X = theta^H (zero padded)
  QTN_CHECK_HIP(hipMemsetAsync(dX, 0, (size_t)ng * x_sz * sizeof(double2), s));
  hipLaunchKernelGGL(gate_epilogue,
                     dim3(grid_for(maxM * maxN / 4), (unsigned)ng), dim3(256),
                     0, s, (const GateDesc*)d_gd, rows_p);
  QTN_CHECK_HIP(hipGetLastError());
  // 3. SVD of every theta.  Mode-1 gates are preconditioned:
  //    X P = Q1 R1 (columns sorted by norm), R1^H = Q2 R2, and the Jacobi
  //    iteration runs on Y = R2^H (M x M, written over X).
  std::vector<QrDims> qd(ng), qd2(ng);
  bool any_pre = false, all_pre = true;
  for (int g = 0; g < ng; ++g) {
    const bool on = gd[g].mode == 1;
    const int64_t M = 2 * gd[g].chiL, N = 2 * gd[g].chiR;
    qd[g] = QrDims{N, M, on ? 1 : 0};
    qd2[g] = QrDims{M, M, on ? 1 : 0};
    any_pre = any_pre || on;
    all_pre = all_pre && on;
  }
  BatchedQR qr1, qr2;
  qr1.slot_base = 50;
  qr2.slot_base = 60;
  if (any_pre) {
    void *dA1, *dA2, *pq;
    QTN_TRY(ctx->buffer(kSlotA1, (size_t)ng * x_sz * sizeof(double2), &dA1));
    QTN_TRY(ctx->buffer(kSlotA2, (size_t)ng * cols_p * cols_p *
                                     sizeof(double2),
                        &dA2));
    QTN_TRY(ctx->buffer(kSlotQd, (size_t)ng * 2 * sizeof(QrDims), &pq));
    QrDims* d_qd = (QrDims*)pq;
    QTN_CHECK_HIP(hipMemcpyAsync(d_qd, qd.data(), ng * sizeof(QrDims),
                                 hipMemcpyHostToDevice, s));
    QTN_CHECK_HIP(hipMemcpyAsync(d_qd + ng, qd2.data(), ng * sizeof(QrDims),
                                 hipMemcpyHostToDevice, s));
    QTN_TRY(qr_sort_columns(ctx, (const double2*)dX, (double2*)dA1, rows_p,
                            cols_p, x_sz, ng, d_qd, kSlotNrm, kSlotPerm2,
                            nullptr));
    QTN_TRY(qr1.factor(ctx, (double2*)dA1, rows_p, x_sz, cols_p, qd, true));
    QTN_TRY(qr_r_to_adjoint(ctx, (const double2*)dA1, rows_p, x_sz,
                            (double2*)dA2, cols_p, cols_p * cols_p, cols_p,
                            cols_p, ng, d_qd));
    QTN_TRY(qr2.factor(ctx, (double2*)dA2, cols_p, cols_p * cols_p, cols_p,
                       qd2, false));
    QTN_TRY(qr_r_to_adjoint(ctx, (const double2*)dA2, cols_p,
                            cols_p * cols_p, (double2*)dX, rows_p, x_sz,
                            rows_p, cols_p, ng, d_qd + ng));
    QTN_CHECK_HIP(hipStreamSynchronize(s));  // qd vectors live on the host
  }
  // The plain form solves every gate with the FP64 engine.  The chosen
  // form first tries FP32 sweeps on the gates whose theta exceeds the cap
  // (min(M, N) > chimax) and keeps that result when the cap binds inside
  // the spectrum FP32 resolves (kept == chimax with every kept value above
  // the FP32 noise floor); every other gate goes to the FP64 engine.
  // (FP32 on every gate measured no gain: on Haar brick-walls the uncapped
  // two-site spectra of the middle layers carry values between the cutoff
  // and the FP32 floor, about 2e-5 |X|_F, which FP32 cannot decide.)
  JacobiEngine eng, eng32;
  eng32.slot_base = kSlotEng32;
  std::vector<int> wantw(ng);
  for (int g = 0; g < ng; ++g) wantw[g] = trans[g] ? 0 : 1;
  int sweeps = 0;
  std::vector<int64_t> kept(ng, 0);
  std::vector<int> st(ng, 0), src(ng, 0), need64(ng, 1);
  std::vector<double> disc(ng, 0.0), total(ng, 0.0), keptw(ng, 0.0);
  auto read_results = [&](JacobiEngine& e, std::vector<int64_t>& k,
                          std::vector<int>& t, std::vector<double>& d,
                          std::vector<double>& tt,
                          std::vector<double>& kw) -> int32_t {
    QTN_CHECK_HIP(hipMemcpyAsync(k.data(), e.kept, ng * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, s));
    QTN_CHECK_HIP(hipMemcpyAsync(t.data(), e.status, ng * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    QTN_CHECK_HIP(hipMemcpyAsync(d.data(), e.disc, ng * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    QTN_CHECK_HIP(hipMemcpyAsync(tt.data(), e.total, ng * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    QTN_CHECK_HIP(hipMemcpyAsync(kw.data(), e.keptw, ng * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    QTN_CHECK_HIP(hipStreamSynchronize(s));
    return QTN_OK;
  };
  // The learned form is the chosen form plus the range finder: gates whose
  // sketch (width hint + 16) stays below the cap and well below the full
  // rank try it; the other capped gates (and rejected sketches where the
  // cap binds) take the FP32 route, and the rest the FP64 engine.  A
  // sketch as wide as the cap would only repeat what the cap decides.
  std::vector<int> attempt(ng, 0);
  std::vector<int64_t> lw(ng, 0);
  bool any_attempt = false;
  if (learned && hint) {
    for (int g = 0; g < ng; ++g) {
      const int64_t l = std::min<int64_t>(hint[g] + 16, nsv[g]);
      if (hint[g] > 0 && l < chis[g] && 4 * l <= 3 * nsv[g]) {
        attempt[g] = 1;
        lw[g] = l;
        any_attempt = true;
      }
    }
  }
  // The sketches run first, so that a gate whose sketch is rejected can
  // still take the FP32 route below.
  std::vector<int> acc(ng, 0);
  if (any_attempt) {
    // each call draws a fresh sketch from the context's generator
    ctx->sketch_seed =
        ctx->sketch_seed * 6364136223846793005ULL + 1442695040888963407ULL;
    const uint64_t seed = ctx->sketch_seed;
    QTN_TRY(rand_range_finder(ctx, ng, gd, d_gd, rows_p, attempt, lw, chis,
                              cutoff, seed, acc, kept, keptw, disc, total));
    for (int g = 0; g < ng; ++g)
      if (acc[g]) {
        src[g] = 2;
        need64[g] = 0;
      }
  }
  void *dX32 = nullptr, *dW32 = nullptr;
  if (chosen || learned) {
    // The FP32 route pays only where the cap binds.  Without a prediction
    // that is taken to be where the bond sat at the cap before the gate or
    // theta is at least twice the cap (a cutoff-bound spectrum below the
    // cap would send the FP32 result back to the FP64 engine).  In the
    // learned form a predicted rank at the cap decides it, a prediction
    // within the sketch margin (16) below the cap is too close to call and
    // falls back to the structural test, as does a rejected sketch.
    std::vector<int> cand(ng, 0);
    bool any32 = false;
    for (int g = 0; g < ng; ++g) {
      const bool structural =
          ms[g]->dims[sites[g]] >= chis[g] || nsv[g] >= 2 * chis[g];
      bool capped;
      if (attempt[g])
        capped = !acc[g] && structural;
      else if (learned && hint && hint[g] > 0)
        capped = hint[g] >= chis[g] || (hint[g] + 16 >= chis[g] && structural);
      else
        capped = structural;
      cand[g] = nsv[g] > chis[g] && capped ? 1 : 0;
      any32 = any32 || cand[g];
    }
    if (any32) {
      // the sweeps run on a complex FP32 copy of X (and W)
      QTN_TRY(ctx->buffer(kSlotX32, (size_t)ng * x_sz * sizeof(float2),
                          &dX32));
      if (any_w)
        QTN_TRY(ctx->buffer(kSlotW32, (size_t)ng * w_sz * sizeof(float2),
                            &dW32));
      hipLaunchKernelGGL(to_single, dim3(grid_for((int64_t)ng * x_sz)),
                         dim3(256), 0, s, (const double2*)dX, (float2*)dX32,
                         (int64_t)ng * x_sz);
      QTN_CHECK_HIP(hipGetLastError());
      int n32 = 0;
      for (int g = 0; g < ng; ++g) n32 += cand[g];
      QTN_TRY(eng32.reserve(ctx, ng, rows_p, cols_p, true, n32));
      eng32.chimax_per = chis;
      QTN_TRY(eng32.run32(ctx, (float2*)dX32, rows_p, x_sz, (float2*)dW32,
                          w_sz, any_w ? wantw.data() : nullptr, nsv.data(),
                          cutoff, 1, 0, 60, &sweeps, cand.data()));
      std::vector<int64_t> k32(ng);
      std::vector<int> st32(ng);
      std::vector<double> d32(ng), t32(ng), kw32(ng);
      QTN_TRY(read_results(eng32, k32, st32, d32, t32, kw32));
      for (int g = 0; g < ng; ++g) {
        if (cand[g] && st32[g] == 0 && k32[g] == chis[g]) {
          src[g] = 1;
          need64[g] = 0;
          kept[g] = k32[g];
          disc[g] = d32[g];
          total[g] = t32[g];
          keptw[g] = kw32[g];
        }
      }
    }
  }
  if (rand_accepted)
    for (int g = 0; g < ng; ++g)
      rand_accepted[g] = src[g] == 2 ? 1 : (attempt[g] ? -1 : 0);
  bool any64 = false;
  for (int g = 0; g < ng; ++g) any64 = any64 || need64[g];
  if (any64) {
    // every problem preconditioned: only the top cols_p rows are non-zero
    int64_t rows_eng = all_pre ? cols_p : rows_p, cols_eng = cols_p;
    if (chosen || learned) {
      // the FP64 engine is padded to the gates it solves, not to the batch
      int64_t r64 = 1, c64 = 1;
      for (int g = 0; g < ng; ++g) {
        if (!need64[g]) continue;
        const int64_t M = 2 * gd[g].chiL, N = 2 * gd[g].chiR;
        r64 = std::max(r64, trans[g] ? N : M);
        c64 = std::max(c64, trans[g] ? M : N);
      }
      rows_eng = round_up64(r64);
      cols_eng = round_up64(c64);
    }
    eng.ldw = cols_p;  // W keeps the batch layout
    int n64 = 0;
    for (int g = 0; g < ng; ++g) n64 += need64[g];
    QTN_TRY(eng.reserve(ctx, ng, rows_eng, cols_eng, false, n64));
    eng.chimax_per = chis;
    QTN_TRY(eng.run(ctx, (double2*)dX, rows_p, x_sz, (double2*)dW, w_sz,
                    any_w ? wantw.data() : nullptr, nsv.data(), cutoff, 1,
                    0, 60, &sweeps,
                    (chosen || learned) ? need64.data() : nullptr));
    std::vector<int64_t> k64(ng);
    std::vector<int> st64(ng);
    std::vector<double> d64(ng), t64(ng), kw64(ng);
    QTN_TRY(read_results(eng, k64, st64, d64, t64, kw64));
    for (int g = 0; g < ng; ++g) {
      if (!need64[g]) continue;
      kept[g] = k64[g];
      st[g] = st64[g];
      disc[g] = d64[g];
      total[g] = t64[g];
      keptw[g] = kw64[g];
    }
  }
  for (int g = 0; g < ng; ++g) {
    if (st[g] != 0) {
      set_error("apply_2q: every singular value is within the truncation "
                "error (zero two-site tensor)");
      return QTN_ERR_TRUNC;
    }
  }
  if (spec_out && ng == 1) {
    if (src[0] == 2) {
      // the range finder only resolves the sketch's values
      if (spec_len) *spec_len = 0;
    } else {
      QTN_CHECK_HIP(hipMemcpy(spec_out, src[0] ? eng32.sig : eng.sig,
                              nsv[0] * sizeof(double),
                              hipMemcpyDeviceToHost));
      if (spec_len) *spec_len = nsv[0];
    }
  }
  // per-gate kept rank and source engine on the device (chosen form)
  int64_t* d_kept = eng.kept;
  int* d_src = nullptr;
  if (chosen || learned) {
    void* pm;
    QTN_TRY(ctx->buffer(kSlotMix, (size_t)ng * (sizeof(int64_t) + sizeof(int)),
                        &pm));
    d_kept = (int64_t*)pm;
    d_src = (int*)(d_kept + ng);
    QTN_CHECK_HIP(hipMemcpyAsync(d_kept, kept.data(), ng * sizeof(int64_t),
                                 hipMemcpyHostToDevice, s));
    QTN_CHECK_HIP(hipMemcpyAsync(d_src, src.data(), ng * sizeof(int),
                                 hipMemcpyHostToDevice, s));
  }
  // 4. V_k, then B_{i+1} = V_k^H and lambda_i
  int64_t maxk = 0;
  for (int g = 0; g < ng; ++g) maxk = std::max(maxk, kept[g]);
  if (chosen || learned) {
    hipLaunchKernelGGL(extract_vk_mixed,
                       dim3(grid_for(maxN * maxk), (unsigned)ng), dim3(256),
                       0, s, (const GateDesc*)d_gd, (const int*)d_src,
                       (const float2*)dX32, x_sz, (const float2*)dW32, w_sz,
                       (const double*)eng32.sig, (const int*)eng32.perm,
                       (const double*)eng.sig, (const int*)eng.perm,
                       (const int64_t*)d_kept, cols_p, eng32.cols_p,
                       eng.cols_p);
    QTN_CHECK_HIP(hipGetLastError());
  }
  bool any_fp32 = false;
  for (int g = 0; g < ng; ++g) any_fp32 = any_fp32 || src[g] == 1;
  if (chosen || (learned && any_fp32)) {
    // Newton-Schulz in FP64: V <- V (3 I - V^H V) / 2, three steps take the
    // FP32 orthogonality error (<~1e-3 on the smallest kept columns) below
    // FP64 rounding.
    void *dG, *dT;
    const int64_t t_sz = maxN * maxk;
    QTN_TRY(ctx->buffer(kSlotNsG, (size_t)ng * maxk * maxk * sizeof(double2),
                        &dG));
    QTN_TRY(ctx->buffer(kSlotNsT, (size_t)ng * t_sz * sizeof(double2), &dT));
    std::vector<GemmDesc<double>> ga(ng), gb(ng);
    for (int g = 0; g < ng; ++g) {
      const int64_t N = 2 * gd[g].chiR, k = kept[g];
      double2* G = (double2*)dG + (size_t)g * maxk * maxk;
      ga[g] = GemmDesc<double>{k, k, N, make_double2(1.0, 0.0),
                               make_double2(0.0, 0.0), gd[g].Vk, N,
                               gd[g].Vk, N, G, k};
      gb[g] = GemmDesc<double>{N, k, k, make_double2(1.0, 0.0),
                               make_double2(0.0, 0.0), gd[g].Vk, N, G, k,
                               (double2*)dT + (size_t)g * t_sz, N};
    }
    GemmDesc<double>* d_ga = d_g2 + ng;
    GemmDesc<double>* d_gb = d_ga + ng;
    QTN_CHECK_HIP(hipMemcpyAsync(d_ga, ga.data(),
                                 ng * sizeof(GemmDesc<double>),
                                 hipMemcpyHostToDevice, s));
    QTN_CHECK_HIP(hipMemcpyAsync(d_gb, gb.data(),
                                 ng * sizeof(GemmDesc<double>),
                                 hipMemcpyHostToDevice, s));
    for (int it = 0; it < 3; ++it) {
      QTN_TRY(launch_zgemm_grouped(s, 2, 0, d_ga, ng, maxk, maxk));
      QTN_TRY(launch_zgemm_grouped(s, 0, 0, d_gb, ng, maxN, maxk));
      hipLaunchKernelGGL(ns_combine, dim3(grid_for(maxN * maxk), (unsigned)ng),
                         dim3(256), 0, s, (const GateDesc*)d_gd,
                         (const double2*)dT, t_sz, (const int64_t*)d_kept);
      QTN_CHECK_HIP(hipGetLastError());
    }
  } else if (!learned) {
    hipLaunchKernelGGL(extract_vk, dim3(grid_for(maxN * maxk), (unsigned)ng),
                       dim3(256), 0, s, (const GateDesc*)d_gd,
                       (const double*)eng.sig, (const int*)eng.perm,
                       (const int64_t*)eng.kept, cols_p);
    QTN_CHECK_HIP(hipGetLastError());
  }
  if (any_pre) {
    // V_k = Q1 [U_k; 0]
    std::vector<int64_t> ldu(ng), ucols(ng);
    for (int g = 0; g < ng; ++g) {
      ldu[g] = 2 * gd[g].chiR;
      ucols[g] = gd[g].mode == 1 ? kept[g] : 0;
    }
    const int64_t vk_stride = maxN * std::max(maxM, maxN);
    QTN_TRY(qr1.apply_q(ctx, (double2*)dVk, ldu, vk_stride, ucols, qd));
  }
  hipLaunchKernelGGL(write_right, dim3(grid_for(maxN * maxk), (unsigned)ng),
                     dim3(256), 0, s, (const GateDesc*)d_gd,
                     (const double*)eng.sig, (const int64_t*)d_kept,
                     (const double*)eng.keptw, cols_p,
                     (chosen || learned) ? 0 : 1);
  QTN_CHECK_HIP(hipGetLastError());
  // 5. B_i = Phi V_k / |S_k|
  std::vector<GemmDesc<double>> g2(ng);
  for (int g = 0; g < ng; ++g) {
    const GateDesc& d = gd[g];
    const double nrm = std::sqrt(keptw[g]);
    GemmDesc<double>& e = g2[g];
    e.M = 2 * d.chiL;
    e.N = kept[g];
    e.K = 2 * d.chiR;
    // the chosen form normalises after recomputing |S_k| in FP64
    e.alpha = make_double2((chosen || learned) ? 1.0
                                               : (nrm > 0 ? 1.0 / nrm : 0.0),
                           0.0);
    e.beta = make_double2(0.0, 0.0);
    e.A = d.phi;
    e.lda = 2 * d.chiL;
    e.B = d.Vk;
    e.ldb = 2 * d.chiR;
    e.C = d.Bl;
    e.ldc = 2 * d.chiL;
  }
  QTN_CHECK_HIP(hipMemcpyAsync(d_g2, g2.data(), ng * sizeof(GemmDesc<double>),
                               hipMemcpyHostToDevice, s));
  QTN_TRY(launch_zgemm_grouped(s, 0, 0, d_g2, ng, maxM, maxk));
  if (chosen || learned) {
    // S_k = column norms of diag(lambda_{i-1}) Phi V_k in FP64; B_i and
    // lambda_i are normalised by |S_k| (overwriting what write_right put in
    // lambda_i from the FP32 values)
    void* dS64;
    QTN_TRY(ctx->buffer(kSlotS64, (size_t)ng * cols_p * sizeof(double),
                        &dS64));
    hipLaunchKernelGGL(left_colnorms,
                       dim3((unsigned)((maxk + 3) / 4), (unsigned)ng),
                       dim3(256), 0, s, (const GateDesc*)d_gd,
                       (const int64_t*)d_kept, (double*)dS64, cols_p);
    hipLaunchKernelGGL(left_finish, dim3(grid_for(maxM * maxk), (unsigned)ng),
                       dim3(256), 0, s, (const GateDesc*)d_gd,
                       (const int64_t*)d_kept, (const double*)dS64, cols_p);
    QTN_CHECK_HIP(hipGetLastError());
  }
  for (int g = 0; g < ng; ++g) {
    ms[g]->dims[sites[g]] = kept[g];
    if (kept_out) kept_out[g] = kept[g];
    if (disc_out) disc_out[g] = total[g] > 0 ? disc[g] / total[g] : 0.0;
  }
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
}

// A layer's gates (on one or many states sharing ctx, disjoint per state)
// in one batched pass per padded column count of X (64, 128, ...), so that
// gates of similar size share a Jacobi batch without padding small
// problems to the largest.  chimax[i] <= 0: the state's capacity.
// hint / accepted (learned form, may be NULL): per-gate rank hints and the
// range finder's outcome, as in apply_gates_core.
int32_t apply_gates_grouped(qtn_ctx* ctx, int32_t ngates,
                            qtn_mps* const* states, const int32_t* sites,
                            const double* gates, const int64_t* chimax,
                            double cutoff, int32_t form, int64_t* kept,
                            double* disc_weight,
                            const int64_t* hint = nullptr,
                            int* accepted = nullptr) {
  std::map<int64_t, std::vector<int>> bucket;
  std::map<const qtn_mps*, std::vector<char>> used;
  for (int g = 0; g < ngates; ++g) {
    const qtn_mps* m = states[g];
    QTN_ARG(m && m->ctx == ctx, "apply_layer: state NULL or on another context");
    const int i = sites[g];
    QTN_ARG(i >= 1 && i < m->n, "apply_2q: site out of range");
    std::vector<char>& u = used[m];
    if (u.empty()) u.assign(m->n + 1, 0);
    QTN_ARG(!u[i] && !u[i + 1], "apply_layer: gates overlap");
    u[i] = u[i + 1] = 1;
    const int64_t M = 2 * m->dims[i - 1], N = 2 * m->dims[i + 1];
    bucket[round_up64(std::min(M, N))].push_back(g);
  }
  // A non-finite gate is refused before any group runs, so such a layer
  // leaves every state untouched.  (Errors that only the sweeps can find --
  // no convergence, a zero two-site tensor -- surface after the other
  // groups have written their sites; the header documents this.)
  for (int64_t e = 0; e < (int64_t)ngates * 32; ++e)
    if (!std::isfinite(gates[e])) {
      set_error("apply_layer: gate contains NaN or Inf");
      return QTN_ERR_NOTFINITE;
    }
  // Groups and their cost (gates x columns^2, the Jacobi work); the
  // costliest is halved while it exceeds a quarter of the total and keeps at
  // least kMinChunk gates per half, so one large size class does not set
  // the critical path alone.
  constexpr int kExec = 4, kMinChunk = 4;
  struct Group {
    std::vector<int> idx;
    double cost;
  };
  std::vector<Group> groups;
  double total = 0.0;
  for (auto& kv : bucket) {
    const double c = (double)kv.second.size() * (double)kv.first *
                     (double)kv.first;
    groups.push_back({kv.second, c});
    total += c;
  }
  for (;;) {
    int big = 0;
    for (int i = 1; i < (int)groups.size(); ++i)
      if (groups[i].cost > groups[big].cost) big = i;
    Group& g = groups[big];
    if (g.cost * kExec <= total || (int)g.idx.size() < 2 * kMinChunk) break;
    const size_t half = g.idx.size() / 2;
    Group h{std::vector<int>(g.idx.begin() + half, g.idx.end()),
            g.cost * (double)(g.idx.size() - half) / (double)g.idx.size()};
    g.cost -= h.cost;
    g.idx.resize(half);
    groups.push_back(std::move(h));
  }
  // One batched pass per group.  The groups are independent (their gates
  // act on disjoint sites and read no bond another group writes), so they
  // run concurrently on up to four executors (four streams match the
  // device's hardware queues): executor 0 is this context and thread,
  // executors 1..3 worker contexts (own stream and scratch) driven by their
  // own host threads.  Groups go costliest first to the least loaded
  // executor.  Per-kernel profiling, when on, covers the workers too.
  const int nexec = std::min<int>(kExec, (int)groups.size());
  std::vector<int> order(groups.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    return groups[a].cost > groups[b].cost;
  });
  std::vector<std::vector<int>> plan(nexec);
  std::vector<double> load(nexec, 0.0);
  for (int gi : order) {
    int e = 0;
    for (int k = 1; k < nexec; ++k)
      if (load[k] < load[e]) e = k;
    plan[e].push_back(gi);
    load[e] += groups[gi].cost;
  }
  auto run_group = [&](qtn_ctx* c, const std::vector<int>& idx) -> int32_t {
    const int nb = (int)idx.size();
    std::vector<qtn_mps*> ms(nb);
    std::vector<int32_t> st(nb);
    std::vector<double> gs((size_t)nb * 32);
    std::vector<int64_t> ch(nb), kp(nb), hn(hint ? nb : 0);
    std::vector<double> dw(nb);
    std::vector<int> ac(nb, 0);
    for (int b = 0; b < nb; ++b) {
      const int g = idx[b];
      ms[b] = states[g];
      st[b] = sites[g];
      std::memcpy(&gs[(size_t)b * 32], gates + (size_t)g * 32,
                  32 * sizeof(double));
      ch[b] = chimax[g] > 0 ? chimax[g] : states[g]->cap;
      if (hint) hn[b] = hint[g];
    }
    QTN_TRY(apply_gates_core(c, nb, ms.data(), st.data(), gs.data(),
                             cutoff, ch.data(), form, kp.data(), dw.data(),
                             nullptr, nullptr, hint ? hn.data() : nullptr,
                             ac.data()));
    for (int b = 0; b < nb; ++b) {
      if (kept) kept[idx[b]] = kp[b];
      if (disc_weight) disc_weight[idx[b]] = dw[b];
      if (accepted) accepted[idx[b]] = ac[b];
    }
    return QTN_OK;
  };
  if (nexec <= 1) {
    for (const auto& g : groups) QTN_TRY(run_group(ctx, g.idx));
    return QTN_OK;
  }
  std::vector<qtn_ctx*> exec(nexec, ctx);
  for (int e = 1; e < nexec; ++e) {
    QTN_TRY(ctx->worker(e - 1, &exec[e]));
    exec[e]->jac_tiling = ctx->jac_tiling;  // the context's tuned tilings
  }
  // the workers' streams start after everything queued on this one
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> status(nexec, QTN_OK);
  std::vector<std::string> message(nexec);
  auto drive = [&](int e) {
    // an exception must not leave a worker thread (std::terminate)
    try {
      if (e > 0 && hipSetDevice(exec[e]->device) != hipSuccess) {
        status[e] = QTN_ERR_HIP;
        message[e] = "hipSetDevice failed in a worker thread";
        return;
      }
      for (int gi : plan[e]) {
        status[e] = run_group(exec[e], groups[gi].idx);
        if (status[e] != QTN_OK) {
          message[e] = qtn_last_error();  // thread-local: carried back below
          return;
        }
      }
    } catch (...) {
      status[e] = abi_exception();
      message[e] = qtn_last_error();
    }
  };
  std::vector<std::thread> threads;
  std::vector<int> here;  // executors whose thread could not be started
  for (int e = 1; e < nexec; ++e) {
    try {
      threads.emplace_back(drive, e);
    } catch (...) {
      here.push_back(e);
    }
  }
  drive(0);
  for (int e : here) drive(e);
  for (auto& t : threads) t.join();
  for (int e = 0; e < nexec; ++e)
    if (status[e] != QTN_OK) {
      set_error(message[e]);
      return status[e];
    }
  return QTN_OK;
}

// All gates on one state with one cap.
int32_t mps_apply_gates(qtn_mps* m, int ng, const int32_t* sites,
                        const double* gates, double cutoff, int64_t chimax,
                        int form, int64_t* kept_out, double* disc_out,
                        double* spec_out, int64_t* spec_len) {
  std::vector<qtn_mps*> ms(ng, m);
  std::vector<int64_t> chis(ng, chimax);
  return apply_gates_core(m->ctx, ng, ms.data(), sites, gates, cutoff,
                          chis.data(), form, kept_out, disc_out, spec_out,
                          spec_len);
}

// ---------------------------------------------------------------------------
// Right environments R_j (chi_j x chi_j), j = 0..n (R_n = 1), stored densely
// at slot offsets; returns the device pointers.
int32_t mps_right_envs(const qtn_mps* m, std::vector<double2*>& R) {
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  size_t total = 0;
  std::vector<size_t> off(m->n + 1);
  for (int j = 0; j <= m->n; ++j) {
    off[j] = total;
    total += (size_t)m->dims[j] * m->dims[j];
  }
  void* base;
  QTN_TRY(ctx->buffer(kSlotEnv, total * sizeof(double2), &base));
  void* t2;
  int64_t maxb = 1;
  for (int j = 0; j <= m->n; ++j) maxb = std::max(maxb, m->dims[j]);
  QTN_TRY(ctx->buffer(kSlotT2, (size_t)2 * maxb * maxb * sizeof(double2), &t2));
  R.resize(m->n + 1);
  for (int j = 0; j <= m->n; ++j) R[j] = (double2*)base + off[j];
  hipLaunchKernelGGL(fill_one, dim3(1), dim3(1), 0, s, R[m->n]);
  for (int j = m->n; j >= 1; --j) {
    const int64_t cl = m->dims[j - 1], cr = m->dims[j];
    // T2 = B_j (2cl x cr) R_j
    QTN_TRY(launch_zgemm(s, 0, 0, 2 * cl, cr, cr, make_double2(1, 0),
                         m->site[j - 1], 2 * cl, 0, R[j], cr, 0,
                         make_double2(0, 0), (double2*)t2, 2 * cl, 0, 1));
    // R_{j-1} = T2 (cl x 2cr) B_j^H (2cr x cl)
    QTN_TRY(launch_zgemm(s, 0, 2, cl, cl, 2 * cr, make_double2(1, 0),
                         (const double2*)t2, cl, 0, m->site[j - 1], cl, 0,
                         make_double2(0, 0), R[j - 1], cl, 0, 1));
  }
  return QTN_OK;
}

}  // namespace qtn

using namespace qtn;

extern "C" {

int32_t qtn_mps_create(qtn_ctx* ctx, int32_t nsites, int64_t chimax_cap,
                       qtn_mps** out) try {
  clear_error();
  QTN_ARG(ctx && out, "qtn_mps_create: NULL argument");
  QTN_ARG(nsites >= 2, "qtn_mps_create: need at least 2 sites");
  QTN_ARG(chimax_cap >= 1, "qtn_mps_create: chimax_cap must be >= 1");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  qtn_mps* m = new qtn_mps();
  m->ctx = ctx;
  m->n = nsites;
  m->cap = chimax_cap;
  m->dims.assign(nsites + 1, 1);
  m->maxb.resize(nsites + 1);
  for (int j = 0; j <= nsites; ++j)
    m->maxb[j] = std::min(pow2_cap(j, chimax_cap), pow2_cap(nsites - j,
                                                            chimax_cap));
  size_t bytes = 0;
  std::vector<size_t> soff(nsites), loff(nsites + 1);
  for (int j = 0; j < nsites; ++j) {
    soff[j] = bytes;
    bytes += (size_t)m->maxb[j] * 2 * m->maxb[j + 1] * sizeof(double2);
    bytes = (bytes + 255) & ~(size_t)255;
  }
  for (int j = 0; j <= nsites; ++j) {
    loff[j] = bytes;
    bytes += (size_t)m->maxb[j] * sizeof(double);
    bytes = (bytes + 255) & ~(size_t)255;
  }
  hipError_t e = hipMalloc(&m->arena, bytes);
  if (e != hipSuccess) {
    delete m;
    set_error(std::string("qtn_mps_create: hipMalloc failed: ") +
              hipGetErrorString(e));
    return QTN_ERR_OOM;
  }
  m->site.resize(nsites);
  m->lam.resize(nsites + 1);
  for (int j = 0; j < nsites; ++j)
    m->site[j] = (double2*)((char*)m->arena + soff[j]);
  for (int j = 0; j <= nsites; ++j)
    m->lam[j] = (double*)((char*)m->arena + loff[j]);
  // |0...0>: every site (1,2,1) = [1, 0], every lambda = [1]
  QTN_CHECK_HIP(hipMemsetAsync(m->arena, 0, bytes, ctx->stream));
  for (int j = 0; j < nsites; ++j)
    hipLaunchKernelGGL(fill_one, dim3(1), dim3(1), 0, ctx->stream, m->site[j]);
  for (int j = 0; j <= nsites; ++j)
    hipLaunchKernelGGL(fill_lambda_one, dim3(1), dim3(1), 0, ctx->stream,
                       m->lam[j]);
  QTN_CHECK_HIP(hipGetLastError());
  QTN_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = m;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_destroy(qtn_mps* m) try {
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

int32_t qtn_mps_copy(const qtn_mps* a, qtn_mps* b) try {
  clear_error();
  QTN_ARG(a && b, "qtn_mps_copy: NULL argument");
  QTN_ARG(a->n == b->n && a->cap == b->cap,
          "qtn_mps_copy: states differ in size or capacity");
  QTN_CHECK_HIP(hipSetDevice(a->ctx->device));
  size_t bytes = 0;
  for (int j = 0; j < a->n; ++j) {
    bytes += (size_t)a->maxb[j] * 2 * a->maxb[j + 1] * sizeof(double2);
    bytes = (bytes + 255) & ~(size_t)255;
  }
  for (int j = 0; j <= a->n; ++j) {
    bytes += (size_t)a->maxb[j] * sizeof(double);
    bytes = (bytes + 255) & ~(size_t)255;
  }
  QTN_CHECK_HIP(hipMemcpyAsync(b->arena, a->arena, bytes,
                               hipMemcpyDeviceToDevice, a->ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
  b->dims = a->dims;
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_bond_dims(const qtn_mps* m, int64_t* dims) try {
  clear_error();
  QTN_ARG(m && dims, "qtn_mps_bond_dims: NULL argument");
  std::copy(m->dims.begin(), m->dims.end(), dims);
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_bond_spectrum(const qtn_mps* m, int32_t bond, double* S,
                              int64_t* k) try {
  clear_error();
  QTN_ARG(m && S && k, "qtn_mps_bond_spectrum: NULL argument");
  QTN_ARG(bond >= 0 && bond <= m->n, "qtn_mps_bond_spectrum: bad bond");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  *k = m->dims[bond];
  QTN_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  QTN_CHECK_HIP(hipMemcpy(S, m->lam[bond], *k * sizeof(double),
                          hipMemcpyDeviceToHost));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_get_site(const qtn_mps* m, int32_t site, double* out,
                         int64_t* dims) try {
  clear_error();
  QTN_ARG(m && dims, "qtn_mps_get_site: NULL argument");
  QTN_ARG(site >= 1 && site <= m->n, "qtn_mps_get_site: bad site");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  dims[0] = m->dims[site - 1];
  dims[1] = 2;
  dims[2] = m->dims[site];
  if (out) {
    QTN_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
    QTN_CHECK_HIP(hipMemcpy(out, m->site[site - 1],
                            (size_t)dims[0] * 2 * dims[2] * sizeof(double2),
                            hipMemcpyDeviceToHost));
  }
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_1q(qtn_mps* m, int32_t site, const double* u) try {
  clear_error();
  QTN_ARG(m && u, "qtn_mps_apply_1q: NULL argument");
  QTN_ARG(site >= 1 && site <= m->n, "qtn_mps_apply_1q: site out of range");
  qtn_ctx* ctx = m->ctx;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  // the gate travels in the kernel arguments: no upload, and no wait for
  // the kernel (later work on the context's stream is ordered after it)
  Mat2 g;
  for (int i = 0; i < 4; ++i) g.u[i] = make_double2(u[2 * i], u[2 * i + 1]);
  const int64_t cl = m->dims[site - 1], cr = m->dims[site];
  hipLaunchKernelGGL(apply_1q_kernel, dim3(grid_for(cl * cr)), dim3(256), 0,
                     ctx->stream, m->site[site - 1], cl, cr, g);
  QTN_CHECK_HIP(hipGetLastError());
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_2q(qtn_mps* m, int32_t site, const double* g,
                         double cutoff, int64_t chimax, int32_t form,
                         int64_t* kept, double* disc_weight, double* spectrum,
                         int64_t* spectrum_len) try {
  clear_error();
  QTN_ARG(m && g, "qtn_mps_apply_2q: NULL argument");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_ARG(form == QTN_FORM_PLAIN || form == QTN_FORM_CHOSEN,
          "qtn_mps_apply_2q: unknown form");
  QTN_ARG(chimax <= m->cap, "qtn_mps_apply_2q: chimax above capacity");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t cm = (chimax > 0) ? chimax : m->cap;
  return mps_apply_gates(m, 1, &site, g, cutoff, cm, form, kept,
                         disc_weight, spectrum, spectrum_len);
} catch (...) {
  return qtn::abi_exception();
}

// ---- SWAP routing on the device state (switch!, permute!, long-range
// gates).  A SWAP of sites (k, k+1) is the two-site update with the SWAP
// gate (index s_k + 2 s_{k+1}: |01> <-> |10>).
namespace {
const double kSwapGate[32] = {1, 0, 0, 0, 0, 0, 0, 0,   // column 0: |00>
                              0, 0, 0, 0, 1, 0, 0, 0,   // column 1 -> row 2
                              0, 0, 1, 0, 0, 0, 0, 0,   // column 2 -> row 1
                              0, 0, 0, 0, 0, 0, 1, 0};  // column 3: |11>

int32_t swap_sites(qtn_mps* m, int k, double cutoff, int64_t cm, int form) {
  return mps_apply_gates(m, 1, &k, kSwapGate, cutoff, cm, form, nullptr,
                         nullptr, nullptr, nullptr);
}

// switch!(mps, i, j) (src/switch.jl:63-86) on the device state: wires are
// numbered from the most significant (wire w is site n - w + 1); site lo
// travels to hi by adjacent swaps and the qubit met last travels back.
int32_t mps_switch(qtn_mps* m, int32_t i, int32_t j) {
  if (!(i > 0 && j > 0)) {
    set_error("Wire indices `i` and `j` must be positive");
    return QTN_ERR_ARG;
  }
  if (!(i <= m->n && j <= m->n)) {
    set_error("Indices to swap `i` and `j` must be less than or equal to "
              "the number of open wires in MPS");
    return QTN_ERR_ARG;
  }
  if (i == j) return QTN_OK;
  const int a = m->n - i + 1, b = m->n - j + 1;
  const int lo = std::min(a, b), hi = std::max(a, b);
  // untruncated (only exact zeros are dropped), up to the state's capacity
  for (int k = lo; k < hi; ++k)
    QTN_TRY(swap_sites(m, k, 0.0, m->cap, QTN_FORM_PLAIN));
  for (int k = hi - 2; k >= lo; --k)
    QTN_TRY(swap_sites(m, k, 0.0, m->cap, QTN_FORM_PLAIN));
  return QTN_OK;
}
}  // namespace

int32_t qtn_mps_switch(qtn_mps* m, int32_t i, int32_t j) try {
  clear_error();
  QTN_ARG(m, "qtn_mps_switch: NULL state");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  return mps_switch(m, i, j);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_permute(qtn_mps* m, const int64_t* order, int64_t len) try {
  clear_error();
  QTN_ARG(m && (len == 0 || order), "qtn_mps_permute: NULL argument");
  const int n = m->n;
  QTN_ARG(len == n, "Given permutation must be same length as number of "
                    "Tensors in MPS");
  std::vector<int64_t> seen(order, order + len);
  std::sort(seen.begin(), seen.end());
  QTN_ARG(std::adjacent_find(seen.begin(), seen.end()) == seen.end(),
          "Permutation order cannot contain repeat values");
  QTN_ARG(seen.empty() || seen.front() > 0,
          "Permutation order can only contain positive values");
  QTN_ARG(seen.empty() || seen.back() == len,
          "Wire numbers in permutation order cannot exceed number of wires "
          "in MPS");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  // Base.permute! (src/switch.jl:94-106): wire order[i] is brought to
  // position i by switch!(mps, i, loc), tracking where each wire sits
  std::vector<int64_t> seq(n);
  for (int q = 0; q < n; ++q) seq[q] = q + 1;
  for (int q = 0; q < n; ++q) {
    const int loc =
        (int)(std::find(seq.begin(), seq.end(), order[q]) - seq.begin());
    QTN_TRY(mps_switch(m, q + 1, loc + 1));
    std::swap(seq[loc], seq[q]);
  }
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_2q_long(qtn_mps* m, int32_t i, int32_t j,
                              const double* g, double cutoff, int64_t chimax,
                              int32_t form, int64_t* kept,
                              double* disc_weight) try {
  clear_error();
  QTN_ARG(m && g, "qtn_mps_apply_2q_long: NULL argument");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_ARG(form == QTN_FORM_PLAIN || form == QTN_FORM_CHOSEN,
          "qtn_mps_apply_2q_long: unknown form");
  QTN_ARG(chimax <= m->cap, "qtn_mps_apply_2q_long: chimax above capacity");
  QTN_ARG(i >= 1 && i <= m->n && j >= 1 && j <= m->n && i != j,
          "qtn_mps_apply_2q_long: sites must be distinct and in 1..n");
  for (int e = 0; e < 32; ++e)
    QTN_ARG(std::isfinite(g[e]), "qtn_mps_apply_2q_long: gate contains NaN "
                                 "or Inf");
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t cm = (chimax > 0) ? chimax : m->cap;
  const int lo = std::min(i, j), hi = std::max(i, j);
  // the gate's index is s_i + 2 s_j; applied to (lo, lo + 1) it must read
  // s_lo + 2 s_hi, so the two bits trade places when i > j
  double gl[32];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const int rr = i < j ? r : ((r & 1) << 1) | (r >> 1);
      const int cc = i < j ? c : ((c & 1) << 1) | (c >> 1);
      gl[2 * (r + 4 * c)] = g[2 * (rr + 4 * cc)];
      gl[2 * (r + 4 * c) + 1] = g[2 * (rr + 4 * cc) + 1];
    }
  // the qubit of site hi travels down to lo + 1, the gate acts on
  // (lo, lo + 1), and the qubit travels back; every step is a two-site
  // update truncated by the same rule
  for (int k = hi - 1; k > lo; --k) QTN_TRY(swap_sites(m, k, cutoff, cm, form));
  int lo_site = lo;
  QTN_TRY(mps_apply_gates(m, 1, &lo_site, gl, cutoff, cm, form, kept,
                          disc_weight, nullptr, nullptr));
  for (int k = lo + 1; k < hi; ++k) QTN_TRY(swap_sites(m, k, cutoff, cm, form));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_layer(qtn_mps* m, int32_t ngates, const int32_t* sites,
                            const double* gates, double cutoff,
                            int64_t chimax, int32_t form, int64_t* kept,
                            double* disc_weight) try {
  clear_error();
  QTN_ARG(m && (ngates == 0 || (sites && gates)),
          "qtn_mps_apply_layer: NULL argument");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_ARG(form == QTN_FORM_PLAIN || form == QTN_FORM_CHOSEN,
          "qtn_mps_apply_layer: unknown form");
  QTN_ARG(chimax <= m->cap, "qtn_mps_apply_layer: chimax above capacity");
  if (ngates == 0) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t cm = (chimax > 0) ? chimax : m->cap;
  std::vector<qtn_mps*> ms(ngates, m);
  std::vector<int64_t> chis(ngates, cm);
  return apply_gates_grouped(m->ctx, ngates, ms.data(), sites, gates,
                             chis.data(), cutoff, form, kept, disc_weight);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_amplitudes(const qtn_mps* m, int64_t nb, const uint8_t* bits,
                           double* out) try {
  clear_error();
  QTN_ARG(m && bits && out, "qtn_mps_amplitudes: NULL argument");
  QTN_ARG(nb >= 1, "qtn_mps_amplitudes: nb must be >= 1");
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  int64_t maxb = 1;
  for (int j = 0; j <= m->n; ++j) maxb = std::max(maxb, m->dims[j]);
  void *dbits, *damp;
  QTN_TRY(ctx->buffer(kSlotBits, (size_t)nb * m->n, &dbits));
  QTN_TRY(ctx->buffer(kSlotAmp, (size_t)2 * nb * maxb * sizeof(double2),
                      &damp));
  QTN_CHECK_HIP(hipMemcpyAsync(dbits, bits, (size_t)nb * m->n,
                               hipMemcpyHostToDevice, s));
  double2* V = (double2*)damp;
  double2* Vn = V + nb * maxb;
  // V starts as nb x 1 of ones
  std::vector<double> ones(2 * nb, 0.0);
  for (int64_t i = 0; i < nb; ++i) ones[2 * i] = 1.0;
  QTN_CHECK_HIP(hipMemcpyAsync(V, ones.data(), nb * sizeof(double2),
                               hipMemcpyHostToDevice, s));
  for (int j = 0; j < m->n; ++j) {
    const int64_t cl = m->dims[j], cr = m->dims[j + 1];
    hipLaunchKernelGGL(amp_step, dim3(grid_for(nb * cr)), dim3(256), 0, s,
                       (const double2*)V, nb, cl, cr,
                       (const double2*)m->site[j], (const uint8_t*)dbits,
                       m->n, j, Vn);
    std::swap(V, Vn);
  }
  QTN_CHECK_HIP(hipGetLastError());
  QTN_CHECK_HIP(hipMemcpyAsync(out, V, nb * sizeof(double2),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_expect_z(const qtn_mps* m, double* out) try {
  clear_error();
  QTN_ARG(m && out, "qtn_mps_expect_z: NULL argument");
  qtn_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  std::vector<double2*> R;
  QTN_TRY(mps_right_envs(m, R));
  int64_t maxb = 1;
  for (int j = 0; j <= m->n; ++j) maxb = std::max(maxb, m->dims[j]);
  void *t1, *ebuf, *red;
  QTN_TRY(ctx->buffer(kSlotT1, (size_t)4 * maxb * maxb * sizeof(double2),
                      &t1));
  QTN_TRY(ctx->buffer(kSlotT2 + 10, (size_t)2 * maxb * maxb * sizeof(double2),
                      &ebuf));
  QTN_TRY(ctx->buffer(kSlotRed, (size_t)4 * m->n * sizeof(double), &red));
  double2* T1 = (double2*)t1;
  double2* Y = T1 + 2 * maxb * maxb;
  double2* E = (double2*)ebuf;
  double2* En = E + maxb * maxb;
  hipLaunchKernelGGL(fill_one, dim3(1), dim3(1), 0, s, E);
  for (int j = 1; j <= m->n; ++j) {
    const int64_t cl = m->dims[j - 1], cr = m->dims[j];
    const double2* B = m->site[j - 1];
    // T1^s = E B^s (batch over s; B^s rows offset s*cl, ld 2cl)
    QTN_TRY(launch_zgemm(s, 0, 0, cl, cr, cl, make_double2(1, 0), E, cl, 0, B,
                         2 * cl, cl, make_double2(0, 0), T1, 2 * cl, cl, 2));
    // Y = T1 R_j
    QTN_TRY(launch_zgemm(s, 0, 0, 2 * cl, cr, cr, make_double2(1, 0), T1,
                         2 * cl, 0, R[j], cr, 0, make_double2(0, 0), Y,
                         2 * cl, 0, 1));
    hipLaunchKernelGGL(env_reduce, dim3(1), dim3(256), 0, s, B,
                       (const double2*)Y, cl, cr, (double*)red + 4 * (j - 1));
    // E_j = B^H T1  (K = 2cl)
    QTN_TRY(launch_zgemm(s, 2, 0, cr, cr, 2 * cl, make_double2(1, 0), B,
                         2 * cl, 0, T1, 2 * cl, 0, make_double2(0, 0), En, cr,
                         0, 1));
    std::swap(E, En);
  }
  QTN_CHECK_HIP(hipGetLastError());
  std::vector<double> h(4 * m->n);
  QTN_CHECK_HIP(hipMemcpyAsync(h.data(), red, 4 * m->n * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < m->n; ++j) out[j] = h[4 * j + 2] / h[4 * j + 0];
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_overlap(const qtn_mps* a, const qtn_mps* b, double* out) try {
  clear_error();
  QTN_ARG(a && b && out, "qtn_mps_overlap: NULL argument");
  QTN_ARG(a->n == b->n, "qtn_mps_overlap: different number of sites");
  QTN_ARG(a->ctx == b->ctx, "qtn_mps_overlap: states from two contexts");
  qtn_ctx* ctx = a->ctx;
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  int64_t mb = 1;
  for (int j = 0; j <= a->n; ++j)
    mb = std::max(mb, std::max(a->dims[j], b->dims[j]));
  void *t1, *ebuf;
  QTN_TRY(ctx->buffer(kSlotT1, (size_t)4 * mb * mb * sizeof(double2), &t1));
  QTN_TRY(ctx->buffer(kSlotT2 + 10, (size_t)2 * mb * mb * sizeof(double2),
                      &ebuf));
  double2* T1 = (double2*)t1;
  double2* E = (double2*)ebuf;
  double2* En = E + mb * mb;
  hipLaunchKernelGGL(fill_one, dim3(1), dim3(1), 0, s, E);
  for (int j = 1; j <= a->n; ++j) {
    const int64_t al = a->dims[j - 1], ar = a->dims[j];
    const int64_t bl = b->dims[j - 1], br = b->dims[j];
    // T1^s = E (al x bl) B^s (bl x br) -> stacked (2al x br)
    QTN_TRY(launch_zgemm(s, 0, 0, al, br, bl, make_double2(1, 0), E, al, 0,
                         b->site[j - 1], 2 * bl, bl, make_double2(0, 0), T1,
                         2 * al, al, 2));
    // E_j = A^H (ar x 2al) T1 (2al x br)
    QTN_TRY(launch_zgemm(s, 2, 0, ar, br, 2 * al, make_double2(1, 0),
                         a->site[j - 1], 2 * al, 0, T1, 2 * al, 0,
                         make_double2(0, 0), En, ar, 0, 1));
    std::swap(E, En);
  }
  QTN_CHECK_HIP(hipMemcpyAsync(out, E, sizeof(double2),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_norm2(const qtn_mps* m, double* out) try {
  double v[2];
  QTN_TRY(qtn_mps_overlap(m, m, v));
  *out = v[0];
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_1q_multi(int32_t ngates, qtn_mps* const* states,
                               const int32_t* sites, const double* us) try {
  clear_error();
  QTN_ARG(ngates >= 0, "qtn_mps_apply_1q_multi: negative count");
  if (ngates == 0) return QTN_OK;
  QTN_ARG(states && sites && us, "qtn_mps_apply_1q_multi: NULL argument");
  qtn_ctx* ctx = states[0] ? states[0]->ctx : nullptr;
  QTN_ARG(ctx != nullptr, "qtn_mps_apply_1q_multi: NULL state");
  std::vector<OneQDesc> dd(ngates);
  int64_t maxt = 1;
  // one launch: a site may carry only one of its gates
  std::set<std::pair<const qtn_mps*, int>> seen;
  for (int g = 0; g < ngates; ++g) {
    const qtn_mps* m = states[g];
    QTN_ARG(m && m->ctx == ctx, "qtn_mps_apply_1q_multi: bad state");
    const int site = sites[g];
    QTN_ARG(site >= 1 && site <= m->n, "qtn_mps_apply_1q_multi: bad site");
    QTN_ARG(seen.insert({m, site}).second,
            "qtn_mps_apply_1q_multi: two gates on one site");
    dd[g].B = m->site[site - 1];
    dd[g].chiL = m->dims[site - 1];
    dd[g].chiR = m->dims[site];
    for (int k = 0; k < 4; ++k)
      dd[g].u[k] = make_double2(us[8 * g + 2 * k], us[8 * g + 2 * k + 1]);
    maxt = std::max(maxt, dd[g].chiL * dd[g].chiR);
  }
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  void* dptr;
  QTN_TRY(ctx->buffer(kSlot1q, (size_t)ngates * sizeof(OneQDesc), &dptr));
  QTN_CHECK_HIP(hipMemcpyAsync(dptr, dd.data(), ngates * sizeof(OneQDesc),
                               hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(apply_1q_multi_kernel,
                     dim3(grid_for(maxt), (unsigned)ngates), dim3(256), 0,
                     ctx->stream, (const OneQDesc*)dptr);
  QTN_CHECK_HIP(hipGetLastError());
  // enqueued without a wait: every later call that reads or rewrites these
  // states (or this descriptor buffer) is ordered after it on the stream
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_layer_multi(int32_t ngates, qtn_mps* const* states,
                                  const int32_t* sites, const double* gates,
                                  const int64_t* chimax, double cutoff,
                                  int32_t form, int64_t* kept,
                                  double* disc_weight) try {
  clear_error();
  QTN_ARG(ngates >= 0, "qtn_mps_apply_layer_multi: negative count");
  if (ngates == 0) return QTN_OK;
  QTN_ARG(states && sites && gates && chimax,
          "qtn_mps_apply_layer_multi: NULL argument");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_ARG(form == QTN_FORM_PLAIN || form == QTN_FORM_CHOSEN,
          "qtn_mps_apply_layer_multi: unknown form");
  qtn_ctx* ctx = states[0] ? states[0]->ctx : nullptr;
  QTN_ARG(ctx != nullptr, "qtn_mps_apply_layer_multi: NULL state");
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  return apply_gates_grouped(ctx, ngates, states, sites, gates, chimax,
                             cutoff, form, kept, disc_weight);
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"

namespace qtn {
namespace {

constexpr int kSlotFeatDesc = 170, kSlotFeat = 171;

// Device feature columns (kGateFeatures x ngates floats, in context slot
// kSlotFeat) of the gates on (sites[i], sites[i] + 1).
int32_t gather_features(qtn_mps* m, int ng, const int32_t* sites,
                        const double* schmidt, int64_t chimax,
                        float** dX_out) {
  std::vector<FeatDesc> fd(ng);
  for (int g = 0; g < ng; ++g) {
    const int w = sites[g];
    QTN_ARG(w >= 1 && w < m->n, "gate features: site out of range");
    FeatDesc& d = fd[g];
    d.lam = m->lam[w];
    d.len = m->dims[w];
    for (int j = 0; j < 4; ++j) d.sch[j] = (float)schmidt[4 * g + j];
    d.pos = (float)((double)w / (double)m->n);
    for (int j = 0; j < 3; ++j)
      d.dims[j] = (float)((double)m->dims[w - 1 + j] / (double)chimax);
  }
  void *pd, *px;
  QTN_TRY(m->ctx->buffer(kSlotFeatDesc, (size_t)ng * sizeof(FeatDesc), &pd));
  QTN_TRY(m->ctx->buffer(kSlotFeat,
                         (size_t)ng * kGateFeatures * sizeof(float), &px));
  hipStream_t s = m->ctx->stream;
  QTN_CHECK_HIP(hipMemcpyAsync(pd, fd.data(), ng * sizeof(FeatDesc),
                               hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(gate_features, dim3(ng), dim3(128), 0, s,
                     (const FeatDesc*)pd, (float*)px);
  QTN_CHECK_HIP(hipGetLastError());
  *dX_out = (float*)px;
  return QTN_OK;
}

}  // namespace
}  // namespace qtn

extern "C" {

int32_t qtn_gate_schmidt(qtn_ctx* ctx, int32_t ngates, const double* gates,
                         double* out) try {
  clear_error();
  QTN_ARG(ctx && (ngates == 0 || (gates && out)),
          "qtn_gate_schmidt: NULL argument");
  QTN_ARG(ngates >= 0, "qtn_gate_schmidt: negative count");
  if (ngates == 0) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(ctx->device));
  void *pg, *po;
  QTN_TRY(ctx->buffer(kSlotFeatDesc, (size_t)ngates * 16 * sizeof(double2),
                      &pg));
  QTN_TRY(ctx->buffer(kSlotFeat, (size_t)ngates * 4 * sizeof(double), &po));
  hipStream_t s = ctx->stream;
  QTN_CHECK_HIP(hipMemcpyAsync(pg, gates,
                               (size_t)ngates * 16 * sizeof(double2),
                               hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(gate_schmidt, dim3((ngates + 63) / 64), dim3(64), 0, s,
                     ngates, (const double2*)pg, (double*)po);
  QTN_CHECK_HIP(hipGetLastError());
  QTN_CHECK_HIP(hipMemcpyAsync(out, po, (size_t)ngates * 4 * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  QTN_CHECK_HIP(hipStreamSynchronize(s));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_gate_features(qtn_mps* m, int32_t ngates,
                              const int32_t* sites, const double* schmidt,
                              int64_t chimax, float* out) try {
  clear_error();
  QTN_ARG(m && (ngates == 0 || (sites && schmidt && out)),
          "qtn_mps_gate_features: NULL argument");
  QTN_ARG(ngates >= 0, "qtn_mps_gate_features: negative count");
  if (ngates == 0) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t cm = chimax > 0 ? chimax : m->cap;
  float* dX;
  QTN_TRY(gather_features(m, ngates, sites, schmidt, cm, &dX));
  QTN_CHECK_HIP(hipMemcpyAsync(out, dX,
                               (size_t)ngates * kGateFeatures * sizeof(float),
                               hipMemcpyDeviceToHost, m->ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_layer_predicted(qtn_mps* m, int32_t ngates,
                                      const int32_t* sites,
                                      const double* gates,
                                      const double* schmidt, double cutoff,
                                      int64_t chimax, qtn_mlp* mlp,
                                      double margin, int64_t* kept,
                                      double* disc_weight, int32_t* accepted,
                                      int64_t* hints) try {
  clear_error();
  QTN_ARG(m && mlp && (ngates == 0 || (sites && gates && schmidt)),
          "qtn_mps_apply_layer_predicted: NULL argument");
  QTN_ARG(mlp_nin(mlp) == kGateFeatures,
          "qtn_mps_apply_layer_predicted: the MLP must take 72 features");
  QTN_ARG(mlp_ctx(mlp) == m->ctx,
          "qtn_mps_apply_layer_predicted: MLP and state on different "
          "contexts");
  QTN_ARG(chimax <= m->cap, "qtn_mps_apply_layer_predicted: chimax above "
                            "capacity");
  if (ngates == 0) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t cm = chimax > 0 ? chimax : m->cap;
  float* dX;
  QTN_TRY(gather_features(m, ngates, sites, schmidt, cm, &dX));
  float* dY;
  QTN_TRY(mlp_predict_device(mlp, ngates, dX, &dY));
  std::vector<float> y(ngates);
  QTN_CHECK_HIP(hipMemcpyAsync(y.data(), dY, ngates * sizeof(float),
                               hipMemcpyDeviceToHost, m->ctx->stream));
  QTN_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  std::vector<int64_t> h(ngates);
  for (int g = 0; g < ngates; ++g) {
    const double v = std::ceil(((double)y[g] + margin) * (double)cm);
    h[g] = std::isfinite(v) ? (int64_t)std::min<double>(
                                  std::max<double>(v, 0.0), (double)cm)
                            : 0;
    if (hints) hints[g] = h[g];
  }
  return qtn_mps_apply_layer_learned(m, ngates, sites, gates, cutoff, chimax,
                                     h.data(), kept, disc_weight, accepted);
} catch (...) {
  return qtn::abi_exception();
}

int32_t qtn_mps_apply_layer_learned(qtn_mps* m, int32_t ngates,
                                    const int32_t* sites, const double* gates,
                                    double cutoff, int64_t chimax,
                                    const int64_t* rank_hint, int64_t* kept,
                                    double* disc_weight, int32_t* accepted) try {
  clear_error();
  QTN_ARG(m && (ngates == 0 || (sites && gates && rank_hint)),
          "qtn_mps_apply_layer_learned: NULL argument");
  QTN_ARG(cutoff >= 0.0, "Error must be positive");
  QTN_ARG(chimax <= m->cap, "qtn_mps_apply_layer_learned: chimax above "
                            "capacity");
  if (ngates == 0) return QTN_OK;
  QTN_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t cm = (chimax > 0) ? chimax : m->cap;
  std::vector<qtn_mps*> ms(ngates, m);
  std::vector<int64_t> chis(ngates, cm);
  std::vector<int> acc(ngates, 0);
  QTN_TRY(apply_gates_grouped(m->ctx, ngates, ms.data(), sites, gates,
                              chis.data(), cutoff, QTN_FORM_LEARNED, kept,
                              disc_weight, rank_hint, acc.data()));
  if (accepted)
    for (int g = 0; g < ngates; ++g) accepted[g] = acc[g];
  return QTN_OK;
} catch (...) {
  return qtn::abi_exception();
}

}  // extern "C"
